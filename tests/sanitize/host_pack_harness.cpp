// The host side of the PairHMM library (region checks, packing, tile runs,
// staging, merging for a serving launch, result unpermuting, read
// preprocessing and post-processing) driven over random and edge-case
// regions, built with AddressSanitizer and UBSan and run on the CPU by
// tests/test_host_sanitize.py. No device call is made: pinned host
// allocation is redirected to malloc at link time (--wrap), and the kernel
// launchers the packing code references are stubs that are never reached.
// Usage: host_pack_harness <iterations> <seed>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#include "phmm_engine.h"
#include "phmm_adapter.h"
using namespace rvc;

extern "C" hipError_t __wrap_hipHostMalloc(void** p, size_t n, unsigned int) {
  *p = aligned_alloc(4096, (n + 4095) & ~size_t(4095));
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
extern "C" hipError_t __wrap_hipHostFree(void* p) { free(p); return hipSuccess; }

struct Reg {
  std::vector<int32_t> hoff, roff;
  std::vector<uint8_t> hb, rb, q, ii, dd, cc, mq;
  rvc_region r{};
};

static void make(std::mt19937& rng, Reg& g, bool bad_byte) {
  auto U = [&](int lo, int hi) { return lo + (int)(rng() % (uint32_t)(hi - lo + 1)); };
  const char* B = "ACGTN";
  const int H = U(1, rng() % 8 == 0 ? 300 : 24), N = U(0, rng() % 8 == 0 ? 200 : 30);
  const int hl_max = rng() % 4 == 0 ? 1500 : 400, rl_max = rng() % 4 == 0 ? 768 : 260;
  g.hoff.assign(1, 0); g.roff.assign(1, 0); g.hb.clear(); g.rb.clear(); g.q.clear(); g.ii.clear(); g.dd.clear();
  g.cc.clear(); g.mq.clear();
  const bool lowgap = rng() % 5 == 0;
  for (int h = 0; h < H; ++h) {
    const int L = rng() % 20 == 0 ? 0 : U(1, hl_max);
    for (int x = 0; x < L; ++x) g.hb.push_back(B[rng() % (rng() % 10 == 0 ? 5 : 4)]);
    g.hoff.push_back((int)g.hb.size());
  }
  for (int r = 0; r < N; ++r) {
    const int L = rng() % 20 == 0 ? 0 : U(1, rl_max);
    for (int x = 0; x < L; ++x) {
      g.rb.push_back(B[rng() % (rng() % 50 == 0 ? 5 : 4)]);
      g.q.push_back((uint8_t)U(0, 60));
      g.ii.push_back((uint8_t)(lowgap ? U(0, 90) : 45));
      g.dd.push_back((uint8_t)(lowgap ? U(0, 90) : 45));
      g.cc.push_back((uint8_t)(lowgap ? U(0, 40) : 10));
    }
    g.roff.push_back((int)g.rb.size());
    g.mq.push_back((uint8_t)U(0, 70));
  }
  if (bad_byte && !g.rb.empty()) g.rb[rng() % g.rb.size()] = "RY=an"[rng() % 5];
  // exactly the bytes the region holds (the C ABI promises nothing past
  // them), in allocations of exactly that size, so that any read past the
  // end is caught; an empty array still gets a valid pointer
  for (auto* v : {&g.hb, &g.rb, &g.q, &g.ii, &g.dd, &g.cc}) {
    if (v->empty()) v->reserve(1);
    else v->shrink_to_fit();
  }
  g.r = rvc_region{(uint32_t)H, (uint32_t)N, g.hoff.data(), g.hb.data(), g.roff.data(), g.rb.data(),
                   g.q.data(), g.ii.data(), g.dd.data(), g.cc.data()};
}

int main(int argc, char** argv) {
  g_engine.n_cu = 256;
  g_engine.max_batch_regions = 4096;
  g_engine.hbm_budget = 64ull << 30;
  const int iters = argc > 1 ? atoi(argv[1]) : 300;
  std::mt19937 rng(argc > 2 ? atoi(argv[2]) : 7);
  int ok = 0, refused = 0, merged = 0;
  for (int it = 0; it < iters; ++it) {
    const int nreg = 1 + (int)(rng() % 6);
    std::vector<Reg> regs(nreg);
    std::vector<rvc_region> rr;
    std::vector<Reg*> src;
    bool all_ok = true;
    for (auto& g : regs) {
      make(rng, g, rng() % 10 == 0);
      const int rc = validate_region(&g.r);
      if (rc != RVC_OK) { ++refused; all_ok = false; continue; }
      rr.push_back(g.r);
      src.push_back(&g);
    }
    if (rr.empty()) continue;
    // one pack of all regions, with each policy, staged as for an upload
    for (int pol_kind = 0; pol_kind < 2; ++pol_kind) {
      HostPack P;
      uint64_t pairs = 0;
      for (auto& x : rr) pairs += (uint64_t)x.n_haps * x.n_reads;
      const FP32Policy pol = pol_kind ? latency_policy(default_policy(), pairs) : default_policy();
      DevicePrep prep{};
      std::vector<uint8_t> mqs;
      const bool use_prep = rr.size() == 1 && rng() % 3 == 0;
      if (use_prep) {
        mqs = src[0]->mq;
        prep.mq = mqs.data();
        prep.min_q = (int)(rng() % 20);
        prep.pcr = (int)(rng() % 3);
      }
      if (build_host_pack(rr.data(), (uint32_t)rr.size(), pol, &P, use_prep ? &prep : nullptr) != RVC_OK) continue;
      split_runs(P);
      const BatchLayout L = batch_layout(pack_sizes(P));
      std::vector<char> img(L.bytes + 64);
      stage_pack(P, L, img.data());
      stage_results(P, L, img.data());
      // fake device results, unpermuted as the sync path does
      std::vector<double> dev(P.n_pairs_slots + 1, -1.0);
      std::vector<uint8_t> dfl(P.n_pairs_slots + 1, 1);
      std::vector<double> out;
      std::vector<uint8_t> fl;
      size_t tot = 0;
      for (auto& x : rr) tot += (size_t)x.n_haps * x.n_reads;
      out.resize(tot + 1);
      fl.resize(tot + 1);
      unpermute_results(P, dev.data(), dfl.data(), out.data(), fl.data());
      ++ok;
    }
    // separate packs merged into one staged launch, as the serving loop does
    if (all_ok && rr.size() > 1) {
      std::vector<HostPack> packs(rr.size());
      std::vector<const HostPack*> pp;
      bool good = true;
      for (size_t x = 0; x < rr.size(); ++x) {
        if (build_host_pack(&rr[x], 1, default_policy(), &packs[x], nullptr) != RVC_OK) { good = false; break; }
        pp.push_back(&packs[x]);
      }
      if (good) {
        HostPack meta;
        BatchLayout L;
        PinnedBuf buf;
        if (stage_merged(pp.data(), (uint32_t)pp.size(), &meta, &L, &buf) == RVC_OK) ++merged;
      }
    }
    // adapter: preprocessing and postprocessing of the first region
    {
      Reg& g = regs[0];
      for (uint32_t r = 0; r < g.r.n_reads; ++r) {
        const int len = g.roff[r + 1] - g.roff[r];
        std::vector<uint8_t> q(len + 1), gop(len + 1), gcp(len + 1);
        preprocess_read(g.rb.data() + g.roff[r], g.q.data() + g.roff[r], len, g.mq[r], (int)(rng() % 20),
                        (int)(rng() % 3), q.data(), gop.data(), gcp.data());
      }
      const uint32_t N = g.r.n_reads, H = g.r.n_haps;
      std::vector<double> rm((size_t)N * H + 1), hm((size_t)N * H + 1);
      std::vector<int32_t> rl(N + 1);
      std::vector<uint32_t> kept(N + 1);
      for (auto& v : rm) v = rng() % 7 == 0 ? std::numeric_limits<double>::quiet_NaN() : -(double)(rng() % 4000) / 10.0;
      for (uint32_t r = 0; r < N; ++r) rl[r] = g.roff[r + 1] - g.roff[r];
      postprocess(N, H, rm.data(), rl.data(), hm.data(), kept.data());
    }
  }
  printf("packs %d, refused regions %d, merged launches %d\n", ok, refused, merged);
  return 0;
}

// kernel-side helpers the packing code references (never launched here)
namespace rvc {
bool fp32_supported(int W, int RPL) {
  return RPL >= 1 && ((W == 16 && RPL <= 16) || (W == 32 && RPL <= 16) || (W == 64 && RPL <= 13) ||
                      (W == 8 && RPL <= 19));
}
int fp32_blocks_per_cu(int, int, bool, int, bool) { return 8; }
int fp64_blocks_per_cu(int, bool, bool) { return 8; }
hipError_t launch_copy(void*, const void*, uint64_t, hipStream_t) { abort(); }
hipError_t launch_begin(uint32_t*, uint32_t, double*, uint64_t, hipStream_t) { abort(); }
hipError_t launch_prep(const PrepArgs&, hipStream_t) { abort(); }
hipError_t launch_fp32(int, int, bool, int, bool, int, hipStream_t, const Fp32Args&) { abort(); }
hipError_t launch_fp64(int, bool, bool, int, hipStream_t, const Fp64Args&) { abort(); }
}  // namespace rvc
