// CPU baseline of the PairHMM benchmark (test and measurement infrastructure,
// never part of the product): Rovaca's region scheme run on all host cores.
//
// Per region, as ph/roDV_schedule.cpp:364-834 does it: pairs sorted by
// (haplotype length, read length) descending; greedy groups of 16 (FP32) /
// 8 (FP64) pairs with idle ratio <= 0.02 (:86-142); pass 1 runs the FP32
// inter-pair SIMD kernel on each group and flags sums below 1e-28; pass 2
// recomputes flagged pairs in FP64 groups; pass 3 runs the intra-pair kernel
// on everything left (N-containing and ungrouped pairs). One region per
// worker thread at a time, as Rovaca's asio pool runs them
// (hc/haplotypecaller_engine.cpp:119).
//
// The inter-pair kernel below restates ph/roDV_inter/pairhmm_inter.cpp:16-246
// (one pair per SIMD lane, row-major, lanes masked beyond their lengths) with
// GCC vector types of 16 floats / 8 doubles, compiled for AVX-512 (or AVX2)
// with the reference's -O3. Pass 3 calls Rovaca's own intra kernel from
// oracle/_ref when a path to it is set, else the scalar restatement.
#include <dlfcn.h>
#include <pthread.h>
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

extern "C" {
// from phmm_oracle.cpp (same library)
void orc_tables_f32(float* ph2pr128, float* mm_tri);
void orc_tables_f64(double* ph2pr128, double* mm_tri);
double orc_intra_log10(const uint8_t* hap, int H, const uint8_t* rd, const uint8_t* q, const uint8_t* i,
                       const uint8_t* d, const uint8_t* c, int R, int use_double);
}

namespace {

typedef float vf16 __attribute__((vector_size(64)));
typedef int32_t vi16 __attribute__((vector_size(64)));
typedef double vd8 __attribute__((vector_size(64)));
typedef int64_t vl8 __attribute__((vector_size(64)));

template <typename T>
struct V;
template <>
struct V<float> {
  using F = vf16;
  using I = vi16;
  using S = int32_t;
  static constexpr int W = 16;
};
template <>
struct V<double> {
  using F = vd8;
  using I = vl8;
  using S = int64_t;
  static constexpr int W = 8;
};

struct Tab {
  float pf[128];
  float mf[128 * 129 / 2];
  double pd[128];
  double md[128 * 129 / 2];
  Tab() {
    orc_tables_f32(pf, mf);
    orc_tables_f64(pd, md);
  }
  template <typename T>
  const T* ph() const;
  template <typename T>
  const T* mm() const;
};
template <>
const float* Tab::ph<float>() const { return pf; }
template <>
const double* Tab::ph<double>() const { return pd; }
template <>
const float* Tab::mm<float>() const { return mf; }
template <>
const double* Tab::mm<double>() const { return md; }

const Tab& tab() {
  static Tab t;
  return t;
}

struct PairRef {
  const uint8_t *hap, *rd, *q, *i, *d, *c;
  uint32_t H, R;
};

// One group of W pairs (ph/roDV_inter/pairhmm_inter.cpp:141-246). Returns raw sums.
template <typename T>
void inter_group(const PairRef* pr, T* results, std::vector<char>& scratch) {
  using F = typename V<T>::F;
  using I = typename V<T>::I;
  using S = typename V<T>::S;
  constexpr int W = V<T>::W;
  const Tab& tb = tab();
  const T* ph = tb.ph<T>();
  const T* mmt = tb.mm<T>();
  uint32_t min_r = UINT32_MAX, max_r = 0, min_h = UINT32_MAX, max_h = 0;
  for (int l = 0; l < W; ++l) {
    min_r = std::min(min_r, pr[l].R); max_r = std::max(max_r, pr[l].R);
    min_h = std::min(min_h, pr[l].H); max_h = std::max(max_h, pr[l].H);
  }
  const size_t need = (size_t)max_r * 8 * sizeof(F) + (size_t)max_r * sizeof(I) + (size_t)max_h * sizeof(I) +
                      3 * (size_t)(max_h + 1) * sizeof(F) + 64;
  if (scratch.size() < need) scratch.resize(need);
  char* p = scratch.data();
  p = (char*)(((uintptr_t)p + 63) & ~(uintptr_t)63);
  F* distm = (F*)p; F* one_m = distm + max_r; F* gapm = one_m + max_r; F* pmm = gapm + max_r;
  F* pmi = pmm + max_r; F* pii = pmi + max_r; F* pmd = pii + max_r; F* pdd = pmd + max_r;
  I* rs = (I*)(pdd + max_r);
  I* hs = rs + max_r;
  F* mrow = (F*)(hs + max_h);
  F* irow = mrow + (max_h + 1);
  F* drow = irow + (max_h + 1);
  const T third = (T)1.0 / (T)3.0;
  for (uint32_t k = 0; k < max_r; ++k)
    for (int l = 0; l < W; ++l) {
      if (k >= pr[l].R) { distm[k][l] = one_m[k][l] = gapm[k][l] = pmm[k][l] = pmi[k][l] = pii[k][l] = pmd[k][l] = pdd[k][l] = 0; rs[k][l] = 0; continue; }
      const int q = pr[l].q[k] & 127, ii = pr[l].i[k] & 127, dd = pr[l].d[k] & 127, cc = pr[l].c[k] & 127;
      const T dist = ph[q];
      distm[k][l] = dist * third;
      one_m[k][l] = (T)1.0 - dist;
      gapm[k][l] = (T)1.0 - ph[cc];
      const int mx = std::max(ii, dd), mn = std::min(ii, dd);
      pmm[k][l] = mmt[((mx * (mx + 1)) >> 1) + mn];
      pmi[k][l] = ph[ii];
      pii[k][l] = ph[cc];
      pmd[k][l] = ph[dd];
      pdd[k][l] = ph[cc];
      rs[k][l] = (S)pr[l].rd[k];
    }
  F initd;
  I hlen, rlen;
  const T C = std::is_same<T, float>::value ? (T)ldexpf(1.f, 120) : (T)ldexp(1.0, 1020);
  for (int l = 0; l < W; ++l) {
    initd[l] = C / (T)pr[l].H;
    hlen[l] = (S)pr[l].H;
    rlen[l] = (S)pr[l].R;
  }
  for (uint32_t j = 0; j < max_h; ++j)
    for (int l = 0; l < W; ++l) hs[j][l] = j < pr[l].H ? (S)pr[l].hap[j] : (S)0;
  const F zero = F{} * 0;
  for (uint32_t j = 0; j <= max_h; ++j) { mrow[j] = zero; irow[j] = zero; drow[j] = initd; }
  for (uint32_t k = 0; k < max_r; ++k) {
    const I rb = rs[k];
    const F dm = distm[k], om = one_m[k], gm = gapm[k], tmm = pmm[k], tmi = pmi[k], tii = pii[k], tmd = pmd[k],
            tdd = pdd[k];
    const I kk = I{} + (S)k;
    const I rmask = rlen > kk;
    F Mj1 = zero, Ij1 = zero, Dj1 = zero;
    F Md = zero, Id = zero, Dd = (k == 0) ? initd : zero;
    F Mu = mrow[0], Iu = irow[0], Du = drow[0];
    for (uint32_t j = 0; j < max_h; ++j) {
      const I m = rb == hs[j];
      const F prior = m ? om : dm;
      F M = ((Md * tmm + Id * gm) + Dd * gm) * prior;
      F Ii = Mu * tmi + Iu * tii;
      F D = Mj1 * tmd + Dj1 * tdd;
      if (k >= min_r || j >= min_h) {
        const I jj = I{} + (S)j;
        const I live = rmask & (hlen > jj);
        M = live ? M : Mu;
        Ii = live ? Ii : Iu;
        D = live ? D : Du;
      }
      Md = Mu; Id = Iu; Dd = Du;
      Mj1 = M; Ij1 = Ii; Dj1 = D;
      mrow[j] = M; irow[j] = Ii; drow[j] = D;
      Mu = mrow[j + 1]; Iu = irow[j + 1]; Du = drow[j + 1];
    }
  }
  F sm = zero, si = zero;
  for (uint32_t j = 0; j < max_h; ++j) { sm += mrow[j]; si += irow[j]; }
  for (int l = 0; l < W; ++l) results[l] = sm[l] + si[l];
}

struct Region {  // same layout as rvc_region (include/rovaca_phmm.h)
  uint32_t n_haps, n_reads;
  const int32_t* hap_off;
  const uint8_t* hap_bases;
  const int32_t* read_off;
  const uint8_t* read_bases;
  const uint8_t *q, *i, *d, *c;
};

typedef double (*IntraFn)(const uint8_t*, int, const uint8_t*, const uint8_t*, const uint8_t*, const uint8_t*,
                          const uint8_t*, int, int);
IntraFn g_intra = nullptr;

struct Pair {
  uint32_t read_idx;
  uint16_t hap_idx, hap_len, read_len;
  uint8_t flags;  // 1 = N, 2 = used
};

std::vector<std::vector<size_t>> grouping(std::vector<Pair>& pairs, double max_idle, uint32_t width) {
  std::vector<std::vector<size_t>> groups;
  const size_t total = pairs.size();
  for (size_t i = 0; i < total; ++i) {
    if (pairs[i].flags & 3) continue;
    std::vector<size_t> g;
    size_t idx = i;
    while (idx < total && g.size() < width) {
      if (!(pairs[idx].flags & 3)) g.push_back(idx);
      ++idx;
    }
    if (g.size() < width) break;
    uint32_t mh = 0, mr = 0;
    uint64_t sum = 0;
    for (size_t x : g) {
      mh = std::max<uint32_t>(mh, pairs[x].hap_len);
      mr = std::max<uint32_t>(mr, pairs[x].read_len);
      sum += (uint64_t)pairs[x].hap_len * pairs[x].read_len;
    }
    const uint64_t cap = (uint64_t)mh * mr * width;
    const double ratio = cap == 0 ? 0.0 : (double)(cap > sum ? cap - sum : 0) / (double)cap;
    if (ratio <= max_idle) {
      for (size_t x : g) pairs[x].flags |= 2;
      groups.push_back(std::move(g));
    }
  }
  return groups;
}

// One region, three passes; out is read-major [n_reads][n_haps] (NaN for empty pairs).
void run_region(const Region& rg, double* out, std::vector<char>& scratch) {
  const size_t M = rg.n_haps, N = rg.n_reads;
  for (size_t x = 0; x < M * N; ++x) out[x] = std::numeric_limits<double>::quiet_NaN();
  std::vector<char> hn(M, 0), rn(N, 0);
  for (size_t h = 0; h < M; ++h)
    for (int32_t x = rg.hap_off[h]; x < rg.hap_off[h + 1]; ++x)
      if (rg.hap_bases[x] == 'N') { hn[h] = 1; break; }
  for (size_t r = 0; r < N; ++r)
    for (int32_t x = rg.read_off[r]; x < rg.read_off[r + 1]; ++x)
      if (rg.read_bases[x] == 'N') { rn[r] = 1; break; }
  std::vector<Pair> pairs;
  pairs.reserve(M * N);
  for (size_t h = 0; h < M; ++h)
    for (size_t r = 0; r < N; ++r) {
      const int hl = rg.hap_off[h + 1] - rg.hap_off[h], rl = rg.read_off[r + 1] - rg.read_off[r];
      if (hl > 0 && rl > 0)
        pairs.push_back(Pair{(uint32_t)r, (uint16_t)h, (uint16_t)hl, (uint16_t)rl, (uint8_t)((hn[h] || rn[r]) ? 1 : 0)});
    }
  std::sort(pairs.begin(), pairs.end(), [](const Pair& a, const Pair& b) {
    if (a.hap_len == b.hap_len) return a.read_len > b.read_len;
    return a.hap_len > b.hap_len;
  });
  auto ref = [&](const Pair& p) {
    PairRef x;
    x.hap = rg.hap_bases + rg.hap_off[p.hap_idx];
    x.H = p.hap_len;
    const int32_t o = rg.read_off[p.read_idx];
    x.rd = rg.read_bases + o; x.q = rg.q + o; x.i = rg.i + o; x.d = rg.d + o; x.c = rg.c + o;
    x.R = p.read_len;
    return x;
  };
  std::vector<char> pf(M * N, 0), pd(M * N, 0);
  {
    auto groups = grouping(pairs, 0.02, 16);
    PairRef prs[16];
    float res[16];
    for (auto& g : groups) {
      for (int l = 0; l < 16; ++l) prs[l] = ref(pairs[g[l]]);
      inter_group<float>(prs, res, scratch);
      for (int l = 0; l < 16; ++l) {
        Pair& p = pairs[g[l]];
        const size_t key = (size_t)p.hap_idx * N + p.read_idx;
        if ((double)res[l] < (double)1e-28f) {
          p.flags &= ~2;
        } else {
          out[(size_t)p.read_idx * M + p.hap_idx] = std::log10((double)res[l]) - (double)36.1236000061f;
          pd[key] = 1;
        }
        pf[key] = 1;
      }
    }
  }
  {
    for (auto& p : pairs)
      if (!pf[(size_t)p.hap_idx * N + p.read_idx]) p.flags |= 2;
    auto groups = grouping(pairs, 0.02, 8);
    PairRef prs[8];
    double res[8];
    for (auto& g : groups) {
      for (int l = 0; l < 8; ++l) prs[l] = ref(pairs[g[l]]);
      inter_group<double>(prs, res, scratch);
      for (int l = 0; l < 8; ++l) {
        const Pair& p = pairs[g[l]];
        out[(size_t)p.read_idx * M + p.hap_idx] = std::log10(res[l]) - 307.050595577260822;
        pd[(size_t)p.hap_idx * N + p.read_idx] = 1;
      }
    }
  }
  IntraFn fn = g_intra ? g_intra : &orc_intra_log10;
  for (size_t h = 0; h < M; ++h)
    for (size_t r = 0; r < N; ++r) {
      if (pd[h * N + r]) continue;
      const int hl = rg.hap_off[h + 1] - rg.hap_off[h];
      const int32_t o = rg.read_off[r];
      const int rl = rg.read_off[r + 1] - o;
      if (hl <= 0 || rl <= 0) continue;
      out[r * M + h] = fn(rg.hap_bases + rg.hap_off[h], hl, rg.read_bases + o, rg.q + o, rg.i + o, rg.d + o,
                          rg.c + o, rl, pf[h * N + r] ? 1 : 0);
    }
}

}  // namespace

extern "C" {

// Uses Rovaca's own intra kernel (oracle/_ref) for pass 3; NULL restores the
// scalar restatement. Returns 0 on success.
int orc_baseline_set_intra(const char* so_path) {
  if (!so_path) {
    g_intra = nullptr;
    return 0;
  }
  void* h = dlopen(so_path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return 1;
  g_intra = (IntraFn)dlsym(h, "ref_intra_log10");
  return g_intra ? 0 : 2;
}

int orc_baseline_isa(void) {
#if defined(__AVX512F__)
  return 512;
#elif defined(__AVX2__)
  return 256;
#else
  return 0;
#endif
}

// Runs regions[0..n) on n_threads pinned workers (one region per worker at a
// time) and returns the wall seconds. out (optional) receives each region's
// read-major matrix, concatenated.
double orc_baseline_run(const Region* regions, uint32_t n, int n_threads, double* out) {
  std::vector<size_t> off(n + 1, 0);
  for (uint32_t r = 0; r < n; ++r) off[r + 1] = off[r] + (size_t)regions[r].n_haps * regions[r].n_reads;
  std::atomic<uint32_t> next(0);
  const unsigned ncpu = std::max(1u, std::thread::hardware_concurrency());
  auto work = [&](int t) {
    cpu_set_t set;
    CPU_ZERO(&set);
    CPU_SET(t % ncpu, &set);
    if (!getenv("ORC_NO_PIN")) pthread_setaffinity_np(pthread_self(), sizeof set, &set);
    std::vector<char> scratch;
    std::vector<double> local;
    for (;;) {
      const uint32_t r = next.fetch_add(1);
      if (r >= n) break;
      double* dst = out ? out + off[r] : nullptr;
      if (!dst) {
        local.resize(off[r + 1] - off[r]);
        dst = local.data();
      }
      run_region(regions[r], dst, scratch);
    }
  };
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> th;
  for (int t = 0; t < n_threads; ++t) th.emplace_back(work, t);
  for (auto& x : th) x.join();
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
