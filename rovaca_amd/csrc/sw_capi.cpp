// Batched Smith-Waterman of reads against haplotypes: the C ABI of
// include/rovaca_sw.h (rvc_sw_align_batch; the served form is in
// sw_server.cpp). Pairs are grouped by sweep shape (read length); the pairs
// sw_narrow() admits are ordered by haplotype length, longest first, and
// swept two per lane by k_sw_fill16 (duos of neighbours in that order), the
// rest by k_sw_fill. Each shape and sweep present is one launch on one of a
// few streams forked from the call's stream, followed on that stream by
// k_sw_best and k_sw_trace for its pairs. Batches whose backtrack bytes
// exceed the engine's budget (12 GiB unless rvc_config says otherwise) run in
// chunks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/rovaca_sw.h"
#include "phmm_engine.h"
#include "sw_internal.h"

namespace rvc {
namespace {

constexpr uint32_t kPackCols = 64;           // CIGAR elements per pair copied back packed
constexpr uint64_t kThreadedUpload = 1u << 20;  // bases from which helper threads upload them
constexpr int kSubStreams = 3;                   // streams the launches of a pass are forked over

// scores of Rovaca's two aligner instances (sw_avx.c:14-15)
constexpr int32_t kGenotype[4] = {10, -15, -30, -5};
constexpr int32_t kAssemble[4] = {200, -150, -260, -11};

inline uint64_t up256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

// A grow-only pinned host buffer (copies from and to it need no staging).
struct PinnedArea {
  void* p = nullptr;
  size_t bytes = 0;
  PinnedArea() = default;
  PinnedArea(const PinnedArea&) = delete;
  PinnedArea& operator=(const PinnedArea&) = delete;
  ~PinnedArea() {
    if (p) hipHostFree(p);
  }
  template <class T>
  T* get(size_t n) {
    const size_t want = std::max<size_t>(n, 1) * sizeof(T);
    if (want > bytes) {
      if (p) hipHostFree(p);
      p = nullptr;
      bytes = 0;
      const size_t grow = want + want / 4;
      if (hipHostMalloc(&p, grow, hipHostMallocDefault) != hipSuccess) {
        p = nullptr;
        return nullptr;
      }
      bytes = grow;
    }
    return static_cast<T*>(p);
  }
};

// Per-thread workspace: the call's stream and its fork/join events, the
// upload and launch streams, a grow-only device buffer for the bases and one
// for the rest, pinned host buffers for the pair table, the work lists and
// the results, and the per-read host arrays, all kept across calls.
// Workspaces are owned by a registry and released by rvc_phmm_shutdown,
// never by thread exit.
struct SwWorkspace {
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {};  // a pass's fork and join
  // the launches of one pass (a sweep, then its pairs' best end and
  // traceback) forked over streams of their own, so that small passes run
  // their shapes at once
  hipStream_t sub[kSubStreams] = {};
  hipEvent_t sub_fill[kSubStreams] = {}, sub_done[kSubStreams] = {};
  void* mem = nullptr;
  size_t bytes = 0;
  // the caller's bases, uploaded from a helper thread on their own stream
  hipStream_t up_stream = nullptr, up_stream2 = nullptr;
  hipEvent_t up_done = nullptr, up_done2 = nullptr;
  void* in_mem = nullptr;
  size_t in_bytes = 0;
  PinnedArea pairs, list, res, cig;
  std::vector<int> shape;
  std::vector<uint8_t> narrow;
  std::vector<uint64_t> bt_est;
  std::vector<uint32_t> keys, ord;
  ~SwWorkspace() {
    if (mem) hipFree(mem);
    if (in_mem) hipFree(in_mem);
    for (hipEvent_t e : ev)
      if (e) hipEventDestroy(e);
    if (up_done) hipEventDestroy(up_done);
    if (up_done2) hipEventDestroy(up_done2);
    for (int k = 0; k < kSubStreams; ++k) {
      if (sub_fill[k]) hipEventDestroy(sub_fill[k]);
      if (sub_done[k]) hipEventDestroy(sub_done[k]);
      if (sub[k]) hipStreamDestroy(sub[k]);
    }
    if (stream) hipStreamDestroy(stream);
    if (up_stream) hipStreamDestroy(up_stream);
    if (up_stream2) hipStreamDestroy(up_stream2);
  }
  int streams() {
    if (stream) return RVC_OK;
    if (hipSetDevice(g_engine.device) != hipSuccess ||
        hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&up_stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&up_stream2, hipStreamNonBlocking) != hipSuccess)
      return set_error(RVC_E_DEVICE, "aligner streams");
    for (hipEvent_t& e : ev)
      if (hipEventCreate(&e) != hipSuccess) return set_error(RVC_E_DEVICE, "aligner events");
    if (hipEventCreateWithFlags(&up_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&up_done2, hipEventDisableTiming) != hipSuccess)
      return set_error(RVC_E_DEVICE, "aligner events");
    for (int k = 0; k < kSubStreams; ++k)
      if (hipStreamCreateWithFlags(&sub[k], hipStreamNonBlocking) != hipSuccess ||
          hipEventCreate(&sub_fill[k]) != hipSuccess ||
          hipEventCreateWithFlags(&sub_done[k], hipEventDisableTiming) != hipSuccess)
        return set_error(RVC_E_DEVICE, "aligner streams");
    return RVC_OK;
  }
  int quiesce() {
    hipStream_t all[3 + kSubStreams] = {stream, up_stream, up_stream2};
    for (int k = 0; k < kSubStreams; ++k) all[3 + k] = sub[k];
    for (hipStream_t x : all)
      if (x && hipStreamSynchronize(x) != hipSuccess) return set_error(RVC_E_DEVICE, "aligner stream");
    return RVC_OK;
  }
  int ensure_in(size_t want) {
    if (want <= in_bytes) return RVC_OK;
    if (in_mem) hipFree(in_mem);
    in_mem = nullptr;
    in_bytes = 0;
    const size_t n = std::max(want + want / 4, (size_t)16 << 20);
    if (hipMalloc(&in_mem, n) != hipSuccess) return set_error(RVC_E_DEVICE, "aligner input buffer of %zu bytes", n);
    in_bytes = n;
    return RVC_OK;
  }
  int ensure(size_t want) {
    if (want <= bytes) return RVC_OK;
    if (mem) hipFree(mem);
    mem = nullptr;
    bytes = 0;
    const size_t n = std::max(want + want / 4, (size_t)64 << 20);
    if (hipMalloc(&mem, n) != hipSuccess) return set_error(RVC_E_DEVICE, "aligner workspace of %zu bytes", n);
    bytes = n;
    return RVC_OK;
  }
};
std::mutex g_sw_mu;
std::vector<SwWorkspace*> g_sw_all;
std::atomic<uint64_t> g_sw_generation{1};
thread_local SwWorkspace* t_sw = nullptr;
thread_local uint64_t t_sw_generation = 0;

SwWorkspace* sw_workspace() {
  if (!t_sw || t_sw_generation != g_sw_generation.load(std::memory_order_acquire)) {
    auto* w = new (std::nothrow) SwWorkspace;
    if (!w) return nullptr;
    std::lock_guard<std::mutex> lk(g_sw_mu);
    g_sw_all.push_back(w);
    t_sw = w;
    t_sw_generation = g_sw_generation.load(std::memory_order_relaxed);
  }
  return t_sw;
}

}  // namespace

void sw_release_workspaces() {
  std::lock_guard<std::mutex> lk(g_sw_mu);
  for (SwWorkspace* w : g_sw_all) delete w;
  g_sw_all.clear();
  g_sw_generation.fetch_add(1, std::memory_order_release);
}

int sw_validate(const uint8_t* hap_bases, const int32_t* hap_off, uint32_t n_haps, const uint8_t* read_bases,
                const int32_t* read_off, uint32_t n_reads, const uint32_t* read_hap, uint32_t max_cigar,
                const uint32_t* cigar, const int32_t* count, const int32_t* offset, const int32_t* score,
                int32_t* longest) {
  if (!g_engine.ready) return set_error(RVC_E_STATE, "rvc_phmm_init has not run");
  if (longest) *longest = 0;
  if (n_reads == 0) return RVC_OK;
  if (!hap_bases || !hap_off || !read_bases || !read_off || !read_hap || !cigar || !count || !offset || !score ||
      n_haps == 0 || max_cigar == 0)
    return set_error(RVC_E_INVALID, "null pointer or empty haplotype list");
  for (uint32_t h = 0; h < n_haps; ++h)
    if (hap_off[h] < 0 || hap_off[h + 1] < hap_off[h]) return set_error(RVC_E_INVALID, "haplotype offsets");
  int32_t most = 0;
  for (uint32_t r = 0; r < n_reads; ++r) {
    if (read_off[r] < 0 || read_off[r + 1] < read_off[r]) return set_error(RVC_E_INVALID, "read offsets");
    if (read_hap[r] >= n_haps) return set_error(RVC_E_INVALID, "read %u names haplotype %u", r, read_hap[r]);
    const int32_t n2 = read_off[r + 1] - read_off[r];
    if (n2 > kSwMaxRead && n2 < kSwMaxLen)  // (none: the largest shape holds every read the reference aligns)
      return set_error(RVC_E_LIMIT, "read %u has %d bases (the device aligner takes up to %d)", r, n2, kSwMaxRead);
    most = std::max(most, n2);
  }
  if (longest) *longest = most;
  return RVC_OK;
}

int sw_align_batch(const uint8_t* hap_bases, const int32_t* hap_off, uint32_t n_haps, const uint8_t* read_bases,
                   const int32_t* read_off, uint32_t n_reads, const uint32_t* read_hap, int32_t assemble,
                   uint32_t max_cigar, uint32_t* cigar, int32_t* count, int32_t* offset, int32_t* score,
                   rvc_sw_stats* stats) {
  const auto t_start = std::chrono::steady_clock::now();
  if (int rc = sw_validate(hap_bases, hap_off, n_haps, read_bases, read_off, n_reads, read_hap, max_cigar, cigar,
                           count, offset, score, nullptr))
    return rc;
  if (n_reads == 0) return RVC_OK;
  const int32_t* sc = assemble ? kAssemble : kGenotype;
  if (stats) std::memset(stats, 0, sizeof *stats);

  SwWorkspace* wp = sw_workspace();
  if (!wp) return set_error(RVC_E_INVALID, "out of host memory");
  SwWorkspace& w = *wp;
  // a call that failed part way may have left copies or kernels in flight
  // on any of the workspace's streams
  if (int rc = w.quiesce()) return rc;
  if (int rc = w.streams()) return rc;
  // the caller's bases (pageable memory, staged by the runtime) go up from a
  // helper thread while this one builds the pair table and the work lists
  const uint64_t hap_bytes = (uint64_t)hap_off[n_haps], read_bytes = (uint64_t)read_off[n_reads];
  const size_t o_read_in = (size_t)up256(hap_bytes + 64);
  if (int rc = w.ensure_in(o_read_in + up256(read_bytes + 64))) return rc;
  char* in = static_cast<char*>(w.in_mem);
  // (small batches, such as one region's reads, go up from this thread: a
  // thread costs more than their copy)
  hipError_t up_err = hipSuccess, up_err2 = hipSuccess;
  const uint64_t half = read_bytes / 2;
  auto upload_first = [&] {
    hipError_t e = hipSetDevice(g_engine.device);
    if (e == hipSuccess) e = hipMemcpyAsync(in, hap_bases, hap_bytes, hipMemcpyHostToDevice, w.up_stream);
    if (e == hipSuccess) e = hipMemcpyAsync(in + o_read_in, read_bases, half, hipMemcpyHostToDevice, w.up_stream);
    if (e == hipSuccess) e = hipEventRecord(w.up_done, w.up_stream);
    up_err = e;
  };
  auto upload_second = [&] {
    hipError_t e = hipSetDevice(g_engine.device);
    if (e == hipSuccess)
      e = hipMemcpyAsync(in + o_read_in + half, read_bases + half, read_bytes - half, hipMemcpyHostToDevice,
                         w.up_stream2);
    if (e == hipSuccess) e = hipEventRecord(w.up_done2, w.up_stream2);
    up_err2 = e;
  };
  const bool threaded = hap_bytes + read_bytes >= kThreadedUpload;
  std::thread uploader, uploader2;
  if (threaded) {
    // (where a thread cannot be had, this thread uploads)
    try {
      uploader = std::thread(upload_first);
    } catch (...) {
      upload_first();
    }
    try {
      uploader2 = std::thread(upload_second);
    } catch (...) {
      upload_second();
    }
  } else {
    upload_first();
    upload_second();
  }
  bool uploads_pending = true;
  struct Join {
    std::thread& t;
    ~Join() {
      if (t.joinable()) t.join();
    }
  } join_uploader{uploader}, join_uploader2{uploader2};
  // one SwPair per read; refused pairs stay out of the sweeps
  SwPair* pairs = w.pairs.get<SwPair>(n_reads);
  uint32_t* list = w.list.get<uint32_t>(2ull * n_reads + 2);
  int32_t* res = w.res.get<int32_t>(3ull * n_reads);
  if (!pairs || !list || !res) return set_error(RVC_E_DEVICE, "aligner pinned host buffers for %u reads", n_reads);
  std::vector<int>& shape = w.shape;
  std::vector<uint8_t>& narrow = w.narrow;
  std::vector<uint64_t>& bt_est = w.bt_est;  // bounds the backtrack bytes a pair adds, alone or in a duo
  shape.assign(n_reads, -1);
  narrow.assign(n_reads, 0);
  bt_est.assign(n_reads, 0);
  for (uint32_t r = 0; r < n_reads; ++r) {
    SwPair& p = pairs[r];
    std::memset(&p, 0, sizeof p);
    const uint32_t h = read_hap[r];
    p.hap_off = (uint32_t)hap_off[h];
    p.hap_len = (uint32_t)(hap_off[h + 1] - hap_off[h]);
    p.read_off = (uint32_t)read_off[r];
    p.read_len = (uint32_t)(read_off[r + 1] - read_off[r]);
    if (p.hap_len == 0 || p.read_len == 0 || p.hap_len >= (uint32_t)kSwMaxLen || p.read_len >= (uint32_t)kSwMaxLen)
      continue;
    int W = 0, RPL = 0;
    shape[r] = sw_shape(p.read_len, &W, &RPL);
    narrow[r] = sw_packed_shape(shape[r]) && sw_narrow(p.hap_len, p.read_len, sc[0], sc[1], sc[2], sc[3]) ? 1 : 0;
    if (stats) stats->cells += (uint64_t)p.hap_len * p.read_len;
  }
  // a shape with too few pairs to fill the device once (or fewer than the
  // engine's sw_min_launch) joins the next larger shape launched anyway (of
  // the same sweep, at most 1.5 times its rows): a launch of its own would
  // cost about one work item's time for a handful of pairs
  int promote[2][kSwBaseShapes];
  // then a pass whose pairs take at most three waves per SIMD on the wide
  // shapes runs on them (a wave holds one pair, or one duo of the packed
  // sweep): such a pass waits for its longest chain of steps, which 64 lanes
  // per pair shorten, and a SIMD that interleaves a few waves hides the
  // latency of each one's dependent instructions. Passes of 1 / 16 / 32 / 48
  // regions of ~126 reads sweep in 0.14 / 0.36 / 0.44 / 0.46 ms on the wide
  // shapes against 0.21 / 0.60 / 0.61 / 0.61 on the base ones; at 64 regions
  // (four waves per SIMD) the two are equal
  bool wide = g_engine.sw_wide == 2;
  bool small = false;  // the pass size that picks the wide shapes by default
  {
    uint64_t cnt[2][kSwBaseShapes] = {};
    for (uint32_t r = 0; r < n_reads; ++r)
      if (shape[r] >= 0) ++cnt[narrow[r]][shape[r]];
    for (int nw = 0; nw < 2; ++nw) {
      int target = -1;  // the smallest larger shape that is launched
      for (int s = kSwBaseShapes - 1; s >= 0; --s) {
        promote[nw][s] = s;
        if (cnt[nw][s] == 0) continue;
        const uint64_t fill =
            g_engine.sw_min_launch ? (uint64_t)g_engine.sw_min_launch
                                   : (uint64_t)g_engine.n_cu * (uint64_t)(nw ? sw16_blocks_per_cu(s) : sw_blocks_per_cu(s)) *
                                         (uint64_t)(64 / kSwW[s]) * (nw ? 2u : 1u);
        if (target >= 0 && cnt[nw][s] < fill && 2 * kSwW[target] * kSwR[target] <= 3 * kSwW[s] * kSwR[s])
          promote[nw][s] = target;
        else
          target = s;
      }
    }
    uint64_t waves = 0;
    for (int s = 0; s < kSwBaseShapes; ++s) waves += cnt[0][s] + (cnt[1][s] + 1) / 2;
    small = waves > 0 && waves <= 12ull * (uint64_t)g_engine.n_cu;
    if (g_engine.sw_wide == 0) wide = small;
  }
  for (uint32_t r = 0; r < n_reads; ++r) {
    if (shape[r] < 0) continue;
    SwPair& p = pairs[r];
    const int sb = promote[narrow[r]][shape[r]];
    const int s = wide ? kSwWide[sb] : sb;
    shape[r] = s;
    p.width = (uint16_t)kSwW[s];
    p.rpl = (uint16_t)kSwR[s];
    // a 16-bit word per 4 rows and step; a duo's steps run to its longer
    // haplotype: at most twice this pair's own share
    bt_est[r] = up256((uint64_t)(p.hap_len + kSwW[s] - 1) * (uint64_t)(kSwW[s] * sw_groups(kSwR[s])) * 2);
    if (narrow[r]) bt_est[r] *= 2;
  }

  uint32_t r0 = 0;
  while (r0 < n_reads) {
    // the chunk: reads r0 .. r1 whose backtrack bytes fit the budget
    uint32_t r1 = r0;
    uint64_t est = 0, hrow_total = 0, hcol_total = 0;
    while (r1 < n_reads && (r1 == r0 || est + bt_est[r1] <= g_engine.sw_bt_budget)) {
      SwPair& p = pairs[r1];
      p.hrow_off = (uint32_t)hrow_total;
      p.hcol_off = (uint32_t)hcol_total;
      est += bt_est[r1];
      if (shape[r1] >= 0) {
        hrow_total += p.hap_len;
        hcol_total += (uint64_t)p.width * p.rpl;
      }
      ++r1;
    }
    const uint32_t n = r1 - r0;
    // order (a counting sort): packed-sweep pairs, then int32-sweep pairs,
    // each by shape and haplotype length, longest first (the persistent
    // waves end on short items)
    constexpr uint32_t kKeys = 2 * kSwShapes * kSwMaxLen;
    auto key = [&](uint32_t x) -> uint32_t {
      const uint32_t r = r0 + x;
      if (shape[r] < 0) return kKeys;
      return ((narrow[r] ? 0u : (uint32_t)kSwShapes) + (uint32_t)shape[r]) * kSwMaxLen +
             (kSwMaxLen - 1 - pairs[r].hap_len);
    };
    std::vector<uint32_t>& cnt = w.keys;
    cnt.assign(kKeys + 1, 0);
    for (uint32_t x = 0; x < n; ++x) {
      const uint32_t k = key(x);
      if (k < kKeys) ++cnt[k + 1];
    }
    for (uint32_t k = 0; k < kKeys; ++k) cnt[k + 1] += cnt[k];
    std::vector<uint32_t>& ord = w.ord;
    ord.resize(cnt[kKeys]);
    for (uint32_t x = 0; x < n; ++x) {
      const uint32_t k = key(x);
      if (k < kKeys) ord[cnt[k]++] = x;
    }
    // work lists and backtrack areas: duos of neighbours in that order (two
    // entries each; a lone last pair gets kSwNoPair), then single pairs
    uint32_t nlist = 0;
    uint32_t dstart[kSwShapes] = {}, dcount[kSwShapes] = {}, sstart[kSwShapes] = {}, scount[kSwShapes] = {};
    uint64_t bt_total = 0;
    size_t i = 0;
    const size_t m = ord.size();
    for (int s = 0; s < kSwShapes; ++s) {
      const uint64_t groups = (uint64_t)kSwW[s] * (uint64_t)sw_groups(kSwR[s]);
      dstart[s] = nlist;
      while (i < m && narrow[r0 + ord[i]] && shape[r0 + ord[i]] == s) {
        const uint32_t xa = ord[i++];
        const uint32_t xb = (i < m && narrow[r0 + ord[i]] && shape[r0 + ord[i]] == s) ? ord[i++] : kSwNoPair;
        SwPair& pa = pairs[r0 + xa];
        uint32_t n1 = pa.hap_len;
        pa.bt_off = bt_total;
        pa.bt_lanes = 2;
        pa.bt_half = 0;
        if (xb != kSwNoPair) {
          SwPair& pb = pairs[r0 + xb];
          n1 = std::max(n1, pb.hap_len);
          pb.bt_off = bt_total;
          pb.bt_lanes = 2;
          pb.bt_half = 1;
        }
        bt_total += up256((uint64_t)(n1 + kSwW[s] - 1) * groups * 4);  // a 32-bit word per 4 rows and step
        list[nlist++] = xa;
        list[nlist++] = xb;
        ++dcount[s];
      }
    }
    for (int s = 0; s < kSwShapes; ++s) {
      const uint64_t groups = (uint64_t)kSwW[s] * (uint64_t)sw_groups(kSwR[s]);
      sstart[s] = nlist;
      while (i < m && !narrow[r0 + ord[i]] && shape[r0 + ord[i]] == s) {
        const uint32_t x = ord[i++];
        SwPair& p = pairs[r0 + x];
        p.bt_off = bt_total;
        p.bt_lanes = 1;
        p.bt_half = 0;
        bt_total += up256((uint64_t)(p.hap_len + kSwW[s] - 1) * groups * 2);
        list[nlist++] = x;
        ++scount[s];
      }
    }
    // device layout
    size_t off = 0;
    auto take = [&](uint64_t b) { const size_t o = off; off += up256(b); return o; };
    const size_t o_pair = take(sizeof(SwPair) * n),
                 o_list = take(4ull * (nlist + 1)), o_head = take(4 * 2 * kSwShapes), o_bt = take(bt_total),
                 o_hrow = take(4 * (hrow_total + 1)), o_hcol = take(4 * (hcol_total + 1)),
                 o_cig = take(4ull * n * max_cigar), o_cnt = take(4ull * n), o_off = take(4ull * n),
                 o_sc = take(4ull * n), o_best = take(8ull * n), o_pack = take(4ull * n * kPackCols);
    int rc = w.ensure(off);
    if (rc) return rc;
    char* base = static_cast<char*>(w.mem);
    hipStream_t st = w.stream;
    hipError_t e = hipMemcpyAsync(base + o_pair, pairs + r0, sizeof(SwPair) * n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && nlist > 0)
      e = hipMemcpyAsync(base + o_list, list, 4ull * nlist, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(base + o_head, 0, 4 * 2 * kSwShapes, st);
    if (e != hipSuccess) return set_error(RVC_E_DEVICE, "aligner upload: %s", hipGetErrorString(e));
    if (uploads_pending) {
      uploads_pending = false;
      if (uploader.joinable()) uploader.join();
      if (uploader2.joinable()) uploader2.join();
      if (up_err == hipSuccess) up_err = up_err2;
      if (up_err == hipSuccess) up_err = hipStreamWaitEvent(st, w.up_done, 0);
      if (up_err == hipSuccess) up_err = hipStreamWaitEvent(st, w.up_done2, 0);
      if (up_err != hipSuccess) return set_error(RVC_E_DEVICE, "aligner upload of the bases: %s", hipGetErrorString(up_err));
    }
    SwArgs a{};
    a.haps = reinterpret_cast<const uint8_t*>(in);
    a.reads = reinterpret_cast<const uint8_t*>(in + o_read_in);
    a.pairs = reinterpret_cast<const SwPair*>(base + o_pair);
    a.n_pairs = n;
    a.bt = reinterpret_cast<uint8_t*>(base + o_bt);
    a.hrow = reinterpret_cast<int32_t*>(base + o_hrow);
    a.hcol = reinterpret_cast<int32_t*>(base + o_hcol);
    a.match = sc[0];
    a.mismatch = sc[1];
    a.open = sc[2];
    a.extend = sc[3];
    a.cigar = reinterpret_cast<uint32_t*>(base + o_cig);
    a.max_cigar = max_cigar;
    a.count = reinterpret_cast<int32_t*>(base + o_cnt);
    a.offset = reinterpret_cast<int32_t*>(base + o_off);
    a.score = reinterpret_cast<int32_t*>(base + o_sc);
    a.best_ij = reinterpret_cast<int32_t*>(base + o_best);
    // a small pass's walks are latency-bound: 8 diagonal cells per load (the
    // traceback of 16-64 regions of ~126 reads: 0.13 against 0.20 ms); a
    // large pass's are not (1024 regions: 0.57-0.65 against 0.29-0.44 ms)
    a.trace_run = small ? 8 : 1;
    // each launch (a shape's sweep, then its pairs' best end and traceback)
    // on the next of the sub streams, forked from and joined back into st
    if (hipEventRecord(w.ev[0], st) != hipSuccess) return set_error(RVC_E_DEVICE, "aligner event");
    bool used[kSubStreams] = {};
    int next_sub = 0;
    for (int packed = 1; packed >= 0; --packed) {
      for (int s = 0; s < kSwShapes; ++s) {
        const uint32_t nl = packed ? dcount[s] : scount[s];
        if (nl == 0) continue;
        const int k = next_sub;
        next_sub = (next_sub + 1) % kSubStreams;
        hipStream_t ss = w.sub[k];
        if (!used[k] && (e = hipStreamWaitEvent(ss, w.ev[0], 0)) != hipSuccess)
          return set_error(RVC_E_DEVICE, "aligner fork: %s", hipGetErrorString(e));
        used[k] = true;
        a.list = reinterpret_cast<const uint32_t*>(base + o_list) + (packed ? dstart[s] : sstart[s]);
        a.n_list = nl;
        a.head = reinterpret_cast<uint32_t*>(base + o_head) + (packed ? s : kSwShapes + s);
        const uint32_t nseg = 64u / (uint32_t)kSwW[s];
        const uint32_t waves = (nl + nseg - 1) / nseg;
        const int per_cu = packed ? sw16_blocks_per_cu(s) : sw_blocks_per_cu(s);
        const int grid = (int)std::max<uint32_t>(1, std::min<uint32_t>(waves, (uint32_t)per_cu * (uint32_t)g_engine.n_cu));
        if ((e = packed ? launch_sw_fill16(s, grid, ss, a) : launch_sw_fill(s, grid, ss, a)) != hipSuccess)
          return set_error(RVC_E_DEVICE, "aligner sweep: %s", hipGetErrorString(e));
        if (hipEventRecord(w.sub_fill[k], ss) != hipSuccess) return set_error(RVC_E_DEVICE, "aligner event");
        a.n_list = packed ? 2 * nl : nl;  // a duo's two entries
        if ((e = launch_sw_trace(ss, a)) != hipSuccess)
          return set_error(RVC_E_DEVICE, "aligner traceback: %s", hipGetErrorString(e));
        if (stats) ++stats->launches;
      }
    }
    for (int k = 0; k < kSubStreams; ++k) {
      if (!used[k]) continue;
      if (hipEventRecord(w.sub_done[k], w.sub[k]) != hipSuccess ||
          hipStreamWaitEvent(st, w.sub_done[k], 0) != hipSuccess)
        return set_error(RVC_E_DEVICE, "aligner join");
    }
    a.list = nullptr;
    a.n_list = 0;
    if (hipEventRecord(w.ev[1], st) != hipSuccess) return set_error(RVC_E_DEVICE, "aligner event");
    // counts first, then of every pair's CIGAR slot only the columns the
    // longest CIGAR of the chunk uses
    // (count, offset and score sit one after another on the device)
    e = hipMemcpyAsync(res, base + o_cnt, 4ull * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(res + n, base + o_off, 4ull * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(res + 2ull * n, base + o_sc, 4ull * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) {
      std::memcpy(count + r0, res, 4ull * n);
      std::memcpy(offset + r0, res + n, 4ull * n);
      std::memcpy(score + r0, res + 2ull * n, 4ull * n);
      // pairs without a sweep: the library refuses a length >= 2048 (count
      // -1, offset 1, sw_core_template.h:356-361); an empty haplotype leaves
      // every read base soft-clipped at offset 0 with score 0; an empty read
      // against n1 bases ends at offset n1 with no CIGAR element and score 0
      // (both empty: score INT32_MIN, the best-end search finds no cell). The
      // library's own instance keeps the previous call's CIGAR count for an
      // empty read (its count is not reset); these are a fresh instance's.
      for (uint32_t x = 0; x < n; ++x) {
        if (shape[r0 + x] >= 0) continue;
        const SwPair& p = pairs[r0 + x];
        if (p.hap_len >= (uint32_t)kSwMaxLen || p.read_len >= (uint32_t)kSwMaxLen) {
          count[r0 + x] = -1;
          offset[r0 + x] = 1;
          score[r0 + x] = 0;
        } else if (p.hap_len == 0 && p.read_len > 0) {
          count[r0 + x] = 1;  // max_cigar >= 1 (validated)
          cigar[(size_t)(r0 + x) * max_cigar] = (p.read_len << 4) | 4u;  // S
          offset[r0 + x] = 0;
          score[r0 + x] = 0;
        } else {
          count[r0 + x] = 0;
          offset[r0 + x] = (int32_t)p.hap_len;
          score[r0 + x] = p.hap_len == 0 ? INT32_MIN : 0;
        }
      }
    }
    int32_t widest = 0;
    for (uint32_t x = 0; x < n && e == hipSuccess; ++x) widest = std::max(widest, count[r0 + x]);
    widest = std::min<int32_t>(widest, (int32_t)max_cigar);
    if (e == hipSuccess && widest > 0) {
      // the used columns packed on the device (pitch 'widest') when they fit
      // the pack area, else the whole slots; one copy, then spread into the
      // caller's slots
      const bool packed = widest <= kPackCols;
      const uint32_t pitch = packed ? (uint32_t)widest : max_cigar;
      if (packed)
        e = hipMemcpy2DAsync(base + o_pack, 4ull * pitch, base + o_cig, 4ull * max_cigar, 4ull * pitch, n,
                             hipMemcpyDeviceToDevice, st);
      uint32_t* host = w.cig.get<uint32_t>((size_t)n * pitch);
      if (!host) return set_error(RVC_E_DEVICE, "aligner pinned host buffer of %zu CIGAR elements", (size_t)n * pitch);
      if (e == hipSuccess)
        e = hipMemcpyAsync(host, base + (packed ? o_pack : o_cig), 4ull * n * pitch, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      for (uint32_t x = 0; x < n && e == hipSuccess; ++x)
        if (count[r0 + x] > 0 && shape[r0 + x] >= 0)
          std::memcpy(cigar + (size_t)(r0 + x) * max_cigar, host + (size_t)x * pitch, 4ull * (uint32_t)count[r0 + x]);
    }
    if (e != hipSuccess) return set_error(RVC_E_DEVICE, "aligner results: %s", hipGetErrorString(e));
    if (stats) {
      // sweeps: fork to the last sweep's end; best end and traceback: the rest
      float f = 0, all = 0;
      for (int k = 0; k < kSubStreams; ++k) {
        float x = 0;
        if (used[k] && hipEventElapsedTime(&x, w.ev[0], w.sub_fill[k]) == hipSuccess) f = std::max(f, x);
      }
      hipEventElapsedTime(&all, w.ev[0], w.ev[1]);
      stats->fill_ms += f;
      stats->trace_ms += std::max(0.f, all - f);
      ++stats->chunks;
    }
    for (uint32_t x = 0; x < n; ++x)
      if (count[r0 + x] == -2) return set_error(RVC_E_LIMIT, "read %u needs more than %u CIGAR elements", r0 + x, max_cigar);
    r0 = r1;
  }
  if (stats)
    stats->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  return RVC_OK;
}

}  // namespace rvc

extern "C" int rvc_sw_align_batch(const uint8_t* hap_bases, const int32_t* hap_off, uint32_t n_haps,
                                  const uint8_t* read_bases, const int32_t* read_off, uint32_t n_reads,
                                  const uint32_t* read_hap, int32_t assemble, uint32_t max_cigar, uint32_t* cigar,
                                  int32_t* count, int32_t* offset, int32_t* score, rvc_sw_stats* stats) {
  return rvc::guarded([&]() -> int {
    return rvc::sw_align_batch(hap_bases, hap_off, n_haps, read_bases, read_off, n_reads, read_hap, assemble, max_cigar,
                               cigar, count, offset, score, stats);
  });
}
