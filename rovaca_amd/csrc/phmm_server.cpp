// Serving loop: continuous batching of active regions submitted by many host
// threads (Rovaca's worker pool calls call_pairhmm once per region from up to
// --nthreads workers, hc/haplotypecaller_engine.cpp:119,254).
//
// submit() packs a region on the caller's thread and returns a ticket at once.
// kSlots launch slots, each with its own thread, stream, pinned staging
// buffer and device arena, take work from one queue: a free slot admits every
// waiting request, in arrival order, whose device footprint still fits the
// HBM budget next to the requests already in flight, merges them straight into
// its pinned staging buffer in device layout, and queues the upload, the FP32
// pass and one copy of results, flags and rescue counts (by a kernel, into a
// pinned result buffer) on its stream; once those counts are back it queues
// only the FP64 lists that hold pairs and the marginalisation of requests
// with sites. Slots work concurrently, so requests that arrive while earlier
// launches are on the device are staged and started while those finish.
// Results land in the slot's pinned result buffer; when the launch finishes,
// the slot thread copies each request's region out (in input read order) and
// the buffer goes back to the pool, so a waiter that comes late holds none.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/rovaca_phmm.h"
#include "phmm_engine.h"
#include "phmm_server.h"

namespace rvc {
namespace {

using Clock = std::chrono::steady_clock;
// Six slots: a slot's host work (staging a launch, handing its results to
// the requests) runs beside the other slots' device work, and several small
// launches on the device overlap each other's ramp and drain. Measured on
// MI355X (served stream of the batched pool, 32 producers): p50/p99 at 24k
// regions/s 1.9/4.8 ms with 3 slots, 1.6/4.5 with 4, 0.8/3.0-3.4 with 6,
// 1.0/3.2-3.5 with 8; 32k/s delivered 7.3-7.4 k GCUPS with 3, 7.6 k with 6.
// Each slot holds its device arena and pinned buffers (about 0.7 GB).
constexpr int kSlots = 6;
// Bounds on one launch: device bytes of its requests and pair slots. A slot
// stops taking queued requests at either (a request larger than both still
// goes alone). The slot buffers are sized for a launch at the bounds when the
// server starts, so that launches under load reallocate nothing; a launch of
// 8M pairs runs for tens of milliseconds, long enough to keep the device busy.
constexpr uint64_t kLaunchBytes = 256ull << 20;
constexpr uint64_t kLaunchPairs = 8ull << 20;

// Pinned host memory for one launch's results in their device layout: log10
// L per pair slot, the flags (rounded up to 8 bytes) and the counters, whose
// last kNumFp64Lists are the FP64 rescue counts.
struct ResultBuf {
  char* mem = nullptr;
  size_t cap = 0;  // pair slots
  size_t n = 0;    // pair slots of the launch holding it
  double* out() const { return reinterpret_cast<double*>(mem); }
  uint8_t* flags() const { return reinterpret_cast<uint8_t*>(mem + n * sizeof(double)); }
  const uint32_t* rescued() const {
    return reinterpret_cast<const uint32_t*>(mem + n * sizeof(double) + ((n + 7) & ~size_t(7))) + kFp32Bins +
           kNumFp64Lists;
  }
  ~ResultBuf() {
    if (mem) hipHostFree(mem);
  }
  // Grows by doubling; the old block goes to 'retired' (freed with the pool:
  // freeing pinned memory waits for the whole device).
  int reserve(size_t n, bool* grew, std::vector<char*>* retired) {
    if (n <= cap && mem) return RVC_OK;
    *grew = true;
    const size_t c = std::max(n + n / 4 + 4096, 2 * cap);
    if (mem) retired->push_back(mem);
    mem = nullptr;
    cap = 0;
    if (hipHostMalloc((void**)&mem, c * (sizeof(double) + 1) + 8 + 4 * kNumCounters + 64, hipHostMallocDefault) !=
        hipSuccess) {
      mem = nullptr;
      return set_error(RVC_E_DEVICE, "pinned result buffer for %zu pairs", n);
    }
    cap = c;
    return RVC_OK;
  }
};

// Result buffers are recycled: a slot takes one for a launch and gives it
// back once its requests have copied their regions out.
struct ResultPool : std::enable_shared_from_this<ResultPool> {
  std::mutex mu;
  std::vector<std::unique_ptr<ResultBuf>> free;
  std::vector<char*> retired;
  ~ResultPool() {
    for (char* p : retired) hipHostFree(p);
  }
  // A buffer of n pair slots in the free list from the start.
  void add(size_t n) {
    auto b = std::unique_ptr<ResultBuf>(new (std::nothrow) ResultBuf);
    bool grew = false;
    if (b && b->reserve(n, &grew, &retired) == RVC_OK) free.push_back(std::move(b));
  }
  std::shared_ptr<ResultBuf> acquire(size_t n, int* rc, bool* grew) {
    std::unique_ptr<ResultBuf> b;
    {
      std::lock_guard<std::mutex> lk(mu);
      size_t best = free.size();
      for (size_t x = 0; x < free.size(); ++x)
        if (free[x]->cap >= n && (best == free.size() || free[x]->cap < free[best]->cap)) best = x;
      if (best == free.size() && !free.empty()) best = free.size() - 1;
      if (best < free.size()) {
        b = std::move(free[best]);
        free.erase(free.begin() + (long)best);
      }
    }
    if (!b) b.reset(new (std::nothrow) ResultBuf);
    if (!b) {
      *rc = set_error(RVC_E_INVALID, "out of host memory");
      return nullptr;
    }
    {
      std::vector<char*> old;
      *rc = b->reserve(n, grew, &old);
      if (!old.empty()) {
        std::lock_guard<std::mutex> lk(mu);
        retired.insert(retired.end(), old.begin(), old.end());
      }
      if (*rc) return nullptr;
    }
    b->n = n;
    std::shared_ptr<ResultPool> self = shared_from_this();
    return std::shared_ptr<ResultBuf>(b.release(), [self](ResultBuf* p) {
      std::lock_guard<std::mutex> lk(self->mu);
      self->free.emplace_back(p);
    });
  }
};

struct Request {
  uint64_t id = 0;
  HostPack pack;  // the region, packed by the submitting thread
  uint64_t resident = 0;
  uint32_t max_hap = 0;
  // the region's read-major results and flags in input read order, copied
  // out of the launch's pinned buffer by the slot thread when it finishes
  std::vector<double> out;
  std::vector<uint8_t> flags;
  // sites to marginalise on the device (optional): alleles per site, the
  // allele of each haplotype per site (-1 = none), and the marginals
  // ([site][allele][read], input read order) copied out like the results
  std::vector<uint32_t> site_alleles;
  std::vector<int16_t> site_map;
  std::vector<double> marg;
  int status = 0;
  bool done = false;
  uint64_t t_submit = 0, t_launch = 0, t_done = 0;
};

uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now().time_since_epoch()).count();
}

struct Slot {
  std::thread thr;
  hipStream_t stream = nullptr;
  hipEvent_t done_ev = nullptr;
  void* mem = nullptr;  // device arena (grow-only)
  size_t bytes = 0;
  std::vector<void*> retired;  // outgrown arenas, freed at stop
  PinnedBuf staging;
  HostPack meta;
  DeviceBatch batch;
  std::shared_ptr<ResultBuf> res;
  std::vector<Request*> reqs;
  uint64_t resident = 0;
  // marginalisation of the requests that carry sites: pinned inputs
  // (regions, read order, site tables) and outputs, device offsets
  PinnedBuf marg_in, marg_out;
  size_t marg_out_bytes = 0;
  std::vector<uint64_t> marg_base;  // per request: first output value
  MargArgs marg{};                  // its launch, queued after the FP64 pass
  size_t marg_in_bytes = 0;
  char* marg_dev = nullptr;
  bool marg_pending = false;
};

size_t up256(size_t x) { return (x + 255) & ~size_t(255); }

// Device bytes a packed region holds while in flight: its share of the
// layout DeviceBatch::upload_staged builds (sequences, tiles, results, rescue
// list, counters), with its tiles counted as if split into single haplotypes.
uint64_t footprint(const HostPack& p) {
  PackSizes z = pack_sizes(p);
  z.n_tiles = 0;
  for (int v = 0; v < kFp32Bins; ++v) z.n_tiles += p.bin_units[v];
  return batch_layout(z).bytes;
}

struct Server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending;
  std::unordered_map<uint64_t, Request*> tickets;
  uint64_t next_id = 1;
  uint64_t resident = 0, peak_resident = 0;
  bool stopping = false;
  Slot slots[kSlots];
  std::shared_ptr<ResultPool> pool = std::make_shared<ResultPool>();
  // statistics
  uint64_t submitted = 0, completed = 0, launches = 0, cells = 0, rescued = 0;
  double busy_ms = 0;
  std::atomic<uint64_t> grows{0};
  std::atomic<uint64_t> stage_ns{0};
  std::atomic<uint32_t> in_flight{0};  // submitted, not yet finished
};

Server* g_srv = nullptr;

// Merges the slot's requests into its staging buffer and queues the launch
// (slot thread, no lock held).
int launch_slot(Slot& s) {
  // longest haplotypes first, so that the persistent waves meet the longest
  // tiles of every shape bin first
  std::stable_sort(s.reqs.begin(), s.reqs.end(),
                   [](const Request* a, const Request* b) { return a->max_hap > b->max_hap; });
  std::vector<const HostPack*> packs;
  packs.reserve(s.reqs.size());
  for (Request* r : s.reqs) packs.push_back(&r->pack);
  BatchLayout L;
  const uint64_t t0 = now_ns();
  const size_t staging_cap = s.staging.cap;
  int rc = stage_merged(packs.data(), (uint32_t)packs.size(), &s.meta, &L, &s.staging);
  if (rc) return rc;
  g_srv->stage_ns += now_ns() - t0;
  if (s.staging.cap != staging_cap) ++g_srv->grows;
  for (Request* r : s.reqs) r->pack = HostPack();  // the staged image is all the launch needs
  // marginalisation inputs of the requests with sites, after the batch in
  // the arena: region records, the batch's read order, per-site tables
  uint32_t n_mreg = 0, n_sites = 0;
  size_t map_entries = 0, marg_values = 0;
  s.marg_base.assign(s.reqs.size(), 0);
  for (uint32_t x = 0; x < (uint32_t)s.reqs.size(); ++x) {
    const Request* r = s.reqs[x];
    if (r->site_alleles.empty()) continue;
    ++n_mreg;
    n_sites += (uint32_t)r->site_alleles.size();
    map_entries += r->site_map.size();
    s.marg_base[x] = marg_values;
    for (uint32_t al : r->site_alleles) marg_values += (size_t)al * s.meta.region[x].n_reads;
  }
  const size_t n_reads_all = s.meta.read_perm.size();
  const size_t o_reg = 0, o_perm = up256(sizeof(MargRegion) * n_mreg), o_nal = o_perm + up256(4 * n_reads_all),
               o_al0 = o_nal + up256(4ull * n_sites), o_map0 = o_al0 + up256(8ull * n_sites),
               o_map = o_map0 + up256(8ull * n_sites), marg_in_bytes = o_map + up256(2 * map_entries + 2);
  const size_t marg_dev = n_mreg ? up256(L.bytes) : 0;
  const size_t need = n_mreg ? marg_dev + marg_in_bytes + 8 * marg_values : L.bytes;
  if (!s.mem || s.bytes < need) {
    // grow by doubling; the old arena is freed at stop (freeing device
    // memory waits for the whole device; its last launch has finished)
    const size_t want = std::max(need + need / 4 + (16u << 20), 2 * s.bytes);
    if (s.mem) s.retired.push_back(s.mem);
    s.mem = nullptr;
    s.bytes = 0;
    if (hipMalloc(&s.mem, want) != hipSuccess) {
      s.mem = nullptr;
      return set_error(RVC_E_DEVICE, "device arena of %zu bytes", want);
    }
    s.bytes = want;
    ++g_srv->grows;
  }
  // the staging buffer is pinned and is not restaged before this launch ends
  if ((rc = s.batch.upload_staged(s.meta, L, s.staging.ptr, s.stream, s.mem, s.bytes, false, true))) return rc;
  const size_t n = s.meta.n_pairs_slots;
  bool grew = false;
  s.res = g_srv->pool->acquire(n, &rc, &grew);
  if (!s.res) return rc;
  if (grew) ++g_srv->grows;
  if ((rc = s.batch.ensure_events())) return rc;
  // the FP32 pass; the FP64 launches follow once its rescue counts are known
  if ((rc = s.batch.enqueue(s.stream, true, kNoFp64))) return rc;
  s.marg_out_bytes = 8 * marg_values;
  if (n_mreg) {
    const size_t cap_in = s.marg_in.cap, cap_out = s.marg_out.cap;
    if ((rc = s.marg_in.reserve(marg_in_bytes)) || (rc = s.marg_out.reserve(8 * marg_values + 8))) return rc;
    if (s.marg_in.cap != cap_in) ++g_srv->grows;
    if (s.marg_out.cap != cap_out) ++g_srv->grows;
    char* h = s.marg_in.ptr;
    char* d = static_cast<char*>(s.mem) + marg_dev;
    MargRegion* mr = reinterpret_cast<MargRegion*>(h + o_reg);
    uint32_t* nal = reinterpret_cast<uint32_t*>(h + o_nal);
    uint64_t* al0 = reinterpret_cast<uint64_t*>(h + o_al0);
    uint64_t* mp0 = reinterpret_cast<uint64_t*>(h + o_map0);
    int16_t* mp = reinterpret_cast<int16_t*>(h + o_map);
    std::memcpy(h + o_perm, s.meta.read_perm.data(), 4 * n_reads_all);
    uint32_t k = 0, site = 0;
    uint64_t moff = 0;
    for (uint32_t x = 0; x < (uint32_t)s.reqs.size(); ++x) {
      const Request* r = s.reqs[x];
      if (r->site_alleles.empty()) continue;
      const RegionLayout& G = s.meta.region[x];
      MargRegion& g = mr[k++];
      g.pair0 = G.pair0;
      g.n_haps = G.n_haps;
      g.n_reads = G.n_reads;
      g.read0 = G.read0;
      g.n_sites = (uint32_t)r->site_alleles.size();
      g.site0 = site;
      g.pad = 0;
      uint64_t vo = s.marg_base[x];
      for (uint32_t q = 0; q < g.n_sites; ++q, ++site) {
        nal[site] = r->site_alleles[q];
        al0[site] = vo;
        vo += (uint64_t)r->site_alleles[q] * G.n_reads;
        mp0[site] = moff + (uint64_t)q * G.n_haps;
      }
      std::memcpy(mp + moff, r->site_map.data(), 2 * r->site_map.size());
      moff += r->site_map.size();
    }
    MargArgs ma;
    ma.results = s.batch.out;
    ma.regions = reinterpret_cast<const MargRegion*>(d + o_reg);
    ma.read_perm = reinterpret_cast<const uint32_t*>(d + o_perm);
    ma.n_alleles = reinterpret_cast<const uint32_t*>(d + o_nal);
    ma.allele0 = reinterpret_cast<const uint64_t*>(d + o_al0);
    ma.map0 = reinterpret_cast<const uint64_t*>(d + o_map0);
    ma.map = reinterpret_cast<const int16_t*>(d + o_map);
    ma.out = reinterpret_cast<double*>(d + marg_in_bytes);
    ma.n_regions = n_mreg;
    s.marg = ma;
    s.marg_in_bytes = marg_in_bytes;
    s.marg_dev = d;
    s.marg_pending = true;
  }
  // results, flags and counters sit one after another: one copy, by a
  // kernel into the pinned result buffer (a DMA copy could wait behind
  // another slot's copy, and so for that slot's kernels)
  if (launch_copy(s.res->out(), s.batch.out, s.batch.fetch_bytes(), s.stream) != hipSuccess ||
      hipEventRecord(s.done_ev, s.stream) != hipSuccess)
    return set_error(RVC_E_DEVICE, "result copy");
  return RVC_OK;
}

// The second phase of a launch, after its FP32 results, flags and rescue
// counts are in the slot's result buffer: the FP64 launches of the lists that
// hold pairs (grids sized by their counts) with a second copy of results and
// flags, then the marginalisation of the requests with sites. *queued says
// whether anything went on the stream (most launches need neither).
int launch_tail(Slot& s, bool* queued) {
  *queued = false;
  uint32_t rescued[kNumFp64Lists];
  std::memcpy(rescued, s.res->rescued(), sizeof rescued);
  bool any = false;
  for (uint32_t c : rescued) any = any || c != 0;
  int rc;
  if (any) {
    if ((rc = s.batch.enqueue_fp64(s.stream, true, rescued))) return rc;
    const size_t n = s.meta.n_pairs_slots;
    if (launch_copy(s.res->out(), s.batch.out, n * (sizeof(double) + 1), s.stream) != hipSuccess)
      return set_error(RVC_E_DEVICE, "result copy");
    *queued = true;
  }
  if (s.marg_pending) {
    s.marg_pending = false;
    if (launch_copy(s.marg_dev, s.marg_in.ptr, s.marg_in_bytes, s.stream) != hipSuccess ||
        launch_marginalize(s.marg, s.stream) != hipSuccess ||
        launch_copy(s.marg_out.ptr, s.marg.out, s.marg_out_bytes, s.stream) != hipSuccess)
      return set_error(RVC_E_DEVICE, "marginalisation launch");
    *queued = true;
  }
  if (*queued && hipEventRecord(s.done_ev, s.stream) != hipSuccess) return set_error(RVC_E_DEVICE, "launch event");
  return RVC_OK;
}

// Waits for the slot's launch without holding a core: the device time of a
// launch is far above the polling interval.
int wait_slot(Slot& s) {
  for (;;) {
    const hipError_t e = hipEventQuery(s.done_ev);
    if (e == hipSuccess) return RVC_OK;
    if (e != hipErrorNotReady) return set_error(RVC_E_DEVICE, "launch failed: %s", hipGetErrorString(e));
    std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

void slot_main(Slot* sp) {
  Server& S = *g_srv;
  Slot& s = *sp;
  hipSetDevice(g_engine.device);
  std::unique_lock<std::mutex> lk(S.mu);
  for (;;) {
    S.cv_work.wait(lk, [&] {
      return (S.stopping && S.pending.empty()) ||
             (!S.pending.empty() && S.resident + S.pending.front()->resident <= g_engine.hbm_budget);
    });
    if (S.pending.empty()) break;  // stopping and drained
    uint64_t add = 0, pairs = 0;
    while (!S.pending.empty() && s.reqs.size() < g_engine.max_batch_regions) {
      Request* r = S.pending.front();
      if (S.resident + add + r->resident > g_engine.hbm_budget) break;
      if (!s.reqs.empty() && (add + r->resident > kLaunchBytes || pairs + r->pack.n_pairs_slots > kLaunchPairs)) break;
      add += r->resident;
      pairs += r->pack.n_pairs_slots;
      s.reqs.push_back(r);
      S.pending.pop_front();
    }
    s.resident = add;
    S.resident += add;
    if (S.resident > S.peak_resident) S.peak_resident = S.resident;
    ++S.launches;
    const uint64_t t = now_ns();
    for (Request* r : s.reqs) r->t_launch = t;
    lk.unlock();
    int rc = RVC_OK;
    try {  // host allocations of a launch: its requests fail, the loop goes on
      s.marg_pending = false;
      rc = launch_slot(s);
      if (rc == RVC_OK) rc = wait_slot(s);
      bool more = false;
      if (rc == RVC_OK) rc = launch_tail(s, &more);
      if (rc == RVC_OK && more) rc = wait_slot(s);
    } catch (...) {
      rc = set_error(RVC_E_INVALID, "out of host memory for a serving launch");
    }
    float ms = 0;
    uint64_t resc = 0;
    if (rc == RVC_OK) {
      // each request takes its own copy (input read order), so the launch's
      // pinned buffer goes back to the pool now rather than when the last
      // waiter has come: a slow waiter holds no launch buffer
      try {
        const ResultBuf& B = *s.res;
        for (uint32_t r = 0; r < (uint32_t)s.reqs.size(); ++r) {
          Request* rq = s.reqs[r];
          const RegionLayout& L = s.meta.region[r];
          rq->out.resize((size_t)L.n_reads * L.n_haps);
          rq->flags.resize((size_t)L.n_reads * L.n_haps);
          for (uint32_t sx = 0; sx < L.n_reads; ++sx) {
            const uint32_t x = s.meta.read_perm[L.read0 + sx];
            const size_t src = L.pair0 + (size_t)sx * L.n_haps, dst = (size_t)x * L.n_haps;
            std::memcpy(rq->out.data() + dst, B.out() + src, L.n_haps * sizeof(double));
            std::memcpy(rq->flags.data() + dst, B.flags() + src, L.n_haps);
          }
          if (!rq->site_alleles.empty()) {
            size_t nv = 0;
            for (uint32_t al : rq->site_alleles) nv += (size_t)al * L.n_reads;
            rq->marg.resize(nv);
            std::memcpy(rq->marg.data(), reinterpret_cast<const double*>(s.marg_out.ptr) + s.marg_base[r],
                        nv * sizeof(double));
          }
        }
        if (hipEventElapsedTime(&ms, s.batch.ev[0], s.batch.ev[2]) != hipSuccess) ms = 0;
        for (int b = 0; b < kNumFp64Lists; ++b) resc += B.rescued()[b];
      } catch (...) {
        rc = set_error(RVC_E_INVALID, "out of host memory for a serving launch's results");
      }
    } else {
      hipStreamSynchronize(s.stream);
    }
    s.res.reset();
    const uint64_t cells = s.meta.n_cells;
    lk.lock();
    const uint64_t t_done = now_ns();
    for (uint32_t r = 0; r < (uint32_t)s.reqs.size(); ++r) {
      Request* rq = s.reqs[r];
      rq->status = rc;
      rq->t_done = t_done;
      rq->done = true;
    }
    S.completed += s.reqs.size();
    S.in_flight.fetch_sub((uint32_t)s.reqs.size(), std::memory_order_relaxed);
    if (rc == RVC_OK) {
      S.busy_ms += ms;
      S.rescued += resc;
      S.cells += cells;
    }
    S.resident -= s.resident;
    s.resident = 0;
    s.reqs.clear();
    S.cv_done.notify_all();
    S.cv_work.notify_all();  // memory came free: another slot may admit the head now
  }
}

}  // namespace

int server_start() {
  if (g_srv) return set_error(RVC_E_STATE, "server already running");
  auto* S = new (std::nothrow) Server;
  if (!S) return set_error(RVC_E_INVALID, "out of host memory");
  for (Slot& s : S->slots) {
    if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s.done_ev, hipEventDisableTiming) != hipSuccess) {
      for (Slot& t : S->slots) {
        if (t.done_ev) hipEventDestroy(t.done_ev);
        if (t.stream) hipStreamDestroy(t.stream);
      }
      delete S;
      return set_error(RVC_E_DEVICE, "server stream");
    }
    s.batch.stream = s.stream;
    s.batch.own_stream = false;
    // sizes that hold a launch at the bounds (the arena as launch_slot would
    // grow it for one), so that launches do not reallocate (each
    // reallocation waits for the device)
    const uint64_t cap = std::min<uint64_t>(g_engine.hbm_budget, kLaunchBytes);
    const size_t arena = cap + cap / 4 + (16u << 20);
    if (hipMalloc(&s.mem, arena) == hipSuccess) s.bytes = arena;
    else s.mem = nullptr;
    s.staging.reserve(cap);
    // a result buffer per slot
    S->pool->add(kLaunchPairs);
  }
  g_srv = S;
  for (Slot& s : S->slots) s.thr = std::thread(slot_main, &s);
  return RVC_OK;
}

void server_stop() {
  Server* S = g_srv;
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stopping = true;
  }
  S->cv_work.notify_all();
  for (Slot& s : S->slots)
    if (s.thr.joinable()) s.thr.join();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (auto& kv : S->tickets) delete kv.second;
    S->tickets.clear();
  }
  S->cv_done.notify_all();
  for (Slot& s : S->slots) {
    s.batch.release();
    s.res.reset();
    s.staging.release();
    if (s.mem) hipFree(s.mem);
    for (void* p : s.retired) hipFree(p);
    if (s.done_ev) hipEventDestroy(s.done_ev);
    if (s.stream) hipStreamDestroy(s.stream);
  }
  g_srv = nullptr;
  delete S;
}

bool server_running() { return g_srv != nullptr; }

int submit_impl(const rvc_region* region, const DevicePrep* prep, const rvc_sites* sites, rvc_ticket* ticket) {
  if (!ticket) return set_error(RVC_E_INVALID, "null ticket");
  if (!g_engine.ready || !g_srv) return set_error(RVC_E_STATE, "serving loop is not running");
  int rc = validate_region(region);
  if (rc) return rc;
  if (sites && sites->n_sites) {
    if (!sites->n_alleles || !sites->hap_allele) return set_error(RVC_E_INVALID, "null site table");
    for (uint32_t s = 0; s < sites->n_sites; ++s) {
      const uint32_t A = sites->n_alleles[s];
      if (A == 0 || A > 32767) return set_error(RVC_E_INVALID, "site %u has %u alleles", s, A);
      for (uint32_t h = 0; h < region->n_haps; ++h) {
        const int32_t v = sites->hap_allele[(size_t)s * region->n_haps + h];
        if (v < -1 || v >= (int32_t)A)
          return set_error(RVC_E_INVALID, "site %u maps haplotype %u to allele %d of %u", s, h, v, A);
      }
    }
  }
  auto* r = new (std::nothrow) Request;
  if (!r) return set_error(RVC_E_INVALID, "out of host memory");
  if (sites && sites->n_sites) {
    r->site_alleles.assign(sites->n_alleles, sites->n_alleles + sites->n_sites);
    r->site_map.resize((size_t)sites->n_sites * region->n_haps);
    for (size_t x = 0; x < r->site_map.size(); ++x) r->site_map[x] = (int16_t)sites->hap_allele[x];
  }
  // with nothing else in flight the region runs alone: shapes for latency
  const bool alone = g_srv->in_flight.load(std::memory_order_relaxed) == 0;
  const uint64_t pairs = (uint64_t)region->n_haps * region->n_reads;
  rc = build_host_pack(region, 1, alone ? latency_policy(default_policy(), pairs) : default_policy(), &r->pack, prep);
  if (rc) {
    delete r;
    return rc;
  }
  r->resident = footprint(r->pack);
  for (uint32_t al : r->site_alleles) r->resident += 8ull * al * region->n_reads + 8 * 3 + 2ull * region->n_haps;
  for (uint32_t h : r->pack.hap_len) r->max_hap = std::max(r->max_hap, h);
  if (r->resident > g_engine.hbm_budget) {
    const uint64_t need = r->resident;
    delete r;
    return set_error(RVC_E_BUDGET, "region needs %llu device bytes, budget is %llu", (unsigned long long)need,
                     (unsigned long long)g_engine.hbm_budget);
  }
  Server& S = *g_srv;
  {
    std::lock_guard<std::mutex> lk(S.mu);
    if (S.stopping) {
      delete r;
      return set_error(RVC_E_SHUTDOWN, "server is stopping");
    }
    r->id = S.next_id++;
    r->t_submit = now_ns();
    S.tickets[r->id] = r;
    S.pending.push_back(r);
    ++S.submitted;
    S.in_flight.fetch_add(1, std::memory_order_relaxed);
    *ticket = r->id;
  }
  S.cv_work.notify_one();
  return RVC_OK;
}

int wait_impl(rvc_ticket ticket, double* out_read_major, uint8_t* out_flags, int64_t timeout_ms,
              rvc_ticket_timing* timing, std::vector<double>* marg) {
  if (!g_srv) return set_error(RVC_E_STATE, "serving loop is not running");
  Server& S = *g_srv;
  std::unique_lock<std::mutex> lk(S.mu);
  auto it = S.tickets.find(ticket);
  if (it == S.tickets.end()) return set_error(RVC_E_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
  Request* r = it->second;
  if (timeout_ms < 0) {
    S.cv_done.wait(lk, [&] { return r->done; });
  } else if (!S.cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return r->done; })) {
    return set_error(RVC_E_TIMEOUT, "ticket %llu not done after %lld ms", (unsigned long long)ticket,
                     (long long)timeout_ms);
  }
  S.tickets.erase(it);
  lk.unlock();
  std::unique_ptr<Request> own(r);
  const int st = r->status;
  if (timing) {
    timing->submit_ns = r->t_submit;
    timing->launch_ns = r->t_launch;
    timing->done_ns = r->t_done;
  }
  if (st == RVC_OK) {
    if (out_read_major) std::memcpy(out_read_major, r->out.data(), r->out.size() * sizeof(double));
    if (out_flags) std::memcpy(out_flags, r->flags.data(), r->flags.size());
    if (marg) marg->swap(r->marg);
  }
  return st == RVC_OK ? RVC_OK : set_error(st, "request failed with code %d", st);
}

}  // namespace rvc

using namespace rvc;

extern "C" {

int rvc_phmm_submit(const rvc_region* region, rvc_ticket* ticket) {
  return rvc::guarded([&]() -> int { return submit_impl(region, nullptr, nullptr, ticket); });
}

int rvc_phmm_wait(rvc_ticket ticket, double* out_read_major, uint8_t* out_flags, int64_t timeout_ms,
                  rvc_ticket_timing* timing) {
  return rvc::guarded([&]() -> int { return wait_impl(ticket, out_read_major, out_flags, timeout_ms, timing, nullptr); });
}

int rvc_phmm_poll(rvc_ticket ticket) {
  return rvc::guarded([&]() -> int {
    if (!g_srv) return -RVC_E_STATE;
    std::lock_guard<std::mutex> lk(g_srv->mu);
    auto it = g_srv->tickets.find(ticket);
    if (it == g_srv->tickets.end()) return -RVC_E_INVALID;
    return it->second->done ? 1 : 0;
  });
}

int rvc_phmm_server_stats(rvc_server_stats* st) {
  return rvc::guarded([&]() -> int {
    if (!st) return set_error(RVC_E_INVALID, "null stats");
    if (!g_srv) return set_error(RVC_E_STATE, "serving loop is not running");
    std::lock_guard<std::mutex> lk(g_srv->mu);
    st->submitted = g_srv->submitted;
    st->completed = g_srv->completed;
    st->launches = g_srv->launches;
    st->cells = g_srv->cells;
    st->rescued_pairs = g_srv->rescued;
    st->resident_bytes = g_srv->resident;
    st->peak_resident_bytes = g_srv->peak_resident;
    st->busy_ms = g_srv->busy_ms;
    st->stage_ms = (double)g_srv->stage_ns.load() * 1e-6;
    st->buffer_grows = g_srv->grows.load();
    return RVC_OK;
  });
}

}  // extern "C"

namespace rvc {
int served_region(const rvc_region* region, double* out, uint8_t* flags, const DevicePrep* prep) {
  rvc_ticket t = 0;
  const int rc = submit_impl(region, prep, nullptr, &t);
  if (rc) return rc;
  return rvc_phmm_wait(t, out, flags, -1, nullptr);
}
}  // namespace rvc
