// Served form of the aligner (rvc_sw_align_served): Rovaca's workers realign
// region by region, each from its own thread (hc/haplotypecaller_engine.cpp:268,
// one call per region from up to --nthreads workers), far too little work per
// call to fill the device. A call is checked on the caller's thread, queued,
// and the caller sleeps until its results are in place. The server thread
// takes every queued request of the oldest request's score set (up to a
// bound on the merged CIGAR slots), merges them into one batch, runs it
// through the batched path (sw_align_batch, on the server thread's own
// workspace and streams) and spreads the results into the callers' slots;
// what queues up meanwhile makes the next pass. One server thread beats two
// here (169 vs 120 GCUPS from 32 callers): two passes at once split the
// waiting calls into smaller passes that contend for the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "phmm_engine.h"
#include "sw_internal.h"

namespace rvc {
namespace {

constexpr int kSwSlots = 1;  // server threads
constexpr uint64_t kMergedCigarBytes = 256ull << 20;  // bound on one pass's CIGAR slots

struct SwRequest {
  const uint8_t* hap_bases;
  const int32_t* hap_off;
  uint32_t n_haps;
  const uint8_t* read_bases;
  const int32_t* read_off;
  uint32_t n_reads;
  const uint32_t* read_hap;
  int32_t assemble;
  uint32_t max_cigar;
  uint32_t* cigar;
  int32_t* count;
  int32_t* offset;
  int32_t* score;
  int32_t longest;  // longest read
  int rc = RVC_OK;
  std::string err;
  bool done = false;
};

// One server for the process, made on first use and never destroyed: it
// outlives every call (a caller may still be waking on its condition variable
// after a pass ends), stopping joins the threads and leaves the object for
// the next start, and a process that exits without rvc_phmm_shutdown ends
// with its idle threads rather than destroying them.
struct SwServer {
  std::mutex mu;
  std::condition_variable work, done;
  std::deque<SwRequest*> queue;
  std::vector<std::thread> threads;
  bool stopping = false;
};
SwServer& server() {
  static SwServer* s = new SwServer;
  return *s;
}

// CIGAR elements that always suffice for a read of n bases
inline uint64_t slot_for(int32_t longest) { return 2ull * (uint64_t)std::max(longest, 1) + 2; }

// Grow-only pinned bytes for a pass's merged bases: the batched path's
// uploads from them go straight to the device, without staging.
struct PinnedBytes {
  uint8_t* p = nullptr;
  size_t cap = 0, n = 0;
  ~PinnedBytes() {
    if (p) hipHostFree(p);
  }
  bool reserve(size_t want) {
    if (want <= cap) return true;
    uint8_t* q = nullptr;
    const size_t c = std::max(want + want / 2, (size_t)1 << 20);
    if (hipHostMalloc((void**)&q, c, hipHostMallocDefault) != hipSuccess) return false;
    if (p) {
      std::memcpy(q, p, n);
      hipHostFree(p);
    }
    p = q;
    cap = c;
    return true;
  }
  bool append(const uint8_t* src, size_t len) {
    if (!reserve(n + len + 1)) return false;
    std::memcpy(p + n, src, len);
    n += len;
    return true;
  }
};

void pass(const std::vector<SwRequest*>& take, int32_t assemble, int32_t longest, PinnedBytes& hb, PinnedBytes& rb,
          std::vector<int32_t>& ho, std::vector<int32_t>& ro, std::vector<int32_t>& cnt, std::vector<int32_t>& off,
          std::vector<int32_t>& sc, std::vector<uint32_t>& rh, std::vector<uint32_t>& cig);

void serve() {
  SwServer& S = server();
  // the merged batch and its results, kept across passes
  PinnedBytes hb, rb;
  std::vector<int32_t> ho, ro, cnt, off, sc;
  std::vector<uint32_t> rh, cig;
  std::vector<SwRequest*> take;
  for (;;) {
    int32_t assemble = 0, longest = 0;
    {
      std::unique_lock<std::mutex> lk(S.mu);
      S.work.wait(lk, [&] { return S.stopping || !S.queue.empty(); });
      if (S.queue.empty()) return;  // stopping, and nothing left to serve
      assemble = S.queue.front()->assemble;
      take.clear();
      uint64_t reads = 0;
      for (auto it = S.queue.begin(); it != S.queue.end();) {
        SwRequest* q = *it;
        const int32_t l = std::max(longest, q->longest);
        if (q->assemble == assemble &&
            (take.empty() || (reads + q->n_reads) * slot_for(l) * sizeof(uint32_t) <= kMergedCigarBytes)) {
          take.push_back(q);
          reads += q->n_reads;
          longest = l;
          it = S.queue.erase(it);
        } else {
          ++it;
        }
      }
      if (!S.queue.empty()) S.work.notify_one();  // the other score set, for another server thread if any
    }
    try {
      pass(take, assemble, longest, hb, rb, ho, ro, cnt, off, sc, rh, cig);
    } catch (...) {  // host memory for the merged pass: its calls fail, the server goes on
      for (SwRequest* q : take) {
        q->rc = RVC_E_INVALID;
        q->err = "out of host memory for a merged aligner pass";
      }
    }
    {
      std::lock_guard<std::mutex> lk(S.mu);
      for (SwRequest* q : take) q->done = true;
    }
    S.done.notify_all();
  }
}

// One merged pass over the requests in 'take' (results into their slots).
void pass(const std::vector<SwRequest*>& take, int32_t assemble, int32_t longest, PinnedBytes& hb, PinnedBytes& rb,
          std::vector<int32_t>& ho, std::vector<int32_t>& ro, std::vector<int32_t>& cnt, std::vector<int32_t>& off,
          std::vector<int32_t>& sc, std::vector<uint32_t>& rh, std::vector<uint32_t>& cig) {
  // merge: the span of each request's bases that its offsets cover,
  // appended; its offsets moved to where that span lands (each request's
  // first haplotype and read start where the previous request's last ones
  // end), its reads' haplotype indexes by the haplotypes before it
  hb.n = 0;
  rb.n = 0;
  ho.assign(1, 0);
  ro.assign(1, 0);
  rh.clear();
  bool room = true;
  for (SwRequest* q : take) {
    const int32_t hbase = (int32_t)hb.n, rbase = (int32_t)rb.n, h0 = (int32_t)ho.size() - 1;
    const int32_t hfirst = q->hap_off[0], rfirst = q->read_off[0];
    room = room && hb.append(q->hap_bases + hfirst, (size_t)(q->hap_off[q->n_haps] - hfirst)) &&
           rb.append(q->read_bases + rfirst, (size_t)(q->read_off[q->n_reads] - rfirst));
    for (uint32_t h = 1; h <= q->n_haps; ++h) ho.push_back(hbase + (q->hap_off[h] - hfirst));
    for (uint32_t r = 1; r <= q->n_reads; ++r) ro.push_back(rbase + (q->read_off[r] - rfirst));
    for (uint32_t r = 0; r < q->n_reads; ++r) rh.push_back((uint32_t)h0 + q->read_hap[r]);
  }
  const uint32_t n_haps = (uint32_t)ho.size() - 1, n_reads = (uint32_t)rh.size();
  const uint32_t slot = (uint32_t)slot_for(longest);
  cnt.resize(std::max<size_t>(n_reads, 1));
  off.resize(cnt.size());
  sc.resize(cnt.size());
  cig.resize((size_t)std::max<uint32_t>(n_reads, 1) * slot);
  room = room && hb.reserve(1) && rb.reserve(1);  // non-empty base arrays
  const int rc = room ? sw_align_batch(hb.p, ho.data(), n_haps, rb.p, ro.data(), n_reads, rh.data(), assemble, slot,
                                       cig.data(), cnt.data(), off.data(), sc.data(), nullptr)
                      : set_error(RVC_E_DEVICE, "pinned host memory for a merged pass");
  const std::string err = rc ? last_error() : std::string();
  // spread the results into the callers' slots
  uint32_t r0 = 0;
  for (SwRequest* q : take) {
    q->rc = rc;
    q->err = err;
    if (rc == RVC_OK) {
      std::memcpy(q->count, cnt.data() + r0, 4ull * q->n_reads);
      std::memcpy(q->offset, off.data() + r0, 4ull * q->n_reads);
      std::memcpy(q->score, sc.data() + r0, 4ull * q->n_reads);
      for (uint32_t r = 0; r < q->n_reads; ++r) {
        const int32_t k = cnt[r0 + r];
        if (k <= 0) continue;
        if ((uint32_t)k > q->max_cigar) {
          q->rc = RVC_E_LIMIT;
          q->err = "read " + std::to_string(r) + " needs more than " + std::to_string(q->max_cigar) +
                   " CIGAR elements";
          continue;
        }
        std::memcpy(q->cigar + (size_t)r * q->max_cigar, cig.data() + (size_t)(r0 + r) * slot, 4ull * (uint32_t)k);
      }
    }
    r0 += q->n_reads;
  }
}

}  // namespace

int sw_align_served(const uint8_t* hap_bases, const int32_t* hap_off, uint32_t n_haps, const uint8_t* read_bases,
                    const int32_t* read_off, uint32_t n_reads, const uint32_t* read_hap, int32_t assemble,
                    uint32_t max_cigar, uint32_t* cigar, int32_t* count, int32_t* offset, int32_t* score) {
  int32_t longest = 0;
  if (int rc = sw_validate(hap_bases, hap_off, n_haps, read_bases, read_off, n_reads, read_hap, max_cigar, cigar,
                           count, offset, score, &longest))
    return rc;
  if (n_reads == 0) return RVC_OK;
  SwRequest q{hap_bases, hap_off, n_haps, read_bases, read_off, n_reads, read_hap, assemble ? 1 : 0,
              max_cigar, cigar,   count,   offset, score,      longest};
  SwServer& S = server();
  {
    std::lock_guard<std::mutex> lk(S.mu);
    if (S.stopping) return set_error(RVC_E_STATE, "the engine is shutting down");
    if (S.threads.empty()) {
      for (int k = 0; k < kSwSlots; ++k) {
        try {
          S.threads.emplace_back(serve);
        } catch (...) {
          return set_error(RVC_E_INVALID, "aligner server thread");
        }
      }
    }
    S.queue.push_back(&q);
  }
  S.work.notify_one();
  {
    std::unique_lock<std::mutex> lk(S.mu);
    S.done.wait(lk, [&] { return q.done; });
  }
  if (q.rc != RVC_OK) return set_error(q.rc, "%s", q.err.c_str());
  return RVC_OK;
}

void sw_server_stop() {
  SwServer& S = server();
  std::vector<std::thread> threads;
  {
    std::lock_guard<std::mutex> lk(S.mu);
    S.stopping = true;
    threads.swap(S.threads);
  }
  S.work.notify_all();
  for (std::thread& t : threads) t.join();  // each drains the queue first
  std::lock_guard<std::mutex> lk(S.mu);
  S.stopping = false;
}

}  // namespace rvc

extern "C" int rvc_sw_align_served(const uint8_t* hap_bases, const int32_t* hap_off, uint32_t n_haps,
                                   const uint8_t* read_bases, const int32_t* read_off, uint32_t n_reads,
                                   const uint32_t* read_hap, int32_t assemble, uint32_t max_cigar, uint32_t* cigar,
                                   int32_t* count, int32_t* offset, int32_t* score) {
  return rvc::guarded([&]() -> int {
    return rvc::sw_align_served(hap_bases, hap_off, n_haps, read_bases, read_off, n_reads, read_hap, assemble,
                                max_cigar, cigar, count, offset, score);
  });
}
