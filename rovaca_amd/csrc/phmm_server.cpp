// Serving loop: continuous batching of active regions submitted by many host
// threads (Rovaca's worker pool calls call_pairhmm once per region from up to
// --nthreads workers, hc/haplotypecaller_engine.cpp:119,254).
//
// submit() copies a region and returns a ticket at once. One server thread
// owns kSlots launch slots, each with its own stream and device arena. Whenever
// a slot is free it admits every waiting request, in arrival order, whose
// device footprint still fits the HBM budget next to the requests already in
// flight, packs them into one batch and queues the FP32 pass, the FP64 rescue
// and the copy of the results on the slot's stream. Slots run concurrently,
// so requests that arrive while earlier launches are on the device start
// while those finish. Completed tickets are released to their waiters one by
// one.
#include <hip/hip_runtime.h>

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
constexpr int kSlots = 3;

// Results of one finished launch, shared by its requests: each waiter copies
// its own region out (rows back in input order) from here.
struct BatchResult {
  std::vector<double> out;
  std::vector<uint8_t> flags;
  std::vector<RegionLayout> region;
  std::vector<uint32_t> read_perm;
};

struct Request {
  uint64_t id = 0;
  HostPack pack;  // the region, packed by the submitting thread
  uint64_t resident = 0;
  std::shared_ptr<const BatchResult> result;
  uint32_t result_region = 0;
  int status = 0;
  bool done = false;
  uint64_t t_submit = 0, t_launch = 0, t_done = 0;
};

uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now().time_since_epoch()).count();
}

struct Slot {
  hipStream_t stream = nullptr;
  void* mem = nullptr;
  size_t bytes = 0;
  double* h_out = nullptr;   // pinned
  uint8_t* h_flags = nullptr;
  size_t h_cap = 0;
  hipEvent_t done_ev = nullptr;
  HostPack pack;
  DeviceBatch batch;
  std::vector<Request*> reqs;
  uint64_t resident = 0;
  bool busy = false;
};

// Device bytes a packed region holds while in flight (sequences, tiles,
// results, rescue list): what DeviceBatch::upload lays out for it.
uint64_t footprint(const HostPack& p) {
  const uint64_t words = p.hap_pack.size() + p.hap_nmask.size() + p.read_pack.size() + p.read_nmask.size();
  const uint64_t index = p.hap_off.size() + p.hap_len.size() + p.read_off.size() + p.read_len.size();
  return 4 * (words + index) + 4 * p.rq.size() + sizeof(Tile) * p.tiles.size() +
         p.n_pairs_slots * (sizeof(double) + 1 + sizeof(RescueEntry)) + 4096;
}

struct Server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending;
  std::unordered_map<uint64_t, Request*> tickets;
  uint64_t next_id = 1;
  uint64_t resident = 0, peak_resident = 0;
  bool running = false, stopping = false;
  std::thread thr;
  Slot slots[kSlots];
  // statistics
  uint64_t submitted = 0, completed = 0, launches = 0, cells = 0, rescued = 0;
  double busy_ms = 0;
};

Server* g_srv = nullptr;

void fail_requests(std::vector<Request*>& reqs, int code) {
  const uint64_t t = now_ns();
  for (Request* r : reqs) {
    r->status = code;
    r->t_done = t;
    r->done = true;
  }
}

// Merges the slot's packed requests into one batch and launches it (server
// thread, no lock held).
int launch_slot(Slot& s) {
  std::vector<const HostPack*> packs;
  packs.reserve(s.reqs.size());
  for (Request* r : s.reqs) packs.push_back(&r->pack);
  int rc = merge_host_packs(packs.data(), (uint32_t)packs.size(), &s.pack);
  if (rc) return rc;
  for (Request* r : s.reqs) r->pack = HostPack();  // merged copy is all the launch needs
  rc = s.batch.upload(s.pack, s.stream, s.mem, s.bytes);
  if (rc) return rc;
  if (s.batch.own_mem) {
    if (s.mem) hipFree(s.mem);
    s.mem = s.batch.mem;
    s.bytes = s.batch.bytes;
    s.batch.own_mem = false;
  }
  const size_t n = s.pack.n_pairs_slots;
  if (n > s.h_cap) {
    if (s.h_out) hipHostFree(s.h_out);
    if (s.h_flags) hipHostFree(s.h_flags);
    s.h_out = nullptr;
    s.h_flags = nullptr;
    size_t cap = n + n / 2 + 1024;
    if (hipHostMalloc((void**)&s.h_out, cap * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&s.h_flags, cap, hipHostMallocDefault) != hipSuccess)
      return set_error(RVC_E_DEVICE, "pinned result buffers");
    s.h_cap = cap;
  }
  if ((rc = s.batch.ensure_events())) return rc;
  if ((rc = s.batch.enqueue(s.stream, true))) return rc;
  if (hipMemcpyAsync(s.h_out, s.batch.out, n * sizeof(double), hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipMemcpyAsync(s.h_flags, s.batch.flags, n, hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipEventRecord(s.done_ev, s.stream) != hipSuccess)
    return set_error(RVC_E_DEVICE, "result copy");
  return RVC_OK;
}

// Hands a finished slot's results to its requests (server thread): one bulk
// copy out of the pinned buffers; each waiter then takes its own rows.
void finish_slot(Slot& s) {
  const HostPack& P = s.pack;
  float ms = 0;
  if (hipEventElapsedTime(&ms, s.batch.ev[0], s.batch.ev[2]) != hipSuccess) ms = 0;
  auto res = std::make_shared<BatchResult>();
  const size_t n = P.n_pairs_slots;
  res->out.assign(s.h_out, s.h_out + n);
  res->flags.assign(s.h_flags, s.h_flags + n);
  res->region = P.region;
  res->read_perm = P.read_perm;
  uint64_t resc = 0;
  for (size_t x = 0; x < n; ++x) resc += (s.h_flags[x] & RVC_PAIR_FP64) ? 1 : 0;
  std::lock_guard<std::mutex> lk(g_srv->mu);
  g_srv->busy_ms += ms;
  g_srv->rescued += resc;
  g_srv->cells += P.n_cells;
  const uint64_t t = now_ns();
  for (uint32_t r = 0; r < (uint32_t)s.reqs.size(); ++r) {
    Request* rq = s.reqs[r];
    rq->result = res;
    rq->result_region = r;
    rq->status = RVC_OK;
    rq->t_done = t;
    rq->done = true;
    ++g_srv->completed;
  }
  g_srv->resident -= s.resident;
  s.resident = 0;
  s.reqs.clear();
  s.busy = false;
  g_srv->cv_done.notify_all();
}

void server_main() {
  Server& S = *g_srv;
  hipSetDevice(g_engine.device);
  std::unique_lock<std::mutex> lk(S.mu);
  for (;;) {
    bool progress = false;
    // 1. launch: fill every free slot with what fits the budget
    for (Slot& s : S.slots) {
      if (s.busy || S.pending.empty()) continue;
      uint64_t add = 0;
      while (!S.pending.empty() && s.reqs.size() < g_engine.max_batch_regions) {
        Request* r = S.pending.front();
        if (S.resident + add + r->resident > g_engine.hbm_budget) break;
        add += r->resident;
        s.reqs.push_back(r);
        S.pending.pop_front();
      }
      if (s.reqs.empty()) break;  // the head does not fit yet: wait for a slot to finish
      s.busy = true;
      s.resident = add;
      const uint64_t t = now_ns();
      for (Request* r : s.reqs) r->t_launch = t;
      S.resident += add;
      if (S.resident > S.peak_resident) S.peak_resident = S.resident;
      ++S.launches;
      lk.unlock();
      const int rc = launch_slot(s);
      lk.lock();
      if (rc) {
        hipStreamSynchronize(s.stream);
        fail_requests(s.reqs, rc);
        S.completed += s.reqs.size();
        S.resident -= s.resident;
        s.resident = 0;
        s.reqs.clear();
        s.busy = false;
        S.cv_done.notify_all();
      }
      progress = true;
    }
    // 2. complete finished slots
    for (Slot& s : S.slots) {
      if (!s.busy) continue;
      if (hipEventQuery(s.done_ev) == hipSuccess) {
        lk.unlock();
        finish_slot(s);
        lk.lock();
        progress = true;
      }
    }
    if (progress) continue;
    bool any_busy = false;
    for (Slot& s : S.slots) any_busy |= s.busy;
    if (S.stopping && !any_busy && S.pending.empty()) break;
    if (any_busy)
      S.cv_work.wait_for(lk, std::chrono::microseconds(20));
    else
      S.cv_work.wait(lk);
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
      delete S;
      return set_error(RVC_E_DEVICE, "server stream");
    }
    s.batch.stream = s.stream;
    s.batch.own_stream = false;
  }
  S->running = true;
  g_srv = S;
  S->thr = std::thread(server_main);
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
  if (S->thr.joinable()) S->thr.join();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (Request* r : S->pending) {
      r->status = RVC_E_SHUTDOWN;
      r->done = true;
    }
    S->pending.clear();
    for (auto& kv : S->tickets) delete kv.second;
    S->tickets.clear();
  }
  S->cv_done.notify_all();
  for (Slot& s : S->slots) {
    s.batch.release();
    if (s.mem) hipFree(s.mem);
    if (s.h_out) hipHostFree(s.h_out);
    if (s.h_flags) hipHostFree(s.h_flags);
    if (s.done_ev) hipEventDestroy(s.done_ev);
    if (s.stream) hipStreamDestroy(s.stream);
  }
  g_srv = nullptr;
  delete S;
}

}  // namespace rvc

using namespace rvc;

extern "C" {

int rvc_phmm_submit(const rvc_region* region, rvc_ticket* ticket) {
  if (!ticket) return set_error(RVC_E_INVALID, "null ticket");
  if (!g_engine.ready || !g_srv) return set_error(RVC_E_STATE, "serving loop is not running");
  int rc = validate_region(region);
  if (rc) return rc;
  auto* r = new (std::nothrow) Request;
  if (!r) return set_error(RVC_E_INVALID, "out of host memory");
  rc = build_host_pack(region, 1, default_policy(), &r->pack);
  if (rc) {
    delete r;
    return rc;
  }
  r->resident = footprint(r->pack);
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
    *ticket = r->id;
  }
  S.cv_work.notify_one();
  return RVC_OK;
}

int rvc_phmm_wait(rvc_ticket ticket, double* out_read_major, uint8_t* out_flags, int64_t timeout_ms,
                  rvc_ticket_timing* timing) {
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
  const int st = r->status;
  if (timing) {
    timing->submit_ns = r->t_submit;
    timing->launch_ns = r->t_launch;
    timing->done_ns = r->t_done;
  }
  if (st == RVC_OK && (out_read_major || out_flags)) {
    const BatchResult& B = *r->result;
    const RegionLayout& L = B.region[r->result_region];
    for (uint32_t sx = 0; sx < L.n_reads; ++sx) {
      const uint32_t x = B.read_perm[L.read0 + sx];
      const size_t src = L.pair0 + (size_t)sx * L.n_haps, dst = (size_t)x * L.n_haps;
      if (out_read_major) std::memcpy(out_read_major + dst, B.out.data() + src, L.n_haps * sizeof(double));
      if (out_flags) std::memcpy(out_flags + dst, B.flags.data() + src, L.n_haps);
    }
  }
  delete r;
  return st == RVC_OK ? RVC_OK : set_error(st, "request failed with code %d", st);
}

int rvc_phmm_poll(rvc_ticket ticket) {
  if (!g_srv) return -RVC_E_STATE;
  std::lock_guard<std::mutex> lk(g_srv->mu);
  auto it = g_srv->tickets.find(ticket);
  if (it == g_srv->tickets.end()) return -RVC_E_INVALID;
  return it->second->done ? 1 : 0;
}

int rvc_phmm_server_stats(rvc_server_stats* st) {
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
  return RVC_OK;
}

}  // extern "C"
