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

struct Request {
  uint64_t id = 0;
  std::vector<int32_t> hap_off, read_off;
  std::vector<uint8_t> hap_bases, read_bases, q, i, d, c;
  rvc_region view{};
  uint64_t resident = 0;
  uint64_t cells = 0;
  std::vector<double> out;
  std::vector<uint8_t> flags;
  int status = 0;
  bool done = false;
};

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

// Device bytes a region holds while in flight (packed sequences, tiles,
// results, rescue list), an upper bound of what DeviceBatch::upload lays out.
uint64_t footprint(const rvc_region& r) {
  const uint64_t H = r.n_haps, N = r.n_reads;
  const uint64_t hb = (uint64_t)r.hap_off[H] + 32 * H, rb = (uint64_t)r.read_off[N] + 32 * N;
  const uint64_t pairs = H * N;
  return hb / 2 + 8 * H + rb * 5 + 12 * N + 16 * (H * ((N + 3) / 4) + 1) + pairs * (8 + 1 + 16) + 4096;
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
  for (Request* r : reqs) {
    r->status = code;
    r->done = true;
  }
}

// Packs and launches the slot's requests (server thread, no lock held).
int launch_slot(Slot& s) {
  std::vector<rvc_region> regs;
  regs.reserve(s.reqs.size());
  for (Request* r : s.reqs) regs.push_back(r->view);
  static FP32Policy pol;
  int rc = build_host_pack(regs.data(), (uint32_t)regs.size(), pol, &s.pack);
  if (rc) return rc;
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

// Scatters a finished slot's results into its requests (server thread).
void finish_slot(Slot& s) {
  const HostPack& P = s.pack;
  float ms = 0;
  if (hipEventElapsedTime(&ms, s.batch.ev[0], s.batch.ev[2]) != hipSuccess) ms = 0;
  for (uint32_t r = 0; r < P.n_regions; ++r) {
    const RegionLayout& L = P.region[r];
    Request* rq = s.reqs[r];
    rq->out.resize((size_t)L.n_haps * L.n_reads);
    rq->flags.resize((size_t)L.n_haps * L.n_reads);
    for (uint32_t sx = 0; sx < L.n_reads; ++sx) {
      const uint32_t x = P.read_perm[L.read0 + sx];
      std::memcpy(&rq->out[(size_t)x * L.n_haps], s.h_out + L.pair0 + (size_t)sx * L.n_haps,
                  L.n_haps * sizeof(double));
      std::memcpy(&rq->flags[(size_t)x * L.n_haps], s.h_flags + L.pair0 + (size_t)sx * L.n_haps, L.n_haps);
    }
  }
  uint64_t resc = 0;
  for (size_t x = 0; x < P.n_pairs_slots; ++x) resc += (s.h_flags[x] & RVC_PAIR_FP64) ? 1 : 0;
  std::lock_guard<std::mutex> lk(g_srv->mu);
  g_srv->busy_ms += ms;
  g_srv->rescued += resc;
  g_srv->cells += P.n_cells;
  for (Request* rq : s.reqs) {
    rq->status = RVC_OK;
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
  const uint32_t H = region->n_haps, N = region->n_reads;
  r->hap_off.assign(region->hap_off, region->hap_off + H + 1);
  r->read_off.assign(region->read_off, region->read_off + N + 1);
  const int32_t hb = region->hap_off[H] - region->hap_off[0], rb = region->read_off[N] - region->read_off[0];
  const int32_t h0 = region->hap_off[0], r0 = region->read_off[0];
  for (auto& v : r->hap_off) v -= h0;
  for (auto& v : r->read_off) v -= r0;
  if (hb > 0) r->hap_bases.assign(region->hap_bases + h0, region->hap_bases + h0 + hb);
  if (rb > 0) {
    r->read_bases.assign(region->read_bases + r0, region->read_bases + r0 + rb);
    r->q.assign(region->q + r0, region->q + r0 + rb);
    r->i.assign(region->i + r0, region->i + r0 + rb);
    r->d.assign(region->d + r0, region->d + r0 + rb);
    r->c.assign(region->c + r0, region->c + r0 + rb);
  }
  r->view.n_haps = H;
  r->view.n_reads = N;
  r->view.hap_off = r->hap_off.data();
  r->view.hap_bases = r->hap_bases.data();
  r->view.read_off = r->read_off.data();
  r->view.read_bases = r->read_bases.data();
  r->view.q = r->q.data();
  r->view.i = r->i.data();
  r->view.d = r->d.data();
  r->view.c = r->c.data();
  r->resident = footprint(r->view);
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
    S.tickets[r->id] = r;
    S.pending.push_back(r);
    ++S.submitted;
    *ticket = r->id;
  }
  S.cv_work.notify_one();
  return RVC_OK;
}

int rvc_phmm_wait(rvc_ticket ticket, double* out_read_major, uint8_t* out_flags, int64_t timeout_ms) {
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
  if (st == RVC_OK) {
    if (out_read_major) std::memcpy(out_read_major, r->out.data(), r->out.size() * sizeof(double));
    if (out_flags) std::memcpy(out_flags, r->flags.data(), r->flags.size());
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
