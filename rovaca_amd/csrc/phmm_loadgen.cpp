// Load generator for the serving loop and the served aligner (see include/rovaca_phmm_loadgen.h).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/rovaca_phmm.h"
#include "../../include/rovaca_phmm_loadgen.h"
#include "phmm_engine.h"

namespace {
using Clock = std::chrono::steady_clock;

uint64_t ns_of(Clock::time_point t) {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(t.time_since_epoch()).count();
}

double pct(std::vector<double> v, double p) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  const size_t i = std::min(v.size() - 1, (size_t)(p * (double)(v.size() - 1) + 0.5));
  return v[i];
}
}  // namespace

extern "C" int rvc_loadgen_run(const rvc_region* regions, uint32_t n_regions, const uint32_t* order,
                               const double* arrival_s, uint32_t n_arrivals, uint32_t n_producers,
                               double* latency_ms, double* const* out_read_major, rvc_loadgen_result* result) {
  return rvc::guarded([&]() -> int {
    using rvc::set_error;
    if (!regions || !arrival_s || !result || n_producers == 0 || n_regions == 0)
      return set_error(RVC_E_INVALID, "bad load generator arguments");
    for (uint32_t i = 0; i < n_arrivals; ++i)
      if (order && order[i] >= n_regions) return set_error(RVC_E_INVALID, "order[%u] out of range", i);
    std::vector<rvc_ticket> tickets(n_arrivals, 0);
    std::vector<int> sub_rc(n_arrivals, 0);
    std::vector<double> late(n_arrivals, 0);
    std::vector<rvc_ticket_timing> timing(n_arrivals);
    std::vector<int> wait_rc(n_arrivals, 0);
    const Clock::time_point start = Clock::now() + std::chrono::milliseconds(20);
    // Each producer submits at its arrival times (open loop: a submission never
    // waits for an earlier result) while a collector thread of its own waits for
    // its tickets in order, as the worker that would consume each result.
    auto producer = [&](uint32_t p) {
      std::mutex mu;
      std::condition_variable cv;
      uint32_t n_sub = 0;  // this producer's arrivals submitted so far
      std::thread collector([&] {
        uint32_t k = 0;
        for (uint32_t i = p; i < n_arrivals; i += n_producers, ++k) {
          {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return n_sub > k; });
          }
          if (sub_rc[i] != RVC_OK) continue;
          wait_rc[i] = rvc_phmm_wait(tickets[i], out_read_major ? out_read_major[i] : nullptr, nullptr, -1,
                                     &timing[i]);
        }
      });
      for (uint32_t i = p; i < n_arrivals; i += n_producers) {
        const Clock::time_point due = start + std::chrono::nanoseconds((int64_t)(arrival_s[i] * 1e9));
        std::this_thread::sleep_until(due);
        late[i] = std::chrono::duration<double, std::milli>(Clock::now() - due).count();
        const rvc_region* rg = &regions[order ? order[i] : i % n_regions];
        sub_rc[i] = rvc_phmm_submit(rg, &tickets[i]);
        {
          std::lock_guard<std::mutex> lk(mu);
          ++n_sub;
        }
        cv.notify_one();
      }
      collector.join();
    };
    std::vector<std::thread> th;
    for (uint32_t p = 0; p < n_producers; ++p) th.emplace_back(producer, p);
    for (auto& t : th) t.join();
    std::memset(result, 0, sizeof *result);
    std::vector<double> lat, q, lt;
    uint64_t first = UINT64_MAX, last = 0;
    for (uint32_t i = 0; i < n_arrivals; ++i) {
      if (sub_rc[i] != RVC_OK) {
        result->failed++;
        if (latency_ms) latency_ms[i] = -1;
        continue;
      }
      result->submitted++;
      if (wait_rc[i] != RVC_OK) {
        result->failed++;
        if (latency_ms) latency_ms[i] = -1;
        continue;
      }
      result->completed++;
      const rvc_region& rg = regions[order ? order[i] : i % n_regions];
      uint64_t hs = 0, rs = 0;
      for (uint32_t h = 0; h < rg.n_haps; ++h) hs += (uint64_t)(rg.hap_off[h + 1] - rg.hap_off[h]);
      for (uint32_t r = 0; r < rg.n_reads; ++r) rs += (uint64_t)(rg.read_off[r + 1] - rg.read_off[r]);
      result->cells += hs * rs;
      const double l = (double)(timing[i].done_ns - timing[i].submit_ns) / 1e6;
      lat.push_back(l);
      q.push_back((double)(timing[i].launch_ns - timing[i].submit_ns) / 1e6);
      lt.push_back(late[i]);
      if (latency_ms) latency_ms[i] = l;
      first = std::min(first, timing[i].submit_ns);
      last = std::max(last, timing[i].done_ns);
    }
    (void)ns_of;
    if (!lat.empty()) {
      result->wall_s = (double)(last - first) / 1e9;
      result->lat_p50_ms = pct(lat, 0.50);
      result->lat_p99_ms = pct(lat, 0.99);
      result->lat_max_ms = pct(lat, 1.0);
      double s = 0;
      for (double x : lat) s += x;
      result->lat_mean_ms = s / (double)lat.size();
      result->queue_p50_ms = pct(q, 0.50);
      result->queue_p99_ms = pct(q, 0.99);
      result->late_p99_ms = pct(lt, 0.99);
    }
    return result->failed ? set_error(RVC_E_DEVICE, "%llu requests failed", (unsigned long long)result->failed)
                          : RVC_OK;
  });
}

extern "C" int rvc_loadgen_region_latency(const rvc_region* region, uint32_t n_calls, int32_t noalloc,
                                          double* latency_ms) {
  return rvc::guarded([&]() -> int {
    using rvc::set_error;
    if (!region || !latency_ms) return set_error(RVC_E_INVALID, "bad arguments");
    std::vector<double> out((size_t)region->n_haps * region->n_reads);
    std::vector<uint8_t> flags(out.size());
    for (uint32_t i = 0; i < n_calls; ++i) {
      const Clock::time_point a = Clock::now();
      const int rc = noalloc ? rvc_phmm_region_sync_noalloc(region, out.data(), flags.data())
                             : rvc_phmm_region_sync(region, out.data(), flags.data());
      latency_ms[i] = std::chrono::duration<double, std::milli>(Clock::now() - a).count();
      if (rc) return rc;
    }
    return RVC_OK;
  });
}

extern "C" int rvc_loadgen_call_pairhmm(const rvc_raw_region* regions, uint32_t n_regions, const uint32_t* order,
                                        uint32_t n_calls, int32_t min_quality_threshold, int32_t pcr_option,
                                        uint32_t n_workers, double* latency_ms, rvc_loadgen_result* result) {
  return rvc::guarded([&]() -> int {
    using rvc::set_error;
    if (!regions || !result || n_workers == 0 || n_regions == 0) return set_error(RVC_E_INVALID, "bad arguments");
    for (uint32_t i = 0; i < n_calls; ++i)
      if (order && order[i] >= n_regions) return set_error(RVC_E_INVALID, "order[%u] out of range", i);
    std::vector<double> lat(n_calls, 0);
    std::vector<int> rcs(n_calls, 0);
    std::vector<uint64_t> t0(n_calls, 0), t1(n_calls, 0);
    auto worker = [&](uint32_t w) {
      std::vector<double> mat;
      std::vector<uint32_t> kept;
      for (uint32_t i = w; i < n_calls; i += n_workers) {
        const rvc_raw_region& rg = regions[order ? order[i] : i % n_regions];
        mat.resize((size_t)rg.n_haps * rg.n_reads);
        kept.resize(rg.n_reads);
        uint32_t nk = 0;
        const Clock::time_point a = Clock::now();
        rcs[i] = rvc_phmm_call_pairhmm(&rg, min_quality_threshold, pcr_option, mat.data(), kept.data(), &nk);
        const Clock::time_point b = Clock::now();
        t0[i] = ns_of(a);
        t1[i] = ns_of(b);
        lat[i] = std::chrono::duration<double, std::milli>(b - a).count();
      }
    };
    std::vector<std::thread> th;
    for (uint32_t w = 0; w < n_workers; ++w) th.emplace_back(worker, w);
    for (auto& t : th) t.join();
    std::memset(result, 0, sizeof *result);
    std::vector<double> ok_lat;
    uint64_t first = UINT64_MAX, last = 0;
    for (uint32_t i = 0; i < n_calls; ++i) {
      result->submitted++;
      if (rcs[i] != RVC_OK) {
        result->failed++;
        if (latency_ms) latency_ms[i] = -1;
        continue;
      }
      result->completed++;
      const rvc_raw_region& rg = regions[order ? order[i] : i % n_regions];
      uint64_t hs = 0, rs = 0;
      for (uint32_t h = 0; h < rg.n_haps; ++h) hs += (uint64_t)(rg.hap_off[h + 1] - rg.hap_off[h]);
      for (uint32_t r = 0; r < rg.n_reads; ++r) rs += (uint64_t)(rg.read_off[r + 1] - rg.read_off[r]);
      result->cells += hs * rs;
      ok_lat.push_back(lat[i]);
      if (latency_ms) latency_ms[i] = lat[i];
      first = std::min(first, t0[i]);
      last = std::max(last, t1[i]);
    }
    if (!ok_lat.empty()) {
      result->wall_s = (double)(last - first) / 1e9;
      result->lat_p50_ms = pct(ok_lat, 0.50);
      result->lat_p99_ms = pct(ok_lat, 0.99);
      result->lat_max_ms = pct(ok_lat, 1.0);
      double sum = 0;
      for (double x : ok_lat) sum += x;
      result->lat_mean_ms = sum / (double)ok_lat.size();
    }
    return result->failed ? set_error(RVC_E_DEVICE, "%llu calls failed", (unsigned long long)result->failed) : RVC_OK;
  });
}

extern "C" int rvc_loadgen_sw_align(const rvc_sw_region* regions, uint32_t n_regions, int32_t assemble, int32_t served,
                                    uint32_t n_workers, double* latency_ms, rvc_loadgen_result* result) {
  return rvc::guarded([&]() -> int {
    using rvc::set_error;
    if (!regions || !result || n_workers == 0) return set_error(RVC_E_INVALID, "bad arguments");
    std::vector<double> lat(n_regions, 0);
    std::vector<int> rcs(n_regions, 0);
    std::vector<uint64_t> t0(n_regions, 0), t1(n_regions, 0);
    std::atomic<uint32_t> next{0};
    auto worker = [&] {
      for (;;) {
        const uint32_t i = next.fetch_add(1, std::memory_order_relaxed);
        if (i >= n_regions) return;
        const rvc_sw_region& g = regions[i];
        const Clock::time_point a = Clock::now();
        rcs[i] = served ? rvc_sw_align_served(g.hap_bases, g.hap_off, g.n_haps, g.read_bases, g.read_off, g.n_reads,
                                              g.read_hap, assemble, g.max_cigar, g.cigar, g.count, g.offset, g.score)
                        : rvc_sw_align_batch(g.hap_bases, g.hap_off, g.n_haps, g.read_bases, g.read_off, g.n_reads,
                                             g.read_hap, assemble, g.max_cigar, g.cigar, g.count, g.offset, g.score,
                                             nullptr);
        const Clock::time_point b = Clock::now();
        t0[i] = ns_of(a);
        t1[i] = ns_of(b);
        lat[i] = std::chrono::duration<double, std::milli>(b - a).count();
      }
    };
    std::vector<std::thread> th;
    for (uint32_t w = 0; w < n_workers; ++w) th.emplace_back(worker);
    for (auto& t : th) t.join();
    std::memset(result, 0, sizeof *result);
    std::vector<double> ok_lat;
    uint64_t first = UINT64_MAX, last = 0;
    for (uint32_t i = 0; i < n_regions; ++i) {
      result->submitted++;
      if (rcs[i] != RVC_OK) {
        result->failed++;
        if (latency_ms) latency_ms[i] = -1;
        continue;
      }
      result->completed++;
      const rvc_sw_region& g = regions[i];
      for (uint32_t r = 0; r < g.n_reads; ++r) {
        const uint32_t h = g.read_hap[r];
        result->cells += (uint64_t)(g.hap_off[h + 1] - g.hap_off[h]) * (uint64_t)(g.read_off[r + 1] - g.read_off[r]);
      }
      ok_lat.push_back(lat[i]);
      if (latency_ms) latency_ms[i] = lat[i];
      first = std::min(first, t0[i]);
      last = std::max(last, t1[i]);
    }
    if (!ok_lat.empty()) {
      result->wall_s = (double)(last - first) / 1e9;
      result->lat_p50_ms = pct(ok_lat, 0.50);
      result->lat_p99_ms = pct(ok_lat, 0.99);
      result->lat_max_ms = pct(ok_lat, 1.0);
      double sum = 0;
      for (double x : ok_lat) sum += x;
      result->lat_mean_ms = sum / (double)ok_lat.size();
    }
    return result->failed ? set_error(RVC_E_DEVICE, "%llu calls failed", (unsigned long long)result->failed) : RVC_OK;
  });
}
