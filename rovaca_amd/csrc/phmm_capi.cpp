// The C ABI of include/rovaca_phmm.h: lifecycle, one-region calls, batches.
// The serving entry points live in phmm_server.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

#include "../../include/rovaca_phmm.h"
#include "phmm_adapter.h"
#include "phmm_engine.h"
#include "phmm_server.h"
#include "sw_internal.h"

using namespace rvc;

struct rvc_batch {
  HostPack pack;
  DeviceBatch dev;
  bool use_graph = false;
  bool ran = false;
};

namespace {

// Per-thread device workspace of the one-region path: a stream and a grow-only
// device allocation reused from call to call. Workspaces are owned by a
// registry and released by rvc_phmm_shutdown, never by thread exit.
struct SyncWorkspace {
  hipStream_t stream = nullptr;
  void* mem = nullptr;
  size_t bytes = 0;
  HostPack pack;
  PinnedBuf staging, results;
  DeviceBatch batch;
};
std::mutex g_ws_mu;
std::vector<SyncWorkspace*> g_ws_all;
thread_local SyncWorkspace* t_ws = nullptr;
uint64_t g_ws_generation = 1;
thread_local uint64_t t_ws_generation = 0;

int get_workspace(SyncWorkspace** out) {
  if (!t_ws || t_ws_generation != g_ws_generation) {
    auto* w = new (std::nothrow) SyncWorkspace;
    if (!w) return set_error(RVC_E_INVALID, "out of host memory");
    hipError_t e = hipSetDevice(g_engine.device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete w;
      return set_error(RVC_E_DEVICE, "workspace stream: %s", hipGetErrorString(e));
    }
    w->batch.stream = w->stream;
    w->batch.own_stream = false;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    g_ws_all.push_back(w);
    t_ws = w;
    t_ws_generation = g_ws_generation;
  }
  *out = t_ws;
  return RVC_OK;
}

void release_workspaces() {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  for (SyncWorkspace* w : g_ws_all) {
    w->batch.release();
    w->staging.release();
    w->results.release();
    if (w->mem) hipFree(w->mem);
    if (w->stream) hipStreamDestroy(w->stream);
    delete w;
  }
  g_ws_all.clear();
  ++g_ws_generation;
}

FP32Policy& policy() { return default_policy(); }

}  // namespace

namespace rvc {
// One region through the device on the calling thread's workspace.
int region_sync_impl(const rvc_region* region, double* out, uint8_t* flags, const DevicePrep* prep) {
  if (!g_engine.ready) return set_error(RVC_E_STATE, "rvc_phmm_init has not run");
  int rc = validate_region(region);
  if (rc) return rc;
  if (!out) return set_error(RVC_E_INVALID, "null output");
  SyncWorkspace* w = nullptr;
  if ((rc = get_workspace(&w))) return rc;
  const FP32Policy pol = latency_policy(policy(), (uint64_t)region->n_haps * region->n_reads);
  if ((rc = build_host_pack(region, 1, pol, &w->pack, prep))) return rc;
  split_runs(w->pack);
  w->batch.graph_exec = nullptr;
  w->batch.graph = nullptr;
  const BatchLayout L = batch_layout(pack_sizes(w->pack));
  if ((rc = w->staging.reserve(L.cnt + 4 * kNumCounters))) return rc;
  stage_pack(w->pack, L, w->staging.ptr);
  stage_results(w->pack, L, w->staging.ptr);  // uploaded with the inputs: no begin kernel
  if ((rc = w->batch.upload_staged(w->pack, L, w->staging.ptr, w->stream, w->mem, w->bytes, true, true))) return rc;
  if (w->batch.own_mem) {  // the workspace was too small: keep the new allocation
    if (w->mem) hipFree(w->mem);
    w->mem = w->batch.mem;
    w->bytes = w->batch.bytes;
    w->batch.own_mem = false;
  }
  // the FP32 pass, one copy of results, flags and counters into pinned
  // memory, one wait; the FP64 launches only where pairs need them
  const size_t n = w->pack.n_pairs_slots;
  if ((rc = w->results.reserve(w->batch.fetch_bytes() + 64))) return rc;
  if ((rc = w->batch.run_and_fetch(w->stream, w->results.ptr, true))) return rc;
  const double* h_out = reinterpret_cast<const double*>(w->results.ptr);
  const uint8_t* h_flags = reinterpret_cast<const uint8_t*>(w->results.ptr + n * sizeof(double));
  unpermute_results(w->pack, h_out, h_flags, out, flags);
  return RVC_OK;
}
}  // namespace rvc

namespace {

// A region-level call after read preprocessing, ready to be computed: the
// region as the engine takes it, with the host-preprocessed arrays it points
// into (host path) or the device-preprocessing arguments (device path).
struct PreparedCall {
  uint32_t H = 0, N = 0;
  std::vector<int32_t> rlen;
  std::vector<uint8_t> q, gop, gcp;
  rvc_region reg{};
  DevicePrep prep{};
  bool on_device = false;
  const DevicePrep* prep_ptr() const { return on_device ? &prep : nullptr; }
};

int prepare_call(const rvc_raw_region* rg, int32_t min_quality_threshold, int32_t pcr_option, PreparedCall* pc) {
  const uint32_t H = rg->n_haps, N = rg->n_reads;
  pc->H = H;
  pc->N = N;
  if (H == 0 || N == 0 || !rg->hap_off || !rg->read_off || !rg->read_qual || !rg->read_mq)
    return set_error(RVC_E_INVALID, "empty or incomplete region");
  if (pcr_option < RVC_PCR_NONE || pcr_option > RVC_PCR_CONSERVATIVE)
    return set_error(RVC_E_INVALID, "unknown PCR indel model %d", pcr_option);
  pc->rlen.resize(N);
  for (uint32_t r = 0; r < N; ++r) {
    const int32_t len = rg->read_off[r + 1] - rg->read_off[r];
    if (len < 0) return set_error(RVC_E_INVALID, "read %u has negative length", r);
    pc->rlen[r] = len;
  }
  // read preprocessing (MQ cap, quality floor, GOP 45 / PCR model, GCP 10):
  // on the device, where the region carries the raw qualities, or here on
  // the calling thread. By default the device takes it when the PCR model
  // is on (its repeat search is the costly part); without it the host steps
  // are a few operations per base.
  const bool on_host = g_engine.host_prep == 1 || (g_engine.host_prep == 0 && pcr_option == RVC_PCR_NONE);
  rvc_region& reg = pc->reg;
  reg.n_haps = H;
  reg.n_reads = N;
  reg.hap_off = rg->hap_off;
  reg.hap_bases = rg->hap_bases;
  reg.read_off = rg->read_off;
  reg.read_bases = rg->read_bases;
  if (on_host) {
    const int32_t total = rg->read_off[N];
    pc->q.resize(total);
    pc->gop.resize(total);
    pc->gcp.resize(total);
    for (uint32_t r = 0; r < N; ++r) {
      const int32_t o = rg->read_off[r];
      preprocess_read(rg->read_bases + o, rg->read_qual + o, pc->rlen[r], rg->read_mq[r], min_quality_threshold,
                      pcr_option, pc->q.data() + o, pc->gop.data() + o, pc->gcp.data() + o);
    }
    reg.q = pc->q.data();
    reg.i = pc->gop.data();
    reg.d = pc->gop.data();
    reg.c = pc->gcp.data();
    pc->on_device = false;
  } else {
    reg.q = rg->read_qual;  // raw qualities; the device writes q, i, d, c
    reg.i = rg->read_qual;  // (not read)
    reg.d = rg->read_qual;
    reg.c = rg->read_qual;
    pc->prep.mq = rg->read_mq;
    pc->prep.min_q = min_quality_threshold;
    pc->prep.pcr = pcr_option;
    pc->on_device = true;
  }
  return RVC_OK;
}

// The reference's error path (ph/pairhmm_internal.cpp:585-588, 605-609):
// every entry NaN and every read kept.
int fail_call(uint32_t H, uint32_t N, double* out_hap_major, uint32_t* kept_reads, uint32_t* n_kept, int code) {
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (size_t x = 0; x < (size_t)H * N; ++x) out_hap_major[x] = nan;
  for (uint32_t r = 0; r < N; ++r) kept_reads[r] = r;
  *n_kept = N;
  return code;
}

// Region-level calls submitted and not yet waited for: what post-processing
// needs at wait (the reads' lengths), by ticket.
struct PendingCall {
  uint32_t H = 0, N = 0;
  std::vector<int32_t> rlen;
  std::vector<uint32_t> site_alleles;  // alleles per site, when the call marginalises
};
std::mutex g_calls_mu;
std::unordered_map<rvc_ticket, PendingCall> g_calls;

}  // namespace

extern "C" {

int rvc_phmm_init(const rvc_config* cfg) {
  return rvc::guarded([&]() -> int {
    int rc = engine_init(cfg);
    if (rc) return rc;
    if ((rc = noalloc_init())) {
      engine_shutdown();
      return rc;
    }
    if (cfg && cfg->server_threads) {
      rc = server_start();
      if (rc) {
        noalloc_shutdown();
        engine_shutdown();
        return rc;
      }
    }
    return RVC_OK;
  });
}

int rvc_phmm_shutdown(void) {
  return rvc::guarded([&]() -> int {
    server_stop();
    {
      std::lock_guard<std::mutex> lk(g_calls_mu);
      g_calls.clear();  // their tickets died with the serving loop
    }
    noalloc_shutdown();
    release_workspaces();
    sw_server_stop();
    sw_release_workspaces();
    return engine_shutdown();
  });
}

const char* rvc_phmm_last_error(void) { return last_error(); }

int rvc_phmm_region_sync(const rvc_region* region, double* out_read_major, uint8_t* out_flags) {
  return rvc::guarded([&]() -> int {
    return region_sync_impl(region, out_read_major, out_flags);
  });
}

int rvc_phmm_call_pairhmm(const rvc_raw_region* rg, int32_t min_quality_threshold, int32_t pcr_option,
                          double* out_hap_major, uint32_t* kept_reads, uint32_t* n_kept) {
  return rvc::guarded([&]() -> int {
    if (!rg || !out_hap_major || !kept_reads || !n_kept) return set_error(RVC_E_INVALID, "null argument");
    PreparedCall pc;
    int rc = prepare_call(rg, min_quality_threshold, pcr_option, &pc);
    if (rc) return fail_call(rg->n_haps, rg->n_reads, out_hap_major, kept_reads, n_kept, rc);
    // with the serving loop running, concurrent callers (Rovaca's worker
    // threads) are batched together; otherwise the region runs on its own
    std::vector<double> read_major((size_t)pc.H * pc.N);
    rc = server_running() ? served_region(&pc.reg, read_major.data(), nullptr, pc.prep_ptr())
                          : region_sync_impl(&pc.reg, read_major.data(), nullptr, pc.prep_ptr());
    if (rc) return fail_call(pc.H, pc.N, out_hap_major, kept_reads, n_kept, rc);
    *n_kept = (uint32_t)postprocess(pc.N, pc.H, read_major.data(), pc.rlen.data(), out_hap_major, kept_reads);
    return RVC_OK;
  });
}

int rvc_phmm_call_pairhmm_submit(const rvc_raw_region* rg, int32_t min_quality_threshold, int32_t pcr_option,
                                 const rvc_sites* sites, rvc_ticket* ticket) {
  return rvc::guarded([&]() -> int {
    if (!rg || !ticket) return set_error(RVC_E_INVALID, "null argument");
    if (!server_running()) return set_error(RVC_E_STATE, "serving loop is not running");
    auto pc = std::make_unique<PreparedCall>();
    int rc = prepare_call(rg, min_quality_threshold, pcr_option, pc.get());
    if (rc) return rc;
    // the region is packed (copied) by submit: the caller's buffers and the
    // host-preprocessed arrays are free once it returns
    rvc_ticket t = 0;
    if ((rc = submit_impl(&pc->reg, pc->prep_ptr(), sites, &t))) return rc;
    PendingCall p;
    p.H = pc->H;
    p.N = pc->N;
    if (sites) p.site_alleles.assign(sites->n_alleles, sites->n_alleles + sites->n_sites);
    p.rlen = std::move(pc->rlen);
    std::lock_guard<std::mutex> lk(g_calls_mu);
    g_calls[t] = std::move(p);
    *ticket = t;
    return RVC_OK;
  });
}

int rvc_phmm_call_pairhmm_wait(rvc_ticket ticket, double* out_hap_major, uint32_t* kept_reads, uint32_t* n_kept,
                               double* out_alleles, int64_t timeout_ms) {
  return rvc::guarded([&]() -> int {
    if (!out_hap_major || !kept_reads || !n_kept) return set_error(RVC_E_INVALID, "null argument");
    PendingCall p;
    {
      std::lock_guard<std::mutex> lk(g_calls_mu);
      auto it = g_calls.find(ticket);
      if (it == g_calls.end())
        return set_error(RVC_E_INVALID, "unknown region-call ticket %llu", (unsigned long long)ticket);
      p.H = it->second.H;
      p.N = it->second.N;
    }
    std::vector<double> read_major((size_t)p.H * p.N), marg;
    const int rc = wait_impl(ticket, read_major.data(), nullptr, timeout_ms, nullptr, &marg);
    if (rc == RVC_E_TIMEOUT) return rc;  // the ticket stays valid: wait again
    {
      std::lock_guard<std::mutex> lk(g_calls_mu);
      auto it = g_calls.find(ticket);
      if (it == g_calls.end())  // another thread waited for the same ticket
        return set_error(RVC_E_INVALID, "region-call ticket %llu already waited for", (unsigned long long)ticket);
      p = std::move(it->second);
      g_calls.erase(it);
    }
    size_t n_marg = 0;
    for (uint32_t al : p.site_alleles) n_marg += (size_t)al * p.N;
    if (rc) {
      if (out_alleles)
        for (size_t x = 0; x < n_marg; ++x) out_alleles[x] = std::numeric_limits<double>::quiet_NaN();
      return fail_call(p.H, p.N, out_hap_major, kept_reads, n_kept, rc);
    }
    *n_kept = (uint32_t)postprocess(p.N, p.H, read_major.data(), p.rlen.data(), out_hap_major, kept_reads);
    if (out_alleles && n_marg) {
      // the kept reads' columns, packed to the front like the matrix's
      const double nan = std::numeric_limits<double>::quiet_NaN();
      size_t row = 0;
      for (uint32_t al : p.site_alleles)
        for (uint32_t a = 0; a < al; ++a, ++row) {
          const double* src = marg.data() + row * p.N;
          double* dst = out_alleles + row * p.N;
          for (uint32_t k = 0; k < *n_kept; ++k) dst[k] = src[kept_reads[k]];
          for (uint32_t k = *n_kept; k < p.N; ++k) dst[k] = nan;
        }
    }
    return RVC_OK;
  });
}

int rvc_phmm_batch_create(const rvc_region* regions, uint32_t n_regions, const rvc_batch_opts* opts,
                          rvc_batch** out) {
  return rvc::guarded([&]() -> int {
    if (!out) return set_error(RVC_E_INVALID, "null output handle");
    *out = nullptr;
    if (!g_engine.ready) return set_error(RVC_E_STATE, "rvc_phmm_init has not run");
    if (!regions || n_regions == 0) return set_error(RVC_E_INVALID, "no regions");
    for (uint32_t r = 0; r < n_regions; ++r) {
      int rc = validate_region(&regions[r]);
      if (rc) return rc;
    }
    auto* b = new (std::nothrow) rvc_batch;
    if (!b) return set_error(RVC_E_INVALID, "out of host memory");
    std::unique_ptr<rvc_batch> guard(b);
    FP32Policy pol = policy();
    if (opts) {
      if (opts->fp32_fold == 1) pol.fold = true;
      if (opts->fp32_fold == 2) pol.fold = false;
      if (opts->rpl_round == 1) pol.rpl_steps_of_3 = false;
      if (opts->rpl_round == 2) pol.rpl_steps_of_3 = true;
      pol.waves_per_cu = opts->persistent_waves_per_cu;
      if (opts->occupancy) pol.force_occ = (uint8_t)std::min(4u, std::max(2u, opts->occupancy));
      if (opts->segment_width == 8 || opts->segment_width == 16 || opts->segment_width == 32 ||
          opts->segment_width == 64)
        pol.force_w = (uint8_t)opts->segment_width;
      if (opts->shape_fold == 1) pol.fold_small_bins = false;
      if (opts->prior_table == 1) pol.prior_table = true;
      if (opts->prior_table == 2) pol.prior_table = false;
      if (opts->hap_run) {
        pol.hap_run = std::min<uint32_t>(kMaxHapRun, opts->hap_run);
        pol.hap_run_fixed = true;
      }
    }
    int rc = build_host_pack(regions, n_regions, pol, &b->pack);
    if (rc) return rc;
    split_runs(b->pack);
    if (hipSetDevice(g_engine.device) != hipSuccess || hipStreamCreateWithFlags(&b->dev.stream, hipStreamNonBlocking) != hipSuccess)
      return set_error(RVC_E_DEVICE, "batch stream");
    b->dev.own_stream = true;
    if ((rc = b->dev.upload(b->pack, b->dev.stream, nullptr, 0))) return rc;
    if (hipStreamSynchronize(b->dev.stream) != hipSuccess) return set_error(RVC_E_DEVICE, "batch upload");
    std::vector<char>().swap(b->pack.staging);
    b->use_graph = opts && opts->use_graph;
    b->dev.untimed = opts && opts->untimed;
    *out = guard.release();
    return RVC_OK;
  });
}

int rvc_phmm_batch_run(rvc_batch* b) {
  return rvc::guarded([&]() -> int {
    if (!b) return set_error(RVC_E_INVALID, "null batch");
    int rc = b->dev.run(b->use_graph);
    if (rc == RVC_OK) b->ran = true;
    return rc;
  });
}

int rvc_phmm_batch_results(rvc_batch* b, double* out, uint8_t* flags) {
  return rvc::guarded([&]() -> int {
    if (!b || !out) return set_error(RVC_E_INVALID, "null argument");
    if (!b->ran) return set_error(RVC_E_STATE, "batch has not run");
    return b->dev.download(out, flags, b->dev.stream);
  });
}

int rvc_phmm_batch_info(const rvc_batch* b, rvc_batch_info* info) {
  return rvc::guarded([&]() -> int {
    if (!b || !info) return set_error(RVC_E_INVALID, "null argument");
    std::memset(info, 0, sizeof *info);
    info->n_pairs = b->pack.n_pairs;
    info->n_cells = b->pack.n_cells;
    info->n_rescued = b->dev.n_rescued;
    info->device_bytes = b->dev.bytes;
    info->fp32_ms = b->dev.fp32_ms;
    info->fp64_ms = b->dev.fp64_ms;
    info->total_ms = b->dev.total_ms;
    info->n_fp32_launches = b->dev.n_fp32_launches;
    info->n_fp64_launches = b->dev.n_fp64_launches;
    return RVC_OK;
  });
}

int rvc_phmm_batch_kernels(const rvc_batch* b, rvc_kernel_time* out, uint32_t max, uint32_t* n) {
  return rvc::guarded([&]() -> int {
    if (!b || !n || (max && !out)) return set_error(RVC_E_INVALID, "null argument");
    const auto& L = b->dev.launches;
    *n = (uint32_t)L.size();
    for (uint32_t x = 0; x < max && x < L.size(); ++x) out[x] = L[x];
    return RVC_OK;
  });
}

int rvc_phmm_batch_destroy(rvc_batch* b) {
  return rvc::guarded([&]() -> int {
    if (!b) return set_error(RVC_E_INVALID, "null batch");
    delete b;
    return RVC_OK;
  });
}

int rvc_phmm_preprocess_read(const uint8_t* bases, const uint8_t* qual, int32_t len, uint8_t mq,
                             int32_t min_quality_threshold, int32_t pcr_option, uint8_t* q, uint8_t* gop,
                             uint8_t* gcp) {
  return rvc::guarded([&]() -> int {
    if (len < 0 || (len > 0 && (!bases || !qual || !q || !gop || !gcp)))
      return set_error(RVC_E_INVALID, "bad read arguments");
    if (pcr_option < RVC_PCR_NONE || pcr_option > RVC_PCR_CONSERVATIVE)
      return set_error(RVC_E_INVALID, "unknown PCR indel model %d", pcr_option);
    preprocess_read(bases, qual, len, mq, min_quality_threshold, pcr_option, q, gop, gcp);
    return RVC_OK;
  });
}

int rvc_phmm_postprocess(uint32_t n_reads, uint32_t n_haps, double* read_major, const int32_t* read_len,
                         double* out_hap_major, uint32_t* kept_reads) {
  return rvc::guarded([&]() -> int {
    if (n_reads == 0 || n_haps == 0 || !read_major || !read_len || !out_hap_major || !kept_reads)
      return -set_error(RVC_E_INVALID, "bad post-processing arguments");
    return postprocess(n_reads, n_haps, read_major, read_len, out_hap_major, kept_reads);
  });
}

int rvc_phmm_tables(float* ph2pr_f32, float* mm_f32, double* ph2pr_f64, double* mm_f64) {
  return rvc::guarded([&]() -> int {
    static HostTables* t = [] {
      auto* p = new HostTables;
      build_tables(p);
      return p;
    }();
    if (ph2pr_f32) std::memcpy(ph2pr_f32, t->ph2pr_f, sizeof t->ph2pr_f);
    if (mm_f32) std::memcpy(mm_f32, t->mm_f, sizeof t->mm_f);
    if (ph2pr_f64) std::memcpy(ph2pr_f64, t->ph2pr_d, sizeof t->ph2pr_d);
    if (mm_f64) std::memcpy(mm_f64, t->mm_d, sizeof t->mm_d);
    return RVC_OK;
  });
}

}  // extern "C"
