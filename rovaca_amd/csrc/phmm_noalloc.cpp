// rvc_phmm_region_sync_noalloc: the one-region call for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc). It runs the same packing and kernels as
// rvc_phmm_region_sync, with the region's device image (packed sequences,
// tiles, results, rescue lists, counters) in a fixed device array instead of
// a workspace that grows with hipMalloc.
//
// The array lives for the whole process on the device the engine uses. Its
// device address is looked up once at rvc_phmm_init (hipGetSymbolAddress);
// host code only ever uses that address, never the symbol itself. Calls are
// serialised by a mutex (they share the one array); every call uploads its
// starting results and zero counters with its inputs (no hipMemsetAsync), so
// nothing of an earlier call's contents is read.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "../../include/rovaca_phmm.h"
#include "phmm_engine.h"
#include "phmm_server.h"

namespace rvc {

// Bytes of the device image the no-allocation call can hold: a region of
// 2000 reads of 250 bases x 128 haplotypes of 1000 bases needs about 10 MiB.
constexpr size_t kNoAllocScratchBytes = 32ull << 20;

// The scratch array. Host code never passes 'scratch' itself to a kernel or
// a copy (on the host it is only a stand-in for the symbol): it uses the
// device address hipGetSymbolAddress returns (g_na.dev).
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {

struct NoAllocState {
  std::mutex mu;
  void* dev = nullptr;          // device address of 'scratch' (per engine device)
  int device = -1;
  hipStream_t stream = nullptr;
  HostPack pack;
  // pinned host buffer of the same size as the device array, allocated at
  // set-up: the staged inputs go up from it and the results come back into
  // it (stream order keeps the two apart), so the copies are asynchronous
  char* host = nullptr;
  DeviceBatch batch;
};
NoAllocState g_na;

}  // namespace

// Set-up at rvc_phmm_init, with the engine's device current: the array's
// device address and the stream of the no-allocation calls.
int noalloc_init() {
  std::lock_guard<std::mutex> lk(g_na.mu);
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_error(RVC_E_DEVICE, "device scratch address: %s", hipGetErrorString(e));
  if (!g_na.stream && hipStreamCreateWithFlags(&g_na.stream, hipStreamNonBlocking) != hipSuccess) {
    g_na.stream = nullptr;
    return set_error(RVC_E_DEVICE, "stream of the no-allocation call");
  }
  if (!g_na.host && hipHostMalloc((void**)&g_na.host, kNoAllocScratchBytes, hipHostMallocDefault) != hipSuccess) {
    g_na.host = nullptr;
    return set_error(RVC_E_DEVICE, "pinned buffer of the no-allocation call");
  }
  g_na.dev = p;
  g_na.device = g_engine.device;
  g_na.batch.stream = g_na.stream;
  g_na.batch.own_stream = false;
  return RVC_OK;
}

void noalloc_shutdown() {
  std::lock_guard<std::mutex> lk(g_na.mu);
  if (g_na.stream) {
    hipStreamSynchronize(g_na.stream);
    hipStreamDestroy(g_na.stream);
  }
  g_na.stream = nullptr;
  if (g_na.host) hipHostFree(g_na.host);
  g_na.host = nullptr;
  g_na.batch.stream = nullptr;
  g_na.dev = nullptr;
  g_na.device = -1;
}

int region_sync_noalloc(const rvc_region* region, double* out, uint8_t* flags) {
  if (!g_engine.ready) return set_error(RVC_E_STATE, "rvc_phmm_init has not run");
  int rc = validate_region(region);
  if (rc) return rc;
  if (!out) return set_error(RVC_E_INVALID, "null output");
  std::lock_guard<std::mutex> lk(g_na.mu);
  if (!g_na.dev || g_na.device != g_engine.device) return set_error(RVC_E_STATE, "device scratch not set up");
  if (hipSetDevice(g_engine.device) != hipSuccess) return set_error(RVC_E_DEVICE, "hipSetDevice");
  const FP32Policy pol = latency_policy(default_policy(), (uint64_t)region->n_haps * region->n_reads);
  if ((rc = build_host_pack(region, 1, pol, &g_na.pack))) return rc;
  split_runs(g_na.pack);
  const BatchLayout L = batch_layout(pack_sizes(g_na.pack));
  if (L.bytes > kNoAllocScratchBytes)
    return set_error(RVC_E_LIMIT, "region needs %zu bytes of device scratch; the no-allocation call holds %zu",
                     L.bytes, kNoAllocScratchBytes);
  // (the staged image and the fetched results both fit: each is at most L.bytes)
  stage_pack(g_na.pack, L, g_na.host);
  stage_results(g_na.pack, L, g_na.host);  // uploaded with the inputs: no begin kernel
  DeviceBatch& b = g_na.batch;
  // the image goes into the scratch array (reuse_mem covers it: no hipMalloc)
  if ((rc = b.upload_staged(g_na.pack, L, g_na.host, g_na.stream, g_na.dev, kNoAllocScratchBytes, true, true)))
    return rc;
  if (b.own_mem) return set_error(RVC_E_INVALID, "device image did not fit the scratch array");
  const size_t n = g_na.pack.n_pairs_slots;
  // the FP32 pass, one copy back of results, flags and counters, one wait;
  // the FP64 launches only where pairs need them
  if ((rc = b.run_and_fetch(g_na.stream, g_na.host, true))) return rc;
  const double* h_out = reinterpret_cast<const double*>(g_na.host);
  const uint8_t* h_flags = reinterpret_cast<const uint8_t*>(g_na.host + n * sizeof(double));
  unpermute_results(g_na.pack, h_out, h_flags, out, flags);
  b.mem = nullptr;  // the batch never owns the scratch array
  return RVC_OK;
}

}  // namespace rvc

extern "C" int rvc_phmm_region_sync_noalloc(const rvc_region* region, double* out_read_major, uint8_t* out_flags) {
  return rvc::guarded([&]() -> int { return rvc::region_sync_noalloc(region, out_read_major, out_flags); });
}
