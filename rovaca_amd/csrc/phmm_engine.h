// Host engine internals shared by the C ABI, the region adapter and the
// serving loop.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/rovaca_phmm.h"
#include "phmm_internal.h"
#include "phmm_tables.h"

namespace rvc {

int set_error(int code, const char* fmt, ...);
const char* last_error();

// Runs the body of a C-ABI entry point: a C++ exception (host memory) must
// not cross into the caller's C or C++ code, so it becomes an error code.
template <class F>
int guarded(F&& body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return set_error(RVC_E_INVALID, "out of host memory");
  } catch (...) {
    return set_error(RVC_E_INVALID, "unexpected host error");
  }
}

struct Engine {
  std::mutex mu;
  bool ready = false;
  int device = 0;
  int n_cu = 0;
  std::unique_ptr<HostTables> host_tables;
  void* dev_tables_mem = nullptr;
  DevTables tables{};
  uint64_t hbm_budget = 0;
  uint32_t max_batch_regions = 0;
  uint32_t host_prep = 0;   // call_pairhmm preprocessing: 0 auto, 1 host, 2 device
  uint32_t sw_min_launch = 0;  // aligner: fewest pairs a sweep shape launches with (0 auto)
  uint64_t sw_bt_budget = 12ull << 30;  // aligner: backtrack bytes per device pass
  uint32_t sw_wide = 0;  // aligner: wide sweep shapes (0 for small passes, 1 never, 2 always)
  std::string tune_source;  // where the kernel-shape table came from ("" = built-in defaults)
};
extern Engine g_engine;

int engine_init(const rvc_config* cfg);
int engine_shutdown();
int validate_region(const rvc_region* r);

// FP32 kernel shapes: W in {16, 32, 64, 8} x RPL in 1..kMaxRpl (W = 64 up to
// 13, W = 16 and 32 up to 16, W = 8 up to 19). Tile bins are [class][shape],
// the class fixed per region: kClassN for haplotypes with N (compare kernel,
// scaled form), kClassTable for those without (prior-table kernel, scaled
// form), kClassPlain for every haplotype of a region whose gap qualities rule
// the scaled form out (compare kernel, plain form). A region's results thus
// never depend on the other regions of its batch. The FP64 rescue lists are
// [class][FP64 bin] in the same way.
constexpr int kClassN = 0, kClassTable = 1, kClassPlain = 2, kNumClasses = 3;
constexpr int kFp32Variants = 4 * kMaxRpl;
constexpr int kFp32Bins = kNumClasses * kFp32Variants;
static_assert(kNumFp64Lists == kNumClasses * kNumFp64Bins, "one FP64 list per class and bin");
inline int bin_class(int v) { return v / kFp32Variants; }
// Scale exponent of the plain form: row -1 at 2^120, nothing to bring back.
constexpr int kPlainExponent = 120;
// Row -1 scale exponent X of a region (2^X in place of 2^120), from the
// range of its gap-open qualities and its smallest gap-continuation quality,
// or kPlainExponent where they rule the scaled form out (or the policy does).
int scale_exponent(int gop_lo, int gop_hi, int gcp_lo, bool allow_scaled);
int fp32_variant_index(int W, int RPL);
void fp32_variant_shape(int v, int* W, int* RPL);
int fp32_blocks_per_cu(int W, int RPL, bool xf, int occ, bool pt);
int fp64_blocks_per_cu(int bin, bool pt, bool scaled);

// Chooses the (W, RPL) kernel for a tile whose longest read has R bases.
// rpl16_override[x] (for reads up to 255 bases, indexed by the minimal
// 16-lane row count x of R + 1 rows) may name another shape as
// (W << 8) | RPL, where RPL = 0 means the fewest rows of that width; it is
// filled by the tile tuner.
struct FP32Policy {
  uint16_t rpl16_override[17] = {};
  // waves per SIMD the 16-lane kernel of each RPL is compiled for (2, 3, 4);
  // rows per lane past 16 (W = 8 only) run at 2
  uint8_t occ16[kMaxRpl + 1] = {4, 4, 4, 4, 4, 4, 4, 4, 3, 3, 3, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};
  bool rpl_steps_of_3 = true;  // 16-lane shapes only in RPL 1, 2, 4, 7, 10, 13, 16
  bool fold = true;            // scaled forms where the gap qualities allow
  uint32_t waves_per_cu = 0;   // persistent grid: 0 = occupancy limit
  // the same for the prior-table kernels (W = 16 and W = 32)
  uint8_t occpt16[kMaxRpl + 1] = {4, 4, 4, 4, 4, 4, 4, 4, 3, 3, 3, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};
  uint8_t occpt32[kMaxRpl + 1] = {4, 4, 4, 4, 4, 4, 4, 4, 4, 3, 3, 3, 3, 2, 2, 2, 2, 2, 2, 2, 2};
  uint8_t occpt8[kMaxRpl + 1] = {4, 4, 4, 4, 4, 4, 4, 4, 4, 3, 3, 3, 3, 4, 4, 4, 4, 2, 2, 2, 2};
  bool prior_table = true;     // use the prior-table kernel where the haplotype has no N
  uint8_t force_occ = 0;       // batch option: 2/3/4 for every 16-lane kernel
  uint8_t force_w = 0;         // batch option: 16/32/64 lanes per read for reads <= 256
  uint32_t hap_run = 16;       // most haplotypes per tile (consecutive, of one region)
  bool hap_run_fixed = false;  // batch option: exactly hap_run per tile, whatever the batch size
  bool fold_small_bins = true; // sweep the tiles of small shape bins with a larger shape (fold_bins)
  // cmp: for the compare-and-select kernel (haplotypes with N), which has no
  // 8-lane shape past 16 rows per lane (it would spill): those tiles keep
  // the 16-lane shape
  void pick(uint32_t R, int* W, int* RPL, bool cmp = false) const;
  int occupancy(int W, int RPL, bool pt) const {
    if (RPL > 16) return 2;
    if (force_occ && (W == 16 || (pt && (W == 32 || W == 8)))) return force_occ;
    if (pt) return W == 16 ? occpt16[RPL] : W == 32 ? occpt32[RPL] : W == 8 ? occpt8[RPL] : 2;
    return W == 16 ? occ16[RPL] : 2;
  }
};

// The process-wide policy (built-in defaults, replaced by a tuned table at init).
FP32Policy& default_policy();

// The policy for a region that runs alone: one too small to give every SIMD
// a wave of 16-lane tiles (up to 2048 pairs) finishes sooner with wider
// segments, fewer rows per lane and so a shorter chain per step (measured on
// 8 x 32 pairs of 150 bases: 59 us at W = 16, 43 us at W = 64).
FP32Policy latency_policy(const FP32Policy& base, uint64_t pairs);
// Reads a table written by rovaca_amd/tune.py; returns RVC_OK or an error code.
int load_tune_table(const char* path, FP32Policy* pol);

// Regions handed over with raw base qualities: the device preprocesses their
// reads (MQ cap, min-quality floor, GOP 45 / PCR model, GCP 10) before the
// sweep. The region's q array holds the raw qualities; i, d, c are unused.
struct DevicePrep {
  const uint8_t* mq;  // per read, in the region's read order
  int32_t min_q;
  int32_t pcr;
};

struct RegionLayout {
  uint32_t n_haps = 0, n_reads = 0;
  uint32_t hap0 = 0, read0 = 0;  // first global haplotype / read index
  uint64_t pair0 = 0;            // first result slot
};

// A batch packed on the host in device format (see DevSeqs).
struct HostPack {
  uint32_t n_regions = 0;
  std::vector<RegionLayout> region;
  std::vector<uint32_t> hap_pack, hap_nmask, hap_off, hap_len;
  std::vector<uint32_t> read_pack, read_nmask, read_off, read_len;
  std::vector<uint8_t> rq, ri, rd, rc;
  std::vector<uint32_t> read_perm;  // device read slot -> input read index within its region
  std::vector<Tile> tiles;          // grouped by FP32 bin
  std::vector<uint32_t> bin_start;  // kFp32Bins + 1
  std::vector<uint32_t> fp64_cap;   // pairs per rescue list (its capacity)
  std::vector<uint32_t> fp64_base;  // kNumFp64Lists + 1
  uint64_t n_pairs_slots = 0;       // sum of n_haps * n_reads
  uint64_t n_pairs = 0;             // pairs with two non-empty sides
  uint64_t n_cells = 0;
  uint32_t waves_per_cu = 0;
  uint8_t occ[kFp32Bins] = {};            // occupancy target per FP32 bin
  uint64_t variant_cells[kFp32Bins] = {};  // cells computed by each FP32 bin
  bool prior_table = true;                 // launch table bins with the prior-table kernel
  uint32_t hap_run = 1;                    // longest haplotype run a tile may keep
  bool hap_run_fixed = false;              // keep runs of hap_run whatever the batch size
  bool no_fold = false;                    // every tile keeps its shape (no fold_bins)
  uint64_t bin_units[kFp32Bins] = {};      // (read group, haplotype) pairs per FP32 bin
  bool any_prep = false;                   // some reads are preprocessed on the device
  std::vector<uint8_t> read_raw;           // their bases as given, at the qualities' offsets
  std::vector<uint32_t> read_prep;         // per read: PrepArgs::prep word
  std::vector<char> staging;               // host image of the device inputs (upload)
  std::vector<double> nan_scratch;
  std::vector<uint8_t> flag_scratch;
  // Back to the state of a new pack, keeping the vectors' storage (packing
  // one region after another on a thread then allocates nothing new).
  void reset();
};

// Element counts of a batch's device arrays.
struct PackSizes {
  uint64_t hap_words = 0, hap_nwords = 0, n_haps = 0;
  uint64_t read_words = 0, read_nwords = 0, qual_bytes = 0, n_reads = 0;
  uint64_t n_tiles = 0, n_pairs_slots = 0, n_rescue = 0;
  uint64_t raw_bytes = 0;  // raw read bases (device preprocessing), 0 if none
};
PackSizes pack_sizes(const HostPack& P);

// Byte offsets of a batch in its device allocation. The inputs come first,
// so one host-to-device copy of [0, in_bytes) uploads a staged batch.
constexpr uint32_t kNumCounters = kFp32Bins + 2 * kNumFp64Lists;
struct BatchLayout {
  size_t hpack = 0, hn = 0, hoff = 0, hlen = 0, rpack = 0, rn = 0, rq = 0, ri = 0, rd = 0, rc = 0;
  size_t roff = 0, rlen = 0, tiles = 0, rbase = 0, raw = 0, prep = 0;
  size_t in_bytes = 0;
  size_t out = 0, flags = 0, resc = 0, cnt = 0;
  size_t bytes = 0;
};
BatchLayout batch_layout(const PackSizes& z);
// Writes the input section of P at dst (in_bytes of L).
void stage_pack(const HostPack& P, const BatchLayout& L, char* dst);
// Writes the starting results (NaN, flags 0) and zero counters behind the
// inputs at dst, so that one upload of [0, out_end) replaces the begin
// kernel; returns out_end.
size_t stage_results(const HostPack& P, const BatchLayout& L, char* dst);

// Grow-only pinned host buffer.
// Pinned host buffer that grows by doubling. Growing keeps the old block
// until release(): freeing pinned memory waits for the whole device, which
// would stall the launches of other streams.
struct PinnedBuf {
  char* ptr = nullptr;
  size_t cap = 0;
  std::vector<char*> retired;
  ~PinnedBuf();
  int reserve(size_t bytes);  // keeps no contents
  void release();
};

// Merges packs built by build_host_pack straight into a staged input image
// (pinned, in device layout), as if the regions had been packed together.
// 'meta' receives the merged bookkeeping (regions, read order, tile bins,
// rescue bins, counts); its sequence arrays stay empty.
int stage_merged(const HostPack* const* packs, uint32_t n, HostPack* meta, BatchLayout* L, PinnedBuf* buf);

// Packs regions; tiles hold runs of up to pol.hap_run haplotypes, in
// largest-first order per bin. Call split_runs before uploading such a pack
// on its own (stage_merged splits on the fly).
int build_host_pack(const rvc_region* regions, uint32_t n_regions, const FP32Policy& pol, HostPack* hp,
                    const DevicePrep* prep = nullptr);
// Persistent grid of FP32 bin v's kernel for this batch.
uint32_t fp32_grid(const HostPack& P, int v);
// Haplotypes per tile for FP32 bin v: long runs amortise a tile's set-up,
// but every persistent wave needs about a dozen tiles so that all end together.
uint32_t run_for_bin(const HostPack& P, int v, uint64_t units);
// Where each FP32 bin's tiles go (to[v]): a bin of few cells into the next
// larger non-empty shape of its class and width, if that has at most a third
// more rows; moves the bins' cell and unit counts along.
void fold_bins(HostPack& P, int (&to)[kFp32Bins]);
// Splits the pack's tiles into runs of run_for_bin haplotypes, largest first
// (after fold_bins).
void split_runs(HostPack& P);

// Copies a batch's results (device order, as downloaded) to input read order.
void unpermute_results(const HostPack& P, const double* dev_out, const uint8_t* dev_flags, double* host_out,
                       uint8_t* host_flags);

// enqueue() argument: queue the FP32 pass only.
inline const uint32_t* const kNoFp64 = reinterpret_cast<const uint32_t*>(uintptr_t(1));

struct DeviceBatch {
  const HostPack* pack = nullptr;
  void* mem = nullptr;
  bool own_mem = false;
  size_t bytes = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  DevSeqs seqs{};
  const Tile* tiles = nullptr;
  double* out = nullptr;
  uint8_t* flags = nullptr;
  RescueEntry* rescue = nullptr;
  const uint32_t* rescue_base = nullptr;
  uint32_t* counters = nullptr;
  uint32_t n_counters_ = 0;
  const uint8_t* read_raw = nullptr;
  const uint32_t* read_prep = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  // per-launch events of timed runs: launch x is bracketed by lev[x], lev[x+1]
  std::vector<hipEvent_t> lev;
  std::vector<rvc_kernel_time> launches;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  double fp32_ms = 0, fp64_ms = 0, total_ms = 0;
  uint64_t n_rescued = 0;
  uint32_t n_fp32_launches = 0, n_fp64_launches = 0;
  bool init_pending = false;  // results not yet set to NaN since the last upload
  bool counters_clear = false;  // the upload brought zeroed counters: no begin kernel
  bool untimed = false;       // eager runs record no per-launch events

  ~DeviceBatch();
  void release();
  BatchLayout layout;
  // Lays the batch out in one device allocation (reuse_mem if it is big
  // enough) and queues the host-to-device copy on s.
  int upload(const HostPack& P, hipStream_t s, void* reuse_mem, size_t reuse_bytes);
  // The same for inputs already staged in device layout at 'host' (pinned
  // memory makes the copy asynchronous); 'meta' describes the batch. Where
  // 'host' is pinned (hipHostMalloc) and the image small, src_pinned lets a
  // kernel on s read it instead of a copy engine: the sweep that follows
  // then starts without waiting for another engine's signal.
  int upload_staged(const HostPack& meta, const BatchLayout& L, const void* host, hipStream_t s, void* reuse_mem,
                    size_t reuse_bytes, bool with_results = false, bool src_pinned = false);
  // Queues counter reset, FP32 launches and FP64 rescue launches on s: one
  // FP64 launch per list that can hold pairs (rescued == nullptr), none
  // (rescued == kNoFp64), or one per list whose count in rescued[] is not 0.
  int enqueue(hipStream_t s, bool timed, const uint32_t* rescued = nullptr);
  // The FP64 launches alone (as enqueue queues them after the FP32 pass).
  int enqueue_fp64(hipStream_t s, bool timed, const uint32_t* rescued);
  // Bytes of results, flags and counters as they sit on the device.
  size_t fetch_bytes() const;
  // The whole launch sequence for a caller that waits for the results: the
  // FP32 pass, one copy of results, flags and counters into host (fetch_bytes
  // of room), a wait; then, only if the counters show pairs below 1e-28, the
  // FP64 launches of those lists and a second copy of results and flags.
  // Returns when host holds the results. host_pinned: host is pinned
  // (hipHostMalloc), and a kernel on s writes the copies there (launch_copy);
  // otherwise they go through a copy engine.
  int run_and_fetch(hipStream_t s, char* host, bool host_pinned);
  int ensure_events();
  // Runs the launch sequence on 'stream' (optionally as a hipGraph) and waits.
  int run(bool use_graph);
  // Copies results back in input read order.
  int download(double* host_out, uint8_t* host_flags, hipStream_t s);
};

}  // namespace rvc
