// Internal types shared by the device kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rvc {

// Limits of the reference scheduler (ph/roDV_schedule.cpp:70-73).
constexpr int kMaxHaps = 4096;
constexpr int kMaxHapLen = 1536;
constexpr int kMaxReadLen = 768;
// Largest rows per lane an FP32 kernel shape may have (8-lane shapes go to 19).
constexpr int kMaxRpl = 20;

// Every packed sequence starts on a 32-base boundary, so its 2-bit words
// (16 bases per uint32) and N-mask words (32 bases per uint32) are aligned.
constexpr uint32_t kSeqAlign = 32;
// LDS bytes for one staged haplotype: 64 leading pad bytes (negative columns
// of the anti-diagonal sweep), up to 1536 bases, 64 + 16 trailing pad bytes
// (columns past the end, read by the last steps and one step of prefetch).
constexpr int kHapBuf = kMaxHapLen + 2 * 64 + 16;

// Phred -> probability tables on the device (built on the host exactly as
// ph/roDV_common/context.h:44-60,92-146 builds them).
struct DevTables {
  const float* ph2pr_f;   // 128
  const float* mm_f;      // 128*129/2, index max*(max+1)/2 + min
  const double* ph2pr_d;
  const double* mm_d;
  const uint8_t* gop_cap;  // PCR indel model caps, [model][repeat length 0..20]
};

// Packed sequences of one batch, all in device memory.
struct DevSeqs {
  const uint32_t* hap_pack;   // 2 bits per base, 16 bases per word, A=0 C=1 G=2 T=3
  const uint32_t* hap_nmask;  // 1 bit per base, set for N (wildcard)
  const uint32_t* hap_off;    // first base of each haplotype (multiple of kSeqAlign)
  const uint32_t* hap_len;
  const uint32_t* read_pack;
  const uint32_t* read_nmask;
  const uint8_t* rq;          // per read base: base, insertion, deletion, gap-continuation quality
  const uint8_t* ri;
  const uint8_t* rd;
  const uint8_t* rc;
  const uint32_t* read_off;
  const uint32_t* read_len;
};

// FP32 work unit: a run of consecutive haplotypes hap .. hap+nh-1 of one
// region against up to 64/W consecutive reads; the reads' row parameters are
// set up once per tile and reused for every haplotype of the run. The result
// of segment s and haplotype hap+h goes to out[out0 + s * stride + h].
struct Tile {
  uint32_t hap;
  uint32_t read0;
  uint32_t out0;
  uint32_t meta;  // bits 0-3: segments in use; bits 4-16: stride; bits 17-23: nh
};
constexpr uint32_t kMaxHapRun = 64;
// x: the region's scale exponent X (row -1 scale 2^X, sum brought back by
// 2^(120 - X)); 120 in the plain form
__host__ __device__ inline uint32_t tile_meta(uint32_t nseg, uint32_t stride, uint32_t nh, uint32_t x) {
  return nseg | (stride << 4) | (nh << 17) | (x << 24);
}
__host__ __device__ inline uint32_t tile_nseg(uint32_t meta) { return meta & 0xfu; }
__host__ __device__ inline uint32_t tile_stride(uint32_t meta) { return (meta >> 4) & 0x1fffu; }
__host__ __device__ inline uint32_t tile_nh(uint32_t meta) { return (meta >> 17) & 0x7fu; }
__host__ __device__ inline uint32_t tile_x(uint32_t meta) { return meta >> 24; }
__host__ __device__ inline uint32_t tile_with_nh(uint32_t meta, uint32_t nh) {
  return (meta & ~(0x7fu << 17)) | (nh << 17);
}

// One pair queued for the FP64 recomputation.
struct RescueEntry {
  uint32_t hap;
  uint32_t read;
  uint32_t out;
  uint32_t x;  // the pair's scale exponent, as in its tile
};

// FP64 bins: (segment width, rows per lane) chosen from the read length so
// that no lane holds more than 12 double rows. Few bins on purpose: the FP64
// pass is one launch per bin that can occur, whether or not any pair of the
// batch needs it. Bins 0-3 take reads of fewer than W * RPL bases, so their
// last row is always a padding row (the prior-table kernel relies on it);
// bin 4 takes reads up to the 768-base limit.
constexpr int kNumFp64Bins = 5;
__host__ __device__ inline int fp64_bin(uint32_t R) {
  if (R < 64) return 0;    // W=16, RPL 4
  if (R < 128) return 1;   // W=16, RPL 8
  if (R < 160) return 2;   // W=32, RPL 5 (150-base reads)
  if (R < 256) return 3;   // W=32, RPL 8 (250-base reads)
  return 4;                // W=64, RPL 12
}
// Rescue lists: per FP64 bin, one for pairs whose haplotype has N (list b),
// one for the others of scaled-form regions (list kNumFp64Bins + b), which
// the prior-table FP64 kernel can take, and one for the pairs of plain-form
// regions (list 2 kNumFp64Bins + b).
constexpr int kNumFp64Lists = 3 * kNumFp64Bins;
inline void fp64_bin_shape(int bin, int* W, int* RPL) {
  static const int kW[kNumFp64Bins] = {16, 16, 32, 32, 64}, kR[kNumFp64Bins] = {4, 8, 5, 8, 12};
  *W = kW[bin];
  *RPL = kR[bin];
}

struct Fp32Args {
  DevSeqs s;
  DevTables t;
  const Tile* tiles;
  uint32_t n_tiles;
  uint32_t* head;            // work-queue counter of this launch
  double* out;
  uint8_t* flags;
  RescueEntry* rescue;       // kNumFp64Lists lists, list l starts at rescue_base[l]
  uint32_t* rescue_count;    // kNumFp64Lists counters
  const uint32_t* rescue_base;
  uint32_t rescue_list0;     // this launch's pairs go to list rescue_list0 + fp64_bin(R)
};

struct Fp64Args {
  DevSeqs s;
  DevTables t;
  const RescueEntry* list;   // this bin's list
  const uint32_t* count;     // this bin's counter
  uint32_t* head;            // this launch's work-queue counter
  double* out;
  uint8_t* flags;
};

// Device read preprocessing (for regions handed over with raw qualities):
// per read, prep[r] = 1 << 31 | pcr << 16 | min_q << 8 | mq, or 0 when the
// read's q/i/d/c are already final.
struct PrepArgs {
  uint8_t* rq;
  uint8_t* ri;
  uint8_t* rd;
  uint8_t* rc;
  const uint8_t* raw;        // read bases as given (ASCII), at the qualities' offsets
  const uint32_t* prep;
  const uint32_t* read_off;
  const uint32_t* read_len;
  const uint8_t* gop_cap;
  uint32_t n_reads;
};
constexpr uint32_t kPrepOn = 1u << 31;

// Haplotype -> allele marginalisation of one region's read-major results on
// the device (allele_likelihoods.hpp:189-217, 701-732, on the matrix that the
// region call normalises, ph/pairhmm_internal.cpp:248-258): per read and site,
// the largest normalised likelihood among the haplotypes mapped to each
// allele, -inf where no haplotype maps to it.
struct MargRegion {
  uint64_t pair0;     // first result slot of the region (device read order)
  uint32_t n_haps, n_reads;
  uint32_t read0;     // the region's entries of read_perm (device slot -> input read)
  uint32_t n_sites;
  uint32_t site0;     // first site of the region in n_alleles / allele0 / map0
  uint32_t pad;
};
struct MargArgs {
  const double* results;     // the batch's read-major results
  const MargRegion* regions;
  const uint32_t* read_perm;
  const uint32_t* n_alleles;  // per site
  const uint64_t* allele0;    // per site: first output row (rows of n_reads values, input read order)
  const uint64_t* map0;       // per site: first entry of its haplotype -> allele map in 'map'
  const int16_t* map;         // allele of each haplotype at the site, -1 = none
  double* out;
  uint32_t n_regions;
};

// Kernel launchers (phmm_kernels.hip).
hipError_t launch_marginalize(const MargArgs& a, hipStream_t stream);
hipError_t launch_prep(const PrepArgs& a, hipStream_t stream);
bool fp32_supported(int W, int RPL);
// scaled: the 5-operation scaled form (else plain); pt: the prior-table kernel
// (scaled form only).
hipError_t launch_fp32(int W, int RPL, bool scaled, int occ, bool pt, int grid, hipStream_t stream,
                       const Fp32Args& a);
// FP64 rescue of one list's bin; pt = the prior-table kernel (pairs without
// N in the haplotype, scaled form only), scaled = the scaled form.
hipError_t launch_fp64(int bin, bool pt, bool scaled, int grid, hipStream_t stream, const Fp64Args& a);
hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t stream);
// n bytes from src to dst by a kernel on stream; each side is device memory
// or pinned host memory (mapped into the device's address space). Unlike a
// DMA copy it waits on no other stream's copies: the copy engines serve
// their queues in order, so a DMA copy queued behind another stream's copy
// that waits on that stream's kernels would wait for those kernels too.
hipError_t launch_copy(void* dst, const void* src, uint64_t n, hipStream_t stream);
// out[0 .. n) = NaN and the n flag bytes behind them (rounded up to 8) = 0.
hipError_t launch_init_results(double* out, uint64_t n, hipStream_t stream);
// The start of a launch sequence: p[0 .. n) = 0 and, for n_out > 0, the
// results as launch_init_results sets them, in one kernel.
hipError_t launch_begin(uint32_t* p, uint32_t n, double* out, uint64_t n_out, hipStream_t stream);

}  // namespace rvc
