// Device-side types of the batched Smith-Waterman (read against haplotype).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/rovaca_sw.h"

namespace rvc {

constexpr int kSwMaxLen = 2048;       // the reference aligner refuses lengths >= this (sw_avx.c:18, sw_core_template.h:358-361)
constexpr int kSwMaxRead = 2048;      // rows of the largest sweep shape (reads up to 2047 bases)
constexpr int kSwBaseShapes = 13;     // the shapes sw_shape() picks from
constexpr int kSwShapes = 18;         // (W, RPL): the base shapes, then the wide ones
constexpr int kSwLongShape = 12;      // 64 x 32 rows: the one shape without a packed sweep
constexpr int kSwLow = -(1 << 28);    // "no gap" at row and column 0
constexpr int kSwBig = -(1 << 26);    // disables a move in the rows past the read
constexpr uint32_t kSwOpM = 0, kSwOpI = 1, kSwOpD = 2, kSwOpS = 4;  // BAM CIGAR operations
constexpr uint32_t kSwNoPair = 0xffffffffu;  // the empty half of a duo

// The packed int16 sweep (two pairs per lane, one in each 16-bit half) takes
// pairs on the 16-bit core whose scores stay small: every H is in
// [-8192, kSwNarrowMax], every E and F above -8192 - 300, so with the
// sentinels below all sums and differences fit in int16.
constexpr int kSwNarrowMax = 12000;
constexpr int kSwLow16 = -16384;  // "no gap" at row and column 0
constexpr int kSwBig16 = -10240;  // disables a move in the rows past the read
inline bool sw_narrow(uint32_t n1, uint32_t n2, int match, int mismatch, int open, int extend) {
  const uint32_t mn = n1 < n2 ? n1 : n2;
  return n2 >= 16 && match > 0 && (int64_t)mn * match <= kSwNarrowMax && mismatch <= 0 && mismatch >= -300 &&
         open <= 0 && open >= -300 && extend <= 0 && extend >= -300;
}

// Sweep shapes (segment width W, rows per lane RPL): 64 .. 2048 rows, with
// steps of 32-64 rows up to 384 so that common read lengths (100, 150, 250
// bases) leave few rows idle. Up to 12 rows per lane a kernel keeps two waves
// per SIMD. The last shape (64 x 32 rows, int32 sweep only) holds the rare
// reads of 1025-2047 bases that the reference still aligns: one wave per
// SIMD, its registers spilled to the accumulation registers (no scratch).
//
// Wide shapes (64 lanes, 2-6 rows each) serve passes too small to fill the
// device: a pair's sweep is a chain of n1 + W - 1 steps, each serial over its
// lane's RPL rows, so the longest pair of such a pass sets its time. Moving a
// pair from 16 or 32 lanes to 64 at the same or a little more capacity cuts
// the rows per step 2-4 times for a few more steps. kSwWide maps each base
// shape to its wide one (itself when it already has 64 lanes).
constexpr int kSwW[kSwShapes] = {16, 16, 16, 16, 16, 32, 32, 32, 32, 64, 64, 64, 64, 64, 64, 64, 64, 64};
constexpr int kSwR[kSwShapes] = {4, 7, 8, 10, 12, 7, 8, 10, 12, 8, 12, 16, 32, 2, 3, 4, 5, 6};
constexpr int kSwWide[kSwBaseShapes] = {13, 13, 13, 14, 14, 15, 15, 16, 17, 9, 10, 11, 12};
inline bool sw_packed_shape(int s) { return s >= 0 && s < kSwShapes && s != kSwLongShape; }
constexpr bool sw_wide_map_ok() {
  for (int s = 0; s < kSwBaseShapes; ++s) {
    const int w = kSwWide[s];
    if (kSwW[w] != 64 || kSwW[w] * kSwR[w] < kSwW[s] * kSwR[s] || (w != s && w < kSwBaseShapes)) return false;
    if (w == s && kSwW[s] != 64) return false;
  }
  return true;
}
static_assert(sw_wide_map_ok(), "a wide shape has 64 lanes and at least its base shape's rows");

// Backtrack groups per lane: 4 rows each, the last one possibly shorter
// (its bits sit where a full group's would).
constexpr int sw_groups(int rpl) { return (rpl + 3) / 4; }

// The shape for a read of n2 bases (the fewest rows with W * RPL >= n2), or -1.
inline int sw_shape(uint32_t n2, int* W, int* RPL) {
  for (int s = 0; s < kSwBaseShapes; ++s)
    if ((uint32_t)(kSwW[s] * kSwR[s]) >= n2) {
      *W = kSwW[s];
      *RPL = kSwR[s];
      return s;
    }
  return -1;
}

struct SwPair {
  uint32_t hap_off, hap_len;
  uint32_t read_off, read_len;
  uint64_t bt_off;      // backtrack words: (hap_len + W - 1) steps x W lanes x sw_groups(RPL)
  uint32_t hrow_off;    // hap_len ints: H of the read's last row per column
  uint32_t hcol_off;    // W * RPL ints: H of the haplotype's last column per row
  uint16_t width, rpl;
  uint8_t bt_lanes;     // 1: the pair's own 16 bits per 4 rows; 2: 32-bit words shared with a duo partner
  uint8_t bt_half;      // with bt_lanes 2: 0 the low, 1 the high 16 bits
  uint16_t pad;
};

struct SwArgs {
  const uint8_t* haps;
  const uint8_t* reads;
  const SwPair* pairs;
  uint32_t n_pairs;
  const uint32_t* list;  // pairs of this launch's shape (duos: two entries each)
  uint32_t n_list;       // pairs (duos)
  uint32_t* head;        // this launch's work-queue counter
  uint8_t* bt;
  int32_t* hrow;
  int32_t* hcol;
  int32_t match, mismatch, open, extend;
  uint32_t* cigar;       // max_cigar elements per pair
  uint32_t max_cigar;
  int32_t* count;
  int32_t* offset;
  int32_t* score;
  int32_t* best_ij;      // per pair: the best end cell (i, j)
  int32_t trace_run;     // traceback: cells of a diagonal loaded at once (1..8)
};

hipError_t launch_sw_fill(int shape, int grid, hipStream_t stream, const SwArgs& a);
hipError_t launch_sw_fill16(int shape, int grid, hipStream_t stream, const SwArgs& a);
// best end and traceback of the pairs a.list names (a.n_list entries,
// kSwNoPair skipped)
hipError_t launch_sw_trace(hipStream_t stream, const SwArgs& a);
int sw_blocks_per_cu(int shape);
int sw16_blocks_per_cu(int shape);
void sw_release_workspaces();  // the aligner's per-thread workspaces (engine shutdown)

// Host side of the aligner (sw_capi.cpp, sw_server.cpp). sw_validate checks
// a call's arguments as the batched entry point does and reports its longest
// read; sw_align_batch is rvc_sw_align_batch; sw_align_served is
// rvc_sw_align_served; sw_server_stop ends the serving threads (engine
// shutdown, before the workspaces are released).
int sw_validate(const uint8_t* hap_bases, const int32_t* hap_off, uint32_t n_haps, const uint8_t* read_bases,
                const int32_t* read_off, uint32_t n_reads, const uint32_t* read_hap, uint32_t max_cigar,
                const uint32_t* cigar, const int32_t* count, const int32_t* offset, const int32_t* score,
                int32_t* longest);
int sw_align_batch(const uint8_t* hap_bases, const int32_t* hap_off, uint32_t n_haps, const uint8_t* read_bases,
                   const int32_t* read_off, uint32_t n_reads, const uint32_t* read_hap, int32_t assemble,
                   uint32_t max_cigar, uint32_t* cigar, int32_t* count, int32_t* offset, int32_t* score,
                   rvc_sw_stats* stats);
int sw_align_served(const uint8_t* hap_bases, const int32_t* hap_off, uint32_t n_haps, const uint8_t* read_bases,
                    const int32_t* read_off, uint32_t n_reads, const uint32_t* read_hap, int32_t assemble,
                    uint32_t max_cigar, uint32_t* cigar, int32_t* count, int32_t* offset, int32_t* score);
void sw_server_stop();

}  // namespace rvc
