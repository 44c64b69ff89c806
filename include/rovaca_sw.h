/* Batched Smith-Waterman of reads against haplotypes on an MI355X: the
 * aligner behind Rovaca's read realignment to the best haplotype.
 *
 * Replaces, one call for a whole batch, the per-read use of Rovaca's
 * aligner library in AlignmentUtils::create_read_aligned_to_ref
 * (hc/genotype/utils/alignment_utils.cpp:573-579: seq1 = haplotype,
 * seq2 = read without soft clips, sw->avx_function(sw), then cigar,
 * cigar_count and alignment_offset), i.e. sw_avx_init(is_assemble) +
 * avx_function (hc/common/base/lib/sw_avx/sw_avx.c:402-451, the ro_sw core
 * common/base/lib/ro_sw/sw_core_template.h:110-361 and sw_cigar.c:20-171),
 * with the same results. Uses the engine initialised by rvc_phmm_init
 * (rovaca_phmm.h) and its error codes. */
#ifndef ROVACA_SW_H_
#define ROVACA_SW_H_

#include <stdint.h>

#include "rovaca_phmm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rvc_sw_stats {
  double fill_ms;    /* device time of the matrix sweeps (HIP events)        */
  double trace_ms;   /* device time of best-end search and traceback          */
  double total_ms;   /* wall time of the call, copies included                */
  uint64_t cells;    /* sum of haplotype length x read length over the pairs  */
  uint32_t launches; /* sweep launches (one per shape and chunk)              */
  uint32_t chunks;   /* device passes the batch was split into                */
} rvc_sw_stats;

/* Aligns read r (read_bases[read_off[r] .. read_off[r+1]), ASCII, as the
 * caller would copy into seq2) against haplotype read_hap[r]
 * (hap_bases[hap_off[h] .. hap_off[h+1])), for r < n_reads, with the
 * genotyping scores (match 10, mismatch -15, gap open -30, extend -5) or,
 * when assemble != 0, the assembly scores (200, -150, -260, -11).
 * Outputs per read: count[r] CIGAR elements at cigar[r * max_cigar ..] in
 * BAM encoding (length << 4 | op; M 0, I 1, D 2, S 4), offset[r] the
 * alignment offset on the haplotype (the library's alignment_offset),
 * score[r] the best score. Where the library refuses the pair (a length
 * >= 2048) count[r] = -1 and offset[r] = 1. An empty haplotype gives one
 * element S<read length>, offset 0, score 0; an empty read against n1 bases
 * gives no element, offset n1, score 0 (both empty: score INT32_MIN), as a
 * fresh instance of the library does (sw_core_template.h:356-361 refuses
 * only the lengths >= 2048).
 * max_cigar = 2 * longest read + 2 always suffices.
 * Reads and haplotypes of up to 2047 bases are aligned (reads of 1025-2047
 * bases take a slow one-wave-per-SIMD sweep; they are rare in short-read
 * data). Returns RVC_OK, RVC_E_INVALID (null pointer, bad index), RVC_E_LIMIT
 * (more CIGAR elements than max_cigar), RVC_E_STATE or RVC_E_DEVICE. stats
 * may be NULL. */
int rvc_sw_align_batch(const uint8_t* hap_bases, const int32_t* hap_off, uint32_t n_haps, const uint8_t* read_bases,
                       const int32_t* read_off, uint32_t n_reads, const uint32_t* read_hap, int32_t assemble,
                       uint32_t max_cigar, uint32_t* cigar, int32_t* count, int32_t* offset, int32_t* score,
                       rvc_sw_stats* stats);

/* The same call for callers that realign one region at a time from many
 * threads, as Rovaca's workers do (realign_reads_to_their_best_haplotype,
 * hc/genotype/utils/assembly_based_caller_utils.h:48, called once per region
 * at hc/haplotypecaller_engine.cpp:268 from each worker thread): same
 * arguments, results and errors as rvc_sw_align_batch. Concurrent calls are
 * merged into shared device passes by the engine's serving threads; the call
 * returns when its own results are in place. A call whose CIGAR does not fit
 * max_cigar fails alone (RVC_E_LIMIT); calls during rvc_phmm_shutdown get
 * RVC_E_STATE. */
int rvc_sw_align_served(const uint8_t* hap_bases, const int32_t* hap_off, uint32_t n_haps, const uint8_t* read_bases,
                        const int32_t* read_off, uint32_t n_reads, const uint32_t* read_hap, int32_t assemble,
                        uint32_t max_cigar, uint32_t* cigar, int32_t* count, int32_t* offset, int32_t* score);

#ifdef __cplusplus
}
#endif

#endif /* ROVACA_SW_H_ */
