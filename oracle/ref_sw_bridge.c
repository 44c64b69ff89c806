// Test bridge to Rovaca's Smith-Waterman library (common/base/lib/sw_avx and
// common/base/lib/ro_sw), compiled from its sources into oracle/_ref: one
// alignment per call through the library's own entry points
// (sw_avx_init / avx_function, genotype/utils/alignment_utils.cpp:573-579).
#include <string.h>

#include "smithwaterman_common.h"

static __thread p_lib_sw_avx g_sw[2];

// seq1 = haplotype (rows), seq2 = read (columns). Returns the CIGAR element
// count (BAM encoding into cigar[]), or -1 when the library refused the pair
// (a length >= its 2048 limit: it sets alignment_offset = 1 and leaves the
// CIGAR as it was).
int ref_sw_align(int is_assemble, const unsigned char* seq1, int len1, const unsigned char* seq2, int len2,
                 unsigned int* cigar, int max_cigar, int* offset, int* score) {
  p_lib_sw_avx* slot = &g_sw[is_assemble ? 1 : 0];
  if (!*slot) *slot = sw_avx_init((uint8_t)(is_assemble ? 1 : 0));
  p_lib_sw_avx sw = *slot;
  if (!sw) return -2;
  if (len1 >= sw->max_seq_len || len2 >= sw->max_seq_len) {
    *offset = 1;
    return -1;
  }
  sw->len1 = (int16_t)len1;
  memcpy(sw->seq1, seq1, (size_t)len1);
  sw->len2 = (int16_t)len2;
  memcpy(sw->seq2, seq2, (size_t)len2);
  sw->avx_function(sw);
  *offset = sw->alignment_offset;
  *score = sw->score;
  int n = (int)sw->cigar_count;
  if (n > max_cigar) n = max_cigar;
  for (int x = 0; x < n; ++x) cigar[x] = sw->cigar[x];
  return n;
}
