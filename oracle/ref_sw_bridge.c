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

// CPU baseline: pairs [0, n) on n_threads threads (pair p: haps + hap_off[hap[p]],
// reads + read_off[p]), each thread with its own aligner instance of the
// genotyping or (is_assemble) the assembly scores. Returns the wall seconds.
#include <pthread.h>
#include <time.h>

typedef struct {
  const unsigned char* haps;
  const int* hap_off;
  const unsigned char* reads;
  const int* read_off;
  const unsigned int* hap;
  int first, step, n, is_assemble;
} sw_job;

static void* sw_worker(void* arg) {
  const sw_job* j = (const sw_job*)arg;
  unsigned int cig[4096];
  int off, sc;
  for (int p = j->first; p < j->n; p += j->step) {
    const unsigned int h = j->hap[p];
    ref_sw_align(j->is_assemble, j->haps + j->hap_off[h], j->hap_off[h + 1] - j->hap_off[h], j->reads + j->read_off[p],
                 j->read_off[p + 1] - j->read_off[p], cig, 4096, &off, &sc);
  }
  return 0;
}

double ref_sw_align_many(int is_assemble, const unsigned char* haps, const int* hap_off, const unsigned char* reads,
                         const int* read_off, const unsigned int* hap, int n, int n_threads) {
  pthread_t th[256];
  sw_job jobs[256];
  if (n_threads > 256) n_threads = 256;
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int t = 0; t < n_threads; ++t) {
    jobs[t].haps = haps; jobs[t].hap_off = hap_off; jobs[t].reads = reads; jobs[t].read_off = read_off;
    jobs[t].hap = hap; jobs[t].first = t; jobs[t].step = n_threads; jobs[t].n = n;
    jobs[t].is_assemble = is_assemble;
    pthread_create(&th[t], 0, sw_worker, &jobs[t]);
  }
  for (int t = 0; t < n_threads; ++t) pthread_join(th[t], 0);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
