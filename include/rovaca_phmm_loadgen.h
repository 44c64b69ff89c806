/*
 * rovaca_phmm_loadgen.h — load generator for the served form.
 *
 * No counterpart in Rovaca: it stands in for Rovaca's worker threads (the
 * boost::asio pool that calls call_pairhmm once per active region,
 * hc/haplotypecaller_engine.cpp:119,254) so the serving loop can be measured
 * under an arrival process, and the aligner's served form under Rovaca's
 * worker pattern. Used by the benchmark and the tests.
 */
#ifndef ROVACA_PHMM_LOADGEN_H_
#define ROVACA_PHMM_LOADGEN_H_

#include <stdint.h>

#include "rovaca_phmm.h"
#include "rovaca_sw.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rvc_loadgen_result {
  uint64_t submitted, completed, failed;
  uint64_t cells;          /* sum of read_len * hap_len over completed regions            */
  double wall_s;           /* first submission to last completion                          */
  double lat_p50_ms, lat_p99_ms, lat_max_ms, lat_mean_ms; /* submit -> results ready      */
  double queue_p50_ms, queue_p99_ms;                      /* submit -> launched           */
  double late_p99_ms;      /* how late producers submitted against the schedule           */
} rvc_loadgen_result;

/* Submits region order[i] (or i, if order is NULL) at arrival_s[i] seconds
 * after the start, for i in [0, n_arrivals), from n_producers threads
 * (arrival i goes to producer i % n_producers). Submissions never wait for
 * results; a collector thread per producer waits for that producer's tickets
 * in order as they are submitted.
 * The serving loop must be running (rvc_config.server_threads = 1).
 * latency_ms (optional, n_arrivals entries) receives each request's
 * submit -> results-ready time. out_read_major (optional, n_arrivals
 * entries, each NULL or room for that arrival's region's n_reads * n_haps
 * doubles) receives each arrival's read-major matrix as rvc_phmm_wait
 * returns it. */
int rvc_loadgen_run(const rvc_region* regions, uint32_t n_regions, const uint32_t* order, const double* arrival_s,
                    uint32_t n_arrivals, uint32_t n_producers, double* latency_ms, double* const* out_read_major,
                    rvc_loadgen_result* result);

/* Rovaca's worker pattern: each of n_workers threads calls
 * rvc_phmm_call_pairhmm (raw reads in, hap-major matrix and kept reads out)
 * for its share of n_calls calls, one region at a time; call i uses region
 * order[i] (or i % n_regions). latency_ms (optional, n_calls entries)
 * receives each call's duration. result->wall_s spans the first call to the
 * last return; queue and lateness fields are 0. */
int rvc_loadgen_call_pairhmm(const rvc_raw_region* regions, uint32_t n_regions, const uint32_t* order, uint32_t n_calls,
                             int32_t min_quality_threshold, int32_t pcr_option, uint32_t n_workers, double* latency_ms,
                             rvc_loadgen_result* result);

/* Latency of the one-region call as a C++ caller sees it: n_calls calls of
 * rvc_phmm_region_sync (noalloc = 0) or rvc_phmm_region_sync_noalloc
 * (noalloc = 1) on one region from this thread, one after another;
 * latency_ms[i] receives call i's duration. */
int rvc_loadgen_region_latency(const rvc_region* region, uint32_t n_calls, int32_t noalloc, double* latency_ms);

/* One region's realignment call: the arguments of rvc_sw_align_batch. */
typedef struct rvc_sw_region {
  const uint8_t* hap_bases;
  const int32_t* hap_off;
  uint32_t n_haps;
  const uint8_t* read_bases;
  const int32_t* read_off;
  uint32_t n_reads;
  const uint32_t* read_hap;
  uint32_t max_cigar;
  uint32_t* cigar;
  int32_t* count;
  int32_t* offset;
  int32_t* score;
} rvc_sw_region;

/* Rovaca's realignment pattern (realign_reads_to_their_best_haplotype once
 * per region, hc/haplotypecaller_engine.cpp:268, from every worker thread):
 * n_workers threads take the n_regions regions one at a time and align each
 * with one call, rvc_sw_align_served (served != 0) or rvc_sw_align_batch,
 * with the genotyping or (assemble != 0) the assembly scores. latency_ms
 * (optional, n_regions entries) receives each call's duration;
 * result->cells sums haplotype length x read length over the aligned pairs,
 * result->wall_s spans the first call to the last return. */
int rvc_loadgen_sw_align(const rvc_sw_region* regions, uint32_t n_regions, int32_t assemble, int32_t served,
                         uint32_t n_workers, double* latency_ms, rvc_loadgen_result* result);

#ifdef __cplusplus
}
#endif

#endif /* ROVACA_PHMM_LOADGEN_H_ */
