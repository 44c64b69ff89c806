/*
 * rovaca_phmm.h — C ABI of the MI355X PairHMM engine for Rovaca's
 * HaplotypeCaller.
 *
 * Plain C: integer return codes, caller-owned host buffers, no C++ or torch
 * types. Paths in the citations are relative to the Rovaca tree, with
 * ph/ = src/haplotypecaller/pairhmm/ and hc/ = src/haplotypecaller/.
 *
 * Numerical contract (identical to Rovaca's PairHMM): per read x haplotype
 * pair the forward algorithm over match/insert/delete states in FP32 with the
 * initial constant 2^120; pairs whose FP32 sum is below 1e-28 are recomputed
 * in FP64 with 2^1020. Results are log10 likelihoods as doubles:
 *   FP32 path: log10((double)S) - 36.1236000061f   (ph/roDV_inter/pairhmm_inter_api.h:10-12)
 *   FP64 path: log10(S) - 307.050595577260822       (ph/adapter.h:25-26)
 * An 'N' in either sequence matches every base (ph/roDV_intra/pairhmm_impl.cpp:13-44).
 * Bases must be A, C, G, T or N: any other byte makes the call return
 * RVC_E_BASES without computing anything.
 */
#ifndef ROVACA_PHMM_H_
#define ROVACA_PHMM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------- */
#define RVC_OK 0
#define RVC_E_INVALID 1   /* null pointer, bad size, zero haplotypes or reads           */
#define RVC_E_LIMIT 2     /* beyond the reference's limits: > 4096 haplotypes, a        */
                          /* haplotype > 1536 bases or a read > 768 bases               */
                          /* (ph/roDV_schedule.cpp:70-73, 374-381, 445-459)             */
#define RVC_E_DEVICE 3    /* a HIP call failed (device lost, out of device memory)      */
#define RVC_E_STATE 4     /* rvc_phmm_init has not run, or ran twice                    */
#define RVC_E_BUDGET 5    /* one region needs more device memory than the whole budget  */
#define RVC_E_TIMEOUT 6   /* rvc_phmm_wait gave up before the result was ready          */
#define RVC_E_SHUTDOWN 7  /* the server stopped before this request finished            */
#define RVC_E_BASES 8     /* a base other than A, C, G, T or N (an IUPAC code, '=', a    */
                          /* lowercase letter). Rovaca's result for such a byte depends */
                          /* on its kernel routing: its inter kernel compares raw bytes */
                          /* (ph/roDV_inter/pairhmm_inter.cpp:167,287), its intra kernel */
                          /* maps B..T to 'A' and reads past its table above 'T'        */
                          /* (ph/roDV_common/common.h:42-56), and which kernel a pair   */
                          /* takes follows its greedy grouping (ph/roDV_schedule.cpp:    */
                          /* 86-142, 417-441). Nothing is computed; the caller keeps    */
                          /* Rovaca's own CPU PairHMM for the region.                   */

/* per-pair flags written beside each result */
#define RVC_PAIR_COMPUTED 0x1u /* the pair had two non-empty sequences               */
#define RVC_PAIR_FP64 0x2u     /* FP32 sum was < 1e-28; the value is the FP64 result  */

/* ---- inputs --------------------------------------------------------------- */
/*
 * One active region after read preprocessing: haplotype bases and read
 * bases as ASCII (A C G T N), and per read base the four Phred arrays that
 * Rovaca's scheduler hands its kernels (ph/roDV_schedule.cpp:522-649):
 * base quality q, insertion and deletion gap-open i, d, gap continuation c.
 * Haplotype h is hap_bases[hap_off[h] .. hap_off[h+1]), read r is
 * read_bases[read_off[r] .. read_off[r+1]) and q/i/d/c share read_off.
 * Replaces the pair list that schedule_pairhmm_with_allocator builds from
 * HaplotypeVector / ReadVector (ph/roDV_schedule.cpp:364-475).
 */
typedef struct rvc_region {
  uint32_t n_haps;
  uint32_t n_reads;
  const int32_t* hap_off;    /* n_haps + 1 entries  */
  const uint8_t* hap_bases;
  const int32_t* read_off;   /* n_reads + 1 entries */
  const uint8_t* read_bases;
  const uint8_t* q;
  const uint8_t* i;
  const uint8_t* d;
  const uint8_t* c;
} rvc_region;

/*
 * A region before read preprocessing, as call_pairhmm receives it
 * (ph/pairhmm_engine.h:14-15): haplotype bases, read bases, raw base
 * qualities and each read's mapping quality. Gap penalties come from the
 * defaults or the PCR indel model inside the engine.
 */
typedef struct rvc_raw_region {
  uint32_t n_haps;
  uint32_t n_reads;
  const int32_t* hap_off;
  const uint8_t* hap_bases;
  const int32_t* read_off;
  const uint8_t* read_bases;
  const uint8_t* read_qual;
  const uint8_t* read_mq;    /* n_reads entries */
} rvc_raw_region;

/* PcrIndelModel (hc/common/enum.h:8) */
#define RVC_PCR_NONE 0
#define RVC_PCR_HOSTILE 1
#define RVC_PCR_AGGRESSIVE 2
#define RVC_PCR_CONSERVATIVE 3

typedef struct rvc_config {
  int32_t device;            /* HIP device ordinal                                  */
  uint32_t server_threads;   /* 0 = no serving loop; 1 = start it at init          */
  uint64_t hbm_budget_bytes; /* serving: bound on resident request buffers (0 = 64 GiB) */
  uint32_t max_batch_regions;/* serving: regions per device launch (0 = 4096)       */
  uint32_t host_prep;        /* where rvc_phmm_call_pairhmm preprocesses reads:     */
                             /* 0 = on the device when the PCR model is on, else   */
                             /* on the calling thread; 1 = always on the calling   */
                             /* thread; 2 = always on the device                   */
  const char* tune_table;    /* kernel-shape table written by the tuner; NULL = the */
                             /* table installed beside the library, if any         */
  uint32_t sw_min_launch_pairs; /* aligner (rovaca_sw.h): a sweep shape with fewer */
                             /* pairs runs with the next larger one launched       */
                             /* anyway (0 = enough to fill the device once;        */
                             /* 1 = every shape launches on its own)               */
  uint32_t sw_bt_budget_mib; /* aligner: backtrack bytes per device pass, in MiB;   */
                             /* larger batches run in several passes (0 = 12 GiB)  */
  uint32_t sw_wide_shapes;   /* aligner: sweep a small pass on 64 lanes per pair,  */
                             /* 2-6 rows per lane (0 = when its pairs take at most */
                             /* three waves per SIMD; 1 = never; 2 = always)       */
  uint32_t reserved[3];
} rvc_config;

/* ---- lifecycle -------------------------------------------------------------- */
/* Replaces init_pairhmm_ptr (ph/pairhmm_engine.cpp:19-38): selects the
 * device, uploads the Phred and match-to-match tables (ph/roDV_common/context.h:44-60,
 * 98-101, 138-141) and, if asked, starts the serving loop. */
int rvc_phmm_init(const rvc_config* cfg);
int rvc_phmm_shutdown(void);
/* Last error text of the calling thread (never NULL). */
const char* rvc_phmm_last_error(void);

/* ---- synchronous, one region per launch ------------------------------------ */
/* Replaces schedule_pairhmm_with_allocator (ph/roDV_schedule.cpp:364-834):
 * out_read_major[r * n_haps + h] receives the log10 likelihood of read r
 * against haplotype h; pairs with an empty side keep NaN
 * (ph/pairhmm_internal.cpp:593). out_flags (optional, same shape) receives
 * RVC_PAIR_* bits. Thread-safe; each call launches its own kernels. */
int rvc_phmm_region_sync(const rvc_region* region, double* out_read_major, uint8_t* out_flags);

/* No counterpart in the reference: the same call and results as
 * rvc_phmm_region_sync for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It allocates no device or pinned host memory: the region's
 * device image lives in a fixed 32 MiB device array, and its host image in a
 * 32 MiB pinned buffer, both set up by rvc_phmm_init; a region whose image
 * exceeds them gets RVC_E_LIMIT. Calls are serialised (they share the
 * array); each call uploads its starting results and zero counters. */
int rvc_phmm_region_sync_noalloc(const rvc_region* region, double* out_read_major, uint8_t* out_flags);

/* Replaces call_roDV_pairhmm_scheduled (ph/pairhmm_internal.cpp:576-623):
 * read preprocessing (MQ cap, min-quality floor, GOP/GCP defaults, PCR indel
 * model: ph/roDV_schedule.cpp:494-649, ph/pairhmm_internal.cpp:260-476),
 * likelihoods, normalize (:248-258), removal of poorly modelled reads
 * (:22-44) and transpose (:499-508). out_hap_major is [n_haps][n_reads]; its
 * first *n_kept columns belong to kept_reads[0 .. *n_kept), the rest stay
 * NaN. On failure every entry is NaN, *n_kept = n_reads, kept_reads lists all
 * reads and the return code says why (the reference logs and returns the
 * all-NaN matrix, :605-609). Thread-safe: with the serving loop running
 * (rvc_config.server_threads = 1) the regions of concurrent callers are
 * batched together on the device; otherwise each call runs its own region. */
int rvc_phmm_call_pairhmm(const rvc_raw_region* region, int32_t min_quality_threshold, int32_t pcr_option,
                          double* out_hap_major, uint32_t* kept_reads, uint32_t* n_kept);

/* A queued request of the serving loop (rvc_phmm_submit and
 * rvc_phmm_call_pairhmm_submit). */
typedef uint64_t rvc_ticket;

/* The variant sites of a region, for haplotype -> allele marginalisation at
 * the genotyping boundary (AlleleLikelihoods::marginalize,
 * hc/genotype/allele_likelihoods.hpp:189-217, 701-732, called per site at
 * hc/genotype/germline_genotying_engine.cpp:143 with the allele mapper of
 * AssemblyBasedCallerUtils::create_allele_mapper, which depends only on the
 * haplotypes). hap_allele[s * n_haps + h] is the allele of haplotype h at
 * site s, or -1 where the haplotype maps to none (each haplotype maps to at
 * most one allele of a site, as old_to_new_allele_index_map requires). */
typedef struct rvc_sites {
  uint32_t n_sites;
  const uint32_t* n_alleles;  /* n_sites entries, each 1 .. 32767 */
  const int32_t* hap_allele;  /* n_sites * n_haps entries */
} rvc_sites;

/* The same region-level call in two halves, so that a worker can go on with
 * its next region (assembly, the CPU steps before PairHMM) while this one is
 * on the device; Rovaca's worker otherwise blocks on call_pairhmm for every
 * region (hc/haplotypecaller_engine.cpp:149-289, the call at :254).
 * rvc_phmm_call_pairhmm_submit preprocesses the reads (on the calling thread
 * or, as configured, on the device), packs the region and queues it on the
 * serving loop (rvc_config.server_threads = 1, else RVC_E_STATE); it returns
 * at once, and the caller's buffers are free when it does. A region the
 * engine cannot take (RVC_E_BASES, RVC_E_LIMIT, RVC_E_BUDGET ...) is refused
 * here, with nothing queued. rvc_phmm_poll(ticket) tells whether it is done. */
int rvc_phmm_call_pairhmm_submit(const rvc_raw_region* region, int32_t min_quality_threshold, int32_t pcr_option,
                                 const rvc_sites* sites, rvc_ticket* ticket);
/* Waits for a submitted region call (timeout_ms < 0: no limit) and finishes
 * it as rvc_phmm_call_pairhmm does: normalize, removal of poorly modelled
 * reads and transpose into out_hap_major / kept_reads / n_kept; on failure
 * all NaN with every read kept. RVC_E_TIMEOUT leaves the ticket valid.
 * With sites given at submit, the device also marginalises the normalised
 * matrix while it is resident (allele_likelihoods.hpp:701-732: per read the
 * largest likelihood of the haplotypes mapped to the allele, -inf for an
 * allele no haplotype maps to, a NaN never the largest) and out_alleles
 * receives, site after site, [n_alleles[s]][n_reads] values whose first
 * *n_kept columns belong to kept_reads (as the reference marginalises the
 * kept reads) and the rest NaN. out_alleles may be NULL. */
int rvc_phmm_call_pairhmm_wait(rvc_ticket ticket, double* out_hap_major, uint32_t* kept_reads, uint32_t* n_kept,
                               double* out_alleles, int64_t timeout_ms);

/* ---- batched: many regions in one launch sequence -------------------------- */
typedef struct rvc_batch rvc_batch;

typedef struct rvc_batch_opts {
  uint32_t persistent_waves_per_cu; /* 0 = as many as fit                             */
  uint32_t use_graph;               /* 1 = capture the launch sequence in a hipGraph  */
  uint32_t fp32_fold;               /* 0 = tuned choice, 1 = rescale I/D (the         */
                                    /* 5-operation scaled form, FP32 and FP64) where  */
                                    /* the gap qualities allow, 2 = never             */
  uint32_t rpl_round;               /* 0 = tuned choice, 1 = exact rows per lane,     */
                                    /* 2 = rows per lane in steps of 3                */
  uint32_t occupancy;               /* 0 = tuned choice, else 2/3/4 waves per SIMD    */
                                    /* for the 16-lane kernels                        */
  uint32_t segment_width;           /* 0 = tuned choice, else 8/16/32/64 lanes per read */
                                    /* for reads up to 256 bases                      */
  uint32_t prior_table;             /* 0 = tuned choice, 1 = prior-table kernel for   */
                                    /* haplotypes without N, 2 = compare-and-select   */
  uint32_t hap_run;                 /* haplotypes per tile: a tile sets up its reads  */
                                    /* once and sweeps a run of haplotypes. 0 = up to */
                                    /* the tuned maximum, fewer for small batches;    */
                                    /* 1..64 = exactly that many                      */
  uint32_t untimed;                 /* 1 = eager runs queue the launches without the  */
                                    /* per-launch events (rvc_phmm_batch_kernels and  */
                                    /* the FP32/FP64 split are then not measured)     */
  uint32_t shape_fold;              /* 0 = tiles of a shape with few cells are swept  */
                                    /* by the next larger shape of their width (at    */
                                    /* most a third more rows; same results), 1 = every */
                                    /* tile keeps its own shape                       */
} rvc_batch_opts;

typedef struct rvc_batch_info {
  uint64_t n_pairs;       /* pairs with two non-empty sides                          */
  uint64_t n_cells;       /* sum over those pairs of read_len * hap_len              */
  uint64_t n_rescued;     /* pairs recomputed in FP64 by the last run                */
  uint64_t rescued_cells;
  uint64_t device_bytes;  /* device memory held by the batch                         */
  double fp32_ms;         /* last run: FP32 pass, device time (HIP events)            */
  double fp64_ms;         /* last run: FP64 rescue pass, device time                  */
  double total_ms;        /* last run: whole launch sequence, device time             */
  uint32_t n_fp32_launches;
  uint32_t n_fp64_launches;
} rvc_batch_info;

/* One kernel launch of the last timed (non-graph) run of a batch. */
typedef struct rvc_kernel_time {
  char name[48];          /* e.g. "k_fp32<16,10,scaled,occ4>"                        */
  uint32_t precision;     /* 32 or 64                                                */
  uint32_t width;         /* lanes per read (W)                                      */
  uint32_t rows_per_lane; /* RPL                                                     */
  uint32_t occupancy;     /* waves per SIMD the kernel was compiled for              */
  uint64_t tiles;         /* work units queued (FP64: rescue capacity)               */
  uint64_t cells;         /* sum of read_len * hap_len of the pairs it computes      */
  double ms;              /* device time between the events around the launch        */
} rvc_kernel_time;

/* Packs regions (2-bit bases + N mask, int8 quals), uploads them and builds
 * the length-binned tile queues. Nothing runs yet. */
int rvc_phmm_batch_create(const rvc_region* regions, uint32_t n_regions, const rvc_batch_opts* opts,
                          rvc_batch** out);
/* Runs the FP32 pass and the FP64 rescue on the device; returns when done. */
int rvc_phmm_batch_run(rvc_batch* b);
/* Region r's read-major matrix starts at out[region_offset(r)], where the
 * offsets are the running sum of n_reads * n_haps. flags may be NULL. */
int rvc_phmm_batch_results(rvc_batch* b, double* out, uint8_t* flags);
int rvc_phmm_batch_info(const rvc_batch* b, rvc_batch_info* info);
/* Per-launch times of the last non-graph run: fills up to max entries and
 * sets *n to the number of launches. */
int rvc_phmm_batch_kernels(const rvc_batch* b, rvc_kernel_time* out, uint32_t max, uint32_t* n);
int rvc_phmm_batch_destroy(rvc_batch* b);

/* ---- served: continuous batching ------------------------------------------- */
/* Requests come from many host threads (Rovaca's worker pool,
 * hc/haplotypecaller_engine.cpp:119,254). Each submit copies the region into
 * pinned staging and returns at once; the serving loop admits requests while
 * their resident buffers fit the HBM budget, launches them together with
 * whatever else is waiting, and completes each ticket on its own. */

/* Steady-clock nanoseconds of a request's life: submitted, launched in a
 * batch, results ready on the host. */
typedef struct rvc_ticket_timing {
  uint64_t submit_ns;
  uint64_t launch_ns;
  uint64_t done_ns;
} rvc_ticket_timing;

int rvc_phmm_submit(const rvc_region* region, rvc_ticket* ticket);
/* Blocks until the ticket completes (timeout_ms < 0: no limit). Copies the
 * read-major matrix and flags out (each optional), fills timing (optional)
 * and releases the ticket. */
int rvc_phmm_wait(rvc_ticket ticket, double* out_read_major, uint8_t* out_flags, int64_t timeout_ms,
                  rvc_ticket_timing* timing);
/* 1 if done, 0 if not yet, negative on a bad ticket. */
int rvc_phmm_poll(rvc_ticket ticket);

typedef struct rvc_server_stats {
  uint64_t submitted, completed, launches;
  uint64_t cells, rescued_pairs;
  uint64_t resident_bytes, peak_resident_bytes;
  double busy_ms;          /* device time of all serving launches */
  double stage_ms;         /* host time spent merging requests into launches */
  uint64_t buffer_grows;   /* device arena and pinned buffer reallocations */
} rvc_server_stats;
int rvc_phmm_server_stats(rvc_server_stats* s);

/* ---- host-side steps of the region call (no device work) ------------------ */
/* Read preprocessing of one read, as ph/roDV_schedule.cpp:522-608 does it
 * before its kernels: q = min(qual, mq), then 6 where below
 * min_quality_threshold (ph/adapter.h:5); gop = 45 ('-'), lowered by the PCR
 * indel model (ph/pairhmm_internal.cpp:391-476, tables ph/adapter.h:13-15);
 * gcp = 10 ('+' - 33, hc/genotype/read_record.cpp:205-210). The deletion
 * gap-open array is the same as gop (ph/roDV_schedule.cpp:589-591). */
int rvc_phmm_preprocess_read(const uint8_t* bases, const uint8_t* qual, int32_t len, uint8_t mq,
                             int32_t min_quality_threshold, int32_t pcr_option, uint8_t* q, uint8_t* gop,
                             uint8_t* gcp);
/* Post-processing of a read-major [n_reads][n_haps] matrix, in place:
 * normalize (ph/pairhmm_internal.cpp:248-258), drop poorly modelled reads
 * (:22-44, :489-497) and transpose into out_hap_major (:499-508). Returns the
 * number of kept reads (listed in kept_reads), or a negative code. */
int rvc_phmm_postprocess(uint32_t n_reads, uint32_t n_haps, double* read_major, const int32_t* read_len,
                         double* out_hap_major, uint32_t* kept_reads);
/* Copies the Phred->probability (128) and match-to-match (128*129/2) tables
 * the engine uploads, built as ph/roDV_common/context.h:44-60, 98, 138 builds
 * them. Any pointer may be NULL. */
int rvc_phmm_tables(float* ph2pr_f32, float* mm_f32, double* ph2pr_f64, double* mm_f64);

#ifdef __cplusplus
}
#endif

#endif /* ROVACA_PHMM_H_ */
