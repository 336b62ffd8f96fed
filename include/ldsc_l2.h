/*
 * ldsc_l2.h — C ABI of the MI355X (gfx950) LD-score engine.
 *
 * This is the drop-in boundary for the `l2` hot path of sharifhsn/ldsc
 * (crate `ldsc`, v0.5.0).  Every entry point below names the reference
 * interface it replaces.  Signatures carry only plain C types: pointers,
 * sizes and POD structs.  All input/output buffers belong to the caller;
 * panels and servers belong to the library.
 *
 * Error convention: every function returning `int` returns LDSC_OK (0) on
 * success and one of the LDSC_ERR_* codes otherwise, with a NUL-terminated
 * message written into `err` (when `err != NULL && err_len > 0`).  This
 * mirrors the reference's `anyhow::Result` / `Err(..)` cases.
 */
#ifndef LDSC_L2_H
#define LDSC_L2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDSC_L2_ABI_VERSION 1

enum {
    LDSC_OK = 0,
    LDSC_ERR_INVALID_ARG = 1,  /* null pointer, chunk_size == 0, bad mode ... */
    LDSC_ERR_BAD_BED = 2,      /* magic / SNP-major mode / length mismatch   */
    LDSC_ERR_EMPTY = 3,        /* zero SNPs or zero individuals              */
    LDSC_ERR_SKETCH_RANGE = 4, /* --sketch d outside [3, N)                  */
    LDSC_ERR_ANNOT = 5,        /* annotation rows != SNP rows                */
    LDSC_ERR_DEVICE = 6,       /* HIP runtime failure                        */
    LDSC_ERR_OOM = 7,          /* device allocation failed / over budget     */
    LDSC_ERR_BUSY = 8,         /* serving queue is full                      */
    LDSC_ERR_NOT_FOUND = 9,    /* unknown request id                         */
    LDSC_ERR_PENDING = 10,     /* request not finished yet (poll)            */
    LDSC_ERR_CANCELLED = 11,   /* request cancelled (ldsc_l2_cancel) or the  */
                               /* server destroyed before it ran             */
    LDSC_ERR_TIMEOUT = 12      /* request queued longer than max_queue_ms    */
};

/* WindowMode (src/l2/window.rs:5-12). */
enum { LDSC_WINDOW_CM = 0, LDSC_WINDOW_KB = 1, LDSC_WINDOW_SNP = 2 };

/* Arithmetic form of the exact path.
 *  EXACT:     exact integer products (G·G, G·M, M·G, M·M) finished in FP64.
 *             Replaces the reference default f64 GEMM path
 *             (compute.rs:1755-1810, 2102-2179).  The products run on the
 *             e2m1 (FP4) MFMA with f32 accumulation, exact for N <= 1,800,000
 *             individuals, and on the int8 MFMA with int32 accumulation
 *             beyond; both give the same integers, hence the same result.
 *  FP32:      standardized FP32 columns on FP32 MFMA, r² finished in FP64.
 *             Replaces `--fast-f32` (compute.rs:1668-1753, 1949-2100).
 *  EXACT_I8:  EXACT with the int8 MFMA kernel at every N.
 * LDSC_PREC_INT8_FP64 is the former name of EXACT.
 * The CountSketch path (sketch_d > 0) always sketches in FP32 like the
 * reference (`use_f32 = fast_f32 || sketch`, mod.rs:393). */
enum {
    LDSC_PREC_EXACT = 0,
    LDSC_PREC_INT8_FP64 = 0,
    LDSC_PREC_FP32 = 1,
    LDSC_PREC_EXACT_I8 = 2
};

/* BimRecord (src/parse.rs:500-507): the fields the l2 path reads. */
typedef struct {
    uint8_t chr;      /* chromosome code, 0 when unparsable            */
    uint8_t pad_[7];
    double cm;        /* genetic position (cM)                         */
    uint32_t bp;      /* physical position (bp)                        */
    uint32_t pad2_;
    uint64_t bed_idx; /* SNP row in the .bed file                      */
} ldsc_snp;

/* L2Config (src/l2/mod.rs:951-988) plus the CLI-only knobs on the path. */
typedef struct {
    int32_t window_mode;       /* LDSC_WINDOW_*                                   */
    int32_t precision;         /* LDSC_PREC_*                                     */
    double window_value;       /* cM, kb, or flanking SNP count                   */
    uint64_t chunk_size;       /* --chunk-size (reference default 200)            */
    uint64_t sketch_d;         /* --sketch d; 0 = exact                           */
    int32_t sketch_maf_aware;  /* --sketch-maf-aware                              */
    int32_t snp_level_masking; /* --snp-level-masking                             */
    double pq_exp;             /* --pq-exp / --per-allele (=1); NaN = none        */
    int32_t yes_really;        /* --yes-really (silences whole-chromosome warning) */
    int32_t per_chromosome;    /* 1 = chunk grid restarts at every chromosome run
                                  (run() default, mod.rs:478-553); 0 = one global
                                  pass (--global-pass / --python-compat)          */
} ldsc_l2_config;

/* L2Perf (src/l2/mod.rs:1016-1034) plus device counters. */
typedef struct {
    uint64_t m;
    uint64_t n_indiv;
    double bed_read_secs;   /* gather of the call's SNP slabs on device         */
    double norm_secs;       /* decode + integer stats + tiling                  */
    double sketch_secs;     /* CountSketch scatter (0 when exact)               */
    double bb_dot_secs;     /* banded contraction + r² finish (BB and AB fused) */
    double ab_dot_secs;     /* always 0: AB is fused into the band kernel       */
    double ring_store_secs; /* per-SNP reduction of the tile partials           */
    double total_secs;      /* wall time of the call                            */
    uint64_t pairs_included; /* unordered pairs (i<j) under the inclusion rule  */
    uint64_t pairs_evaluated;/* pairs the band kernel computed                  */
    uint64_t tiles;          /* 64x64 SNP tiles launched                        */
    int32_t any_full_chr_window;
    int32_t sketch_int8;     /* 1: the CountSketch contraction ran on exact int8
                                operands; 0: on f32 values (or no sketch)       */
} ldsc_l2_perf;

/* L2Progress (src/l2/mod.rs:22-39). */
typedef struct {
    uint64_t chunks_done;
    uint64_t chunks_total;
    uint64_t snps_done;
    uint64_t snps_total;
    double chunk_wall_ms;
} ldsc_l2_progress;

typedef void (*ldsc_progress_cb)(const ldsc_l2_progress* progress, void* ctx);

typedef struct ldsc_panel ldsc_panel;

/* ---------------------------------------------------------------------- */
/* Panels: a packed PLINK .bed resident in HBM.                            */
/* ---------------------------------------------------------------------- */

/* Replaces Bed::from_bytes / Bed::from_source / BedBuilder::build
 * (src/bed.rs:70-128, 238-297) and ChunkReader's --keep positions
 * (src/bed.rs:455-468).  Validates the 3-byte header (6c 1b 01) and the
 * length 3 + ceil(n_indiv_fam/4)*m_bim, then uploads the SNP slabs to the
 * device once.  `keep_idx` (nullable) selects FAM rows in the given order
 * (load_individual_indices, src/l2/io.rs:53-91); N becomes n_keep. */
int ldsc_l2_open_panel(const uint8_t* bed_bytes, uint64_t bed_len,
                       uint64_t n_indiv_fam, uint64_t m_bim,
                       const int64_t* keep_idx, uint64_t n_keep,
                       ldsc_panel** out, char* err, uint64_t err_len);

/* Releases the device copy.  NULL is accepted. */
int ldsc_l2_close_panel(ldsc_panel* panel);

/* Individuals N used by computations on this panel (after --keep). */
uint64_t ldsc_l2_panel_n_indiv(const ldsc_panel* panel);

/* ---------------------------------------------------------------------- */
/* The hot path.                                                            */
/* ---------------------------------------------------------------------- */

/* Replaces compute_ldscore_global (src/l2/compute.rs:1009-2447) and, with
 * cfg->per_chromosome = 1, the per-chromosome fan-out of run()
 * (src/l2/mod.rs:478-553).
 *   snps    m BIM records (the SNP list of the call, in order)
 *   annot   m x K column-major annotation matrix, NULL = all-ones (K must be 1)
 *   l2_out  m x K column-major LD scores (caller-owned)
 *   maf_out m MAFs (caller-owned, nullable)
 *   perf    nullable
 *   cb      nullable: progress while the call runs (L2Progress,
 *           src/l2/mod.rs:22-39; called per chunk by compute.rs:2391-2398).
 *           The reference's chunks are chunk_size SNPs on the call's grid
 *           (restarted at every chromosome run when per_chromosome).  With a
 *           callback the band runs in up to 8 pieces of SNP rows; after each
 *           piece that completes further chunks, cb receives the number of
 *           chunks whose SNPs are complete (monotone), the end of the last of
 *           them in snps_done and the wall time since the previous call.  A
 *           last call with chunks_done == chunks_total and snps_done == m
 *           comes once the results are back: at most 9 calls in all, on the
 *           calling thread. */
int ldsc_l2_compute(const ldsc_panel* panel, const ldsc_snp* snps, uint64_t m,
                    const ldsc_l2_config* cfg, const double* annot, uint64_t K,
                    double* l2_out, double* maf_out, ldsc_l2_perf* perf,
                    ldsc_progress_cb cb, void* cb_ctx, char* err, uint64_t err_len);

/* ldsc_l2_compute without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels
 * and results as ldsc_l2_compute (bit-identical), but every device buffer
 * of the call is carved from one fixed 256 MiB device array per GPU (held
 * from the first call on that GPU until the process ends), and host staging
 * is pageable memory: no hipMalloc, hipFree or hipHostMalloc after the first
 * call on a device, which sets up the array's address and the call's
 * streams.  A call whose buffers would exceed 256 MiB fails with
 * LDSC_ERR_OOM (message names the scratch size); the CountSketch path
 * (sketch_d > 0) fails with LDSC_ERR_INVALID_ARG.  Calls on one GPU share its
 * array and are serialized by a lock inside the library. */
int ldsc_l2_compute_noalloc(const ldsc_panel* panel, const ldsc_snp* snps, uint64_t m,
                            const ldsc_l2_config* cfg, const double* annot, uint64_t K,
                            double* l2_out, double* maf_out, ldsc_l2_perf* perf,
                            ldsc_progress_cb cb, void* cb_ctx, char* err, uint64_t err_len);

/* ---------------------------------------------------------------------- */
/* Sub-steps exposed for bit-exact checks and for callers that need them.  */
/* ---------------------------------------------------------------------- */

/* Replaces get_block_lefts_by_chr (src/l2/window.rs:32-90).  Host code.
 * Returns 1 in *any_full_chr_window when some chromosome run lies in one
 * window (compute.rs:1059-1065). */
int ldsc_l2_block_left(const ldsc_snp* snps, uint64_t m, int32_t window_mode,
                       double window_value, uint64_t* block_left_out,
                       int32_t* any_full_chr_window, char* err, uint64_t err_len);

/* Per-SNP left partner bound lo[j]: pair (i<j) enters l2 iff lo[j] <= i.
 * Restates the chunk/window inclusion of compute.rs:1164-1174 (block
 * sizes), 1283-1299 (a_left eviction), 1855-1867 and 2244-2257 (masking).
 * Host code. */
int ldsc_l2_pair_bounds(const ldsc_snp* snps, uint64_t m, const ldsc_l2_config* cfg,
                        uint64_t* lo_out, char* err, uint64_t err_len);

/* Replaces decode_column (src/bed.rs:478-521) with the LUT [2, NaN, 1, 0]
 * (build_lut, src/bed.rs:470-476) and snp_stats_from_bytes
 * (src/l2/compute.rs:363-427) on device.  For the m SNPs at `bed_idx`:
 *   dosage   N x m int8 column-major, 0/1/2 and -1 for missing (nullable)
 *   miss     ceil(N/32) x m uint32 column-major missing bit masks (nullable)
 *   stats    m x 4 uint32 row-major: sum, count, sum_sq, het (nullable)
 *   std_f32  N x m float column-major standardized columns in the
 *            reference's f32 form (normalize.rs:115-154) (nullable) */
int ldsc_l2_decode(const ldsc_panel* panel, const uint64_t* bed_idx, uint64_t m,
                   int8_t* dosage, uint32_t* miss, uint32_t* stats, float* std_f32,
                   char* err, uint64_t err_len);

/* CountSketchProj::new (src/l2/compute.rs:179-192): fastrand 2.3.0 wyrand
 * stream seeded with `seed`; N bucket draws u32(..d) then N sign draws.
 * Host code. */
int ldsc_l2_countsketch_hash(uint64_t n, uint64_t d, uint64_t seed, uint32_t* bucket,
                             int8_t* sign, char* err, uint64_t err_len);

/* Integer CountSketch buckets of the fused decode->scatter
 * (countsketch_fused_project_f32, compute.rs:601-770 and
 * scatter_one_column_f32_scalar, wasm_simd.rs:1691-1784) on device:
 *   sg[b, j] = sum_{i: bucket(i)=b} sign_i * g_ij   (missing g = 0)
 *   sm[b, j] = sum_{i: bucket(i)=b} sign_i * [i present in j]
 * d x m int32 column-major each.  Uses the seed-42 hash of the reference. */
int ldsc_l2_sketch_buckets(const ldsc_panel* panel, const uint64_t* bed_idx, uint64_t m,
                           uint64_t d, int32_t* sg, int32_t* sm, char* err,
                           uint64_t err_len);

/* ---------------------------------------------------------------------- */
/* Serving: continuous batching of LD-score requests over resident panels. */
/* No counterpart in the reference (nearest: the per-chromosome Web-Worker  */
/* calls of ldsc-web/src/worker.rs:664-872 into compute_l2_from_bed).       */
/* ---------------------------------------------------------------------- */

typedef struct ldsc_server ldsc_server;

typedef struct {
    uint64_t hbm_budget_bytes; /* device bytes the workers' workspaces may  */
                               /* hold, buffers kept between requests       */
                               /* included (see ldsc_server_stats)          */
    uint32_t max_inflight;     /* requests resident on the device at once   */
    uint32_t queue_capacity;   /* pending requests accepted before BUSY     */
    uint32_t batch_window_us;  /* 0: one band launch per request.  Else the */
                               /* band tiles of requests prepared together  */
                               /* share one launch; a request waits at most */
                               /* this long for others still being prepared */
    uint32_t max_queue_ms;     /* 0: none.  Else a request still queued     */
                               /* after this long fails with TIMEOUT        */
} ldsc_server_config;

/* Counters of a server since it was created.  The budget is charged with
 * each worker's workspace: the device bytes it holds while idle, plus the
 * running request's estimate (an upper bound of what it can add) while busy.
 * hbm_charged is the current sum, hbm_charged_max its largest value; both
 * stay within hbm_budget_bytes. */
typedef struct {
    uint64_t requests_done;       /* finished, in error or cancelled too         */
    uint64_t band_launches;       /* band kernel launches, batched or not        */
    uint64_t batched_requests;    /* requests whose band shared a launch         */
    uint32_t max_batch;           /* most requests in one band launch            */
    uint32_t reserved;
    uint64_t requests_cancelled;
    uint64_t requests_timed_out;
    uint64_t hbm_charged;
    uint64_t hbm_charged_max;
    uint64_t workspace_trims;     /* times a worker freed its buffers            */
} ldsc_server_stats;

int ldsc_server_create(const ldsc_server_config* cfg, ldsc_server** out, char* err,
                       uint64_t err_len);
/* Finishes the running requests, abandons the queued ones and frees the
 * server.  No other call on the server may run during or after it. */
int ldsc_server_destroy(ldsc_server* server);
int ldsc_server_stats_get(ldsc_server* server, ldsc_server_stats* out);

/* Enqueue one request.  The snps/annot arrays are copied; l2_out/maf_out
 * must stay valid until the request is reported done by poll/wait. */
int ldsc_l2_submit(ldsc_server* server, const ldsc_panel* panel, const ldsc_snp* snps,
                   uint64_t m, const ldsc_l2_config* cfg, const double* annot, uint64_t K,
                   double* l2_out, double* maf_out, uint64_t* request_id, char* err,
                   uint64_t err_len);

/* LDSC_OK when done (latency_secs = arrival -> result), LDSC_ERR_PENDING
 * while queued or running, or the request's own error code. */
int ldsc_l2_poll(ldsc_server* server, uint64_t request_id, double* latency_secs,
                 char* err, uint64_t err_len);

/* Blocks until the request finishes; same return convention as poll. */
int ldsc_l2_wait(ldsc_server* server, uint64_t request_id, double* latency_secs,
                 char* err, uint64_t err_len);

/* Cancels a request.  A queued one is done at once; a running one stops at
 * the next phase boundary of its computation (after tiling and planning,
 * after the band contraction) once its device work has drained, or
 * finishes normally when it is past the last boundary.  poll / wait then
 * report LDSC_ERR_CANCELLED (or the request's own result); the contents of
 * l2_out / maf_out of a cancelled request are unspecified.  Returns
 * LDSC_ERR_NOT_FOUND for an unknown or already collected id.  Cancelling a
 * finished request changes nothing. */
int ldsc_l2_cancel(ldsc_server* server, uint64_t request_id, char* err, uint64_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* LDSC_L2_H */
