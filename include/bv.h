/*
 * bv.h -- C ABI of the MI355X Block Vecchia log-likelihood engine (libbv.so).
 *
 * Plain pointers and sizes only; every function is callable from C, from an
 * NLopt-style optimizer, from ctypes or from cgo.  All arithmetic is FP64.
 *
 * Reference interfaces these entry points replace (paths relative to the
 * blockVecchia repository):
 *   bv_create      the allocation + upload block of the driver
 *                  (src/test_Xvecchia_batch.cpp:459-689) that fills `llh_data`
 *                  (include/vecchia_helper_data.h:69-146)
 *   bv_llh /
 *   llh_Xvecchia_batch
 *                  T llh_Xvecchia_batch(unsigned n, const T *localtheta, T *grad,
 *                  void *f_data)   (include/llh_Xvecchia_batch.h:22-24), the
 *                  objective registered with nlopt_set_max_objective
 *                  (src/test_Xvecchia_batch.cpp:704)
 *   bv_destroy     the frees at src/test_Xvecchia_batch.cpp:759-815
 *   bv_partition   the host pre-processing of src/test_Xvecchia_batch.cpp:240-404
 *                  (k-means, block ordering, firstbatch, cluster_combine,
 *                  cluster_to_batch, findNearestPoints)
 *   bv_matern      core_dcmg (src/ckernel.cpp:1057-1093) on the device
 */
#ifndef BV_H
#define BV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define BV_OK              0
#define BV_ERR_NOT_PD      1   /* a block's covariance is not positive definite  */
#define BV_ERR_ARG        -1   /* invalid argument (sizes, pointers, theta)       */
#define BV_ERR_HIP        -2   /* a HIP runtime call failed                       */
#define BV_ERR_NO_DEVICE  -3   /* no usable gfx950 device                         */
#define BV_ERR_UNSUPPORTED -4  /* kernel / metric combination not implemented, or a
                                  fused factorization forced (BV_FACTOR_KERNEL or
                                  BV_PATH) on blocks of padded augmented size above
                                  1024 (the other forms take any size)            */
#define BV_ERR_BUSY       -5   /* serving queue full / workspace exhausted        */
#define BV_ERR_SCRATCH    -6   /* dataset too large for bv_llh_noalloc's scratch  */

/* ---- covariance kernels (ids of include/vecchia_helper_data.h:482-521) --- */
#define BV_KERNEL_MATERN            1  /* univariate_matern_stationary_no_nugget   */
#define BV_KERNEL_POWEXP            2  /* univariate_powexp_stationary_no_nugget:
                                          evaluates Matern, as the reference does */
#define BV_KERNEL_POWEXP_NUGGET     3  /* univariate_powexp_nugget_stationary: Matern,
                                          4 parameters, nugget never added        */
#define BV_KERNEL_SPACETIME         4  /* univariate_matern_spacetime_stationary  */

typedef struct bv_ctx bv_ctx;

/* One dataset, already partitioned and in block order (the arrays the
 * reference keeps in llh_data: locations_new, h_obs_new, batchNum,
 * batchNumAccum, and the conditioning sets).  The conditioning set of block i
 * (i >= 1) is given EITHER as neighbour indices into the block-ordered arrays
 * (nbr[i*cs + k], the kNN order of include/nearest_neighbor.h:115-147) OR as
 * copies (con_x/con_y/con_z/con_obs, B*cs each, the reference's locations_con
 * and h_obs_conditioning).  Row 0 is ignored: block 0 is evaluated as an
 * exact dense Gaussian likelihood, exactly as the reference's chain leaves it
 * (its conditioning correction is skipped, include/llh_Xvecchia_batch.h:407). */
typedef struct bv_problem {
    int64_t       n;               /* number of locations                      */
    int32_t       batch_count;     /* B: number of blocks                      */
    int32_t       cs;              /* conditioning size (0 = independent blocks)*/
    const int32_t *batch_num;      /* [B] block sizes                          */
    const double *x, *y, *z;       /* [n] block-ordered coordinates; z may be NULL */
    const double *obs;             /* [n] block-ordered observations           */
    const int32_t *nbr;            /* [B*cs] neighbour indices, or NULL        */
    const double *con_x, *con_y, *con_z, *con_obs; /* [B*cs] copies if nbr == NULL */
    int32_t       kernel;          /* BV_KERNEL_*                              */
    int32_t       distance_metric; /* 0 Euclidean, 1 great-circle (2-D only)   */
    int32_t       time_flag;       /* 1: use z (3-D / space-time)              */
    double        dist_scale;      /* divides the scaled distance (default 1)  */
    int32_t       device;          /* HIP device ordinal                       */
    int32_t       verbose;         /* 1: print the reference's per-iteration line */
} bv_problem;

/* Upload one dataset (HBM-resident until bv_destroy) and size the workspace.
 * Returns BV_OK and *out, or a negative status. */
int bv_create(const bv_problem *p, bv_ctx **out);

/* NLopt objective: log-likelihood at theta[0..n-1] (3 parameters for kernels
 * 1/2, 4 for kernel 3, 7 for kernel 4).  grad is ignored (derivative-free).
 * f_data is the bv_ctx*.  Returns -INFINITY when a block is not positive
 * definite (see bv_last_status) and NaN on argument/runtime errors. */
double bv_llh(unsigned n, const double *theta, double *grad, void *f_data);

/* Same symbol name and signature as the reference objective, for a drop-in
 * nlopt_set_max_objective(opt, llh_Xvecchia_batch, ctx). */
double llh_Xvecchia_batch(unsigned n, const double *theta, double *grad, void *f_data);

/* Evaluate k parameter vectors (row-major k x nparam) of the same dataset in
 * shared launches; out[k] receives the log-likelihoods, status[k] (optional)
 * the per-evaluation status. */
int bv_llh_many(bv_ctx *ctx, int k, int nparam, const double *thetas, double *out,
                int *status);

/* Per-block terms of the last evaluation: logdet[b] = sum_j log(Ls_jj^2) and
 * dot[b] = ||Ls^{-1}(y_B - mu)||^2 (the reference's logdet_result_h and
 * norm2_result_h^2, include/llh_Xvecchia_batch.h:468-478).  Either may be NULL. */
int bv_block_terms(bv_ctx *ctx, double *logdet, double *dot, int32_t *info);

/* Status of the last evaluation: BV_OK, BV_ERR_NOT_PD (first_bad_block set),
 * or a negative runtime status. */
int bv_last_status(bv_ctx *ctx, int *first_bad_block);

/* Device time in milliseconds of the last evaluation's kernels (HIP events). */
double bv_last_kernel_ms(bv_ctx *ctx);

/* The factorization configuration in use: cfg[0] = kernel variant (0: four
 * wavefronts per block, 1: one, 2: two), cfg[1] = 1 for the stepped factorization
 * (one launch per panel step over all blocks), 2 for the stepped factorization
 * with lookahead (the next panel's diagonal block factored one launch
 * early), 3 for the paired-panel factorization (one launch per pair of
 * panels, the next pair's 64x64 diagonal block factored one launch early),
 * 0 for the fused kernel, cfg[2] = block chunks, cfg[3] = streams the chunks
 * cycle over (0: one stream, chunk c's step j in the same launch as chunk
 * 0's step j + c), cfg[4] = 1 when single evaluations replay a HIP graph
 * (cfg holds 5 entries).  No counterpart in the reference.  The first
 * evaluation of a context times the candidate configurations on its own data
 * and keeps the fastest, then times graph replay against direct launches for
 * single evaluations (unless BV_FACTOR_KERNEL, BV_PATH or BV_CHUNKS fix the
 * configuration, or BV_AUTOTUNE=0); before that, returns BV_ERR_BUSY.  The
 * choice is cached per device and workload in a text file (BV_TUNE_CACHE,
 * default $HOME/.cache/blockvecchia_amd/autotune.txt; BV_TUNE_CACHE=0 turns
 * it off), so later contexts on the machine take it without timing.  Blocks
 * of padded augmented size above 1024 rule out the fused kernels (cfg[1] = 0);
 * without autotuning such a dataset uses the lookahead form. */
int bv_factor_config(bv_ctx *ctx, int32_t *cfg);

/* The autotuner's candidate configurations: writes up to max_count rows of
 * 4 entries (cfg[0..3] of bv_factor_config) and returns the candidate count.
 * No counterpart in the reference; for tests that pin every candidate. */
int bv_factor_candidates(int32_t *cfgs, int max_count);

/* Single evaluations (bv_llh, bv_llh_many with k = 1) may replay a HIP graph
 * of their kernel launches, captured on the first such call after the
 * factorization configuration is fixed and again whenever the configuration
 * or the buffers change; the autotuner keeps it only where it times faster
 * (many launches: the stepped factorization over several streams).
 * enable = 0 launches them one by one (as calls with k > 1 always do); the
 * results are the same.  Default on; BV_GRAPH=0 in the environment at
 * bv_create turns it off.  No counterpart in the reference. */
int bv_set_graph(bv_ctx *ctx, int enable);

/* Number of iterations (objective calls) so far; the reference's
 * data->iterations counter. */
int bv_iterations(bv_ctx *ctx);

void bv_destroy(bv_ctx *ctx);

/* ---- request server ---------------------------------------------------
 * No counterpart in the reference, which evaluates one theta at a time for
 * one NLopt session.  For callers that serve many optimizer sessions on
 * different datasets: each dataset is a bv_ctx from bv_create (resident in
 * HBM); requests queue in arrival order; bv_server_poll admits them into idle
 * lanes (one stream each) while their workspace bytes fit the lane budget,
 * merges their blocks into shared launches, and returns finished results.
 * Every evaluation runs the factorization form its dataset's autotuner chose
 * (a dataset not yet tuned is tuned synchronously on its first request), so
 * a served result equals bv_llh on the same context and parameters bit for
 * bit, whatever batch it ran in.  Merged blocks are grouped by covariance
 * kind and form: fused-form blocks share launches with one wavefront per
 * block (BV_SERVER_1W_MAXMP=M in the environment at creation limits that to
 * groups of padded size <= M; larger ones use four wavefronts, with the same
 * results), each stepped form gets its own step launches over its blocks.
 * A request whose dataset is tuned to a stepped form and needs at least
 * BV_SERVER_OWN_BYTES (default 2e8) of workspace runs alone on the
 * dataset's own launches and stream, one at a time per dataset; it counts as
 * the context's latest evaluation for bv_block_terms.  Calls on one server
 * are not thread-safe. */
typedef struct bv_server bv_server;

/* nlanes concurrent batches (1..16), each with ws_bytes_per_lane of
 * workspace (a larger single request still runs, alone) and at most
 * max_evals_per_batch evaluations. */
int bv_server_create(int device, int nlanes, double ws_bytes_per_lane,
                     int max_evals_per_batch, bv_server **out);

/* Register a dataset (the context stays owned by the caller and must outlive
 * the server's use of it).  Idle lanes are sized here for a full batch
 * (ws_bytes_per_lane of workspace, max_evals_per_batch evaluations of the
 * largest attached dataset), so that serving does not grow them later. */
int bv_server_attach(bv_server *s, bv_ctx *dataset, int *dataset_id);

/* Queue one evaluation; tag is returned with its result. */
int bv_server_submit(bv_server *s, int dataset_id, int nparam, const double *theta,
                     int64_t tag);

/* Collect finished evaluations (up to max_out: tags, llh, status) and start
 * queued ones on idle lanes.  wait != 0 blocks until at least one result is
 * ready (or nothing is queued or running).  Returns the number of results,
 * or a negative status. */
int bv_server_poll(bv_server *s, int wait, int max_out, int64_t *tags, double *llh,
                   int *status);

/* Requests queued or running. */
int bv_server_pending(bv_server *s);

void bv_server_destroy(bv_server *s);

/* Draw y (n values, block order) from the Block Vecchia model of the
 * context's partition at theta: block 0 from its own covariance, each block
 * i >= 1 from y_B | y_N(i), in dependency levels on the device.  z: n
 * standard normal deviates (block order); nugget >= 0 is added to every
 * covariance diagonal (the likelihood never adds one; a small value such as
 * 1e-12 sigma2 keeps smooth kernels on clustered points numerically positive
 * definite).  *levels (optional) receives the number of levels.  Needs
 * neighbour indices (not copies) and blocks of padded size <= 1024.  Returns
 * BV_OK, BV_ERR_NOT_PD (a block's covariance failed: y is incomplete) or an
 * error.  No counterpart in the reference, which reads pre-simulated data
 * (src/test_Xvecchia_batch.cpp:182-237); it generates the inputs of the
 * large configurations. */
int bv_simulate(bv_ctx *ctx, unsigned n, const double *theta, const double *z, double nugget,
                double *y, int *levels);

/* ---- evaluation without device allocation -------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): one log-likelihood evaluation, as bv_llh computes it with
 * the fused factorization (four wavefronts per block; equal bit for bit to
 * bv_llh on a context fixed to that form), that calls no hipMalloc or
 * hipFree.  The dataset stays the caller's: in *p, x, y, z, obs and nbr (or
 * con_x, con_y, con_z, con_obs) are DEVICE pointers on p->device, while
 * batch_num is a HOST array.  Everything else the kernels need lives in a
 * fixed device array of BV_NOALLOC_SCRATCH_BYTES per device, held from the
 * library's load on a device until the process ends.  A dataset whose
 * descriptors, per-block terms and factor workspace do not fit returns
 * BV_ERR_SCRATCH: 128 MiB hold about 3,100 blocks at padded size 96 (cs = 60
 * with 13-point blocks: n ~ 40,000), 840 at 192 (cs = 150), 310 at 320
 * (cs = 300); the n = 2,000, cs = 60 workload takes 4.2 MB.  Blocks of padded
 * augmented size above 1024 return BV_ERR_UNSUPPORTED.  Runs on
 * the device's null stream and returns when the result is on the host.
 * Calls that overlap on one device share the array and race: serialise them.
 * Returns BV_OK, BV_ERR_NOT_PD (*llh = -INFINITY, *first_bad_block set) or an
 * error. */
#define BV_NOALLOC_SCRATCH_BYTES (128ull << 20)
int bv_llh_noalloc(const bv_problem *p, unsigned n, const double *theta, double *llh,
                   int *first_bad_block);

/* Dense Matern block A[i + j*m] = K(l1[i], l2[j]) computed on the device with
 * the engine's covariance code (core_dcmg, src/ckernel.cpp:1057-1093).
 * Host arrays in, host array out.  Used for entrywise parity tests. */
int bv_matern(int m, int nn, const double *x1, const double *y1, const double *z1,
              const double *x2, const double *y2, const double *z2,
              const double *theta, int kernel, int distance_metric, double dist_scale,
              double *A, int device);

/* Result of the host pre-processing; arrays are allocated by the library and
 * released with bv_blocks_free. */
typedef struct bv_blocks {
    int64_t  n;
    int32_t  batch_count;     /* B = nclusters - firstClusterCount + 1          */
    int32_t  cs;
    int32_t  first_cluster_count; /* clusters merged into block 0 (firstbatch) */
    int32_t  nclusters;
    double  *x, *y, *z, *obs; /* [n] block-ordered (z NULL in 2-D)             */
    int32_t *batch_num;       /* [B]                                           */
    int32_t *nbr;             /* [B*cs] kNN indices into the block order;
                                 row 0 = 0..cs-1 (the reference's dummy set)   */
    int64_t *perm;            /* [n] original index of each block-ordered point */
    double  *cx, *cy, *cz;    /* [nclusters] centroids in block order (after the
                                 random permutation); block i uses entry
                                 i + first_cluster_count - 1                   */
} bv_blocks;

/* Host pre-processing: partition n points into blocks and pick conditioning
 * sets, as src/test_Xvecchia_batch.cpp:240-404 does: k-means (classic Vecchia
 * when n < 2 bc), random centroid order (srand(42) Fisher-Yates), reorderIndex,
 * firstbatch, cluster_combine, cluster_to_batch and findNearestPoints.  The
 * k-means is a deterministic restatement (zeroed accumulators, z = 0 in 2-D,
 * sums in point order).  kmeans_epochs: 50 in the reference; at least 1
 * when k-means runs (n >= 2 bc), else BV_ERR_ARG.  threads: host threads
 * (<= 0: all).  Returns BV_OK and fills *out. */
int bv_partition(int64_t n, const double *x, const double *y, const double *z,
                 const double *obs, int32_t bc, int32_t cs, int32_t seed,
                 int32_t kmeans_epochs, int32_t time_flag, int32_t distance_metric,
                 int32_t threads, bv_blocks *out);

/* bv_partition with the k-means assignment step (every point against every
 * centroid, each epoch) on GPU `device`; everything else as bv_partition, on
 * the host.  The output is bv_partition's, bit for bit: the device uses the
 * host's rounding and tie rule.  No counterpart in the reference, whose
 * k-means runs on the host (src/kmeans.cpp:59-145); for large n it dominates
 * the set-up. */
int bv_partition_gpu(int32_t device, int64_t n, const double *x, const double *y,
                     const double *z, const double *obs, int32_t bc, int32_t cs, int32_t seed,
                     int32_t kmeans_epochs, int32_t time_flag, int32_t distance_metric,
                     int32_t threads, bv_blocks *out);

void bv_blocks_free(bv_blocks *b);

/* Checks that the library's matrix-core lane layout assumption holds on the
 * current device (one 16x16x4 FP64 product against a host product). */
int bv_device_selftest(int device);

/* FP64 pipe probe of the device (measurement of the kernels' limiter; no
 * counterpart in the reference).  One workgroup per compute unit; each
 * wavefront times 2000 repetitions of a step; cycles[] receives the median
 * cycles per step:
 *   [0] 2 independent 16x16x4 FP64 matrix-core products,
 *   [1] 8 independent FP64 fused multiply-adds (64 lanes),
 *   [2] both in one wavefront,
 *   [3] 2 products and 32 multiply-adds in one wavefront,
 *   [4], [5] body [2] split over two wavefronts of one SIMD: the product
 *            wavefront, the multiply-add wavefront.
 * Measured on MI355X: [0] 128, [1] 35, [2] 169, [3] 289, [4] 128, [5] 35
 * cycles: within one wavefront the multiply-adds add to the products' time,
 * from two wavefronts of one SIMD they overlap completely. */
int bv_fp64_pipe_probe(int device, double *cycles);

/* Wavefront trace of the factorization kernels, for launch-occupancy
 * analysis.  In a library built with BV_TRACE (`python -m
 * blockvecchia_amd.build --trace` writes lib/libbv_trace.so), every wavefront
 * of the fused, stepped, lookahead and paired factorization kernels appends
 * one record of three 64-bit words when it leaves the kernel: its start and
 * end on the device's 100 MHz real-time counter, and (kernel << 56) |
 * ((step + 1) << 48) | (workgroup << 8) | wavefront, kernel 0 fused, 1
 * stepped, 2 lookahead, 3 paired; at most 4,194,304 records per device.
 * bv_trace_reset empties the device's trace; bv_trace_read copies up to
 * `capacity` records and sets *count to the number appended since the reset
 * (which may exceed what was kept).  Other builds return BV_ERR_UNSUPPORTED.
 * No counterpart in the reference. */
int bv_trace_reset(int device);
int bv_trace_read(int device, uint64_t *records, int64_t capacity, int64_t *count);

#ifdef __cplusplus
}
#endif
#endif /* BV_H */
