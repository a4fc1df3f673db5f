/*
 * bv_api.hip -- host side of libbv: dataset upload, evaluation launches and
 * the C ABI of include/bv.h.
 *
 * One bv_ctx holds one dataset resident in HBM (block-ordered locations,
 * observations, block sizes, conditioning indices or copies), the launch
 * descriptors of its blocks, and a workspace for the factor tiles of every
 * block of up to `cap_evals` simultaneous evaluations.  An evaluation is two
 * launches on the context's stream (bv_block_kernel over all blocks, then
 * bv_reduce_kernel) and one 32-byte device-to-host copy of the result.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "bv.h"
#include "bv_prep.h"
#include "bv_engine.h"
#include "bv_device.h"

hipError_t bv_launch_block(int kind, unsigned nblk, size_t lds, hipStream_t st,
                           const bv_block_desc *desc, const bv_eval_slot *slots, double *ws);
hipError_t bv_launch_block_2w(int kind, unsigned nblk, size_t lds, hipStream_t st,
                              const bv_block_desc *desc, const bv_eval_slot *slots, double *ws);
hipError_t bv_launch_block_1w(int kind, unsigned nblk, size_t lds, hipStream_t st,
                              const bv_block_desc *desc, const bv_eval_slot *slots, double *ws);
template <int KIND>
hipError_t bv_launch_res_k(int nr, unsigned nblk, hipStream_t st, const bv_block_desc *desc,
                           const bv_eval_slot *slots, double *ws);

/* The register-resident kernels (bv_resident.hip, one object per kind). */
static hipError_t bv_launch_res(int kind, int nr, unsigned nblk, hipStream_t st,
                                const bv_block_desc *desc, const bv_eval_slot *slots,
                                double *ws) {
    switch (kind) {
    case BV_KIND_NU05: return bv_launch_res_k<BV_KIND_NU05>(nr, nblk, st, desc, slots, ws);
    case BV_KIND_NU15: return bv_launch_res_k<BV_KIND_NU15>(nr, nblk, st, desc, slots, ws);
    case BV_KIND_NU25: return bv_launch_res_k<BV_KIND_NU25>(nr, nblk, st, desc, slots, ws);
    case BV_KIND_GENERAL: return bv_launch_res_k<BV_KIND_GENERAL>(nr, nblk, st, desc, slots, ws);
    default: return bv_launch_res_k<BV_KIND_LOADA>(nr, nblk, st, desc, slots, ws);
    }
}
hipError_t bv_launch_gen(int kind, unsigned max_tiles, unsigned nblk, hipStream_t st,
                         const bv_block_desc *desc, const bv_eval_slot *slots, double *ws);
hipError_t bv_launch_reduce(unsigned nevals, hipStream_t st, const bv_eval_slot *slots);
hipError_t bv_launch_cheb_build(unsigned nslots, hipStream_t st, const bv_eval_slot *slots);
hipError_t bv_launch_step(int kind, unsigned nitems, hipStream_t st,
                          const bv_step_item *items, const bv_block_desc *desc,
                          const bv_eval_slot *slots, double *ws, double *part_logdet,
                          double *part_dot, int32_t *part_info, int pmax);
hipError_t bv_launch_step2(int kind, unsigned nitems, hipStream_t st,
                           const bv_step_item *items, const bv_block_desc *desc,
                           const bv_eval_slot *slots, double *ws, double *part_logdet,
                           double *part_dot, int32_t *part_info, int pmax);
hipError_t bv_launch_pair(int kind, unsigned nitems, hipStream_t st,
                          const bv_step_item *items, const bv_block_desc *desc,
                          const bv_eval_slot *slots, double *ws, double *part_logdet,
                          double *part_dot, int32_t *part_info, int pmax);
hipError_t bv_launch_step_parts(int ndesc, hipStream_t st, const bv_block_desc *desc,
                                const bv_eval_slot *slots, const double *part_logdet,
                                const double *part_dot, const int32_t *part_info, int pmax);
hipError_t bv_launch_matern(int m, int n, const double *x1, const double *y1, const double *z1,
                            const double *x2, const double *y2, const double *z2,
                            const bv_kparams &kp, int kind, double *A);
hipError_t bv_launch_selftest(const double *a, const double *b, double *d);
hipError_t bv_launch_pipe_probe(int mode, int ncu, int iters, double *out,
                                unsigned long long *cyc);
hipError_t bv_launch_sim_finish(unsigned nblk, hipStream_t st, const bv_block_desc *desc,
                                const bv_eval_slot *slots, const double *ws, const double *z);
hipError_t bv_launch_kmeans_assign(int64_t n, const double *px, const double *py, const double *pz,
                                   int k, const double *cx, const double *cy, const double *cz,
                                   int32_t *cluster, hipStream_t st);
hipError_t bv_launch_knn(const double *x, const double *y, const double *z, const int64_t *accum,
                         const double *bcx, const double *bcy, const double *bcz, int64_t B,
                         int k, int time_flag, int32_t *nbr, int32_t *overflow, hipStream_t st);

#define BV_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            fprintf(stderr, "libbv: %s failed at %s:%d: %s\n", #call, __FILE__, __LINE__, \
                    hipGetErrorString(e_));                                             \
            return BV_ERR_HIP;                                                          \
        }                                                                               \
    } while (0)

struct bv_ctx {
    int device = 0;
    int64_t n = 0;
    int32_t B = 0, cs = 0, kernel = 1, metric = 0, time_flag = 0, verbose = 0;
    double dist_scale = 1.0;
    std::vector<int32_t> h_batch_num;
    std::vector<int64_t> h_accum;
    /* device dataset */
    double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr, *d_obs = nullptr;
    int32_t *d_nbr = nullptr;
    double *d_cx = nullptr, *d_cy = nullptr, *d_cz = nullptr, *d_cobs = nullptr;
    int32_t *d_batch_num = nullptr;
    /* launch layout of one evaluation */
    std::vector<bv_block_desc> base_desc;   /* sorted by padded size, descending */
    int64_t ws_per_eval = 0;                /* doubles                            */
    int32_t max_mp = 0;
    /* per-capacity buffers */
    int cap_evals = 0;
    bv_block_desc *d_desc = nullptr;
    int desc_evals = 0;                     /* evaluations d_desc is laid out for */
    bv_eval_slot *d_slots = nullptr, *h_slots = nullptr;
    double *d_ws = nullptr;
    double *d_logdet = nullptr, *d_dot = nullptr, *d_result = nullptr, *h_result = nullptr;
    double *d_cheb = nullptr;               /* general-nu tables, one per evaluation */
    int32_t *d_info = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    /* state of the last evaluation */
    int last_status = BV_OK;
    int last_bad = -1;
    int last_k = 0;
    double last_ms = 0.0;
    int iterations = 0;
    int variant = 0; /* factorization kernel: wavefronts per block 0: four, 1: one, 2: two;
                        3: register-resident (bv_resident.hip) */
    int stepped = 0; /* one launch per panel step over all blocks: 1, or 2 with lookahead;
                        3: one launch per pair of panels with lookahead */
    /* stepped factorization: work items of every step for desc_evals evaluations */
    bv_step_item *d_items = nullptr;
    std::vector<int64_t> step_off;          /* [chunk][pmax + 1] item offsets      */
    int nchunks = 1, nstreams = 1;          /* block chunks, and streams they cycle over */
    int look_first = 1;                     /* lookahead items dispatched first */
    int srv_busy = 0;                       /* a server lane runs an evaluation on it */
    int chunk_split = 1;                    /* chunks: contiguous by cost (1) or round-robin */
    int autotune = 1, tuned = 0;            /* choose the configuration on first use */
    int fused_ok = 1;                       /* every block fits the fused kernels */
    int res_ok = 0;                         /* ... and the register-resident one */
    hipStream_t cstream[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t cev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int pmax = 0;
    double *d_part_logdet = nullptr, *d_part_dot = nullptr;
    int32_t *d_part_info = nullptr;
    /* single evaluations replay a captured graph of their launches when
       allowed (use_graph) and, once tuned, when that timed faster */
    int use_graph = 1, graph_wins = 1;
    hipGraphExec_t graph = nullptr;
    int64_t graph_key[12] = {};
};

static int bv_nparams(int kernel) {
    return kernel == BV_KERNEL_SPACETIME ? 7 : (kernel == BV_KERNEL_POWEXP_NUGGET ? 4 : 3);
}

template <class T>
static int bv_upload(T **dst, const T *src, size_t count, hipStream_t s) {
    if (count == 0) { *dst = nullptr; return BV_OK; }
    BV_HIP(hipMalloc((void **)dst, count * sizeof(T)));
    BV_HIP(hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    return BV_OK;
}

static void bv_free_eval_buffers(bv_ctx *c) {
    if (c->d_desc) (void)hipFree(c->d_desc);
    if (c->d_slots) (void)hipFree(c->d_slots);
    if (c->h_slots) (void)hipHostFree(c->h_slots);
    if (c->d_ws) (void)hipFree(c->d_ws);
    if (c->d_logdet) (void)hipFree(c->d_logdet);
    if (c->d_dot) (void)hipFree(c->d_dot);
    if (c->d_info) (void)hipFree(c->d_info);
    if (c->d_result) (void)hipFree(c->d_result);
    if (c->h_result) (void)hipHostFree(c->h_result);
    if (c->d_items) (void)hipFree(c->d_items);
    if (c->d_part_logdet) (void)hipFree(c->d_part_logdet);
    if (c->d_part_dot) (void)hipFree(c->d_part_dot);
    if (c->d_part_info) (void)hipFree(c->d_part_info);
    if (c->d_cheb) (void)hipFree(c->d_cheb);
    c->d_cheb = nullptr;
    c->d_items = nullptr; c->d_part_logdet = nullptr; c->d_part_dot = nullptr;
    c->d_part_info = nullptr;
    c->d_desc = nullptr; c->d_slots = nullptr; c->h_slots = nullptr; c->d_ws = nullptr;
    c->d_logdet = nullptr; c->d_dot = nullptr; c->d_info = nullptr;
    c->d_result = nullptr; c->h_result = nullptr;
    c->cap_evals = 0; c->desc_evals = 0;
}

/* Make room for k simultaneous evaluations. */
static int bv_reserve(bv_ctx *c, int k) {
    if (k <= c->cap_evals) return BV_OK;
    bv_free_eval_buffers(c);
    const size_t B = (size_t)c->B;
    BV_HIP(hipMalloc((void **)&c->d_desc, sizeof(bv_block_desc) * B * k));
    BV_HIP(hipMalloc((void **)&c->d_slots, sizeof(bv_eval_slot) * k));
    BV_HIP(hipHostMalloc((void **)&c->h_slots, sizeof(bv_eval_slot) * k, hipHostMallocDefault));
    BV_HIP(hipMalloc((void **)&c->d_ws, sizeof(double) * (size_t)c->ws_per_eval * k));
    BV_HIP(hipMalloc((void **)&c->d_logdet, sizeof(double) * B * k));
    BV_HIP(hipMalloc((void **)&c->d_dot, sizeof(double) * B * k));
    BV_HIP(hipMalloc((void **)&c->d_info, sizeof(int32_t) * B * k));
    BV_HIP(hipMalloc((void **)&c->d_result, sizeof(double) * 4 * k));
    BV_HIP(hipHostMalloc((void **)&c->h_result, sizeof(double) * 4 * k, hipHostMallocDefault));
    c->pmax = c->max_mp / BV_NB;
    BV_HIP(hipMalloc((void **)&c->d_part_logdet, sizeof(double) * B * k * c->pmax));
    BV_HIP(hipMalloc((void **)&c->d_part_dot, sizeof(double) * B * k * c->pmax));
    BV_HIP(hipMalloc((void **)&c->d_part_info, sizeof(int32_t) * B * k * c->pmax));
    BV_HIP(hipMalloc((void **)&c->d_cheb, sizeof(double) * BV_CHEB_SIZE * k));
    c->cap_evals = k;
    return BV_OK;
}

/* Work items of the stepped factorization `stepped` (1: one launch per panel
   step, 2: with lookahead, 3: paired panels) over the descriptors all[0..nd).
   Per step j, for every descriptor with more than j panels, its rows 2j..T-1
   in items of BV_STEP_ROWS row tiles (paired: rows 4j.. in items of four; the
   lookahead forms' launch index 0 holds the first diagonal block alone).
   Workgroups are dealt to the XCDs round-robin (workgroup w runs on XCD
   w % 8), so the items of one descriptor get ids of one residue mod 8: they
   share the panel-row and W tiles through that XCD's L2.  Descriptors go to
   the least loaded XCD; shorter queues are padded with empty items.  Items
   that factor a diagonal block ahead (lookahead forms) are dispatched before
   every other item of their launch when look_first is set, so that their
   serial factorizations start first.  Chunks: contiguous ranges of equal
   factorization cost (Mp^3) over the size-sorted list (chunk_split) or
   round-robin; with skew, launch L holds chunk c's items of step L - c (the
   chunks' steps overlap inside each launch).  step_off[ch * (pmax + 1) + j]
   is the first item of chunk ch's launch j (skewed: step_off[L]).  Which
   launch or stream runs an item does not change its arithmetic: a block's
   result is the same under any chunking. */
static void bv_build_items(const bv_block_desc *all, size_t nd, int stepped, int NC, bool skew,
                           int chunk_split, int look_first, int pmax,
                           std::vector<bv_step_item> &items, std::vector<int64_t> &step_off) {
    const int NX = 8;
    typedef std::vector<std::vector<bv_step_item>> bv_queues;
    items.clear();
    std::vector<int> chunk_of(nd, 0);
    {
        double tot = 0.0, acc = 0.0;
        for (size_t i = 0; i < nd; ++i) tot += (double)all[i].mp * all[i].mp * all[i].mp;
        for (size_t i = 0; i < nd; ++i) {
            const double cst = (double)all[i].mp * all[i].mp * all[i].mp;
            if (chunk_split)
                chunk_of[i] = std::min(NC - 1, (int)((acc + 0.5 * cst) / tot * NC));
            else
                chunk_of[i] = (int)(i % (size_t)NC);
            acc += cst;
        }
    }
    auto build = [&](int ch, int j, bv_queues &qa, bv_queues &qb) {
        for (size_t i = 0; i < nd; ++i) {
            if (chunk_of[i] != ch) continue;
            const int T = all[i].mp / BV_TILE, P = all[i].mp / BV_NB;
            const int rows = (stepped == 3) ? 4 : BV_STEP_ROWS;
            const int r0 = (stepped == 3) ? 4 * j : 2 * j;
            if (stepped == 3 ? r0 >= T : j >= P) continue;
            size_t x = 0;
            for (int t = 1; t < NX; ++t)
                if (qa[t].size() + qb[t].size() < qa[x].size() + qb[x].size()) x = t;
            const int rend = (stepped >= 2 && j == 0) ? std::min(stepped == 3 ? 4 : 2, T) : T;
            for (int r = r0; r < rend; r += rows) {
                bv_step_item it;
                it.desc = (int32_t)i;
                it.row0 = r;
                it.nrows = std::min(rows, rend - r);
                it.step = (stepped >= 2) ? j - 1 : j;
                const bool ahead = look_first && stepped >= 2 && r == r0;
                (ahead ? qa : qb)[x].push_back(it);
            }
        }
    };
    auto emit = [&](const bv_queues &q) {
        size_t len = 0;
        for (int t = 0; t < NX; ++t) len = std::max(len, q[t].size());
        for (size_t w = 0; w < len * NX; ++w) {
            const std::vector<bv_step_item> &qq = q[w % NX];
            const size_t k2 = w / NX;
            if (k2 < qq.size()) items.push_back(qq[k2]);
            else {
                bv_step_item e;
                e.desc = -1; e.row0 = 0; e.nrows = 0; e.step = 0;
                items.push_back(e);
            }
        }
    };
    step_off.assign((size_t)NC * (pmax + 1), 0);
    if (skew) {
        const int nl = pmax + NC - 1; /* <= NC (pmax + 1) - 1 offsets */
        for (int L = 0; L < nl; ++L) {
            step_off[L] = (int64_t)items.size();
            bv_queues qa(NX), qb(NX);
            for (int ch = 0; ch < NC; ++ch)
                if (L - ch >= 0 && L - ch < pmax) build(ch, L - ch, qa, qb);
            emit(qa);
            emit(qb);
        }
        step_off[nl] = (int64_t)items.size();
    } else {
        for (int ch = 0; ch < NC; ++ch) {
            for (int j = 0; j < pmax; ++j) {
                step_off[(size_t)ch * (pmax + 1) + j] = (int64_t)items.size();
                bv_queues qa(NX), qb(NX);
                build(ch, j, qa, qb);
                emit(qa);
                emit(qb);
            }
            step_off[(size_t)ch * (pmax + 1) + pmax] = (int64_t)items.size();
        }
    }
}

/* Lay out the descriptors of k evaluations, interleaved so that the largest
   blocks of every evaluation start first, and the step items of the
   configured stepped form (bv_build_items). */
static int bv_layout(bv_ctx *c, int k) {
    if (c->desc_evals == k) return BV_OK;
    std::vector<bv_block_desc> all;
    all.reserve((size_t)c->B * k);
    for (const bv_block_desc &d0 : c->base_desc)
        for (int e = 0; e < k; ++e) {
            bv_block_desc d = d0;
            d.eval = e;
            d.ws_off = d0.ws_off + (int64_t)e * c->ws_per_eval;
            all.push_back(d);
        }
    BV_HIP(hipMemcpyAsync(c->d_desc, all.data(), sizeof(bv_block_desc) * all.size(),
                          hipMemcpyHostToDevice, c->stream));
    std::vector<bv_step_item> items;
    const bool skew = c->nstreams == 0 && c->nchunks > 1;
    bv_build_items(all.data(), all.size(), c->stepped, c->nchunks, skew, c->chunk_split,
                   c->look_first, c->pmax, items, c->step_off);
    if (c->d_items) (void)hipFree(c->d_items);
    c->d_items = nullptr;
    if (!items.empty()) {
        BV_HIP(hipMalloc((void **)&c->d_items, sizeof(bv_step_item) * items.size()));
        BV_HIP(hipMemcpyAsync(c->d_items, items.data(), sizeof(bv_step_item) * items.size(),
                              hipMemcpyHostToDevice, c->stream));
    }
    BV_HIP(hipStreamSynchronize(c->stream));
    c->desc_evals = k;
    return BV_OK;
}

/* Host preparation of the covariance parameters, as core_dcmg computes its
   constant (src/ckernel.cpp:1070-1074). */
static int bv_make_kparams(const bv_ctx *c, const double *theta, bv_kparams *kp, int *kind) {
    const int np = bv_nparams(c->kernel);
    for (int i = 0; i < np; ++i)
        if (!isfinite(theta[i])) return BV_ERR_ARG;
    memset(kp, 0, sizeof(*kp));
    kp->sigma2 = theta[0];
    kp->beta = theta[1];
    kp->nu = theta[2];
    double con = pow(2, (theta[2] - 1)) * tgamma(theta[2]);
    con = 1.0 / con;
    kp->con = theta[0] * con;
    kp->dist_scale = c->dist_scale;
    kp->inv_scale = 1.0 / (theta[1] * c->dist_scale);
    kp->metric = c->metric;
    kp->use_z = (c->time_flag && c->d_z) ? 1 : 0;
    /* the reference's box bounds are < 100 (include/vecchia_helper_data.h:684-700);
       the K_nu order recurrence runs nu + 1/2 steps */
    if (!(theta[2] <= 100.0)) return BV_ERR_ARG;
    if (c->kernel == BV_KERNEL_SPACETIME) {
        kp->beta_t = theta[3]; kp->nu_t = theta[4]; kp->sep = theta[5]; kp->aux = theta[6];
        *kind = BV_KIND_SPACETIME;
        if (!c->d_z) return BV_ERR_ARG;
        return BV_OK;
    }
    if (!(theta[1] > 0.0) || !(theta[2] > 0.0)) return BV_ERR_ARG;
    if (theta[2] == 0.5) *kind = BV_KIND_NU05;
    else if (theta[2] == 1.5) *kind = BV_KIND_NU15;
    else if (theta[2] == 2.5) *kind = BV_KIND_NU25;
    else *kind = BV_KIND_GENERAL;
    return BV_OK;
}

/* The side streams (and their events) of the chunked stepped factorization,
   created on first use. */
static int bv_side_streams(bv_ctx *c) {
    for (int i = 0; i < c->nstreams; ++i)
        if (!c->cstream[i] && hipStreamCreateWithFlags(&c->cstream[i], hipStreamNonBlocking) != hipSuccess)
            return BV_ERR_HIP;
    for (int i = 0; i <= c->nstreams; ++i)
        if (!c->cev[i] && hipEventCreateWithFlags(&c->cev[i], hipEventDisableTiming) != hipSuccess)
            return BV_ERR_HIP;
    return BV_OK;
}

/* one side stream and its fork/join events for the register-resident form */
static int bv_res_streams(bv_ctx *c) {
    if (!c->cstream[0] && hipStreamCreateWithFlags(&c->cstream[0], hipStreamNonBlocking) != hipSuccess)
        return BV_ERR_HIP;
    for (int i = 0; i < 2; ++i)
        if (!c->cev[i] && hipEventCreateWithFlags(&c->cev[i], hipEventDisableTiming) != hipSuccess)
            return BV_ERR_HIP;
    return BV_OK;
}

extern "C" int bv_create(const bv_problem *p, bv_ctx **out) {
    if (!p || !out) return BV_ERR_ARG;
    *out = nullptr;
    if (p->n <= 0 || p->batch_count <= 0 || p->cs < 0 || !p->batch_num || !p->x || !p->y ||
        !p->obs)
        return BV_ERR_ARG;
    if (p->kernel < 1 || p->kernel > 4) return BV_ERR_ARG;
    if ((p->time_flag || p->kernel == BV_KERNEL_SPACETIME) && !p->z) return BV_ERR_ARG;
    if (p->distance_metric == 1 && p->time_flag) return BV_ERR_UNSUPPORTED;
    if (p->cs > 0 && p->batch_count > 1 && !p->nbr && !(p->con_x && p->con_y && p->con_obs))
        return BV_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || p->device >= ndev)
        return BV_ERR_NO_DEVICE;
    BV_HIP(hipSetDevice(p->device));

    bv_ctx *c = new bv_ctx();
    c->device = p->device;
    c->n = p->n;
    c->B = p->batch_count;
    c->cs = p->cs;
    c->kernel = p->kernel;
    c->metric = p->distance_metric;
    c->time_flag = p->time_flag || p->kernel == BV_KERNEL_SPACETIME;
    c->dist_scale = p->dist_scale > 0 ? p->dist_scale : 1.0;
    c->verbose = p->verbose;
    /* factorization kernel: BV_FACTOR_KERNEL=1w selects the one-wavefront one */
    if (const char *v = getenv("BV_FACTOR_KERNEL"))
        c->variant = (strcmp(v, "1w") == 0)    ? 1
                     : (strcmp(v, "2w") == 0)  ? 2
                     : (strcmp(v, "res") == 0) ? 3
                                               : 0;
    if (const char *v = getenv("BV_PATH"))
        c->stepped = (strcmp(v, "stepped") == 0)    ? 1
                     : (strcmp(v, "stepped2") == 0) ? 2
                     : (strcmp(v, "paired") == 0)   ? 3
                                                    : 0;
    if (const char *v = getenv("BV_CHUNKS")) c->nchunks = std::max(1, std::min(64, atoi(v)));
    if (const char *v = getenv("BV_CHUNK_STREAMS")) c->nstreams = std::max(0, std::min(4, atoi(v)));
    if (getenv("BV_FACTOR_KERNEL") || getenv("BV_PATH") || getenv("BV_CHUNKS")) c->autotune = 0;
    if (const char *v = getenv("BV_AUTOTUNE")) c->autotune = c->autotune && atoi(v) != 0;
    if (const char *v = getenv("BV_GRAPH")) c->use_graph = atoi(v) != 0;
    if (const char *v = getenv("BV_LOOK_FIRST")) c->look_first = atoi(v) != 0;
    if (const char *v = getenv("BV_CHUNK_SPLIT")) c->chunk_split = atoi(v) != 0;
    c->h_batch_num.assign(p->batch_num, p->batch_num + p->batch_count);
    c->h_accum.assign((size_t)c->B + 1, 0);
    for (int i = 0; i < c->B; ++i) {
        if (p->batch_num[i] <= 0) { delete c; return BV_ERR_ARG; }
        c->h_accum[i + 1] = c->h_accum[i] + p->batch_num[i];
    }
    if (c->h_accum[c->B] != c->n) { delete c; return BV_ERR_ARG; }
    if (p->nbr && c->cs > 0) {
        for (int64_t i = (int64_t)c->cs; i < (int64_t)c->B * c->cs; ++i)
            if (p->nbr[i] < 0 || p->nbr[i] >= c->n) { delete c; return BV_ERR_ARG; }
    }

    /* launch descriptors: block sizes, padding, workspace offsets */
    std::vector<bv_block_desc> desc((size_t)c->B);
    int64_t off = 0;
    for (int i = 0; i < c->B; ++i) {
        bv_block_desc &d = desc[i];
        memset(&d, 0, sizeof(d));
        d.block = i;
        d.cs0 = (i == 0 || c->cs == 0) ? 0 : c->cs;
        d.m = d.cs0 + p->batch_num[i];
        d.mp = ((d.m + 1 + BV_NB - 1) / BV_NB) * BV_NB;
        d.obs_off = c->h_accum[i];
        const int64_t T = d.mp / BV_TILE;
        d.ws_off = off;
        off += T * (T + 1) / 2 * 256;
        c->max_mp = std::max(c->max_mp, d.mp);
    }
    /* the fused kernels stage a block's points in LDS (4 Mp doubles): they
       take padded sizes up to BV_MAX_MP; the stepped forms have no limit */
    c->fused_ok = c->max_mp <= BV_MAX_MP;
    /* the register-resident kernel: padded sizes up to 16 * 24, Euclidean */
    c->res_ok = c->max_mp <= BV_RES_MAX_MP && c->metric == 0;
    if (c->variant == 3 && !c->res_ok) {
        fprintf(stderr,
                "libbv: the register-resident factorization takes padded sizes up to %d and "
                "Euclidean distances (largest padded size %d, metric %d)\n",
                BV_RES_MAX_MP, c->max_mp, c->metric);
        delete c;
        return BV_ERR_UNSUPPORTED;
    }
    if (!c->fused_ok && c->stepped == 0) {
        if (getenv("BV_FACTOR_KERNEL") || getenv("BV_PATH")) {
            fprintf(stderr,
                    "libbv: a block's padded augmented size %d exceeds %d, the limit of the "
                    "fused factorization kernels; use a stepped form (BV_PATH=stepped2)\n",
                    c->max_mp, BV_MAX_MP);
            delete c;
            return BV_ERR_UNSUPPORTED;
        }
        c->stepped = 2; /* default without autotuning */
    }
    c->ws_per_eval = off;
    std::stable_sort(desc.begin(), desc.end(),
                     [](const bv_block_desc &a, const bv_block_desc &b) { return a.mp > b.mp; });
    c->base_desc = desc;

    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        bv_destroy(c);
        return BV_ERR_HIP;
    }
    if (c->stepped && c->nstreams > 1 && bv_side_streams(c) != BV_OK) {
        bv_destroy(c);
        return BV_ERR_HIP;
    }
    if (!c->stepped && c->variant == 3 && bv_res_streams(c) != BV_OK) {
        bv_destroy(c);
        return BV_ERR_HIP;
    }
    int rc = BV_OK;
    const size_t n = (size_t)c->n;
    if ((rc = bv_upload(&c->d_x, p->x, n, c->stream)) ||
        (rc = bv_upload(&c->d_y, p->y, n, c->stream)) ||
        (rc = bv_upload(&c->d_obs, p->obs, n, c->stream)) ||
        (rc = bv_upload(&c->d_batch_num, p->batch_num, (size_t)c->B, c->stream))) {
        bv_destroy(c);
        return rc;
    }
    if (c->time_flag && (rc = bv_upload(&c->d_z, p->z, n, c->stream))) { bv_destroy(c); return rc; }
    const size_t ncon = (size_t)c->B * c->cs;
    if (c->cs > 0) {
        if (p->nbr) {
            if ((rc = bv_upload(&c->d_nbr, p->nbr, ncon, c->stream))) { bv_destroy(c); return rc; }
        } else {
            if ((rc = bv_upload(&c->d_cx, p->con_x, ncon, c->stream)) ||
                (rc = bv_upload(&c->d_cy, p->con_y, ncon, c->stream)) ||
                (rc = bv_upload(&c->d_cobs, p->con_obs, ncon, c->stream))) {
                bv_destroy(c);
                return rc;
            }
            if (c->time_flag && (rc = bv_upload(&c->d_cz, p->con_z, ncon, c->stream))) {
                bv_destroy(c);
                return rc;
            }
        }
    }
    if ((rc = bv_reserve(c, 1)) || (rc = bv_layout(c, 1))) { bv_destroy(c); return rc; }
    if (hipStreamSynchronize(c->stream) != hipSuccess) { bv_destroy(c); return BV_ERR_HIP; }
    *out = c;
    return BV_OK;
}

/* The register-resident builds of nblk descriptors sorted by padded size,
   largest first: runs of equal bv_res_nr (at most BV_RES_NR_MAX - 1), build
   nr[i] for cnt[i] descriptors; returns the number of runs. */
static int bv_res_runs(const bv_block_desc *h_desc, unsigned nblk, int *nr, unsigned *cnt) {
    int n = 0;
    for (unsigned i = 0; i < nblk; ++i) {
        const int r = bv_res_nr(h_desc[i].mp);
        if (n == 0 || nr[n - 1] != r) {
            nr[n] = r;
            cnt[n] = 0;
            ++n;
        }
        ++cnt[n - 1];
    }
    return n;
}

/* The register-resident kernel over nblk descriptors sorted by padded size,
   largest first: each run with its build (the fewest registers that hold its
   blocks), on one stream. */
static hipError_t bv_launch_res_sorted(int kind, const bv_block_desc *h_desc,
                                       const bv_block_desc *d_desc, unsigned nblk,
                                       hipStream_t st, const bv_eval_slot *slots, double *ws) {
    int nr[BV_RES_NR_MAX];
    unsigned cnt[BV_RES_NR_MAX], o = 0;
    const int n = bv_res_runs(h_desc, nblk, nr, cnt);
    for (int i = 0; i < n; ++i) {
        const hipError_t e = bv_launch_res(kind, nr[i], cnt[i], st, d_desc + o, slots, ws);
        if (e != hipSuccess) return e;
        o += cnt[i];
    }
    return hipSuccess;
}

/* The kernels of one call after its slots are on the device: covariance
   generation where it is tabulated beforehand, the factorization in the
   configuration chosen, and the stepped form's partial sums.  Everything is
   ordered on c->stream (side streams fork from it and join back). */
static int bv_enqueue(bv_ctx *c, unsigned nblk, int kinds_used) {
    const size_t lds = sizeof(double) * 4 * (size_t)c->max_mp;
    for (int kind = 0; kind < 8; ++kind) {
        if (!(kinds_used & (1 << kind))) continue;
        if (kind == BV_KIND_GENERAL)
            BV_HIP(bv_launch_cheb_build((unsigned)c->desc_evals, c->stream, c->d_slots));
        if (kind == BV_KIND_SPACETIME) {
            const unsigned T = (unsigned)(c->max_mp / BV_TILE);
            BV_HIP(bv_launch_gen(kind, T * (T + 1) / 2, nblk, c->stream, c->d_desc, c->d_slots,
                                 c->d_ws));
        }
        if (c->stepped) {
            /* chunks of blocks cycle over the side streams: a few chunks are
               active at a time, so their workspace stays in the last-level
               cache while their steps interleave on the device */
            const bool side = c->nstreams > 1;
            if (side) {
                BV_HIP(hipEventRecord(c->cev[c->nstreams], c->stream));
                for (int i = 0; i < c->nstreams; ++i)
                    BV_HIP(hipStreamWaitEvent(c->cstream[i], c->cev[c->nstreams], 0));
            }
            /* step j of every chunk before step j + 1 of any: each stream
               gets its first launch at once, so the chunks' chains overlap
               from the start (on one stream, chunk by chunk was slower: the
               launches lose parallelism and gain no cache reuse) */
            const bool skew = c->nstreams == 0 && c->nchunks > 1;
            const int nl = skew ? c->pmax + c->nchunks - 1 : c->pmax;
            const int nch = skew ? 1 : c->nchunks;
            for (int j = 0; j < nl; ++j) {
                for (int ch = 0; ch < nch; ++ch) {
                    hipStream_t st = side ? c->cstream[ch % c->nstreams] : c->stream;
                    const int64_t *off = c->step_off.data() + (size_t)ch * (c->pmax + 1);
                    const int64_t o = off[j], cnt = off[j + 1] - o;
                    if (cnt <= 0) continue;
                    /* every item carries its own step */
                    if (c->stepped == 3)
                        BV_HIP(bv_launch_pair(kind, (unsigned)cnt, st, c->d_items + o,
                                              c->d_desc, c->d_slots, c->d_ws, c->d_part_logdet,
                                              c->d_part_dot, c->d_part_info, c->pmax));
                    else if (c->stepped == 2)
                        BV_HIP(bv_launch_step2(kind, (unsigned)cnt, st, c->d_items + o,
                                               c->d_desc, c->d_slots, c->d_ws, c->d_part_logdet,
                                               c->d_part_dot, c->d_part_info, c->pmax));
                    else
                        BV_HIP(bv_launch_step(kind, (unsigned)cnt, st, c->d_items + o,
                                              c->d_desc, c->d_slots, c->d_ws, c->d_part_logdet,
                                              c->d_part_dot, c->d_part_info, c->pmax));
                }
            }
            if (side) {
                for (int i = 0; i < c->nstreams; ++i) {
                    BV_HIP(hipEventRecord(c->cev[i], c->cstream[i]));
                    BV_HIP(hipStreamWaitEvent(c->stream, c->cev[i], 0));
                }
            }
            BV_HIP(bv_launch_step_parts((int)nblk, c->stream, c->d_desc, c->d_slots,
                                        c->d_part_logdet, c->d_part_dot, c->d_part_info,
                                        c->pmax));
        } else if (c->variant == 3) {
            /* descriptors of k evaluations: each block's k copies in a row,
               blocks sorted by padded size */
            const unsigned k = (unsigned)c->desc_evals;
            int nr[BV_RES_NR_MAX];
            unsigned cnt[BV_RES_NR_MAX], o = 0;
            const int n = bv_res_runs(c->base_desc.data(), (unsigned)c->B, nr, cnt);
            const bool side = n > 1 && c->cstream[0];
            if (side) {
                /* the run of the largest blocks on a side stream beside the others */
                BV_HIP(hipEventRecord(c->cev[1], c->stream));
                BV_HIP(hipStreamWaitEvent(c->cstream[0], c->cev[1], 0));
            }
            for (int i = 0; i < n; ++i) {
                const unsigned m = cnt[i] * k;
                BV_HIP(bv_launch_res(kind, nr[i], m, (side && i == 0) ? c->cstream[0] : c->stream,
                                     c->d_desc + o, c->d_slots, c->d_ws));
                o += m;
            }
            if (side) {
                BV_HIP(hipEventRecord(c->cev[0], c->cstream[0]));
                BV_HIP(hipStreamWaitEvent(c->stream, c->cev[0], 0));
            }
        } else if (c->variant == 1)
            BV_HIP(bv_launch_block_1w(kind, nblk, lds, c->stream, c->d_desc, c->d_slots, c->d_ws));
        else if (c->variant == 2)
            BV_HIP(bv_launch_block_2w(kind, nblk, lds, c->stream, c->d_desc, c->d_slots, c->d_ws));
        else
            BV_HIP(bv_launch_block(kind, nblk, lds, c->stream, c->d_desc, c->d_slots, c->d_ws));
    }
    return BV_OK;
}

/* Identity of the launches bv_enqueue makes: the configuration, the kinds and
   every device pointer and count baked into them. */
static void bv_graph_key(const bv_ctx *c, int kinds_used, int64_t *key) {
    key[0] = kinds_used;
    key[1] = c->variant | (c->stepped << 4) | (c->nchunks << 8) | ((int64_t)c->nstreams << 16);
    key[2] = (int64_t)(intptr_t)c->d_ws;
    key[3] = (int64_t)(intptr_t)c->d_desc;
    key[4] = (int64_t)(intptr_t)c->d_items;
    key[5] = (int64_t)(intptr_t)c->d_slots;
    key[6] = (int64_t)(intptr_t)c->d_part_logdet;
    key[7] = c->desc_evals;
    key[8] = c->pmax;
    key[9] = c->max_mp;
    key[10] = (int64_t)(intptr_t)c->d_part_dot;
    key[11] = (int64_t)(intptr_t)c->d_part_info;
}

/* bv_enqueue through a graph captured on first use and replayed while its
   key holds.  The graph holds kernel launches only (no memory set or copy
   nodes); the per-call parameters reach it through the slots, copied
   before the launch. */
static int bv_enqueue_graph(bv_ctx *c, unsigned nblk, int kinds_used) {
    int64_t key[12];
    bv_graph_key(c, kinds_used, key);
    if (c->graph && memcmp(key, c->graph_key, sizeof(key)) != 0) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    if (!c->graph) {
        BV_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = bv_enqueue(c, nblk, kinds_used);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        BV_HIP(ec);
        const hipError_t ei = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        BV_HIP(ei);
        memcpy(c->graph_key, key, sizeof(key));
    }
    BV_HIP(hipGraphLaunch(c->graph, c->stream));
    return BV_OK;
}

/* Launch k evaluations (thetas row-major, nparam each) with the current
   factorization configuration and wait for them. */
static int bv_run_once(bv_ctx *c, int k, int nparam, const double *thetas, double *out,
                       int *status) {
    if (k <= 0 || !thetas || !out) return BV_ERR_ARG;
    if (nparam != bv_nparams(c->kernel)) return BV_ERR_ARG;
    BV_HIP(hipSetDevice(c->device));
    int rc;
    if ((rc = bv_reserve(c, k)) || (rc = bv_layout(c, k))) return rc;
    const size_t B = (size_t)c->B;
    for (int e = 0; e < k; ++e) {
        bv_eval_slot &s = c->h_slots[e];
        memset(&s, 0, sizeof(s));
        if ((rc = bv_make_kparams(c, thetas + (size_t)e * nparam, &s.kp, &s.kind))) return rc;
        s.x = c->d_x; s.y = c->d_y; s.z = c->d_z; s.obs = c->d_obs;
        s.nbr = c->d_nbr; s.cx = c->d_cx; s.cy = c->d_cy; s.cz = c->d_cz; s.cobs = c->d_cobs;
        s.cs = c->cs;
        s.nblocks = c->B;
        s.batch_num = c->d_batch_num;
        s.logdet = c->d_logdet + e * B;
        s.dot = c->d_dot + e * B;
        s.info = c->d_info + e * B;
        s.result = c->d_result + 4 * e;
        s.log2pi = log(2 * 3.141592653589793);
        s.kp.cheb = c->d_cheb + (size_t)e * BV_CHEB_SIZE;
    }
    BV_HIP(hipEventRecord(c->ev0, c->stream));
    /* one kind per launch: evaluations of a call share the kernel when their
       smoothness matches; otherwise one launch per kind over its own slots */
    int kinds_used = 0;
    for (int e = 0; e < k; ++e) kinds_used |= 1 << c->h_slots[e].kind;
    if (kinds_used & (kinds_used - 1)) {
        /* smoothness differs across the call: all go through the general path */
        for (int e = 0; e < k; ++e) c->h_slots[e].kind = BV_KIND_GENERAL;
        kinds_used = 1 << BV_KIND_GENERAL;
    }
    BV_HIP(hipMemcpyAsync(c->d_slots, c->h_slots, sizeof(bv_eval_slot) * k,
                          hipMemcpyHostToDevice, c->stream));
    const unsigned nblk = (unsigned)(B * k);
    if (c->use_graph && c->graph_wins && k == 1 && (c->tuned || !c->autotune))
        rc = bv_enqueue_graph(c, nblk, kinds_used);
    else
        rc = bv_enqueue(c, nblk, kinds_used);
    if (rc) return rc;
    BV_HIP(hipEventRecord(c->ev1, c->stream));
    BV_HIP(bv_launch_reduce((unsigned)k, c->stream, c->d_slots));
    BV_HIP(hipMemcpyAsync(c->h_result, c->d_result, sizeof(double) * 4 * k,
                          hipMemcpyDeviceToHost, c->stream));
    BV_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->last_ms = ms;
    c->last_k = k;
    int worst = BV_OK;
    c->last_bad = -1;
    for (int e = 0; e < k; ++e) {
        const double llh = c->h_result[4 * e];
        const int bad = (int)c->h_result[4 * e + 1];
        int st = BV_OK;
        if (bad >= 0) {
            st = BV_ERR_NOT_PD;
            if (c->last_bad < 0) c->last_bad = bad;
            out[e] = -INFINITY;
        } else {
            out[e] = llh;
        }
        if (status) status[e] = st;
        if (st != BV_OK) worst = st;
    }
    c->last_status = worst;
    return BV_OK;
}

/* Factorization configurations the autotuner chooses from: the fused kernel
   with four, one or two wavefronts per block, and the stepped
   factorization, whole or in two or three chunks over as many streams, with
   lookahead in two or three chunks, and by pairs of panels whole or in two
   or three chunks.  Which is fastest depends on the block sizes and the
   covariance kind. */
struct bv_cfg {
    int variant, stepped, nchunks, nstreams;
};
static const bv_cfg bv_cfgs[] = {{0, 0, 1, 1}, {1, 0, 1, 1}, {2, 0, 1, 1}, {0, 1, 1, 1},
                                 {0, 1, 2, 2}, {0, 1, 3, 3}, {0, 2, 2, 2}, {0, 2, 3, 3},
                                 {0, 3, 1, 1}, {0, 3, 2, 2}, {0, 3, 3, 3}, {0, 2, 2, 0},
                                 {0, 3, 2, 0}, {3, 0, 1, 1}};

/* Whether the autotuner may keep configuration g (a context's blocks fit its
   kernels).  The register-resident form is offered where it holds every
   block and some block is above BV_RES_TUNE_MIN_MP: on datasets of only tiny
   blocks (padded size <= 128) a four-wavefront workgroup per block loses in
   the request server's merged launches to the one-wavefront fused kernel
   (small-dataset serving at 0.81x the throughput), although one evaluation
   alone is faster; it can still be forced (BV_FACTOR_KERNEL=res). */
static bool bv_cfg_ok(const bv_ctx *c, const bv_cfg &g) {
    if (g.stepped) return true;
    if (g.variant == 3) return c->res_ok != 0 && c->max_mp >= BV_RES_TUNE_MIN_MP;
    return c->fused_ok != 0;
}

static int bv_apply_cfg(bv_ctx *c, const bv_cfg &g) {
    /* the step items depend on the chunks and on whether they are skewed */
    const bool relayout = g.nchunks != c->nchunks || g.stepped != c->stepped ||
                          (g.nstreams == 0) != (c->nstreams == 0);
    c->variant = g.variant;
    c->stepped = g.stepped;
    c->nchunks = g.nchunks;
    c->nstreams = g.nstreams;
    if (relayout) c->desc_evals = 0; /* the step items depend on the chunks */
    if (c->stepped && c->nstreams > 1) return bv_side_streams(c);
    if (!c->stepped && c->variant == 3) return bv_res_streams(c);
    return BV_OK;
}

/* ---- autotune cache ----------------------------------------------------
   The choice of the autotuner is kept in a text file, one line per workload:
   "<key> <candidate index> <graph>", the last line of a key winning.  The key
   names the device (architecture, compute units), the dataset's shape (block
   count, conditioning size, kernel, metric, 3-D flag, a hash of every
   block's padded size), the evaluations per call, the covariance kind and
   the candidate list.  Path: BV_TUNE_CACHE, else
   $HOME/.cache/blockvecchia_amd/autotune.txt; BV_TUNE_CACHE=0 or empty
   disables it.  A context that finds its key skips the timing runs, so the
   configuration (and with it the last bits of the results) is the same in
   every process on the machine. */
static std::string bv_tune_cache_path() {
    const char *v = getenv("BV_TUNE_CACHE");
    if (v) return (v[0] == 0 || strcmp(v, "0") == 0) ? std::string() : std::string(v);
    const char *h = getenv("HOME");
    if (!h || !h[0]) return std::string();
    return std::string(h) + "/.cache/blockvecchia_amd/autotune.txt";
}

static std::string bv_tune_key(const bv_ctx *c, int k, int kinds_used) {
    hipDeviceProp_t prop;
    char arch[64] = "unknown";
    int cus = 0;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) {
        snprintf(arch, sizeof(arch), "%s", prop.gcnArchName);
        for (char *q = arch; *q; ++q)
            if (*q == ' ' || *q == ':') *q = '_';
        cus = prop.multiProcessorCount;
    }
    uint64_t h = 1469598103934665603ULL; /* FNV-1a over the padded sizes */
    for (const bv_block_desc &d : c->base_desc) {
        h ^= (uint64_t)(uint32_t)d.mp;
        h *= 1099511628211ULL;
    }
    for (const bv_cfg &g : bv_cfgs) {
        h ^= (uint64_t)(g.variant | g.stepped << 4 | g.nchunks << 8 | g.nstreams << 16);
        h *= 1099511628211ULL;
    }
    char buf[256];
    snprintf(buf, sizeof(buf), "%s/%d/B%d/cs%d/k%d/m%d/t%d/e%d/kind%d/%016llx", arch, cus, c->B,
             c->cs, c->kernel, c->metric, c->time_flag, k, kinds_used, (unsigned long long)h);
    return std::string(buf);
}

static bool bv_tune_cache_get(const std::string &path, const std::string &key, int *idx,
                              int *graph) {
    if (path.empty()) return false;
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return false;
    char line[512], kbuf[300];
    bool found = false;
    while (fgets(line, sizeof(line), f)) {
        int i = -1, g = 0;
        if (sscanf(line, "%299s %d %d", kbuf, &i, &g) == 3 && key == kbuf) {
            *idx = i;
            *graph = g;
            found = true;
        }
    }
    fclose(f);
    return found;
}

static void bv_tune_cache_put(const std::string &path, const std::string &key, int idx,
                              int graph) {
    if (path.empty()) return;
    const size_t sl = path.find_last_of('/');
    if (sl != std::string::npos && sl > 0) {
        const std::string dir = path.substr(0, sl);
        const size_t sl2 = dir.find_last_of('/');
        if (sl2 != std::string::npos && sl2 > 0) (void)mkdir(dir.substr(0, sl2).c_str(), 0755);
        (void)mkdir(dir.c_str(), 0755);
    }
    char line[400];
    const int len = snprintf(line, sizeof(line), "%s %d %d\n", key.c_str(), idx, graph);
    const int fd = open(path.c_str(), O_WRONLY | O_APPEND | O_CREAT, 0644);
    if (fd < 0) return;
    /* one write of a short line to an O_APPEND file: lines of concurrent
       writers do not interleave */
    if (write(fd, line, (size_t)len) != len) { /* a lost cache line only costs a retune */ }
    close(fd);
}

/* bv_run_once, choosing the configuration on the first call: the cached
   choice for this workload if there is one, else each candidate runs once to
   warm up and twice timed (device time of the factorization), the fastest
   is kept for the life of the context (and cached), and the call's results
   come from a final run with it.  Blocks above BV_MAX_MP exclude the fused
   candidates. */
static int bv_run(bv_ctx *c, int k, int nparam, const double *thetas, double *out, int *status) {
    if (c->autotune && !c->tuned) {
        const int ncfg = (int)(sizeof(bv_cfgs) / sizeof(bv_cfgs[0]));
        int kinds_used = 0;
        for (int e = 0; e < k; ++e) {
            bv_kparams kp;
            int kind = 0;
            if (bv_make_kparams(c, thetas + (size_t)e * nparam, &kp, &kind) == BV_OK)
                kinds_used |= 1 << kind;
        }
        const std::string path = bv_tune_cache_path();
        const std::string key = bv_tune_key(c, k, kinds_used);
        int idx = -1, graph = 1;
        if (bv_tune_cache_get(path, key, &idx, &graph) && idx >= 0 && idx < ncfg &&
            bv_cfg_ok(c, bv_cfgs[idx])) {
            int rc = bv_apply_cfg(c, bv_cfgs[idx]);
            if (rc) return rc;
            c->tuned = 1;
            c->graph_wins = graph != 0;
            return bv_run_once(c, k, nparam, thetas, out, status);
        }
        int best = -1, best_fused = -1;
        double best_ms = INFINITY, fused_ms = INFINITY;
        for (int i = 0; i < ncfg; ++i) {
            if (!bv_cfg_ok(c, bv_cfgs[i])) continue;
            int rc = bv_apply_cfg(c, bv_cfgs[i]);
            if (rc) return rc;
            double ms = INFINITY;
            for (int r = 0; r < 3; ++r) {
                if ((rc = bv_run_once(c, k, nparam, thetas, out, status))) return rc;
                if (r > 0) ms = std::min(ms, c->last_ms);
            }
            if (best < 0 || ms < best_ms) { best_ms = ms; best = i; }
            if (!bv_cfgs[i].stepped && (best_fused < 0 || ms < fused_ms)) {
                fused_ms = ms;
                best_fused = i;
            }
        }
        /* a fused form within 10% of the fastest is kept instead: its blocks
           merge with other datasets' fused-form blocks into one launch in the
           request server (a stepped form needs its own step launches) */
        if (best_fused >= 0 && fused_ms <= 1.1 * best_ms) best = best_fused;
        int rc = bv_apply_cfg(c, bv_cfgs[best]);
        if (rc) return rc;
        c->tuned = 1;
        if (k == 1 && c->use_graph) {
            /* a graph launch costs more than one kernel launch and less than
               many: replay only if it times faster than launching directly */
            double t[2] = {INFINITY, INFINITY};
            for (int g = 0; g < 2; ++g) {
                c->graph_wins = g;
                for (int r = 0; r < 3; ++r) {
                    if ((rc = bv_run_once(c, k, nparam, thetas, out, status))) return rc;
                    if (r > 0) t[g] = std::min(t[g], c->last_ms);
                }
            }
            c->graph_wins = t[1] < t[0];
        }
        bv_tune_cache_put(path, key, best, c->graph_wins);
    }
    return bv_run_once(c, k, nparam, thetas, out, status);
}

extern "C" int bv_factor_candidates(int32_t *cfgs, int max_count) {
    const int ncfg = (int)(sizeof(bv_cfgs) / sizeof(bv_cfgs[0]));
    if (cfgs) {
        for (int i = 0; i < ncfg && i < max_count; ++i) {
            cfgs[4 * i + 0] = bv_cfgs[i].variant;
            cfgs[4 * i + 1] = bv_cfgs[i].stepped;
            cfgs[4 * i + 2] = bv_cfgs[i].nchunks;
            cfgs[4 * i + 3] = bv_cfgs[i].nstreams;
        }
    }
    return ncfg;
}

extern "C" int bv_factor_config(bv_ctx *c, int32_t *cfg) {
    if (!c || !cfg) return BV_ERR_ARG;
    cfg[0] = c->variant;
    cfg[1] = c->stepped;
    cfg[2] = c->nchunks;
    cfg[3] = c->nstreams;
    cfg[4] = c->use_graph && c->graph_wins;
    return c->tuned || !c->autotune ? BV_OK : BV_ERR_BUSY;
}

extern "C" int bv_set_graph(bv_ctx *c, int enable) {
    if (!c) return BV_ERR_ARG;
    c->use_graph = enable != 0;
    return BV_OK;
}

extern "C" double bv_llh(unsigned n, const double *theta, double *grad, void *f_data) {
    (void)grad;
    bv_ctx *c = (bv_ctx *)f_data;
    if (!c || !theta || (int)n != bv_nparams(c->kernel)) return NAN;
    double llh = NAN;
    int st = 0;
    const int rc = bv_run(c, 1, (int)n, theta, &llh, &st);
    if (rc != BV_OK) { c->last_status = rc; return NAN; }
    if (c->verbose) {
        if (c->kernel == 1 || c->kernel == 2)
            printf("%dth Model Parameters (Variance, range, smoothness): (%1.8lf, %1.8lf, %1.8lf) -> Loglik: %.18lf \n",
                   c->iterations, theta[0], theta[1], theta[2], llh);
        else if (c->kernel == 3)
            printf("%dth Model Parameters (Variance, range, smoothness, nugget): (%1.8lf, %1.8lf, %1.8lf, %1.8lf) -> Loglik: %.18lf \n",
                   c->iterations, theta[0], theta[1], theta[2], theta[3], llh);
        else
            printf("%dth Model Parameters (Variance, beta, nu, beta_time, nu_time, sep): (%lf, %lf, %lf, %lf, %lf, %lf) -> Loglik: %lf \n",
                   c->iterations, theta[0], theta[1], theta[2], theta[3], theta[4], theta[5], llh);
    }
    c->iterations += 1;
    return llh;
}

extern "C" double llh_Xvecchia_batch(unsigned n, const double *theta, double *grad, void *f_data) {
    return bv_llh(n, theta, grad, f_data);
}

extern "C" int bv_llh_many(bv_ctx *c, int k, int nparam, const double *thetas, double *out,
                           int *status) {
    if (!c) return BV_ERR_ARG;
    const int rc = bv_run(c, k, nparam, thetas, out, status);
    if (rc == BV_OK) c->iterations += k;
    else c->last_status = rc;
    return rc;
}

extern "C" int bv_block_terms(bv_ctx *c, double *logdet, double *dot, int32_t *info) {
    if (!c || c->last_k <= 0) return BV_ERR_ARG;
    BV_HIP(hipSetDevice(c->device));
    const size_t B = (size_t)c->B;
    if (logdet) BV_HIP(hipMemcpy(logdet, c->d_logdet, sizeof(double) * B, hipMemcpyDeviceToHost));
    if (dot) BV_HIP(hipMemcpy(dot, c->d_dot, sizeof(double) * B, hipMemcpyDeviceToHost));
    if (info) BV_HIP(hipMemcpy(info, c->d_info, sizeof(int32_t) * B, hipMemcpyDeviceToHost));
    return BV_OK;
}

extern "C" int bv_last_status(bv_ctx *c, int *first_bad_block) {
    if (!c) return BV_ERR_ARG;
    if (first_bad_block) *first_bad_block = c->last_bad;
    return c->last_status;
}

extern "C" double bv_last_kernel_ms(bv_ctx *c) { return c ? c->last_ms : -1.0; }

extern "C" int bv_iterations(bv_ctx *c) { return c ? c->iterations : -1; }

extern "C" void bv_destroy(bv_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    for (int i = 0; i < 4; ++i)
        if (c->cstream[i]) { (void)hipStreamSynchronize(c->cstream[i]); (void)hipStreamDestroy(c->cstream[i]); }
    for (int i = 0; i < 5; ++i)
        if (c->cev[i]) (void)hipEventDestroy(c->cev[i]);
    bv_free_eval_buffers(c);
    double *bufs[] = {c->d_x, c->d_y, c->d_z, c->d_obs, c->d_cx, c->d_cy, c->d_cz, c->d_cobs};
    for (double *b : bufs)
        if (b) (void)hipFree(b);
    if (c->d_nbr) (void)hipFree(c->d_nbr);
    if (c->d_batch_num) (void)hipFree(c->d_batch_num);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int bv_matern(int m, int nn, const double *x1, const double *y1, const double *z1,
                         const double *x2, const double *y2, const double *z2,
                         const double *theta, int kernel, int distance_metric, double dist_scale,
                         double *A, int device) {
    if (m <= 0 || nn <= 0 || !x1 || !y1 || !x2 || !y2 || !theta || !A) return BV_ERR_ARG;
    if (kernel < 1 || kernel > 4) return BV_ERR_ARG;
    BV_HIP(hipSetDevice(device));
    bv_ctx tmp;
    tmp.kernel = kernel;
    tmp.metric = distance_metric;
    tmp.dist_scale = dist_scale > 0 ? dist_scale : 1.0;
    const bool use_z = (z1 && z2);
    tmp.time_flag = use_z ? 1 : 0;
    tmp.d_z = use_z ? (double *)1 : nullptr; /* only tested for non-null */
    bv_kparams kp;
    int kind = 0;
    int rc = bv_make_kparams(&tmp, theta, &kp, &kind);
    if (rc) return rc;
    double *d[7] = {nullptr};
    const size_t sm = sizeof(double) * m, sn = sizeof(double) * nn;
    BV_HIP(hipMalloc(&d[0], sm)); BV_HIP(hipMalloc(&d[1], sm));
    BV_HIP(hipMalloc(&d[3], sn)); BV_HIP(hipMalloc(&d[4], sn));
    BV_HIP(hipMalloc(&d[6], sm * nn));
    BV_HIP(hipMemcpy(d[0], x1, sm, hipMemcpyHostToDevice));
    BV_HIP(hipMemcpy(d[1], y1, sm, hipMemcpyHostToDevice));
    BV_HIP(hipMemcpy(d[3], x2, sn, hipMemcpyHostToDevice));
    BV_HIP(hipMemcpy(d[4], y2, sn, hipMemcpyHostToDevice));
    if (use_z) {
        BV_HIP(hipMalloc(&d[2], sm)); BV_HIP(hipMalloc(&d[5], sn));
        BV_HIP(hipMemcpy(d[2], z1, sm, hipMemcpyHostToDevice));
        BV_HIP(hipMemcpy(d[5], z2, sn, hipMemcpyHostToDevice));
    }
    double *d_cheb = nullptr;
    bv_eval_slot *d_sl = nullptr;
    if (kind == BV_KIND_GENERAL) {
        /* the evaluation's x^nu K_nu table, as the likelihood builds it */
        BV_HIP(hipMalloc(&d_cheb, sizeof(double) * BV_CHEB_SIZE));
        BV_HIP(hipMalloc(&d_sl, sizeof(bv_eval_slot)));
        kp.cheb = d_cheb;
        bv_eval_slot sl;
        memset(&sl, 0, sizeof(sl));
        sl.kp = kp;
        sl.kind = kind;
        BV_HIP(hipMemcpy(d_sl, &sl, sizeof(sl), hipMemcpyHostToDevice));
        BV_HIP(bv_launch_cheb_build(1u, 0, d_sl));
    }
    BV_HIP(bv_launch_matern(m, nn, d[0], d[1], d[2], d[3], d[4], d[5], kp, kind, d[6]));
    BV_HIP(hipMemcpy(A, d[6], sm * nn, hipMemcpyDeviceToHost));
    for (double *p : d)
        if (p) (void)hipFree(p);
    if (d_cheb) (void)hipFree(d_cheb);
    if (d_sl) (void)hipFree(d_sl);
    return BV_OK;
}

extern "C" int bv_device_selftest(int device) {
    BV_HIP(hipSetDevice(device));
    double ha[64], hb[64], hd[256];
    for (int l = 0; l < 64; ++l) {
        ha[l] = 1.0 + l;            /* A[l%16][l/16] */
        hb[l] = 0.5 * (l % 7) - 1;  /* B[l/16][l%16] */
    }
    double *d = nullptr;
    BV_HIP(hipMalloc(&d, sizeof(double) * 384));
    BV_HIP(hipMemcpy(d, ha, sizeof(ha), hipMemcpyHostToDevice));
    BV_HIP(hipMemcpy(d + 64, hb, sizeof(hb), hipMemcpyHostToDevice));
    BV_HIP(bv_launch_selftest(d, d + 64, d + 128));
    BV_HIP(hipMemcpy(hd, d + 128, sizeof(hd), hipMemcpyDeviceToHost));
    (void)hipFree(d);
    /* expected D[i][j] = sum_k A[i][k] B[k][j], lane l holds D[4g + l/16][l%16] */
    for (int l = 0; l < 64; ++l)
        for (int g = 0; g < 4; ++g) {
            const int i = 4 * g + l / 16, j = l % 16;
            double ref = 0.0;
            for (int k = 0; k < 4; ++k) ref += ha[i + 16 * k] * hb[j + 16 * k];
            if (fabs(ref - hd[4 * l + g]) > 1e-12 * (1.0 + fabs(ref))) return 1;
        }
    return BV_OK;
}

extern "C" int bv_fp64_pipe_probe(int device, double *cycles) {
    if (!cycles) return BV_ERR_ARG;
    BV_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    BV_HIP(hipGetDeviceProperties(&prop, device));
    const int ncu = prop.multiProcessorCount, iters = 125; /* 2000 steps */
    double *out = nullptr;
    unsigned long long *cyc = nullptr;
    BV_HIP(hipMalloc(&out, sizeof(double) * 512 * (size_t)ncu));
    if (hipMalloc(&cyc, sizeof(unsigned long long) * 8 * (size_t)ncu) != hipSuccess) {
        (void)hipFree(out);
        return BV_ERR_HIP;
    }
    std::vector<unsigned long long> h(8 * (size_t)ncu);
    int rc = BV_OK;
    for (int mode = 0; mode < 5 && rc == BV_OK; ++mode) {
        /* one warm-up launch, then the timed one */
        if (bv_launch_pipe_probe(mode, ncu, 10, out, cyc) != hipSuccess ||
            bv_launch_pipe_probe(mode, ncu, iters, out, cyc) != hipSuccess ||
            hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(h.data(), cyc, sizeof(unsigned long long) * h.size(),
                      hipMemcpyDeviceToHost) != hipSuccess) {
            rc = BV_ERR_HIP;
            break;
        }
        /* median over the compute units of wavefront 0 (and, split, of
           wavefront 1: the multiply-add one) */
        for (int w = 0; w < (mode == 4 ? 2 : 1); ++w) {
            std::vector<unsigned long long> v;
            for (int c = 0; c < ncu; ++c) v.push_back(h[8 * (size_t)c + w]);
            std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
            cycles[mode + w] = (double)v[v.size() / 2] / (16.0 * iters);
        }
    }
    (void)hipFree(out);
    (void)hipFree(cyc);
    return rc;
}

/* ======================================================================
 * Request server: many datasets, many concurrent evaluations.
 *
 * Each dataset is a bv_ctx (resident in HBM).  Requests (dataset, theta, tag)
 * queue in arrival order.  A lane is one stream with its own workspace,
 * descriptor, item, partial-sum, slot and result buffers; when a lane is
 * idle, bv_server_poll admits queued requests into it while their estimated
 * workspace bytes fit the lane's budget (at least one request per batch),
 * merges the blocks of all admitted evaluations -- across datasets -- into
 * shared launches and starts them without waiting.  Several lanes run at
 * once, so requests that arrive while a batch is on the GPU are admitted as
 * soon as a lane frees, and their work overlaps on the device.
 *
 * A served result equals the same evaluation run alone by bv_llh, bit for
 * bit: every evaluation of a dataset runs the factorization form the
 * dataset's autotuner chose (tuned on its first request if it was not
 * already).  Blocks are grouped by (covariance kind, form): the fused
 * kernels (four, two and one wavefronts per block add the same terms in the
 * same order, so one wavefront is used for merged batches), and each
 * stepped form with its own work items over the group's blocks (a block's
 * arithmetic in a stepped form does not depend on the launch or stream its
 * items run in).  The per-block terms and the per-evaluation sum run in the
 * fixed orders bv_llh uses.  An evaluation that fills the GPU on its own (a
 * stepped form and at least BV_SERVER_OWN_BYTES of workspace) runs alone on
 * the dataset's own launches (streams, chunks, graph), as bv_llh does.
 * ====================================================================== */
struct bv_request {
    int dataset;
    int64_t tag;
    double theta[8];
};

struct bv_lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<bv_request> reqs;            /* the batch in flight          */
    bv_ctx *own = nullptr;                   /* set: one evaluation on its dataset's own
                                                tuned path, results in h_result */
    /* capacities (sized when datasets are attached, grown while idle) */
    int64_t cap_ws = 0, cap_blocks = 0, cap_items = 0;
    int cap_evals = 0, cap_pmax = 0;
    double *d_ws = nullptr;
    bv_block_desc *d_desc = nullptr, *h_desc = nullptr;
    bv_step_item *d_items = nullptr, *h_items = nullptr;
    double *d_part_logdet = nullptr, *d_part_dot = nullptr;
    int32_t *d_part_info = nullptr;
    bv_eval_slot *d_slots = nullptr, *h_slots = nullptr;
    double *d_logdet = nullptr, *d_dot = nullptr, *d_result = nullptr, *h_result = nullptr;
    double *d_cheb = nullptr;                /* general-nu tables, one per evaluation */
    int32_t *d_info = nullptr;
};

struct bv_server {
    int device = 0;
    double ws_budget = 0.0;                  /* bytes of workspace per lane  */
    int max_evals = 64;                      /* evaluations per batch        */
    /* fused groups up to this padded size use the one-wavefront kernel:
       merged batches hold many blocks, and one wavefront per block keeps
       more of them resident (measured faster at every size tried); the
       results are those of the four-wavefront kernel */
    int64_t one_wave_max_mp = INT32_MAX;
    /* an evaluation whose dataset's workspace is at least this many bytes,
       and whose autotuned configuration is a stepped form, fills the GPU on
       its own: it runs alone on its dataset's own launches */
    double own_min_bytes = 2e8;
    std::vector<bv_ctx *> datasets;
    std::vector<bv_request> queue;           /* FIFO (front = oldest)        */
    size_t qhead = 0;
    std::vector<bv_lane> lanes;
};

static void bv_lane_free(bv_lane &L) {
    if (L.d_ws) (void)hipFree(L.d_ws);
    if (L.d_desc) (void)hipFree(L.d_desc);
    if (L.h_desc) (void)hipHostFree(L.h_desc);
    if (L.d_items) (void)hipFree(L.d_items);
    if (L.h_items) (void)hipHostFree(L.h_items);
    if (L.d_part_logdet) (void)hipFree(L.d_part_logdet);
    if (L.d_part_dot) (void)hipFree(L.d_part_dot);
    if (L.d_part_info) (void)hipFree(L.d_part_info);
    if (L.d_slots) (void)hipFree(L.d_slots);
    if (L.h_slots) (void)hipHostFree(L.h_slots);
    if (L.d_logdet) (void)hipFree(L.d_logdet);
    if (L.d_dot) (void)hipFree(L.d_dot);
    if (L.d_info) (void)hipFree(L.d_info);
    if (L.d_result) (void)hipFree(L.d_result);
    if (L.h_result) (void)hipHostFree(L.h_result);
    if (L.d_cheb) (void)hipFree(L.d_cheb);
    L.d_cheb = nullptr;
    L.d_ws = nullptr; L.d_desc = nullptr; L.h_desc = nullptr; L.d_items = nullptr;
    L.h_items = nullptr; L.d_part_logdet = nullptr; L.d_part_dot = nullptr;
    L.d_part_info = nullptr; L.d_slots = nullptr;
    L.h_slots = nullptr; L.d_logdet = nullptr; L.d_dot = nullptr; L.d_info = nullptr;
    L.d_result = nullptr; L.h_result = nullptr;
    L.cap_ws = L.cap_blocks = L.cap_items = 0;
    L.cap_evals = L.cap_pmax = 0;
}

static int bv_lane_reserve(bv_lane &L, int64_t ws, int64_t blocks, int evals, int64_t nitems,
                           int pmax) {
    if (ws <= L.cap_ws && blocks <= L.cap_blocks && evals <= L.cap_evals &&
        nitems <= L.cap_items && pmax <= L.cap_pmax)
        return BV_OK;
    ws = std::max(ws, L.cap_ws);
    blocks = std::max(blocks, L.cap_blocks);
    evals = std::max(evals, L.cap_evals);
    nitems = std::max<int64_t>(std::max(nitems, L.cap_items), 1);
    pmax = std::max(std::max(pmax, L.cap_pmax), 1);
    bv_lane_free(L);
    BV_HIP(hipMalloc((void **)&L.d_ws, sizeof(double) * (size_t)std::max<int64_t>(ws, 1)));
    BV_HIP(hipMalloc((void **)&L.d_desc, sizeof(bv_block_desc) * (size_t)blocks));
    BV_HIP(hipHostMalloc((void **)&L.h_desc, sizeof(bv_block_desc) * (size_t)blocks,
                         hipHostMallocDefault));
    BV_HIP(hipMalloc((void **)&L.d_items, sizeof(bv_step_item) * (size_t)nitems));
    BV_HIP(hipHostMalloc((void **)&L.h_items, sizeof(bv_step_item) * (size_t)nitems,
                         hipHostMallocDefault));
    const size_t np = (size_t)blocks * (size_t)pmax;
    BV_HIP(hipMalloc((void **)&L.d_part_logdet, sizeof(double) * np));
    BV_HIP(hipMalloc((void **)&L.d_part_dot, sizeof(double) * np));
    BV_HIP(hipMalloc((void **)&L.d_part_info, sizeof(int32_t) * np));
    BV_HIP(hipMalloc((void **)&L.d_slots, sizeof(bv_eval_slot) * evals));
    BV_HIP(hipHostMalloc((void **)&L.h_slots, sizeof(bv_eval_slot) * evals, hipHostMallocDefault));
    BV_HIP(hipMalloc((void **)&L.d_logdet, sizeof(double) * (size_t)blocks));
    BV_HIP(hipMalloc((void **)&L.d_dot, sizeof(double) * (size_t)blocks));
    BV_HIP(hipMalloc((void **)&L.d_info, sizeof(int32_t) * (size_t)blocks));
    BV_HIP(hipMalloc((void **)&L.d_result, sizeof(double) * 4 * evals));
    BV_HIP(hipHostMalloc((void **)&L.h_result, sizeof(double) * 4 * evals, hipHostMallocDefault));
    BV_HIP(hipMalloc((void **)&L.d_cheb, sizeof(double) * BV_CHEB_SIZE * evals));
    L.cap_ws = ws; L.cap_blocks = blocks; L.cap_evals = evals;
    L.cap_items = nitems; L.cap_pmax = pmax;
    return BV_OK;
}

/* Upper bound of the step items one evaluation of c adds to a merged batch
   (its blocks' items; the per-launch padding is bounded separately). */
static int64_t bv_items_per_eval(const bv_ctx *c) {
    int64_t n = 0;
    for (const bv_block_desc &d : c->base_desc) {
        const int64_t T = d.mp / BV_TILE;
        n += T * T / 4 + T + 1;
    }
    return n;
}

extern "C" int bv_server_create(int device, int nlanes, double ws_bytes_per_lane,
                                int max_evals_per_batch, bv_server **out) {
    if (!out || nlanes < 1 || nlanes > 16 || !(ws_bytes_per_lane > 0) || max_evals_per_batch < 1)
        return BV_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev)
        return BV_ERR_NO_DEVICE;
    BV_HIP(hipSetDevice(device));
    bv_server *s = new bv_server();
    s->device = device;
    s->ws_budget = ws_bytes_per_lane;
    s->max_evals = max_evals_per_batch;
    if (const char *v = getenv("BV_SERVER_1W_MAXMP")) s->one_wave_max_mp = atoi(v);
    if (const char *v = getenv("BV_SERVER_OWN_BYTES")) s->own_min_bytes = atof(v);
    s->lanes.resize(nlanes);
    for (bv_lane &L : s->lanes) {
        if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess) {
            bv_server_destroy(s);
            return BV_ERR_HIP;
        }
    }
    *out = s;
    return BV_OK;
}

/* Capacities a lane needs for a full batch of the attached datasets. */
static int bv_server_size_lanes(bv_server *s) {
    int64_t max_b = 0, max_items = 0, ws = (int64_t)(s->ws_budget / 8.0);
    int pmax = 1;
    for (const bv_ctx *c : s->datasets) {
        max_b = std::max<int64_t>(max_b, c->B);
        max_items = std::max<int64_t>(max_items, bv_items_per_eval(c));
        ws = std::max<int64_t>(ws, c->ws_per_eval);
        pmax = std::max(pmax, c->max_mp / BV_NB);
    }
    /* padding as in bv_server_launch */
    const int64_t items = max_items * s->max_evals + (int64_t)(pmax + 1) * 16 * 15;
    BV_HIP(hipSetDevice(s->device));
    for (bv_lane &L : s->lanes) {
        if (L.busy) continue;
        const int rc = bv_lane_reserve(L, ws, max_b * s->max_evals, s->max_evals, items, pmax);
        if (rc != BV_OK) return rc;
    }
    return BV_OK;
}

extern "C" int bv_server_attach(bv_server *s, bv_ctx *dataset, int *dataset_id) {
    if (!s || !dataset || !dataset_id) return BV_ERR_ARG;
    if (dataset->device != s->device) return BV_ERR_ARG;
    s->datasets.push_back(dataset);
    *dataset_id = (int)s->datasets.size() - 1;
    /* size the idle lanes for a full batch now: growing a lane later frees
       device memory, which waits for every lane's work */
    return bv_server_size_lanes(s);
}

extern "C" int bv_server_submit(bv_server *s, int dataset_id, int nparam, const double *theta,
                                int64_t tag) {
    if (!s || !theta || dataset_id < 0 || dataset_id >= (int)s->datasets.size()) return BV_ERR_ARG;
    const bv_ctx *c = s->datasets[dataset_id];
    if (nparam != bv_nparams(c->kernel)) return BV_ERR_ARG;
    bv_request r;
    memset(&r, 0, sizeof(r));
    r.dataset = dataset_id;
    r.tag = tag;
    for (int i = 0; i < nparam; ++i) r.theta[i] = theta[i];
    /* reject bad parameters now, so a batch never fails for one request */
    bv_kparams kp;
    int kind;
    if (bv_make_kparams(c, r.theta, &kp, &kind) != BV_OK) return BV_ERR_ARG;
    if (s->qhead > 0 && (s->qhead == s->queue.size() || s->qhead > 4096)) {
        s->queue.erase(s->queue.begin(), s->queue.begin() + (ptrdiff_t)s->qhead);
        s->qhead = 0;
    }
    s->queue.push_back(r);
    return BV_OK;
}

static bool bv_server_own(const bv_server *s, const bv_ctx *c) {
    return c->stepped > 0 && 8.0 * (double)c->ws_per_eval >= s->own_min_bytes;
}

/* The dataset's factorization configuration is fixed before any of its
   requests is served: an untuned dataset runs its autotuner now, on the
   request's parameters (synchronously, on the dataset's own stream and
   buffers, which no lane uses). */
static int bv_server_tune(bv_ctx *c, const bv_request &r) {
    if (!c->autotune || c->tuned) return BV_OK;
    double v = 0.0;
    int st = 0;
    return bv_run(c, 1, bv_nparams(c->kernel), r.theta, &v, &st);
}

/* One evaluation alone on its dataset's tuned path, enqueued on the dataset's
   stream without waiting (as bv_run_once, with the lane's pinned slot and
   result buffers so that other calls on the dataset cannot touch them). */
static int bv_server_launch_own(bv_server *s, bv_lane &L, const bv_request &r) {
    bv_ctx *c = s->datasets[r.dataset];
    int rc;
    if ((rc = bv_lane_reserve(L, 0, 0, 1, 0, 0))) return rc;
    if ((rc = bv_reserve(c, 1)) || (rc = bv_layout(c, 1))) return rc;
    bv_eval_slot &sl = L.h_slots[0];
    memset(&sl, 0, sizeof(sl));
    if ((rc = bv_make_kparams(c, r.theta, &sl.kp, &sl.kind))) return rc;
    sl.x = c->d_x; sl.y = c->d_y; sl.z = c->d_z; sl.obs = c->d_obs;
    sl.nbr = c->d_nbr; sl.cx = c->d_cx; sl.cy = c->d_cy; sl.cz = c->d_cz; sl.cobs = c->d_cobs;
    sl.cs = c->cs;
    sl.nblocks = c->B;
    sl.batch_num = c->d_batch_num;
    sl.logdet = c->d_logdet;
    sl.dot = c->d_dot;
    sl.info = c->d_info;
    sl.result = c->d_result;
    sl.log2pi = log(2 * 3.141592653589793);
    sl.kp.cheb = c->d_cheb;
    BV_HIP(hipSetDevice(s->device));
    BV_HIP(hipMemcpyAsync(c->d_slots, L.h_slots, sizeof(bv_eval_slot), hipMemcpyHostToDevice,
                          c->stream));
    const int kinds_used = 1 << sl.kind;
    rc = c->use_graph && c->graph_wins ? bv_enqueue_graph(c, (unsigned)c->B, kinds_used)
                                       : bv_enqueue(c, (unsigned)c->B, kinds_used);
    if (rc) return rc;
    BV_HIP(bv_launch_reduce(1u, c->stream, c->d_slots));
    BV_HIP(hipMemcpyAsync(L.h_result, c->d_result, sizeof(double) * 4, hipMemcpyDeviceToHost,
                          c->stream));
    BV_HIP(hipEventRecord(L.done, c->stream));
    c->srv_busy = 1;
    L.own = c;
    L.reqs.push_back(r);
    L.busy = true;
    return BV_OK;
}

/* Admit queued requests into the idle lane L and launch them. */
static int bv_server_launch(bv_server *s, bv_lane &L) {
    L.reqs.clear();
    L.own = nullptr;
    if (s->qhead < s->queue.size()) {
        const bv_request &r = s->queue[s->qhead];
        bv_ctx *c = s->datasets[r.dataset];
        int rc = bv_server_tune(c, r);
        if (rc) return rc;
        if (bv_server_own(s, c)) {
            if (c->srv_busy) return BV_OK; /* its dataset's path is in use: wait */
            ++s->qhead;
            return bv_server_launch_own(s, L, r);
        }
    }
    int64_t ws = 0, blocks = 0, items_bound = 0;
    int pmax_need = 1;
    while (s->qhead < s->queue.size() && (int)L.reqs.size() < s->max_evals) {
        const bv_request &r = s->queue[s->qhead];
        bv_ctx *c = s->datasets[r.dataset];
        if (!c->tuned && c->autotune) {
            if (!L.reqs.empty()) break; /* tune it at the head of the next batch */
            int rc = bv_server_tune(c, r);
            if (rc) return rc;
        }
        if (bv_server_own(s, c)) break; /* runs alone on its own path */
        const double bytes = 8.0 * (double)(ws + c->ws_per_eval);
        if (!L.reqs.empty() && bytes > s->ws_budget) break; /* admission by workspace bytes */
        ws += c->ws_per_eval;
        blocks += c->B;
        items_bound += bv_items_per_eval(c);
        pmax_need = std::max(pmax_need, c->max_mp / BV_NB);
        L.reqs.push_back(r);
        ++s->qhead;
    }
    const int k = (int)L.reqs.size();
    if (k == 0) return BV_OK;
    int rc;
    /* padding: at most 2 * 8 empty items per launch, at most 5 kinds x 3
       stepped forms of at most pmax_need + 1 launches each */
    items_bound += (int64_t)(pmax_need + 1) * 16 * 15;
    if ((rc = bv_lane_reserve(L, ws, blocks, k, items_bound, pmax_need))) return rc;
    /* groups: (kind, form), form = the dataset's factorization (0 fused,
       1..3 stepped forms, 4 register-resident) */
    const int NF = 5, NG = 8 * NF;
    int64_t gcount[NG] = {0};
    int gpmax[NG] = {0};
    std::vector<int> form_of(k);
    int64_t blk_off = 0;
    for (int e = 0; e < k; ++e) {
        const bv_request &r = L.reqs[e];
        const bv_ctx *c = s->datasets[r.dataset];
        bv_eval_slot &sl = L.h_slots[e];
        memset(&sl, 0, sizeof(sl));
        if ((rc = bv_make_kparams(c, r.theta, &sl.kp, &sl.kind))) return rc;
        sl.x = c->d_x; sl.y = c->d_y; sl.z = c->d_z; sl.obs = c->d_obs;
        sl.nbr = c->d_nbr; sl.cx = c->d_cx; sl.cy = c->d_cy; sl.cz = c->d_cz; sl.cobs = c->d_cobs;
        sl.cs = c->cs;
        sl.nblocks = c->B;
        sl.batch_num = c->d_batch_num;
        sl.logdet = L.d_logdet + blk_off;
        sl.dot = L.d_dot + blk_off;
        sl.info = L.d_info + blk_off;
        sl.result = L.d_result + 4 * e;
        sl.log2pi = log(2 * 3.141592653589793);
        sl.kp.cheb = L.d_cheb + (size_t)e * BV_CHEB_SIZE;
        form_of[e] = c->stepped ? c->stepped : (c->variant == 3 ? 4 : 0);
        const int g = sl.kind * NF + form_of[e];
        gcount[g] += c->B;
        gpmax[g] = std::max(gpmax[g], c->max_mp / BV_NB);
        blk_off += c->B;
    }
    /* descriptors grouped, each group sorted by padded size (largest first) */
    int64_t start[NG], fill[NG], gmax_mp[NG];
    int64_t acc = 0;
    for (int g = 0; g < NG; ++g) { start[g] = acc; fill[g] = acc; acc += gcount[g]; gmax_mp[g] = 0; }
    int64_t ws_off = 0;
    for (int e = 0; e < k; ++e) {
        const bv_ctx *c = s->datasets[L.reqs[e].dataset];
        const int g = L.h_slots[e].kind * NF + form_of[e];
        for (const bv_block_desc &d0 : c->base_desc) {
            bv_block_desc d = d0;
            d.eval = e;
            d.ws_off = d0.ws_off + ws_off;
            L.h_desc[fill[g]++] = d;
            gmax_mp[g] = std::max<int64_t>(gmax_mp[g], d0.mp);
        }
        ws_off += c->ws_per_eval;
    }
    for (int g = 0; g < NG; ++g)
        std::stable_sort(L.h_desc + start[g], L.h_desc + fill[g],
                         [](const bv_block_desc &a, const bv_block_desc &b) { return a.mp > b.mp; });
    /* step items of the stepped groups, one list per group, one chunk */
    std::vector<bv_step_item> items;
    std::vector<int64_t> goff[NG];
    int64_t gitem0[NG] = {0};
    int64_t nitems = 0;
    for (int g = 0; g < NG; ++g) {
        if (!gcount[g] || g % NF == 0 || g % NF == 4) continue;
        bv_build_items(L.h_desc + start[g], (size_t)gcount[g], g % NF, 1, false, 1, 1, gpmax[g],
                       items, goff[g]);
        if (nitems + (int64_t)items.size() > L.cap_items) return BV_ERR_BUSY; /* not sized */
        memcpy(L.h_items + nitems, items.data(), sizeof(bv_step_item) * items.size());
        gitem0[g] = nitems;
        nitems += (int64_t)items.size();
    }
    BV_HIP(hipSetDevice(s->device));
    BV_HIP(hipMemcpyAsync(L.d_slots, L.h_slots, sizeof(bv_eval_slot) * k, hipMemcpyHostToDevice,
                          L.stream));
    BV_HIP(hipMemcpyAsync(L.d_desc, L.h_desc, sizeof(bv_block_desc) * (size_t)blocks,
                          hipMemcpyHostToDevice, L.stream));
    if (nitems > 0)
        BV_HIP(hipMemcpyAsync(L.d_items, L.h_items, sizeof(bv_step_item) * (size_t)nitems,
                              hipMemcpyHostToDevice, L.stream));
    const int pst = L.cap_pmax; /* partial-sum stride */
    bool general = false;
    for (int g = 0; g < NG; ++g) general = general || (gcount[g] && g / NF == BV_KIND_GENERAL);
    if (general) BV_HIP(bv_launch_cheb_build((unsigned)k, L.stream, L.d_slots));
    for (int g = 0; g < NG; ++g) {
        const unsigned nblk = (unsigned)gcount[g];
        if (!nblk) continue;
        const int kd = g / NF, form = g % NF;
        const bv_block_desc *dd = L.d_desc + start[g];
        if (kd == BV_KIND_SPACETIME) {
            const unsigned T = (unsigned)(gmax_mp[g] / BV_TILE);
            BV_HIP(bv_launch_gen(kd, T * (T + 1) / 2, nblk, L.stream, dd, L.d_slots, L.d_ws));
        }
        if (form == 4) {
            BV_HIP(bv_launch_res_sorted(kd, L.h_desc + start[g], dd, nblk, L.stream, L.d_slots,
                                        L.d_ws));
            continue;
        }
        if (form == 0) {
            const size_t lds = sizeof(double) * 4 * (size_t)gmax_mp[g];
            if (gmax_mp[g] <= s->one_wave_max_mp)
                BV_HIP(bv_launch_block_1w(kd, nblk, lds, L.stream, dd, L.d_slots, L.d_ws));
            else
                BV_HIP(bv_launch_block(kd, nblk, lds, L.stream, dd, L.d_slots, L.d_ws));
            continue;
        }
        double *pl = L.d_part_logdet + start[g] * pst, *pd = L.d_part_dot + start[g] * pst;
        int32_t *pi = L.d_part_info + start[g] * pst;
        const bv_step_item *it0 = L.d_items + gitem0[g];
        for (int j = 0; j < gpmax[g]; ++j) {
            const int64_t o = goff[g][j], cnt = goff[g][j + 1] - o;
            if (cnt <= 0) continue;
            if (form == 3)
                BV_HIP(bv_launch_pair(kd, (unsigned)cnt, L.stream, it0 + o, dd, L.d_slots, L.d_ws,
                                      pl, pd, pi, pst));
            else if (form == 2)
                BV_HIP(bv_launch_step2(kd, (unsigned)cnt, L.stream, it0 + o, dd, L.d_slots,
                                       L.d_ws, pl, pd, pi, pst));
            else
                BV_HIP(bv_launch_step(kd, (unsigned)cnt, L.stream, it0 + o, dd, L.d_slots, L.d_ws,
                                      pl, pd, pi, pst));
        }
        BV_HIP(bv_launch_step_parts((int)nblk, L.stream, dd, L.d_slots, pl, pd, pi, pst));
    }
    BV_HIP(bv_launch_reduce((unsigned)k, L.stream, L.d_slots));
    BV_HIP(hipMemcpyAsync(L.h_result, L.d_result, sizeof(double) * 4 * k, hipMemcpyDeviceToHost,
                          L.stream));
    BV_HIP(hipEventRecord(L.done, L.stream));
    L.busy = true;
    return BV_OK;
}

extern "C" int bv_server_poll(bv_server *s, int wait, int max_out, int64_t *tags, double *llh,
                              int *status) {
    if (!s || max_out < 0 || (max_out > 0 && (!tags || !llh))) return BV_ERR_ARG;
    BV_HIP(hipSetDevice(s->device));
    int nout = 0;
    for (;;) {
        /* collect finished lanes (only whole batches, while there is room) */
        for (bv_lane &L : s->lanes) {
            if (!L.busy || nout + (int)L.reqs.size() > max_out) continue;
            const hipError_t q = hipEventQuery(L.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) return BV_ERR_HIP;
            for (size_t e = 0; e < L.reqs.size(); ++e) {
                const int bad = (int)L.h_result[4 * e + 1];
                tags[nout] = L.reqs[e].tag;
                llh[nout] = bad >= 0 ? -INFINITY : L.h_result[4 * e];
                if (status) status[nout] = bad >= 0 ? BV_ERR_NOT_PD : BV_OK;
                ++nout;
            }
            L.busy = false;
            L.reqs.clear();
            if (L.own) L.own->srv_busy = 0;
            L.own = nullptr;
        }
        /* start queued work on idle lanes */
        for (bv_lane &L : s->lanes) {
            if (L.busy || s->qhead >= s->queue.size()) continue;
            const int rc = bv_server_launch(s, L);
            if (rc != BV_OK) return rc;
        }
        if (nout > 0 || !wait) return nout;
        /* wait for the oldest busy lane that fits in the output */
        bv_lane *wl = nullptr;
        for (bv_lane &L : s->lanes)
            if (L.busy && nout + (int)L.reqs.size() <= max_out) { wl = &L; break; }
        if (!wl) return nout; /* nothing in flight (or no room to return it) */
        BV_HIP(hipEventSynchronize(wl->done));
    }
}

extern "C" int bv_server_pending(bv_server *s) {
    if (!s) return BV_ERR_ARG;
    int n = (int)(s->queue.size() - s->qhead);
    for (const bv_lane &L : s->lanes)
        if (L.busy) n += (int)L.reqs.size();
    return n;
}

extern "C" void bv_server_destroy(bv_server *s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    for (bv_lane &L : s->lanes) {
        if (L.busy && L.done) (void)hipEventSynchronize(L.done); /* own-path work too */
        if (L.own) L.own->srv_busy = 0;
        if (L.stream) (void)hipStreamSynchronize(L.stream);
        bv_lane_free(L);
        if (L.done) (void)hipEventDestroy(L.done);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    delete s;
}

/* ---- Block Vecchia simulation ------------------------------------------
   bv_simulate draws y from the Block Vecchia model of the context's
   partition: block 0 from its own covariance, block i from y_B | y_N(i).
   Blocks are grouped into dependency levels (a block's level is one more
   than the highest level of a block holding one of its conditioning
   points), so that every block of a level conditions on values of earlier
   levels only.  Per level: the fused factorization in sample mode of the
   augmented matrix [S_NN, S_NB; S_BN, S_BB] with row m = [y_N, 0] (general
   nu: generation first), then bv_sim_finish_kernel writes
   y_B = L_BN (L_NN^-1 y_N) + L_S z_B. */
extern "C" int bv_simulate(bv_ctx *c, unsigned n, const double *theta, const double *z,
                           double nugget, double *y, int *levels) {
    if (!c || !theta || !z || !y || (int)n != bv_nparams(c->kernel) || !(nugget >= 0.0))
        return BV_ERR_ARG;
    if (!c->fused_ok) {
        fprintf(stderr, "libbv: bv_simulate factors with the fused kernel, which takes padded "
                        "augmented sizes up to %d (this dataset needs %d)\n", BV_MAX_MP,
                c->max_mp);
        return BV_ERR_UNSUPPORTED;
    }
    BV_HIP(hipSetDevice(c->device));
    int rc;
    if ((rc = bv_reserve(c, 1))) return rc;
    bv_eval_slot sl;
    memset(&sl, 0, sizeof(sl));
    if ((rc = bv_make_kparams(c, theta, &sl.kp, &sl.kind))) return rc;
    sl.kp.nugget = nugget;
    /* dependency levels */
    const int B = c->B, cs = c->cs;
    std::vector<int32_t> level((size_t)B, 0), nbr;
    if (cs > 0 && B > 1) {
        if (!c->d_nbr) return BV_ERR_UNSUPPORTED; /* needs neighbour indices */
        nbr.resize((size_t)B * cs);
        BV_HIP(hipMemcpy(nbr.data(), c->d_nbr, sizeof(int32_t) * nbr.size(),
                         hipMemcpyDeviceToHost));
        std::vector<int32_t> owner((size_t)c->n);
        for (int i = 0; i < B; ++i)
            for (int64_t p = c->h_accum[i]; p < c->h_accum[i + 1]; ++p) owner[(size_t)p] = i;
        for (int i = 1; i < B; ++i) {
            int lv = 0;
            for (int k = 0; k < cs; ++k) {
                const int o = owner[(size_t)nbr[(size_t)i * cs + k]];
                if (o >= i) {
                    fprintf(stderr, "libbv: bv_simulate: block %d conditions on block %d, not an "
                                    "earlier one\n", i, o);
                    return BV_ERR_ARG;
                }
                lv = std::max(lv, level[(size_t)o] + 1);
            }
            level[(size_t)i] = lv;
        }
    }
    const int nlev = B ? *std::max_element(level.begin(), level.end()) + 1 : 0;
    /* descriptors ordered by level, largest padded size first inside one */
    std::vector<bv_block_desc> all(c->base_desc);
    std::stable_sort(all.begin(), all.end(), [&](const bv_block_desc &a, const bv_block_desc &b) {
        return level[(size_t)a.block] < level[(size_t)b.block];
    });
    std::vector<int64_t> lstart((size_t)nlev + 1, 0);
    std::vector<int32_t> lmaxmp((size_t)nlev, 0);
    for (const bv_block_desc &d : all) {
        const int lv = level[(size_t)d.block];
        lstart[(size_t)lv + 1] += 1;
        lmaxmp[(size_t)lv] = std::max(lmaxmp[(size_t)lv], d.mp);
    }
    for (int l = 0; l < nlev; ++l) lstart[(size_t)l + 1] += lstart[(size_t)l];
    /* device buffers of the draw: output (read back as y_N), z, descriptors */
    double *d_y = nullptr, *d_z = nullptr;
    bv_block_desc *d_desc = nullptr;
    const size_t nb = sizeof(double) * (size_t)c->n;
    rc = BV_OK;
    auto done = [&](int code) {
        c->last_status = code;
        if (d_y) (void)hipFree(d_y);
        if (d_z) (void)hipFree(d_z);
        if (d_desc) (void)hipFree(d_desc);
        return code;
    };
    if (hipMalloc((void **)&d_y, nb) != hipSuccess || hipMalloc((void **)&d_z, nb) != hipSuccess ||
        hipMalloc((void **)&d_desc, sizeof(bv_block_desc) * all.size()) != hipSuccess)
        return done(BV_ERR_HIP);
    if (hipMemsetAsync(d_y, 0, nb, c->stream) != hipSuccess ||
        hipMemcpyAsync(d_z, z, nb, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(d_desc, all.data(), sizeof(bv_block_desc) * all.size(),
                       hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return done(BV_ERR_HIP);
    sl.x = c->d_x; sl.y = c->d_y; sl.z = c->d_z; sl.obs = d_y;
    sl.nbr = c->d_nbr; sl.cx = c->d_cx; sl.cy = c->d_cy; sl.cz = c->d_cz; sl.cobs = c->d_cobs;
    sl.cs = c->cs;
    sl.nblocks = c->B;
    sl.batch_num = c->d_batch_num;
    sl.logdet = c->d_logdet;
    sl.dot = c->d_dot;
    sl.info = c->d_info;
    sl.result = c->d_result;
    sl.log2pi = log(2 * 3.141592653589793);
    sl.flags = BV_SLOT_SAMPLE;
    sl.kp.cheb = c->d_cheb;
    c->h_slots[0] = sl;
    if (hipMemcpyAsync(c->d_slots, c->h_slots, sizeof(bv_eval_slot), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess)
        return done(BV_ERR_HIP);
    if (sl.kind == BV_KIND_GENERAL &&
        bv_launch_cheb_build(1u, c->stream, c->d_slots) != hipSuccess)
        return done(BV_ERR_HIP);
    for (int l = 0; l < nlev; ++l) {
        const bv_block_desc *dd = d_desc + lstart[(size_t)l];
        const unsigned cnt = (unsigned)(lstart[(size_t)l + 1] - lstart[(size_t)l]);
        if (!cnt) continue;
        const int mp = lmaxmp[(size_t)l];
        hipError_t e = hipSuccess;
        if (sl.kind == BV_KIND_SPACETIME) {
            const unsigned T = (unsigned)(mp / BV_TILE);
            e = bv_launch_gen(sl.kind, T * (T + 1) / 2, cnt, c->stream, dd, c->d_slots, c->d_ws);
        }
        if (e == hipSuccess)
            e = bv_launch_block(sl.kind, cnt, sizeof(double) * 4 * (size_t)mp, c->stream, dd,
                                c->d_slots, c->d_ws);
        if (e == hipSuccess) e = bv_launch_sim_finish(cnt, c->stream, dd, c->d_slots, c->d_ws, d_z);
        if (e != hipSuccess) {
            fprintf(stderr, "libbv: bv_simulate level %d: %s\n", l, hipGetErrorString(e));
            return done(BV_ERR_HIP);
        }
    }
    std::vector<int32_t> info((size_t)B);
    if (hipMemcpyAsync(y, d_y, nb, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(info.data(), c->d_info, sizeof(int32_t) * (size_t)B,
                       hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return done(BV_ERR_HIP);
    if (levels) *levels = nlev;
    c->last_k = 0; /* the per-block terms are not a likelihood evaluation's */
    for (int i = 0; i < B; ++i)
        if (info[(size_t)i] != 0) {
            c->last_bad = i;
            return done(BV_ERR_NOT_PD);
        }
    return done(BV_OK);
}

/* ---- evaluation without allocation --------------------------------------
   bv_llh_noalloc: one evaluation of a dataset that the caller keeps in device
   memory, with every buffer the kernels need beyond the inputs and the
   result (launch descriptors, the evaluation slot, block sizes, per-block
   terms, the factor workspace) carved out of `scratch`, a file-scope device
   array of BV_NOALLOC_SCRATCH_BYTES.  Each device holds its own copy of the
   array, at its own address, from the library's load on that device until
   the process ends; the host looks that address up once per device
   (hipGetSymbolAddress, under a lock) and passes it to the kernels.  Every
   call on a device uses the same array: calls that overlap on one device
   race.  Nothing in it needs to start at zero: the kernels write every tile,
   term and result before reading it. */
static __device__ __align__(256) unsigned char scratch[BV_NOALLOC_SCRATCH_BYTES];

#define BV_MAX_DEVICES 64
static unsigned char *bv_scratch_dev[BV_MAX_DEVICES];
static std::mutex bv_scratch_mu;

static int bv_scratch_ptr(int device, unsigned char **out) {
    std::lock_guard<std::mutex> lk(bv_scratch_mu);
    if (device < 0 || device >= BV_MAX_DEVICES) return BV_ERR_ARG;
    if (!bv_scratch_dev[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            fprintf(stderr, "libbv: hipGetSymbolAddress(scratch) failed on device %d: %s\n",
                    device, hipGetErrorString(e));
            return BV_ERR_HIP;
        }
        bv_scratch_dev[device] = (unsigned char *)ptr;
    }
    *out = bv_scratch_dev[device];
    return BV_OK;
}

static size_t bv_align256(size_t v) { return (v + 255) & ~(size_t)255; }

extern "C" int bv_llh_noalloc(const bv_problem *p, unsigned n, const double *theta, double *llh,
                              int *first_bad_block) {
    if (!p || !theta || !llh) return BV_ERR_ARG;
    if (first_bad_block) *first_bad_block = -1;
    *llh = NAN;
    if (p->n <= 0 || p->batch_count <= 0 || p->cs < 0 || !p->batch_num || !p->x || !p->y ||
        !p->obs)
        return BV_ERR_ARG;
    if (p->kernel < 1 || p->kernel > 4) return BV_ERR_ARG;
    if ((p->time_flag || p->kernel == BV_KERNEL_SPACETIME) && !p->z) return BV_ERR_ARG;
    if (p->distance_metric == 1 && p->time_flag) return BV_ERR_UNSUPPORTED;
    if (p->cs > 0 && p->batch_count > 1 && !p->nbr && !(p->con_x && p->con_y && p->con_obs))
        return BV_ERR_ARG;
    if ((int)n != bv_nparams(p->kernel)) return BV_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || p->device < 0 ||
        p->device >= ndev)
        return BV_ERR_NO_DEVICE;
    BV_HIP(hipSetDevice(p->device));
    unsigned char *base = nullptr;
    int rc = bv_scratch_ptr(p->device, &base);
    if (rc) return rc;

    /* covariance parameters, as bv_create + bv_make_kparams do */
    bv_ctx tmp;
    tmp.kernel = p->kernel;
    tmp.metric = p->distance_metric;
    tmp.time_flag = p->time_flag || p->kernel == BV_KERNEL_SPACETIME;
    tmp.dist_scale = p->dist_scale > 0 ? p->dist_scale : 1.0;
    tmp.d_z = tmp.time_flag ? (double *)p->z : nullptr;
    bv_eval_slot slot;
    memset(&slot, 0, sizeof(slot));
    if ((rc = bv_make_kparams(&tmp, theta, &slot.kp, &slot.kind))) return rc;

    /* descriptors (as bv_create), largest padded size first */
    const int B = p->batch_count;
    std::vector<bv_block_desc> desc((size_t)B);
    int64_t off = 0, acc = 0;
    int max_mp = 0;
    for (int i = 0; i < B; ++i) {
        if (p->batch_num[i] <= 0) return BV_ERR_ARG;
        bv_block_desc &d = desc[i];
        memset(&d, 0, sizeof(d));
        d.block = i;
        d.cs0 = (i == 0 || p->cs == 0) ? 0 : p->cs;
        d.m = d.cs0 + p->batch_num[i];
        d.mp = ((d.m + 1 + BV_NB - 1) / BV_NB) * BV_NB;
        d.obs_off = acc;
        acc += p->batch_num[i];
        const int64_t T = d.mp / BV_TILE;
        d.ws_off = off;
        off += T * (T + 1) / 2 * 256;
        max_mp = std::max(max_mp, d.mp);
    }
    if (acc != p->n) return BV_ERR_ARG;
    if (max_mp > BV_MAX_MP) {
        fprintf(stderr, "libbv: bv_llh_noalloc factors with the fused kernel, which takes padded "
                        "augmented sizes up to %d (this dataset needs %d)\n", BV_MAX_MP, max_mp);
        return BV_ERR_UNSUPPORTED;
    }
    std::stable_sort(desc.begin(), desc.end(),
                     [](const bv_block_desc &a, const bv_block_desc &b) { return a.mp > b.mp; });

    /* scratch layout */
    const size_t o_slot = 0;
    const size_t o_res = o_slot + bv_align256(sizeof(bv_eval_slot));
    const size_t o_bn = o_res + 256;
    const size_t o_desc = o_bn + bv_align256(sizeof(int32_t) * (size_t)B);
    const size_t o_ld = o_desc + bv_align256(sizeof(bv_block_desc) * (size_t)B);
    const size_t o_dot = o_ld + bv_align256(sizeof(double) * (size_t)B);
    const size_t o_info = o_dot + bv_align256(sizeof(double) * (size_t)B);
    const size_t o_cheb = o_info + bv_align256(sizeof(int32_t) * (size_t)B);
    const size_t o_ws = o_cheb + bv_align256(sizeof(double) * BV_CHEB_SIZE);
    const size_t need = o_ws + sizeof(double) * (size_t)off;
    if (need > (size_t)BV_NOALLOC_SCRATCH_BYTES) {
        fprintf(stderr, "libbv: bv_llh_noalloc needs %zu bytes of scratch for this dataset, more "
                        "than the %llu it holds\n", need,
                (unsigned long long)BV_NOALLOC_SCRATCH_BYTES);
        return BV_ERR_SCRATCH;
    }
    bv_eval_slot *d_slot = (bv_eval_slot *)(base + o_slot);
    double *d_res = (double *)(base + o_res);
    int32_t *d_bn = (int32_t *)(base + o_bn);
    bv_block_desc *d_desc = (bv_block_desc *)(base + o_desc);
    double *d_ws = (double *)(base + o_ws);
    slot.x = p->x; slot.y = p->y; slot.z = tmp.d_z; slot.obs = p->obs;
    if (p->cs > 0 && p->nbr) {
        slot.nbr = p->nbr;
    } else if (p->cs > 0) {
        slot.cx = p->con_x; slot.cy = p->con_y; slot.cobs = p->con_obs;
        slot.cz = tmp.time_flag ? p->con_z : nullptr;
    }
    slot.cs = p->cs;
    slot.nblocks = B;
    slot.batch_num = d_bn;
    slot.logdet = (double *)(base + o_ld);
    slot.dot = (double *)(base + o_dot);
    slot.info = (int32_t *)(base + o_info);
    slot.result = d_res;
    slot.log2pi = log(2 * 3.141592653589793);
    slot.kp.cheb = (const double *)(base + o_cheb);
    const hipStream_t st = 0;
    BV_HIP(hipMemcpyAsync(d_slot, &slot, sizeof(slot), hipMemcpyHostToDevice, st));
    BV_HIP(hipMemcpyAsync(d_bn, p->batch_num, sizeof(int32_t) * (size_t)B, hipMemcpyHostToDevice,
                          st));
    BV_HIP(hipMemcpyAsync(d_desc, desc.data(), sizeof(bv_block_desc) * (size_t)B,
                          hipMemcpyHostToDevice, st));
    if (slot.kind == BV_KIND_GENERAL) BV_HIP(bv_launch_cheb_build(1u, st, d_slot));
    if (slot.kind == BV_KIND_SPACETIME) {
        const unsigned T = (unsigned)(max_mp / BV_TILE);
        BV_HIP(bv_launch_gen(slot.kind, T * (T + 1) / 2, (unsigned)B, st, d_desc, d_slot, d_ws));
    }
    BV_HIP(bv_launch_block(slot.kind, (unsigned)B, sizeof(double) * 4 * (size_t)max_mp, st,
                           d_desc, d_slot, d_ws));
    BV_HIP(bv_launch_reduce(1u, st, d_slot));
    double res[4];
    BV_HIP(hipMemcpyAsync(res, d_res, sizeof(res), hipMemcpyDeviceToHost, st));
    BV_HIP(hipStreamSynchronize(st));
    const int bad = (int)res[1];
    if (bad >= 0) {
        if (first_bad_block) *first_bad_block = bad;
        *llh = -INFINITY;
        return BV_ERR_NOT_PD;
    }
    *llh = res[0];
    return BV_OK;
}

/* ---- pre-processing with the k-means assignment step on the device ------ */

struct bv_km_dev {
    hipStream_t st = nullptr;
    int64_t n = 0;
    int kcap = 0;
    double *px = nullptr, *py = nullptr, *pz = nullptr;
    double *cx = nullptr, *cy = nullptr, *cz = nullptr;
    int32_t *cl = nullptr;
};

static int bv_km_assign(void *user, int k, const double *cx, const double *cy, const double *cz,
                        int32_t *cluster) {
    bv_km_dev &D = *(bv_km_dev *)user;
    if (k < 1 || k > D.kcap) return BV_ERR_ARG;
    BV_HIP(hipMemcpyAsync(D.cx, cx, sizeof(double) * k, hipMemcpyHostToDevice, D.st));
    BV_HIP(hipMemcpyAsync(D.cy, cy, sizeof(double) * k, hipMemcpyHostToDevice, D.st));
    BV_HIP(hipMemcpyAsync(D.cz, cz, sizeof(double) * k, hipMemcpyHostToDevice, D.st));
    BV_HIP(bv_launch_kmeans_assign(D.n, D.px, D.py, D.pz, k, D.cx, D.cy, D.cz, D.cl, D.st));
    BV_HIP(hipMemcpyAsync(cluster, D.cl, sizeof(int32_t) * D.n, hipMemcpyDeviceToHost, D.st));
    BV_HIP(hipStreamSynchronize(D.st));
    return BV_OK;
}

static void bv_km_free(bv_km_dev &D) {
    double *b[] = {D.px, D.py, D.pz, D.cx, D.cy, D.cz};
    for (double *p : b)
        if (p) (void)hipFree(p);
    if (D.cl) (void)hipFree(D.cl);
    if (D.st) (void)hipStreamDestroy(D.st);
    D = bv_km_dev();
}

static int bv_km_setup(bv_km_dev &D, int64_t n, const double *x, const double *y,
                       const double *z, int32_t time_flag, int kcap) {
    D.n = n;
    D.kcap = kcap;
    BV_HIP(hipStreamCreateWithFlags(&D.st, hipStreamNonBlocking));
    const size_t nb = sizeof(double) * (size_t)n, kb = sizeof(double) * (size_t)kcap;
    BV_HIP(hipMalloc((void **)&D.px, nb));
    BV_HIP(hipMalloc((void **)&D.py, nb));
    BV_HIP(hipMalloc((void **)&D.pz, nb));
    BV_HIP(hipMalloc((void **)&D.cx, kb));
    BV_HIP(hipMalloc((void **)&D.cy, kb));
    BV_HIP(hipMalloc((void **)&D.cz, kb));
    BV_HIP(hipMalloc((void **)&D.cl, sizeof(int32_t) * (size_t)n));
    BV_HIP(hipMemcpyAsync(D.px, x, nb, hipMemcpyHostToDevice, D.st));
    BV_HIP(hipMemcpyAsync(D.py, y, nb, hipMemcpyHostToDevice, D.st));
    if (time_flag) {
        BV_HIP(hipMemcpyAsync(D.pz, z, nb, hipMemcpyHostToDevice, D.st));
    } else {
        BV_HIP(hipMemsetAsync(D.pz, 0, nb, D.st)); /* z = 0 in 2-D, as on the host */
    }
    BV_HIP(hipStreamSynchronize(D.st));
    return BV_OK;
}

/* The conditioning sets on the device (bv_knn_kernel), for bv_partition_run. */
static int bv_knn_dev(void *user, int64_t n, const double *x, const double *y, const double *z,
                      int64_t B, const int64_t *accum, const double *cx, const double *cy,
                      const double *cz, int32_t cs, int32_t time_flag, int32_t *nbr,
                      int32_t *redo) {
    bv_km_dev &D = *(bv_km_dev *)user;
    const size_t nb = sizeof(double) * (size_t)n, bb = sizeof(double) * (size_t)B;
    double *dx = nullptr, *dy = nullptr, *dz = nullptr, *dcx = nullptr, *dcy = nullptr,
           *dcz = nullptr;
    int64_t *dacc = nullptr;
    int32_t *dnbr = nullptr, *dredo = nullptr;
    int rc = BV_ERR_HIP;
    const size_t nn = (size_t)B * (size_t)cs;
    if (hipMalloc((void **)&dx, nb) == hipSuccess && hipMalloc((void **)&dy, nb) == hipSuccess &&
        (!time_flag || hipMalloc((void **)&dz, nb) == hipSuccess) &&
        hipMalloc((void **)&dcx, bb) == hipSuccess && hipMalloc((void **)&dcy, bb) == hipSuccess &&
        hipMalloc((void **)&dcz, bb) == hipSuccess &&
        hipMalloc((void **)&dacc, sizeof(int64_t) * (size_t)(B + 1)) == hipSuccess &&
        hipMalloc((void **)&dnbr, sizeof(int32_t) * nn) == hipSuccess &&
        hipMalloc((void **)&dredo, sizeof(int32_t) * (size_t)B) == hipSuccess &&
        hipMemcpyAsync(dx, x, nb, hipMemcpyHostToDevice, D.st) == hipSuccess &&
        hipMemcpyAsync(dy, y, nb, hipMemcpyHostToDevice, D.st) == hipSuccess &&
        (!time_flag || hipMemcpyAsync(dz, z, nb, hipMemcpyHostToDevice, D.st) == hipSuccess) &&
        hipMemcpyAsync(dcx, cx, bb, hipMemcpyHostToDevice, D.st) == hipSuccess &&
        hipMemcpyAsync(dcy, cy, bb, hipMemcpyHostToDevice, D.st) == hipSuccess &&
        hipMemcpyAsync(dcz, cz, bb, hipMemcpyHostToDevice, D.st) == hipSuccess &&
        hipMemcpyAsync(dacc, accum, sizeof(int64_t) * (size_t)(B + 1), hipMemcpyHostToDevice,
                       D.st) == hipSuccess &&
        hipMemsetAsync(dredo, 0, sizeof(int32_t) * (size_t)B, D.st) == hipSuccess &&
        bv_launch_knn(dx, dy, dz, dacc, dcx, dcy, dcz, B, cs, time_flag, dnbr, dredo, D.st) ==
            hipSuccess &&
        /* row 0 is the host's (0..cs-1); rows 1..B-1 come back */
        hipMemcpyAsync(nbr + cs, dnbr + cs, sizeof(int32_t) * (nn - (size_t)cs),
                       hipMemcpyDeviceToHost, D.st) == hipSuccess &&
        hipMemcpyAsync(redo, dredo, sizeof(int32_t) * (size_t)B, hipMemcpyDeviceToHost, D.st) ==
            hipSuccess &&
        hipStreamSynchronize(D.st) == hipSuccess)
        rc = BV_OK;
    void *bufs[] = {dx, dy, dz, dcx, dcy, dcz, dacc, dnbr, dredo};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    return rc;
}

extern "C" int bv_partition_gpu(int32_t device, int64_t n, const double *x, const double *y,
                                const double *z, const double *obs, int32_t bc, int32_t cs,
                                int32_t seed, int32_t epochs, int32_t time_flag, int32_t metric,
                                int32_t threads, bv_blocks *out) {
    if (!out || n <= 0 || n > 2147483647LL || !x || !y || !obs || bc <= 0) return BV_ERR_ARG;
    if (time_flag && !z) return BV_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev)
        return BV_ERR_NO_DEVICE;
    BV_HIP(hipSetDevice(device));
    bv_km_dev D;
    const bool km = n >= 2 * (int64_t)bc; /* else classic Vecchia: no k-means */
    int rc = BV_OK;
    if (km) rc = bv_km_setup(D, n, x, y, z, time_flag, bc);
    else if (hipStreamCreateWithFlags(&D.st, hipStreamNonBlocking) != hipSuccess)
        rc = BV_ERR_HIP;
    /* the kNN on the device unless BV_KNN_HOST=1 (great-circle distances stay
       on the host: their library sin/cos/asin round differently) */
    const char *kh = getenv("BV_KNN_HOST");
    const bool dev_knn = !(kh && atoi(kh) != 0);
    if (rc == BV_OK)
        rc = bv_partition_run(n, x, y, z, obs, bc, cs, seed, epochs, time_flag, metric, threads,
                              km ? bv_km_assign : nullptr, &D, out,
                              dev_knn ? bv_knn_dev : nullptr);
    bv_km_free(D);
    return rc;
}
