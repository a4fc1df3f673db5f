/*
 * bv_resident.hip -- the register-resident factorization of the Block Vecchia
 * blocks: one workgroup of four wavefronts per block, no factor in memory.
 *
 * bv_res_kernel<KIND, NR> factors the augmented matrix of bv_engine.h of one
 * block with T <= 4 NR tile rows (padded size Mp = 16 T) by a left-looking
 * Cholesky one 16-column tile column at a time.  Wavefront w owns the tile
 * rows r = w + 4 i (i < NR).  Every factor tile L(r, k) stays in that
 * wavefront's registers from its solve (column k) until the column of its own
 * diagonal, the last one that reads it; nothing of the factor is written to
 * memory.  Column c (a compile-time constant: the column loop is unrolled, so
 * every register index is static):
 *   - C(r, c) = A(r, c) - sum_k L(r, k) L(c, k)^T for the own rows r >= c,
 *     FP64 matrix-core products with the own tile as one operand and pivot
 *     row c's tile (LDS, negated) as the other;
 *   - the owner of row c factors the 16x16 diagonal C(c, c) in registers
 *     (L_cc, W_c = L_cc^-1, pivots, row m's squares) and publishes W_c in
 *     LDS; its matrix-core jobs of the column run between the sweep's steps;
 *   - L(r, c) = C(r, c) W_c^T for the own rows below (rows c+1 and c+2 right
 *     after W_c, so their owners can publish them; the others at the start of
 *     column c+1, where they also give C(r, c+1) its last term).
 * Lookahead: C(r, c+1) over k < c is formed during column c (seeded with
 * A(r, c+1) from LDS), so each column costs one diagonal sweep, the last
 * term of the next pivot block and two barriers, the products filling the
 * rest.  The wavefronts that do not own column c's diagonal generate the
 * covariance tiles of column c+2 into LDS.
 *
 * LDS: two slots of 4 NR tiles (slot j % 2: pivot row j's tiles at t < j,
 * column j's covariance tiles at t >= j), two W slots, the staged points
 * (4 x 64 NR doubles), the exponential's table: 42, 61, 79, 97 and 113 KB for
 * NR = 2 .. 6.  NR = 2 and 3 (padded sizes up to 128 and 192) run two
 * workgroups per compute unit (two wavefronts per SIMD, up to 256 registers
 * each: one block's vector work, covariance generation and diagonal sweeps,
 * overlaps the other's matrix-core products); NR = 4 .. 6 one (a wavefront's
 * live tiles need up to 512 registers).  Rows a wavefront holds only
 * statically (finished, or beyond the block) are skipped by wave-uniform
 * branches or carry values nobody reads.
 */
#include <hip/hip_runtime.h>
#include "bv_engine.h"
#include "bv_device.h"

#define BV_RES_THREADS 256

namespace {

template <int NR>
struct bv_res_state {
    bv_d4 L[NR][4 * NR]; /* L(w + 4 i, k): own tile row i, tile column k      */
    bv_d4 acc[NR];       /* C(w + 4 i, c) of the column being factored         */
    bv_d4 nxt[NR];       /* C(w + 4 i, c + 1), formed over the columns k < c   */
    double lmant;        /* this lane's owned pivots (lane = pivot): product of */
    int lexp;            /* their mantissas in [1/2, 1) and sum of exponents    */
    double dot;          /* this lane's squares of row m's solved entries       */
    int info;            /* first failing column + 1, 0 if none (uniform)       */
};

struct bv_res_block {
    int T, NC, m, cs0, w, lane;
    const double *sx, *sy, *sz, *so;  /* staged points of the block (LDS)        */
    const double *etab;               /* 2^(j/64) as hi, lo pairs (LDS)          */
    const double *cheb;               /* general nu: the x^nu K_nu table (LDS)   */
    double *buf;                      /* two slots of pivot-row / column tiles   */
    double *sW;                       /* two W slots (LDS)                       */
    const double *ws;                 /* LOADA: the block's A tiles (workspace)  */
};

__device__ __forceinline__ bv_d4 bv_lds_tile(const double *t, int lane) {
    return *(const bv_d4 *)(t + 4 * lane);
}

__device__ __forceinline__ void bv_lds_put(double *t, int lane, const bv_d4 &v) {
    *(bv_d4 *)(t + 4 * lane) = v;
}

/* Euclidean distance in the plane or in 3-D (dz = 0 in the plane adds an
   exact zero), products and sums rounded separately as the reference's
   sqrt(pow(dx,2) + pow(dy,2) [+ pow(dz,2)]). */
__device__ __forceinline__ double bv_res_dist(double xp, double yp, double zp, double xq,
                                              double yq, double zq, bool use_z) {
#pragma clang fp contract(off)
    const double dx = xq - xp, dy = yq - yp;
    const double dz = use_z ? zq - zp : 0.0;
    double s = dx * dx + dy * dy;
    s = s + dz * dz;
    return bv_sqrt_nr(s);
}

/* The covariance at distance d: the closed forms of bv_matern_closed with
   the table exponential, or general nu's entries of bv_matern_planar, with
   the tables in LDS. */
template <int KIND>
__device__ __forceinline__ double bv_res_entry(double d, const bv_res_block &B,
                                               const bv_kparams &kp) {
    const double t = d * kp.inv_scale;
    if (KIND == BV_KIND_GENERAL) return bv_matern_general_x(t, kp, B.cheb, B.etab);
    const double x = bv_expneg_tab(t, B.etab);
    if (KIND == BV_KIND_NU05) return kp.sigma2 * x;
    if (KIND == BV_KIND_NU15) return kp.sigma2 * (1.0 + t) * x;
    return kp.sigma2 * (1.0 + t + t * t * (1.0 / 3.0)) * x;
}

/* A(r, c) in accumulator layout (lane: row 16 r + lane % 16, columns
   16 c + 4 g + lane / 16).  Entries outside the block (rows or columns >= m)
   follow the augmented matrix: row m holds the observations, the padding is
   the identity; every choice is a select.  Euclidean distances only (the
   great-circle metric takes the other factorization forms). */
template <int KIND>
__device__ __forceinline__ bv_d4 bv_res_gen(const bv_res_block &B, int r, int c,
                                            const bv_kparams &kp) {
    const int lr = B.lane & 15, lq = B.lane >> 4;
    const int p = 16 * r + lr;
    bv_d4 v;
    if (KIND == BV_KIND_LOADA) {
        if (r >= c && r < B.T) return *(const bv_d4 *)(B.ws + bv_tile_off(r, c, B.T) + 4 * B.lane);
#pragma unroll
        for (int g = 0; g < 4; ++g) v[g] = (p == 16 * c + 4 * g + lq) ? 1.0 : 0.0;
        return v;
    } else {
        const double xp = B.sx[p], yp = B.sy[p], zp = B.sz[p];
        const bool use_z = kp.use_z != 0;
        if (!use_z && 16 * r + 15 < B.m && 16 * c + 15 < B.m) {
            /* a tile inside the block, planar distances (wave-uniform): the
               same entries without the augmented matrix's selects */
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int q = 16 * c + 4 * g + lq;
                v[g] = bv_res_entry<KIND>(bv_res_dist(xp, yp, 0.0, B.sx[q], B.sy[q], 0.0, false),
                                          B, kp);
            }
            return v;
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int q = 16 * c + 4 * g + lq;
            const double d = bv_res_dist(xp, yp, zp, B.sx[q], B.sy[q], B.sz[q], use_z);
            const double e = bv_res_entry<KIND>(d, B, kp);
            const double pad = (p == B.m && q < B.m) ? B.so[q] : ((p == q) ? 1.0 : 0.0);
            v[g] = (p < B.m && q < B.m) ? e : pad;
        }
        return v;
    }
}

/* Squares of row m's entries of the solved tile x = L(r, k) (r > k), for the
   lanes holding row m, in the block columns [cs0, m). */
__device__ __forceinline__ double bv_res_msq(const bv_d4 &x, int r, int k, const bv_res_block &B) {
    const int lr = B.lane & 15, lq = B.lane >> 4;
    const bool rowm = (16 * r + lr == B.m) && r > k;
    double t = 0.0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int q = 16 * k + 4 * g + lq;
        t += (rowm && q >= B.cs0 && q < B.m) ? x[g] * x[g] : 0.0;
    }
    return t;
}

/* Step K of the register sweep of a 16x16 diagonal tile (the sweep of
   bv_kernels.hip without branches or LDS): l_KK = sqrt(a_KK) from a
   readlane, the column-K factors of the other rows taken from row K by DPP
   row broadcast (the tile is exactly symmetric where it matters), W = L^-1
   built in the same step.  The pivot a_KK lands in lane K of piv. */
template <int K>
__device__ __forceinline__ void bv_res_sweep_step(bv_d4 &a, bv_d4 &w, int lane, int q0, int m,
                                                  int cs0, double &piv, double &dacc, int &info) {
    constexpr int kq = K & 3, kg = K >> 2;
    const int lr = lane & 15, lq = lane >> 4;
    const int q = q0 + K;
    const int kml = m - q0;
    const double akk = bv_readlane(a[kg], K + 16 * kq);
    /* 1/l_KK: hardware reciprocal square root and two Newton steps; columns
       at or beyond m (row m, padding) keep a pivot inverse of 1 */
    const double ha = 0.5 * akk;
    double y = __builtin_amdgcn_rsq(akk);
    y = fma(y, fma(-ha * y, y, 0.5), y);
    y = fma(y, fma(-ha * y, y, 0.5), y);
    const double pinv = (q < m) ? y : 1.0;
    info = (q < m && !(akk > 0.0) && info == 0) ? q + 1 : info;
    piv = (lane == K) ? akk : piv;
    const double t = a[kg] * pinv;
    const double lrk = __shfl(t, lr + 16 * kq, 64);
    dacc += (lr == kml && K < kml && q >= cs0 && lq == 0) ? lrk * lrk : 0.0;
    const double lre = (lr > K) ? lrk : 0.0;
    const double ws = (lr == K) ? pinv : 1.0;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const double lk = bv_row_bcast<K>(a[g]) * pinv;
        const double wk = bv_row_bcast<K>(w[g]) * pinv;
        a[g] -= lrk * lk;
        w[g] = fma(-lre, wk, w[g] * ws);
    }
}

template <int K>
__device__ __forceinline__ void bv_res_sweep_steps(bv_d4 &a, bv_d4 &w, int lane, int q0, int m,
                                                   int cs0, double &piv, double &dacc,
                                                   int &info) {
    bv_res_sweep_step<K>(a, w, lane, q0, m, cs0, piv, dacc, info);
    if constexpr (K + 1 < 16) bv_res_sweep_steps<K + 1>(a, w, lane, q0, m, cs0, piv, dacc, info);
}

/* LDS tile (slot b, index t): slot j % 2 holds pivot row j's finished tiles
   L(j, k), negated, at t = k < j, and column j's covariance tiles A(r, j) at
   t = r >= j (generated two columns ahead by the wavefronts that do not own
   the diagonal).  Row j is copied there after column j - 2's diagonal, column
   j is generated during column j - 2, and each is dead before the other half
   of the slot is rewritten. */
__device__ __forceinline__ double *bv_res_tile(const bv_res_block &B, int slot, int t) {
    return B.buf + ((slot * B.NC + t) << 8);
}

/* Jobs of column C that do not depend on its diagonal factorization, in
   dependency order: for each own row i > q (q = C / 4) its solve of column
   C-1 (unless done at the end of column C-1) and the last term of C(i, C);
   then for each own row its accumulator of column C+1 from A(., C+1) (LDS);
   then for each tile column k < C the products with pivot row C+1's tile k.
   The owner of row C runs them between the steps of its diagonal sweep, the
   other wavefronts in order, followed by their share of the generation of
   column C+2. */
template <int NR, int C>
struct bv_res_jobs {
    static constexpr int NC = 4 * NR;
    static constexpr int q = C / 4, s = C % 4;
    static constexpr int ilo = (s == 3) ? q + 1 : q; /* first row that can be >= C+1 */
    static constexpr int nprep = NR - q - 1;
    static constexpr int ninit = (C + 1 < NC) ? 1 : 0;
    static constexpr int ngemm = (C + 1 < NC) ? C : 0;
    static constexpr int n = nprep + ninit + ngemm;
};

template <int KIND, int NR, int C, int J>
__device__ __forceinline__ void bv_res_job(bv_res_state<NR> &S, const bv_res_block &B,
                                           const bv_d4 &Wp) {
    using Jb = bv_res_jobs<NR, C>;
    constexpr int q = Jb::q;
    const int lane = B.lane, w = B.w;
    if constexpr (J < Jb::nprep) {
        constexpr int i = q + 1 + J;
        if constexpr (C > 0) {
            constexpr int cp = C - 1, qp = cp / 4, sp = cp % 4;
            constexpr int i1p = qp + (sp + 1) / 4, i2p = qp + (sp + 2) / 4;
            if constexpr (i != i1p && i != i2p) {
                S.L[i][cp] = bv_rjt(S.acc[i], Wp, (bv_d4){0.0, 0.0, 0.0, 0.0});
                S.dot += bv_res_msq(S.L[i][cp], w + 4 * i, cp, B);
            }
            const bv_d4 Bt = bv_lds_tile(bv_res_tile(B, C & 1, cp), lane);
            S.acc[i] = bv_rjt(S.L[i][cp], Bt, S.nxt[i]);
        } else {
            S.acc[i] = S.nxt[i];
        }
    } else if constexpr (J < Jb::nprep + Jb::ninit) {
#pragma unroll
        for (int i = Jb::ilo; i < NR; ++i)
            S.nxt[i] = bv_lds_tile(bv_res_tile(B, (C + 1) & 1, w + 4 * i), lane);
    } else if constexpr (J < Jb::n) {
        /* row q is past the pivot on the wavefronts w <= s: skipped there */
        constexpr int k = J - Jb::nprep - Jb::ninit;
        const bv_d4 Bk = bv_lds_tile(bv_res_tile(B, (C + 1) & 1, k), lane);
#pragma unroll
        for (int i = Jb::ilo; i < NR; ++i) {
            if (i == q && w <= Jb::s) continue;
            S.nxt[i] = bv_rjt(S.L[i][k], Bk, S.nxt[i]);
        }
    }
}

template <int KIND, int NR, int C, int J0, int J1>
__device__ __forceinline__ void bv_res_jobs_run(bv_res_state<NR> &S, const bv_res_block &B,
                                                const bv_d4 &Wp) {
    if constexpr (J0 < J1) {
        bv_res_job<KIND, NR, C, J0>(S, B, Wp);
        bv_res_jobs_run<KIND, NR, C, J0 + 1, J1>(S, B, Wp);
    }
}

/* The owner's diagonal sweep with the column's jobs spread over it: the
   first share of jobs before step 0 (matrix-core products that keep the
   pivot block's accumulation, just issued, well apart from the sweep's first
   read of it), then after each step K the share (K + 1) of 17; a scheduling
   barrier after each share keeps the jobs spread (the scheduler fills a
   step's latency gaps with them) and bounds what is live. */
template <int KIND, int NR, int C, int K>
__device__ __forceinline__ void bv_res_sweep_jobs(bv_res_state<NR> &S, const bv_res_block &B,
                                                  const bv_d4 &Wp, bv_d4 &a, bv_d4 &wt,
                                                  double &piv, double &dd, int &inf) {
    constexpr int n = bv_res_jobs<NR, C>::n;
    if constexpr (K == 0) {
        bv_res_jobs_run<KIND, NR, C, 0, n / 17>(S, B, Wp);
        __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (K < 16) {
        bv_res_sweep_step<K>(a, wt, B.lane, 16 * C, B.m, B.cs0, piv, dd, inf);
        bv_res_jobs_run<KIND, NR, C, ((K + 1) * n) / 17, ((K + 2) * n) / 17>(S, B, Wp);
        __builtin_amdgcn_sched_barrier(0);
        bv_res_sweep_jobs<KIND, NR, C, K + 1>(S, B, Wp, a, wt, piv, dd, inf);
    }
}

/* Covariance tiles A(r, c) for r = r0, r0 + step, ... < 4 NR into slot c % 2
   (rows at or beyond the block's padded size get the identity padding). */
template <int KIND>
__device__ __forceinline__ void bv_res_gen_col(const bv_res_block &B, const bv_kparams &kp, int c,
                                               int r0, int step) {
    for (int r = r0; r < B.NC; r += step)
        bv_lds_put(bv_res_tile(B, c & 1, r), B.lane, bv_res_gen<KIND>(B, r, c, kp));
}

/* Column C of the schedule in the header comment (C, its group q = C / 4 and
   slot s = C % 4 are compile-time constants, so every register index is). */
template <int KIND, int NR, int C>
__device__ __forceinline__ void bv_res_col(bv_res_state<NR> &S, const bv_res_block &B,
                                           const bv_kparams &kp) {
    constexpr int NC = 4 * NR; /* columns (and rows) the kernel is built for */
    if constexpr (C < NC) {
        if (C >= B.T) return;
        using Jb = bv_res_jobs<NR, C>;
        constexpr int q = C / 4, s = C % 4;
        const int lane = B.lane, w = B.w;
        const bv_d4 zero = (bv_d4){0.0, 0.0, 0.0, 0.0};

        /* row q first: it holds the pivot block on the owner (its solve of
           column C-1 was done at the end of column C-1) */
        const bv_d4 Wp = (C > 0) ? bv_lds_tile(B.sW + ((C - 1) & 1) * 256, lane) : zero;
        if constexpr (C > 0) {
            const bv_d4 Bt = bv_lds_tile(bv_res_tile(B, C & 1, C - 1), lane);
            S.acc[q] = bv_rjt(S.L[q][C - 1], Bt, S.nxt[q]);
        } else {
            S.acc[q] = S.nxt[q];
        }
        /* nothing moves across: the pivot block's accumulation ends here */
        __builtin_amdgcn_sched_barrier(0);

        if (w == s) {
            /* the owner: the diagonal sweep with the column's jobs between its
               steps; then its pivots, row m's squares, first failure and W_C */
            double piv = 1.0, dd = 0.0;
            int inf = 0;
            bv_d4 a = S.acc[q], wt;
            {
                const int lr = lane & 15, lq = lane >> 4;
#pragma unroll
                for (int g = 0; g < 4; ++g) wt[g] = (4 * g + lq == lr) ? 1.0 : 0.0;
            }
            bv_res_sweep_jobs<KIND, NR, C, 0>(S, B, Wp, a, wt, piv, dd, inf);
            const int qd = 16 * C + lane;
            const double pv = (lane < 16 && qd >= B.cs0 && qd < B.m) ? piv : 1.0;
            S.lmant *= __builtin_amdgcn_frexp_mant(pv);
            S.lexp += __builtin_amdgcn_frexp_exp(pv);
            S.lexp += __builtin_amdgcn_frexp_exp(S.lmant);
            S.lmant = __builtin_amdgcn_frexp_mant(S.lmant);
            S.dot += dd;
            if (S.info == 0) S.info = inf;
            bv_lds_put(B.sW + (C & 1) * 256, lane, wt);
        } else {
            bv_res_jobs_run<KIND, NR, C, 0, Jb::n>(S, B, Wp);
            /* column C+2's covariance tiles, split over the three wavefronts
               that do not factor this column's diagonal */
            if constexpr (C + 2 < NC) {
                const int rank = (w - s + 3) & 3; /* 0, 1, 2 */
                bv_res_gen_col<KIND>(B, kp, C + 2, C + 2 + rank, 3);
            }
        }
        __syncthreads();

        /* rows C+1 and C+2 of column C: their owners copy them to the
           buffers of the next two pivot rows (row C+2 with its older tiles) */
        if constexpr (C + 1 < NC) {
            constexpr int i1 = q + (s + 1) / 4, i2 = q + (s + 2) / 4;
            const bv_d4 W = bv_lds_tile(B.sW + (C & 1) * 256, lane);
            if constexpr (i1 < NR) {
                if (w + 4 * i1 > C) {
                    S.L[i1][C] = bv_rjt(S.acc[i1], W, zero);
                    S.dot += bv_res_msq(S.L[i1][C], w + 4 * i1, C, B);
                }
                if (w == (s + 1) % 4) bv_lds_put(bv_res_tile(B, (C + 1) & 1, C), lane, -S.L[i1][C]);
            }
            if constexpr (i2 != i1 && i2 < NR) {
                if (w + 4 * i2 > C) {
                    S.L[i2][C] = bv_rjt(S.acc[i2], W, zero);
                    S.dot += bv_res_msq(S.L[i2][C], w + 4 * i2, C, B);
                }
            }
            if constexpr (C + 2 < NC && i2 < NR) {
                if (w == (s + 2) % 4 && C + 2 < B.T) {
#pragma unroll
                    for (int k = 0; k <= C; ++k)
                        bv_lds_put(bv_res_tile(B, C & 1, k), lane, -S.L[i2][k]);
                }
            }
        }
        __syncthreads();
        bv_res_col<KIND, NR, C + 1>(S, B, kp);
    }
}

} // namespace

template <int KIND, int NR>
__global__ void __launch_bounds__(BV_RES_THREADS, (NR <= 3) ? 2 : 1)
bv_res_kernel(const bv_block_desc *__restrict__ desc, const bv_eval_slot *__restrict__ slots,
              const double *ws_base) {
    constexpr int NC = 4 * NR;
    constexpr int NP = 16 * NC; /* staged points */
    __shared__ __align__(16) double s_buf[2 * NC * 256];
    __shared__ __align__(16) double s_W[2 * 256];
    __shared__ double s_x[NP], s_y[NP], s_z[NP], s_o[NP];
    __shared__ double s_red[4][2];
    __shared__ double s_etab[2 * BV_EXP_TAB];
    __shared__ double s_cheb[KIND == BV_KIND_GENERAL ? BV_CHEB_SIZE : 1];
    __shared__ int s_redi[4];

    const bv_block_desc d = desc[blockIdx.x];
    const bv_eval_slot &E = slots[d.eval];
    const int tid = threadIdx.x;
    __shared__ bv_kparams s_kp; /* read where used, not held in registers */
    const bv_kparams &kp = s_kp;

    bv_res_block B;
    B.T = d.mp / BV_TILE;
    B.NC = NC;
    B.m = d.m;
    B.cs0 = d.cs0;
    B.w = __builtin_amdgcn_readfirstlane(tid >> 6);
    B.lane = tid & 63;
    B.sx = s_x; B.sy = s_y; B.sz = s_z; B.so = s_o;
    B.etab = s_etab;
    B.cheb = s_cheb;
    B.buf = s_buf;
    B.sW = s_W;
    B.ws = ws_base + d.ws_off;
    if (B.T > NC) {
        /* not a block for this kernel (the host never sends one) */
        if (tid == 0) {
            E.logdet[d.block] = __builtin_nan("");
            E.dot[d.block] = __builtin_nan("");
            E.info[d.block] = 1;
        }
        return;
    }
    if (tid < 2 * BV_EXP_TAB) s_etab[tid] = bv_exp_tab_entry(tid);
    if (KIND == BV_KIND_GENERAL)
        for (int i = tid; i < BV_CHEB_SIZE; i += BV_RES_THREADS) s_cheb[i] = E.kp.cheb[i];
    if (tid == 0) s_kp = E.kp;
    if (KIND != BV_KIND_LOADA) {
        for (int p = tid; p < NP; p += BV_RES_THREADS) {
            double xv, yv, zv, ov;
            bv_point(d, E, p, xv, yv, zv, ov);
            s_x[p] = xv; s_y[p] = yv; s_z[p] = zv; s_o[p] = ov;
        }
    }
    __syncthreads();
    /* columns 0 and 1 of A, all four wavefronts */
    bv_res_gen_col<KIND>(B, kp, 0, B.w, 4);
    bv_res_gen_col<KIND>(B, kp, 1, 1 + B.w, 4);
    __syncthreads();

    bv_res_state<NR> S;
    S.lmant = 1.0;
    S.lexp = 0;
    S.dot = 0.0;
    S.info = 0;
#pragma unroll
    for (int i = 0; i < NR; ++i) S.nxt[i] = bv_lds_tile(bv_res_tile(B, 0, B.w + 4 * i), B.lane);
    __syncthreads(); /* column 0 overwrites slot 0 with column 2 */
    bv_res_col<KIND, NR, 0>(S, B, kp);

    /* per-block results: each wavefront's sums in a fixed tree, then the four
       wavefronts in order */
    /* log of this lane's pivot product: log(mantissa) + exponent log 2 */
    double ld = log(S.lmant) + (double)S.lexp * 0.6931471805599453;
    double dot = S.dot;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        ld += __shfl_xor(ld, off, 64);
        dot += __shfl_xor(dot, off, 64);
    }
    if (B.lane == 0) {
        s_red[B.w][0] = ld;
        s_red[B.w][1] = dot;
        s_redi[B.w] = S.info;
    }
    __syncthreads();
    if (tid == 0) {
        double l = 0.0, t = 0.0;
        int info = 0;
        for (int k = 0; k < 4; ++k) {
            l += s_red[k][0];
            t += s_red[k][1];
            const int f = s_redi[k];
            if (f != 0 && (info == 0 || f < info)) info = f;
        }
        E.logdet[d.block] = l;
        E.dot[d.block] = t;
        E.info[d.block] = info;
    }
}

/* The kernels are compiled once per covariance kind: this file is built with
   -DBV_RES_KIND=k for each kind k (separate objects, compiled in parallel),
   each defining bv_launch_res_k<k>; bv_launch_res dispatches (bv_api.hip). */
#ifndef BV_RES_KIND
#error "build bv_resident.hip with -DBV_RES_KIND=<kind>"
#endif
template <int KIND>
hipError_t bv_launch_res_k(int nr, unsigned nblk, hipStream_t st, const bv_block_desc *desc,
                           const bv_eval_slot *slots, double *ws);

/* Launch over nblk descriptors whose blocks all have at most 4 NR tile rows
   (padded size <= 64 NR), NR = 2 .. 6. */
template <>
hipError_t bv_launch_res_k<BV_RES_KIND>(int nr, unsigned nblk, hipStream_t st,
                                        const bv_block_desc *desc, const bv_eval_slot *slots,
                                        double *ws) {
    if (nblk == 0) return hipSuccess;
    const dim3 grid(nblk), blk(BV_RES_THREADS);
    switch (nr) {
    case 2:
        hipLaunchKernelGGL((bv_res_kernel<BV_RES_KIND, 2>), grid, blk, 0, st, desc, slots, ws);
        break;
    case 3:
        hipLaunchKernelGGL((bv_res_kernel<BV_RES_KIND, 3>), grid, blk, 0, st, desc, slots, ws);
        break;
    case 4:
        hipLaunchKernelGGL((bv_res_kernel<BV_RES_KIND, 4>), grid, blk, 0, st, desc, slots, ws);
        break;
    case 5:
        hipLaunchKernelGGL((bv_res_kernel<BV_RES_KIND, 5>), grid, blk, 0, st, desc, slots, ws);
        break;
    case 6:
        hipLaunchKernelGGL((bv_res_kernel<BV_RES_KIND, 6>), grid, blk, 0, st, desc, slots, ws);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
