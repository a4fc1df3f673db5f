/*
 * bv_kernels.hip -- gfx950 kernels of the Block Vecchia likelihood.
 *
 * bv_block_kernel<KIND>: one workgroup (4 wavefronts) per block.  It factors
 * the block's augmented covariance (see bv_engine.h) with a left-looking
 * blocked Cholesky:
 *   for each 32-column panel j:
 *     0: A(rows >= 32j, panel) generated into the panel's workspace tiles
 *        (closed-form Matern kinds; general nu and space-time entries are
 *        written beforehand by bv_gen_kernel)
 *     1: C = A - L(rows, 0:32j) L(panel, 0:32j)^T (FP64 matrix-core 16x16x4
 *        products, accumulators in registers), stored over A
 *     2: diagonal 32x32 block of C -> LDS; one sweep computes its Cholesky
 *        factor L_jj and W = L_jj^-1 (forward elimination of I, one barrier
 *        per step)
 *     3: L(rows below, panel) = C W^T (matrix-core products), stored over C
 * The log-determinant over the block rows and the squared norm of the solved
 * observation row are reduced in-kernel (fixed order) to one pair per block.
 *
 * Storage of the factor: 16x16 tiles, packed lower triangle, tile-column
 * major.  Inside a tile, lane l of a wavefront owns the 4 consecutive doubles
 * at offset 4l: row (l % 16), columns (l / 16) + 4g for g = 0..3.  This is the
 * accumulator layout of v_mfma_f64_16x16x4f64 for the transposed product
 * (its D[4g + l/16][l % 16] sits in lane l, register g), so a panel is stored
 * with one 32-byte store per lane and read back as matrix-core operands with
 * one 32-byte load per lane: the 16-term contraction over a tile's columns is
 * split into 4 products, the g-th using column 4g + l/16 of every lane, which
 * is where both operands hold it.
 *
 * bv_reduce_kernel: llh = sum_i -1/2 (dot_i + logdet_i + bs_i log(2 pi)) in
 * block order (include/llh_Xvecchia_batch.h:480-494), plus the first block
 * whose factorization failed.
 */
#include <hip/hip_runtime.h>
#include "bv_engine.h"

typedef double bv_d4 __attribute__((ext_vector_type(4)));

#define BV_KIND_LOADA 7 /* factor kernel reads A from the workspace tiles */

__device__ __forceinline__ bv_d4 bv_mfma(double a, double b, bv_d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

/* tile (rt, ct), rt >= ct, of a factor with T tile rows; packed lower
   triangle, tile-column major; in doubles */
__device__ __forceinline__ int64_t bv_tile_off(int rt, int ct, int T) {
    return ((int64_t)(ct * T - (ct * (ct - 1)) / 2 + (rt - ct))) << 8;
}

/* Entry (p, q) of the augmented matrix of bv_engine.h. */
template <int KIND>
__device__ __forceinline__ double bv_aug_entry(int p, int q, int m, double xp, double yp,
                                               double zp, double xq, double yq, double zq,
                                               double oq, const bv_kparams &kp) {
    if (p < m && q < m) return bv_cov<KIND>(xp, yp, zp, xq, yq, zq, kp);
    if (p == m && q < m) return oq;
    return (p == q) ? 1.0 : 0.0;
}

/* Location / observation of augmented index p of block d. */
__device__ __forceinline__ void bv_point(const bv_block_desc &d, const bv_eval_slot &E, int p,
                                         double &xv, double &yv, double &zv, double &ov) {
    xv = yv = zv = ov = 0.0;
    if (p < d.cs0) {
        const int64_t o = (int64_t)d.block * E.cs + p;
        if (E.nbr) {
            const int64_t idx = E.nbr[o];
            xv = E.x[idx]; yv = E.y[idx]; ov = E.obs[idx];
            if (E.z) zv = E.z[idx];
        } else {
            xv = E.cx[o]; yv = E.cy[o]; ov = E.cobs[o];
            if (E.cz) zv = E.cz[o];
        }
    } else if (p < d.m) {
        const int64_t idx = d.obs_off + (p - d.cs0);
        xv = E.x[idx]; yv = E.y[idx]; ov = E.obs[idx];
        if (E.z) zv = E.z[idx];
    }
}

/* acc[s][ct] -= L(panel rows ct) L(row tile s)^T over the first 2j tile
   columns, for the C row tiles rt0 + wave + 4 (i0 + s) (acc holds A on entry;
   the panel operand is negated, so the products subtract).  C is a compile-time
   count so the loop is straight-line matrix-core code; the loads of tile
   column kt+1 are in flight while the products of tile column kt issue. */
template <int C>
__device__ __forceinline__ void bv_gemm_step(const bv_d4 &j0, const bv_d4 &j1, const bv_d4 (&r)[C],
                                             bv_d4 (&acc)[BV_NRT][2]) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
#pragma unroll
        for (int s = 0; s < C; ++s) {
            acc[s][0] = bv_mfma(j0[g], r[s][g], acc[s][0]);
            acc[s][1] = bv_mfma(j1[g], r[s][g], acc[s][1]);
        }
    }
}

template <int C>
__device__ __forceinline__ void bv_gemm_load(const double *colp, int rt0, int wave, int i0, int lo,
                                             bv_d4 &j0, bv_d4 &j1, bv_d4 (&r)[C]) {
    j0 = *(const bv_d4 *)(colp + ((int64_t)rt0 << 8) + lo);
    j1 = *(const bv_d4 *)(colp + ((int64_t)(rt0 + 1) << 8) + lo);
#pragma unroll
    for (int s = 0; s < C; ++s)
        r[s] = *(const bv_d4 *)(colp + ((int64_t)(rt0 + wave + 4 * (i0 + s)) << 8) + lo);
}

/* acc += L(panel rows ct) L(row tile s)^T over the first 2j tile columns, for
   the C row tiles rt0 + wave + 4 (i0 + s).  C is a compile-time count so the
   loop is straight-line matrix-core code.  Two operand buffers alternate: the
   loads of tile column kt+1 are in flight while the products of column kt
   issue, with no register copies between iterations.  Tile (rt, kt) sits at
   ws + (kt T - kt(kt-1)/2 + rt - kt) * 256: the column part is wave-uniform and
   advances by T - kt - 1 tiles per column, the lane part is 4 * lane. */
template <int C>
__device__ __forceinline__ void bv_panel_gemm(const double *__restrict__ ws, int T, int j,
                                              int wave, int lane, int i0,
                                              bv_d4 (&acc)[BV_NRT][2]) {
    const int rt0 = 2 * j;
    if (rt0 == 0) return;
    const int lo = 4 * lane;
    const double *colp = ws;
    bv_d4 ja0, ja1, ra[C], jb0, jb1, rb[C];
    bv_gemm_load<C>(colp, rt0, wave, i0, lo, ja0, ja1, ra);
    for (int kt = 0; kt < rt0; kt += 2) {
        /* column kt+1 (rt0 is even, so it always exists here) */
        const double *colb = colp + ((int64_t)(T - kt - 1) << 8);
        bv_gemm_load<C>(colb, rt0, wave, i0, lo, jb0, jb1, rb);
        bv_gemm_step<C>(ja0, ja1, ra, acc);
        /* column kt+2, or column kt+1 again after the last pair */
        const double *cola = (kt + 2 < rt0) ? colb + ((int64_t)(T - kt - 2) << 8) : colb;
        bv_gemm_load<C>(cola, rt0, wave, i0, lo, ja0, ja1, ra);
        bv_gemm_step<C>(jb0, jb1, rb, acc);
        colp = cola;
    }
}

/* Phase 0 of a panel: A(rt, panel) for this wave's row tiles rt0 + wave + 4i,
   generated from the staged locations and written to the tiles (rt, rt0),
   (rt, rt0+1) in the lane layout phase 1 reads them back in (each lane reads
   only what it wrote, so no barrier is needed between the two phases). */
template <int KIND>
__device__ __forceinline__ void bv_panel_generate(double *__restrict__ ws, int T, int j, int wave,
                                                  int lane, int m, const double *sx,
                                                  const double *sy, const double *sz,
                                                  const double *so, const bv_kparams &kp) {
    const int J0 = j * BV_NB;
    const int rt0 = 2 * j;
    const int lr = lane & 15, lq = lane >> 4;
    for (int rt = rt0 + wave; rt < T; rt += 4) {
        const int p = rt * BV_TILE + lr;
        const double xp = sx[p], yp = sy[p], zp = sz[p];
        for (int ct = 0; ct < 2; ++ct) {
            if (rt == rt0 && ct == 1) continue; /* upper tile of the diagonal block */
            bv_d4 v;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int q = J0 + ct * BV_TILE + 4 * g + lq;
                v[g] = bv_aug_entry<KIND>(p, q, m, xp, yp, zp, sx[q], sy[q], sz[q], so[q], kp);
            }
            *(bv_d4 *)(ws + bv_tile_off(rt, rt0 + ct, T) + 4 * lane) = v;
        }
    }
}

/* Phase 1 of a panel for C row tiles: acc = A - products, stored as C into
   the tiles (rt, rt0) and (rt, rt0+1) (the slots L will take). */
template <int C>
__device__ __forceinline__ void bv_pass_compute_c(double *__restrict__ ws, int T, int j, int wave,
                                                  int lane, int i0) {
    bv_d4 acc[BV_NRT][2];
    const int rt0 = 2 * j;
    /* acc = -A; the products are added; C = -acc (negations are exact) */
#pragma unroll
    for (int s = 0; s < C; ++s) {
        const int rt = rt0 + wave + 4 * (i0 + s);
        acc[s][0] = -*(const bv_d4 *)(ws + bv_tile_off(rt, rt0, T) + 4 * lane);
        acc[s][1] = (rt > rt0) ? -*(const bv_d4 *)(ws + bv_tile_off(rt, rt0 + 1, T) + 4 * lane)
                               : (bv_d4){0.0, 0.0, 0.0, 0.0};
    }
    bv_panel_gemm<C>(ws, T, j, wave, lane, i0, acc);
#pragma unroll
    for (int s = 0; s < C; ++s) {
        const int rt = rt0 + wave + 4 * (i0 + s);
        *(bv_d4 *)(ws + bv_tile_off(rt, rt0, T) + 4 * lane) = -acc[s][0];
        if (rt > rt0) *(bv_d4 *)(ws + bv_tile_off(rt, rt0 + 1, T) + 4 * lane) = -acc[s][1];
    }
}

__device__ __forceinline__ void bv_pass_compute(int c, double *__restrict__ ws, int T, int j,
                                                int wave, int lane, int i0) {
    switch (c) {
    case 1: bv_pass_compute_c<1>(ws, T, j, wave, lane, i0); break;
    case 2: bv_pass_compute_c<2>(ws, T, j, wave, lane, i0); break;
    case 3: bv_pass_compute_c<3>(ws, T, j, wave, lane, i0); break;
    case 4: bv_pass_compute_c<4>(ws, T, j, wave, lane, i0); break;
    default: break;
    }
}

/* Phase 3 for one row tile below the diagonal block: X = C W^T (matrix-core
   products with the three W tiles), stored over C; squares of the solved
   observation row m accumulated into dot_lane. */
__device__ __forceinline__ void bv_tile_solve(double *__restrict__ ws, int T, int j, int rt,
                                              int lane, int m, int cs0, const bv_d4 &w00,
                                              const bv_d4 &w10, const bv_d4 &w11,
                                              double &dot_lane) {
    const int J0 = j * BV_NB;
    const int rt0 = 2 * j;
    double *t0 = ws + bv_tile_off(rt, rt0, T) + 4 * lane;
    double *t1 = ws + bv_tile_off(rt, rt0 + 1, T) + 4 * lane;
    const bv_d4 c0 = *(const bv_d4 *)t0;
    const bv_d4 c1 = *(const bv_d4 *)t1;
    bv_d4 x0 = (bv_d4){0.0, 0.0, 0.0, 0.0};
    bv_d4 x1 = (bv_d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        x0 = bv_mfma(w00[g], c0[g], x0);
        x1 = bv_mfma(w10[g], c0[g], x1);
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) x1 = bv_mfma(w11[g], c1[g], x1);
    *(bv_d4 *)t0 = x0;
    *(bv_d4 *)t1 = x1;
    const int lr = lane & 15, lq = lane >> 4;
    if (rt * BV_TILE + lr == m) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int q0 = J0 + 4 * g + lq, q1 = q0 + BV_TILE;
            if (q0 >= cs0 && q0 < m) dot_lane += x0[g] * x0[g];
            if (q1 >= cs0 && q1 < m) dot_lane += x1[g] * x1[g];
        }
    }
}

/* Cholesky of the 32x32 diagonal block in sA (column-major, ld 33; lower part
   used) and, in the same sweep, W = L_jj^-1 by forward elimination of the
   identity held in sR (row-major residual rows, ld 33).  W is written to sW
   as three 16x16 tiles in accumulator layout: (0,0), (1,0), (1,1).  Step k,
   element (r, c) of the 32x32 grid:
     c > k, r >= c :  A[r][c] -= l_rk l_ck          (Cholesky trailing update)
     r == k, c <= k:  W[k][c]  = R[k][c] / l_kk
     r >  k, c <= k:  R[r][c] -= l_rk R[k][c] / l_kk
   with l_rk = A[r][k] / l_kk read before any write of step k.  One barrier per
   step.  Also accumulates the log-determinant over block rows, the first
   failing column, and the solved observations of row m if it lies here. */
__device__ __forceinline__ void bv_diag_factor(int tid, int J0, int m, int cs0, double *sA,
                                               double *sR, double *sW, double *s_logdet,
                                               double *s_dotd, int *s_info) {
    constexpr int LD = 33;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int e = tid + BV_THREADS * t;
        const int r = e & 31, cc = e >> 5;
        sR[r * LD + cc] = (r == cc) ? 1.0 : 0.0;
    }
    __syncthreads();
    const int km = m - J0; /* local index of the observation row, if in range */
    for (int k = 0; k < BV_NB; ++k) {
        const int qk = J0 + k;
        const double akk = sA[k * LD + k];
        const double piv = (qk >= m) ? 1.0 : sqrt(akk);
        const double pinv = 1.0 / piv;
        if (tid == 0) {
            if (qk < m) {
                if (!(akk > 0.0) && *s_info == 0) *s_info = qk + 1;
                if (qk >= cs0) *s_logdet += log(piv * piv);
            }
            if (km >= 0 && km < BV_NB && k < km && qk >= cs0) {
                const double v = sA[k * LD + km] * pinv;
                *s_dotd += v * v;
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int e = tid + BV_THREADS * t;
            const int r = e & 31, cc = e >> 5;
            if (cc > k) {
                if (r >= cc) sA[cc * LD + r] -= (sA[k * LD + r] * pinv) * (sA[k * LD + cc] * pinv);
            } else if (r == k) {
                const double w = sR[k * LD + cc] * pinv;
                const int tileid = (k < 16) ? 0 : (cc < 16 ? 1 : 2);
                const int rl = k & 15, cl = cc & 15;
                sW[tileid * 256 + 4 * (rl + 16 * (cl & 3)) + (cl >> 2)] = w;
            } else if (r > k) {
                sR[r * LD + cc] -= (sA[k * LD + r] * pinv) * (sR[k * LD + cc] * pinv);
            }
        }
        __syncthreads();
    }
    /* upper parts of the W tiles (0,0) and (1,1) are zero */
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int e = tid + BV_THREADS * t; /* 0..511: two 16x16 tiles */
        const int tile = e >> 8, rl = (e >> 4) & 15, cl = e & 15;
        if (cl > rl) sW[(tile ? 2 : 0) * 256 + 4 * (rl + 16 * (cl & 3)) + (cl >> 2)] = 0.0;
    }
    __syncthreads();
}

template <int KIND>
__global__ void __launch_bounds__(BV_THREADS, 2)
bv_block_kernel(const bv_block_desc *__restrict__ desc, const bv_eval_slot *__restrict__ slots,
                double *ws_base) {
    extern __shared__ double smem[];
    __shared__ double sA[BV_NB * 33];
    __shared__ double sR[BV_NB * 33];
    __shared__ double sW[3 * 256];
    __shared__ double s_logdet, s_dotd, s_wdot[BV_THREADS / 64];
    __shared__ int s_info;

    const bv_block_desc d = desc[blockIdx.x];
    const bv_eval_slot &E = slots[d.eval];
    const int tid0 = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid0 >> 6); /* wave-uniform */
    const int m = d.m, cs0 = d.cs0, Mp = d.mp;
    const int T = Mp / BV_TILE, P = Mp / BV_NB;
    double *sx = smem, *sy = sx + Mp, *sz = sy + Mp, *so = sz + Mp;
    const bv_kparams kp = E.kp;

    if (KIND != BV_KIND_LOADA) {
        /* stage the block's points: conditioning set first, then the block */
        for (int p = tid0; p < Mp; p += BV_THREADS) {
            double xv, yv, zv, ov;
            bv_point(d, E, p, xv, yv, zv, ov);
            sx[p] = xv; sy[p] = yv; sz[p] = zv; so[p] = ov;
        }
    }
    if (tid0 == 0) { s_logdet = 0.0; s_dotd = 0.0; s_info = 0; }
    __syncthreads();

    double *ws = ws_base + d.ws_off;
    double dot_lane = 0.0;

    for (int j = 0; j < P; ++j) {
        /* thread indices re-derived per panel: an empty asm hides that they are
           loop-invariant, so what depends on them is recomputed here instead
           of being kept in registers across the whole loop */
        int tid = tid0;
        asm volatile("" : "+v"(tid));
        const int lane = tid & 63;
        const int J0 = j * BV_NB;
        const int rt0 = 2 * j;
        const int nrt = T - rt0;
        /* phase 1: row tiles rt0 + wave + 4i, in passes of at most BV_NRT */
        const int cnt = (nrt > wave) ? (nrt - wave + 3) / 4 : 0;
        const int ngrp = (cnt + BV_NRT - 1) / BV_NRT;
        const int gsz = ngrp ? (cnt + ngrp - 1) / ngrp : 0;
        if (KIND != BV_KIND_LOADA)
            bv_panel_generate<KIND>(ws, T, j, wave, lane, m, sx, sy, sz, so, kp);
        for (int gi = 0; gi < ngrp; ++gi) {
            const int i0 = gi * gsz;
            const int c = (cnt - i0) < gsz ? (cnt - i0) : gsz;
            bv_pass_compute(c, ws, T, j, wave, lane, i0);
        }
        __syncthreads();
        /* phase 2: the diagonal block's C tiles -> sA, factor and invert */
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int e = tid + BV_THREADS * t;       /* 0..767: tiles (0,0),(1,0),(1,1) */
            const int tl = e >> 8, off = e & 255;
            const int l = off >> 2, g = off & 3;
            const int r = (l & 15) + (tl ? 16 : 0);
            const int cc = 4 * g + (l >> 4) + (tl == 2 ? 16 : 0);
            const int trt = rt0 + (tl ? 1 : 0), tct = rt0 + (tl == 2 ? 1 : 0);
            sA[cc * 33 + r] = ws[bv_tile_off(trt, tct, T) + off];
        }
        __syncthreads();
        bv_diag_factor(tid, J0, m, cs0, sA, sR, sW, &s_logdet, &s_dotd, &s_info);
        /* phase 3: rows below the diagonal block */
        {
            const bv_d4 w00 = *(const bv_d4 *)(sW + 0 * 256 + 4 * lane);
            const bv_d4 w10 = *(const bv_d4 *)(sW + 1 * 256 + 4 * lane);
            const bv_d4 w11 = *(const bv_d4 *)(sW + 2 * 256 + 4 * lane);
            for (int rt = rt0 + 2 + wave; rt < T; rt += 4)
                bv_tile_solve(ws, T, j, rt, lane, m, cs0, w00, w10, w11, dot_lane);
        }
        __syncthreads();
    }

    /* fixed-order reduction of the solved-observation squares */
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) dot_lane += __shfl_xor(dot_lane, off, 64);
    if ((tid0 & 63) == 0) s_wdot[wave] = dot_lane;
    __syncthreads();
    if (tid0 == 0) {
        double dsum = s_dotd;
        for (int w = 0; w < BV_THREADS / 64; ++w) dsum += s_wdot[w];
        E.logdet[d.block] = s_logdet;
        E.dot[d.block] = dsum;
        E.info[d.block] = s_info;
    }
}

/* Writes the augmented matrix A of every block (lower tiles, accumulator
   layout) into the workspace, for the kinds whose entries are too costly to
   recompute inside the factorization (general nu, space-time).  grid.x: tile
   index within the block, grid.y: descriptor. */
template <int KIND>
__global__ void __launch_bounds__(256)
bv_gen_kernel(const bv_block_desc *__restrict__ desc, const bv_eval_slot *__restrict__ slots,
              double *ws_base) {
    const bv_block_desc d = desc[blockIdx.y];
    const int T = d.mp / BV_TILE;
    const int ntile = T * (T + 1) / 2;
    int tix = blockIdx.x;
    if (tix >= ntile) return;
    int ct = 0;
    while (tix >= T - ct) { tix -= T - ct; ++ct; }
    const int rt = ct + tix;
    const bv_eval_slot &E = slots[d.eval];
    const int e = threadIdx.x, lane = e >> 2, g = e & 3;
    const int p = rt * BV_TILE + (lane & 15);
    const int q = ct * BV_TILE + 4 * g + (lane >> 4);
    double xp, yp, zp, op, xq, yq, zq, oq;
    bv_point(d, E, p, xp, yp, zp, op);
    bv_point(d, E, q, xq, yq, zq, oq);
    const double v = bv_aug_entry<KIND>(p, q, d.m, xp, yp, zp, xq, yq, zq, oq, E.kp);
    ws_base[d.ws_off + bv_tile_off(rt, ct, T) + e] = v;
}

/* One workgroup per evaluation: llh and first failing block, fixed order. */
extern "C" __global__ void __launch_bounds__(BV_THREADS)
bv_reduce_kernel(const bv_eval_slot *__restrict__ slots) {
    const bv_eval_slot &E = slots[blockIdx.x];
    __shared__ double s_part[BV_THREADS];
    __shared__ int s_bad[BV_THREADS];
    const int tid = threadIdx.x;
    const int B = E.nblocks;
    const int per = (B + BV_THREADS - 1) / BV_THREADS;
    const int lo = tid * per, hi = (lo + per < B) ? lo + per : B;
    double acc = 0.0;
    int bad = -1;
    for (int k = lo; k < hi; ++k) {
        acc += -(E.dot[k] + E.logdet[k] + (double)E.batch_num[k] * E.log2pi) * 0.5;
        if (bad < 0 && E.info[k] != 0) bad = k;
    }
    s_part[tid] = acc;
    s_bad[tid] = bad;
    __syncthreads();
    if (tid == 0) {
        double llh = 0.0;
        int first = -1;
        for (int t = 0; t < BV_THREADS; ++t) {
            llh += s_part[t];
            if (first < 0 && s_bad[t] >= 0) first = s_bad[t];
        }
        E.result[0] = llh;
        E.result[1] = (double)first;
    }
}

/* Dense covariance block A[i + j*m] = K(l1[i], l2[j]) (core_dcmg layout). */
extern "C" __global__ void __launch_bounds__(256)
bv_matern_kernel(int m, int n, const double *x1, const double *y1, const double *z1,
                 const double *x2, const double *y2, const double *z2, bv_kparams kp, int kind,
                 double *A) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)m * n) return;
    const int i = (int)(e % m), jj = (int)(e / m);
    const double za = z1 ? z1[i] : 0.0, zb = z2 ? z2[jj] : 0.0;
    double v;
    switch (kind) {
    case BV_KIND_NU05: v = bv_cov<BV_KIND_NU05>(x1[i], y1[i], za, x2[jj], y2[jj], zb, kp); break;
    case BV_KIND_NU15: v = bv_cov<BV_KIND_NU15>(x1[i], y1[i], za, x2[jj], y2[jj], zb, kp); break;
    case BV_KIND_NU25: v = bv_cov<BV_KIND_NU25>(x1[i], y1[i], za, x2[jj], y2[jj], zb, kp); break;
    case BV_KIND_SPACETIME:
        v = bv_cov<BV_KIND_SPACETIME>(x1[i], y1[i], za, x2[jj], y2[jj], zb, kp);
        break;
    default: v = bv_cov<BV_KIND_GENERAL>(x1[i], y1[i], za, x2[jj], y2[jj], zb, kp); break;
    }
    A[e] = v;
}

/* Matrix-core layout check: one 16x16x4 product of known operands.
   in: a[64], b[64] as lane values; out: d[256] in lane order (4 per lane). */
extern "C" __global__ void __launch_bounds__(64)
bv_selftest_kernel(const double *a, const double *b, double *d) {
    const int lane = threadIdx.x;
    bv_d4 acc = (bv_d4){0.0, 0.0, 0.0, 0.0};
    acc = bv_mfma(a[lane], b[lane], acc);
    *(bv_d4 *)(d + 4 * lane) = acc;
}

/* ---- host-side launch wrappers (called from bv_api.hip) ---- */
hipError_t bv_launch_block(int kind, unsigned nblk, size_t lds, hipStream_t st,
                           const bv_block_desc *desc, const bv_eval_slot *slots, double *ws) {
    switch (kind) {
    case BV_KIND_NU05:
        hipLaunchKernelGGL(bv_block_kernel<BV_KIND_NU05>, dim3(nblk), dim3(BV_THREADS), lds, st,
                           desc, slots, ws);
        break;
    case BV_KIND_NU15:
        hipLaunchKernelGGL(bv_block_kernel<BV_KIND_NU15>, dim3(nblk), dim3(BV_THREADS), lds, st,
                           desc, slots, ws);
        break;
    case BV_KIND_NU25:
        hipLaunchKernelGGL(bv_block_kernel<BV_KIND_NU25>, dim3(nblk), dim3(BV_THREADS), lds, st,
                           desc, slots, ws);
        break;
    default:
        hipLaunchKernelGGL(bv_block_kernel<BV_KIND_LOADA>, dim3(nblk), dim3(BV_THREADS), 0, st,
                           desc, slots, ws);
        break;
    }
    return hipGetLastError();
}

hipError_t bv_launch_gen(int kind, unsigned max_tiles, unsigned nblk, hipStream_t st,
                         const bv_block_desc *desc, const bv_eval_slot *slots, double *ws) {
    if (kind == BV_KIND_SPACETIME)
        hipLaunchKernelGGL(bv_gen_kernel<BV_KIND_SPACETIME>, dim3(max_tiles, nblk), dim3(256), 0,
                           st, desc, slots, ws);
    else
        hipLaunchKernelGGL(bv_gen_kernel<BV_KIND_GENERAL>, dim3(max_tiles, nblk), dim3(256), 0,
                           st, desc, slots, ws);
    return hipGetLastError();
}

hipError_t bv_launch_reduce(unsigned nevals, hipStream_t st, const bv_eval_slot *slots) {
    hipLaunchKernelGGL(bv_reduce_kernel, dim3(nevals), dim3(BV_THREADS), 0, st, slots);
    return hipGetLastError();
}

hipError_t bv_launch_matern(int m, int n, const double *x1, const double *y1, const double *z1,
                            const double *x2, const double *y2, const double *z2,
                            const bv_kparams &kp, int kind, double *A) {
    const int64_t tot = (int64_t)m * n;
    hipLaunchKernelGGL(bv_matern_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0,
                       m, n, x1, y1, z1, x2, y2, z2, kp, kind, A);
    return hipGetLastError();
}

hipError_t bv_launch_selftest(const double *a, const double *b, double *d) {
    hipLaunchKernelGGL(bv_selftest_kernel, dim3(1), dim3(64), 0, 0, a, b, d);
    return hipGetLastError();
}
