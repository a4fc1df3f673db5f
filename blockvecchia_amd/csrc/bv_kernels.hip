/*
 * bv_kernels.hip -- gfx950 kernels of the Block Vecchia likelihood.
 *
 * bv_block_kernel<KIND>: one workgroup (4 wavefronts) per block.  It factors
 * the block's augmented covariance (see bv_engine.h) with a left-looking
 * blocked Cholesky:
 *   for each 32-column panel j:
 *     0: A(rows >= 32j, panel) generated into the panel's workspace tiles
 *        (closed-form Matern kinds: for panel j+1 by wavefronts 1-3 while
 *        wavefront 0 runs phase 2 of panel j; general nu and space-time
 *        entries are written beforehand by bv_gen_kernel)
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
 * bv_reduce_kernel: llh = sum_i -1/2 (dot_i + logdet_i + bs_i log(2 pi))
 * (include/llh_Xvecchia_batch.h:480-494 sums it in block order; here in a
 * fixed pairwise order, equal to rounding), plus the first block whose
 * factorization failed.
 */
#include <hip/hip_runtime.h>
#include "bv_engine.h"
#include "bv_device.h"


/* Wavefront trace (builds with BV_TRACE only; see bv_trace_read in bv.h):
   each wavefront of a factorization kernel appends {start, end, id} when it
   leaves the kernel, start and end from the 100 MHz real-time counter, id =
   (kernel << 56) | ((step + 1) << 48) | (workgroup << 8) | wavefront.  The
   scope object's destructor runs on every return path. */
#ifdef BV_TRACE
#define BV_TRACE_CAP (1u << 22)
__device__ unsigned long long bv_trace_rec[3ull * BV_TRACE_CAP];
__device__ unsigned int bv_trace_n;
struct bv_trace_wave {
    unsigned long long t0, id;
    __device__ __forceinline__ bv_trace_wave(int kernel, int step) {
        t0 = __builtin_amdgcn_s_memrealtime();
        id = ((unsigned long long)(unsigned)kernel << 56) |
             ((unsigned long long)(unsigned)((step + 1) & 0xff) << 48) |
             ((unsigned long long)blockIdx.x << 8) | (threadIdx.x >> 6);
    }
    __device__ __forceinline__ ~bv_trace_wave() {
        if ((threadIdx.x & 63) == 0) {
            const unsigned long long t1 = __builtin_amdgcn_s_memrealtime();
            const unsigned int i = atomicAdd(&bv_trace_n, 1u);
            if (i < BV_TRACE_CAP) {
                bv_trace_rec[3ull * i] = t0;
                bv_trace_rec[3ull * i + 1] = t1;
                bv_trace_rec[3ull * i + 2] = id;
            }
        }
    }
};
#define BV_TRACE_WAVE(kernel, step) bv_trace_wave bv_trace_scope_((kernel), (step))
#else
#define BV_TRACE_WAVE(kernel, step) ((void)0)
#endif

/* acc[s][ct] -= L(panel rows ct) L(row tile s)^T over the first 2j tile
   columns, for the C row tiles rt0 + wave + 4 (i0 + s) (acc holds A on entry;
   the panel operand is negated, so the products subtract).  C is a compile-time
   count so the loop is straight-line matrix-core code; the loads of tile
   column kt+1 are in flight while the products of tile column kt issue. */
template <int C, int N>
__device__ __forceinline__ void bv_gemm_step(const bv_d4 &j0, const bv_d4 &j1, const bv_d4 (&r)[C],
                                             bv_d4 (&acc)[N][2]) {
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
__device__ __forceinline__ void bv_gemm_load(const double *colp, int rt0, int rbase, int rstride,
                                             int lo, bv_d4 &j0, bv_d4 &j1, bv_d4 (&r)[C]) {
    j0 = *(const bv_d4 *)(colp + ((int64_t)rt0 << 8) + lo);
    j1 = *(const bv_d4 *)(colp + ((int64_t)(rt0 + 1) << 8) + lo);
#pragma unroll
    for (int s = 0; s < C; ++s)
        r[s] = *(const bv_d4 *)(colp + ((int64_t)(rbase + rstride * s) << 8) + lo);
}

/* acc += L(panel rows ct) L(row tile s)^T over the first 2j tile columns, for
   the C row tiles rbase + rstride s.  C is a compile-time count so the
   loop is straight-line matrix-core code.  Two operand buffers alternate: the
   loads of tile column kt+1 are in flight while the products of column kt
   issue, with no register copies between iterations.  Tile (rt, kt) sits at
   ws + (kt T - kt(kt-1)/2 + rt - kt) * 256: the column part is wave-uniform and
   advances by T - kt - 1 tiles per column, the lane part is 4 * lane. */
template <int C, int N>
__device__ __forceinline__ void bv_panel_gemm(const double *__restrict__ ws, int T, int j,
                                              int rbase, int rstride, int lane,
                                              bv_d4 (&acc)[N][2]) {
    const int rt0 = 2 * j;
    if (rt0 == 0) return;
    const int lo = 4 * lane;
    const double *colp = ws;
    bv_d4 ja0, ja1, ra[C], jb0, jb1, rb[C];
    bv_gemm_load<C>(colp, rt0, rbase, rstride, lo, ja0, ja1, ra);
    for (int kt = 0; kt < rt0; kt += 2) {
        /* column kt+1 (rt0 is even, so it always exists here) */
        const double *colb = colp + ((int64_t)(T - kt - 1) << 8);
        bv_gemm_load<C>(colb, rt0, rbase, rstride, lo, jb0, jb1, rb);
        bv_gemm_step<C, N>(ja0, ja1, ra, acc);
        /* column kt+2, or column kt+1 again after the last pair */
        const double *cola = (kt + 2 < rt0) ? colb + ((int64_t)(T - kt - 2) << 8) : colb;
        bv_gemm_load<C>(cola, rt0, rbase, rstride, lo, ja0, ja1, ra);
        bv_gemm_step<C, N>(jb0, jb1, rb, acc);
        colp = cola;
    }
}

/* A(rt, panel j) for the row tiles rt = rtb, rtb + rstep, ..., generated from
   the staged locations and written to the tiles (rt, 2j), (rt, 2j+1) in
   accumulator layout (etab: the exponential's table, in LDS). */
template <int KIND>
__device__ __forceinline__ void bv_panel_generate(double *__restrict__ ws, int T, int j, int rtb,
                                                  int rstep, int lane, int m, const double *sx,
                                                  const double *sy, const double *sz,
                                                  const double *so, const bv_kparams &kp,
                                                  const double *etab) {
    const int J0 = j * BV_NB;
    const int rt0 = 2 * j;
    const int lr = lane & 15, lq = lane >> 4;
    /* planar Euclidean Matern entries: tiles away from the observation row
       and the padding take a short path (wave-uniform test) */
    const bool planar = bv_planar_ok<KIND>(kp);
    for (int rt = rtb; rt < T; rt += rstep) {
        const int p = rt * BV_TILE + lr;
        const double xp = sx[p], yp = sy[p], zp = sz[p];
        for (int ct = 0; ct < 2; ++ct) {
            if (rt == rt0 && ct == 1) continue; /* upper tile of the diagonal block */
            bv_d4 v;
            const int q0 = J0 + ct * BV_TILE;
            if (planar && rt * BV_TILE + 15 < m && q0 + 15 < m) {
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int q = q0 + 4 * g + lq;
                    v[g] = bv_matern_planar<KIND>(xp, yp, sx[q], sy[q], kp, kp.cheb, etab);
                }
            } else {
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int q = q0 + 4 * g + lq;
                    v[g] = bv_aug_entry<KIND>(p, q, m, xp, yp, zp, sx[q], sy[q], sz[q], so[q], kp);
                }
            }
            if (rt == rt0 + ct && kp.nugget != 0.0) {
#pragma unroll
                for (int g = 0; g < 4; ++g)
                    if (4 * g + lq == lr && p < m) v[g] += kp.nugget;
            }
            *(bv_d4 *)(ws + bv_tile_off(rt, rt0 + ct, T) + 4 * lane) = v;
            /* one tile row at a time: bounds the generation's live registers */
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

/* Phase 1 of a panel for C row tiles: acc = A - products, stored as C into
   the tiles (rt, rt0) and (rt, rt0+1) (the slots L will take). */
template <int C, int N>
__device__ __forceinline__ void bv_pass_compute_c(double *__restrict__ ws, int T, int j, int rbase,
                                                  int rstride, int lane) {
    bv_d4 acc[N][2];
    const int rt0 = 2 * j;
    /* acc = -A; the products are added; C = -acc (negations are exact) */
#pragma unroll
    for (int s = 0; s < C; ++s) {
        const int rt = rbase + rstride * s;
        acc[s][0] = -*(const bv_d4 *)(ws + bv_tile_off(rt, rt0, T) + 4 * lane);
        acc[s][1] = (rt > rt0) ? -*(const bv_d4 *)(ws + bv_tile_off(rt, rt0 + 1, T) + 4 * lane)
                               : (bv_d4){0.0, 0.0, 0.0, 0.0};
    }
    bv_panel_gemm<C, N>(ws, T, j, rbase, rstride, lane, acc);
#pragma unroll
    for (int s = 0; s < C; ++s) {
        const int rt = rbase + rstride * s;
        *(bv_d4 *)(ws + bv_tile_off(rt, rt0, T) + 4 * lane) = -acc[s][0];
        if (rt > rt0) *(bv_d4 *)(ws + bv_tile_off(rt, rt0 + 1, T) + 4 * lane) = -acc[s][1];
    }
}

/* c row tiles (c <= N, the most a pass holds) */
template <int N>
__device__ __forceinline__ void bv_pass_compute(int c, double *__restrict__ ws, int T, int j,
                                                int rbase, int rstride, int lane) {
    switch (c) {
    case 1: bv_pass_compute_c<1, N>(ws, T, j, rbase, rstride, lane); break;
    case 2: if (N >= 2) bv_pass_compute_c<(N >= 2 ? 2 : 1), N>(ws, T, j, rbase, rstride, lane); break;
    case 3: if (N >= 3) bv_pass_compute_c<(N >= 3 ? 3 : 1), N>(ws, T, j, rbase, rstride, lane); break;
    case 4: if (N >= 4) bv_pass_compute_c<(N >= 4 ? 4 : 1), N>(ws, T, j, rbase, rstride, lane); break;
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

/* Transpose of a 16x16 tile in accumulator layout, through an LDS scratch of
   16 x 17 doubles (one wavefront). */
__device__ __forceinline__ bv_d4 bv_transpose16(const bv_d4 &v, int lane, double *sT) {
    const int lr = lane & 15, lq = lane >> 4;
    bv_wave_sync(); /* earlier reads of sT are done */
#pragma unroll
    for (int g = 0; g < 4; ++g) sT[lr * 17 + 4 * g + lq] = v[g];
    bv_wave_sync();
    bv_d4 t;
#pragma unroll
    for (int g = 0; g < 4; ++g) t[g] = sT[(4 * g + lq) * 17 + lr];
    return t;
}

/* Unblocked Cholesky of a 16x16 tile held in accumulator layout (lane l holds
   row l%16, columns 4g + l/16), and in the same sweep W = L^-1 by forward
   elimination of the identity.  Step k:
     l_kk = sqrt(a_kk)                     (readlane from lane k + 16 (k%4))
     l_rk = a_rk / l_kk                    (from the lane of row r holding column k)
     a_rc -= l_rk l_ck                     (column k broadcast through sCol)
     W_k = R_k / l_kk; R_r -= l_rk W_k     (r > k; row k broadcast through sRow)
   Entries with c <= k or r <= k are updated too but never read again; W stays
   lower triangular.  The lanes exchange values through LDS with a wavefront
   release/acquire (bv_wave_sync) between the writes and the reads.  kbase:
   global index of local column 0.  The squared pivots a_kk = l_kk^2 go to
   sPiv[pivoff + k]; squares of row m's factor entries in columns >= cs0 add to
   dacc; the first failing column (global, + 1) to info. */
#ifndef BV_SWEEP_DPP
#define BV_SWEEP_DPP 1
#endif

/* Step K of the sweep below without LDS: the column-K factors of the other
   rows that the update needs are taken from row K (the trailing block is
   exactly symmetric: A and every product subtracted from it are, entry for
   entry; the observation row's column, where the augmented matrix is not
   symmetric, only meets pivots of 1 and zero factors), broadcast within each
   16-lane row by DPP, as is row K of W.  Only the row's own column-K factor
   crosses rows (one lane permutation). */
template <int K>
__device__ __forceinline__ void bv_sweep_step(bv_d4 &a, bv_d4 &w, int lane, int kbase, int m,
                                              int cs0, double *sPiv, int pivoff, double &dacc,
                                              int &info) {
    constexpr int kq = K & 3, kg = K >> 2;
    const int lr = lane & 15, lq = lane >> 4;
    const int kml = m - kbase;
    const int q = kbase + K;
    const double akk = bv_readlane(a[kg], K + 16 * kq);
    double pinv = 1.0;
    if (q < m) {
        const double ha = 0.5 * akk;
        double y = __builtin_amdgcn_rsq(akk);
        y = fma(y, fma(-ha * y, y, 0.5), y);
        y = fma(y, fma(-ha * y, y, 0.5), y);
        pinv = y;
    }
    if (q < m && !(akk > 0.0) && info == 0) info = q + 1;
    const double t = a[kg] * pinv;
    const double lrk = __shfl(t, lr + 16 * kq, 64);
    if (lr == kml && K < kml && q >= cs0 && lq == 0) dacc += lrk * lrk;
    if (lane == 0) sPiv[pivoff + K] = akk; /* l_kk^2: log-determinant term */
    /* W: row K scaled by 1/l_KK (for its own lanes the broadcast is the lane's
       own value, so w * pinv there is exactly wk), the rows below it minus
       l_rK times the scaled row */
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
__device__ __forceinline__ void bv_sweep_steps(bv_d4 &a, bv_d4 &w, int lane, int kbase, int m,
                                               int cs0, double *sPiv, int pivoff, double &dacc,
                                               int &info) {
    bv_sweep_step<K>(a, w, lane, kbase, m, cs0, sPiv, pivoff, dacc, info);
    if constexpr (K + 1 < 16)
        bv_sweep_steps<K + 1>(a, w, lane, kbase, m, cs0, sPiv, pivoff, dacc, info);
}

__device__ __forceinline__ bv_d4 bv_sweep16(bv_d4 a, int lane, int kbase, int m, int cs0,
                                            double *sCol, double *sRow, double *sPiv,
                                            int pivoff, double &dacc, int &info) {
#if BV_SWEEP_DPP
    (void)sCol;
    (void)sRow;
    const int lr0 = lane & 15, lq0 = lane >> 4;
    bv_d4 w0;
#pragma unroll
    for (int g = 0; g < 4; ++g) w0[g] = (4 * g + lq0 == lr0) ? 1.0 : 0.0;
    bv_sweep_steps<0>(a, w0, lane, kbase, m, cs0, sPiv, pivoff, dacc, info);
    return w0;
#else
    /* the same sweep with the column factors and W's row K through LDS */
    const int lr = lane & 15, lq = lane >> 4;
    const int kml = m - kbase;
    bv_d4 w;
#pragma unroll
    for (int g = 0; g < 4; ++g) w[g] = (4 * g + lq == lr) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int kq = k & 3, kg = k >> 2;
        const int q = kbase + k;
        const double akk = bv_readlane(a[kg], k + 16 * kq);
        /* 1/l_kk = 1/sqrt(a_kk): hardware reciprocal square root refined by two
           Newton steps y <- y + y (1/2 - (a/2) y^2) (each doubles the correct
           bits; the hardware estimate carries more than 2^-27) */
        double pinv = 1.0;
        if (q < m) {
            const double ha = 0.5 * akk;
            double y = __builtin_amdgcn_rsq(akk);
            y = fma(y, fma(-ha * y, y, 0.5), y);
            y = fma(y, fma(-ha * y, y, 0.5), y);
            pinv = y;
        }
        if (q < m && !(akk > 0.0) && info == 0) info = q + 1;
        const double t = a[kg] * pinv;
        const double lrk = __shfl(t, lr + 16 * kq, 64);
        if (lr == kml && k < kml && q >= cs0 && lq == 0) dacc += lrk * lrk;
        if (lr == k) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                w[g] *= pinv;
                sRow[4 * g + lq] = w[g];
            }
        }
        if (lq == kq) sCol[lr] = lrk;
        if (lane == 0) sPiv[pivoff + k] = akk; /* l_kk^2: log-determinant term */
        bv_wave_sync();
        const double lre = (lr > k) ? lrk : 0.0;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            a[g] -= lrk * sCol[4 * g + lq];
            w[g] -= lre * sRow[4 * g + lq];
        }
        bv_wave_sync(); /* reads done before the next step's writes */
    }
    return w;
#endif
}

/* Log-determinant terms of one panel: log(l_kk^2) for the block rows
   (cs0 <= J0 + k < m), one per lane 0..31, summed in a fixed tree order and
   added to *s_logdet by lane 0.  One wavefront. */
__device__ __forceinline__ void bv_panel_logdet(int lane, int J0, int m, int cs0,
                                                const double *sPiv, double *s_logdet) {
    bv_wave_sync();
    double t = 0.0;
    if (lane < BV_NB) {
        const int q = J0 + lane;
        if (q >= cs0 && q < m) t = log(sPiv[lane]);
    }
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane == 0) *s_logdet += t;
}

/* Phase 2, one wavefront: factor the 32x32 diagonal block whose C tiles are
   (rt0,rt0), (rt0+1,rt0), (rt0+1,rt0+1) of the workspace, as 2x2 blocks of
   16x16:  L11, W11 = sweep(A11);  L21 = A21 W11^T;  A22' = A22 - L21 L21^T;
   L22, W22 = sweep(A22');  W21 = -W22 (L21 W11).  W = L_jj^-1 goes to sW as
   the tiles (0,0), (1,0), (1,1) in accumulator layout; pivots to sPiv. */
__device__ __forceinline__ void bv_diag_factor_tiles(const double *t11, const double *t21,
                                                     const double *t22, int lane, int J0, int m,
                                                     int cs0, double *sCol, double *sRow,
                                                     double *sT, double *sW, double *sPiv,
                                                     double *s_dotd, int *s_info) {
    const int lr = lane & 15, lq = lane >> 4;
    const bv_d4 a11 = *(const bv_d4 *)(t11 + 4 * lane);
    const bv_d4 a21 = *(const bv_d4 *)(t21 + 4 * lane);
    const bv_d4 a22 = *(const bv_d4 *)(t22 + 4 * lane);
    double dacc = 0.0;
    int info = 0;
    const bv_d4 zero = (bv_d4){0.0, 0.0, 0.0, 0.0};
    const bv_d4 w11 = bv_sweep16(a11, lane, J0, m, cs0, sCol, sRow, sPiv, 0, dacc, info);
    const bv_d4 l21 = bv_rjt(a21, w11, zero);
    const bv_d4 a22p = -bv_rjt(l21, l21, -a22);
    const int km = m - J0;
    if (km >= 16 && km < 32 && lr == km - 16) {
#pragma unroll
        for (int g = 0; g < 4; ++g)
            if (J0 + 4 * g + lq >= cs0) dacc += l21[g] * l21[g];
    }
    const bv_d4 w22 = bv_sweep16(a22p, lane, J0 + 16, m, cs0, sCol, sRow, sPiv, 16, dacc, info);
    const bv_d4 w11t = bv_transpose16(w11, lane, sT);
    const bv_d4 pm = bv_rjt(l21, w11t, zero);          /* L21 W11 */
    const bv_d4 pmt = bv_transpose16(pm, lane, sT);
    const bv_d4 w21 = -bv_rjt(w22, pmt, zero);         /* -W22 L21 W11 */
    *(bv_d4 *)(sW + 0 * 256 + 4 * lane) = w11;
    *(bv_d4 *)(sW + 1 * 256 + 4 * lane) = w21;
    *(bv_d4 *)(sW + 2 * 256 + 4 * lane) = w22;
    /* fixed-order reduction of row m's squares over the wavefront */
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) dacc += __shfl_xor(dacc, off, 64);
    if (lane == 0) {
        *s_dotd += dacc;
        if (info != 0 && *s_info == 0) *s_info = info;
    }
}

/* The same for the tiles (rt0,rt0), (rt0+1,rt0), (rt0+1,rt0+1) of a workspace
   with T tile rows. */
__device__ __forceinline__ void bv_diag_factor_wave(const double *__restrict__ ws, int T, int rt0,
                                                    int lane, int J0, int m, int cs0,
                                                    double *sCol, double *sRow, double *sT,
                                                    double *sW, double *sPiv, double *s_dotd,
                                                    int *s_info) {
    bv_diag_factor_tiles(ws + bv_tile_off(rt0, rt0, T), ws + bv_tile_off(rt0 + 1, rt0, T),
                         ws + bv_tile_off(rt0 + 1, rt0 + 1, T), lane, J0, m, cs0, sCol, sRow,
                         sT, sW, sPiv, s_dotd, s_info);
}

/* NW wavefronts per block: 4 (BV_WAVES_PER_SIMD per SIMD) or 2
   (BV_2W_WAVES_PER_SIMD: fewer registers, so more blocks resident at once). */
#ifndef BV_2W_WAVES_PER_SIMD
#define BV_2W_WAVES_PER_SIMD 3
#endif
#ifndef BV_2W_NRT
#define BV_2W_NRT 2
#endif
template <int KIND, int NW>
__global__ void __launch_bounds__(64 * NW, NW == 4 ? BV_WAVES_PER_SIMD : BV_2W_WAVES_PER_SIMD)
bv_block_kernel(const bv_block_desc *__restrict__ desc, const bv_eval_slot *__restrict__ slots,
                double *ws_base) {
    extern __shared__ double smem[];
    __shared__ double sW[3 * 256];
    __shared__ double sT[16 * 17];
    __shared__ double sCol[16], sRow[16], sPiv[BV_NB];
    __shared__ double s_logdet, s_dotd;
    __shared__ int s_info;
    __shared__ bv_kparams s_kp; /* read where used: not held in registers */
    __shared__ double s_etab[2 * BV_EXP_TAB]; /* bv_exp_tab_dev, copied */

    const bv_block_desc d = desc[blockIdx.x];
    BV_TRACE_WAVE(0, -1);
    const bv_eval_slot &E = slots[d.eval];
    const int tid0 = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid0 >> 6); /* wave-uniform */
    const int m = d.m, cs0 = d.cs0, Mp = d.mp;
    const int T = Mp / BV_TILE, P = Mp / BV_NB;
    double *sx = smem, *sy = sx + Mp, *sz = sy + Mp, *so = sz + Mp;
    const bv_kparams &kp = s_kp;
    constexpr int NRT = (NW == 4) ? BV_NRT : BV_2W_NRT; /* row tiles per GEMM pass */

    if (KIND != BV_KIND_LOADA) {
        /* stage the block's points: conditioning set first, then the block */
        for (int p = tid0; p < Mp; p += 64 * NW) {
            double xv, yv, zv, ov;
            bv_point(d, E, p, xv, yv, zv, ov);
            sx[p] = xv; sy[p] = yv; sz[p] = zv; so[p] = ov;
        }
    }
    if (tid0 == 0) { s_logdet = 0.0; s_dotd = 0.0; s_info = 0; s_kp = E.kp; }
    for (int i = tid0; i < 2 * BV_EXP_TAB; i += 64 * NW) s_etab[i] = bv_exp_tab_dev[i];
    __syncthreads();

    double *ws = ws_base + d.ws_off;

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
        /* phase 1: row tiles rt0 + wave + NW i, in passes of at most NRT */
        const int cnt = (nrt > wave) ? (nrt - wave + NW - 1) / NW : 0;
        const int ngrp = (cnt + NRT - 1) / NRT;
        const int gsz = ngrp ? (cnt + ngrp - 1) / ngrp : 0;
        /* phase 0 for the first panel (later panels are generated during the
           previous panel's phase 2) */
        if (KIND != BV_KIND_LOADA && j == 0) {
            bv_panel_generate<KIND>(ws, T, 0, wave, NW, lane, m, sx, sy, sz, so, kp, s_etab);
            __syncthreads();
        }
        for (int gi = 0; gi < ngrp; ++gi) {
            const int i0 = gi * gsz;
            const int c = (cnt - i0) < gsz ? (cnt - i0) : gsz;
            bv_pass_compute<NRT>(c, ws, T, j, rt0 + wave + NW * i0, NW, lane);
        }
        __syncthreads();
        /* phase 2 (wavefront 0): factor and invert the diagonal block, while
           the other wavefronts generate the next panel's covariance tiles */
        if (KIND != BV_KIND_LOADA && wave > 0 && j + 1 < P)
            bv_panel_generate<KIND>(ws, T, j + 1, rt0 + 2 + wave - 1, NW - 1, lane, m, sx, sy, sz,
                                    so, kp, s_etab);
        if (wave == 0) {
            bv_diag_factor_wave(ws, T, rt0, lane, J0, m, cs0, sCol, sRow, sT, sW, sPiv, &s_dotd,
                                &s_info);
            bv_panel_logdet(lane, J0, m, cs0, sPiv, &s_logdet);
        }
        __syncthreads();
        /* phase 3: rows below the diagonal block */
        {
            const bv_d4 w00 = *(const bv_d4 *)(sW + 0 * 256 + 4 * lane);
            const bv_d4 w10 = *(const bv_d4 *)(sW + 1 * 256 + 4 * lane);
            const bv_d4 w11 = *(const bv_d4 *)(sW + 2 * 256 + 4 * lane);
            double dot_lane = 0.0;
            for (int rt = rt0 + 2 + wave; rt < T; rt += NW)
                bv_tile_solve(ws, T, j, rt, lane, m, cs0, w00, w10, w11, dot_lane);
            if ((E.flags & BV_SLOT_SAMPLE) && wave < 2) {
                /* the diagonal block's rows too: L_jj = C_jj W^T, so the
                   workspace ends with the complete factor (their squares of
                   row m were counted by the diagonal factorization) */
                double unused = 0.0;
                if (wave == 1) {
                    bv_tile_solve(ws, T, j, rt0 + 1, lane, m, cs0, w00, w10, w11, unused);
                } else {
                    double *t0 = ws + bv_tile_off(rt0, rt0, T) + 4 * lane;
                    const bv_d4 c0 = *(const bv_d4 *)t0;
                    bv_d4 x0 = (bv_d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int g = 0; g < 4; ++g) x0 = bv_mfma(w00[g], c0[g], x0);
                    *(bv_d4 *)t0 = x0;
                }
            }
            /* the panel's row-m squares go to s_dotd in panel order, reduced by
               the one wavefront that solved row m's tile: the same sequence of
               additions for any number of wavefronts per block */
            const int rtm = m / BV_TILE;
            if (rtm >= rt0 + 2 && (rtm - rt0 - 2) % NW == wave) {
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) dot_lane += __shfl_xor(dot_lane, off, 64);
                if (lane == 0) s_dotd += dot_lane;
            }
        }
        __syncthreads();
    }

    if (tid0 == 0) {
        E.logdet[d.block] = s_logdet;
        E.dot[d.block] = s_dotd;
        E.info[d.block] = s_info;
    }
}

/* bv_block_kernel_1w: one wavefront per block.  The same four phases as
   bv_block_kernel, all run by the one wavefront: in the tile layout every lane
   reads back only the workspace values it wrote itself (the cross-lane mixing
   happens inside the matrix-core products), so no barrier is needed; the
   diagonal factorization exchanges values through LDS with wavefront syncs.
   Many blocks per SIMD run concurrently, so one block's serial diagonal
   factorization overlaps other blocks' matrix-core work. */
#ifndef BV_1W_WAVES_PER_SIMD
#define BV_1W_WAVES_PER_SIMD 2
#endif
template <int KIND>
__global__ void __launch_bounds__(64, BV_1W_WAVES_PER_SIMD)
bv_block_kernel_1w(const bv_block_desc *__restrict__ desc, const bv_eval_slot *__restrict__ slots,
                   double *ws_base) {
    extern __shared__ double smem[];
    __shared__ double sW[3 * 256];
    __shared__ double sT[16 * 17];
    __shared__ double sCol[16], sRow[16], sPiv[BV_NB];
    __shared__ double s_logdet, s_dotd;
    __shared__ int s_info;
    __shared__ bv_kparams s_kp;
    __shared__ double s_etab[2 * BV_EXP_TAB]; /* bv_exp_tab_dev, copied */

    const bv_block_desc d = desc[blockIdx.x];
    BV_TRACE_WAVE(0, -1);
    const bv_eval_slot &E = slots[d.eval];
    const int lane0 = threadIdx.x;
    const int m = d.m, cs0 = d.cs0, Mp = d.mp;
    const int T = Mp / BV_TILE, P = Mp / BV_NB;
    double *sx = smem, *sy = sx + Mp, *sz = sy + Mp, *so = sz + Mp;
    const bv_kparams &kp = s_kp;

    if (KIND != BV_KIND_LOADA) {
        for (int p = lane0; p < Mp; p += 64) {
            double xv, yv, zv, ov;
            bv_point(d, E, p, xv, yv, zv, ov);
            sx[p] = xv; sy[p] = yv; sz[p] = zv; so[p] = ov;
        }
    }
    if (lane0 == 0) { s_logdet = 0.0; s_dotd = 0.0; s_info = 0; s_kp = E.kp; }
    s_etab[lane0] = bv_exp_tab_dev[lane0];
    s_etab[lane0 + 64] = bv_exp_tab_dev[lane0 + 64];
    bv_wave_sync();

    double *ws = ws_base + d.ws_off;
    for (int j = 0; j < P; ++j) {
        int lane = lane0;
        asm volatile("" : "+v"(lane)); /* recompute lane-derived values per panel */
        const int J0 = j * BV_NB;
        const int rt0 = 2 * j;
        const int nrt = T - rt0;
        if (KIND != BV_KIND_LOADA)
            bv_panel_generate<KIND>(ws, T, j, rt0, 1, lane, m, sx, sy, sz, so, kp, s_etab);
        const int ngrp = (nrt + BV_NRT - 1) / BV_NRT;
        const int gsz = (nrt + ngrp - 1) / ngrp;
        for (int gi = 0; gi < ngrp; ++gi) {
            const int i0 = gi * gsz;
            const int c = (nrt - i0) < gsz ? (nrt - i0) : gsz;
            bv_pass_compute<BV_NRT>(c, ws, T, j, rt0 + i0, 1, lane);
        }
        bv_diag_factor_wave(ws, T, rt0, lane, J0, m, cs0, sCol, sRow, sT, sW, sPiv, &s_dotd,
                            &s_info);
        bv_panel_logdet(lane, J0, m, cs0, sPiv, &s_logdet);
        const bv_d4 w00 = *(const bv_d4 *)(sW + 0 * 256 + 4 * lane);
        const bv_d4 w10 = *(const bv_d4 *)(sW + 1 * 256 + 4 * lane);
        const bv_d4 w11 = *(const bv_d4 *)(sW + 2 * 256 + 4 * lane);
        double dot_lane = 0.0;
        for (int rt = rt0 + 2; rt < T; ++rt)
            bv_tile_solve(ws, T, j, rt, lane, m, cs0, w00, w10, w11, dot_lane);
        /* row m's squares of this panel, in panel order (as bv_block_kernel) */
        if (m / BV_TILE >= rt0 + 2) {
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) dot_lane += __shfl_xor(dot_lane, off, 64);
            if (lane == 0) s_dotd += dot_lane;
        }
        bv_wave_sync();
    }
    if (lane0 == 0) {
        E.logdet[d.block] = s_logdet;
        E.dot[d.block] = s_dotd;
        E.info[d.block] = s_info;
    }
}

/* ======================================================================
 * Stepped factorization: one launch per panel step j over every block.
 *
 * A work item is one wavefront on BV_STEP_ROWS consecutive row tiles R of
 * one block (all rows >= 2j).  At step j it
 *   (a) solves panel j-1 for its rows, L(R, j-1) = C(R, j-1) W_{j-1}^T, with
 *       C and W_{j-1} from step j-1, stores L (except rows 2j, 2j+1, whose C
 *       the other items read in this step and whose L no later step uses)
 *       and adds row m's squares;
 *   (b) forms C(R, j) = A(R, j) - sum_k L(R, k) L(2j..2j+1, k)^T: A generated
 *       in registers (closed forms) or read (LOADA); the last tile-column
 *       pair k = 2j-2, 2j-1 from registers -- its own rows from (a), the
 *       panel rows recomputed from C(2j..2j+1, j-1) and W_{j-1} -- and the
 *       older columns from the workspace;
 *   (c) stores C(R, j), or, for the item holding rows 2j and 2j+1, factors
 *       the diagonal block (LDS) and stores W_j over its three tiles, with
 *       the panel's log-determinant, row-m squares and first failure into
 *       the per-(block, panel) partial arrays.
 * Items never wait for each other inside a launch: everything they read was
 * written by earlier launches.  No barriers; one wavefront per workgroup.
 * ====================================================================== */
#ifndef BV_STEP_WAVES_PER_SIMD
#define BV_STEP_WAVES_PER_SIMD 3
#endif
#ifndef BV_LOOK_PRIO
#define BV_LOOK_PRIO 1 /* lookahead items at raised wave priority (0: off) */
#endif

/* acc += L(rows, k) L(panel rows, k)^T over tile columns [0, kend) from the
   workspace, for C consecutive row tiles rbase + s (ping-pong operands). */
template <int C, int N>
__device__ __forceinline__ void bv_step_gemm(const double *__restrict__ ws, int T, int j, int kend,
                                             int rbase, int lane, bv_d4 (&acc)[N][2]) {
    const int rt0 = 2 * j;
    if (kend <= 0) return;
    const int lo = 4 * lane;
    const double *colp = ws;
    bv_d4 ja0, ja1, ra[C], jb0, jb1, rb[C];
    bv_gemm_load<C>(colp, rt0, rbase, 1, lo, ja0, ja1, ra);
    for (int kt = 0; kt < kend; kt += 2) {
        const double *colb = colp + ((int64_t)(T - kt - 1) << 8);
        bv_gemm_load<C>(colb, rt0, rbase, 1, lo, jb0, jb1, rb);
        bv_gemm_step<C, N>(ja0, ja1, ra, acc);
        const double *cola = (kt + 2 < kend) ? colb + ((int64_t)(T - kt - 2) << 8) : colb;
        bv_gemm_load<C>(cola, rt0, rbase, 1, lo, ja0, ja1, ra);
        bv_gemm_step<C, N>(jb0, jb1, rb, acc);
        colp = cola;
    }
}

/* bv_step_gemm for two row tiles rbase, rbase + 1 that also accumulate
   their own lower 2x2 tile block (the SYRK of the next panel's diagonal
   block: d00 += r0 r0^T, d10 += r1 r0^T, d11 += r1 r1^T) from the same
   loaded tiles. */
template <int N>
__device__ __forceinline__ void bv_step_gemm_syrk(const double *__restrict__ ws, int T, int j,
                                                  int kend, int rbase, int lane,
                                                  bv_d4 (&acc)[N][2], bv_d4 &d00, bv_d4 &d10,
                                                  bv_d4 &d11) {
    const int rt0 = 2 * j;
    if (kend <= 0) return;
    const int lo = 4 * lane;
    const double *colp = ws;
    bv_d4 ja0, ja1, ra[2], jb0, jb1, rb[2];
    bv_gemm_load<2>(colp, rt0, rbase, 1, lo, ja0, ja1, ra);
    for (int kt = 0; kt < kend; kt += 2) {
        const double *colb = colp + ((int64_t)(T - kt - 1) << 8);
        bv_gemm_load<2>(colb, rt0, rbase, 1, lo, jb0, jb1, rb);
        bv_gemm_step<2, N>(ja0, ja1, ra, acc);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            d00 = bv_mfma(ra[0][g], ra[0][g], d00);
            d10 = bv_mfma(ra[0][g], ra[1][g], d10);
            d11 = bv_mfma(ra[1][g], ra[1][g], d11);
        }
        const double *cola = (kt + 2 < kend) ? colb + ((int64_t)(T - kt - 2) << 8) : colb;
        bv_gemm_load<2>(cola, rt0, rbase, 1, lo, ja0, ja1, ra);
        bv_gemm_step<2, N>(jb0, jb1, rb, acc);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            d00 = bv_mfma(rb[0][g], rb[0][g], d00);
            d10 = bv_mfma(rb[0][g], rb[1][g], d10);
            d11 = bv_mfma(rb[1][g], rb[1][g], d11);
        }
        colp = cola;
    }
}

/* X = C W^T for one row tile (two column tiles), W = the three tiles of a
   diagonal block's inverse. */
__device__ __forceinline__ void bv_solve_pair(const bv_d4 &c0, const bv_d4 &c1, const bv_d4 &w00,
                                              const bv_d4 &w10, const bv_d4 &w11, bv_d4 &x0,
                                              bv_d4 &x1) {
    x0 = (bv_d4){0.0, 0.0, 0.0, 0.0};
    x1 = (bv_d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        x0 = bv_mfma(w00[g], c0[g], x0);
        x1 = bv_mfma(w10[g], c0[g], x1);
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) x1 = bv_mfma(w11[g], c1[g], x1);
}

template <int KIND>
__global__ void __launch_bounds__(64, BV_STEP_WAVES_PER_SIMD)
bv_step_kernel(const bv_step_item *__restrict__ items, int nitems,
               const bv_block_desc *__restrict__ desc, const bv_eval_slot *__restrict__ slots,
               double *ws_base, double *part_logdet, double *part_dot, int32_t *part_info,
               int pmax) {
    constexpr int G = BV_STEP_ROWS;
    __shared__ double sW[3 * 256]; /* diagonal C tiles in, W out */
    __shared__ double sT[16 * 17];
    __shared__ double sCol[16], sRow[16], sPiv[BV_NB];
    __shared__ double s_logdet, s_dotd;
    __shared__ int s_info;
    __shared__ double qx[G * BV_TILE + BV_NB], qy[G * BV_TILE + BV_NB], qz[G * BV_TILE + BV_NB],
        qo[G * BV_TILE + BV_NB];
    __shared__ bv_kparams s_kp;
    __shared__ double s_etab[2 * BV_EXP_TAB]; /* bv_exp_tab_dev, copied */

    if ((int)blockIdx.x >= nitems) return;
    const bv_step_item it = items[blockIdx.x];
    if (it.desc < 0) return;
    BV_TRACE_WAVE(1, it.step);
    const int j = it.step;
    const bv_block_desc d = desc[it.desc];
    const bv_eval_slot &E = slots[d.eval];
    const int lane = threadIdx.x;
    const int m = d.m, cs0 = d.cs0;
    const int T = d.mp / BV_TILE, P = d.mp / BV_NB;
    const int rt0 = 2 * j;
    const int row0 = it.row0, nr = it.nrows;
    const bool lead = (row0 == rt0);
    double *ws = ws_base + d.ws_off;
    const int lr = lane & 15, lq = lane >> 4;
    double dot_lane = 0.0;
    bool has_m = false; /* this item holds the row tile of row m */
#pragma unroll
    for (int s = 0; s < G; ++s)
        if (s < nr && (row0 + s) * BV_TILE <= m && m < (row0 + s + 1) * BV_TILE) has_m = true;

    /* (a) solve panel j-1 for the own rows; keep L(R, 2j-2..2j-1) */
    bv_d4 xl[G][2];
    bv_d4 jl[2][2]; /* L(2j + r, 2j-2 + c), the panel rows of the last pair */
    if (j >= 1) {
        const int p0 = rt0 - 2;
        const bv_d4 w00 = *(const bv_d4 *)(ws + bv_tile_off(p0, p0, T) + 4 * lane);
        const bv_d4 w10 = *(const bv_d4 *)(ws + bv_tile_off(p0 + 1, p0, T) + 4 * lane);
        const bv_d4 w11 = *(const bv_d4 *)(ws + bv_tile_off(p0 + 1, p0 + 1, T) + 4 * lane);
#pragma unroll
        for (int s = 0; s < G; ++s) {
            if (s < nr) {
                const int rt = row0 + s;
                const bv_d4 c0 = *(const bv_d4 *)(ws + bv_tile_off(rt, p0, T) + 4 * lane);
                const bv_d4 c1 = *(const bv_d4 *)(ws + bv_tile_off(rt, p0 + 1, T) + 4 * lane);
                bv_solve_pair(c0, c1, w00, w10, w11, xl[s][0], xl[s][1]);
                /* rows 2j, 2j+1 (the lead's) keep C: other items of this step
                   read it to recompute the panel rows, and their L is used
                   only in this step */
                if (rt >= rt0 + 2) {
                    *(bv_d4 *)(ws + bv_tile_off(rt, p0, T) + 4 * lane) = xl[s][0];
                    *(bv_d4 *)(ws + bv_tile_off(rt, p0 + 1, T) + 4 * lane) = xl[s][1];
                }
                if (rt * BV_TILE + lr == m) {
                    const int J0p = p0 * BV_TILE;
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const int q0 = J0p + 4 * g + lq, q1 = q0 + BV_TILE;
                        if (q0 >= cs0 && q0 < m) dot_lane += xl[s][0][g] * xl[s][0][g];
                        if (q1 >= cs0 && q1 < m) dot_lane += xl[s][1][g] * xl[s][1][g];
                    }
                }
            }
        }
        if (j < P) {
            if (lead) {
                jl[0][0] = xl[0][0]; jl[0][1] = xl[0][1];
                jl[1][0] = xl[1][0]; jl[1][1] = xl[1][1];
            } else {
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const bv_d4 c0 = *(const bv_d4 *)(ws + bv_tile_off(rt0 + r, p0, T) + 4 * lane);
                    const bv_d4 c1 =
                        *(const bv_d4 *)(ws + bv_tile_off(rt0 + r, p0 + 1, T) + 4 * lane);
                    bv_solve_pair(c0, c1, w00, w10, w11, jl[r][0], jl[r][1]);
                }
            }
        }
        if (has_m) {
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) dot_lane += __shfl_xor(dot_lane, off, 64);
            if (lane == 0) part_dot[(int64_t)it.desc * pmax + (j - 1)] = dot_lane;
        }
    }
    if (j >= P) return; /* (only panels < P exist) */

    /* (b) C(R, j): -A, the last column pair from registers, the older
       columns from the workspace */
    if (KIND != BV_KIND_LOADA) {
        /* stage the item's row points and the panel's column points */
        const int pr = (lane < G * BV_TILE) ? row0 * BV_TILE + lane : -1;
        if (pr >= 0) {
            double xv, yv, zv, ov;
            bv_point(d, E, pr, xv, yv, zv, ov);
            qx[lane] = xv; qy[lane] = yv; qz[lane] = zv; qo[lane] = ov;
        }
        if (lane < BV_NB) {
            double xv, yv, zv, ov;
            bv_point(d, E, rt0 * BV_TILE + lane, xv, yv, zv, ov);
            const int t = G * BV_TILE + lane;
            qx[t] = xv; qy[t] = yv; qz[t] = zv; qo[t] = ov;
        }
        if (lane == 0) s_kp = E.kp;
        s_etab[lane] = bv_exp_tab_dev[lane];
        s_etab[lane + 64] = bv_exp_tab_dev[lane + 64];
        bv_wave_sync();
    }
    bv_d4 acc[G][2];
    const bv_d4 zero = (bv_d4){0.0, 0.0, 0.0, 0.0};
    {
        const bv_kparams &kp = s_kp;
        const bool planar = bv_planar_ok<KIND>(kp);
#pragma unroll
        for (int s = 0; s < G; ++s) {
            acc[s][0] = zero;
            acc[s][1] = zero;
            if (s >= nr) continue;
            const int rt = row0 + s;
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) {
                if (rt == rt0 && ct == 1) continue; /* upper tile of the diagonal block */
                if (KIND == BV_KIND_LOADA) {
                    acc[s][ct] = -*(const bv_d4 *)(ws + bv_tile_off(rt, rt0 + ct, T) + 4 * lane);
                } else {
                    const int p = rt * BV_TILE + lr;
                    const int ip = s * BV_TILE + lr;
                    const int q0 = (rt0 + ct) * BV_TILE;
                    bv_d4 v;
                    if (planar && rt * BV_TILE + 15 < m && q0 + 15 < m) {
#pragma unroll
                        for (int g = 0; g < 4; ++g) {
                            const int iq = G * BV_TILE + ct * BV_TILE + 4 * g + lq;
                            v[g] = bv_matern_planar<KIND>(qx[ip], qy[ip], qx[iq], qy[iq], kp,
                                                          kp.cheb, s_etab);
                        }
                    } else {
#pragma unroll
                        for (int g = 0; g < 4; ++g) {
                            const int iq = G * BV_TILE + ct * BV_TILE + 4 * g + lq;
                            v[g] = bv_aug_entry<KIND>(p, q0 + 4 * g + lq, m, qx[ip], qy[ip],
                                                      qz[ip], qx[iq], qy[iq], qz[iq], qo[iq], kp);
                        }
                    }
                    acc[s][ct] = -v;
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    if (j >= 1) {
        /* k = 2j-2, 2j-1 */
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
#pragma unroll
                for (int s = 0; s < G; ++s) {
                    if (s < nr) {
                        acc[s][0] = bv_mfma(jl[0][kk][g], xl[s][kk][g], acc[s][0]);
                        acc[s][1] = bv_mfma(jl[1][kk][g], xl[s][kk][g], acc[s][1]);
                    }
                }
            }
        }
    }
    if (j >= 2) {
        switch (nr) {
        case 1: bv_step_gemm<1, G>(ws, T, j, rt0 - 2, row0, lane, acc); break;
#if BV_STEP_ROWS >= 2
        case 2: bv_step_gemm<2, G>(ws, T, j, rt0 - 2, row0, lane, acc); break;
#endif
#if BV_STEP_ROWS >= 3
        case 3: bv_step_gemm<3, G>(ws, T, j, rt0 - 2, row0, lane, acc); break;
#endif
#if BV_STEP_ROWS >= 4
        case 4: bv_step_gemm<4, G>(ws, T, j, rt0 - 2, row0, lane, acc); break;
#endif
        default: break;
        }
    }

    /* (c) store C, or factor the diagonal block */
    if (!lead) {
#pragma unroll
        for (int s = 0; s < G; ++s) {
            if (s < nr) {
                const int rt = row0 + s;
                *(bv_d4 *)(ws + bv_tile_off(rt, rt0, T) + 4 * lane) = -acc[s][0];
                *(bv_d4 *)(ws + bv_tile_off(rt, rt0 + 1, T) + 4 * lane) = -acc[s][1];
            }
        }
        return;
    }
    /* lead: rows rt0, rt0+1 are s = 0, 1 (nr >= 2: the row count 2P - 2j is
       even); its further rows store C as the other items do */
#pragma unroll
    for (int s = 2; s < G; ++s) {
        if (s < nr) {
            const int rt = row0 + s;
            *(bv_d4 *)(ws + bv_tile_off(rt, rt0, T) + 4 * lane) = -acc[s][0];
            *(bv_d4 *)(ws + bv_tile_off(rt, rt0 + 1, T) + 4 * lane) = -acc[s][1];
        }
    }
    *(bv_d4 *)(sW + 0 * 256 + 4 * lane) = -acc[0][0];
    *(bv_d4 *)(sW + 1 * 256 + 4 * lane) = -acc[1][0];
    *(bv_d4 *)(sW + 2 * 256 + 4 * lane) = -acc[1][1];
    if (lane == 0) { s_logdet = 0.0; s_dotd = 0.0; s_info = 0; }
    bv_wave_sync();
    bv_diag_factor_tiles(sW, sW + 256, sW + 512, lane, rt0 * BV_TILE, m, cs0, sCol, sRow, sT, sW,
                         sPiv, &s_dotd, &s_info);
    bv_panel_logdet(lane, rt0 * BV_TILE, m, cs0, sPiv, &s_logdet);
    bv_wave_sync();
    *(bv_d4 *)(ws + bv_tile_off(rt0, rt0, T) + 4 * lane) = *(const bv_d4 *)(sW + 0 * 256 + 4 * lane);
    *(bv_d4 *)(ws + bv_tile_off(rt0 + 1, rt0, T) + 4 * lane) =
        *(const bv_d4 *)(sW + 1 * 256 + 4 * lane);
    *(bv_d4 *)(ws + bv_tile_off(rt0 + 1, rt0 + 1, T) + 4 * lane) =
        *(const bv_d4 *)(sW + 2 * 256 + 4 * lane);
    if (lane == 0) {
        const int64_t o = (int64_t)it.desc * pmax + j;
        part_logdet[o] = s_logdet;
        part_dot[o] = s_dotd; /* row m in the diagonal block; else the solve of
                                 step j+1 overwrites it */
        part_info[o] = s_info;
    }
}

/* ======================================================================
 * Stepped factorization with lookahead: launches j = -1, 0, ..., P-2.
 *
 * Launch j (j >= 0): a work item is one wavefront on up to BV_STEP_ROWS row
 * tiles R >= 2j+2 of one block.  It
 *   (a) forms C(R, j) = A(R, j) - sum_{k<2j} L(R, k) L(2j..2j+1, k)^T (A
 *       generated in registers or read for LOADA; every L from earlier
 *       launches),
 *   (b) solves L(R, j) = C W_j^T with W_j, the inverse of panel j's diagonal
 *       factor, written by launch j-1, stores L(R, j) and adds row m's squares
 *       to the (block, panel j) partial;
 *   (c) if it holds rows 2j+2 and 2j+3 (panel j+1's diagonal block), forms
 *       that block's C from its own rows (older columns from the workspace,
 *       columns 2j, 2j+1 from (b) in registers), factors it in LDS, stores
 *       W_{j+1} over its three tiles and the panel's log-determinant, row-m
 *       squares and first failure into the (block, panel j+1) partials.
 * Launch -1 is (c) alone for panel 0.  What an item reads was written by an
 * earlier launch and what it writes no other item of its launch touches, so
 * items never wait for each other.  Compared with bv_step_kernel, C never
 * goes through the workspace and no item recomputes the panel rows.
 * ====================================================================== */
template <int KIND>
__global__ void __launch_bounds__(64, BV_STEP_WAVES_PER_SIMD)
bv_step2_kernel(const bv_step_item *__restrict__ items, int nitems,
                const bv_block_desc *__restrict__ desc, const bv_eval_slot *__restrict__ slots,
                double *ws_base, double *part_logdet, double *part_dot, int32_t *part_info,
                int pmax) {
    constexpr int G = BV_STEP_ROWS;
    static_assert(G >= 2, "the lookahead item holds two row tiles");
    __shared__ double sW[3 * 256]; /* diagonal C tiles in, W out */
    __shared__ double sT[16 * 17];
    __shared__ double sCol[16], sRow[16], sPiv[BV_NB];
    __shared__ double s_logdet, s_dotd;
    __shared__ int s_info;
    __shared__ double qx[G * BV_TILE + BV_NB], qy[G * BV_TILE + BV_NB], qz[G * BV_TILE + BV_NB],
        qo[G * BV_TILE + BV_NB];
    __shared__ bv_kparams s_kp;
    __shared__ double s_etab[2 * BV_EXP_TAB]; /* bv_exp_tab_dev, copied */

    if ((int)blockIdx.x >= nitems) return;
    const bv_step_item it = items[blockIdx.x];
    if (it.desc < 0) return;
    BV_TRACE_WAVE(2, it.step);
    const int j = it.step;
    const bv_block_desc d = desc[it.desc];
    const bv_eval_slot &E = slots[d.eval];
    const int lane = threadIdx.x;
    const int m = d.m, cs0 = d.cs0;
    const int T = d.mp / BV_TILE, P = d.mp / BV_NB;
    const int row0 = it.row0, nr = it.nrows;
    const int jn = j + 1;                         /* the panel this launch factors */
    const int rn0 = 2 * jn;                       /* its diagonal block's first row tile */
    const bool look = (row0 == rn0) && (jn < P) && nr >= 2;
#if BV_LOOK_PRIO
    /* the lookahead item is the launch's longest serial chain: its
       instructions issue first on its SIMD */
    if (look) __builtin_amdgcn_s_setprio(3);
#endif
    double *ws = ws_base + d.ws_off;
    const int lr = lane & 15, lq = lane >> 4;

    if (KIND != BV_KIND_LOADA) {
        /* the item's row points, then panel j's column points */
        if (lane < G * BV_TILE) {
            double xv, yv, zv, ov;
            bv_point(d, E, row0 * BV_TILE + lane, xv, yv, zv, ov);
            qx[lane] = xv; qy[lane] = yv; qz[lane] = zv; qo[lane] = ov;
        }
        if (j >= 0 && lane < BV_NB) {
            double xv, yv, zv, ov;
            bv_point(d, E, 2 * j * BV_TILE + lane, xv, yv, zv, ov);
            const int t = G * BV_TILE + lane;
            qx[t] = xv; qy[t] = yv; qz[t] = zv; qo[t] = ov;
        }
        if (lane == 0) s_kp = E.kp;
        s_etab[lane] = bv_exp_tab_dev[lane];
        s_etab[lane + 64] = bv_exp_tab_dev[lane + 64];
        bv_wave_sync();
    }
    const bv_kparams &kp = s_kp;
    const bool planar = bv_planar_ok<KIND>(kp);
    const bv_d4 zero = (bv_d4){0.0, 0.0, 0.0, 0.0};

    bv_d4 xl[G][2]; /* L(R, panel j) */
    /* the lookahead item's products for panel jn's diagonal block over the
       columns before 2j, accumulated in (a)'s pass over its own rows */
    bv_d4 d00 = zero, d10 = zero, d11 = zero;
    const bool fused = look && nr == 2 && j >= 1;
    if (j >= 0) {
        const int rt0 = 2 * j;
        bv_d4 acc[G][2];
        /* (a) acc = -A(R, j), then the products of the older columns */
#pragma unroll
        for (int s = 0; s < G; ++s) {
            acc[s][0] = zero;
            acc[s][1] = zero;
            if (s >= nr) continue;
            const int rt = row0 + s;
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) {
                if (KIND == BV_KIND_LOADA) {
                    acc[s][ct] = -*(const bv_d4 *)(ws + bv_tile_off(rt, rt0 + ct, T) + 4 * lane);
                } else {
                    const int p = rt * BV_TILE + lr;
                    const int ip = s * BV_TILE + lr;
                    const int q0 = (rt0 + ct) * BV_TILE;
                    bv_d4 v;
                    if (planar && rt * BV_TILE + 15 < m && q0 + 15 < m) {
#pragma unroll
                        for (int g = 0; g < 4; ++g) {
                            const int iq = G * BV_TILE + ct * BV_TILE + 4 * g + lq;
                            v[g] = bv_matern_planar<KIND>(qx[ip], qy[ip], qx[iq], qy[iq], kp,
                                                          kp.cheb, s_etab);
                        }
                    } else {
#pragma unroll
                        for (int g = 0; g < 4; ++g) {
                            const int iq = G * BV_TILE + ct * BV_TILE + 4 * g + lq;
                            v[g] = bv_aug_entry<KIND>(p, q0 + 4 * g + lq, m, qx[ip], qy[ip],
                                                      qz[ip], qx[iq], qy[iq], qz[iq], qo[iq], kp);
                        }
                    }
                    acc[s][ct] = -v;
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (fused) {
            bv_step_gemm_syrk<G>(ws, T, j, rt0, row0, lane, acc, d00, d10, d11);
        } else if (j >= 1) {
            switch (nr) {
            case 1: bv_step_gemm<1, G>(ws, T, j, rt0, row0, lane, acc); break;
#if BV_STEP_ROWS >= 2
            case 2: bv_step_gemm<2, G>(ws, T, j, rt0, row0, lane, acc); break;
#endif
#if BV_STEP_ROWS >= 3
            case 3: bv_step_gemm<3, G>(ws, T, j, rt0, row0, lane, acc); break;
#endif
#if BV_STEP_ROWS >= 4
            case 4: bv_step_gemm<4, G>(ws, T, j, rt0, row0, lane, acc); break;
#endif
            default: break;
            }
        }
        /* (b) L(R, j) = C W_j^T; row m's squares */
        const bv_d4 w00 = *(const bv_d4 *)(ws + bv_tile_off(rt0, rt0, T) + 4 * lane);
        const bv_d4 w10 = *(const bv_d4 *)(ws + bv_tile_off(rt0 + 1, rt0, T) + 4 * lane);
        const bv_d4 w11 = *(const bv_d4 *)(ws + bv_tile_off(rt0 + 1, rt0 + 1, T) + 4 * lane);
        double dot_lane = 0.0;
        bool has_m = false;
#pragma unroll
        for (int s = 0; s < G; ++s) {
            if (s >= nr) continue;
            const int rt = row0 + s;
            bv_solve_pair(-acc[s][0], -acc[s][1], w00, w10, w11, xl[s][0], xl[s][1]);
            *(bv_d4 *)(ws + bv_tile_off(rt, rt0, T) + 4 * lane) = xl[s][0];
            *(bv_d4 *)(ws + bv_tile_off(rt, rt0 + 1, T) + 4 * lane) = xl[s][1];
            if (rt * BV_TILE <= m && m < (rt + 1) * BV_TILE) has_m = true;
            if (rt * BV_TILE + lr == m) {
                const int J0 = rt0 * BV_TILE;
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int q0 = J0 + 4 * g + lq, q1 = q0 + BV_TILE;
                    if (q0 >= cs0 && q0 < m) dot_lane += xl[s][0][g] * xl[s][0][g];
                    if (q1 >= cs0 && q1 < m) dot_lane += xl[s][1][g] * xl[s][1][g];
                }
            }
        }
        if (has_m) {
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) dot_lane += __shfl_xor(dot_lane, off, 64);
            if (lane == 0) part_dot[(int64_t)it.desc * pmax + j] = dot_lane;
        }
    }
    if (!look) return;

    /* (c) panel jn's diagonal block from the own rows s = 0, 1 */
    bv_d4 acc2[2][2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            acc2[s][ct] = zero;
            if (s == 0 && ct == 1) continue; /* upper tile */
            const int rt = rn0 + s;
            if (KIND == BV_KIND_LOADA) {
                acc2[s][ct] = -*(const bv_d4 *)(ws + bv_tile_off(rt, rn0 + ct, T) + 4 * lane);
            } else {
                const int p = rt * BV_TILE + lr;
                const int ip = s * BV_TILE + lr;
                const int q0 = (rn0 + ct) * BV_TILE;
                bv_d4 v;
                if (planar && rt * BV_TILE + 15 < m && q0 + 15 < m) {
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const int iq = ct * BV_TILE + 4 * g + lq; /* own rows are the columns */
                        v[g] = bv_matern_planar<KIND>(qx[ip], qy[ip], qx[iq], qy[iq], kp,
                                                      kp.cheb, s_etab);
                    }
                } else {
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const int iq = ct * BV_TILE + 4 * g + lq;
                        v[g] = bv_aug_entry<KIND>(p, q0 + 4 * g + lq, m, qx[ip], qy[ip], qz[ip],
                                                  qx[iq], qy[iq], qz[iq], qo[iq], kp);
                    }
                }
                acc2[s][ct] = -v;
            }
        }
    }
    if (fused) {
        acc2[0][0] += d00;
        acc2[1][0] += d10;
        acc2[1][1] += d11;
    } else if (j >= 1) {
        bv_step_gemm<2, 2>(ws, T, jn, 2 * j, rn0, lane, acc2);
    }
    if (j >= 0) {
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    acc2[s][0] = bv_mfma(xl[0][kk][g], xl[s][kk][g], acc2[s][0]);
                    acc2[s][1] = bv_mfma(xl[1][kk][g], xl[s][kk][g], acc2[s][1]);
                }
            }
        }
    }
    *(bv_d4 *)(sW + 0 * 256 + 4 * lane) = -acc2[0][0];
    *(bv_d4 *)(sW + 1 * 256 + 4 * lane) = -acc2[1][0];
    *(bv_d4 *)(sW + 2 * 256 + 4 * lane) = -acc2[1][1];
    if (lane == 0) { s_logdet = 0.0; s_dotd = 0.0; s_info = 0; }
    bv_wave_sync();
    bv_diag_factor_tiles(sW, sW + 256, sW + 512, lane, rn0 * BV_TILE, m, cs0, sCol, sRow, sT, sW,
                         sPiv, &s_dotd, &s_info);
    bv_panel_logdet(lane, rn0 * BV_TILE, m, cs0, sPiv, &s_logdet);
    bv_wave_sync();
    *(bv_d4 *)(ws + bv_tile_off(rn0, rn0, T) + 4 * lane) = *(const bv_d4 *)(sW + 0 * 256 + 4 * lane);
    *(bv_d4 *)(ws + bv_tile_off(rn0 + 1, rn0, T) + 4 * lane) =
        *(const bv_d4 *)(sW + 1 * 256 + 4 * lane);
    *(bv_d4 *)(ws + bv_tile_off(rn0 + 1, rn0 + 1, T) + 4 * lane) =
        *(const bv_d4 *)(sW + 2 * 256 + 4 * lane);
    if (lane == 0) {
        const int64_t o = (int64_t)it.desc * pmax + jn;
        part_logdet[o] = s_logdet;
        part_dot[o] = s_dotd; /* row m in the diagonal block; else the solve of
                                 launch jn overwrites it */
        part_info[o] = s_info;
    }
}

/* ======================================================================
 * Paired-panel factorization: one launch per pair of 32-column panels
 * (2s, 2s+1), i.e. tile columns c0 = 4s .. 4s+3, with a lookahead of one
 * pair.  A workgroup is two wavefronts on four consecutive row tiles of one
 * block (wave w: rows row0 + 2w, row0 + 2w + 1), all rows >= 4s + 4.  At
 * launch s, with W = L_pp^-1 of panels 2s and 2s+1 and L(panel 2s+1 rows,
 * panel 2s) from launch s-1, each wave
 *   (a) forms C(R, pair) = A(R, pair) - sum_{k < c0} L(R, k) L(pair rows, k)^T
 *       (four output column tiles per load of its own row tiles: twice the
 *       products per byte of the one-panel steps),
 *   (b) solves panel 2s, L(R, 2s) = C(R, 2s) W_2s^T, subtracts
 *       L(R, 2s) L(panel 2s+1 rows, 2s)^T from C(R, 2s+1) and solves panel
 *       2s+1; stores L(R, pair) and adds row m's squares per panel.
 * The workgroup on rows 4s+4 .. 4s+7 (the next pair's rows Q = A u B) also
 * factors the next pair's 64x64 diagonal block D = A(Q,Q) - L(Q, <c0+4)
 * L(Q, <c0+4)^T: wave A's 2x2 tile block is accumulated with its pass (a),
 * wave B's own block likewise and its cross block D_BA by a second pass over
 * the columns < c0 (the pair's columns from wave A's L tiles, stored before a
 * workgroup barrier) while wave A factors D_AA (W_A); then wave B forms
 * L_BA = D_BA W_A^T, D_BB - L_BA L_BA^T and factors it (W_B).  W_A, L_BA and
 * W_B go to the workspace with the panels' log-determinants, row-m squares
 * and first failures in the (block, panel) partials.  Launch -1 is that
 * factorization alone for pair 0.  Nothing a workgroup reads is written by
 * another workgroup of the same launch: workgroups never wait for each other.
 * ====================================================================== */
#ifndef BV_PAIR_WAVES_PER_SIMD
#define BV_PAIR_WAVES_PER_SIMD 2
#endif
#ifndef BV_PAIR_PINGPONG
#define BV_PAIR_PINGPONG 1 /* two operand buffers in the column loop */
#endif

/* The pair's four row tiles and two own row tiles of one tile column. */
__device__ __forceinline__ void bv_pair_load(const double *colp, int c0, int rbase, int lo,
                                             bv_d4 (&p)[4], bv_d4 (&r)[2]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) p[c] = *(const bv_d4 *)(colp + ((int64_t)(c0 + c) << 8) + lo);
#pragma unroll
    for (int i = 0; i < 2; ++i) r[i] = *(const bv_d4 *)(colp + ((int64_t)(rbase + i) << 8) + lo);
}

template <bool SYRK>
__device__ __forceinline__ void bv_pair_step(const bv_d4 (&p)[4], const bv_d4 (&r)[2],
                                             bv_d4 (&acc)[2][4], bv_d4 &d00, bv_d4 &d10,
                                             bv_d4 &d11) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[i][c] = bv_mfma(p[c][g], r[i][g], acc[i][c]);
        }
        if (SYRK) {
            d00 = bv_mfma(r[0][g], r[0][g], d00);
            d10 = bv_mfma(r[0][g], r[1][g], d10);
            d11 = bv_mfma(r[1][g], r[1][g], d11);
        }
    }
}

/* acc[i][c] += L(rbase + i, k) L(c0 + c, k)^T over the tile columns k < kend
   (kend even), and with SYRK the own lower 2x2 tile block from the same
   loaded tiles.  Two operand buffers alternate: the loads of column k+1 are in
   flight while column k's products issue. */
template <bool SYRK>
__device__ __forceinline__ void bv_pair_gemm(const double *__restrict__ ws, int T, int c0, int kend,
                                             int rbase, int lane, bv_d4 (&acc)[2][4], bv_d4 &d00,
                                             bv_d4 &d10, bv_d4 &d11) {
    if (kend <= 0) return;
    const int lo = 4 * lane;
    const double *colp = ws;
#if BV_PAIR_PINGPONG
    bv_d4 pa[4], ra[2], pb[4], rb[2];
    bv_pair_load(colp, c0, rbase, lo, pa, ra);
    for (int kt = 0; kt < kend; kt += 2) {
        const double *colb = colp + ((int64_t)(T - kt - 1) << 8);
        bv_pair_load(colb, c0, rbase, lo, pb, rb);
        bv_pair_step<SYRK>(pa, ra, acc, d00, d10, d11);
        const double *cola = (kt + 2 < kend) ? colb + ((int64_t)(T - kt - 2) << 8) : colb;
        bv_pair_load(cola, c0, rbase, lo, pa, ra);
        bv_pair_step<SYRK>(pb, rb, acc, d00, d10, d11);
        colp = cola;
    }
#else
    for (int kt = 0; kt < kend; ++kt) {
        bv_d4 pa[4], ra[2];
        bv_pair_load(colp, c0, rbase, lo, pa, ra);
        bv_pair_step<SYRK>(pa, ra, acc, d00, d10, d11);
        colp += ((int64_t)(T - kt - 1) << 8);
    }
#endif
}

/* Tile (rt, ct) of the augmented matrix: generated from the points staged in
   LDS (row points from index ip0, column points from iq0) or, for LOADA, read
   from the workspace. */
template <int KIND>
__device__ __forceinline__ bv_d4 bv_pair_tile(const double *__restrict__ ws, int T, int rt, int ct,
                                              int m, int ip0, int iq0, int lane,
                                              const double *qx, const double *qy,
                                              const double *qz, const double *qo,
                                              const bv_kparams &kp, bool planar,
                                              const double *cheb, const double *etab) {
    if (KIND == BV_KIND_LOADA) return *(const bv_d4 *)(ws + bv_tile_off(rt, ct, T) + 4 * lane);
    const int lr = lane & 15, lq = lane >> 4;
    const int p = rt * BV_TILE + lr, ip = ip0 + lr, q0 = ct * BV_TILE;
    bv_d4 v;
    if (planar && rt * BV_TILE + 15 < m && q0 + 15 < m) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int iq = iq0 + 4 * g + lq;
            v[g] = bv_matern_planar<KIND>(qx[ip], qy[ip], qx[iq], qy[iq], kp, cheb, etab);
        }
    } else {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int iq = iq0 + 4 * g + lq;
            v[g] = bv_aug_entry<KIND>(p, q0 + 4 * g + lq, m, qx[ip], qy[ip], qz[ip], qx[iq], qy[iq],
                                      qz[iq], qo[iq], kp);
        }
    }
    return v;
}

/* Squares of row m's entries in columns [J0, J0 + 16 NC) that are >= cs0 and
   < m, for the lanes holding row m of tile rt (else 0). */
template <int NC>
__device__ __forceinline__ double bv_row_m_sq(const bv_d4 (&x)[NC], int rt, int J0, int m, int cs0,
                                              int lane) {
    const int lr = lane & 15, lq = lane >> 4;
    double t = 0.0;
    if (rt * BV_TILE + lr == m) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int q = J0 + c * BV_TILE + 4 * g + lq;
                if (q >= cs0 && q < m) t += x[c][g] * x[c][g];
            }
        }
    }
    return t;
}

/* d += L(r, pair) L(r', pair)^T for the own rows r, r' in {rbase, rbase + 1}
   (lower 2x2 tile block), the pair's four tile columns read back from the
   workspace where this wavefront stored them. */
__device__ __forceinline__ void bv_pair_syrk_own(const double *__restrict__ ws, int T, int rbase,
                                                 int c0, int lane, bv_d4 &d00, bv_d4 &d10,
                                                 bv_d4 &d11) {
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        const bv_d4 x0 = *(const bv_d4 *)(ws + bv_tile_off(rbase, c0 + kk, T) + 4 * lane);
        const bv_d4 x1 = *(const bv_d4 *)(ws + bv_tile_off(rbase + 1, c0 + kk, T) + 4 * lane);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            d00 = bv_mfma(x0[g], x0[g], d00);
            d10 = bv_mfma(x0[g], x1[g], d10);
            d11 = bv_mfma(x1[g], x1[g], d11);
        }
    }
}

__device__ __forceinline__ double bv_wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <int KIND>
__global__ void __launch_bounds__(128, BV_PAIR_WAVES_PER_SIMD)
bv_pair_kernel(const bv_step_item *__restrict__ items, int nitems,
               const bv_block_desc *__restrict__ desc, const bv_eval_slot *__restrict__ slots,
               double *ws_base, double *part_logdet, double *part_dot, int32_t *part_info,
               int pmax) {
    __shared__ double qx[128], qy[128], qz[128], qo[128]; /* row points, then the pair's columns */
    __shared__ double sW[2][3 * 256];                     /* each wave's diagonal block, W out */
    __shared__ double sT[2][16 * 17];
    __shared__ double sCol[2][16], sRow[2][16], sPiv[2][BV_NB];
    __shared__ double s_logdet[2], s_dotd[2];
    __shared__ int s_info[2];
    __shared__ bv_kparams s_kp;
    __shared__ double s_cheb[KIND == BV_KIND_GENERAL ? BV_CHEB_SIZE : 1]; /* kp.cheb, copied */
    __shared__ double s_etab[2 * BV_EXP_TAB];                             /* bv_exp_tab_dev   */

    if ((int)blockIdx.x >= nitems) return;
    const bv_step_item it = items[blockIdx.x];
    if (it.desc < 0) return;
    BV_TRACE_WAVE(3, it.step);
    const int s = it.step;
    const bv_block_desc d = desc[it.desc];
    const bv_eval_slot &E = slots[d.eval];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int m = d.m, cs0 = d.cs0;
    const int T = d.mp / BV_TILE;
    const int row0 = it.row0, nr = it.nrows;
    const int c0 = 4 * s;
    const bool look = row0 == 4 * s + 4; /* this workgroup holds the next pair's rows */
#if BV_LOOK_PRIO
    if (look) __builtin_amdgcn_s_setprio(3); /* the launch's longest serial chain */
#endif
    const int myr = row0 + 2 * w;
    const bool mine = 2 * w < nr;
    double *ws = ws_base + d.ws_off;

    if (KIND != BV_KIND_LOADA) {
        if (tid < 64 || s >= 0) {
            const int p = (tid < 64) ? row0 * BV_TILE + tid : c0 * BV_TILE + (tid - 64);
            double xv, yv, zv, ov;
            bv_point(d, E, p, xv, yv, zv, ov);
            qx[tid] = xv; qy[tid] = yv; qz[tid] = zv; qo[tid] = ov;
        }
        if (tid == 0) s_kp = E.kp;
        if (KIND == BV_KIND_GENERAL)
            for (int i = tid; i < BV_CHEB_SIZE; i += 128) s_cheb[i] = E.kp.cheb[i];
        s_etab[tid] = bv_exp_tab_dev[tid]; /* 128 threads, 128 entries */
    }
    __syncthreads();
    const bv_kparams &kp = s_kp;
    const bool planar = bv_planar_ok<KIND>(kp);
    const bv_d4 zero = (bv_d4){0.0, 0.0, 0.0, 0.0};

    bv_d4 d00 = zero, d10 = zero, d11 = zero; /* own rows' SYRK over k < c0 */
    if (s >= 0 && mine) {
        /* -A(R, pair), written out: the generation is too large for the
           unroller, and a rolled loop would index the accumulators
           dynamically (scratch memory) */
        auto gen = [&](int i, int c) {
            return -bv_pair_tile<KIND>(ws, T, myr + i, c0 + c, m, (2 * w + i) * BV_TILE,
                                       64 + c * BV_TILE, lane, qx, qy, qz, qo, kp, planar, s_cheb,
                                       s_etab);
        };
        bv_d4 acc[2][4] = {{gen(0, 0), gen(0, 1), gen(0, 2), gen(0, 3)},
                           {gen(1, 0), gen(1, 1), gen(1, 2), gen(1, 3)}};
        if (look) bv_pair_gemm<true>(ws, T, c0, c0, myr, lane, acc, d00, d10, d11);
        else bv_pair_gemm<false>(ws, T, c0, c0, myr, lane, acc, d00, d10, d11);
        const int lo = 4 * lane;
        bv_d4 x[2][2]; /* L(own rows, one panel of the pair) */
        bool has_m = false;
        double dp0 = 0.0, dp1 = 0.0;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int rt = myr + i;
            if (rt * BV_TILE <= m && m < (rt + 1) * BV_TILE) has_m = true;
        }
        {
            const bv_d4 w00 = *(const bv_d4 *)(ws + bv_tile_off(c0, c0, T) + lo);
            const bv_d4 w10 = *(const bv_d4 *)(ws + bv_tile_off(c0 + 1, c0, T) + lo);
            const bv_d4 w11 = *(const bv_d4 *)(ws + bv_tile_off(c0 + 1, c0 + 1, T) + lo);
#pragma unroll
            for (int i = 0; i < 2; ++i)
                bv_solve_pair(-acc[i][0], -acc[i][1], w00, w10, w11, x[i][0], x[i][1]);
        }
        {
            /* C(R, 2s+1) -= L(R, 2s) L(panel 2s+1 rows, 2s)^T, then L(R, 2s) is stored */
            const bv_d4 l00 = *(const bv_d4 *)(ws + bv_tile_off(c0 + 2, c0, T) + lo);
            const bv_d4 l01 = *(const bv_d4 *)(ws + bv_tile_off(c0 + 2, c0 + 1, T) + lo);
            const bv_d4 l10 = *(const bv_d4 *)(ws + bv_tile_off(c0 + 3, c0, T) + lo);
            const bv_d4 l11 = *(const bv_d4 *)(ws + bv_tile_off(c0 + 3, c0 + 1, T) + lo);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    acc[i][2] = bv_mfma(l00[g], x[i][0][g], acc[i][2]);
                    acc[i][2] = bv_mfma(l01[g], x[i][1][g], acc[i][2]);
                    acc[i][3] = bv_mfma(l10[g], x[i][0][g], acc[i][3]);
                    acc[i][3] = bv_mfma(l11[g], x[i][1][g], acc[i][3]);
                }
            }
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int rt = myr + i;
                *(bv_d4 *)(ws + bv_tile_off(rt, c0, T) + lo) = x[i][0];
                *(bv_d4 *)(ws + bv_tile_off(rt, c0 + 1, T) + lo) = x[i][1];
                dp0 += bv_row_m_sq<2>(x[i], rt, c0 * BV_TILE, m, cs0, lane);
            }
        }
        {
            const bv_d4 w00 = *(const bv_d4 *)(ws + bv_tile_off(c0 + 2, c0 + 2, T) + lo);
            const bv_d4 w10 = *(const bv_d4 *)(ws + bv_tile_off(c0 + 3, c0 + 2, T) + lo);
            const bv_d4 w11 = *(const bv_d4 *)(ws + bv_tile_off(c0 + 3, c0 + 3, T) + lo);
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int rt = myr + i;
                bv_solve_pair(-acc[i][2], -acc[i][3], w00, w10, w11, x[i][0], x[i][1]);
                *(bv_d4 *)(ws + bv_tile_off(rt, c0 + 2, T) + lo) = x[i][0];
                *(bv_d4 *)(ws + bv_tile_off(rt, c0 + 3, T) + lo) = x[i][1];
                dp1 += bv_row_m_sq<2>(x[i], rt, (c0 + 2) * BV_TILE, m, cs0, lane);
            }
        }
        if (has_m) {
            dp0 = bv_wave_sum(dp0);
            dp1 = bv_wave_sum(dp1);
            if (lane == 0) {
                part_dot[(int64_t)it.desc * pmax + 2 * s] = dp0;
                part_dot[(int64_t)it.desc * pmax + 2 * s + 1] = dp1;
            }
        }
    }
    if (!look) return;

    /* the next pair's diagonal block: Q = rows qn .. qn+3 (wave A: qn, qn+1;
       wave B: qn+2, qn+3 when they exist); accumulators hold -D */
    const int qn = row0;
    const bool hasB = qn + 2 < T;
    const int pa = qn / 2;
    bv_d4 db0 = zero, db1 = zero, db2 = zero; /* wave B: D_BB */
    bv_d4 e[2][2];                            /* wave B: D_BA */
    if (w == 0) {
        bv_d4 da0 = d00 - bv_pair_tile<KIND>(ws, T, qn, qn, m, 0, 0, lane, qx, qy, qz, qo, kp,
                                             planar, s_cheb, s_etab);
        bv_d4 da1 = d10 - bv_pair_tile<KIND>(ws, T, qn + 1, qn, m, 16, 0, lane, qx, qy, qz, qo, kp,
                                             planar, s_cheb, s_etab);
        bv_d4 da2 = d11 - bv_pair_tile<KIND>(ws, T, qn + 1, qn + 1, m, 16, 16, lane, qx, qy, qz, qo,
                                             kp, planar, s_cheb, s_etab);
        if (s >= 0) bv_pair_syrk_own(ws, T, myr, c0, lane, da0, da1, da2);
        *(bv_d4 *)(sW[0] + 0 * 256 + 4 * lane) = -da0;
        *(bv_d4 *)(sW[0] + 1 * 256 + 4 * lane) = -da1;
        *(bv_d4 *)(sW[0] + 2 * 256 + 4 * lane) = -da2;
        if (lane == 0) { s_logdet[0] = 0.0; s_dotd[0] = 0.0; s_info[0] = 0; }
    } else if (hasB) {
        db0 = d00 - bv_pair_tile<KIND>(ws, T, qn + 2, qn + 2, m, 32, 32, lane, qx, qy, qz, qo, kp,
                                       planar, s_cheb, s_etab);
        db1 = d10 - bv_pair_tile<KIND>(ws, T, qn + 3, qn + 2, m, 48, 32, lane, qx, qy, qz, qo, kp,
                                       planar, s_cheb, s_etab);
        db2 = d11 - bv_pair_tile<KIND>(ws, T, qn + 3, qn + 3, m, 48, 48, lane, qx, qy, qz, qo, kp,
                                       planar, s_cheb, s_etab);
        if (s >= 0) bv_pair_syrk_own(ws, T, myr, c0, lane, db0, db1, db2);
        if (lane == 0) { s_logdet[1] = 0.0; s_dotd[1] = 0.0; s_info[1] = 0; }
    }
    __syncthreads(); /* sW[0]; wave A's L(Q_A, pair) tiles in the workspace */
    if (w == 0) {
        bv_diag_factor_tiles(sW[0], sW[0] + 256, sW[0] + 512, lane, qn * BV_TILE, m, cs0, sCol[0],
                             sRow[0], sT[0], sW[0], sPiv[0], &s_dotd[0], &s_info[0]);
        bv_panel_logdet(lane, qn * BV_TILE, m, cs0, sPiv[0], &s_logdet[0]);
        bv_wave_sync();
        *(bv_d4 *)(ws + bv_tile_off(qn, qn, T) + 4 * lane) = *(const bv_d4 *)(sW[0] + 4 * lane);
        *(bv_d4 *)(ws + bv_tile_off(qn + 1, qn, T) + 4 * lane) =
            *(const bv_d4 *)(sW[0] + 256 + 4 * lane);
        *(bv_d4 *)(ws + bv_tile_off(qn + 1, qn + 1, T) + 4 * lane) =
            *(const bv_d4 *)(sW[0] + 512 + 4 * lane);
    } else if (hasB) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int c = 0; c < 2; ++c)
                e[i][c] = -bv_pair_tile<KIND>(ws, T, qn + 2 + i, qn + c, m, (2 + i) * BV_TILE,
                                              c * BV_TILE, lane, qx, qy, qz, qo, kp, planar,
                                              s_cheb, s_etab);
        }
        if (s >= 0) {
            bv_step_gemm<2, 2>(ws, T, qn / 2, c0, qn + 2, lane, e);
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                /* stored by wave A before the barrier */
                const bv_d4 la0 = *(const bv_d4 *)(ws + bv_tile_off(qn, c0 + kk, T) + 4 * lane);
                const bv_d4 la1 = *(const bv_d4 *)(ws + bv_tile_off(qn + 1, c0 + kk, T) + 4 * lane);
                const bv_d4 lb0 = *(const bv_d4 *)(ws + bv_tile_off(qn + 2, c0 + kk, T) + 4 * lane);
                const bv_d4 lb1 = *(const bv_d4 *)(ws + bv_tile_off(qn + 3, c0 + kk, T) + 4 * lane);
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    e[0][0] = bv_mfma(la0[g], lb0[g], e[0][0]);
                    e[0][1] = bv_mfma(la1[g], lb0[g], e[0][1]);
                    e[1][0] = bv_mfma(la0[g], lb1[g], e[1][0]);
                    e[1][1] = bv_mfma(la1[g], lb1[g], e[1][1]);
                }
            }
        }
    }
    __syncthreads(); /* W_A in sW[0]; wave A's partials */
    if (w == 1 && hasB) {
        const bv_d4 wa00 = *(const bv_d4 *)(sW[0] + 4 * lane);
        const bv_d4 wa10 = *(const bv_d4 *)(sW[0] + 256 + 4 * lane);
        const bv_d4 wa11 = *(const bv_d4 *)(sW[0] + 512 + 4 * lane);
        bv_d4 lba[2][2];
        double dba = 0.0;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            bv_solve_pair(-e[i][0], -e[i][1], wa00, wa10, wa11, lba[i][0], lba[i][1]);
            const int rt = qn + 2 + i;
            *(bv_d4 *)(ws + bv_tile_off(rt, qn, T) + 4 * lane) = lba[i][0];
            *(bv_d4 *)(ws + bv_tile_off(rt, qn + 1, T) + 4 * lane) = lba[i][1];
            dba += bv_row_m_sq<2>(lba[i], rt, qn * BV_TILE, m, cs0, lane);
        }
        dba = bv_wave_sum(dba);
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                db0 = bv_mfma(lba[0][kk][g], lba[0][kk][g], db0);
                db1 = bv_mfma(lba[0][kk][g], lba[1][kk][g], db1);
                db2 = bv_mfma(lba[1][kk][g], lba[1][kk][g], db2);
            }
        }
        *(bv_d4 *)(sW[1] + 0 * 256 + 4 * lane) = -db0;
        *(bv_d4 *)(sW[1] + 1 * 256 + 4 * lane) = -db1;
        *(bv_d4 *)(sW[1] + 2 * 256 + 4 * lane) = -db2;
        bv_wave_sync();
        bv_diag_factor_tiles(sW[1], sW[1] + 256, sW[1] + 512, lane, (qn + 2) * BV_TILE, m, cs0,
                             sCol[1], sRow[1], sT[1], sW[1], sPiv[1], &s_dotd[1], &s_info[1]);
        bv_panel_logdet(lane, (qn + 2) * BV_TILE, m, cs0, sPiv[1], &s_logdet[1]);
        bv_wave_sync();
        *(bv_d4 *)(ws + bv_tile_off(qn + 2, qn + 2, T) + 4 * lane) =
            *(const bv_d4 *)(sW[1] + 4 * lane);
        *(bv_d4 *)(ws + bv_tile_off(qn + 3, qn + 2, T) + 4 * lane) =
            *(const bv_d4 *)(sW[1] + 256 + 4 * lane);
        *(bv_d4 *)(ws + bv_tile_off(qn + 3, qn + 3, T) + 4 * lane) =
            *(const bv_d4 *)(sW[1] + 512 + 4 * lane);
        if (lane == 0) {
            const int64_t o = (int64_t)it.desc * pmax + pa;
            part_logdet[o] = s_logdet[0];
            part_dot[o] = s_dotd[0] + dba;
            part_info[o] = s_info[0];
            part_logdet[o + 1] = s_logdet[1];
            part_dot[o + 1] = s_dotd[1];
            part_info[o + 1] = s_info[1];
        }
    } else if (w == 0 && !hasB && lane == 0) {
        const int64_t o = (int64_t)it.desc * pmax + pa;
        part_logdet[o] = s_logdet[0];
        part_dot[o] = s_dotd[0];
        part_info[o] = s_info[0];
    }
}

/* Per-block totals of the stepped factorization's partials, fixed order. */
__global__ void __launch_bounds__(256)
bv_step_parts_kernel(int ndesc, const bv_block_desc *__restrict__ desc,
                     const bv_eval_slot *__restrict__ slots, const double *part_logdet,
                     const double *part_dot, const int32_t *part_info, int pmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndesc) return;
    const bv_block_desc d = desc[i];
    const bv_eval_slot &E = slots[d.eval];
    const int P = d.mp / BV_NB;
    double ld = 0.0, dt = 0.0;
    int info = 0;
    for (int j = 0; j < P; ++j) {
        const int64_t o = (int64_t)i * pmax + j;
        ld += part_logdet[o];
        dt += part_dot[o];
        if (info == 0 && part_info[o] != 0) info = part_info[o];
    }
    E.logdet[d.block] = ld;
    E.dot[d.block] = dt;
    E.info[d.block] = info;
}

/* Writes the augmented matrix A of every block (lower tiles, accumulator
   layout) into the workspace, for the kinds whose entries are too costly to
   recompute inside the factorization (general nu, space-time).  grid.x: tile
   index within the block, grid.y: descriptor. */
#ifndef BV_GEN_WAVES_PER_SIMD
#define BV_GEN_WAVES_PER_SIMD 1
#endif
template <int KIND>
__global__ void __launch_bounds__(256, BV_GEN_WAVES_PER_SIMD)
bv_gen_kernel(const bv_block_desc *__restrict__ desc, const bv_eval_slot *__restrict__ slots,
              double *ws_base) {
    const bv_block_desc d = desc[blockIdx.y];
    const int T = d.mp / BV_TILE;
    const int ntile = T * (T + 1) / 2;
    int tix = blockIdx.x;
    if (tix >= ntile) return; /* whole workgroup */
    const bv_eval_slot &E = slots[d.eval];
    constexpr bool TAB = (KIND == BV_KIND_GENERAL || KIND == BV_KIND_SPACETIME);
    __shared__ bv_knu_tab s_tab;
    if (TAB) {
        bv_knu_tab_fill(E.kp.nu, &s_tab, (int)threadIdx.x, 256);
        __syncthreads();
    }
    int ct = 0;
    while (tix >= T - ct) { tix -= T - ct; ++ct; }
    const int rt = ct + tix;
    const int e = threadIdx.x, lane = e >> 2, g = e & 3;
    const int p = rt * BV_TILE + (lane & 15);
    const int q = ct * BV_TILE + 4 * g + (lane >> 4);
    double xp, yp, zp, op, xq, yq, zq, oq;
    bv_point(d, E, p, xp, yp, zp, op);
    bv_point(d, E, q, xq, yq, zq, oq);
    double v;
    if (TAB && p < d.m && q < d.m) v = bv_cov_tab<KIND>(xp, yp, zp, xq, yq, zq, E.kp, s_tab);
    else v = bv_aug_entry<KIND>(p, q, d.m, xp, yp, zp, xq, yq, zq, oq, E.kp);
    if (p == q && p < d.m) v += E.kp.nugget; /* simulation only (0 otherwise: exact) */
    ws_base[d.ws_off + bv_tile_off(rt, ct, T) + e] = v;
}

/* Block Vecchia simulation, after the factorization of a level's blocks in
   sample mode (the workspace holds the complete lower factor L of each
   block's augmented matrix, whose row m carries the simulated values y_N of
   the conditioning set): y_B = L_BN w + L_S z_B with w = L_NN^-1 y_N (row m's
   first cs0 entries) and L_S the block rows' own factor.  One workgroup per
   block: v = [w; z_B] staged in LDS, then one thread per block row b sums
   L(cs0 + b, k) v_k over k <= cs0 + b in order.  Writes y_B into the slot's
   observation array (the simulation's output, read as y_N by later
   levels). */
__device__ __forceinline__ double bv_factor_at(const double *ws, int T, int p, int q) {
    const int rt = p >> 4, ct = q >> 4, r = p & 15, c = q & 15;
    return ws[bv_tile_off(rt, ct, T) + 4 * (r + 16 * (c & 3)) + (c >> 2)];
}

extern "C" __global__ void __launch_bounds__(256)
bv_sim_finish_kernel(const bv_block_desc *__restrict__ desc, const bv_eval_slot *__restrict__ slots,
                     const double *__restrict__ ws_base, const double *__restrict__ z) {
    __shared__ double sv[BV_MAX_MP];
    const bv_block_desc d = desc[blockIdx.x];
    const bv_eval_slot &E = slots[d.eval];
    const double *ws = ws_base + d.ws_off;
    const int T = d.mp / BV_TILE, m = d.m, cs0 = d.cs0;
    for (int k = threadIdx.x; k < m; k += blockDim.x)
        sv[k] = (k < cs0) ? bv_factor_at(ws, T, m, k) : z[d.obs_off + (k - cs0)];
    __syncthreads();
    double *y = (double *)E.obs;
    for (int b = threadIdx.x; b < m - cs0; b += blockDim.x) {
        const int p = cs0 + b;
        double acc = 0.0;
        for (int k = 0; k <= p; ++k) acc = fma(bv_factor_at(ws, T, p, k), sv[k], acc);
        y[d.obs_off + b] = acc;
    }
}

hipError_t bv_launch_sim_finish(unsigned nblk, hipStream_t st, const bv_block_desc *desc,
                                const bv_eval_slot *slots, const double *ws, const double *z) {
    if (nblk == 0) return hipSuccess;
    hipLaunchKernelGGL(bv_sim_finish_kernel, dim3(nblk), dim3(256), 0, st, desc, slots, ws, z);
    return hipGetLastError();
}

/* The general-nu table of every evaluation slot of a call (bv_xnu_knu_cheb):
   one workgroup per (half-octave, slot); slots of other kinds are left
   alone.  The node values (bv_cheb_node_g) come from the accurate routine,
   the Chebyshev coefficients from their cosine sums, the monomial ones from
   those (bv_cheb_mono).  Written through kp.cheb, the slot's own table. */
extern "C" __global__ void __launch_bounds__(64)
bv_cheb_build_kernel(const bv_eval_slot *__restrict__ slots) {
    const bv_eval_slot &E = slots[blockIdx.y];
    if (E.kind != BV_KIND_GENERAL) return;
    __shared__ double gv[BV_CHEB_N], cv[BV_CHEB_N];
    const int j = blockIdx.x, t = threadIdx.x;
    if (t < BV_CHEB_N) gv[t] = bv_cheb_node_g(E.kp.nu, j, t);
    __syncthreads();
    if (t < BV_CHEB_N) cv[t] = bv_cheb_coef(gv, t);
    __syncthreads();
    if (t == 0) {
        double a[BV_CHEB_N];
#pragma unroll
        for (int i = 0; i < BV_CHEB_N; ++i) a[i] = cv[i];
        bv_cheb_mono(a);
#pragma unroll
        for (int i = 0; i < BV_CHEB_N; ++i) ((double *)E.kp.cheb)[j * BV_CHEB_N + i] = a[i];
    }
}

hipError_t bv_launch_cheb_build(unsigned nslots, hipStream_t st, const bv_eval_slot *slots) {
    if (nslots == 0) return hipSuccess;
    hipLaunchKernelGGL(bv_cheb_build_kernel, dim3(BV_CHEB_NJ, nslots), dim3(64), 0, st, slots);
    return hipGetLastError();
}

/* findNearestPoints (include/nearest_neighbor.h:76-148) on the device: for
   block b >= 1 (one workgroup each), the cs points among positions
   0..accum[b]-1 of the block-ordered points nearest to the block's centroid,
   in the order the reference's selection sort leaves them.  Its step s takes
   the first minimum (strict <) of the distances at positions s..N-1 and
   swaps it to position s, so the result is: repeatedly the (distance,
   current position)-smallest remaining point, where a swap moves the point
   that sat at position s to the taken point's old position.  Only points
   with distance <= the cs-th smallest can be taken; the kernel finds that
   distance by a radix select over the distances' IEEE bits (non-negative, so
   ordered as integers), gathers those candidates in position order, and
   replays the sort's swaps on them.  Distances are the host's expression,
   products and sums rounded separately, with the correctly rounded square
   root: sqrt(dx dx + dy dy), and for 3-D sqrt(d d + dz dz).  A block with
   more than BV_KNN_CAND tied candidates sets its flag; the host redoes it. */
#define BV_KNN_CAND 2048
#define BV_KNN_BITS 11

__device__ __forceinline__ double bv_knn_dist(const double *__restrict__ x,
                                              const double *__restrict__ y,
                                              const double *__restrict__ z, int64_t i, double cx,
                                              double cy, double cz, int time_flag) {
#pragma clang fp contract(off)
    const double dx = x[i] - cx, dy = y[i] - cy;
    double d = sqrt(dx * dx + dy * dy);
    if (time_flag) {
        const double dz = cz - z[i];
        d = sqrt(d * d + dz * dz);
    }
    return d;
}

extern "C" __global__ void __launch_bounds__(256)
bv_knn_kernel(const double *__restrict__ x, const double *__restrict__ y,
              const double *__restrict__ z, const int64_t *__restrict__ accum,
              const double *__restrict__ bcx, const double *__restrict__ bcy,
              const double *__restrict__ bcz, int k, int time_flag, int32_t *__restrict__ nbr,
              int32_t *__restrict__ overflow) {
    __shared__ uint32_t hist[4][1 << BV_KNN_BITS];
    __shared__ double cdist[BV_KNN_CAND];
    __shared__ int32_t cpos[BV_KNN_CAND];
    __shared__ int32_t corig[BV_KNN_CAND];
    __shared__ uint64_t s_prefix, s_mask;
    __shared__ int32_t s_rem, s_ncand, s_wcount[4];
    const int b = blockIdx.x + 1;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t N = accum[b];
    const double cx = bcx[b], cy = bcy[b], cz = time_flag ? bcz[b] : 0.0;
    if (N < k) { /* the reference's "not enough points": never after firstbatch */
        if (tid == 0) overflow[b] = 1;
        return;
    }
    if (tid == 0) { s_prefix = 0; s_mask = 0; s_rem = k; }
    __syncthreads();
    /* radix select of the k-th smallest key, BV_KNN_BITS bits per pass from
       the top (the last pass takes the remaining low bits) */
    for (int hi = 64; hi > 0; hi -= BV_KNN_BITS) {
        const int width = hi >= BV_KNN_BITS ? BV_KNN_BITS : hi;
        const int shift = hi - width;
        const uint64_t dmask = ((uint64_t)1 << width) - 1;
        for (int t = tid; t < 4 * (1 << BV_KNN_BITS); t += 256) (&hist[0][0])[t] = 0;
        __syncthreads();
        const uint64_t prefix = s_prefix, mask = s_mask;
        for (int64_t i = tid; i < N; i += 256) {
            const uint64_t key =
                __builtin_bit_cast(uint64_t, bv_knn_dist(x, y, z, i, cx, cy, cz, time_flag));
            if ((key & mask) == prefix) atomicAdd(&hist[wave][(key >> shift) & dmask], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            int rem = s_rem;
            int digit = 0;
            for (int dgt = 0; dgt <= (int)dmask; ++dgt) {
                const int cnt =
                    (int)(hist[0][dgt] + hist[1][dgt] + hist[2][dgt] + hist[3][dgt]);
                if (rem <= cnt) { digit = dgt; break; }
                rem -= cnt;
            }
            s_rem = rem;
            s_prefix = prefix | ((uint64_t)digit << shift);
            s_mask = mask | (dmask << shift);
        }
        __syncthreads();
    }
    const uint64_t kth = s_prefix; /* bits of the k-th smallest distance */
    /* candidates (key <= kth) in position order */
    if (tid == 0) s_ncand = 0;
    __syncthreads();
    for (int64_t base = 0; base < N; base += 256) {
        const int64_t i = base + tid;
        double d = 0.0;
        bool c = false;
        if (i < N) {
            d = bv_knn_dist(x, y, z, i, cx, cy, cz, time_flag);
            c = __builtin_bit_cast(uint64_t, d) <= kth;
        }
        const uint64_t bal = __ballot(c);
        if (lane == 0) s_wcount[wave] = __popcll(bal);
        __syncthreads();
        int off = s_ncand;
        for (int w = 0; w < wave; ++w) off += s_wcount[w];
        off += __popcll(bal & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
        if (c && off < BV_KNN_CAND) {
            cdist[off] = d;
            cpos[off] = (int32_t)i;
            corig[off] = (int32_t)i;
        }
        __syncthreads();
        if (tid == 0) s_ncand += s_wcount[0] + s_wcount[1] + s_wcount[2] + s_wcount[3];
        __syncthreads();
    }
    const int nc = s_ncand;
    if (nc > BV_KNN_CAND) {
        if (tid == 0) overflow[b] = 1;
        return;
    }
    if (wave != 0) return;
    /* the selection sort's steps, replayed on the candidates by one wavefront;
       a taken candidate's position becomes < every later step, so it is
       marked by cpos = -1 - s instead */
    int32_t *out = nbr + (int64_t)b * k;
    for (int st = 0; st < k; ++st) {
        double bd = 0.0;
        int bp = INT32_MAX, bc = -1;
        for (int c = lane; c < nc; c += 64) {
            const int p = cpos[c];
            if (p < 0) continue;
            const double d = cdist[c];
            if (bc < 0 || d < bd || (d == bd && p < bp)) { bd = d; bp = p; bc = c; }
        }
        /* wavefront argmin of (distance, position) */
        for (int o = 32; o >= 1; o >>= 1) {
            const double od = __shfl_xor(bd, o, 64);
            const int op = __shfl_xor(bp, o, 64), oc = __shfl_xor(bc, o, 64);
            if (oc >= 0 && (bc < 0 || od < bd || (od == bd && op < bp))) {
                bd = od; bp = op; bc = oc;
            }
        }
        const int from = bp;
        /* the candidate sitting at position st moves to `from` */
        for (int c = lane; c < nc; c += 64)
            if (c != bc && cpos[c] == st) cpos[c] = from;
        bv_wave_sync();
        if (lane == 0) {
            cpos[bc] = -1 - st;
            out[st] = corig[bc];
        }
        bv_wave_sync();
    }
}

hipError_t bv_launch_knn(const double *x, const double *y, const double *z, const int64_t *accum,
                         const double *bcx, const double *bcy, const double *bcz, int64_t B,
                         int k, int time_flag, int32_t *nbr, int32_t *overflow,
                         hipStream_t st) {
    if (B <= 1) return hipSuccess;
    hipLaunchKernelGGL(bv_knn_kernel, dim3((unsigned)(B - 1)), dim3(256), 0, st, x, y, z, accum,
                       bcx, bcy, bcz, k, time_flag, nbr, overflow);
    return hipGetLastError();
}

/* The k-means assignment step of the host pre-processing (bv_prep.cpp) on
   the device: one thread per point, the centroids staged through LDS in
   tiles, the host's rounding (d = ((0 + dx*dx) + dy*dy) + dz*dz with
   dx = point - centroid, no contraction) and its tie rule (the first minimum
   wins), so the assignments are the host loop's. */
#define BV_KM_TILE 1024
extern "C" __global__ void __launch_bounds__(256)
bv_kmeans_assign_kernel(int64_t n, const double *__restrict__ px, const double *__restrict__ py,
                        const double *__restrict__ pz, int k, const double *__restrict__ cx,
                        const double *__restrict__ cy, const double *__restrict__ cz,
                        int32_t *__restrict__ cluster) {
#pragma clang fp contract(off)
    __shared__ double sx[BV_KM_TILE], sy[BV_KM_TILE], sz[BV_KM_TILE];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool live = i < n;
    const double ax = live ? px[i] : 0.0, ay = live ? py[i] : 0.0, az = live ? pz[i] : 0.0;
    double best = 0.0;
    int bi = 0;
    for (int base = 0; base < k; base += BV_KM_TILE) {
        const int cnt = min(BV_KM_TILE, k - base);
        __syncthreads();
        for (int t = threadIdx.x; t < cnt; t += 256) {
            sx[t] = cx[base + t];
            sy[t] = cy[base + t];
            sz[t] = cz[base + t];
        }
        __syncthreads();
        int c0 = 0;
        if (base == 0) {
            const double dx = ax - sx[0], dy = ay - sy[0], dz = az - sz[0];
            double d = 0.0;
            d += dx * dx;
            d += dy * dy;
            d += dz * dz;
            best = d;
            c0 = 1;
        }
#pragma unroll 8
        for (int c = c0; c < cnt; ++c) {
            const double dx = ax - sx[c], dy = ay - sy[c], dz = az - sz[c];
            double d = 0.0;
            d += dx * dx;
            d += dy * dy;
            d += dz * dz;
            if (d < best) {
                best = d;
                bi = base + c;
            }
        }
    }
    if (live) cluster[i] = bi;
}

/* One workgroup per evaluation: llh and first failing block, fixed order. */
extern "C" __global__ void __launch_bounds__(BV_THREADS)
bv_reduce_kernel(const bv_eval_slot *__restrict__ slots) {
    const bv_eval_slot &E = slots[blockIdx.x];
    __shared__ double s_part[BV_THREADS];
    __shared__ int s_bad[BV_THREADS];
    const int tid = threadIdx.x;
    const int B = E.nblocks;
    /* thread t: blocks t, t + 256, ... (coalesced), then a fixed pairwise
       tree: the order of the sum depends only on B */
    double acc = 0.0;
    int bad = INT32_MAX;
    for (int k = tid; k < B; k += BV_THREADS) {
        acc += -(E.dot[k] + E.logdet[k] + (double)E.batch_num[k] * E.log2pi) * 0.5;
        if (bad == INT32_MAX && E.info[k] != 0) bad = k;
    }
    s_part[tid] = acc;
    s_bad[tid] = bad;
    __syncthreads();
#pragma unroll
    for (int off = BV_THREADS / 2; off >= 1; off >>= 1) {
        if (tid < off) {
            s_part[tid] += s_part[tid + off];
            s_bad[tid] = s_bad[tid] < s_bad[tid + off] ? s_bad[tid] : s_bad[tid + off];
        }
        __syncthreads();
    }
    if (tid == 0) {
        E.result[0] = s_part[0];
        E.result[1] = s_bad[0] == INT32_MAX ? -1.0 : (double)s_bad[0];
    }
}

/* Dense covariance block A[i + j*m] = K(l1[i], l2[j]) (core_dcmg layout). */
extern "C" __global__ void __launch_bounds__(256)
bv_matern_kernel(int m, int n, const double *x1, const double *y1, const double *z1,
                 const double *x2, const double *y2, const double *z2, bv_kparams kp, int kind,
                 double *A) {
    __shared__ bv_knu_tab s_tab;
    if (kind == BV_KIND_SPACETIME) {
        bv_knu_tab_fill(kp.nu, &s_tab, (int)threadIdx.x, (int)blockDim.x);
        __syncthreads();
    }
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
        v = bv_cov_tab<BV_KIND_SPACETIME>(x1[i], y1[i], za, x2[jj], y2[jj], zb, kp, s_tab);
        break;
    default:
        v = bv_cov<BV_KIND_GENERAL>(x1[i], y1[i], za, x2[jj], y2[jj], zb, kp);
        break;
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

/* FP64 pipe probe: each wavefront repeats `iters` times a body of 16 steps,
   each step NM independent 16x16x4 FP64 matrix-core products (two
   accumulators) and NV independent FP64 fused multiply-adds (eight
   registers), and stamps its cycle count.  With `split`, the even
   wavefronts of a workgroup run only the products and the odd ones only the
   multiply-adds (two wavefronts per SIMD), so the two kinds come from
   different wavefronts.  Results are per step. */
template <int NM, int NV>
__device__ __forceinline__ void bv_pipe_body(bv_d4 &a0, bv_d4 &a1, double (&v)[8], double x,
                                             double y, int iters) {
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
#pragma unroll
            for (int t = 0; t < NM; t += 2) {
                a0 = bv_mfma(x, y, a0);
                a1 = bv_mfma(y, x, a1);
            }
#pragma unroll
            for (int r = 0; r < NV; ++r) v[r & 7] = __builtin_fma(v[r & 7], y, x);
        }
    }
}

template <int NM, int NV>
__global__ void __launch_bounds__(512)
bv_pipe_probe_kernel(double *out, unsigned long long *cyc, int iters, int split) {
    const int wave = threadIdx.x >> 6;
    bv_d4 a0 = (bv_d4){0.0, 0.0, 0.0, 0.0}, a1 = a0;
    const double x = threadIdx.x * 1e-3, y = 1.0 + threadIdx.x * 1e-4;
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = x + j;
    const unsigned long long t0 = __builtin_readcyclecounter();
    if (!split) bv_pipe_body<NM, NV>(a0, a1, v, x, y, iters);
    else if (!(wave & 1)) bv_pipe_body<NM, 0>(a0, a1, v, x, y, iters);
    else bv_pipe_body<0, NV>(a0, a1, v, x, y, iters);
    const unsigned long long t1 = __builtin_readcyclecounter();
    double t = a0[0] + a1[1];
#pragma unroll
    for (int j = 0; j < 8; ++j) t += v[j];
    out[blockIdx.x * blockDim.x + threadIdx.x] = t;
    if ((threadIdx.x & 63) == 0) cyc[blockIdx.x * 8 + wave] = t1 - t0;
}

/* mode 0: 2 products; 1: 8 multiply-adds; 2: both; 3: 2 products and 32
   multiply-adds; 4: mode 2 split over two wavefronts per SIMD.  One
   workgroup per compute unit (ncu of them), 256 threads (512 split). */
hipError_t bv_launch_pipe_probe(int mode, int ncu, int iters, double *out,
                                unsigned long long *cyc) {
    const dim3 g(ncu), b(mode == 4 ? 512 : 256);
    switch (mode) {
    case 0: hipLaunchKernelGGL((bv_pipe_probe_kernel<2, 0>), g, b, 0, 0, out, cyc, iters, 0); break;
    case 1: hipLaunchKernelGGL((bv_pipe_probe_kernel<0, 8>), g, b, 0, 0, out, cyc, iters, 0); break;
    case 2: hipLaunchKernelGGL((bv_pipe_probe_kernel<2, 8>), g, b, 0, 0, out, cyc, iters, 0); break;
    case 3: hipLaunchKernelGGL((bv_pipe_probe_kernel<2, 32>), g, b, 0, 0, out, cyc, iters, 0); break;
    default: hipLaunchKernelGGL((bv_pipe_probe_kernel<2, 8>), g, b, 0, 0, out, cyc, iters, 1); break;
    }
    return hipGetLastError();
}

/* ---- host-side launch wrappers (called from bv_api.hip) ---- */
hipError_t bv_launch_block_1w(int kind, unsigned nblk, size_t lds, hipStream_t st,
                              const bv_block_desc *desc, const bv_eval_slot *slots, double *ws) {
    switch (kind) {
    case BV_KIND_NU05:
        hipLaunchKernelGGL(bv_block_kernel_1w<BV_KIND_NU05>, dim3(nblk), dim3(64), lds, st, desc,
                           slots, ws);
        break;
    case BV_KIND_NU15:
        hipLaunchKernelGGL(bv_block_kernel_1w<BV_KIND_NU15>, dim3(nblk), dim3(64), lds, st, desc,
                           slots, ws);
        break;
    case BV_KIND_NU25:
        hipLaunchKernelGGL(bv_block_kernel_1w<BV_KIND_NU25>, dim3(nblk), dim3(64), lds, st, desc,
                           slots, ws);
        break;
    case BV_KIND_GENERAL:
        hipLaunchKernelGGL(bv_block_kernel_1w<BV_KIND_GENERAL>, dim3(nblk), dim3(64), lds, st,
                           desc, slots, ws);
        break;
    default:
        hipLaunchKernelGGL(bv_block_kernel_1w<BV_KIND_LOADA>, dim3(nblk), dim3(64), 0, st, desc,
                           slots, ws);
        break;
    }
    return hipGetLastError();
}

template <int NW>
static hipError_t bv_launch_block_nw(int kind, unsigned nblk, size_t lds, hipStream_t st,
                                     const bv_block_desc *desc, const bv_eval_slot *slots,
                                     double *ws) {
    const dim3 blk(64 * NW);
    switch (kind) {
    case BV_KIND_NU05:
        hipLaunchKernelGGL((bv_block_kernel<BV_KIND_NU05, NW>), dim3(nblk), blk, lds, st, desc,
                           slots, ws);
        break;
    case BV_KIND_NU15:
        hipLaunchKernelGGL((bv_block_kernel<BV_KIND_NU15, NW>), dim3(nblk), blk, lds, st, desc,
                           slots, ws);
        break;
    case BV_KIND_NU25:
        hipLaunchKernelGGL((bv_block_kernel<BV_KIND_NU25, NW>), dim3(nblk), blk, lds, st, desc,
                           slots, ws);
        break;
    case BV_KIND_GENERAL:
        hipLaunchKernelGGL((bv_block_kernel<BV_KIND_GENERAL, NW>), dim3(nblk), blk, lds, st, desc,
                           slots, ws);
        break;
    default:
        hipLaunchKernelGGL((bv_block_kernel<BV_KIND_LOADA, NW>), dim3(nblk), blk, 0, st, desc,
                           slots, ws);
        break;
    }
    return hipGetLastError();
}

hipError_t bv_launch_block(int kind, unsigned nblk, size_t lds, hipStream_t st,
                           const bv_block_desc *desc, const bv_eval_slot *slots, double *ws) {
    return bv_launch_block_nw<4>(kind, nblk, lds, st, desc, slots, ws);
}

hipError_t bv_launch_block_2w(int kind, unsigned nblk, size_t lds, hipStream_t st,
                              const bv_block_desc *desc, const bv_eval_slot *slots, double *ws) {
    return bv_launch_block_nw<2>(kind, nblk, lds, st, desc, slots, ws);
}

hipError_t bv_launch_gen(int kind, unsigned max_tiles, unsigned nblk, hipStream_t st,
                         const bv_block_desc *desc, const bv_eval_slot *slots, double *ws) {
    /* grid.y (descriptors) is limited to 65535: launch in slices */
    for (unsigned off = 0; off < nblk; off += 65535u) {
        const unsigned cnt = (nblk - off) < 65535u ? (nblk - off) : 65535u;
        if (kind == BV_KIND_SPACETIME)
            hipLaunchKernelGGL(bv_gen_kernel<BV_KIND_SPACETIME>, dim3(max_tiles, cnt), dim3(256),
                               0, st, desc + off, slots, ws);
        else
            hipLaunchKernelGGL(bv_gen_kernel<BV_KIND_GENERAL>, dim3(max_tiles, cnt), dim3(256),
                               0, st, desc + off, slots, ws);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t bv_launch_step(int kind, unsigned nitems, hipStream_t st,
                          const bv_step_item *items, const bv_block_desc *desc,
                          const bv_eval_slot *slots, double *ws, double *part_logdet,
                          double *part_dot, int32_t *part_info, int pmax) {
    if (nitems == 0) return hipSuccess;
    switch (kind) {
    case BV_KIND_NU05:
        hipLaunchKernelGGL(bv_step_kernel<BV_KIND_NU05>, dim3(nitems), dim3(64), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    case BV_KIND_NU15:
        hipLaunchKernelGGL(bv_step_kernel<BV_KIND_NU15>, dim3(nitems), dim3(64), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    case BV_KIND_NU25:
        hipLaunchKernelGGL(bv_step_kernel<BV_KIND_NU25>, dim3(nitems), dim3(64), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    case BV_KIND_GENERAL:
        hipLaunchKernelGGL(bv_step_kernel<BV_KIND_GENERAL>, dim3(nitems), dim3(64), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    default:
        hipLaunchKernelGGL(bv_step_kernel<BV_KIND_LOADA>, dim3(nitems), dim3(64), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    }
    return hipGetLastError();
}

hipError_t bv_launch_step2(int kind, unsigned nitems, hipStream_t st,
                          const bv_step_item *items, const bv_block_desc *desc,
                          const bv_eval_slot *slots, double *ws, double *part_logdet,
                          double *part_dot, int32_t *part_info, int pmax) {
    if (nitems == 0) return hipSuccess;
    switch (kind) {
    case BV_KIND_NU05:
        hipLaunchKernelGGL(bv_step2_kernel<BV_KIND_NU05>, dim3(nitems), dim3(64), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    case BV_KIND_NU15:
        hipLaunchKernelGGL(bv_step2_kernel<BV_KIND_NU15>, dim3(nitems), dim3(64), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    case BV_KIND_NU25:
        hipLaunchKernelGGL(bv_step2_kernel<BV_KIND_NU25>, dim3(nitems), dim3(64), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    case BV_KIND_GENERAL:
        hipLaunchKernelGGL(bv_step2_kernel<BV_KIND_GENERAL>, dim3(nitems), dim3(64), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    default:
        hipLaunchKernelGGL(bv_step2_kernel<BV_KIND_LOADA>, dim3(nitems), dim3(64), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    }
    return hipGetLastError();
}

hipError_t bv_launch_pair(int kind, unsigned nitems, hipStream_t st,
                          const bv_step_item *items, const bv_block_desc *desc,
                          const bv_eval_slot *slots, double *ws, double *part_logdet,
                          double *part_dot, int32_t *part_info, int pmax) {
    if (nitems == 0) return hipSuccess;
    switch (kind) {
    case BV_KIND_NU05:
        hipLaunchKernelGGL(bv_pair_kernel<BV_KIND_NU05>, dim3(nitems), dim3(128), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    case BV_KIND_NU15:
        hipLaunchKernelGGL(bv_pair_kernel<BV_KIND_NU15>, dim3(nitems), dim3(128), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    case BV_KIND_NU25:
        hipLaunchKernelGGL(bv_pair_kernel<BV_KIND_NU25>, dim3(nitems), dim3(128), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    case BV_KIND_GENERAL:
        hipLaunchKernelGGL(bv_pair_kernel<BV_KIND_GENERAL>, dim3(nitems), dim3(128), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    default:
        hipLaunchKernelGGL(bv_pair_kernel<BV_KIND_LOADA>, dim3(nitems), dim3(128), 0, st, items,
                           (int)nitems, desc, slots, ws, part_logdet, part_dot, part_info, pmax);
        break;
    }
    return hipGetLastError();
}

hipError_t bv_launch_step_parts(int ndesc, hipStream_t st, const bv_block_desc *desc,
                                const bv_eval_slot *slots, const double *part_logdet,
                                const double *part_dot, const int32_t *part_info, int pmax) {
    hipLaunchKernelGGL(bv_step_parts_kernel, dim3((ndesc + 255) / 256), dim3(256), 0, st, ndesc,
                       desc, slots, part_logdet, part_dot, part_info, pmax);
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

hipError_t bv_launch_kmeans_assign(int64_t n, const double *px, const double *py, const double *pz,
                                   int k, const double *cx, const double *cy, const double *cz,
                                   int32_t *cluster, hipStream_t st) {
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(bv_kmeans_assign_kernel, dim3(grid), dim3(256), 0, st, n, px, py, pz, k,
                       cx, cy, cz, cluster);
    return hipGetLastError();
}

/* Wavefront trace readers (bv.h). */
#include "bv.h"
#ifdef BV_TRACE
extern "C" int bv_trace_reset(int device) {
    if (hipSetDevice(device) != hipSuccess) return BV_ERR_HIP;
    const unsigned int zero = 0;
    return hipMemcpyToSymbol(HIP_SYMBOL(bv_trace_n), &zero, sizeof(zero)) == hipSuccess ? BV_OK
                                                                                       : BV_ERR_HIP;
}

extern "C" int bv_trace_read(int device, uint64_t *records, int64_t capacity, int64_t *count) {
    if (!records || !count || capacity < 0) return BV_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return BV_ERR_HIP;
    unsigned int n = 0;
    if (hipMemcpyFromSymbol(&n, HIP_SYMBOL(bv_trace_n), sizeof(n)) != hipSuccess) return BV_ERR_HIP;
    int64_t k = n < BV_TRACE_CAP ? (int64_t)n : (int64_t)BV_TRACE_CAP;
    if (k > capacity) k = capacity;
    if (k > 0 && hipMemcpyFromSymbol(records, HIP_SYMBOL(bv_trace_rec),
                                     (size_t)k * 3 * sizeof(uint64_t)) != hipSuccess)
        return BV_ERR_HIP;
    *count = n;
    return BV_OK;
}
#else
extern "C" int bv_trace_reset(int device) {
    (void)device;
    return BV_ERR_UNSUPPORTED;
}

extern "C" int bv_trace_read(int device, uint64_t *records, int64_t capacity, int64_t *count) {
    (void)device; (void)records; (void)capacity; (void)count;
    return BV_ERR_UNSUPPORTED;
}
#endif
