/*
 * bv_device.h -- device helpers shared by the factorization kernels
 * (bv_kernels.hip, bv_resident.hip): the 16x16 FP64 matrix-core tile type and
 * product, the packed tile offset of the workspace, the augmented matrix entry,
 * the point loader and the wavefront lane exchanges.
 */
#ifndef BV_DEVICE_H
#define BV_DEVICE_H

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

/* Makes one wavefront's LDS writes visible to its other lanes: a wavefront-
   scope release/acquire around the wave barrier, so the compiler neither keeps
   stale LDS values in registers nor moves accesses across it. */
__device__ __forceinline__ void bv_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* 64-bit readlane: the value of v in lane `src` (wave-uniform). */
__device__ __forceinline__ double bv_readlane(double v, int src) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), src);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

/* Y = R J^T for 16x16 tiles in accumulator layout (matrix-core products). */
__device__ __forceinline__ bv_d4 bv_rjt(const bv_d4 &r, const bv_d4 &jm, bv_d4 acc) {
#pragma unroll
    for (int g = 0; g < 4; ++g) acc = bv_mfma(jm[g], r[g], acc);
    return acc;
}

/* The value of v in lane K of this lane's row of 16 lanes: one 64-bit DPP
   move (row_newbcast writes every lane, so the destination needs no initial
   value), no LDS round trip. */
template <int K>
__device__ __forceinline__ double bv_row_bcast(double v) {
    return __builtin_amdgcn_mov_dpp(v, 0x150 + K, 0xF, 0xF, false);
}

/* sqrt(s), s >= 0: the hardware reciprocal square root and the Newton
   corrections of the library square root, without its rescaling of inputs
   below 2^-767 (squared distances of distinct points are far above) */
__device__ __forceinline__ double bv_sqrt_nr(double s) {
    const double y = __builtin_amdgcn_rsq(s);
    double g = s * y, h = 0.5 * y;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double d = fma(-g, g, s);
    g = fma(d, h, g);
    d = fma(-g, g, s);
    g = fma(d, h, g);
    return (s == 0.0) ? 0.0 : g;
}

/* the exponential's table (bv_expneg_tab in bv_matern.h), one copy per code
   object in device memory */
static __device__ const double bv_exp_tab_dev[2 * BV_EXP_TAB] = {
    BV_EXP_TAB_HI_LIST, BV_EXP_TAB_LO_LIST};

/* Whether a kind's entries take bv_matern_planar: Euclidean distances in the
   plane, closed forms or general nu (the other kinds, 3-D and great-circle
   distances take bv_cov). */
template <int KIND>
__device__ __forceinline__ bool bv_planar_ok(const bv_kparams &kp) {
    return (KIND == BV_KIND_NU05 || KIND == BV_KIND_NU15 || KIND == BV_KIND_NU25 ||
            KIND == BV_KIND_GENERAL) && kp.metric == 0 && !kp.use_z;
}

/* Matern entry between two points of the plane: the distance as
   bv_dist_euclid rounds it, the square root and the exponential above
   (within one ulp each); the forms of bv_matern_closed (nu = 1/2, 3/2, 5/2)
   or, for general nu, bv_matern_general_x (cheb: the evaluation's Chebyshev
   table of x^nu e^x K_nu(x), kp.cheb or a copy in LDS; etab: the
   exponential's table, bv_exp_tab_dev or a copy in LDS). */
template <int KIND>
__device__ __forceinline__ double bv_matern_planar(double xa, double ya, double xb, double yb,
                                                   const bv_kparams &kp, const double *cheb,
                                                   const double *etab) {
#pragma clang fp contract(off)
    const double dx = xb - xa, dy = yb - ya;
    const double d = bv_sqrt_nr(dx * dx + dy * dy);
    const double t = d * kp.inv_scale;
    if (KIND == BV_KIND_GENERAL) return bv_matern_general_x(t, kp, cheb, etab);
    const double x = bv_expneg_tab(t, etab);
    if (KIND == BV_KIND_NU05) return kp.sigma2 * x;
    if (KIND == BV_KIND_NU15) return kp.sigma2 * (1.0 + t) * x;
    return kp.sigma2 * (1.0 + t + t * t * (1.0 / 3.0)) * x;
}

#endif /* BV_DEVICE_H */
