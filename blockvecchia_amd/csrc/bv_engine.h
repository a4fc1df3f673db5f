/*
 * bv_engine.h -- structures shared by the device kernels (bv_kernels.hip)
 * and the host engine (bv_api.hip).
 *
 * Per block i the engine factors the augmented matrix
 *
 *        [ S22  S21  0 ]   rows 0..cs0-1     conditioning points N_i
 *   A =  [ S12  S11  0 ]   rows cs0..m-1     block points B_i
 *        [ yN'  yB'  * ]   row  m            observations (right-hand side)
 *        [ 0    0    I ]   rows m+1..Mp-1    padding
 *
 * with one Cholesky sweep.  Its trailing factor on rows cs0..m-1 is
 * chol(S11 - S12 S22^-1 S21) and row m holds [L22^-1 yN ; Ls^-1 (yB - mu)],
 * which is exactly what the reference's chain (potrf S22, trsm, gemm, geadd,
 * potrf, trsm; include/llh_Xvecchia_batch.h:295-466) produces for the
 * log-determinant and the quadratic form.  Block 0 has cs0 = 0.
 */
#ifndef BV_ENGINE_H
#define BV_ENGINE_H

#include <stdint.h>
#include "bv_matern.h"

#define BV_NB 32          /* panel width (columns factored per left-looking step) */
#define BV_TILE 16        /* matrix-core tile edge                               */
#define BV_THREADS 256    /* 4 wavefronts of 64 lanes per block                  */
#ifndef BV_NRT
#define BV_NRT 4          /* row tiles held in accumulators per wave per pass    */
#endif
#ifndef BV_WAVES_PER_SIMD
#define BV_WAVES_PER_SIMD 2 /* occupancy target of the factorization kernel       */
#endif
#define BV_MAX_MP 1024    /* largest padded augmented size handled (LDS staging)  */
#define BV_RES_MAX_MP 384 /* register-resident kernel: padded sizes up to 16 * 24 */
#define BV_RES_NR_MAX 6   /* ... in builds of 4 NR tile rows, NR = 2 .. 6           */
#define BV_RES_TUNE_MIN_MP 129 /* ... offered by the autotuner above this size */

/* The register-resident build for a block of padded size mp: the smallest
   NR >= 2 with 64 NR >= mp. */
static inline int bv_res_nr(int mp) { return mp <= 128 ? 2 : (mp + 63) / 64; }

/* One block of one evaluation, in launch order. */
struct bv_block_desc {
    int32_t block;     /* block index within its dataset                 */
    int32_t m;         /* cs0 + bs: augmented size without the RHS row    */
    int32_t cs0;       /* conditioning rows (0 for block 0 or cs == 0)    */
    int32_t mp;        /* padded size: multiple of BV_NB, >= m + 1        */
    int32_t eval;      /* evaluation slot (index into bv_eval_slot[])     */
    int32_t pad_;
    int64_t ws_off;    /* offset in doubles of this block's factor tiles  */
    int64_t obs_off;   /* offset of the block's first point (batchNumAccum) */
};

/* One work item of a panel step of the stepped factorization: BV_STEP_ROWS
 * consecutive row tiles of one block (descriptor); desc < 0 marks padding. */
#ifndef BV_STEP_ROWS
#define BV_STEP_ROWS 2
#endif
struct bv_step_item {
    int32_t desc;      /* index into the descriptor array, or -1          */
    int32_t row0;      /* first row tile (2j for the item holding the diagonal block) */
    int32_t nrows;     /* row tiles in this item, 1..BV_STEP_ROWS (paired: 1..4) */
    int32_t step;      /* the item's panel step (paired: pair step); one launch
                          may hold items of several steps of different blocks */
};

/* Everything one evaluation needs: dataset pointers, covariance parameters
 * and where its per-block results go. */
struct bv_eval_slot {
    const double *x, *y, *z, *obs;
    const int32_t *nbr;                 /* [B*cs] or NULL               */
    const double *cx, *cy, *cz, *cobs;  /* [B*cs] copies when nbr == NULL */
    int32_t cs;
    int32_t nblocks;
    const int32_t *batch_num;           /* [B] (device)                  */
    double *logdet, *dot;               /* [B] per-block results         */
    int32_t *info;                      /* [B] 0 or first failing column + 1 */
    double *result;                     /* [4] llh, first bad block, ... */
    bv_kparams kp;
    int32_t kind;
    int32_t flags;                      /* BV_SLOT_SAMPLE: store the whole factor */
    double log2pi;                      /* log(2*PI) with PI = 3.141592653589793 */
};

/* bv_eval_slot.flags: the fused factorization also solves the diagonal
   blocks' rows (L_jj = C_jj W_jj^T), so the workspace holds the complete
   lower factor, as the simulation needs. */
#define BV_SLOT_SAMPLE 1

#endif /* BV_ENGINE_H */
