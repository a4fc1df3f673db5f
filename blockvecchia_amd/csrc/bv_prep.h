/*
 * bv_prep.h -- internal interface between the host pre-processing
 * (bv_prep.cpp) and the device k-means assignment step (bv_api.hip).
 */
#ifndef BV_PREP_H
#define BV_PREP_H

#include <stdint.h>

#include "bv.h"

/* One k-means assignment step: for every point i (the points bv_partition_run
 * was given, z = 0 in 2-D), cluster[i] = the first c minimising
 * ((0 + dx^2) + dy^2) + dz^2 over the k centroids (cx, cy, cz), products and
 * sums rounded separately.  Returns BV_OK or an error code. */
typedef int (*bv_assign_fn)(void *user, int k, const double *cx, const double *cy,
                            const double *cz, int32_t *cluster);

/* The conditioning sets of blocks 1..B-1 (findNearestPoints, Euclidean
 * distance): for block b, nbr[b*cs .. b*cs+cs) = the positions among
 * 0..accum[b]-1 of the block-ordered points (x, y, z) nearest to the block's
 * centroid (cx[b], cy[b], cz[b]) in the reference's selection-sort order.
 * Blocks it could not finish are flagged in redo[b] (1) and left to the host.
 * Returns BV_OK or an error code. */
typedef int (*bv_knn_fn)(void *user, int64_t n, const double *x, const double *y,
                         const double *z, int64_t B, const int64_t *accum, const double *cx,
                         const double *cy, const double *cz, int32_t cs, int32_t time_flag,
                         int32_t *nbr, int32_t *redo);

/* bv_partition with the assignment step supplied by `assign` and the
 * Euclidean conditioning sets by `knn` (nullptr: the host loops).  The result
 * does not depend on which one runs. */
int bv_partition_run(int64_t n, const double *x, const double *y, const double *z,
                     const double *obs, int32_t bc, int32_t cs, int32_t seed, int32_t epochs,
                     int32_t time_flag, int32_t metric, int32_t threads, bv_assign_fn assign,
                     void *user, bv_blocks *out, bv_knn_fn knn = nullptr);

#endif /* BV_PREP_H */
