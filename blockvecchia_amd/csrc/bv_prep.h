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

/* bv_partition with the assignment step supplied by `assign` (nullptr: the
 * host loop).  The result does not depend on which one runs. */
int bv_partition_run(int64_t n, const double *x, const double *y, const double *z,
                     const double *obs, int32_t bc, int32_t cs, int32_t seed, int32_t epochs,
                     int32_t time_flag, int32_t metric, int32_t threads, bv_assign_fn assign,
                     void *user, bv_blocks *out);

#endif /* BV_PREP_H */
