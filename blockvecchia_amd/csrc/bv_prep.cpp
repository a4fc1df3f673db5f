/*
 * bv_prep.cpp -- host pre-processing of the Block Vecchia method: partition the
 * points into blocks, order the blocks, choose each block's conditioning set.
 *
 * Follows src/test_Xvecchia_batch.cpp:240-404:
 *   - classic Vecchia (one point per cluster) when n < 2 bc (:240-252), else
 *     k-means: random_initializer (include/initializer.h:35-55) and kmean_par
 *     (src/kmeans.cpp:59-145) for `epochs` Lloyd steps;
 *   - centroids permuted by random_reordering (include/ckernel.h:177-202, and
 *     random_reordering_3d :811-840): srand(42), j = rand() % (i+1), i = n-1..1;
 *   - reorderIndex (include/vecchia_helper_method.h:130-157): the original
 *     cluster of every permuted centroid, by nearest centroid (first minimum);
 *   - firstbatch / cluster_combine (:37-108): the first clusters, in order, are
 *     merged into block 0 until it has >= cs points;
 *   - cluster_to_batch (:60-90): points of each block in original index order;
 *   - findNearestPoints (include/nearest_neighbor.h:76-148): for block i >= 1,
 *     the cs points among blocks 0..i-1 nearest to the block's (permuted)
 *     centroid, in the order the reference's selection sort (strict <) leaves
 *     them; block 0 gets points 0..cs-1.
 *
 * The reference's k-means reads uninitialised accumulators and coordinates
 * (include/Point.h:29-37: to_zero iterates by value; coordinates[2] is never
 * set in 2-D) and merges per-thread sums in a nondeterministic order.  Here the
 * accumulators start at zero, z = 0 in 2-D and the sums run in point order, so
 * the partition is a deterministic function of (points, bc, cs, seed).
 * random_initializer's distribution includes n (an out-of-range index); such a
 * draw is discarded and drawn again.
 *
 * The kNN selection sort is reproduced exactly, including its tie order, in
 * O(N + K^2) per block: only points with distance <= the K-th smallest can be
 * selected, and the selection picks the minimum of (distance, position) among
 * the remaining positions; positions change only through the sort's swaps,
 * which are replayed on the candidate set.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <float.h>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bv.h"
#include "bv_prep.h"

namespace {

struct Pt {
    double c[3];
    int cluster;
};

template <class F>
void parallel_for(int64_t n, int threads, F f) {
    if (threads <= 1 || n < 1024) {
        f((int64_t)0, n);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        pool.emplace_back([=]() { f(lo, hi); });
    }
    for (auto &th : pool) th.join();
}

inline double sqdist3(const double *a, const double *b) {
    double d = 0.0;
    for (int i = 0; i < 3; ++i) d += (a[i] - b[i]) * (a[i] - b[i]);
    return d;
}

/* random_initializer: distinct points drawn with mt19937(seed) */
std::vector<Pt> init_centroids(const std::vector<Pt> &pts, int k, int seed) {
    std::vector<Pt> cen;
    std::mt19937 gen(seed);
    std::uniform_int_distribution<> distrib(0, (int)pts.size());
    int i = 0;
    while ((int)cen.size() < k) {
        const int idx = distrib(gen);
        if (idx >= (int)pts.size()) continue;
        Pt p = pts[idx];
        bool dup = false;
        for (const Pt &q : cen)
            if (q.c[0] == p.c[0] && q.c[1] == p.c[1] && q.c[2] == p.c[2]) { dup = true; break; }
        if (!dup) {
            p.cluster = i++;
            cen.push_back(p);
        }
    }
    return cen;
}

/* kmean_par with a fixed accumulation order; the assignment step on the host
   or through `assign` */
int kmeans(std::vector<Pt> &pts, std::vector<Pt> &cen, int epochs, int k, int threads,
           bv_assign_fn assign, void *user) {
    const int64_t n = (int64_t)pts.size();
    std::vector<double> sum((size_t)k * 3);
    std::vector<int> card(k);
    std::vector<double> ccx, ccy, ccz;
    std::vector<int32_t> cl;
    if (assign) {
        ccx.resize(k); ccy.resize(k); ccz.resize(k);
        cl.resize((size_t)n);
    }
    for (int ep = 0; ep < epochs; ++ep) {
        if (assign) {
            for (int c = 0; c < k; ++c) {
                ccx[c] = cen[c].c[0]; ccy[c] = cen[c].c[1]; ccz[c] = cen[c].c[2];
            }
            const int rc = assign(user, k, ccx.data(), ccy.data(), ccz.data(), cl.data());
            if (rc != BV_OK) return rc;
            for (int64_t i = 0; i < n; ++i) pts[i].cluster = cl[i];
        } else parallel_for(n, threads, [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i) {
                Pt &p = pts[i];
                double best = sqdist3(p.c, cen[0].c);
                int bi = 0;
                for (int c = 1; c < k; ++c) {
                    const double t = sqdist3(p.c, cen[c].c);
                    if (t < best) { best = t; bi = c; }
                }
                p.cluster = bi;
            }
        });
        std::fill(sum.begin(), sum.end(), 0.0);
        std::fill(card.begin(), card.end(), 0);
        for (int64_t i = 0; i < n; ++i) {
            const Pt &p = pts[i];
            for (int d = 0; d < 3; ++d) sum[(size_t)p.cluster * 3 + d] += p.c[d];
            card[p.cluster]++;
        }
        for (int c = 0; c < k; ++c) {
            if (card[c] == 0) {
                /* empty cluster: the reference re-seeds mt19937(42) each time */
                std::mt19937 g2(42);
                std::uniform_int_distribution<> distr(0, (int)n - 1);
                const int r = distr(g2);
                for (int d = 0; d < 3; ++d) sum[(size_t)c * 3 + d] += pts[r].c[d];
                pts[r].cluster = c;
                card[c] += 1;
            }
            for (int d = 0; d < 3; ++d) cen[c].c[d] = sum[(size_t)c * 3 + d] / (double)card[c];
            cen[c].cluster = -1;
        }
    }
    return BV_OK;
}

inline double deg2rad(double d) { return d * 3.141592653589793 / 180; }

inline double earth_km(double lat1d, double lon1d, double lat2d, double lon2d) {
    const double lat1r = deg2rad(lat1d), lon1r = deg2rad(lon1d);
    const double lat2r = deg2rad(lat2d), lon2r = deg2rad(lon2d);
    const double u = sin((lat2r - lat1r) / 2), v = sin((lon2r - lon1r) / 2);
    return 2.0 * 6371.0 * asin(sqrt(u * u + cos(lat1r) * cos(lat2r) * v * v));
}

/* findNearestPoints for one block: the first k entries of the reference's
   selection sort over positions 0..N-1 */
void knn_block(const double *x, const double *y, const double *z, int64_t N, double cxv,
               double cyv, double czv, int k, int metric, int time_flag, int32_t *out,
               std::vector<double> &dist, std::vector<double> &tmp) {
    dist.resize((size_t)N);
    for (int64_t i = 0; i < N; ++i) {
        double d;
        if (metric == 1) {
            d = earth_km(cxv, cyv, x[i], y[i]);
        } else {
            const double dx = x[i] - cxv, dy = y[i] - cyv;
            d = sqrt(dx * dx + dy * dy);
        }
        if (time_flag) d = sqrt(d * d + (czv - z[i]) * (czv - z[i]));
        dist[i] = d;
    }
    if (k > N) k = (int)N;
    /* K-th smallest distance */
    tmp.assign(dist.begin(), dist.end());
    std::nth_element(tmp.begin(), tmp.begin() + (k - 1), tmp.end());
    const double dk = tmp[k - 1];
    /* candidates: every position with distance <= dk (ties included) */
    std::vector<int64_t> cpos;           /* current position of candidate c   */
    std::vector<int64_t> corig;          /* original index of candidate c     */
    std::unordered_map<int64_t, int> at; /* position -> candidate             */
    for (int64_t i = 0; i < N; ++i)
        if (dist[i] <= dk) {
            at[i] = (int)cpos.size();
            cpos.push_back(i);
            corig.push_back(i);
        }
    std::vector<char> taken(cpos.size(), 0);
    for (int s = 0; s < k; ++s) {
        int best = -1;
        for (size_t c = 0; c < cpos.size(); ++c) {
            if (taken[c]) continue;
            if (best < 0) { best = (int)c; continue; }
            const double dc = dist[corig[c]], db = dist[corig[best]];
            if (dc < db || (dc == db && cpos[c] < cpos[best])) best = (int)c;
        }
        /* swap positions s and cpos[best] */
        const int64_t from = cpos[best];
        auto it = at.find(s);
        if (it != at.end() && it->second != best) {
            const int other = it->second;
            cpos[other] = from;
            at[from] = other;
        } else if (from != s) {
            at.erase(from);
        }
        cpos[best] = s;
        at[s] = best;
        taken[best] = 1;
        out[s] = (int32_t)corig[best];
    }
}

}  // namespace

extern "C" void bv_blocks_free(bv_blocks *b) {
    if (!b) return;
    free(b->x); free(b->y); free(b->z); free(b->obs);
    free(b->batch_num); free(b->nbr); free(b->perm);
    free(b->cx); free(b->cy); free(b->cz);
    memset(b, 0, sizeof(*b));
}

extern "C" int bv_partition(int64_t n, const double *x, const double *y, const double *z,
                            const double *obs, int32_t bc, int32_t cs, int32_t seed,
                            int32_t epochs, int32_t time_flag, int32_t metric, int32_t threads,
                            bv_blocks *out) {
    return bv_partition_run(n, x, y, z, obs, bc, cs, seed, epochs, time_flag, metric, threads,
                            nullptr, nullptr, out);
}

int bv_partition_run(int64_t n, const double *x, const double *y, const double *z,
                     const double *obs, int32_t bc, int32_t cs, int32_t seed, int32_t epochs,
                     int32_t time_flag, int32_t metric, int32_t threads, bv_assign_fn assign,
                     void *user, bv_blocks *out, bv_knn_fn knn) {
    if (!out || n <= 0 || !x || !y || !obs || bc <= 0 || cs < 0 || epochs < 0) return BV_ERR_ARG;
    if (time_flag && !z) return BV_ERR_ARG;
    if (n > 2147483647LL) return BV_ERR_ARG;
    memset(out, 0, sizeof(*out));
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    const int64_t nclusters = (n < 2 * (int64_t)bc) ? n : bc;

    std::vector<Pt> pts((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        pts[i].c[0] = x[i];
        pts[i].c[1] = y[i];
        pts[i].c[2] = time_flag ? z[i] : 0.0;
        pts[i].cluster = -1;
    }
    std::vector<Pt> cen;
    std::vector<int> clusterNum;
    if (n < 2 * (int64_t)bc) {
        for (int64_t i = 0; i < n; ++i) pts[i].cluster = (int)i;
        cen = pts;
        clusterNum.assign((size_t)n, 1);
    } else {
        /* k-means assigns every point to a cluster only from its first epoch on */
        if (epochs < 1) return BV_ERR_ARG;
        cen = init_centroids(pts, (int)nclusters, seed);
        const int krc = kmeans(pts, cen, epochs, (int)nclusters, threads, assign, user);
        if (krc != BV_OK) return krc;
        int maxid = 0;
        for (const Pt &p : pts) maxid = std::max(maxid, p.cluster);
        clusterNum.assign((size_t)maxid + 1, 0);
        for (const Pt &p : pts) clusterNum[p.cluster]++;
        if ((int64_t)clusterNum.size() < nclusters) clusterNum.resize((size_t)nclusters, 0);
    }

    /* centroid copies, permuted with srand(42) / rand() */
    std::vector<double> lx((size_t)nclusters), ly((size_t)nclusters), lz((size_t)nclusters);
    for (int64_t i = 0; i < nclusters; ++i) {
        lx[i] = cen[i].c[0];
        ly[i] = cen[i].c[1];
        lz[i] = cen[i].c[2];
    }
    srand(42);
    for (int64_t i = nclusters - 1; i > 0; i--) {
        const int64_t j = rand() % (i + 1);
        std::swap(lx[i], lx[j]);
        std::swap(ly[i], ly[j]);
        if (time_flag) std::swap(lz[i], lz[j]);
    }
    /* reorderIndex: nearest original centroid of each permuted centroid */
    std::vector<int> perm((size_t)nclusters);
    parallel_for(nclusters, threads, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            double best = DBL_MAX;
            int bi = 0;
            for (int64_t jj = 0; jj < nclusters; ++jj) {
                const double dx = lx[i] - cen[jj].c[0], dy = ly[i] - cen[jj].c[1];
                double d = dx * dx + dy * dy;
                if (time_flag) {
                    const double dz = lz[i] - cen[jj].c[2];
                    d += dz * dz;
                }
                if (d < best) { best = d; bi = (int)jj; }
            }
            perm[i] = bi;
        }
    });
    /* firstbatch */
    int firstSize = 0, firstCount = 0;
    {
        int64_t acc = 0;
        for (int64_t i = 0; i < nclusters; ++i) {
            acc += clusterNum[perm[i]];
            if (acc >= cs) { firstSize = (int)acc; firstCount = (int)i + 1; break; }
        }
    }
    if (firstCount <= 0 || firstSize <= 0) return BV_ERR_ARG; /* fewer than cs points */
    /* cluster_combine */
    for (int i = 1; i < firstCount; ++i) {
        for (int64_t kk = 0; kk < n; ++kk)
            if (pts[kk].cluster == perm[i]) pts[kk].cluster = perm[0];
        perm[i] = perm[0];
        clusterNum[perm[i]] = firstSize;
    }
    const int64_t B = nclusters - firstCount + 1;
    const int *batchIndex = perm.data() + firstCount - 1;
    std::vector<int32_t> bnum((size_t)B);
    std::vector<int64_t> accum((size_t)B + 1, 0);
    for (int64_t i = 0; i < B; ++i) {
        bnum[i] = clusterNum[batchIndex[i]];
        accum[i + 1] = accum[i] + bnum[i];
    }
    if (accum[B] != n) return BV_ERR_ARG;

    out->n = n;
    out->batch_count = (int32_t)B;
    out->cs = cs;
    out->first_cluster_count = firstCount;
    out->nclusters = (int32_t)nclusters;
    out->x = (double *)malloc(sizeof(double) * n);
    out->y = (double *)malloc(sizeof(double) * n);
    out->z = time_flag ? (double *)malloc(sizeof(double) * n) : nullptr;
    out->obs = (double *)malloc(sizeof(double) * n);
    out->perm = (int64_t *)malloc(sizeof(int64_t) * n);
    out->batch_num = (int32_t *)malloc(sizeof(int32_t) * B);
    out->nbr = (int32_t *)malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(1, B * cs));
    out->cx = (double *)malloc(sizeof(double) * nclusters);
    out->cy = (double *)malloc(sizeof(double) * nclusters);
    out->cz = (double *)malloc(sizeof(double) * nclusters);
    if (!out->x || !out->y || !out->obs || !out->perm || !out->batch_num || !out->nbr ||
        !out->cx || !out->cy || !out->cz || (time_flag && !out->z)) {
        bv_blocks_free(out);
        return BV_ERR_ARG;
    }
    memcpy(out->batch_num, bnum.data(), sizeof(int32_t) * B);
    memcpy(out->cx, lx.data(), sizeof(double) * nclusters);
    memcpy(out->cy, ly.data(), sizeof(double) * nclusters);
    memcpy(out->cz, lz.data(), sizeof(double) * nclusters);

    /* cluster_to_batch: bucket the points by cluster in index order */
    {
        std::unordered_map<int, int64_t> slot;
        slot.reserve((size_t)B * 2);
        for (int64_t i = 0; i < B; ++i) slot[batchIndex[i]] = accum[i];
        for (int64_t kk = 0; kk < n; ++kk) {
            auto it = slot.find(pts[kk].cluster);
            if (it == slot.end()) continue;
            const int64_t o = it->second++;
            out->x[o] = x[kk];
            out->y[o] = y[kk];
            if (time_flag) out->z[o] = z[kk];
            out->obs[o] = obs[kk];
            out->perm[o] = kk;
        }
        for (int64_t i = 0; i < B; ++i)
            if (slot[batchIndex[i]] != accum[i + 1]) { bv_blocks_free(out); return BV_ERR_ARG; }
    }

    /* conditioning sets */
    if (cs > 0) {
        for (int t = 0; t < cs; ++t) out->nbr[t] = t;
        const double *bx = out->x, *by = out->y, *bz = out->z;
        int32_t *nbr = out->nbr;
        const int fc = firstCount;
        std::vector<int32_t> redo((size_t)B, 1);
        if (knn && metric == 0 && B > 1) {
            /* block b's centroid is the permuted centroid b + firstCount - 1 */
            std::vector<double> bcx((size_t)B, 0.0), bcy((size_t)B, 0.0), bcz((size_t)B, 0.0);
            for (int64_t i = 1; i < B; ++i) {
                bcx[i] = lx[i + fc - 1];
                bcy[i] = ly[i + fc - 1];
                bcz[i] = time_flag ? lz[i + fc - 1] : 0.0;
            }
            std::fill(redo.begin(), redo.end(), 0);
            const int rc = knn(user, n, bx, by, bz, B, accum.data(), bcx.data(), bcy.data(),
                               bcz.data(), cs, time_flag, nbr, redo.data());
            if (rc != BV_OK) { bv_blocks_free(out); return rc; }
        }
        parallel_for(B - 1, threads, [&](int64_t lo, int64_t hi) {
            std::vector<double> dist, tmp;
            for (int64_t ii = lo; ii < hi; ++ii) {
                const int64_t i = ii + 1;
                if (!redo[i]) continue;
                const int64_t ci = i + fc - 1;
                knn_block(bx, by, bz, accum[i], lx[ci], ly[ci], time_flag ? lz[ci] : 0.0, cs,
                          metric, time_flag, nbr + i * cs, dist, tmp);
            }
        });
    }
    return BV_OK;
}
