"""Plain-Python restatement of the reference's host pre-processing
(src/test_Xvecchia_batch.cpp:240-404), the checker for libbv's bv_partition.
TEST INFRASTRUCTURE ONLY (imported by tests/ only).

Follows, step by step:
  classic Vecchia when n < 2 bc                  src/test_Xvecchia_batch.cpp:240-252
  random_initializer (mt19937(seed), uniform_int_distribution(0, n))
                                                 include/initializer.h:35-55
  kmean_par, `epochs` Lloyd steps                src/kmeans.cpp:59-145
  random_reordering (srand(42), rand() % (i+1))  include/ckernel.h:177-202
  reorderIndex (first nearest centroid)          include/vecchia_helper_method.h:130-157
  firstbatch / cluster_combine / cluster_to_batch include/vecchia_helper_method.h:37-108
  findNearestPoints: literal selection sort      include/nearest_neighbor.h:76-148
with the deterministic conventions of the restatement: zeroed accumulators,
z = 0 in 2-D, sums in point order, an out-of-range initializer draw redrawn.
glibc rand() is called through ctypes; mt19937 and libstdc++'s
uniform_int_distribution (Lemire's method, GCC >= 11) are restated here, so the
random streams are the ones the reference's build draws.
"""
import ctypes
import math

import numpy as np

_libc = ctypes.CDLL("libc.so.6")
_libc.rand.restype = ctypes.c_int
_libc.srand.argtypes = [ctypes.c_uint]


def _mt19937_uniform_ints(seed, lo, hi, count):
    """count draws of std::uniform_int_distribution<int>(lo, hi) over
    std::mt19937(seed), as libstdc++ computes them."""
    # mt19937 (Matsumoto & Nishimura 1998), 32-bit
    mt = [0] * 624
    mt[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        mt[i] = (1812433253 * (mt[i - 1] ^ (mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
    idx = [624]

    def nxt():
        if idx[0] >= 624:
            for i in range(624):
                yv = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7FFFFFFF)
                mt[i] = mt[(i + 397) % 624] ^ (yv >> 1) ^ (0x9908B0DF if yv & 1 else 0)
            idx[0] = 0
        yv = mt[idx[0]]
        idx[0] += 1
        yv ^= yv >> 11
        yv ^= (yv << 7) & 0x9D2C5680
        yv ^= (yv << 15) & 0xEFC60000
        yv ^= yv >> 18
        return yv & 0xFFFFFFFF

    # libstdc++ (GCC >= 11) uniform_int_distribution over a 32-bit engine:
    # Lemire's multiply-and-reject ("nearly divisionless") method with 64-bit
    # products, range = hi - lo + 1
    rng_ = hi - lo + 1
    threshold = ((1 << 32) - rng_) % rng_
    out = []
    while len(out) < count:
        prod = nxt() * rng_
        low = prod & 0xFFFFFFFF
        if low < rng_:
            while low < threshold:
                prod = nxt() * rng_
                low = prod & 0xFFFFFFFF
        out.append(lo + (prod >> 32))
    return out


class _Gen:
    """Lazily extended stream of uniform_int_distribution draws."""

    def __init__(self, seed, lo, hi):
        self.seed, self.lo, self.hi = seed, lo, hi
        self.buf = []
        self.pos = 0

    def __call__(self):
        if self.pos >= len(self.buf):
            self.buf = _mt19937_uniform_ints(self.seed, self.lo, self.hi, 2 * len(self.buf) + 64)
        v = self.buf[self.pos]
        self.pos += 1
        return v


def partition(x, y, obs, bc, cs, seed=0, epochs=50):
    n = len(x)
    pts = np.column_stack([x, y, np.zeros(n)])
    if n < 2 * bc:
        ncl = n
        cluster = np.arange(n)
        cen = pts.copy()
        clusterNum = np.ones(n, dtype=np.int64)
    else:
        ncl = bc
        g = _Gen(seed, 0, n)
        cen_l = []
        while len(cen_l) < ncl:
            i = g()
            if i >= n:
                continue
            p = pts[i]
            if not any((q == p).all() for q in cen_l):
                cen_l.append(p.copy())
        cen = np.array(cen_l)
        cluster = np.zeros(n, dtype=np.int64)
        empty_pick = _mt19937_uniform_ints(42, 0, n - 1, 1)[0]
        for _ in range(epochs):
            for i in range(n):
                # first minimum with strict '<', computed term by term like the reference
                best = 0
                bd = ((pts[i][0] - cen[0][0]) * (pts[i][0] - cen[0][0]) +
                      (pts[i][1] - cen[0][1]) * (pts[i][1] - cen[0][1]) +
                      (pts[i][2] - cen[0][2]) * (pts[i][2] - cen[0][2]))
                for c in range(1, ncl):
                    t = ((pts[i][0] - cen[c][0]) * (pts[i][0] - cen[c][0]) +
                         (pts[i][1] - cen[c][1]) * (pts[i][1] - cen[c][1]) +
                         (pts[i][2] - cen[c][2]) * (pts[i][2] - cen[c][2]))
                    if t < bd:
                        bd, best = t, c
                cluster[i] = best
            sums = np.zeros((ncl, 3))
            card = np.zeros(ncl, dtype=np.int64)
            for i in range(n):
                for k in range(3):
                    sums[cluster[i], k] += pts[i, k]
                card[cluster[i]] += 1
            for c in range(ncl):
                if card[c] == 0:
                    for k in range(3):
                        sums[c, k] += pts[empty_pick, k]
                    cluster[empty_pick] = c
                    card[c] += 1
                for k in range(3):
                    cen[c, k] = sums[c, k] / float(card[c])
        clusterNum = np.bincount(cluster, minlength=ncl)
    lx, ly = list(cen[:, 0]), list(cen[:, 1])
    _libc.srand(42)
    for i in range(ncl - 1, 0, -1):
        j = _libc.rand() % (i + 1)
        lx[i], lx[j] = lx[j], lx[i]
        ly[i], ly[j] = ly[j], ly[i]
    perm = []
    for i in range(ncl):
        best, bd = 0, math.inf
        for j in range(ncl):
            dx, dy = lx[i] - cen[j, 0], ly[i] - cen[j, 1]
            d = dx * dx + dy * dy
            if d < bd:
                bd, best = d, j
        perm.append(best)
    clusterNum = list(clusterNum)
    acc, fsize, fcount = 0, 0, 0
    for i in range(ncl):
        acc += clusterNum[perm[i]]
        if acc >= cs:
            fsize, fcount = acc, i + 1
            break
    cluster = list(cluster)
    for i in range(1, fcount):
        for k in range(n):
            if cluster[k] == perm[i]:
                cluster[k] = perm[0]
        perm[i] = perm[0]
        clusterNum[perm[i]] = fsize
    B = ncl - fcount + 1
    bidx = perm[fcount - 1:]
    bnum = [clusterNum[bidx[i]] for i in range(B)]
    accum = np.concatenate([[0], np.cumsum(bnum)])
    order = []
    for i in range(B):
        order += [k for k in range(n) if cluster[k] == bidx[i]]
    order = np.array(order)
    bx, by, bo = x[order], y[order], obs[order]
    nbr = np.zeros(B * cs, dtype=np.int32)
    if cs > 0:
        nbr[:cs] = np.arange(cs)
        for i in range(1, B):
            cxv, cyv = lx[i + fcount - 1], ly[i + fcount - 1]
            N = int(accum[i])
            dist = [math.sqrt((bx[t] - cxv) * (bx[t] - cxv) + (by[t] - cyv) * (by[t] - cyv))
                    for t in range(N)]
            ind = list(range(N))
            for s in range(cs):
                mi = s
                for t in range(s + 1, N):
                    if dist[t] < dist[mi]:
                        mi = t
                dist[s], dist[mi] = dist[mi], dist[s]
                ind[s], ind[mi] = ind[mi], ind[s]
            nbr[i * cs:(i + 1) * cs] = ind[:cs]
    return dict(x=bx, y=by, obs=bo, batch_num=np.array(bnum, dtype=np.int32), nbr=nbr,
                perm=order, first_cluster_count=fcount, cx=np.array(lx), cy=np.array(ly))
