"""The pre-processing with the k-means assignment steps on the GPU
(bv_partition_gpu) gives bv_partition's blocks bit for bit: the same block
sizes, point order, observations and conditioning sets, in 2-D, 3-D (time
flag), on clustered points with empty clusters, and in the classic-Vecchia
case that runs no k-means."""
import numpy as np
import pytest

import blockvecchia_amd as bv
from blockvecchia_amd import datagen

pytestmark = pytest.mark.gpu


def _same(a, b):
    for f in ("x", "y", "z", "obs", "batch_num", "nbr", "perm", "cx", "cy"):
        va, vb = getattr(a, f, None), getattr(b, f, None)
        if va is None and vb is None:
            continue
        assert np.array_equal(np.asarray(va), np.asarray(vb)), f


@pytest.mark.parametrize("n,bc,cs,clustered,time_flag", [
    (5000, 300, 40, False, False),
    (30000, 2000, 60, True, False),
    (8000, 500, 30, False, True),
    (1000, 800, 20, False, False),   # n < 2 bc: classic Vecchia
])
def test_gpu_kmeans_partition_identical(n, bc, cs, clustered, time_flag):
    print(f"partition n={n} bc={bc} cs={cs} clustered={clustered} 3d={time_flag}", flush=True)
    if clustered:
        x, y = datagen.clustered_locations(n, 4)
    else:
        x, y = datagen.uniform_locations(n, 4)
    z = np.random.default_rng(5).uniform(0, 1, n) if time_flag else None
    obs = np.random.default_rng(6).standard_normal(n)
    a = bv.partition(x, y, obs, bc, cs, seed=2, z=z, time_flag=time_flag)
    b = bv.partition(x, y, obs, bc, cs, seed=2, z=z, time_flag=time_flag, device=0)
    _same(a, b)


def _grid(n, seed):
    """n distinct points of an 80 x 80 lattice in random order: many equal
    distances to a centroid, so the selection sort's tie order decides the
    conditioning sets."""
    rng = np.random.default_rng(seed)
    cells = rng.permutation(80 * 80)[:n]
    return (cells % 80) / 80.0, (cells // 80) / 80.0


@pytest.mark.parametrize("n,bc,cs,points,time_flag", [
    (6000, 300, 100, "grid", False),      # heavy ties
    (4000, 4000, 30, "grid", False),      # classic Vecchia on a lattice: ties at every step
    (6000, 300, 60, "grid", True),        # 3-D distance with ties
    (200000, 10000, 400, "clustered", False),   # the uneven-batching shape
])
def test_gpu_knn_tie_order_identical(n, bc, cs, points, time_flag):
    """The conditioning sets from the device kNN (radix select, then the
    reference's selection-sort swaps replayed on the candidates) equal the
    host's, entry for entry, including the order of tied distances."""
    print(f"knn n={n} bc={bc} cs={cs} points={points} 3d={time_flag}", flush=True)
    if points == "grid":
        x, y = _grid(n, 7)
    else:
        x, y = datagen.clustered_locations(n, 0)
    z = (np.random.default_rng(8).integers(0, 4, n) / 4.0) if time_flag else None
    obs = np.zeros(n)
    b = bv.partition(x, y, obs, bc, cs, seed=0, z=z, time_flag=time_flag, device=0)
    import os
    os.environ["BV_KNN_HOST"] = "1"
    try:
        a = bv.partition(x, y, obs, bc, cs, seed=0, z=z, time_flag=time_flag, device=0)
    finally:
        del os.environ["BV_KNN_HOST"]
    _same(a, b)
