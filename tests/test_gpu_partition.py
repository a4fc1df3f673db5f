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
