"""Host pre-processing (libbv bv_partition, C++) against the plain-Python
restatement of the reference's steps (oracle/partition.py), bit-exactly, plus
the invariants every partition must satisfy.  CPU only."""
import numpy as np
import pytest

import blockvecchia_amd as bv
from blockvecchia_amd import datagen
from oracle import partition as opart


@pytest.mark.parametrize("n,bc,cs,seed", [(400, 20, 30, 0), (300, 40, 25, 3), (60, 40, 10, 1)])
def test_partition_matches_restatement(n, bc, cs, seed):
    x, y = datagen.uniform_locations(n, seed)
    obs = np.arange(n, dtype=np.float64)
    b = bv.partition(x, y, obs, bc, cs, seed=seed, epochs=10)
    r = opart.partition(x, y, obs, bc, cs, seed=seed, epochs=10)
    assert b.first_cluster_count == r["first_cluster_count"]
    np.testing.assert_array_equal(b.batch_num, r["batch_num"])
    np.testing.assert_array_equal(b.perm, r["perm"])
    np.testing.assert_array_equal(b.x, r["x"])
    np.testing.assert_array_equal(b.obs, r["obs"])
    np.testing.assert_array_equal(b.nbr, r["nbr"])
    np.testing.assert_array_equal(b.cx, r["cx"])


def test_knn_ties_follow_selection_sort():
    """A lattice has many equal distances: the tie order of the reference's
    selection sort must be reproduced."""
    g = np.arange(12) / 11.0
    x, y = np.meshgrid(g, g)
    x, y = x.ravel(), y.ravel()
    obs = np.zeros(len(x))
    b = bv.partition(x, y, obs, 200, 20, seed=0, epochs=5)  # classic Vecchia (n < 2 bc)
    r = opart.partition(x, y, obs, 200, 20, seed=0, epochs=5)
    np.testing.assert_array_equal(b.nbr, r["nbr"])
    np.testing.assert_array_equal(b.batch_num, r["batch_num"])


def test_partition_invariants():
    x, y = datagen.uniform_locations(3000, 5)
    obs = np.random.default_rng(0).standard_normal(3000)
    cs = 50
    b = bv.partition(x, y, obs, 150, cs, seed=5)
    acc = np.concatenate([[0], np.cumsum(b.batch_num)])
    assert acc[-1] == 3000
    assert b.batch_num[0] >= cs
    assert sorted(b.perm.tolist()) == list(range(3000))
    np.testing.assert_array_equal(b.x, x[b.perm])
    np.testing.assert_array_equal(b.obs, obs[b.perm])
    np.testing.assert_array_equal(b.nbr[:cs], np.arange(cs))
    fc = b.first_cluster_count
    for i in range(1, b.batch_count):
        nb = b.nbr[i * cs:(i + 1) * cs]
        assert nb.max() < acc[i] and nb.min() >= 0      # neighbours precede the block
        assert len(set(nb.tolist())) == cs
        c = (b.cx[i + fc - 1], b.cy[i + fc - 1])
        d = np.sqrt((b.x[nb] - c[0]) ** 2 + (b.y[nb] - c[1]) ** 2)
        assert np.all(np.diff(d) >= 0)                  # ascending distance
        dall = np.sqrt((b.x[:acc[i]] - c[0]) ** 2 + (b.y[:acc[i]] - c[1]) ** 2)
        assert d[-1] <= np.sort(dall)[cs - 1]           # the cs nearest


def test_classic_vecchia_one_point_blocks():
    x, y = datagen.uniform_locations(100, 2)
    b = bv.partition(x, y, np.zeros(100), 80, 10, seed=2)
    assert b.batch_num[0] >= 10
    assert np.all(b.batch_num[1:] == 1)


def test_zero_kmeans_epochs_rejected():
    """k-means with no epoch assigns no point to a cluster: an argument
    error, not a partition (the classic-Vecchia path needs no epoch)."""
    x, y = datagen.uniform_locations(400, 0)
    obs = np.zeros(400)
    with pytest.raises(RuntimeError, match="bv_partition failed"):
        bv.partition(x, y, obs, 20, 30, seed=0, epochs=0)
    b = bv.partition(x, y, obs, 300, 30, seed=0, epochs=0)  # n < 2 bc: one point per cluster
    assert int(np.sum(b.batch_num)) == 400
