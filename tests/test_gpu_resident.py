"""The register-resident factorization (bv_res_kernel, BV_FACTOR_KERNEL=res):
one workgroup per block with the factor held in registers.  Against the
oracle on the shapes its builds cover (padded sizes up to 128, 192, 256, 320
and 384; several builds in one evaluation, the largest blocks' on a second
stream), every covariance kind it takes (closed forms, general nu, 3-D
distances, space-time read from the workspace), the non-positive-definite
report, and its limits."""
import math
import sys

import numpy as np
import pytest

import blockvecchia_amd as bv
import oracle
from blockvecchia_amd import datagen
from conftest import LLH_RTOL, rel

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[resident] " + msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(autouse=True)
def _resident(monkeypatch):
    for k in ("BV_PATH", "BV_CHUNKS", "BV_CHUNK_STREAMS"):
        monkeypatch.delenv(k, raising=False)
    monkeypatch.setenv("BV_FACTOR_KERNEL", "res")


def _check(blocks, theta, kernel=1, block_rtol=1e-9, **kw):
    eng = bv.Engine(blocks, kernel=kernel, **kw)
    try:
        assert eng.factor_config()[0] == 3
        v = eng.llh(theta)
        st, bad = eng.status()
        ld, dt, info = eng.block_terms()
    finally:
        eng.close()
    ov, old, odt, oinfo = oracle.llh(blocks, theta, kernel=kernel, metric=kw.get("metric", 0),
                                     time_flag=kw.get("time_flag", False),
                                     dist_scale=kw.get("dist_scale", 1.0), per_block=True)
    _say("  gpu %.17g oracle %.17g rel %.3e" % (v, ov, rel(v, ov)))
    assert st == bv.BV_OK and np.all(info == 0) and np.all(oinfo == 0), (st, bad)
    assert rel(v, ov) <= LLH_RTOL
    np.testing.assert_allclose(ld, old, rtol=block_rtol, atol=1e-9)
    np.testing.assert_allclose(dt, odt, rtol=block_rtol, atol=1e-9)


def _max_mp(blocks, cs):
    m = blocks.batch_num.astype(np.int64) + cs
    m[0] = blocks.batch_num[0]
    return int(((m + 1 + 31) // 32 * 32).max())


@pytest.mark.parametrize("cs,theta,block_rtol", [
    (330, [1.0, 0.1, 0.5], 1e-9),   # every block above 320 rows: the 24-row build alone
    (300, [1.0, 0.1, 1.5], 1e-9),   # 320 and 352 rows: both builds on two streams
    # nu = 5/2: a single block's quadratic form carries ~1e-9 of rounding
    # from its conditioning (the llh still within 1e-11)
    (250, [1.0, 0.05, 2.5], 1e-8),
    (120, [1.0, 0.08, 0.8], 1e-9),  # general nu (Chebyshev table)
])
def test_resident_shapes(cs, theta, block_rtol):
    x, y = datagen.uniform_locations(1500, 4)
    obs = datagen.sample_dense(x, y, *theta, seed=4)
    b = bv.partition(x, y, obs, 100, cs, seed=0)
    _say("case: n=1500 bc=100 cs=%d theta=%s padded <= %d" % (cs, theta, _max_mp(b, cs)))
    _check(b, theta, block_rtol=block_rtol)


def _builds(blocks, cs):
    m = blocks.batch_num.astype(np.int64) + cs
    m[0] = blocks.batch_num[0]
    mp = (m + 1 + 31) // 32 * 32
    nr = np.where(mp <= 128, 2, (mp + 63) // 64)
    u, c = np.unique(nr, return_counts=True)
    return dict(zip(u.tolist(), c.tolist()))


@pytest.mark.parametrize("n,bc,cs,theta", [
    (1500, 100, 30, [1.0, 0.1, 0.5]),    # 8-row build (two workgroups per compute unit)
    (1500, 60, 60, [1.2, 0.08, 1.5]),    # 8-row build, larger blocks
    (2500, 125, 150, [1.0, 0.08, 0.8]),  # 12-row build (and 16 for a few), general nu
    (2500, 100, 200, [1.0, 0.1, 0.5]),   # 16-row build (and 12 / 20 around it)
])
def test_resident_builds(n, bc, cs, theta):
    """Each build of the resident kernel, from 8 to 16 tile rows, against the
    oracle, with the blocks of one evaluation split over several builds."""
    x, y = datagen.uniform_locations(n, 8)
    obs = datagen.sample_dense(x, y, *theta, seed=8)
    b = bv.partition(x, y, obs, bc, cs, seed=0)
    _say("case: n=%d bc=%d cs=%d theta=%s blocks per build (4 NR rows) %s"
         % (n, bc, cs, theta, _builds(b, cs)))
    _check(b, theta)


def test_resident_many_and_graph_over_builds():
    """An evaluation whose blocks span several builds (one launch per build,
    the largest blocks' on a second stream): bv_llh_many equals bv_llh bit
    for bit (each build's descriptors of k evaluations contiguous), HIP graph
    replays equal direct launches, and the oracle agrees."""
    x, y = datagen.uniform_locations(4000, 9)
    obs = np.random.default_rng(9).standard_normal(4000)
    b = bv.partition(x, y, obs, 60, 150, seed=0)
    builds = _builds(b, 150)
    _say("case: n=4000 bc=60 cs=150 blocks per build %s" % builds)
    assert len(builds) >= 3
    thetas = [[1.0, 0.1, 0.5], [1.2, 0.08, 0.5], [0.9, 0.12, 0.5]]
    eng = bv.Engine(b)
    try:
        assert eng.factor_config()[0] == 3
        eng.set_graph(False)
        single = [eng.llh(t) for t in thetas]
        many, st = eng.llh_many(np.array(thetas))
        assert np.all(st == bv.BV_OK) and list(many) == single, (list(many), single)
        eng.set_graph(True)
        for rep in range(3):
            for i, t in enumerate(thetas):
                got = eng.llh(t)
                assert got == single[i], (rep, i, got, single[i])
    finally:
        eng.close()
    ref = oracle.llh(b, thetas[0])
    _say("  gpu %.17g oracle %.17g rel %.3e" % (single[0], ref, rel(single[0], ref)))
    assert rel(single[0], ref) <= LLH_RTOL


def test_resident_not_tuned_for_tiny_blocks(monkeypatch):
    """The autotuner does not keep the resident form on a dataset whose blocks
    are all at most 128 rows padded (merged serving is faster with the
    one-wavefront fused kernel there); forced, it still runs and agrees with
    the oracle."""
    monkeypatch.delenv("BV_FACTOR_KERNEL", raising=False)
    monkeypatch.setenv("BV_TUNE_CACHE", "0")
    x, y = datagen.uniform_locations(1500, 10)
    obs = np.random.default_rng(10).standard_normal(1500)
    b = bv.partition(x, y, obs, 100, 40, seed=0)
    assert _max_mp(b, 40) <= 128
    eng = bv.Engine(b)
    try:
        v = eng.llh([1.0, 0.1, 0.5])
        cfg = eng.factor_config()
    finally:
        eng.close()
    _say("case: tiny blocks (padded <= %d): autotuned %s" % (_max_mp(b, 40), cfg))
    assert cfg[0] != 3
    ref = oracle.llh(b, [1.0, 0.1, 0.5])
    assert rel(v, ref) <= LLH_RTOL
    monkeypatch.setenv("BV_FACTOR_KERNEL", "res")
    _check(b, [1.0, 0.1, 0.5])


def test_resident_3d_and_dist_scale():
    x, y = datagen.uniform_locations(800, 5)
    z = np.random.default_rng(5).random(800) * 0.3
    obs = np.random.default_rng(6).standard_normal(800)
    b = bv.partition(x, y, obs, 40, 80, seed=0, z=z, time_flag=True)
    _say("case: 3-D distance, dist_scale 2")
    _check(b, [1.5, 0.1, 0.5], time_flag=True, dist_scale=2.0)


def test_resident_spacetime():
    rng = np.random.default_rng(12)
    n = 400
    x, y = rng.random(n), rng.random(n)
    t = (rng.integers(1, 5, n)) / 4.0
    obs = rng.standard_normal(n)
    b = bv.partition(x, y, obs, 20, 40, seed=0, z=t, time_flag=True)
    _say("case: space-time kernel (A read from the workspace)")
    _check(b, [1.0, 0.1, 0.5, 2.2, 0.5, 0.5, 0.0], kernel=4, time_flag=True)


def test_resident_not_positive_definite():
    x = np.array([0.1, 0.1, 0.5, 0.7, 0.2, 0.9])
    y = np.array([0.2, 0.2, 0.5, 0.1, 0.8, 0.3])
    b = bv.Blocks(x, y, np.ones(6), np.array([3, 3], np.int32), np.zeros(0, np.int32), 0)
    eng = bv.Engine(b)
    v = eng.llh([1.0, 0.1, 0.5])
    st, bad = eng.status()
    _, _, info = eng.block_terms()
    eng.close()
    _say("case: non-PD block: llh %r status %d block %d info %s" % (v, st, bad, info))
    assert v == -math.inf and st == bv.BV_ERR_NOT_PD and bad == 0 and info[0] == 2


def test_resident_limits():
    """Refused when forced on blocks above 384 rows or great-circle distances
    (the autotuner does not offer it there)."""
    x, y = datagen.uniform_locations(1200, 7)
    obs = np.random.default_rng(7).standard_normal(1200)
    with pytest.raises(RuntimeError):
        bv.Engine(bv.partition(x, y, obs, 60, 400, seed=0))
    lat, lon = 20 + 30 * x[:300], -120 + 50 * y[:300]
    with pytest.raises(RuntimeError):
        bv.Engine(bv.partition(lat, lon, obs[:300], 15, 40, seed=0, metric=1), metric=1)


def test_resident_ill_conditioned_within_floor():
    """nu = 5/2, range 0.1 on 1,500 points (cs = 250): the blocks'
    conditioning amplifies rounding, so the bound is stated against the
    measured difference of two correct FP64 CPU implementations (the oracle
    and a numpy/LAPACK chain): the GPU stays within 10x of it."""
    from test_oracle import _numpy_chain
    th = [1.0, 0.1, 2.5]
    x, y = datagen.uniform_locations(1500, 4)
    obs = datagen.sample_dense(x, y, *th, seed=4)
    b = bv.partition(x, y, obs, 100, 250, seed=0)
    eng = bv.Engine(b)
    v = eng.llh(th)
    eng.close()
    ov = oracle.llh(b, th)
    floor = rel(_numpy_chain(b, th), ov)
    _say("case: ill-conditioned nu=2.5: gpu %.17g oracle %.17g rel %.3e floor %.3e"
         % (v, ov, rel(v, ov), floor))
    assert rel(v, ov) <= max(10 * floor, LLH_RTOL)
