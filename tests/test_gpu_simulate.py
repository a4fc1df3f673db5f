"""Block Vecchia simulation on the GPU (bv_simulate, dependency levels)
against the CPU sampler (datagen.sample_block_vecchia, scipy block by block,
itself pinned by the Monte Carlo covariance test in test_datagen.py) on the
same standard normals: the same draw to rounding.  Runs through the C ABI."""
import sys

import numpy as np
import pytest

import blockvecchia_amd as bv
from blockvecchia_amd import datagen

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[simulate] " + msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,bc,cs,theta", [
    (2000, 100, 60, [1.5, 0.1, 0.5]),
    (2000, 100, 60, [1.0, 0.1, 1.5]),
    (1500, 75, 40, [1.0, 0.08, 0.8]),      # general nu: generation launch first
    (1200, 1200, 30, [1.0, 0.1, 1.5]),     # classic Vecchia: one point per block
    (900, 45, 0, [1.0, 0.1, 0.5]),         # independent blocks: one level
])
def test_simulate_equals_cpu_sampler(n, bc, cs, theta):
    x, y = datagen.uniform_locations(n, 31)
    b = bv.partition(x, y, np.zeros(n), bc, cs, seed=0)
    z = np.random.default_rng(32).standard_normal(n)
    eng = bv.Engine(b)
    try:
        yg, rc, levels = eng.simulate(theta, z)
    finally:
        eng.close()
    yc = np.asarray(datagen.sample_block_vecchia(b, *theta, threads=8, z=z))
    err = np.max(np.abs(yg - yc)) / np.max(np.abs(yc))
    _say("n=%d bc=%d cs=%d theta=%s: %d blocks, %d levels, max diff %.2e" % (
        n, bc, cs, theta, b.batch_count, levels, err))
    assert rc == bv.BV_OK
    assert (levels == 1) if cs == 0 else (1 < levels <= b.batch_count)
    assert err < 1e-10, err


def test_simulate_jitter_on_clustered_smooth_field():
    """nu = 5/2 on tightly clustered points is not numerically positive
    definite without a nugget: the first attempt reports it, the sampler
    retries with the smallest jitter that works, and the draw is finite."""
    n = 20000
    x, y = datagen.clustered_locations(n, 3, frac_uniform=0.3, sd=0.005)
    out = datagen.sample_vecchia(x, y, 1.0, 0.05, 2.5, cs=200, bc=n // 20, seed=3)
    _say("clustered nu=5/2: nugget %g, %d levels" % (out.nugget, out.levels))
    assert np.all(np.isfinite(np.asarray(out)))
    assert out.nugget > 0.0


def test_sample_vecchia_large_levels():
    """The C3 generator shape at a quarter of its size: n = 50,000 clustered,
    cs = 400, bc = 2,500; finite draw with unit-scale marginal variance."""
    n = 50000
    x, y = datagen.clustered_locations(n, 0)
    out = np.asarray(datagen.sample_vecchia(x, y, 1.0, 0.05, 1.5, cs=400, bc=n // 20, seed=0))
    _say("n=%d cs=400: var %.3f" % (n, out.var()))
    assert np.all(np.isfinite(out)) and 0.5 < out.var() < 2.0
