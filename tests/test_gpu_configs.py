"""The shapes of the uneven-batching and general-nu configurations on the GPU,
through the C ABI, in every factorization form:

  uneven: clustered points (70% uniform, 30% in 20 Gaussian clusters of sd
          0.02), cs = 400 (padded block sizes 416-448), nu = 3/2 and 5/2, at
          n = 20,000 against the oracle block by block, and at the full n =
          200,000 through size-independent properties.  Smooth kernels on
          clustered points are not all numerically positive definite: every
          block that fails on the GPU is shown singular in FP64 (numpy's
          eigenvalues), and blocks that are well conditioned never fail.
  general nu: uniform points, blocks of ~20, cs = 150, theta = (1, 0.08,
          0.8), observations from the Vecchia simulation with cs = 400: the
          likelihood against the oracle at 1e-10 in every form (general nu
          generates the covariance by its own launch, which the factorization
          then reads), and the driver's maximum-likelihood run with all three
          parameters free from 0.1:0.1:0.1 against scipy's optimizers on the
          same GPU objective."""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import blockvecchia_amd as bv
import oracle
from blockvecchia_amd import datagen
from conftest import LLH_RTOL, rel

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "blockvecchia_amd", "bin", "test_dvecchia_batch")

FORMS = {"fused": {"BV_PATH": "fused", "BV_FACTOR_KERNEL": "4w"},
         "stepped": {"BV_PATH": "stepped", "BV_CHUNKS": "2", "BV_CHUNK_STREAMS": "2"},
         "stepped2": {"BV_PATH": "stepped2", "BV_CHUNKS": "2", "BV_CHUNK_STREAMS": "0"},
         "paired": {"BV_PATH": "paired", "BV_CHUNKS": "3", "BV_CHUNK_STREAMS": "3"}}
# the register-resident form takes padded sizes up to 384: the general-nu
# shape (cs = 150), not the uneven one (cs = 400)
RES_FORM = {"BV_FACTOR_KERNEL": "res"}


def _say(msg):
    print("[configs] " + msg, flush=True)
    sys.stdout.flush()


def _engine(blocks, form, monkeypatch):
    for k in ("BV_PATH", "BV_FACTOR_KERNEL", "BV_CHUNKS", "BV_CHUNK_STREAMS"):
        monkeypatch.delenv(k, raising=False)
    for k, v in dict(FORMS, res=RES_FORM)[form].items():
        monkeypatch.setenv(k, v)
    return bv.Engine(blocks)


_C3 = {}


def _c3(nu):
    """n = 20,000 clustered points, bc = 1,000, cs = 400; observations from
    the Vecchia simulation over the same partition at (1, 0.05, nu); the
    oracle's per-block terms (cached per nu)."""
    if nu not in _C3:
        n = 20000
        x, y = datagen.clustered_locations(n, 7)
        b0 = bv.partition(x, y, np.zeros(n), 1000, 400, seed=0, device=0)
        obs = datagen.sample_partition(b0, 1.0, 0.05, nu, seed=7)
        b = bv.Blocks(b0.x, b0.y, np.asarray(obs, dtype=np.float64), b0.batch_num, b0.nbr, 400)
        th = [1.0, 0.05, nu]
        ov, old, odt, oinfo = oracle.llh(b, th, per_block=True)
        _C3[nu] = (b, th, old, odt, oinfo, obs.nugget)
    return _C3[nu]


def _lam_ratio(b, th, i):
    """lambda_min / lambda_max of block i's covariance over its conditioning
    set and its own points (numpy, dense)."""
    acc = np.concatenate([[0], np.cumsum(b.batch_num)])
    p = np.arange(acc[i], acc[i + 1])
    q = np.asarray(b.nbr).reshape(-1, b.cs)[i] if i > 0 else np.zeros(0, np.int64)
    u = np.concatenate([q, p]).astype(np.int64)
    bx, by = np.asarray(b.x), np.asarray(b.y)
    w = np.linalg.eigvalsh(datagen.matern_cov(bx[u], by[u], bx[u], by[u], *th))
    return w[0] / w[-1]


@pytest.mark.parametrize("nu", [1.5, 2.5])
@pytest.mark.parametrize("form", sorted(FORMS))
def test_uneven_shape_vs_oracle(form, nu, monkeypatch):
    b, th, old, odt, oinfo, nug = _c3(nu)
    mp = ((b.cs + np.asarray(b.batch_num[1:]) + 1 + 31) // 32) * 32
    assert mp.max() > 352 and mp.min() >= 416
    eng = _engine(b, form, monkeypatch)
    try:
        v = eng.llh(th)
        ld, dt, info = eng.block_terms()
        cfg = eng.factor_config()
    finally:
        eng.close()
    gfail = np.flatnonzero(info != 0)
    ofail = np.flatnonzero(oinfo != 0)
    both = (info == 0) & (oinfo == 0)
    dld = np.abs(ld - old) / np.maximum(1.0, np.abs(old))
    ddt = np.abs(dt - odt) / np.maximum(1.0, np.abs(odt))
    off = np.flatnonzero(both & ((dld > 1e-8) | (ddt > 1e-8)))
    _say("uneven nu=%g %s %s: %d blocks (padded %d..%d), GPU failed %d, oracle failed %d, "
         "PD in both %d, of which %d differ beyond 1e-8 (sim nugget %g)" % (
             nu, form, cfg, b.batch_count, mp.min(), mp.max(), len(gfail), len(ofail),
             both.sum(), len(off), nug))
    # every block that fails on the GPU or in the oracle is numerically
    # singular in FP64: condition number above 1e12 (a Cholesky of a 400-point
    # block loses all digits near cond ~ 1 / (400 eps) ~ 1e13)
    ratios = {}
    for i in sorted(set(gfail) | set(ofail))[:40]:
        ratios[i] = _lam_ratio(b, th, i)
        assert ratios[i] < 1e-12, (i, ratios[i])
    # blocks positive definite in both agree to 1e-8 unless ill-conditioned
    # (a relative difference d needs cond > d / (100 eps), i.e. > 1e6 here)
    for i in off[:40]:
        r = _lam_ratio(b, th, i)
        assert r < 1e-6, (i, r, dld[i], ddt[i])
    if ratios:
        _say("  failing blocks: lambda_min/lambda_max %.1e .. %.1e" % (min(ratios.values()),
                                                                       max(ratios.values())))
    # and most blocks are well conditioned and agree
    assert both.sum() >= 0.9 * b.batch_count
    if len(gfail) == 0 and len(ofail) == 0:
        # the total against the oracle, within 10x the disagreement of two
        # correct FP64 CPU chains (the oracle and MKL) on this ill-conditioned
        # data (clustered points, cs = 400)
        ov = float(np.sum(-(odt + old + np.asarray(b.batch_num) * math.log(2 * math.pi)) / 2))
        floor = rel(oracle.lapack_llh(b, th, threads=8), ov) if oracle.lapack_available() \
            else 1e-9
        _say("  total: gpu rel %.2e, CPU floor (MKL vs oracle) %.2e" % (rel(v, ov), floor))
        assert rel(v, ov) <= max(10 * floor, LLH_RTOL)


def test_uneven_full_size_properties(monkeypatch):
    """n = 200,000 clustered, bc = 10,000, cs = 400, nu = 3/2: per-block
    sigma^2 law with c = 4, a power of two, so every covariance entry and
    every step of the factorization scales exactly (logdet(4 s2) = logdet(s2)
    + bs log 4 up to the rounding of each log, dot(4 s2) = dot(s2) / 4 and
    the same blocks failing), k evaluations in one call equal to single ones
    bit for bit, and the failing blocks singular in FP64 whichever form
    factors them."""
    n = 200000
    x, y = datagen.clustered_locations(n, 0)
    b0 = bv.partition(x, y, np.zeros(n), 10000, 400, seed=0, device=0)
    obs = datagen.sample_partition(b0, 1.0, 0.05, 1.5, seed=0)
    b = bv.Blocks(b0.x, b0.y, np.asarray(obs, dtype=np.float64), b0.batch_num, b0.nbr, 400)
    bsz = np.asarray(b.batch_num, dtype=np.float64)
    fails = {}
    for form in ("paired", "stepped2"):
        eng = _engine(b, form, monkeypatch)
        try:
            eng.llh([1.0, 0.05, 1.5])
            ld1, dt1, i1 = eng.block_terms()
            eng.llh([4.0, 0.05, 1.5])
            ld3, dt3, i3 = eng.block_terms()
            thetas = np.array([[1.0, 0.05, 1.5], [1.1, 0.05, 1.5], [0.9, 0.05, 1.5]])
            single = np.array([eng.llh(t) for t in thetas])
            many, _ = eng.llh_many(thetas)
        finally:
            eng.close()
        ok = (i1 == 0) & (i3 == 0)
        _say("full size %s: %d blocks, %d failed at s2=1, %d at s2=4" % (
            form, b.batch_count, np.count_nonzero(i1), np.count_nonzero(i3)))
        np.testing.assert_array_equal(i1, i3)
        np.testing.assert_allclose(ld3[ok], ld1[ok] + bsz[ok] * math.log(4.0), rtol=1e-13,
                                   atol=1e-10)
        np.testing.assert_array_equal(dt3[ok], dt1[ok] / 4.0)
        np.testing.assert_array_equal(single, many)
        assert np.count_nonzero(i1) < 0.01 * b.batch_count
        fails[form] = set(np.flatnonzero(i1))
    # the failing blocks are the singular ones, whichever form factors them
    for i in sorted(fails["paired"] ^ fails["stepped2"])[:20]:
        assert _lam_ratio(b, [1.0, 0.05, 1.5], i) < 1e-12


_C4 = {}


def _c4():
    if not _C4:
        n = 4000
        x, y = datagen.uniform_locations(n, 4)
        obs = np.asarray(datagen.sample_vecchia(x, y, 1.0, 0.08, 0.8, cs=400, bc=n // 20,
                                                seed=4))
        b = bv.partition(x, y, obs, n // 20, 150, seed=0)
        th = [1.0, 0.08, 0.8]
        _C4.update(x=x, y=y, obs=obs, b=b, th=th, ov=oracle.llh(b, th))
    return _C4


@pytest.mark.parametrize("form", sorted(FORMS) + ["res"])
def test_general_nu_shape_vs_oracle(form, monkeypatch):
    c = _c4()
    eng = _engine(c["b"], form, monkeypatch)
    try:
        v = eng.llh(c["th"])
        st, _ = eng.status()
        cfg = eng.factor_config()
    finally:
        eng.close()
    _say("general nu %s %s: gpu %.17g oracle %.17g rel %.2e" % (form, cfg, v, c["ov"],
                                                                rel(v, c["ov"])))
    assert st == bv.BV_OK and rel(v, c["ov"]) <= LLH_RTOL


def test_general_nu_mle_vs_scipy(tmp_path):
    """The driver from the reference's start 0.1:0.1:0.1, xtol 1e-4, all three
    parameters free (--ikernel ?:?:?), against scipy's Powell maximum of the
    same likelihood through the C ABI (bounds 0.01..3, as the driver)."""
    from scipy.optimize import minimize
    c = _c4()
    n = len(c["x"])
    datagen.write_xy_csv(str(tmp_path / "xy.csv"), c["x"], c["y"])
    datagen.write_obs_csv(str(tmp_path / "obs.csv"), c["obs"])
    _say("driver MLE n=%d bc=%d cs=150 from 0.1:0.1:0.1" % (n, n // 20))
    r = subprocess.run([DRIVER, "--xy_path", str(tmp_path / "xy.csv"), "--obs_path",
                        str(tmp_path / "obs.csv"), "--num_loc", str(n), "--vecchia_bc",
                        str(n // 20), "--vecchia_cs", "150", "--kernel_init", "0.1:0.1:0.1",
                        "--ikernel", "?:?:?", "--tol", "4", "--omp_threads", "8"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    logs = os.listdir(os.path.join(tmp_path, "log"))
    vals = [v.strip() for v in
            open(os.path.join(tmp_path, "log", logs[0])).read().strip().splitlines()[1].split(",")]
    iters, est, llh = int(vals[0]), [float(v) for v in vals[1:4]], float(vals[4])
    eng = bv.Engine(c["b"])
    try:
        f = lambda t: -eng.llh(np.clip(t, 0.01, 3.0))
        best = None
        for x0 in ([0.1, 0.1, 0.1], est):
            res = minimize(f, x0, method="Powell", bounds=[(0.01, 3.0)] * 3,
                           options={"xtol": 1e-9, "ftol": 1e-13, "maxfev": 20000})
            if best is None or res.fun < best.fun:
                best = res
    finally:
        eng.close()
    _say("driver: %d evaluations, %s llh %.12g; scipy Powell: %s llh %.12g" % (
        iters, est, llh, list(best.x), -best.fun))
    assert iters > 10
    assert abs(llh - (-best.fun)) <= 1e-6 * abs(best.fun), (llh, -best.fun)
    np.testing.assert_allclose(est, best.x, rtol=2e-2)
