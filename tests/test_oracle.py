"""Pins of the CPU oracle (oracle/oracle.c) -- each check is chosen so that a
plausible slip in the restatement (a dropped term, a sign, an index, a
transposed operand) fails it.  CPU only."""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg as sla
from scipy.special import gamma, kv

import oracle
from blockvecchia_amd import Blocks, datagen

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden():
    with open(os.path.join(HERE, "golden", "besselk_golden.json")) as f:
        return json.load(f)["rows"]


def test_besselk_golden():
    """K_nu by quadrature against 40-digit mpmath values."""
    worst = 0.0
    for nu, x, ref in _golden():
        ref = float(ref)
        v = oracle.besselk(nu, x)
        if ref == 0.0:
            assert v == 0.0
            continue
        worst = max(worst, abs(v - ref) / ref)
    assert worst < 2e-15, worst


def test_besselk_closed_forms():
    """K_{1/2}(x) = sqrt(pi/2x) e^-x; K_{3/2} = K_{1/2} (1 + 1/x)."""
    for x in [1e-4, 0.1, 1.0, 3.7, 40.0]:
        k12 = math.sqrt(math.pi / (2 * x)) * math.exp(-x)
        assert abs(oracle.besselk(0.5, x) - k12) <= 4e-16 * k12
        k32 = k12 * (1 + 1 / x)
        assert abs(oracle.besselk(1.5, x) - k32) <= 4e-16 * k32


def test_matern_closed_forms():
    """core_dcmg's con * x^nu K_nu(x) equals the textbook closed forms."""
    rng = np.random.default_rng(3)
    x1, y1, x2, y2 = rng.random(30), rng.random(30), rng.random(20), rng.random(20)
    for nu, f in [(0.5, lambda e: np.exp(-e)), (1.5, lambda e: (1 + e) * np.exp(-e)),
                  (2.5, lambda e: (1 + e + e * e / 3) * np.exp(-e))]:
        A = oracle.matern(x1, y1, x2, y2, [1.7, 0.13, nu])
        d = np.sqrt((x2[None, :] - x1[:, None]) ** 2 + (y2[None, :] - y1[:, None]) ** 2)
        ref = 1.7 * f(d / 0.13)
        np.testing.assert_allclose(A, ref, rtol=2e-15, atol=0)


def test_matern_general_nu_vs_scipy():
    rng = np.random.default_rng(4)
    x1, y1, x2, y2 = rng.random(25), rng.random(25), rng.random(25), rng.random(25)
    nu, s2, beta, ds = 0.8, 1.0, 0.08, 2.0
    A = oracle.matern(x1, y1, x2, y2, [s2, beta, nu], dist_scale=ds)
    e = np.sqrt((x2[None, :] - x1[:, None]) ** 2 + (y2[None, :] - y1[:, None]) ** 2) / beta / ds
    ref = s2 / (2 ** (nu - 1) * gamma(nu)) * e ** nu * kv(nu, e)
    np.testing.assert_allclose(A, ref, rtol=1e-13)


def test_matern_zero_distance_and_earth():
    A = oracle.matern([0.3], [0.4], [0.3], [0.4], [2.5, 0.1, 0.9])
    assert A[0, 0] == 2.5
    # great circle: one degree of latitude = 2 pi R / 360 km
    A = oracle.matern([10.0], [20.0], [11.0], [20.0], [1.0, 100.0, 0.5], metric=1)
    d = 2 * math.pi * 6371.0 / 360.0
    assert abs(A[0, 0] - math.exp(-d / 100.0)) < 1e-12


def test_matern_3d_distance():
    """--3D (use_z): the Euclidean distance over x, y, z, checked with numpy's
    norm through the nu = 1/2 closed form sigma2 exp(-|a - b| / beta)."""
    rng = np.random.default_rng(11)
    P, Q = rng.random((17, 3)), rng.random((13, 3))
    A = oracle.matern(P[:, 0], P[:, 1], Q[:, 0], Q[:, 1], [1.4, 0.3, 0.5], z1=P[:, 2],
                      z2=Q[:, 2])
    d = np.linalg.norm(P[:, None, :] - Q[None, :, :], axis=2)
    np.testing.assert_allclose(A, 1.4 * np.exp(-d / 0.3), rtol=2e-15, atol=0)
    # without z the same points give the planar distance: the z term matters
    A2 = oracle.matern(P[:, 0], P[:, 1], Q[:, 0], Q[:, 1], [1.4, 0.3, 0.5])
    assert np.max(np.abs(A2 - A)) > 1e-3


def _st(P, Q, theta):
    return oracle.matern(P[:, 0], P[:, 1], Q[:, 0], Q[:, 1], theta, kernel=4, z1=P[:, 2],
                         z2=Q[:, 2])


def test_spacetime_equal_times_is_spatial_matern():
    """Kernel 4 with dt = 0: both temporal factors are 1, so the entry is the
    spatial Matern of the (x, y) distance.  Closed forms at nu = 1/2 and 3/2
    with numpy distances; a dropped factor or a wrong exponent in the
    temporal terms would still pass here, the tests below catch those."""
    rng = np.random.default_rng(12)
    P, Q = rng.random((9, 3)), rng.random((11, 3))
    P[:, 2] = 0.25
    Q[:, 2] = 0.25
    d = np.linalg.norm(P[:, None, :2] - Q[None, :, :2], axis=2)
    for nu, f in [(0.5, lambda e: np.exp(-e)), (1.5, lambda e: (1 + e) * np.exp(-e))]:
        th = [1.7, 0.2, nu, 0.8, 0.35, 0.6, 0.4]
        np.testing.assert_allclose(_st(P, Q, th), 1.7 * f(d / 0.2), rtol=4e-15, atol=0)


def test_spacetime_no_separability_terms_is_3d_matern():
    """Kernel 4 with theta5 = theta6 = 0: (1 + |dt|^{2a}/c)^0 = 1, so at any
    dt the entry is the Matern of the full x, y, z distance (the reference's
    distance in this kernel includes z)."""
    rng = np.random.default_rng(13)
    P, Q = rng.random((8, 3)), rng.random((10, 3))
    d = np.linalg.norm(P[:, None, :] - Q[None, :, :], axis=2)
    th = [0.9, 0.25, 0.5, 0.6, 0.7, 0.0, 0.0]
    np.testing.assert_allclose(_st(P, Q, th), 0.9 * np.exp(-d / 0.25), rtol=4e-15, atol=0)


def test_spacetime_entry_by_hand():
    """One entry with dt != 0 against a value computed at 40 digits with
    mpmath from the Gneiting-type form C = s2 psi^-(b + d) M_nu(h / (beta
    psi^(b/2))), psi = 1 + |dt|^(2 alpha) / a, h the x, y, z distance,
    M_{3/2}(u) = (1 + u) e^-u: a = (0.1, 0.2, 0), b = (0.3, 0.25, 0.5),
    theta = (1.3, 0.2, 1.5, 0.7, 0.45, 0.6, 0.3)."""
    P = np.array([[0.1, 0.2, 0.0]])
    Q = np.array([[0.3, 0.25, 0.5]])
    v = _st(P, Q, [1.3, 0.2, 1.5, 0.7, 0.45, 0.6, 0.3])[0, 0]
    assert abs(v - 0.2614444787599308693036548) <= 2e-15 * 0.2614
    # symmetric in the two points, and sigma2 at zero distance and dt = 0
    assert _st(Q, P, [1.3, 0.2, 1.5, 0.7, 0.45, 0.6, 0.3])[0, 0] == v
    assert _st(P, P, [1.3, 0.2, 1.5, 0.7, 0.45, 0.6, 0.3])[0, 0] == 1.3


def test_spacetime_zero_space_distance():
    """Zero distance with dt != 0 cannot happen (the distance includes z);
    the temporal decay alone: same x, y, different z.  psi^-(b + d) times the
    Matern of |dt| / (beta psi^(b/2)) at nu = 1/2."""
    P = np.array([[0.4, 0.4, 0.0]])
    Q = np.array([[0.4, 0.4, 0.3]])
    th = [2.0, 0.5, 0.5, 1.5, 0.5, 0.8, 0.2]
    psi = 1 + 0.3 ** 1.0 / 1.5
    ref = 2.0 * math.exp(-0.3 / (0.5 * psi ** 0.4)) / psi ** 1.0
    assert abs(_st(P, Q, th)[0, 0] - ref) <= 4e-15 * ref


def _blocks_from(x, y, obs, sizes, nbr, cs):
    return Blocks(x, y, obs, np.array(sizes, np.int32), np.array(nbr, np.int32).ravel(), cs)


def test_single_block_is_dense_gp():
    """One block (cs = 0): the exact Gaussian log-likelihood."""
    x, y = datagen.uniform_locations(150, 1)
    obs = datagen.sample_dense(x, y, 1.2, 0.2, 1.5, seed=1)
    b = _blocks_from(x, y, obs, [150], np.zeros(0), 0)
    th = [1.2, 0.2, 1.5]
    v = oracle.llh(b, th)
    S = datagen.matern_cov(x, y, x, y, 1.2, 0.2, 1.5)
    c, low = sla.cho_factor(S, lower=True)
    ref = -0.5 * (obs @ sla.cho_solve((c, low), obs) + 2 * np.log(np.diag(c)).sum()
                  + 150 * math.log(2 * math.pi))
    assert abs(v - ref) < 1e-11 * abs(ref)


def test_two_blocks_full_conditioning_is_exact():
    """Block 1 conditions on all of block 0 (bs0 = cs): the chain telescopes
    to the exact joint likelihood of both blocks."""
    x, y = datagen.uniform_locations(90, 2)
    obs = datagen.sample_dense(x, y, 1.0, 0.15, 0.5, seed=2)
    cs = 60
    nbr = np.concatenate([np.arange(cs), np.arange(cs)[::-1]])  # any order of block 0
    b = _blocks_from(x, y, obs, [60, 30], nbr, cs)
    th = [1.0, 0.15, 0.5]
    v = oracle.llh(b, th)
    ref = oracle.dense_llh(x, y, obs, th)
    S = datagen.matern_cov(x, y, x, y, 1.0, 0.15, 0.5)
    c, low = sla.cho_factor(S, lower=True)
    ref2 = -0.5 * (obs @ sla.cho_solve((c, low), obs) + 2 * np.log(np.diag(c)).sum()
                   + 90 * math.log(2 * math.pi))
    assert abs(v - ref2) < 1e-10 * abs(ref2)
    assert abs(ref - ref2) < 1e-10 * abs(ref2)


def _numpy_chain(b, th):
    """Brute-force restatement with scipy: every block's conditional density
    N(y_B | S12 S22^-1 y_N, S11 - S12 S22^-1 S21)."""
    acc = np.concatenate([[0], np.cumsum(b.batch_num)])
    total = 0.0
    for i in range(len(b.batch_num)):
        sl = slice(acc[i], acc[i + 1])
        xb, yb, ob = b.x[sl], b.y[sl], b.obs[sl]
        S11 = datagen.matern_cov(xb, yb, xb, yb, *th)
        mu = np.zeros(len(xb))
        if i >= 1 and b.cs > 0:
            nb = b.nbr[i * b.cs:(i + 1) * b.cs]
            xn, yn, on = b.x[nb], b.y[nb], b.obs[nb]
            S22 = datagen.matern_cov(xn, yn, xn, yn, *th)
            S21 = datagen.matern_cov(xn, yn, xb, yb, *th)
            K = np.linalg.solve(S22, S21)
            S11 = S11 - S21.T @ K
            mu = K.T @ on
        r = ob - mu
        sign, ld = np.linalg.slogdet(S11)
        total += -0.5 * (r @ np.linalg.solve(S11, r) + ld + len(xb) * math.log(2 * math.pi))
    return total


def test_chain_vs_bruteforce(c1_small):
    import blockvecchia_amd as bv
    x, y, obs = c1_small
    b = bv.partition(x, y, obs, 30, 40, seed=0)
    for th in ([1.5, 0.1, 0.5], [0.9, 0.05, 1.5], [2.0, 0.2, 0.8]):
        v = oracle.llh(b, th)
        ref = _numpy_chain(b, th)
        assert abs(v - ref) < 1e-9 * abs(ref), (th, v, ref)


def test_sigma_scaling_law(c1_small):
    """llh(c sigma2) = -1/2 (Q/c + L + n log c + n log 2pi) where Q, L are
    the quadratic form and log-determinant at sigma2 = 1 (every block)."""
    import blockvecchia_amd as bv
    x, y, obs = c1_small
    b = bv.partition(x, y, obs, 30, 40, seed=0)
    n = len(x)
    _, ld1, dt1, _ = oracle.llh(b, [1.0, 0.1, 0.5], per_block=True)
    Q, L = dt1.sum(), ld1.sum()
    for c in (0.3, 2.7):
        v = oracle.llh(b, [c, 0.1, 0.5])
        ref = -0.5 * (Q / c + L + n * math.log(c) + n * math.log(2 * math.pi))
        assert abs(v - ref) < 1e-11 * abs(ref)


def test_zero_observations_logdet_only(c1_small):
    """With y = 0 (the reference's --perf mode) the quadratic form vanishes."""
    import blockvecchia_amd as bv
    x, y, obs = c1_small
    b = bv.partition(x, y, np.zeros_like(obs), 30, 40, seed=0)
    v, ld, dt, info = oracle.llh(b, [1.5, 0.1, 0.5], per_block=True)
    assert np.all(dt == 0.0) and np.all(info == 0)
    assert abs(v - (-0.5 * (ld.sum() + len(x) * math.log(2 * math.pi)))) < 1e-9


def test_not_positive_definite_reports_info():
    """Duplicate locations without a nugget make S11 singular."""
    x = np.array([0.1, 0.1, 0.5, 0.7])
    y = np.array([0.2, 0.2, 0.5, 0.1])
    b = _blocks_from(x, y, np.ones(4), [4], np.zeros(0), 0)
    v, ld, dt, info = oracle.llh(b, [1.0, 0.1, 0.5], per_block=True)
    assert info[0] == 2 and math.isnan(v)


def test_lapack_chain_equals_oracle(c1_small):
    """The CPU baseline (the reference chain through MKL per block) gives the
    oracle's log-likelihood and per-block terms."""
    import blockvecchia_amd as bv
    if not oracle.lapack_available():
        pytest.skip("MKL (libmkl_rt.so) not available")
    x, y, obs = c1_small
    b = bv.partition(x, y, obs, 30, 40, seed=0)
    # well-conditioned parameters (the data are a nu = 1/2 draw): a smooth
    # misspecified kernel amplifies rounding far past 1e-11
    for th in ([1.5, 0.1, 0.5], [0.9, 0.05, 1.5], [1.1, 0.03, 2.5]):
        v, ld, dt = oracle.lapack_llh(b, th, threads=4, per_block=True)
        ov, old, odt, _ = oracle.llh(b, th, per_block=True)
        assert abs(v - ov) <= 1e-11 * abs(ov), (th, v, ov)
        np.testing.assert_allclose(ld, old, rtol=1e-10, atol=1e-10)
        np.testing.assert_allclose(dt, odt, rtol=1e-10, atol=1e-10)
    assert oracle.lapack_library()
