"""The driver's derivative-free optimizer (blockvecchia_amd/csrc/bv_optim.h)
compiled for the host and checked on functions with known maximisers, and on
the Vecchia log-likelihood (the oracle's, on the CPU) against an independent
optimizer (scipy's L-BFGS-B) on the same objective.  NLopt, the reference's
optimizer, is not available: its estimates are pinned only through "both
optimizers reach the same maximiser within the tolerance"."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

SHIM = r'''
#include "bv_optim.h"
typedef double (*objfn)(int, const double *);
extern "C" int shim_maximize(int n, const double *x0, const double *lb, const double *ub,
                             double xtol, int maxeval, objfn f, double *xout, double *fout,
                             int *evals) {
    std::vector<double> x(x0, x0 + n), l(lb, lb + n), u(ub, ub + n);
    bvopt::Result r = bvopt::maximize(
        [&](const std::vector<double> &t) { return f(n, t.data()); }, x, l, u, xtol, maxeval);
    for (int i = 0; i < n; ++i) xout[i] = r.x[i];
    *fout = r.f;
    *evals = r.evals;
    return r.status;
}
'''

OBJ = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double))


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    d = tmp_path_factory.mktemp("optim")
    src = d / "shim.cpp"
    src.write_text(SHIM)
    so = d / "shim.so"
    subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17",
                    "-I" + os.path.join(ROOT, "blockvecchia_amd", "csrc"), str(src), "-o", str(so)],
                   check=True)
    lib = ctypes.CDLL(str(so))
    dp = ctypes.POINTER(ctypes.c_double)
    lib.shim_maximize.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_double, ctypes.c_int, OBJ,
                                  dp, dp, ctypes.POINTER(ctypes.c_int)]
    lib.shim_maximize.restype = ctypes.c_int
    return lib


def maximize(lib, f, x0, lb, ub, xtol=1e-6, maxeval=2000):
    n = len(x0)
    arr = lambda v: (ctypes.c_double * n)(*v)
    cb = OBJ(lambda m, p: float(f(np.array([p[i] for i in range(m)]))))
    xo = (ctypes.c_double * n)()
    fo = ctypes.c_double()
    ev = ctypes.c_int()
    st = lib.shim_maximize(n, arr(x0), arr(lb), arr(ub), xtol, maxeval, cb, xo, ctypes.byref(fo),
                           ctypes.byref(ev))
    return np.array(xo[:]), fo.value, ev.value, st


def test_concave_quadratic_interior(shim):
    t = np.array([0.7, 0.2, 1.3])
    w = np.array([1.0, 30.0, 0.5])
    f = lambda x: -np.sum(w * (x - t) ** 2) + 0.3 * (x[0] - t[0]) * (x[1] - t[1])
    x, fx, ev, st = maximize(shim, f, [0.01, 0.01, 0.01], [0.01] * 3, [3.0] * 3)
    assert st == 1 and ev < 300
    assert np.allclose(x, t, rtol=1e-4, atol=1e-5), x


def test_maximiser_on_the_bound(shim):
    t = np.array([-1.0, 0.5])
    f = lambda x: -np.sum((x - t) ** 2)
    x, fx, ev, st = maximize(shim, f, [1.0, 1.0], [0.01, 0.01], [3.0, 3.0])
    assert abs(x[0] - 0.01) < 1e-9 and abs(x[1] - 0.5) < 1e-4, x


def test_rosenbrock(shim):
    f = lambda x: -(100.0 * (x[1] - x[0] ** 2) ** 2 + (1.0 - x[0]) ** 2)
    x, fx, ev, st = maximize(shim, f, [-1.2, 1.0], [-2.0, -2.0], [2.0, 2.0], xtol=1e-8,
                             maxeval=5000)
    assert np.allclose(x, [1.0, 1.0], atol=1e-4), (x, ev)


def test_fixed_coordinate_and_single_evaluation(shim):
    calls = []

    def f(x):
        calls.append(x.copy())
        return -np.sum((x - 0.5) ** 2)

    x, fx, ev, st = maximize(shim, f, [0.2, 0.3, 0.0], [0.01, 0.01, 0.0], [3.0, 3.0, 0.0])
    assert all(c[2] == 0.0 for c in calls)
    assert np.allclose(x[:2], [0.5, 0.5], atol=1e-4)
    calls.clear()
    x, fx, ev, st = maximize(shim, f, [0.2, 0.3, 0.0], [0.01, 0.01, 0.0], [3.0, 3.0, 0.0],
                             maxeval=1)
    assert ev == 1 and len(calls) == 1 and np.allclose(calls[0], [0.2, 0.3, 0.0])


def test_failed_evaluations_are_rejected(shim):
    # -inf (a factorization failure) outside a region: the maximiser avoids it
    def f(x):
        if x[0] > 1.0:
            return -np.inf
        return -((x[0] - 0.9) ** 2 + (x[1] - 0.4) ** 2)

    x, fx, ev, st = maximize(shim, f, [0.3, 0.3], [0.01, 0.01], [3.0, 3.0])
    assert np.allclose(x, [0.9, 0.4], atol=1e-4), x


def test_vecchia_mle_agrees_with_lbfgsb(shim):
    """Maximum-likelihood estimate of (sigma^2, range) on the oracle's Vecchia
    likelihood (n = 400, 30 blocks, cs = 20, nu = 1/2 fixed by the bounds):
    the same maximiser as scipy's L-BFGS-B, within the tolerance."""
    from scipy.optimize import minimize

    import blockvecchia_amd as bv
    from blockvecchia_amd import datagen
    import oracle

    x, y = datagen.uniform_locations(400, 3)
    obs = datagen.sample_dense(x, y, 1.5, 0.1, 0.5, seed=3)
    blocks = bv.partition(x, y, obs, 30, 20, seed=0)
    f = lambda t: oracle.llh(blocks, [t[0], t[1], t[2]])
    xo, fo, ev, st = maximize(shim, f, [0.5, 0.05, 0.5], [0.01, 0.01, 0.5], [3.0, 3.0, 0.5],
                              xtol=1e-7, maxeval=400)
    ref = minimize(lambda t: -f([t[0], t[1], 0.5]), [0.5, 0.05], method="L-BFGS-B",
                   bounds=[(0.01, 3.0), (0.01, 3.0)], options={"ftol": 1e-14, "gtol": 1e-9})
    assert abs(fo - (-ref.fun)) <= 1e-8 * abs(ref.fun), (fo, -ref.fun)
    assert np.allclose(xo[:2], ref.x, rtol=2e-4), (xo, ref.x)


def _recorded(f, lb, ub):
    pts = []

    def g(x):
        pts.append(x.copy())
        return f(x)
    return g, pts


@pytest.mark.parametrize("start", [[0.1, 0.1, 0.1], [2.9, 0.05, 1.7], [0.3, 2.5, 0.02]])
def test_every_evaluated_point_in_bounds(shim, start):
    """Rebuilding the interpolation set after a rejected step must start from
    a copy of the best point (the set it is taken from is overwritten):
    every evaluated point stays finite and inside the box, on a curved
    valley that forces rejected steps and rebuilds."""
    lb, ub = [0.01] * 3, [3.0] * 3

    def valley(x):
        # a banana in (x0, x1) with a steep wall in x2
        return -(50.0 * (x[1] - x[0] ** 2) ** 2 + (1.0 - x[0]) ** 2 + 20.0 * (x[2] - 0.8) ** 2
                 + 5.0 * np.sin(7.0 * x[2]) ** 2)
    g, pts = _recorded(valley, lb, ub)
    x, fx, ev, st = maximize(shim, g, start, lb, ub, xtol=1e-6, maxeval=1500)
    P = np.array(pts)
    assert np.isfinite(P).all()
    assert (P >= np.array(lb)).all() and (P <= np.array(ub)).all(), P[(P < lb).any(1) | (P > ub).any(1)][:3]
    assert ev == len(pts)


def test_general_nu_mle_from_default_start(shim):
    """All three parameters free, general smoothness, from the reference's
    default start 0.1:0.1:0.1 (n = 300, 23 blocks, cs = 30, observations
    drawn from the Block Vecchia model at (1.0, 0.08, 0.8)).  The maximum of
    the oracle's likelihood, found by scipy's Nelder-Mead (300 evaluations)
    and Powell (1306) from the same start, both to 1e-7 in x:
    (0.70817772, 0.06900015, 0.78882709), llh = -207.94945712489.  This start
    once made the optimizer cycle on the initial lattice (a rejected step,
    then the same design rebuilt at the same radius, until maxeval)."""
    import blockvecchia_amd as bv
    from blockvecchia_amd import datagen
    import oracle

    n = 300
    x, y = datagen.uniform_locations(n, 0)
    b = bv.partition(x, y, np.zeros(n), n // 13, 30, seed=0)
    b.obs = np.ascontiguousarray(np.asarray(datagen.sample_block_vecchia(b, 1.0, 0.08, 0.8, seed=0)))
    f = lambda t: oracle.llh(b, [t[0], t[1], t[2]])
    xo, fo, ev, st = maximize(shim, f, [0.1, 0.1, 0.1], [0.01] * 3, [3.0] * 3, xtol=1e-7,
                              maxeval=1000)
    assert st == 1 and ev < 1000, (st, ev, xo)
    assert abs(fo - (-207.94945712489)) <= 1e-9 * 207.95, fo
    assert np.allclose(xo, [0.70817772, 0.06900015, 0.78882709], rtol=2e-4), xo
