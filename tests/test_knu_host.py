"""The engine's K_nu (blockvecchia_amd/csrc/bv_matern.h, Temme series +
backward-recurrence CF2) compiled for the host and checked against the
40-digit mpmath table and the oracle's independent quadrature.  CPU only;
the same source runs on the device (checked in test_gpu_parity.py)."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SHIM = r'''
#include "bv_matern.h"
extern "C" double shim_besselk(double nu, double x) { return bv_besselk(nu, x); }
extern "C" double shim_xnu_knu_tab(double nu, double x) {
    static bv_knu_tab t;
    bv_knu_tab_fill(nu, &t, 0, 1);
    return bv_xnu_knu_tab(x, t);
}
extern "C" int shim_cheb_size() { return BV_CHEB_SIZE; }
extern "C" void shim_cheb_table(double nu, double *tab) {
    double gv[BV_CHEB_N];
    for (int j = 0; j < BV_CHEB_NJ; ++j) {
        for (int k = 0; k < BV_CHEB_N; ++k) gv[k] = bv_cheb_node_g(nu, j, k);
        for (int t = 0; t < BV_CHEB_N; ++t) tab[j * BV_CHEB_N + t] = bv_cheb_coef(gv, t);
        bv_cheb_mono(tab + j * BV_CHEB_N);
    }
}
extern "C" double shim_xnu_knu_cheb(double nu, double x, const double *tab) {
    return bv_xnu_knu_cheb(x, tab, nu);
}
extern "C" double shim_xnu_knu_cheb_t(double nu, double x, const double *tab) {
    static double et[2 * BV_EXP_TAB];
    static int init = 0;
    if (!init) {
        for (int i = 0; i < 2 * BV_EXP_TAB; ++i) et[i] = bv_exp_tab_entry(i);
        init = 1;
    }
    return bv_xnu_knu_cheb_t(x, tab, nu, et);
}
extern "C" double shim_exp_tab_entry(int i) { return bv_exp_tab_entry(i); }
extern "C" void shim_expneg_tab(const double *e, double *out, long n) {
    double tab[2 * BV_EXP_TAB];
    for (int i = 0; i < 2 * BV_EXP_TAB; ++i) tab[i] = bv_exp_tab_entry(i);
    for (long i = 0; i < n; ++i) out[i] = bv_expneg_tab(e[i], tab);
}
extern "C" void shim_expl(const double *e, long double *out, long n) {
    for (long i = 0; i < n; ++i) out[i] = expl(-(long double)e[i]);
}
extern "C" double shim_cov(int kind, double x1, double y1, double x2, double y2, double s2,
                           double beta, double nu, double con) {
    bv_kparams kp = {};
    kp.sigma2 = s2; kp.beta = beta; kp.nu = nu; kp.con = con; kp.dist_scale = 1.0;
    kp.inv_scale = 1.0 / beta;
    switch (kind) {
    case 0: return bv_cov<BV_KIND_NU05>(x1, y1, 0, x2, y2, 0, kp);
    case 1: return bv_cov<BV_KIND_NU15>(x1, y1, 0, x2, y2, 0, kp);
    case 2: return bv_cov<BV_KIND_NU25>(x1, y1, 0, x2, y2, 0, kp);
    default: {
        /* the general kind reads the evaluation's x^nu K_nu table */
        static double tab[BV_CHEB_SIZE];
        static double tab_nu = -1.0;
        if (tab_nu != nu) {
            shim_cheb_table(nu, tab);
            tab_nu = nu;
        }
        kp.cheb = tab;
        return bv_cov<BV_KIND_GENERAL>(x1, y1, 0, x2, y2, 0, kp);
    }
    }
}
'''


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    d = tmp_path_factory.mktemp("knu")
    src = d / "shim.cpp"
    src.write_text(SHIM)
    so = d / "shim.so"
    subprocess.run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
                    "-Wno-unknown-pragmas", "-I" + os.path.join(ROOT, "blockvecchia_amd", "csrc"),
                    str(src), "-o", str(so)], check=True)
    lib = ctypes.CDLL(str(so))
    lib.shim_besselk.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.shim_besselk.restype = ctypes.c_double
    lib.shim_xnu_knu_tab.argtypes = [ctypes.c_double, ctypes.c_double]
    lib.shim_xnu_knu_tab.restype = ctypes.c_double
    lib.shim_cheb_size.restype = ctypes.c_int
    lib.shim_cheb_table.argtypes = [ctypes.c_double, ctypes.c_void_p]
    lib.shim_cheb_table.restype = None
    lib.shim_xnu_knu_cheb.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
    lib.shim_xnu_knu_cheb.restype = ctypes.c_double
    lib.shim_xnu_knu_cheb_t.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
    lib.shim_xnu_knu_cheb_t.restype = ctypes.c_double
    lib.shim_exp_tab_entry.argtypes = [ctypes.c_int]
    lib.shim_exp_tab_entry.restype = ctypes.c_double
    lib.shim_expneg_tab.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    lib.shim_expneg_tab.restype = None
    lib.shim_expl.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    lib.shim_expl.restype = None
    lib.shim_cov.argtypes = [ctypes.c_int] + [ctypes.c_double] * 8
    lib.shim_cov.restype = ctypes.c_double
    return lib


def test_engine_besselk_golden(shim):
    with open(os.path.join(HERE, "golden", "besselk_golden.json")) as f:
        rows = json.load(f)["rows"]
    worst = 0.0
    for nu, x, ref in rows:
        ref = float(ref)
        v = shim.shim_besselk(nu, x)
        if ref == 0.0:
            assert v == 0.0
            continue
        worst = max(worst, abs(v - ref) / ref)
    assert worst < 1e-14, worst


def test_engine_besselk_vs_oracle_random(shim):
    rng = np.random.default_rng(11)
    for _ in range(400):
        nu = rng.uniform(0.01, 3.0)
        x = 10 ** rng.uniform(-7, np.log10(650))
        a, b = shim.shim_besselk(nu, x), oracle.besselk(nu, x)
        assert abs(a - b) <= 1e-14 * b, (nu, x, a, b)


def test_engine_closed_forms_vs_general(shim):
    """Closed forms for nu = 1/2, 3/2, 5/2 agree with con x^nu K_nu(x)."""
    from scipy.special import gamma
    rng = np.random.default_rng(12)
    for kind, nu in [(0, 0.5), (1, 1.5), (2, 2.5)]:
        con = 1.3 / (2 ** (nu - 1) * gamma(nu))
        for _ in range(50):
            p = rng.random(4)
            a = shim.shim_cov(kind, *p, 1.3, 0.07, nu, con)
            b = shim.shim_cov(3, *p, 1.3, 0.07, nu, con)
            assert abs(a - b) <= 1e-14 * abs(b)


def test_tabulated_knu_golden_and_random(shim):
    """x^nu K_nu(x) from the nu-tables (the path the covariance generation
    uses for general nu) against the mpmath table and the oracle."""
    import math
    with open(os.path.join(HERE, "golden", "besselk_golden.json")) as f:
        rows = json.load(f)["rows"]
    worst = 0.0
    for nu, x, ref in rows:
        ref = float(ref)
        if ref == 0.0 or nu == 0.0:
            continue
        v = shim.shim_xnu_knu_tab(nu, x)
        r = math.exp(nu * math.log(x)) * ref
        if r == 0.0 or not math.isfinite(r):
            continue
        worst = max(worst, abs(v - r) / r)
    assert worst < 2e-14, worst
    rng = np.random.default_rng(21)
    for _ in range(400):
        nu = rng.uniform(0.01, 3.0)
        x = 10 ** rng.uniform(-7, np.log10(650))
        a = shim.shim_xnu_knu_tab(nu, x)
        b = math.exp(nu * math.log(x)) * oracle.besselk(nu, x)
        assert abs(a - b) <= 2e-14 * b, (nu, x, a, b)


def _cheb(shim, nu, cache={}):
    if nu not in cache:
        tab = np.zeros(shim.shim_cheb_size())
        shim.shim_cheb_table(nu, tab.ctypes.data)
        cache[nu] = tab
    return cache[nu]


CHEB_FNS = ["shim_xnu_knu_cheb", "shim_xnu_knu_cheb_t"]


@pytest.mark.parametrize("fn", CHEB_FNS)
def test_chebyshev_knu_golden_and_random(shim, fn):
    """x^nu K_nu(x) from the per-nu Chebyshev table (the path every general-nu
    entry takes: 14 terms on each half-octave of [2^-24, 2^10), coefficients
    from the accurate routine at the Chebyshev nodes) against the 40-digit
    mpmath table and, on random (nu, x), the oracle's quadrature K_nu.  The
    table's coefficients here are built on the host by the same functions the
    device's bv_cheb_build_kernel calls.  Both forms: e^-x by exp
    (bv_xnu_knu_cheb) and by the table-driven exponential (bv_xnu_knu_cheb_t,
    the factorization kernels' planar entries)."""
    knu = getattr(shim, fn)
    import math
    with open(os.path.join(HERE, "golden", "besselk_golden.json")) as f:
        rows = json.load(f)["rows"]
    worst, n = 0.0, 0
    for nu, x, ref in rows:
        ref = float(ref)
        if ref == 0.0 or nu == 0.0:
            continue
        r = math.exp(nu * math.log(x)) * ref
        if r == 0.0 or not math.isfinite(r) or r < 1e-300:
            continue
        v = knu(nu, x, _cheb(shim, nu).ctypes.data)
        worst = max(worst, abs(v - r) / r)
        n += 1
    assert n > 300
    assert worst < 2e-14, worst
    rng = np.random.default_rng(23)
    for _ in range(300):
        nu = float(rng.uniform(0.01, 3.0))
        x = float(10 ** rng.uniform(-9, np.log10(650)))
        tab = _cheb(shim, nu)
        a = knu(nu, x, tab.ctypes.data)
        b = math.exp(nu * math.log(x)) * oracle.besselk(nu, x)
        assert abs(a - b) <= 2e-14 * b, (nu, x, a, b)


@pytest.mark.parametrize("fn", CHEB_FNS)
def test_chebyshev_knu_half_integer_closed_forms(shim, fn):
    """For nu = 1/2, 3/2, 5/2 the table must reproduce the elementary forms
    x^nu K_nu(x) = sqrt(pi/2) e^-x p(x), p = 1, 1 + x, 3 + 3x + x^2 (a
    dropped or misplaced node or coefficient fails this on whole
    half-octaves)."""
    import math
    forms = {0.5: lambda x: 1.0, 1.5: lambda x: 1.0 + x, 2.5: lambda x: 3.0 + 3.0 * x + x * x}
    for nu, p in forms.items():
        tab = _cheb(shim, nu)
        for x in np.geomspace(2.0 ** -24, 700.0, 997):
            x = float(x)
            ref = math.sqrt(math.pi / 2) * math.exp(-x) * p(x)
            if ref < 1e-300:
                continue
            v = getattr(shim, fn)(nu, x, tab.ctypes.data)
            assert abs(v - ref) <= 1e-14 * ref, (nu, x, v, ref)


def test_exp_table_is_2_pow_j_over_64(shim):
    """The table of the register-resident kernel's exponential: T_j is
    2^(j/64) rounded to nearest and T_j + L_j equals 2^(j/64) to ~2^-106,
    against 40-digit mpmath values."""
    import mpmath as mp
    mp.mp.dps = 40
    for j in range(64):
        v = mp.power(2, mp.mpf(j) / 64)
        hi, lo = shim.shim_exp_tab_entry(j), shim.shim_exp_tab_entry(64 + j)
        assert hi == float(v)
        assert abs((mp.mpf(hi) + mp.mpf(lo) - v) / v) < mp.mpf(2) ** -104


def test_expneg_tab_within_one_ulp(shim):
    """e^-e by the table-driven method against the host's extended-precision
    expl over [0, 745] (closed-form Matern arguments d / range), sampled
    densely near 0 and across the whole range; exact at 0, 0 beyond the
    smallest subnormal."""
    rng = np.random.default_rng(0)
    e = np.concatenate([np.linspace(0.0, 1.0, 200001), rng.uniform(0.0, 745.0, 400000),
                        10.0 ** rng.uniform(-12, 0, 100000)])
    n = e.size
    got = np.empty(n)
    ref = np.empty(n, dtype=np.longdouble)
    shim.shim_expneg_tab(e.ctypes.data_as(ctypes.c_void_p), got.ctypes.data_as(ctypes.c_void_p), n)
    shim.shim_expl(e.ctypes.data_as(ctypes.c_void_p), ref.ctypes.data_as(ctypes.c_void_p), n)
    normal = ref > np.longdouble(2.3e-308)
    r = ref[normal].astype(np.float64)
    ulp = np.spacing(r)
    err = np.abs((got[normal].astype(np.longdouble) - ref[normal]) / ulp.astype(np.longdouble))
    assert float(err.max()) < 1.0, float(err.max())
    one = np.empty(1)
    shim.shim_expneg_tab(np.zeros(1).ctypes.data_as(ctypes.c_void_p),
                         one.ctypes.data_as(ctypes.c_void_p), 1)
    assert one[0] == 1.0
    big = np.array([746.0, 800.0, 1e4])
    out = np.empty(3)
    shim.shim_expneg_tab(big.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), 3)
    assert np.all(out == 0.0)
