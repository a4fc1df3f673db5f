"""The test_dvecchia_batch driver end to end on the GPU: --perf (one
evaluation at the true parameters, y = 0, generated locations) against the
oracle on the same inputs, and a maximum-likelihood run from CSV files against
scipy's L-BFGS-B maximum of the same likelihood."""
import ctypes
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import blockvecchia_amd as bv
from blockvecchia_amd import datagen
import oracle

from conftest import LLH_RTOL

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "blockvecchia_amd", "bin", "test_dvecchia_batch")
_libc = ctypes.CDLL("libc.so.6")
_libc.rand.restype = ctypes.c_int
_libc.srand.argtypes = [ctypes.c_uint]


def generate_xy(n, seed):
    """GenerateXYLoc (src/ckernel.cpp:43-76): jittered grid with glibc rand,
    16-bit quantisation, Morton order -- restated for the check."""
    _libc.srand(seed)
    s = int(math.ceil(math.sqrt(n)))
    x, y = [], []
    for i in range(s):
        for j in range(s):
            if len(x) == n:
                break
            x.append(((i + 1) - 0.5 + (_libc.rand() / (1.0 + 2147483647)) * 0.8 - 0.4) / s)
            y.append(((j + 1) - 0.5 + (_libc.rand() / (1.0 + 2147483647)) * 0.8 - 0.4) / s)
    def spread(v):
        v &= 0xFFFF
        v = (v ^ (v << 8)) & 0x00FF00FF
        v = (v ^ (v << 4)) & 0x0F0F0F0F
        v = (v ^ (v << 2)) & 0x33333333
        v = (v ^ (v << 1)) & 0x55555555
        return v
    def gather(v):
        v &= 0x55555555
        v = (v ^ (v >> 1)) & 0x33333333
        v = (v ^ (v >> 2)) & 0x0F0F0F0F
        v = (v ^ (v >> 4)) & 0x00FF00FF
        v = (v ^ (v >> 8)) & 0x0000FFFF
        return v
    codes = sorted(((spread(int(b * 65535.0 + 0.5) & 0xFFFF) << 1) +
                    spread(int(a * 65535.0 + 0.5) & 0xFFFF)) & 0xFFFFFFFF for a, b in zip(x, y))
    xs = np.array([gather(c) / 65535.0 for c in codes])
    ys = np.array([gather(c >> 1) / 65535.0 for c in codes])
    return xs, ys


def _summary_llh(cwd):
    logs = os.listdir(os.path.join(cwd, "log"))
    assert len(logs) == 1
    lines = open(os.path.join(cwd, "log", logs[0])).read().strip().splitlines()
    assert lines[0].startswith("Iterations, variance, range, smoothness, log-likelihood, time")
    vals = [v.strip() for v in lines[1].split(",")]
    return int(vals[0]), [float(v) for v in vals[1:4]], float(vals[4])


def test_perf_mode_matches_oracle(tmp_path):
    n, bc, cs = 2000, 100, 60
    print("[driver] perf n=%d bc=%d cs=%d" % (n, bc, cs), flush=True)
    r = subprocess.run([DRIVER, "--num_loc", str(n), "--vecchia_bc", str(bc), "--vecchia_cs",
                        str(cs), "--ikernel", "1.5:0.1:0.5", "--kernel",
                        "univariate_matern_stationary_no_nugget", "--perf", "--seed", "0"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    m = re.search(r"Found maximum at f\(1\.5, 0\.1, 0\.5\) = (\S+)", r.stdout)
    assert m, r.stdout
    iters, theta, llh = _summary_llh(tmp_path)
    assert iters == 1 and theta == [1.5, 0.1, 0.5]
    x, y = generate_xy(n, 0)
    blocks = bv.partition(x, y, np.zeros(n), bc, cs, seed=0)
    ref = oracle.llh(blocks, [1.5, 0.1, 0.5])
    assert abs(llh - ref) <= LLH_RTOL * abs(ref), (llh, ref)
    assert abs(float(m.group(1)) - ref) <= 1e-9 * abs(ref)


def test_mle_run_matches_lbfgsb(tmp_path):
    from scipy.optimize import minimize
    n, bc, cs = 1000, 60, 30
    x, y = datagen.uniform_locations(n, 5)
    obs = datagen.sample_dense(x, y, 1.5, 0.1, 0.5, seed=5)
    datagen.write_xy_csv(str(tmp_path / "xy.csv"), x, y)
    datagen.write_obs_csv(str(tmp_path / "obs.csv"), obs)
    print("[driver] mle n=%d bc=%d cs=%d" % (n, bc, cs), flush=True)
    r = subprocess.run([DRIVER, "--xy_path", str(tmp_path / "xy.csv"), "--obs_path",
                        str(tmp_path / "obs.csv"), "--num_loc", str(n), "--vecchia_bc", str(bc),
                        "--vecchia_cs", str(cs), "--kernel_init", "0.5:0.05:0.5",
                        "--ikernel", "1.5:0.1:0.5", "--tol", "7", "--maxiter", "600"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    iters, theta, llh = _summary_llh(tmp_path)
    assert 1 < iters <= 600
    # the same likelihood through the C ABI, maximised by L-BFGS-B
    blocks = bv.partition(x, y, obs, bc, cs, seed=0)
    eng = bv.Engine(blocks)
    ref = minimize(lambda t: -eng.llh(t), [0.5, 0.05, 0.5], method="L-BFGS-B",
                   bounds=[(0.01, 3.0)] * 3, options={"ftol": 1e-15, "gtol": 1e-10,
                                                       "maxiter": 2000})
    eng.close()
    assert llh >= -ref.fun - 1e-7 * abs(ref.fun), (llh, -ref.fun)
    assert abs(llh - (-ref.fun)) <= 1e-6 * abs(ref.fun), (llh, -ref.fun, theta, ref.x)
