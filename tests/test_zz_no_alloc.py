"""bv_llh_noalloc: the evaluation entry point that allocates no device memory
(its buffers live in the library's fixed scratch array) against bv_llh on a
context fixed to the same factorization form (the fused kernel), through the
C ABI: equal bit for bit on two calls in a row, the second on a smaller
dataset, and for general nu; a dataset too large for the scratch is refused
with BV_ERR_SCRATCH.  The dataset lives in device memory owned by the caller
(allocated here through the HIP runtime)."""
import ctypes
import sys

import numpy as np
import pytest

import blockvecchia_amd as bv
from blockvecchia_amd import datagen

pytestmark = pytest.mark.gpu


def _say(msg):
    print("[noalloc] " + msg, flush=True)
    sys.stdout.flush()


class _DeviceArrays:
    """Device copies made with the HIP runtime through ctypes (the caller's
    own allocations; the entry point under test allocates nothing)."""

    def __init__(self):
        try:
            self.hip = ctypes.CDLL("libamdhip64.so")
        except OSError:
            self.hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def copy(self, a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(ptr), max(a.nbytes, 8)) == 0
        self.ptrs.append(ptr)
        assert self.hip.hipMemcpy(ptr, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return ptr.value

    def free(self):
        for ptr in self.ptrs:
            self.hip.hipFree(ptr)
        self.ptrs = []


def _device_problem(blocks, mem, keep):
    """bv_problem with device copies of the dataset and the block sizes on
    the host."""
    p = bv.BvProblem()
    p.n = blocks.n
    p.batch_count = blocks.batch_count
    p.cs = blocks.cs
    bn = np.ascontiguousarray(blocks.batch_num, dtype=np.int32)
    keep.append(bn)
    p.batch_num = bn.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    dp = ctypes.POINTER(ctypes.c_double)
    p.x = ctypes.cast(mem.copy(blocks.x, np.float64), dp)
    p.y = ctypes.cast(mem.copy(blocks.y, np.float64), dp)
    p.obs = ctypes.cast(mem.copy(blocks.obs, np.float64), dp)
    p.z = None
    if blocks.cs > 0:
        p.nbr = ctypes.cast(mem.copy(blocks.nbr, np.int32), ctypes.POINTER(ctypes.c_int32))
    p.kernel = 1
    p.distance_metric = 0
    p.time_flag = 0
    p.dist_scale = 1.0
    p.device = 0
    return p


def _noalloc(p, theta):
    lib = bv.load_library()
    th = np.ascontiguousarray(theta, dtype=np.float64)
    out = ctypes.c_double(0.0)
    bad = ctypes.c_int(-2)
    rc = lib.bv_llh_noalloc(ctypes.byref(p), len(th),
                            th.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            ctypes.byref(out), ctypes.byref(bad))
    return rc, out.value, bad.value


def _alone(blocks, theta, monkeypatch):
    monkeypatch.setenv("BV_PATH", "fused")
    monkeypatch.setenv("BV_FACTOR_KERNEL", "4w")
    eng = bv.Engine(blocks)
    try:
        return eng.llh(theta)
    finally:
        eng.close()


def test_noalloc_equals_bv_llh_two_calls(monkeypatch):
    x, y = datagen.uniform_locations(2000, 0)
    obs = datagen.sample_dense(x, y, 1.5, 0.1, 0.5, seed=0)
    big = bv.partition(x, y, obs, 100, 60, seed=0)
    xs, ys = datagen.uniform_locations(600, 1)
    obs_s = datagen.sample_dense(xs, ys, 1.0, 0.1, 1.5, seed=1)
    small = bv.partition(xs, ys, obs_s, 30, 40, seed=0)
    keep = []
    mem = _DeviceArrays()
    cases = [(big, [1.5, 0.1, 0.5]), (small, [1.0, 0.1, 1.5]), (small, [1.0, 0.08, 0.8]),
             (big, [1.3, 0.12, 0.5])]
    try:
        for blocks, th in cases:
            _say("case: n=%d B=%d cs=%d theta=%s" % (blocks.n, blocks.batch_count, blocks.cs,
                                                      th))
            p = _device_problem(blocks, mem, keep)
            rc, v, bad = _noalloc(p, th)
            ref = _alone(blocks, th, monkeypatch)
            _say("  noalloc %.17g bv_llh %.17g" % (v, ref))
            assert rc == bv.BV_OK and bad == -1, (rc, bad)
            assert v == ref, (v, ref)
    finally:
        mem.free()


def test_noalloc_rejects_oversize_and_bad_arguments():
    n = 30000
    x, y = datagen.uniform_locations(n, 2)
    obs = np.random.default_rng(2).standard_normal(n)
    blocks = bv.partition(x, y, obs, n // 13, 150, seed=0, device=0)
    keep = []
    mem = _DeviceArrays()
    try:
        p = _device_problem(blocks, mem, keep)
        _say("case: oversize n=%d cs=150" % n)
        rc, v, bad = _noalloc(p, [1.0, 0.1, 0.5])
        assert rc == bv.BV_ERR_SCRATCH and v != v
        rc, v, bad = _noalloc(p, [1.0, -0.1, 0.5])
        assert rc == bv.BV_ERR_ARG
    finally:
        mem.free()
    lib = bv.load_library()
    assert lib.bv_llh_noalloc(None, 3, None, None, None) == bv.BV_ERR_ARG
