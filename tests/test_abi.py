"""The C ABI library: it loads, exports every function include/bv.h declares,
and rejects bad arguments before touching a device.  CPU only (no compute)."""
import ctypes
import os
import re

import numpy as np

import blockvecchia_amd as bv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "bv.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^[a-z_][A-Za-z0-9_ \*]*?\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "for", "while", "return", "sizeof")))


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(bv.LIB_PATH)
    names = _declared()
    assert "bv_create" in names and "llh_Xvecchia_batch" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(bv.EXPORTED) <= set(names)


def test_create_rejects_bad_arguments():
    lib = bv.load_library()
    p = bv.BvProblem()
    ctx = ctypes.c_void_p()
    assert lib.bv_create(None, ctypes.byref(ctx)) == bv.BV_ERR_ARG
    p.n = 0
    assert lib.bv_create(ctypes.byref(p), ctypes.byref(ctx)) == bv.BV_ERR_ARG
    x = np.zeros(4)
    bn = np.array([2, 2], np.int32)
    p.n = 4
    p.batch_count = 2
    p.cs = 2
    p.batch_num = bn.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    p.x = p.y = p.obs = x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    p.kernel = 1
    # cs > 0 with neither neighbour indices nor copies
    assert lib.bv_create(ctypes.byref(p), ctypes.byref(ctx)) == bv.BV_ERR_ARG
    p.kernel = 9
    assert lib.bv_create(ctypes.byref(p), ctypes.byref(ctx)) == bv.BV_ERR_ARG


def test_llh_on_null_context_is_nan():
    lib = bv.load_library()
    th = np.array([1.0, 0.1, 0.5])
    v = lib.bv_llh(3, th.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None, None)
    assert v != v


def test_partition_argument_errors():
    lib = bv.load_library()
    res = bv.BvBlocks()
    assert lib.bv_partition(0, None, None, None, None, 10, 5, 0, 50, 0, 0, 1,
                            ctypes.byref(res)) == bv.BV_ERR_ARG
