"""CPU oracle of the Block Vecchia log-likelihood (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
import this package, as the checker.  The product (blockvecchia_amd, libbv)
never imports or links it.

  oracle.c      the likelihood chain op by op (include/llh_Xvecchia_batch.h),
                core_dcmg (src/ckernel.cpp:1057-1093), K_nu by quadrature
  partition.py  the host pre-processing of src/test_Xvecchia_batch.cpp:240-404
                restated in numpy / plain loops

Parity pins (tests/test_oracle.py): K_nu against mpmath golden values and
closed forms; the chain against the exact dense Gaussian likelihood for one
block, for two blocks whose conditioning set is the whole first block, the
sigma^2 scaling law and a brute-force numpy restatement on small inputs.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from . import build
            build.build_oracle()
        lib = ctypes.CDLL(LIB)
        lib.oracle_besselk.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.oracle_besselk.restype = ctypes.c_double
        lib.oracle_matern.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp,
                                      _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]
        lib.oracle_matern.restype = ctypes.c_int
        lib.oracle_llh.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp,
                                   _dp, _dp, _ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_int, _dp, _dp, _ip]
        lib.oracle_llh.restype = ctypes.c_double
        lib.oracle_dense_llh.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double]
        lib.oracle_dense_llh.restype = ctypes.c_double
        _lib = lib
    return _lib


def _d(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


def besselk(nu, x):
    return _load().oracle_besselk(float(nu), float(x))


def matern(x1, y1, x2, y2, theta, kernel=1, metric=0, dist_scale=1.0, z1=None, z2=None):
    x1, y1, x2, y2 = [np.ascontiguousarray(a, dtype=np.float64) for a in (x1, y1, x2, y2)]
    z1 = None if z1 is None else np.ascontiguousarray(z1, dtype=np.float64)
    z2 = None if z2 is None else np.ascontiguousarray(z2, dtype=np.float64)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    m, n = len(x1), len(x2)
    out = np.empty(m * n)
    _load().oracle_matern(m, n, _d(x1), _d(y1), _d(z1), _d(x2), _d(y2), _d(z2), _d(th), kernel,
                          metric, dist_scale, out.ctypes.data_as(_dp))
    return out.reshape(n, m).T.copy()


def llh(blocks, theta, kernel=1, metric=0, time_flag=False, dist_scale=1.0, threads=0,
        per_block=False):
    """Reference chain on the CPU.  blocks: anything with x, y, z, obs,
    batch_num, nbr, cs (blockvecchia_amd.Blocks)."""
    b = blocks
    B = len(b.batch_num)
    bn = np.ascontiguousarray(b.batch_num, dtype=np.int32)
    nbr = np.ascontiguousarray(b.nbr, dtype=np.int32) if b.cs > 0 else np.zeros(1, np.int32)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    ld, dt, info = np.empty(B), np.empty(B), np.empty(B, np.int32)
    keep = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
            for a in (b.x, b.y, b.z, b.obs)]
    v = _load().oracle_llh(len(b.x), B, b.cs, bn.ctypes.data_as(_ip), _d(keep[0]), _d(keep[1]),
                           _d(keep[2]), _d(keep[3]), nbr.ctypes.data_as(_ip), _d(th), kernel,
                           metric, 1 if time_flag else 0, dist_scale, threads,
                           ld.ctypes.data_as(_dp), dt.ctypes.data_as(_dp),
                           info.ctypes.data_as(_ip))
    if per_block:
        return v, ld, dt, info
    return v


def dense_llh(x, y, obs, theta, kernel=1, metric=0, dist_scale=1.0, z=None):
    return _load().oracle_dense_llh(len(x), _d(x), _d(y), _d(z), _d(obs), _d(theta), kernel,
                                    metric, dist_scale)


# ---- the LAPACK (MKL) chain: the CPU baseline ------------------------------
LAPACK_LIB = os.path.join(HERE, "liblapackchain.so")
_lap = None


def lapack_available():
    """True when liblapackchain.so loads and finds MKL (libmkl_rt)."""
    global _lap
    if _lap is None:
        if not os.path.exists(LAPACK_LIB):
            from . import build
            build.build_oracle()
        lib = ctypes.CDLL(LAPACK_LIB)
        lib.lapack_chain_open.argtypes = [ctypes.c_char_p]
        lib.lapack_chain_open.restype = ctypes.c_int
        lib.lapack_chain_library.argtypes = []
        lib.lapack_chain_library.restype = ctypes.c_char_p
        lib.lapack_chain_llh.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, _ip, _dp,
                                         _dp, _dp, _ip, _dp, ctypes.c_int, _dp, _dp]
        lib.lapack_chain_llh.restype = ctypes.c_double
        _lap = lib
    return _lap.lapack_chain_open(os.environ.get("BV_MKL_PATH", "").encode()) == 0


def lapack_library():
    return _lap.lapack_chain_library().decode() if lapack_available() else None


def lapack_llh(blocks, theta, threads=1, per_block=False):
    """The reference chain through MKL's dpotrf/dtrsm/dgemm/dnrm2 per block
    (OpenMP over blocks, one MKL thread each); closed-form nu only."""
    if not lapack_available():
        raise OSError("MKL (libmkl_rt.so) not found")
    b = blocks
    B = len(b.batch_num)
    bn = np.ascontiguousarray(b.batch_num, dtype=np.int32)
    nbr = np.ascontiguousarray(b.nbr, dtype=np.int32) if b.cs > 0 else np.zeros(1, np.int32)
    th = np.ascontiguousarray(theta, dtype=np.float64)
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (b.x, b.y, b.obs)]
    ld, dt = np.empty(B), np.empty(B)
    v = _lap.lapack_chain_llh(len(b.x), B, b.cs, bn.ctypes.data_as(_ip), _d(keep[0]),
                              _d(keep[1]), _d(keep[2]), nbr.ctypes.data_as(_ip), _d(th),
                              int(threads), ld.ctypes.data_as(_dp), dt.ctypes.data_as(_dp))
    return (v, ld, dt) if per_block else v
