"""blockvecchia_amd -- MI355X Block Vecchia log-likelihood engine.

Python binding of libbv.so (include/bv.h) with ctypes.  The engine itself is
HIP/C++; this module only moves arrays across the C ABI.  Loading fails
loudly when the library has not been built (python -m blockvecchia_amd.build).

Mirrors the reference's objects:
  Engine(...)             llh_data + the buffers the driver preallocates
                          (src/test_Xvecchia_batch.cpp:459-689)
  Engine.llh(theta)       llh_Xvecchia_batch(n, theta, grad, &data)
                          (include/llh_Xvecchia_batch.h:22-522)
  partition(...)          the host pre-processing of
                          src/test_Xvecchia_batch.cpp:240-404
  matern(...)             core_dcmg (src/ckernel.cpp:1057-1093) on the device
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libbv.so")

BV_OK = 0
BV_ERR_NOT_PD = 1
BV_ERR_ARG = -1
BV_ERR_HIP = -2
BV_ERR_NO_DEVICE = -3
BV_ERR_UNSUPPORTED = -4
BV_ERR_BUSY = -5
BV_ERR_SCRATCH = -6

KERNELS = {
    "univariate_matern_stationary_no_nugget": 1,
    "univariate_powexp_stationary_no_nugget": 2,
    "univariate_powexp_nugget_stationary": 3,
    "univariate_matern_spacetime_stationary": 4,
}
NPARAMS = {1: 3, 2: 3, 3: 4, 4: 7}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)


class BvProblem(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("batch_count", ctypes.c_int32),
        ("cs", ctypes.c_int32),
        ("batch_num", _ip),
        ("x", _dp), ("y", _dp), ("z", _dp),
        ("obs", _dp),
        ("nbr", _ip),
        ("con_x", _dp), ("con_y", _dp), ("con_z", _dp), ("con_obs", _dp),
        ("kernel", ctypes.c_int32),
        ("distance_metric", ctypes.c_int32),
        ("time_flag", ctypes.c_int32),
        ("dist_scale", ctypes.c_double),
        ("device", ctypes.c_int32),
        ("verbose", ctypes.c_int32),
    ]


class BvBlocks(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("batch_count", ctypes.c_int32),
        ("cs", ctypes.c_int32),
        ("first_cluster_count", ctypes.c_int32),
        ("nclusters", ctypes.c_int32),
        ("x", _dp), ("y", _dp), ("z", _dp), ("obs", _dp),
        ("batch_num", _ip),
        ("nbr", _ip),
        ("perm", _lp),
        ("cx", _dp), ("cy", _dp), ("cz", _dp),
    ]


EXPORTED = [
    "bv_create", "bv_llh", "llh_Xvecchia_batch", "bv_llh_many", "bv_block_terms",
    "bv_last_status", "bv_last_kernel_ms", "bv_iterations", "bv_destroy", "bv_matern",
    "bv_partition", "bv_blocks_free", "bv_device_selftest", "bv_fp64_pipe_probe",
    "bv_trace_reset", "bv_trace_read",
    "bv_server_create", "bv_server_attach", "bv_server_submit", "bv_server_poll",
    "bv_server_pending", "bv_server_destroy", "bv_factor_config", "bv_set_graph",
    "bv_partition_gpu", "bv_factor_candidates", "bv_llh_noalloc", "bv_simulate",
]

_lib = None


def load_library(path=None):
    """Load libbv.so (raises OSError when it is missing)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError("libbv.so not built: run `python -m blockvecchia_amd.build` (looked for %s)" % p)
    lib = ctypes.CDLL(p)
    lib.bv_create.argtypes = [ctypes.POINTER(BvProblem), ctypes.POINTER(ctypes.c_void_p)]
    lib.bv_create.restype = ctypes.c_int
    lib.bv_llh.argtypes = [ctypes.c_uint, _dp, _dp, ctypes.c_void_p]
    lib.bv_llh.restype = ctypes.c_double
    lib.llh_Xvecchia_batch.argtypes = [ctypes.c_uint, _dp, _dp, ctypes.c_void_p]
    lib.llh_Xvecchia_batch.restype = ctypes.c_double
    lib.bv_llh_many.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                ctypes.POINTER(ctypes.c_int)]
    lib.bv_llh_many.restype = ctypes.c_int
    lib.bv_block_terms.argtypes = [ctypes.c_void_p, _dp, _dp, _ip]
    lib.bv_block_terms.restype = ctypes.c_int
    lib.bv_last_status.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    lib.bv_last_status.restype = ctypes.c_int
    lib.bv_last_kernel_ms.argtypes = [ctypes.c_void_p]
    lib.bv_last_kernel_ms.restype = ctypes.c_double
    lib.bv_iterations.argtypes = [ctypes.c_void_p]
    lib.bv_iterations.restype = ctypes.c_int
    lib.bv_destroy.argtypes = [ctypes.c_void_p]
    lib.bv_destroy.restype = None
    lib.bv_matern.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                              ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int]
    lib.bv_matern.restype = ctypes.c_int
    lib.bv_partition.argtypes = [ctypes.c_int64, _dp, _dp, _dp, _dp, ctypes.c_int32,
                                 ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                 ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(BvBlocks)]
    lib.bv_partition.restype = ctypes.c_int
    lib.bv_partition_gpu.argtypes = [ctypes.c_int32] + list(lib.bv_partition.argtypes)
    lib.bv_partition_gpu.restype = ctypes.c_int
    lib.bv_blocks_free.argtypes = [ctypes.POINTER(BvBlocks)]
    lib.bv_blocks_free.restype = None
    lib.bv_device_selftest.argtypes = [ctypes.c_int]
    lib.bv_device_selftest.restype = ctypes.c_int
    lib.bv_fp64_pipe_probe.argtypes = [ctypes.c_int, _dp]
    lib.bv_fp64_pipe_probe.restype = ctypes.c_int
    lib.bv_factor_config.argtypes = [ctypes.c_void_p, _ip]
    lib.bv_factor_config.restype = ctypes.c_int
    lib.bv_set_graph.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.bv_set_graph.restype = ctypes.c_int
    lib.bv_factor_candidates.argtypes = [_ip, ctypes.c_int]
    lib.bv_factor_candidates.restype = ctypes.c_int
    lib.bv_llh_noalloc.argtypes = [ctypes.POINTER(BvProblem), ctypes.c_uint, _dp, _dp,
                                   ctypes.POINTER(ctypes.c_int)]
    lib.bv_llh_noalloc.restype = ctypes.c_int
    lib.bv_simulate.argtypes = [ctypes.c_void_p, ctypes.c_uint, _dp, _dp, ctypes.c_double, _dp,
                                ctypes.POINTER(ctypes.c_int)]
    lib.bv_simulate.restype = ctypes.c_int
    lib.bv_server_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_void_p)]
    lib.bv_server_create.restype = ctypes.c_int
    lib.bv_server_attach.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    lib.bv_server_attach.restype = ctypes.c_int
    lib.bv_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp,
                                     ctypes.c_int64]
    lib.bv_server_submit.restype = ctypes.c_int
    lib.bv_server_poll.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int64), _dp, ctypes.POINTER(ctypes.c_int)]
    lib.bv_server_poll.restype = ctypes.c_int
    lib.bv_server_pending.argtypes = [ctypes.c_void_p]
    lib.bv_server_pending.restype = ctypes.c_int
    lib.bv_server_destroy.argtypes = [ctypes.c_void_p]
    lib.bv_server_destroy.restype = None
    if path is None:
        _lib = lib
    return lib


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Blocks:
    """Block-ordered dataset: what the reference keeps after its pre-processing
    (locations_new, h_obs_new, batchNum, conditioning sets as indices)."""

    def __init__(self, x, y, obs, batch_num, nbr, cs, z=None, perm=None, cx=None, cy=None,
                 first_cluster_count=1):
        self.x, self.y, self.z = _f64(x), _f64(y), _f64(z)
        self.obs = _f64(obs)
        self.batch_num = np.ascontiguousarray(batch_num, dtype=np.int32)
        self.nbr = np.ascontiguousarray(nbr, dtype=np.int32)
        self.cs = int(cs)
        self.perm = perm
        self.cx, self.cy = cx, cy
        self.first_cluster_count = first_cluster_count

    @property
    def n(self):
        return len(self.x)

    @property
    def batch_count(self):
        return len(self.batch_num)


def partition(x, y, obs, bc, cs, seed=0, z=None, epochs=50, time_flag=False, metric=0,
              threads=0, device=None):
    """Host pre-processing (k-means / classic Vecchia, random block order,
    firstbatch, kNN conditioning sets) in libbv.  device=None: all on the
    host (bv_partition); device=d: the k-means assignment steps on GPU d
    (bv_partition_gpu), with the same result.  Returns Blocks."""
    lib = load_library()
    x, y, obs, z = _f64(x), _f64(y), _f64(obs), _f64(z)
    res = BvBlocks()
    if device is None:
        rc = lib.bv_partition(len(x), _d(x), _d(y), _d(z), _d(obs), bc, cs, seed, epochs,
                              1 if time_flag else 0, metric, threads, ctypes.byref(res))
    else:
        rc = lib.bv_partition_gpu(int(device), len(x), _d(x), _d(y), _d(z), _d(obs), bc, cs,
                                  seed, epochs, 1 if time_flag else 0, metric, threads,
                                  ctypes.byref(res))
    if rc != BV_OK:
        raise RuntimeError("bv_partition failed: %d" % rc)
    try:
        n, B = res.n, res.batch_count

        def arr(ptr, count, dt):
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dt, copy=True)
        blocks = Blocks(
            arr(res.x, n, np.float64), arr(res.y, n, np.float64), arr(res.obs, n, np.float64),
            arr(res.batch_num, B, np.int32),
            arr(res.nbr, B * cs, np.int32) if cs > 0 else np.zeros(0, np.int32), cs,
            z=arr(res.z, n, np.float64) if time_flag else None,
            perm=arr(res.perm, n, np.int64),
            cx=arr(res.cx, res.nclusters, np.float64), cy=arr(res.cy, res.nclusters, np.float64),
            first_cluster_count=res.first_cluster_count)
    finally:
        lib.bv_blocks_free(ctypes.byref(res))
    return blocks


class Engine:
    """One dataset resident in HBM; llh(theta) evaluates the Block Vecchia
    log-likelihood on the GPU."""

    def __init__(self, blocks, kernel=1, metric=0, time_flag=False, dist_scale=1.0, device=0,
                 verbose=False, use_copies=False):
        lib = load_library()
        self._lib = lib
        self.blocks = blocks
        self.kernel = int(kernel)
        self.nparams = NPARAMS[self.kernel]
        b = blocks
        p = BvProblem()
        p.n = b.n
        p.batch_count = b.batch_count
        p.cs = b.cs
        p.batch_num = b.batch_num.ctypes.data_as(_ip)
        p.x, p.y, p.z, p.obs = _d(b.x), _d(b.y), _d(b.z), _d(b.obs)
        self._keep = []
        if b.cs > 0 and use_copies:
            idx = b.nbr.astype(np.int64)
            cx, cy, co = b.x[idx].copy(), b.y[idx].copy(), b.obs[idx].copy()
            cz = b.z[idx].copy() if b.z is not None else None
            self._keep += [cx, cy, co, cz]
            p.con_x, p.con_y, p.con_obs, p.con_z = _d(cx), _d(cy), _d(co), _d(cz)
            p.nbr = None
        else:
            p.nbr = b.nbr.ctypes.data_as(_ip) if b.cs > 0 else None
        p.kernel = self.kernel
        p.distance_metric = metric
        p.time_flag = 1 if (time_flag or self.kernel == 4) else 0
        p.dist_scale = dist_scale
        p.device = device
        p.verbose = 1 if verbose else 0
        ctx = ctypes.c_void_p()
        rc = lib.bv_create(ctypes.byref(p), ctypes.byref(ctx))
        if rc != BV_OK:
            raise RuntimeError("bv_create failed: %d" % rc)
        self.ctx = ctx

    def llh(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        return self._lib.bv_llh(len(th), _d(th), None, self.ctx)

    def llh_many(self, thetas):
        th = np.ascontiguousarray(thetas, dtype=np.float64)
        k, npar = th.shape
        out = np.empty(k, np.float64)
        st = (ctypes.c_int * k)()
        rc = self._lib.bv_llh_many(self.ctx, k, npar, _d(th), _d(out), st)
        if rc != BV_OK:
            raise RuntimeError("bv_llh_many failed: %d" % rc)
        return out, np.array(st[:], dtype=np.int32)

    def block_terms(self):
        B = self.blocks.batch_count
        ld, dt, info = np.empty(B), np.empty(B), np.empty(B, np.int32)
        rc = self._lib.bv_block_terms(self.ctx, _d(ld), _d(dt), info.ctypes.data_as(_ip))
        if rc != BV_OK:
            raise RuntimeError("bv_block_terms failed: %d" % rc)
        return ld, dt, info

    def status(self):
        bad = ctypes.c_int(-1)
        st = self._lib.bv_last_status(self.ctx, ctypes.byref(bad))
        return st, bad.value

    def last_kernel_ms(self):
        return self._lib.bv_last_kernel_ms(self.ctx)

    def iterations(self):
        return self._lib.bv_iterations(self.ctx)

    def factor_config(self):
        """(variant, stepped, chunks, streams, graph) of the factorization in
        use, or None before the first evaluation has chosen it."""
        cfg = np.zeros(5, np.int32)
        rc = self._lib.bv_factor_config(self.ctx, cfg.ctypes.data_as(_ip))
        return tuple(int(v) for v in cfg) if rc == BV_OK else None

    def simulate(self, theta, z, nugget=0.0):
        """A draw from the Block Vecchia model of this dataset's partition at
        theta (bv_simulate): z standard normals in block order.  Returns
        (y in block order, status, number of dependency levels)."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        z = np.ascontiguousarray(z, dtype=np.float64)
        if len(z) != self.blocks.n:
            raise ValueError("z must have n entries")
        y = np.empty(self.blocks.n, np.float64)
        lv = ctypes.c_int(0)
        rc = self._lib.bv_simulate(self.ctx, len(th), _d(th), _d(z), float(nugget), _d(y),
                                   ctypes.byref(lv))
        if rc not in (BV_OK, BV_ERR_NOT_PD):
            raise RuntimeError("bv_simulate failed: %d" % rc)
        return y, rc, lv.value

    def set_graph(self, enable):
        """Replay single evaluations from a captured HIP graph (default) or
        launch their kernels one by one."""
        rc = self._lib.bv_set_graph(self.ctx, 1 if enable else 0)
        if rc != BV_OK:
            raise RuntimeError("bv_set_graph failed: %d" % rc)

    def close(self):
        if getattr(self, "ctx", None):
            self._lib.bv_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request server over several resident datasets (Engine objects):
    submit(dataset_id, theta, tag) queues an evaluation; poll() starts queued
    work on idle lanes and returns finished (tag, llh, status) triples."""

    def __init__(self, lanes=2, ws_bytes_per_lane=2 << 30, max_evals_per_batch=64, device=0):
        self._lib = load_library()
        srv = ctypes.c_void_p()
        rc = self._lib.bv_server_create(device, lanes, float(ws_bytes_per_lane),
                                        max_evals_per_batch, ctypes.byref(srv))
        if rc != BV_OK:
            raise RuntimeError("bv_server_create failed: %d" % rc)
        self.srv = srv
        self.engines = []
        self._cap = 4096
        self._tags = (ctypes.c_int64 * self._cap)()
        self._llh = np.empty(self._cap, np.float64)
        self._st = (ctypes.c_int * self._cap)()

    def attach(self, engine):
        did = ctypes.c_int(-1)
        rc = self._lib.bv_server_attach(self.srv, engine.ctx, ctypes.byref(did))
        if rc != BV_OK:
            raise RuntimeError("bv_server_attach failed: %d" % rc)
        self.engines.append(engine)
        return did.value

    def submit(self, dataset_id, theta, tag):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        rc = self._lib.bv_server_submit(self.srv, dataset_id, len(th), _d(th), int(tag))
        if rc != BV_OK:
            raise RuntimeError("bv_server_submit failed: %d" % rc)

    def poll(self, wait=False):
        n = self._lib.bv_server_poll(self.srv, 1 if wait else 0, self._cap, self._tags,
                                     _d(self._llh), self._st)
        if n < 0:
            raise RuntimeError("bv_server_poll failed: %d" % n)
        return [(int(self._tags[i]), float(self._llh[i]), int(self._st[i])) for i in range(n)]

    def pending(self):
        return self._lib.bv_server_pending(self.srv)

    def close(self):
        if getattr(self, "srv", None):
            self._lib.bv_server_destroy(self.srv)
            self.srv = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def matern(x1, y1, x2, y2, theta, kernel=1, metric=0, dist_scale=1.0, z1=None, z2=None,
           device=0):
    """Dense covariance block A[i, j] = K(l1[i], l2[j]) computed on the GPU."""
    lib = load_library()
    x1, y1, x2, y2, z1, z2 = map(_f64, (x1, y1, x2, y2, z1, z2))
    th = _f64(theta)
    m, n = len(x1), len(x2)
    out = np.empty(m * n, np.float64)
    rc = lib.bv_matern(m, n, _d(x1), _d(y1), _d(z1), _d(x2), _d(y2), _d(z2), _d(th), kernel,
                       metric, dist_scale, _d(out), device)
    if rc != BV_OK:
        raise RuntimeError("bv_matern failed: %d" % rc)
    return out.reshape(n, m).T.copy()


def device_selftest(device=0):
    return load_library().bv_device_selftest(device)


def fp64_pipe_probe(device=0):
    """Median cycles per loop repetition of the device's FP64 pipe probe
    (bv_fp64_pipe_probe in include/bv.h): 2 matrix-core products, 8 FP64
    multiply-adds, both in one wavefront, 2 products + 32 multiply-adds, and
    the products / multiply-adds wavefronts of the split body."""
    c = np.zeros(6)
    rc = load_library().bv_fp64_pipe_probe(device, _d(c))
    if rc != BV_OK:
        raise RuntimeError("bv_fp64_pipe_probe failed: %d" % rc)
    keys = ["mfma2", "fma8", "mfma2_fma8", "mfma2_fma32", "split_mfma2", "split_fma8"]
    return dict(zip(keys, c.tolist()))


def factor_candidates():
    """The autotuner's candidate configurations, as (variant, stepped,
    chunks, streams) tuples (bv_factor_candidates)."""
    lib = load_library()
    n = lib.bv_factor_candidates(None, 0)
    buf = np.zeros(4 * n, np.int32)
    lib.bv_factor_candidates(buf.ctypes.data_as(_ip), n)
    return [tuple(int(v) for v in buf[4 * i:4 * i + 4]) for i in range(n)]
