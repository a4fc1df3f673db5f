"""Synthetic datasets for the benchmark and the tests (the reference reads
external CSV files, src/test_Xvecchia_batch.cpp:182-237; these are the
build's own generator and are written in the same CSV formats).

  uniform_locations(n, seed)       iid U[0,1]^2 (numpy PCG64, seed)
  clustered_locations(...)         uniform / Gaussian-cluster mixture
  matern_cov(...)                  dense Matern covariance (closed forms for
                                   nu = 1/2, 3/2, 5/2; scipy K_nu otherwise)
  sample_dense(...)                y = L z, L = chol(Sigma) in FP64
  sample_vecchia(...)              Vecchia sequential simulation on the GPU
                                   (bv_simulate, dependency levels)
  sample_partition(...)            the same over a given partition (GPU)
  sample_block_vecchia(...)        the same model drawn on the CPU over a
                                   given partition (scipy, block by block)
"""
import numpy as np


def uniform_locations(n, seed=0):
    rng = np.random.default_rng(seed)
    return rng.random(n), rng.random(n)


def clustered_locations(n, seed=0, frac_uniform=0.7, nclusters=20, sd=0.02):
    rng = np.random.default_rng(seed)
    nu = int(round(n * frac_uniform))
    x = np.empty(n)
    y = np.empty(n)
    x[:nu] = rng.random(nu)
    y[:nu] = rng.random(nu)
    cen = rng.random((nclusters, 2))
    lab = rng.integers(0, nclusters, n - nu)
    x[nu:] = np.clip(cen[lab, 0] + sd * rng.standard_normal(n - nu), 0.0, 1.0)
    y[nu:] = np.clip(cen[lab, 1] + sd * rng.standard_normal(n - nu), 0.0, 1.0)
    return x, y


def _matern_of_scaled(e, sigma2, nu):
    out = np.empty_like(e)
    zero = e == 0
    ez = np.where(zero, 1.0, e)
    if nu == 0.5:
        v = sigma2 * np.exp(-ez)
    elif nu == 1.5:
        v = sigma2 * (1.0 + ez) * np.exp(-ez)
    elif nu == 2.5:
        v = sigma2 * (1.0 + ez + ez * ez / 3.0) * np.exp(-ez)
    else:
        from scipy.special import gamma, kv
        con = sigma2 / (2.0 ** (nu - 1.0) * gamma(nu))
        v = con * ez ** nu * kv(nu, ez)
    out[:] = np.where(zero, sigma2, v)
    return out


def matern_cov(x1, y1, x2, y2, sigma2, beta, nu, dist_scale=1.0):
    d = np.sqrt((x2[None, :] - x1[:, None]) ** 2 + (y2[None, :] - y1[:, None]) ** 2)
    return _matern_of_scaled(d / beta / dist_scale, sigma2, nu)


def sample_dense(x, y, sigma2, beta, nu, seed=0, chunk=4096, device=None):
    """y = L z with L the FP64 Cholesky of the full covariance.  device='cuda'
    uses torch on the GPU for the factorization (data preparation only)."""
    n = len(x)
    rng = np.random.default_rng(seed + 1)
    z = rng.standard_normal(n)
    if device is not None:
        import torch
        xt = torch.as_tensor(x, dtype=torch.float64, device=device)
        yt = torch.as_tensor(y, dtype=torch.float64, device=device)
        S = torch.empty((n, n), dtype=torch.float64, device=device)
        for i0 in range(0, n, chunk):
            i1 = min(n, i0 + chunk)
            d = torch.sqrt((xt[None, :] - xt[i0:i1, None]) ** 2 + (yt[None, :] - yt[i0:i1, None]) ** 2)
            e = d / beta
            S[i0:i1] = torch.as_tensor(_matern_of_scaled(e.cpu().numpy(), sigma2, nu), device=device)
        L = torch.linalg.cholesky(S)
        del S
        out = (L @ torch.as_tensor(z, dtype=torch.float64, device=device)).cpu().numpy()
        del L
        return out
    S = np.empty((n, n))
    for i0 in range(0, n, chunk):
        i1 = min(n, i0 + chunk)
        S[i0:i1] = matern_cov(x[i0:i1], y[i0:i1], x, y, sigma2, beta, nu)
    L = np.linalg.cholesky(S)
    del S
    return L @ z


def sample_vecchia(x, y, sigma2, beta, nu, cs=400, bc=None, seed=0, device=0, z=None):
    """Vecchia sequential simulation on the GPU (bv_simulate): the points are
    partitioned as the likelihood partitions them (k-means into bc blocks,
    random block order, the cs nearest earlier points of each block's
    centroid as its conditioning set; bc >= n/2 gives one point per block,
    the classic point-wise Vecchia order), then each block is drawn from its
    Gaussian conditional given the values already drawn, block by block in
    dependency levels.  Exact in distribution for that Vecchia model.  A
    covariance that is not numerically positive definite in FP64 (smooth
    kernels on tightly clustered points) is drawn again with the smallest
    diagonal jitter 1e-12 sigma2 times a power of 10 that succeeds, on every
    block.  Returns the draw in the points' original order; `nugget` and
    `levels` on the result say which jitter and how many levels were used.
    z: n standard normals in the partition's block order (default: seeded)."""
    import blockvecchia_amd as bv
    n = len(x)
    if bc is None:
        bc = max(1, n // 20)
    blocks = bv.partition(x, y, np.zeros(n), bc, cs, seed=0, device=device)
    yb = sample_partition(blocks, sigma2, beta, nu, seed=seed, device=device, z=z)
    out = np.empty(n)
    out[np.asarray(blocks.perm)] = yb
    out = out.view(_Sample)
    out.nugget = yb.nugget
    out.levels = yb.levels
    out.blocks = blocks
    return out


def sample_partition(blocks, sigma2, beta, nu, seed=0, device=0, z=None):
    """A draw from the Block Vecchia model of a given partition on the GPU
    (bv_simulate), in block order, with the jitter rule of sample_vecchia.
    z: standard normals in block order (default: seeded as sample_vecchia)."""
    import blockvecchia_amd as bv
    n = blocks.n
    if z is None:
        z = np.random.default_rng(seed + 1).standard_normal(n)
    eng = bv.Engine(bv.Blocks(blocks.x, blocks.y, np.zeros(n), blocks.batch_num, blocks.nbr,
                              blocks.cs), device=device)
    try:
        for k in [None] + list(range(-12, 0)):
            nug = 0.0 if k is None else sigma2 * 10.0 ** k
            yb, rc, levels = eng.simulate([sigma2, beta, nu], z, nugget=nug)
            if rc == bv.BV_OK:
                break
        else:
            raise RuntimeError("covariance not positive definite even with 0.1 sigma2 jitter")
    finally:
        eng.close()
    out = np.ascontiguousarray(yb).view(_Sample)
    out.nugget = nug
    out.levels = levels
    return out


def write_xy_csv(path, x, y, z=None):
    with open(path, "w") as f:
        for i in range(len(x)):
            if z is None:
                f.write("%.17g,%.17g\n" % (x[i], y[i]))
            else:
                f.write("%.17g,%.17g,%.17g\n" % (x[i], y[i], z[i]))


def write_obs_csv(path, obs):
    with open(path, "w") as f:
        for v in obs:
            f.write("%.17g\n" % v)


def sample_block_vecchia(blocks, sigma2, beta, nu, seed=0, threads=None, chunk=256, z=None):
    """A draw from the Block Vecchia model of a partition (bv.partition's
    Blocks, points in block order): block 0 from its own covariance, block
    i >= 1 from y_B | y_N(i) with N(i) its conditioning set.  Returns the
    observations in block order.  Exact in distribution for that model.  The
    per-block factorizations do not depend on y and run in a thread pool; only
    the conditional means run in block order.  A covariance that is not
    numerically positive definite in FP64 (smooth kernels on tightly clustered
    points) is factored with the smallest diagonal jitter, 1e-12 sigma2 times a
    power of 10, that succeeds; `blocks_jittered` on the result counts them.
    z: standard normals in block order (default: seeded).  The CPU
    restatement of bv_simulate (the same draw for the same z when no block
    needs jitter)."""
    from concurrent.futures import ThreadPoolExecutor

    from scipy.linalg import LinAlgError, solve_triangular
    from scipy.linalg import cholesky as _chol

    jittered = []

    def cholesky(a, lower=True):
        try:
            return _chol(a, lower=lower)
        except LinAlgError:
            pass
        for k in range(-12, 0):
            try:
                r = _chol(a + (sigma2 * 10.0 ** k) * np.eye(len(a)), lower=lower)
                jittered.append(k)
                return r
            except LinAlgError:
                continue
        raise LinAlgError("covariance not positive definite even with 0.1 sigma2 jitter")

    x, y = blocks.x, blocks.y
    bn = np.asarray(blocks.batch_num, dtype=np.int64)
    cs = int(blocks.cs)
    acc = np.concatenate([[0], np.cumsum(bn)])
    nbr = np.asarray(blocks.nbr).reshape(-1, cs) if cs > 0 else None
    B = len(bn)
    eps = (np.random.default_rng(seed + 1).standard_normal(len(x)) if z is None
           else np.asarray(z, dtype=np.float64))
    out = np.empty(len(x))

    def prep(i):
        p = slice(acc[i], acc[i + 1])
        sbb = matern_cov(x[p], y[p], x[p], y[p], sigma2, beta, nu)
        if i == 0 or cs == 0:
            return None, None, cholesky(sbb, lower=True)
        q = nbr[i]
        ln = cholesky(matern_cov(x[q], y[q], x[q], y[q], sigma2, beta, nu), lower=True)
        a = solve_triangular(ln, matern_cov(x[q], y[q], x[p], y[p], sigma2, beta, nu), lower=True)
        return ln, a, cholesky(sbb - a.T @ a, lower=True)

    with ThreadPoolExecutor(threads) as ex:
        for c0 in range(0, B, chunk):
            parts = list(ex.map(prep, range(c0, min(B, c0 + chunk))))
            for i, (ln, a, lc) in zip(range(c0, c0 + len(parts)), parts):
                p = slice(acc[i], acc[i + 1])
                mean = 0.0 if ln is None else a.T @ solve_triangular(ln, out[nbr[i]], lower=True)
                out[p] = mean + lc @ eps[p]
    out = out.view(_Sample)
    out.blocks_jittered = len(jittered)
    return out


class _Sample(np.ndarray):
    """float64 observations with the count of jittered factorizations (CPU
    sampler) or the jitter and dependency levels used (GPU sampler)."""
    blocks_jittered = 0
    nugget = 0.0
    levels = 0
    blocks = None
