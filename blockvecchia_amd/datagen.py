"""Synthetic datasets for the benchmark and the tests (the reference reads
external CSV files, src/test_Xvecchia_batch.cpp:182-237; these are the
build's own generator and are written in the same CSV formats).

  uniform_locations(n, seed)       iid U[0,1]^2 (numpy PCG64, seed)
  clustered_locations(...)         uniform / Gaussian-cluster mixture
  matern_cov(...)                  dense Matern covariance (closed forms for
                                   nu = 1/2, 3/2, 5/2; scipy K_nu otherwise)
  sample_dense(...)                y = L z, L = chol(Sigma) in FP64
  sample_vecchia(...)              Vecchia sequential simulation
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


def sample_vecchia(x, y, sigma2, beta, nu, cs=400, seed=0):
    """Sequential simulation y_i | y_{N(i)} with N(i) the cs nearest earlier
    points (in index order); exact in distribution for the Vecchia
    approximation of the Matern field."""
    from scipy.spatial import cKDTree
    n = len(x)
    rng = np.random.default_rng(seed + 1)
    z = rng.standard_normal(n)
    out = np.empty(n)
    pts = np.column_stack([x, y])
    for i in range(n):
        if i == 0:
            out[0] = np.sqrt(sigma2) * z[0]
            continue
        k = min(cs, i)
        if i <= cs:
            nb = np.arange(i)
        else:
            tree_pts = pts[:i]
            d = np.sum((tree_pts - pts[i]) ** 2, axis=1)
            nb = np.argpartition(d, k - 1)[:k]
        S22 = matern_cov(x[nb], y[nb], x[nb], y[nb], sigma2, beta, nu)
        s21 = matern_cov(x[nb], y[nb], x[i:i + 1], y[i:i + 1], sigma2, beta, nu)[:, 0]
        L = np.linalg.cholesky(S22)
        a = np.linalg.solve(L, s21)
        w = np.linalg.solve(L, out[nb])
        mu = a @ w
        var = sigma2 - a @ a
        out[i] = mu + np.sqrt(max(var, 0.0)) * z[i]
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
