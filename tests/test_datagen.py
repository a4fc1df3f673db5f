"""The synthetic-data generators (CPU): the Block Vecchia sampler reduces to
the dense Cholesky draw when the partition is a single block, and its
per-block conditional draws have the model's conditional moments."""
import numpy as np

import blockvecchia_amd as bv
from blockvecchia_amd import datagen


def test_single_block_equals_dense_draw():
    n = 300
    x, y = datagen.uniform_locations(n, 2)
    b = bv.partition(x, y, np.zeros(n), 1, 0, seed=0)   # one k-means cluster, points in index order
    assert b.batch_count == 1 and np.array_equal(np.asarray(b.perm), np.arange(n))
    for nu in (0.5, 1.5, 0.8):
        s = np.asarray(datagen.sample_block_vecchia(b, 1.3, 0.2, nu, seed=4))
        d = datagen.sample_dense(x, y, 1.3, 0.2, nu, seed=4)
        np.testing.assert_allclose(s, d, rtol=1e-10, atol=1e-12)


def test_conditional_moments_of_a_block():
    """Block i's draw is mean + chol(cond. cov) eps with the Gaussian
    conditional moments given its conditioning set: recompute them densely
    and check the draw reproduces eps."""
    n = 400
    x, y = datagen.uniform_locations(n, 3)
    b = bv.partition(x, y, np.zeros(n), 25, 20, seed=0)
    s = np.asarray(datagen.sample_block_vecchia(b, 1.0, 0.1, 1.5, seed=1))
    eps = np.random.default_rng(2).standard_normal(n)
    acc = np.concatenate([[0], np.cumsum(b.batch_num)])
    nbr = np.asarray(b.nbr).reshape(-1, 20)
    bx, by = np.asarray(b.x), np.asarray(b.y)
    for i in (1, 7, b.batch_count - 1):
        p = np.arange(acc[i], acc[i + 1])
        q = nbr[i]
        assert q.max() < acc[i]          # conditioning points precede the block
        S = lambda u, v: datagen.matern_cov(bx[u], by[u], bx[v], by[v], 1.0, 0.1, 1.5)
        K = np.linalg.solve(S(q, q), S(q, p))
        mean = K.T @ s[q]
        cov = S(p, p) - S(q, p).T @ K
        L = np.linalg.cholesky(cov)
        np.testing.assert_allclose(np.linalg.solve(L, s[p] - mean), eps[p], rtol=1e-6, atol=1e-6)
    assert datagen.sample_block_vecchia(b, 1.0, 0.1, 1.5, seed=1).blocks_jittered == 0


def _block_vecchia_covariance(b, theta):
    """The joint covariance the Block Vecchia model of partition b implies:
    y_B(i) = K_i y_N(i) + e_i, e_i ~ N(0, S_i), i.e. y = (I - K)^-1 e, built
    with dense solves (np.linalg.solve, no Cholesky)."""
    n = b.n
    acc = np.concatenate([[0], np.cumsum(b.batch_num)])
    nbr = np.asarray(b.nbr).reshape(-1, b.cs) if b.cs > 0 else None
    bx, by = np.asarray(b.x), np.asarray(b.y)
    S = lambda u, v: datagen.matern_cov(bx[u], by[u], bx[v], by[v], *theta)
    K = np.zeros((n, n))
    D = np.zeros((n, n))
    for i in range(b.batch_count):
        p = np.arange(acc[i], acc[i + 1])
        if i == 0 or b.cs == 0:
            D[np.ix_(p, p)] = S(p, p)
            continue
        q = nbr[i]
        Ki = np.linalg.solve(S(q, q), S(q, p)).T
        K[np.ix_(p, q)] = Ki
        D[np.ix_(p, p)] = S(p, p) - Ki @ S(q, p)
    M = np.linalg.inv(np.eye(n) - K)
    return M @ D @ M.T


def test_block_vecchia_draws_have_the_model_covariance():
    """Monte Carlo: 6,000 draws of a 36-point, 6-block model (cs = 8) have
    the empirical covariance of the Block Vecchia model within sampling
    error (standard error of a covariance entry: sqrt((s_ii s_jj + s_ij^2) /
    N)); a sampler that dropped the conditional mean, or used the marginal
    instead of the conditional covariance, is off by many standard errors."""
    n, N = 36, 6000
    x, y = datagen.uniform_locations(n, 8)
    b = bv.partition(x, y, np.zeros(n), 6, 8, seed=0)
    theta = (1.0, 0.25, 1.5)
    C = _block_vecchia_covariance(b, theta)
    rng = np.random.default_rng(9)
    Y = np.empty((N, n))
    for k in range(N):
        Y[k] = datagen.sample_block_vecchia(b, *theta, threads=1, z=rng.standard_normal(n))
    E = Y.T @ Y / N
    se = np.sqrt((np.outer(np.diag(C), np.diag(C)) + C ** 2) / N)
    zsc = np.abs(E - C) / se
    assert zsc.max() < 5.0, zsc.max()
    # the conditional mean matters: the model's cross-block covariances are
    # far from zero, so a draw without it would miss them by many errors
    acc = np.concatenate([[0], np.cumsum(b.batch_num)])
    off = C[acc[1]:, :acc[1]]
    assert np.abs(off).max() / se[acc[1]:, :acc[1]].max() > 20
