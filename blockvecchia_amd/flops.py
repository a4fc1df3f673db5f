"""Algorithmic FP64 operation count of one Block Vecchia evaluation, from the
reference's FLOP macros (include/flops.h, taken from MAGMA's testing code):
  FLOPS_DPOTRF(n)      = n^3/3 + n^2/2 + n/6               (flops.h:93-94, 289)
  FLOPS_DTRSM(L, m, n) = n m^2                             (flops.h:66-75, 266)
  FLOPS_DSYRK(k, n)    = k n (n + 1)                       (flops.h:57-58, 248)
  FLOPS_DGEMM(m, n, k) = 2 m n k                           (flops.h:49-50, 232)
Per evaluation (blocks i >= 1 condition on cs points; block 0 is a dense
factorization; the reference's full A^T A GEMM is counted as the SYRK it
mathematically is):
  F = sum_{i>=1} [POTRF(cs) + TRSM(cs, bs_i) + TRSM(cs, 1) + SYRK(cs, bs_i)
                  + GEMM(bs_i, 1, cs) + bs_i^2 + bs_i + POTRF(bs_i)
                  + TRSM(bs_i, 1) + 2 bs_i]
      + POTRF(bs_0) + TRSM(bs_0, 1) + 2 bs_0
"""
import numpy as np


def potrf(n):
    n = np.asarray(n, dtype=np.float64)
    return n * (((1.0 / 6.0) * n + 0.5) * n + 1.0 / 3.0) + n * (((1.0 / 6.0) * n) * n - 1.0 / 6.0)


def trsm_left(m, n):
    m = np.asarray(m, dtype=np.float64)
    n = np.asarray(n, dtype=np.float64)
    return 0.5 * n * m * (m + 1) + 0.5 * n * m * (m - 1)


def syrk(k, n):
    k = np.asarray(k, dtype=np.float64)
    n = np.asarray(n, dtype=np.float64)
    return 2.0 * 0.5 * k * n * (n + 1)


def gemm(m, n, k):
    return 2.0 * np.float64(m) * n * k


def eval_flops(batch_num, cs):
    bs = np.asarray(batch_num, dtype=np.float64)
    b0, bi = bs[0], bs[1:]
    f = potrf(b0) + trsm_left(b0, 1) + 2 * b0
    if cs > 0:
        f += np.sum(potrf(cs) + trsm_left(cs, bi) + trsm_left(cs, 1) + syrk(cs, bi)
                    + gemm(bi, 1, cs) + bi * bi + bi)
    f += np.sum(potrf(bi) + trsm_left(bi, 1) + 2 * bi)
    return float(f)


def compulsory_bytes(n, batch_count, cs, dims=2):
    """Bytes an evaluation must read once: coordinates, observations, block
    sizes and neighbour indices (plus the per-block results written)."""
    return 8.0 * n * (dims + 1) + 4.0 * batch_count * (cs + 1) + 8.0 * 2 * batch_count
