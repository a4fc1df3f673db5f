"""The launch schedule of the stepped factorization with lookahead
(bv_step2_kernel), emulated on the CPU with 16x16 numpy tiles: launches
j = -1..P-2, launch j's items on rows >= 2j+2 in groups of G (the first also
factors panel j+1's diagonal block), launch -1 the first diagonal block
alone.  Items run in a shuffled order with every write deferred to the end of
the launch, and each launch's read and write sets are checked disjoint
between items (what the kernel relies on: items never wait for each other).
The tiles left behind (L below the diagonal blocks, W = L_pp^-1 on them) and
the observation row's partial sums are checked against numpy's Cholesky of
the same augmented matrix, with the engine's pivot 1 for rows >= m."""
import numpy as np
import pytest


def _run(Mp, m, cs0, G, seed):
    T = Mp // 16; P = Mp // 32
    rng = np.random.default_rng(seed)
    n = m
    B = rng.standard_normal((n, n)); S = B @ B.T / n + np.eye(n)
    y = rng.standard_normal(n)
    A = np.eye(Mp); A[:n, :n] = S; A[n, :n] = y
    # the engine's pivots: rows >= m use pivot 1 (row m: the observation row)
    ws = {(rt, ct): A[rt*16:(rt+1)*16, ct*16:(ct+1)*16].copy() for rt in range(T) for ct in range(rt+1)}
    Lref = np.zeros((Mp, Mp))
    # reference: Cholesky of the leading m x m, row m solved, padding identity
    L11 = np.linalg.cholesky(S)
    Lref[:n, :n] = L11
    Lref[n, :n] = np.linalg.solve(L11, y)
    Lref[n, n] = 1.0  # the engine's pivot for row m
    for r in range(n+1, Mp): Lref[r, r] = 1.0
    part_dot = {}
    def tile(rt, ct, acc, mode):
        acc.append((mode, (rt, ct)))
        return ws[(rt, ct)]
    def diag_factor(C32, J0):
        # C32: 32x32 lower block; rows >= m: pivot 1 (identity rows below m are already identity)
        Lb = np.zeros((32, 32))
        Cw = C32.copy()
        for k in range(32):
            r = J0 + k
            if r < m:
                piv = np.sqrt(Cw[k, k])
            else:
                piv = 1.0
            Lb[k, k] = piv
            for i in range(k+1, 32):
                Lb[i, k] = Cw[i, k] / piv
            for i in range(k+1, 32):
                for jj in range(k+1, i+1):
                    Cw[i, jj] -= Lb[i, k] * Lb[jj, k]
        W = np.linalg.inv(Lb)
        # dot of row m inside the block
        d = 0.0
        if J0 <= m < J0 + 32:
            k = m - J0
            q = np.arange(J0, J0 + 32)
            msk = (q >= cs0) & (q < m)
            d = np.sum(Lb[k, msk] ** 2)
        return Lb, W, d
    for j in range(-1, P - 1):
        if j == -1:
            items = [(0, 2)]
        else:
            items = [(r, min(G, T - r)) for r in range(2*j + 2, T, G)]
        rng.shuffle(items)
        log = []
        writes = []
        for row0, nr in items:
            acc = []
            look = row0 == 2*j + 2
            xl = {}
            if j >= 0:
                W = np.zeros((32, 32))
                W[:16, :16] = tile(2*j, 2*j, acc, 'r'); W[16:, :16] = tile(2*j+1, 2*j, acc, 'r'); W[16:, 16:] = tile(2*j+1, 2*j+1, acc, 'r')
                for s in range(nr):
                    rt = row0 + s
                    C = A[rt*16:(rt+1)*16, 2*j*16:(2*j+2)*16].copy()
                    for k in range(0, 2*j):
                        C -= tile(rt, k, acc, 'r') @ np.vstack([tile(2*j, k, acc, 'r'), tile(2*j+1, k, acc, 'r')]).T
                    X = C @ W.T
                    xl[s] = X
                    for r in range(16):
                        if rt*16 + r == m:
                            q = np.arange(2*j*16, 2*j*16 + 32)
                            msk = (q >= cs0) & (q < m)
                            part_dot[j] = np.sum(X[r, msk] ** 2)
                    acc.append(('w', (rt, 2*j))); acc.append(('w', (rt, 2*j+1)))
                    writes.append(((rt, 2*j), X[:, :16].copy())); writes.append(((rt, 2*j+1), X[:, 16:].copy()))
            if look:
                jp = j + 1
                r0 = 2*jp
                C = A[r0*16:(r0+2)*16, r0*16:(r0+2)*16].copy()
                for k in range(0, 2*j if j >= 0 else 0):
                    Lk = np.vstack([tile(r0, k, acc, 'r'), tile(r0+1, k, acc, 'r')])
                    C -= Lk @ Lk.T
                if j >= 0:
                    Lk = np.vstack([xl[0], xl[1]])
                    C -= Lk @ Lk.T
                Lb, Wn, d = diag_factor(np.tril(C), r0*16)
                part_dot[jp] = d
                for (a, b) in [(r0, r0), (r0+1, r0), (r0+1, r0+1)]:
                    acc.append(('w', (a, b)))
                    writes.append(((a, b), Wn[(a-r0)*16:(a-r0+1)*16, (b-r0)*16:(b-r0+1)*16].copy()))
            log.append(acc)
        # hazards: a tile written by one item and read or written by another
        for i, ai in enumerate(log):
            wi = {t for mo, t in ai if mo == 'w'}
            for k, ak in enumerate(log):
                if k == i: continue
                tk = {t for mo, t in ak}
                assert not (wi & tk), ("hazard in launch", j, wi & tk)
        for t, v in writes:   # all writes land after the launch (items see only earlier launches)
            ws[t] = v
    # check: L below the diagonal blocks and W on them
    err = 0.0
    for rt in range(T):
        for ct in range(rt + 1):
            p = ct // 2
            if rt // 2 == p:   # diagonal block of panel p: W = inv(L_pp)
                Lpp = Lref[2*p*16:(2*p+2)*16, 2*p*16:(2*p+2)*16]
                Wr = np.linalg.inv(Lpp)
                ref = Wr[(rt-2*p)*16:(rt-2*p+1)*16, (ct-2*p)*16:(ct-2*p+1)*16]
            else:
                ref = Lref[rt*16:(rt+1)*16, ct*16:(ct+1)*16]
            err = max(err, np.max(np.abs(ws[(rt, ct)] - ref)))
    dot = sum(part_dot.get(p, 0.0) for p in range(P))
    dref = np.sum(Lref[m, cs0:m] ** 2)
    return err, abs(dot - dref) / dref



@pytest.mark.parametrize("Mp,m,cs0,G", [(64, 50, 20, 2), (128, 100, 60, 2), (320, 313, 300, 2),
                                        (320, 300, 0, 2), (96, 70, 30, 3), (160, 140, 100, 4),
                                        (64, 33, 10, 2), (96, 95, 40, 2)])
def test_lookahead_schedule(Mp, m, cs0, G):
    for seed in range(3):
        err, dot_err = _run(Mp, m, cs0, G, seed)
        assert err < 1e-9 and dot_err < 1e-9, (seed, err, dot_err)
