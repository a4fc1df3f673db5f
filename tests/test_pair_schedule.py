"""The launch schedule of the paired-panel factorization (bv_pair_kernel),
emulated on the CPU with 16x16 numpy tiles.

Launch s = -1, 0, 1, ... processes the panel pair (2s, 2s+1) (tile columns
4s..4s+3) for the rows 4s+4.. in workgroups of four row tiles (two waves of
two), and the workgroup holding rows 4s+4..4s+7 also factors the next pair's
64x64 diagonal block: panel 2s+2's 32x32 block by wave A, then wave B's rows
of panel 2s+2 (L_BA = D_BA W_A^T) and panel 2s+3's block.  Launch -1 is that
factorization alone for pair 0.  Workgroups run in a shuffled order with every
write deferred to the end of the launch, and each launch's read and write
sets are checked disjoint between workgroups (the kernel's workgroups never
wait for each other).  The tiles left behind (L below the diagonal blocks,
W = L_pp^-1 on them) and the observation row's per-panel partial sums are
checked against numpy's Cholesky of the same augmented matrix."""
import numpy as np
import pytest


def _diag_factor(C32, J0, m, cs0):
    """32x32 lower Cholesky with the engine's pivot 1 for rows >= m, its
    inverse, and row m's squares inside the block (columns >= cs0)."""
    Lb = np.zeros((32, 32))
    Cw = C32.copy()
    for k in range(32):
        piv = np.sqrt(Cw[k, k]) if J0 + k < m else 1.0
        Lb[k, k] = piv
        for i in range(k + 1, 32):
            Lb[i, k] = Cw[i, k] / piv
        for i in range(k + 1, 32):
            for jj in range(k + 1, i + 1):
                Cw[i, jj] -= Lb[i, k] * Lb[jj, k]
    d = 0.0
    if J0 <= m < J0 + 32:
        q = np.arange(J0, J0 + 32)
        msk = (q >= cs0) & (q < m)
        d = np.sum(Lb[m - J0, msk] ** 2)
    return Lb, np.linalg.inv(Lb), d


def _row_m_dot(X, rt, J0, m, cs0):
    for r in range(16):
        if rt * 16 + r == m:
            q = np.arange(J0, J0 + X.shape[1])
            msk = (q >= cs0) & (q < m)
            return np.sum(X[r, msk] ** 2)
    return None


def _run(Mp, m, cs0, seed):
    T = Mp // 16
    P = Mp // 32
    rng = np.random.default_rng(seed)
    n = m
    B = rng.standard_normal((n, n))
    S = B @ B.T / n + np.eye(n)
    y = rng.standard_normal(n)
    A = np.eye(Mp)
    A[:n, :n] = S
    A[n, :n] = y
    ws = {(rt, ct): A[rt * 16:(rt + 1) * 16, ct * 16:(ct + 1) * 16].copy()
          for rt in range(T) for ct in range(rt + 1)}
    Lref = np.eye(Mp)
    L11 = np.linalg.cholesky(S)
    Lref[:n, :n] = L11
    Lref[n, :n] = np.linalg.solve(L11, y)
    part_dot = {}
    s = -1
    nlaunch = 0
    while True:
        if s == -1:
            groups = [(0, min(4, T))]
        else:
            groups = [(r, min(4, T - r)) for r in range(4 * s + 4, T, 4)]
        if not groups:
            break
        nlaunch += 1
        rng.shuffle(groups)
        log, writes = [], []
        for row0, nr in groups:
            acc = []

            def rd(rt, ct):
                acc.append(('r', (rt, ct)))
                return ws[(rt, ct)]

            def wr(rt, ct, v):
                acc.append(('w', (rt, ct)))
                writes.append(((rt, ct), v.copy()))

            c0 = 4 * s
            X = {}
            if s >= 0:
                for w in range(2):
                    rows = [row0 + 2 * w + i for i in range(2) if 2 * w + i < nr]
                    if not rows:
                        continue
                    W0 = np.zeros((32, 32))
                    W0[:16, :16] = rd(c0, c0); W0[16:, :16] = rd(c0 + 1, c0)
                    W0[16:, 16:] = rd(c0 + 1, c0 + 1)
                    W1 = np.zeros((32, 32))
                    W1[:16, :16] = rd(c0 + 2, c0 + 2); W1[16:, :16] = rd(c0 + 3, c0 + 2)
                    W1[16:, 16:] = rd(c0 + 3, c0 + 3)
                    Lp = np.block([[rd(c0 + 2, c0), rd(c0 + 2, c0 + 1)],
                                   [rd(c0 + 3, c0), rd(c0 + 3, c0 + 1)]])
                    for rt in rows:
                        C = A[rt * 16:(rt + 1) * 16, c0 * 16:(c0 + 4) * 16].copy()
                        for k in range(c0):
                            C -= rd(rt, k) @ np.vstack([rd(c0 + c, k) for c in range(4)]).T
                        X0 = C[:, :32] @ W0.T
                        X1 = (C[:, 32:] - X0 @ Lp.T) @ W1.T
                        for p, Xp in ((2 * s, X0), (2 * s + 1, X1)):
                            dv = _row_m_dot(Xp, rt, p * 32, m, cs0)
                            if dv is not None:
                                part_dot[p] = dv
                        Xr = np.hstack([X0, X1])
                        X[rt] = Xr
                        for c in range(4):
                            wr(rt, c0 + c, Xr[:, c * 16:(c + 1) * 16])
            look = row0 == 4 * s + 4
            if look:
                qn = row0
                nq = min(4, T - qn)
                Q = range(qn, qn + nq)
                D = A[qn * 16:(qn + nq) * 16, qn * 16:(qn + nq) * 16].copy()
                for k in range(max(c0, 0)):
                    Lk = np.vstack([rd(q, k) for q in Q])
                    D -= Lk @ Lk.T
                if s >= 0:
                    Lk = np.vstack([X[q] for q in Q])
                    D -= Lk @ Lk.T
                pa = qn // 2
                LA, WA, dA = _diag_factor(np.tril(D[:32, :32]), qn * 16, m, cs0)
                for a, b in [(0, 0), (1, 0), (1, 1)]:
                    wr(qn + a, qn + b, WA[a * 16:(a + 1) * 16, b * 16:(b + 1) * 16])
                part_dot[pa] = dA
                if nq == 4:
                    LBA = D[32:, :32] @ WA.T
                    for r in range(2):
                        dv = _row_m_dot(LBA[r * 16:(r + 1) * 16], qn + 2 + r, qn * 16, m, cs0)
                        if dv is not None:
                            part_dot[pa] += dv
                        for c in range(2):
                            wr(qn + 2 + r, qn + c, LBA[r * 16:(r + 1) * 16, c * 16:(c + 1) * 16])
                    D11 = D[32:, 32:] - LBA @ LBA.T
                    LB, WB, dB = _diag_factor(np.tril(D11), (qn + 2) * 16, m, cs0)
                    for a, b in [(0, 0), (1, 0), (1, 1)]:
                        wr(qn + 2 + a, qn + 2 + b, WB[a * 16:(a + 1) * 16, b * 16:(b + 1) * 16])
                    part_dot[pa + 1] = dB
            log.append(acc)
        for i, ai in enumerate(log):
            wi = {t for mo, t in ai if mo == 'w'}
            for k, ak in enumerate(log):
                if k == i:
                    continue
                tk = {t for mo, t in ak}
                assert not (wi & tk), ("hazard in launch", s, wi & tk)
        for t, v in writes:
            ws[t] = v
        s += 1
    assert nlaunch == (P + 1) // 2
    assert sorted(part_dot) == list(range(P)), "every panel's partial written"
    err = 0.0
    for rt in range(T):
        for ct in range(rt + 1):
            p = ct // 2
            if rt // 2 == p:
                Wr = np.linalg.inv(Lref[2 * p * 16:(2 * p + 2) * 16, 2 * p * 16:(2 * p + 2) * 16])
                ref = Wr[(rt - 2 * p) * 16:(rt - 2 * p + 1) * 16, (ct - 2 * p) * 16:(ct - 2 * p + 1) * 16]
            else:
                ref = Lref[rt * 16:(rt + 1) * 16, ct * 16:(ct + 1) * 16]
            err = max(err, np.max(np.abs(ws[(rt, ct)] - ref)))
    dot = sum(part_dot[p] for p in range(P))
    dref = np.sum(Lref[m, cs0:m] ** 2)
    return err, abs(dot - dref) / dref


@pytest.mark.parametrize("Mp,m,cs0", [(32, 20, 5), (64, 50, 20), (96, 70, 30), (128, 100, 60),
                                      (160, 140, 100), (192, 150, 100), (320, 313, 300),
                                      (320, 300, 0), (352, 330, 300), (64, 33, 10),
                                      (96, 95, 40), (128, 64, 40)])
def test_pair_schedule(Mp, m, cs0):
    for seed in range(2):
        err, dot_err = _run(Mp, m, cs0, seed)
        assert err < 1e-9 and dot_err < 1e-9, (seed, err, dot_err)
