"""The blocked banded L D L^T of k_geo_factor (band_factor_blocked in
lbmib_kernels.hip), restated on the host with the kernel's index formulas:
blocks of FNB columns, each factored by one wavefront (lane a = row k0 + a,
register c = column k0 + c, values of other lanes by readlane), then the
trailing band updated by pairs (a, c) in two parts (T1, the next panel's
columns, right after the panel; T2, the rest, beside the next panel), rows
entering the LDS
window (slot r mod 64) one group of FNB rows per block after the panel,
two blocks after their loads.  Every entry must receive its updates in increasing k with the
same operations as the right-looking algorithm the oracle uses
(oracle/lbmib_oracle.c `ldlt`), so the factors agree bit for bit; the window
slots must never be overwritten while a row in them is still read, which
the restatement checks as it goes.
"""
import numpy as np
import pytest

FNB = 8
WIN = 64
T = 256


def right_looking(A, bw):
    """a(i,j) -= l_i * t_j in increasing k, l_i = t_i * (1/d_k) (dense, lower)."""
    a = A.copy()
    n = a.shape[0]
    dinv = np.zeros(n)
    for k in range(n):
        d = a[k, k]
        r = 1.0 / d
        dinv[k] = r
        iend = min(k + bw, n - 1)
        t = [a[i, k] for i in range(k + 1, iend + 1)]
        for i in range(k + 1, iend + 1):
            a[i, k] = t[i - k - 1] * r
        for j in range(k + 1, iend + 1):
            tj = t[j - k - 1]
            for i in range(j, iend + 1):
                a[i, j] = a[i, j] - a[i, k] * tj
    return a, dinv


def blocked(A, bw):
    """The kernel's schedule over a 64-slot window of band rows."""
    n = A.shape[0]
    W = bw + 1
    assert bw + 2 * FNB <= WIN
    win = np.full((WIN, WIN), np.nan)
    owner = [-1] * WIN  # the row each slot holds
    retired = set()     # rows no longer read

    def ga(r, d):
        return A[r, r - d] if r - d >= 0 else 0.0

    def put(r, d, v):
        s = r % WIN
        if owner[s] != r:
            assert owner[s] < 0 or owner[s] in retired, "slot of row %d still in use" % owner[s]
            owner[s] = r
        win[s, d] = v

    def get(r, d):
        assert owner[r % WIN] == r, "row %d not in the window" % r
        return win[r % WIN, d]

    for r in range(min(n, FNB + bw)):
        for d in range(W):
            put(r, d, ga(r, d))
    GQ = (FNB * (WIN - 2 * FNB + 1) + T - 1) // T

    def load_group(g):
        out = []
        for q in range(GQ):
            for tid in range(T):
                e = tid + q * T
                r, d = g * FNB + bw + e // W, e % W
                if e < FNB * W and r < n:
                    out.append((r, d, ga(r, d)))
        return out

    dinv = np.zeros(n)
    L = np.zeros((n, n))
    pend = {1: load_group(1), 2: load_group(2)}
    lt_buf = [None, None]

    def trailing(b, nb, buf, part):
        """pairs (a, c), c <= a < bw, of the selected columns, each with the block's updates in increasing k"""
        k0 = b * FNB
        lcol, tcol = buf
        for c in range(bw):
            if not part(c):
                continue
            for a in range(c, bw):
                i, j = k0 + FNB + a, k0 + FNB + c
                if i >= n:
                    continue
                v = get(i, i - j)
                for q in range(FNB):
                    if q < nb and i - k0 - q <= bw:
                        v = v - lcol[i - k0, q] * tcol[j - k0, q]
                put(i, i - j, v)

    def finish(b, nb):
        """the block's columns are final: read L as the coefficients do, then retire its rows"""
        k0 = b * FNB
        for c in range(nb):
            k = k0 + c
            m = min(k + bw, n - 1) - k
            for e in range(1, m + 1):
                L[k + e, k] = get(k + e, e)
            for e in range(1, min(bw, k) + 1):
                assert get(k, e) == L[k, k - e]
        for c in range(nb):
            retired.add(k0 + c)

    nblk = (n + FNB - 1) // FNB
    for b in range(nblk):
        k0 = b * FNB
        nb = min(FNB, n - k0)
        # panel: lane a = row k0 + a
        P = np.zeros((64, FNB))
        Tt = np.zeros((64, FNB))
        live = [a < FNB + bw and k0 + a < n for a in range(64)]
        for a in range(64):
            for c in range(FNB):
                dd = a - c
                P[a, c] = get(k0 + a, dd) if (live[a] and 0 <= dd <= bw) else 0.0
        for c in range(FNB):
            if c >= nb:
                break
            d = P[c, c]
            r = 1.0 / d
            dinv[k0 + c] = r
            t = P[:, c].copy()
            l = t * r
            below = [live[a] and 1 <= a - c <= bw for a in range(64)]
            for a in range(64):
                if below[a]:
                    P[a, c] = l[a]
                    Tt[a, c] = t[a]
            for c2 in range(c + 1, FNB):
                tj = t[c2]
                for a in range(64):
                    if below[a] and a >= c2 and c2 - c <= bw:
                        P[a, c2] = P[a, c2] - l[a] * tj
        for a in range(64):
            for c in range(FNB):
                dd = a - c
                if live[a] and 0 <= dd <= bw:
                    put(k0 + a, dd, P[a, c])
        lt_buf[b % 2] = (P.copy(), Tt.copy())
        # beside the panel: T2 of the block before (columns after its next
        # panel's) and its coefficients, then its rows retire
        if b > 0:
            trailing(b - 1, FNB, lt_buf[(b - 1) % 2], lambda c: c >= FNB)
            finish(b - 1, FNB)
        # after the panel: T1 (the next panel's columns), rows of group b+1
        # into the window, loads of group b+3
        trailing(b, nb, lt_buf[b % 2], lambda c: c < FNB)
        for (r, d, v) in pend.pop(b + 1, []):
            put(r, d, v)
        pend[b + 3] = load_group(b + 3)
    if nblk:
        finish(nblk - 1, n - (nblk - 1) * FNB)
    return L, dinv


def spd_band(n, bw, seed):
    rng = np.random.default_rng(seed)
    B = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - bw), i + 1):
            B[i, j] = rng.uniform(-1, 1)
    A = B @ B.T  # bandwidth up to 2*bw; keep the lower bw band only, then make it diagonally dominant
    for i in range(n):
        for j in range(n):
            if abs(i - j) > bw:
                A[i, j] = 0.0
        A[i, i] = abs(A[i]).sum() + 1.0
    return A


@pytest.mark.parametrize("n,bw", [(1, 0), (5, 3), (8, 1), (37, 5), (64, 28), (131, 28), (100, 48), (17, 40)])
def test_blocked_factor_equals_right_looking(n, bw):
    A = spd_band(n, bw, 7 + n + bw)
    a, dinv = right_looking(A, bw)
    L, dinv_b = blocked(A, bw)
    assert np.array_equal(dinv, dinv_b)
    for k in range(n):
        for i in range(k + 1, min(k + bw, n - 1) + 1):
            assert a[i, k] == L[i, k], (i, k)

