"""The separator form of the immersed-boundary solve, restated in numpy with
the kernels' own index formulas (k_geo_factor2, k_geo_schur, k_crit_solve2):
band storage G[p][d] = a(p, p-d); top half rows 0..mT-1, separator rows
mT..mT+b-1, bottom half mT+b..n-1 taken in reverse row order; each half's
L D L^T; Y = L^-1 C from the factor's last b rows; S' = A_S - sum Y^T D^-1 Y;
forward sweeps; x_S = M v with M = [S'^-1, -S'^-1 Y_T^T, -S'^-1 Y_B^T] and
v = [f_S; both halves' scaled forward results over their last b rows]; tail
corrections N x_S with N = [D_T Y_T; D_B Y_B]; backward sweeps.

Checked against a dense solve of random banded SPD systems, and on a
marker-ring system D D^T in zig-zag order (the IB systems' structure) by its
backward error, so that a wrong index in any of those formulas fails here on
the CPU."""
import numpy as np
import pytest


def band_of(A, b):
    n = A.shape[0]
    G = np.zeros((n, b + 1))
    for p in range(n):
        for d in range(min(b, p) + 1):
            G[p, d] = A[p, p - d]
    return G


def ldlt_band(ga, n, b):
    """L D L^T of the band read by ga(r, d) = a(r, r-d): returns L (dense,
    unit lower), dinv, and ltail[i][d] = row n-b+i's (d_k, l(k,k-1), ..)."""
    A = np.zeros((n, n))
    for r in range(n):
        for d in range(min(b, r) + 1):
            A[r, r - d] = A[r - d, r] = ga(r, d)
    L = np.eye(n)
    dv = np.zeros(n)
    W = A.copy()
    for k in range(n):
        dv[k] = W[k, k]
        for i in range(k + 1, min(n, k + b + 1)):
            L[i, k] = W[i, k] / dv[k]
        for i in range(k + 1, min(n, k + b + 1)):
            for j in range(k + 1, i + 1):
                W[i, j] -= L[i, k] * W[j, k]
    ltail = np.zeros((b, b + 1))
    for i in range(b):
        k = n - b + i
        ltail[i, 0] = dv[k]
        for d in range(1, b + 1):
            ltail[i, d] = L[k, k - d] if d <= k else 0.0
    return L, 1.0 / dv, ltail


def separator_solve(G, b, f):
    n = G.shape[0]
    W = b + 1
    mT = (n - b) // 2
    mB = n - b - mT
    gT = lambda r, d: G[r, d]
    gB = lambda r, d: G[n - 1 - r + d, d] if d <= r else 0.0
    LT, dT, ltT = ldlt_band(gT, mT, b)
    LB, dB, ltB = ldlt_band(gB, mB, b)
    Y = np.zeros((2, b, b))
    for h, lt in ((0, ltT), (1, ltB)):
        for s in range(b):
            for i in range(b):
                if h == 0:
                    d = b + s - i
                    cv = G[mT + s, d] if d <= b else 0.0
                else:
                    d = 2 * b - 1 - i - s
                    cv = G[mT + 2 * b - 1 - i, d] if d <= b else 0.0
                for j in range(i):
                    cv -= lt[i, i - j] * Y[h, j, s]
                Y[h, i, s] = cv
    Sm = np.zeros((b, b))
    for s in range(b):
        for t in range(b):
            hi, lo = max(s, t), min(s, t)
            v = G[mT + hi, hi - lo]
            v -= sum(Y[0, i, s] * dT[mT - b + i] * Y[0, i, t] for i in range(b))
            v -= sum(Y[1, i, s] * dB[mB - b + i] * Y[1, i, t] for i in range(b))
            Sm[s, t] = v
    Sinv = np.linalg.inv(Sm)
    # the separator matrices of k_geo_schur: x_S = M v, tails -= N x_S
    M = np.hstack([Sinv, -Sinv @ Y[0].T, -Sinv @ Y[1].T])
    N = np.vstack([dT[mT - b:, None] * Y[0], dB[mB - b:, None] * Y[1]])
    # forward sweeps and scaling
    fT = f[:mT].copy()
    fB = f[::-1][:mB].copy()  # bottom half reversed: row p' = n-1-p'
    zT = np.linalg.solve(LT, fT) * dT
    zB = np.linalg.solve(LB, fB) * dB
    v = np.concatenate([f[mT:mT + b], zT[mT - b:], zB[mB - b:]])
    xs = M @ v
    corr = N @ xs
    zT[mT - b:] -= corr[:b]
    zB[mB - b:] -= corr[b:]
    xT = np.linalg.solve(LT.T, zT)
    xB = np.linalg.solve(LB.T, zB)
    x = np.zeros(n)
    x[:mT] = xT
    x[mT:mT + b] = xs
    x[mT + b:] = xB[::-1]
    return x


def random_band_spd(n, b, rng):
    # D D^T with D banded-ish: SPD with lower bandwidth b
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - b), i + 1):
            A[i, j] = A[j, i] = rng.uniform(-1, 1) * np.exp(-0.3 * (i - j))
        A[i, i] = 0.0
    A += np.diag(np.abs(A).sum(axis=1) + 0.5)  # diagonally dominant: SPD
    return A


@pytest.mark.parametrize("n,b", [(40, 3), (97, 5), (160, 12)])
def test_separator_solve_matches_dense(n, b):
    rng = np.random.default_rng(n * 100 + b)
    A = random_band_spd(n, b, rng)
    G = band_of(A, b)
    f = rng.standard_normal(n)
    x = separator_solve(G, b, f)
    ref = np.linalg.solve(A, f)
    assert np.max(np.abs(x - ref)) < 1e-11 * np.max(np.abs(ref))


def test_separator_solve_on_an_ib_shaped_system():
    """D D^T of a ring of markers with the smoothed 3-point kernel's support
    (5 nodes per axis), in zig-zag order: bandwidth as in the solver."""
    rng = np.random.default_rng(3)
    ns = 160  # about 0.55 lattice units apart, as the workloads' markers
    t = np.linspace(0, 2 * np.pi, ns, endpoint=False)
    X = 40 + 20 * np.cos(t) + 0.01 * rng.standard_normal(ns)
    Y = 40 + 3 * np.sin(t)
    # zig-zag order from t = 0 outwards along both sides
    order, lo, hi = [], 0, ns - 1
    while lo <= hi:
        order.append(lo); lo += 1
        if lo <= hi:
            order.append(hi); hi -= 1
    order = np.array(order)

    def phi(r):
        r = np.abs(r)
        return np.where(r <= 1.5, np.maximum(0, 1 - r / 1.5) ** 2 + 0.1, 0.0) * (r <= 2.5)

    nodes = [(i, j) for i in range(10, 70) for j in range(30, 50)]
    Dm = np.array([[phi(X[k] - i) * phi(Y[k] - j) for (i, j) in nodes] for k in order])
    A = Dm @ Dm.T + 1e-9 * np.eye(ns)
    b = 0
    for p in range(ns):
        nz = np.nonzero(np.abs(A[p, :p + 1]) > 0)[0]
        if len(nz):
            b = max(b, p - nz[0])
    assert 4 <= b <= 40
    G = band_of(A, b)
    f = rng.standard_normal(ns)
    x = separator_solve(G, b, f)
    # backward error: the solve is exact for a system within round-off of A
    res = np.linalg.norm(A @ x - f) / (np.linalg.norm(A, 2) * np.linalg.norm(x))
    assert res < 1e-13
    ref = np.linalg.solve(A, f)
    cond = np.linalg.cond(A)
    assert np.linalg.norm(x - ref) < 1e-14 * cond * np.linalg.norm(ref)
