"""Pins the oracle's outer boundary, boundary_condition1
(boundary_condition.f90:288-858), against textbook forms and its own wiring.

* The characteristic update is the locally one-dimensional inviscid (LODI)
  treatment of Thompson and of Poinsot & Lele for an isothermal gas with
  p = c^2 rho: wave amplitudes L1 = (u-c)(p_x - rho c u_x),
  L5 = (u+c)(p_x + rho c u_x), L3 = u v_x from second-order one-sided
  differences, the amplitudes of waves entering the domain set to zero, and
  rho -= (L1+L5)/(2c^2), u -= (L5-L1)/(2 rho c), v -= L3.  The reference
  divides the velocity update by the already updated density; the check
  allows that (1% of the update) and nothing larger.
* The non-equilibrium extrapolation (NEE) keeps the neighbour's
  non-equilibrium moments scaled by 1 - s per moment around the edge's
  equilibrium, whose moments are those of the second-order D2Q9 equilibrium.
* The wiring: wall nodes take the neighbour row's density at rest, edge
  columns the characteristic state, each from the neighbour's streamed
  populations.

A sign, a swapped wave, a wrong side rule, a transposed M or a misplaced s
fails at least one check.
"""
import numpy as np
import pytest

from fortran_loadfile_mitiwing_amd import cases

EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=np.float64)
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.float64)
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
M = np.array([[1, 1, 1, 1, 1, 1, 1, 1, 1],
              [-4, -1, -1, -1, -1, 2, 2, 2, 2],
              [4, -2, -2, -2, -2, 1, 1, 1, 1],
              [0, 1, 0, -1, 0, 1, -1, -1, 1],
              [0, -2, 0, 2, 0, 1, -1, -1, 1],
              [0, 0, 1, 0, -1, 1, 1, -1, -1],
              [0, 0, -2, 0, 2, 1, 1, -1, -1],
              [0, 1, -1, 1, -1, 0, 0, 0, 0],
              [0, 0, 0, 0, 0, 1, -1, 1, -1]], dtype=np.float64)


def feq(rho, u, v):
    eu = EX * u + EY * v
    return W * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * (u * u + v * v))


def test_nee_of_an_equilibrium_neighbour_is_the_edge_equilibrium(oracle_mod):
    rng = np.random.default_rng(5)
    s = cases.relaxation_rates(0.05, 0.01, 1000.0)
    for _ in range(50):
        re, rn = rng.uniform(0.95, 1.05, 2)
        ue, ve, un, vn = rng.uniform(-0.1, 0.1, 4)
        out = oracle_mod.nee_node(s, (re, ue, ve), (rn, un, vn), feq(rn, un, vn))
        assert np.allclose(out, feq(re, ue, ve), rtol=0, atol=2e-15)


def test_nee_keeps_each_non_equilibrium_moment_times_one_minus_s(oracle_mod):
    rng = np.random.default_rng(6)
    for _ in range(50):
        s = rng.uniform(0.5, 1.9, 9)
        re, rn = rng.uniform(0.95, 1.05, 2)
        ue, ve, un, vn = rng.uniform(-0.1, 0.1, 4)
        f0n = feq(rn, un, vn) + rng.normal(0.0, 1e-3, 9)
        out = oracle_mod.nee_node(s, (re, ue, ve), (rn, un, vn), f0n)
        lhs = M @ out - M @ feq(re, ue, ve)
        rhs = (1.0 - s) * (M @ f0n - M @ feq(rn, un, vn))
        assert np.allclose(lhs, rhs, rtol=0, atol=1e-14)


def lodi(side, c, states):
    (r0, u0, v0), (r1, u1, v1), (r2, u2, v2) = states
    sgn = -1.0 if side == 0 else 1.0  # x grows inwards on the inflow side, outwards on the outflow side
    d = lambda a0, a1, a2: sgn * (3.0 * a0 - 4.0 * a1 + a2) * 0.5
    px, ux, vx = c * c * d(r0, r1, r2), d(u0, u1, u2), d(v0, v1, v2)
    lam = {1: u0 - c, 5: u0 + c, 3: u0}
    L = {1: lam[1] * (px - r0 * c * ux), 5: lam[5] * (px + r0 * c * ux), 3: lam[3] * vx}
    for k in L:  # a wave enters through the inflow side when lambda > 0, the outflow side when lambda < 0
        if (side == 0 and lam[k] > 0.0) or (side == 1 and lam[k] < 0.0):
            L[k] = 0.0
    return np.array([r0 - (L[1] + L[5]) / (2.0 * c * c), u0 - (L[5] - L[1]) / (2.0 * r0 * c), v0 - L[3]])


@pytest.mark.parametrize("side", [0, 1])
def test_characteristic_update_is_lodi(oracle_mod, side):
    rng = np.random.default_rng(7 + side)
    c = oracle_mod.cs()
    for _ in range(200):
        # a smooth boundary layer: three nodes close to one state
        base = np.array([rng.uniform(0.97, 1.03), rng.uniform(-0.1, 0.1), rng.uniform(-0.05, 0.05)])
        states = [tuple(base + rng.normal(0.0, 1e-3, 3)) for _ in range(3)]
        got = oracle_mod.char_update(side, c, states)
        want = lodi(side, c, states)
        old = np.array(states[0])
        step = np.abs(want - old)
        assert np.all(np.abs(got - want) <= 1e-2 * step + 1e-15), (states, got, want)
    # a still, uniform state is left alone
    assert np.array_equal(oracle_mod.char_update(side, c, [(1.0, 0.05, 0.0)] * 3), np.array([1.0, 0.05, 0.0]))


@pytest.fixture(scope="module")
def stepped(oracle_mod):
    case = cases.config(0, n_steps=10, scale_down=2)
    o = oracle_mod.Oracle(case)
    o.step(7)
    return case, o


def test_walls_take_the_neighbour_density_at_rest(stepped, oracle_mod):
    case, o = stepped
    fl = o.fields()
    s = case.params.s
    rho, u, v, f, f0 = fl["rho"], fl["u"], fl["v"], fl["f"], fl["f0"]
    ny, nx = rho.shape
    for je, jn in ((0, 1), (ny - 1, ny - 2)):
        assert np.array_equal(rho[je, 1:-1], rho[jn, 1:-1])
        assert np.all(u[je, 1:-1] == 0.0) and np.all(v[je, 1:-1] == 0.0)
        for i in (1, nx // 3, nx - 2):
            want = oracle_mod.nee_node(s, (rho[jn, i], 0.0, 0.0), (rho[jn, i], u[jn, i], v[jn, i]), f0[:, jn, i])
            assert np.array_equal(f[:, je, i], want), (je, i)


def test_edge_columns_take_the_characteristic_state(stepped, oracle_mod):
    case, o = stepped
    fl = o.fields()
    s = case.params.s
    rho, u, v, f, f0 = fl["rho"], fl["u"], fl["v"], fl["f"], fl["f0"]
    ny, nx = rho.shape
    c = oracle_mod.cs()
    for side, ie, i1, i2 in ((0, 0, 1, 2), (1, nx - 1, nx - 2, nx - 3)):
        for j in (1, ny // 2, ny - 2):
            # the edge's own state before the update: moments of its stale f0
            g = f0[:, j, ie]
            r0 = g.sum()
            u0, v0 = (EX * g).sum() / r0, (EY * g).sum() / r0
            st = [(r0, u0, v0), (rho[j, i1], u[j, i1], v[j, i1]), (rho[j, i2], u[j, i2], v[j, i2])]
            assert np.allclose([rho[j, ie], u[j, ie], v[j, ie]], oracle_mod.char_update(side, c, st),
                               rtol=1e-13, atol=1e-16)
            want = oracle_mod.nee_node(s, (rho[j, ie], u[j, ie], v[j, ie]), (rho[j, i1], u[j, i1], v[j, i1]),
                                       f0[:, j, i1])
            assert np.array_equal(f[:, j, ie], want), (side, j)
