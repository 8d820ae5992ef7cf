"""The FP32 deviation form's collision (collide_dev_f32: M g from sums and
differences of opposite populations, m (1 - s) + s m_eq, the rational
inverse from shared partial sums) against the FP64 collision of the same
node (collide_node, the reference's operation order, which the GPU path
matches bit for bit), both built for the host from the library's header.

A wrong coefficient anywhere in the FP32 transforms or constants moves the
result by the size of the deviation itself; FP32 round-off moves it by about
1e-7 of it.  The bar is 1e-6 of the largest deviation before or after the
collision, per node (with a non-zero rate on the density the collision pulls
the deviation towards -rho0, far from the input's size).  The
relaxation rates are the reference's (s = 0, 1.64, 1.54, 0, 1.9, 0, 1.9,
s_nu, s_nu) and a set with every rate non-zero, so no term hides behind a
zero rate.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("fp32c") / "libfp32c.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas", "-shared", "-fPIC",
                    "-o", out, os.path.join(HERE, "fp32_collision_shim.cpp")], check=True)
    L = C.CDLL(out)
    dp, fp = C.POINTER(C.c_double), C.POINTER(C.c_float)
    L.collide_both.argtypes = [C.c_int, dp, dp, fp, fp, dp]
    return L


def weights():
    # the reference's REAL(4) weights as the library holds them
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, dtype=np.float32)
    return w.astype(np.float64)


def run(L, s, g):
    W = weights()
    n = g.shape[0]
    g = np.ascontiguousarray(g, dtype=np.float32)
    g32 = np.empty_like(g)
    g64 = np.empty((n, 9))
    s = np.asarray(s, dtype=np.float64)
    L.collide_both(n, s.ctypes.data_as(C.POINTER(C.c_double)), W.ctypes.data_as(C.POINTER(C.c_double)),
                   g.ctypes.data_as(C.POINTER(C.c_float)), g32.ctypes.data_as(C.POINTER(C.c_float)),
                   g64.ctypes.data_as(C.POINTER(C.c_double)))
    return g32, g64


def states(n, seed, scale):
    """Deviations of flow states near rest: an equilibrium of density 1 + dr
    and velocity (u, v) (|u| up to 0.1), plus a non-equilibrium part."""
    rng = np.random.default_rng(seed)
    W = weights()
    ex = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
    ey = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
    dr = rng.uniform(-0.02, 0.02, n)[:, None]
    u = rng.uniform(-0.1, 0.1, n)[:, None]
    v = rng.uniform(-0.1, 0.1, n)[:, None]
    cu = 3 * (ex * u + ey * v)
    feq = W * (1 + dr) * (1 + cu + 0.5 * cu ** 2 - 1.5 * (u ** 2 + v ** 2))
    g = feq - W + scale * W * rng.standard_normal((n, 9))
    return g.astype(np.float32)


@pytest.mark.parametrize("s", [
    [0.0, 1.64, 1.54, 0.0, 1.9, 0.0, 1.9, 1.2, 1.2],
    [0.3, 1.1, 1.3, 0.7, 1.5, 0.9, 1.7, 1.95, 0.6],
])
@pytest.mark.parametrize("scale", [1e-3, 1e-2])
def test_fp32_collision_matches_the_fp64_collision(shim, s, scale):
    g = states(4000, 7, scale)
    g32, g64 = run(shim, s, g)
    mag = np.maximum(np.abs(g.astype(np.float64)).max(axis=1, keepdims=True),
                     np.abs(g64).max(axis=1, keepdims=True))
    err = np.abs(g32.astype(np.float64) - g64) / mag
    assert err.max() < 1e-6, err.max()
    # the collision moves the state by far more than the bar: it is tested
    moved = np.abs(g64 - g.astype(np.float64)).max(axis=1) / mag[:, 0]
    assert np.median(moved) > 1e-2


def test_rest_state_is_a_fixed_point(shim):
    # g = 0 is the rest equilibrium f = W: the collision keeps it (to FP32
    # round-off of the rest state's moment constants)
    g32, g64 = run(shim, [0.0, 1.64, 1.54, 0.0, 1.9, 0.0, 1.9, 1.2, 1.2], np.zeros((1, 9), np.float32))
    assert np.abs(g32).max() < 1e-7
    assert np.abs(g64).max() < 1e-15


def test_conserved_moments(shim):
    # with s = 0 on density and momentum the collision conserves them: the
    # sums of g and e g before and after agree to FP32 round-off (1e-6 of the
    # sum of |g|)
    g = states(2000, 11, 1e-2)
    g32, _ = run(shim, [0.0, 1.64, 1.54, 0.0, 1.9, 0.0, 1.9, 1.2, 1.2], g)
    ex = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
    ey = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
    g, g32 = g.astype(np.float64), g32.astype(np.float64)
    size = np.abs(g).sum(axis=1)
    for w in (np.ones(9), ex, ey):
        assert (np.abs(g32 @ w - g @ w) / size).max() < 1e-6
