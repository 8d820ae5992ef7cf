"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblbmib_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int),
                ("ox", C.c_double), ("oy", C.c_double), ("scale", C.c_double), ("U", C.c_double),
                ("xc", C.c_double), ("yc", C.c_double), ("xp", C.c_double), ("yp", C.c_double),
                ("s", C.c_double * 9), ("smooth", C.c_int), ("n_wings", C.c_int),
                ("pts_per_wing", C.c_int), ("solver", C.c_int), ("dense_ib", C.c_int),
                ("n_threads", C.c_int)]


class OrcDynamics(C.Structure):
    _fields_ = [("enabled", C.c_int), ("rk4", C.c_int), ("closed_boundary", C.c_int), ("reserved", C.c_int),
                ("solid_den", C.c_double), ("area_airfoil", C.c_double), ("par1", C.c_double),
                ("par2", C.c_double), ("alphad", C.c_double)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcParams), _dp, _dp, _dp, C.c_int, C.c_int, C.c_char_p]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_steps_done.argtypes = [C.c_void_p]
        L.orc_max_steps.argtypes = [C.c_void_p]
        L.orc_get_fields.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.orc_set_f.argtypes = [C.c_void_p, _dp]
        L.orc_history_cols.argtypes = [C.c_void_p]
        L.orc_get_history.argtypes = [C.c_void_p, _dp, C.c_int]
        L.orc_err.argtypes = [C.c_void_p]
        L.orc_err.restype = C.c_double
        L.orc_ib_count.argtypes = [C.c_void_p]
        L.orc_ib_nodes.argtypes = [C.c_void_p, _ip, _ip, _ip, _dp, _dp, _dp]
        L.orc_n_segments.argtypes = [C.c_void_p]
        L.orc_markers.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.orc_build_ib_set.argtypes = [C.c_void_p, _dp, _dp, _ip, _ip, _ip, C.c_int]
        L.orc_weights.argtypes = [_dp]
        L.orc_minv.argtypes = [_dp]
        L.orc_asin.argtypes = [C.c_double]
        L.orc_asin.restype = C.c_double
        L.orc_delta.argtypes = [C.c_double, C.c_int]
        L.orc_delta.restype = C.c_double
        L.orc_cs.restype = C.c_double
        L.orc_feq.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_feq.restype = C.c_double
        L.orc_motion_step_table.argtypes = [_dp, C.c_int, C.c_int, C.c_double, _dp, C.c_int]
        L.orc_solve.argtypes = [_dp, _dp, _dp, C.c_int, _ip, C.c_int, _dp]
        L.orc_collide_node.argtypes = [_dp, C.c_double, C.c_double, C.c_double, _dp, C.c_double, C.c_double, _dp]
        L.orc_set_dynamics.argtypes = [C.c_void_p, C.POINTER(OrcDynamics)]
        L.orc_body_state.argtypes = [C.c_void_p, _dp]
        L.orc_motion_rows.argtypes = [C.c_void_p, _dp, C.c_int]
        L.orc_move_mesh.argtypes = [C.c_void_p, C.c_int]
        L.orc_nee_node.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                   _dp, _dp]
        L.orc_char_update.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.orc_vorticity.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_vorticity_rate.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.orc_surface_pressure.argtypes = [C.c_void_p, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def weights():
    w = np.zeros(9)
    lib().orc_weights(_d(w))
    return w


def minv():
    m = np.zeros(81)
    lib().orc_minv(_d(m))
    return m.reshape(9, 9)


def asin(x: float) -> float:
    return lib().orc_asin(float(x))


def delta(r: float, smooth: int = 1) -> float:
    return lib().orc_delta(float(r), int(smooth))


def cs() -> float:
    return lib().orc_cs()


def feq(rh, u, v, k):
    return lib().orc_feq(rh, u, v, k)


def collide_node(f, rho: float, u: float, v: float, s, gfx: float = 0.0, gfy: float = 0.0) -> np.ndarray:
    f = np.ascontiguousarray(f, dtype=np.float64)
    s = np.ascontiguousarray(s, dtype=np.float64)
    out = np.zeros(9)
    lib().orc_collide_node(_d(f), rho, u, v, _d(s), gfx, gfy, _d(out))
    return out


def motion_step_table(motion: np.ndarray, dt: float) -> np.ndarray:
    m = np.ascontiguousarray(motion, dtype=np.float64)
    rows, ncols = m.shape
    n = lib().orc_motion_step_table(_d(m), rows, ncols, dt, None, 0)
    out = np.zeros((n, ncols))
    rc = lib().orc_motion_step_table(_d(m), rows, ncols, dt, _d(out), n)
    assert rc == n, rc
    return out


def nee_node(s, edge, nb, f0n) -> np.ndarray:
    """One NEE edge node: edge = (rho, u, v) of the edge, nb = that of the
    neighbour whose populations are f0n."""
    sv = np.ascontiguousarray(s, dtype=np.float64)
    f0 = np.ascontiguousarray(f0n, dtype=np.float64)
    out = np.zeros(9)
    lib().orc_nee_node(_d(sv), edge[0], edge[1], edge[2], nb[0], nb[1], nb[2], _d(f0), _d(out))
    return out


def char_update(side: int, cs_: float, states) -> np.ndarray:
    """Characteristic update of an edge node; states = 3 x (rho, u, v) from
    the edge inwards; side 0 inflow, 1 outflow."""
    a = np.ascontiguousarray(np.asarray(states, dtype=np.float64).reshape(9))
    out = np.zeros(3)
    lib().orc_char_update(side, cs_, _d(a), _d(out))
    return out


def vorticity(u: np.ndarray, v: np.ndarray, U: float, scale: float) -> np.ndarray:
    """output's vorticity column (solution.f90:181-210) of (ny, nx) fields."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    ny, nx = u.shape
    out = np.zeros((ny, nx))
    lib().orc_vorticity(nx, ny, U, scale, _d(u), _d(v), _d(out))
    return out


class VorticityRate:
    """vor_mom_rate's state (solution.f90:1841-1900): vorticity of the last
    two calls, zero before the first, and the rate of the last call."""

    def __init__(self, nx: int, ny: int, U: float, scale: float):
        self.nx, self.ny, self.U, self.scale = nx, ny, U, scale
        self.vor = np.zeros((ny, nx))
        self.vor1 = np.zeros((ny, nx))
        self.dvordt = np.zeros((ny, nx))

    def update(self, u: np.ndarray, v: np.ndarray) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        lib().orc_vorticity_rate(self.nx, self.ny, self.U, self.scale, _d(u), _d(v),
                                 _d(self.vor), _d(self.vor1), _d(self.dvordt))
        return self.dvordt.copy()


def solve(G: np.ndarray, S: np.ndarray, b: np.ndarray, perm: np.ndarray, solver: int) -> np.ndarray:
    G = np.ascontiguousarray(G, dtype=np.float64)
    n = G.shape[0]
    S = np.ascontiguousarray(S, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64).reshape(2, n)
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    y = np.zeros((2, n))
    rc = lib().orc_solve(_d(G), _d(S), _d(b), n, _i(perm), solver, _d(y))
    if rc != 0:
        raise RuntimeError("orc_solve failed %d" % rc)
    return y


class Oracle:
    """One oracle case; mirrors the product handle's observable state."""

    def __init__(self, case, solver: int = 0, dense_ib: int = 0, n_threads: int = 1):
        p = case.params
        self.p = p
        prm = OrcParams()
        prm.nx, prm.ny = p.nx, p.ny
        prm.ox, prm.oy, prm.scale, prm.U = p.ox, p.oy, p.scale, p.U
        prm.xc, prm.yc, prm.xp, prm.yp = p.xc, p.yc, p.xp, p.yp
        for k in range(9):
            prm.s[k] = p.s[k]
        prm.smooth, prm.n_wings, prm.pts_per_wing = p.smooth, p.n_wings, p.pts_per_wing
        prm.solver, prm.dense_ib, prm.n_threads = solver, dense_ib, n_threads
        self._xb = np.ascontiguousarray(case.xb0, dtype=np.float64)
        self._yb = np.ascontiguousarray(case.yb0, dtype=np.float64)
        self._mt = np.ascontiguousarray(case.motion, dtype=np.float64)
        err = C.create_string_buffer(256)
        self.h = lib().orc_create(C.byref(prm), _d(self._xb), _d(self._yb), _d(self._mt),
                                  self._mt.shape[0], self._mt.shape[1], err)
        if not self.h:
            raise ValueError(err.value.decode())
        self.N = p.nx * p.ny

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        rc = lib().orc_step(self.h, n)
        if rc != 0:
            raise RuntimeError("oracle step failed: %d" % rc)

    @property
    def steps_done(self) -> int:
        return lib().orc_steps_done(self.h)

    def max_steps(self) -> int:
        return lib().orc_max_steps(self.h)

    def fields(self):
        nx, ny = self.p.nx, self.p.ny
        rho = np.zeros(self.N); u = np.zeros(self.N); v = np.zeros(self.N)
        f = np.zeros(9 * self.N); f0 = np.zeros(9 * self.N)
        lib().orc_get_fields(self.h, _d(rho), _d(u), _d(v), _d(f), _d(f0))
        sh = (ny, nx)
        return dict(rho=rho.reshape(sh), u=u.reshape(sh), v=v.reshape(sh),
                    f=f.reshape(9, ny, nx), f0=f0.reshape(9, ny, nx))

    def set_f(self, f: np.ndarray):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        lib().orc_set_f(self.h, _d(f))

    def history(self) -> np.ndarray:
        cols = lib().orc_history_cols(self.h)
        n = lib().orc_get_history(self.h, None, 0)
        out = np.zeros((n, cols))
        lib().orc_get_history(self.h, _d(out), n)
        return out

    def err(self) -> float:
        return lib().orc_err(self.h)

    def ib_nodes(self):
        n = lib().orc_ib_count(self.h)
        pi = np.zeros(n, np.int32); pj = np.zeros(n, np.int32); w = np.zeros(n, np.int32)
        gx = np.zeros(n); gy = np.zeros(n); ra = np.zeros(n)
        lib().orc_ib_nodes(self.h, _i(pi), _i(pj), _i(w), _d(gx), _d(gy), _d(ra))
        return dict(i=pi, j=pj, wing=w, glx=gx, gly=gy, rho=ra)

    def markers(self):
        n = lib().orc_n_segments(self.h)
        a = [np.zeros(n) for _ in range(5)]
        lib().orc_markers(self.h, *[_d(x) for x in a])
        return dict(xb2=a[0], yb2=a[1], ub=a[2], vb=a[3], arc_l=a[4])

    def set_dynamics(self, dyn):
        """Switch on the body dynamics (a cases.Dynamics) before the first step."""
        d = OrcDynamics()
        d.enabled, d.rk4, d.closed_boundary = int(dyn.enabled), int(dyn.rk4), int(dyn.closed_boundary)
        d.solid_den, d.area_airfoil = dyn.solid_den, dyn.area_airfoil
        d.par1, d.par2, d.alphad = dyn.par1, dyn.par2, dyn.alphad
        if lib().orc_set_dynamics(self.h, C.byref(d)) != 0:
            raise ValueError("set_dynamics refused")

    def body_state(self) -> np.ndarray:
        """11 per-wing arrays in the restart file's order, then movenum."""
        out = np.zeros(11 * self.p.n_wings + 1)
        lib().orc_body_state(self.h, _d(out))
        return out

    def move_mesh(self, dr: int):
        """Test hook: one domain shift of the current state."""
        lib().orc_move_mesh(self.h, dr)

    def motion_rows(self) -> np.ndarray:
        cols = 1 + 3 * self.p.n_wings
        n = lib().orc_motion_rows(self.h, None, 0)
        out = np.zeros((n, cols))
        lib().orc_motion_rows(self.h, _d(out), n)
        return out

    def surface_pressure(self, pre=None) -> np.ndarray:
        """cal_sur_cp on the last step's markers and IB nodes; ``pre`` (ny, nx)
        replaces the case's own rho/3."""
        n = lib().orc_n_segments(self.h)
        out = np.zeros(n)
        if pre is not None:
            pre = np.ascontiguousarray(pre, dtype=np.float64).reshape(-1)
            assert pre.size == self.N
        lib().orc_surface_pressure(self.h, _d(pre) if pre is not None else None, _d(out))
        return out

    def build_ib_set(self, xb2, yb2, max_nodes=1 << 20):
        xb2 = np.ascontiguousarray(xb2, dtype=np.float64)
        yb2 = np.ascontiguousarray(yb2, dtype=np.float64)
        pi = np.zeros(max_nodes, np.int32); pj = np.zeros(max_nodes, np.int32)
        w = np.zeros(max_nodes, np.int32)
        n = lib().orc_build_ib_set(self.h, _d(xb2), _d(yb2), _i(pi), _i(pj), _i(w), max_nodes)
        if n < 0:
            raise RuntimeError("build_ib_set failed %d" % n)
        return pi[:n], pj[:n], w[:n]
