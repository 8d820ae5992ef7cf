"""Python binding of liblbmib (ctypes over the C ABI in include/lbmib.h).

The class :class:`LBMIB` mirrors the reference solver's per-timestep interface
(module state + subroutines, reference ``main.f90`` / ``solution.f90``):

=====================  ==========================================================
reference              this binding
=====================  ==========================================================
ini_mesh + init +      ``LBMIB(case)`` -> ``lbmib_create`` (+ the polygon of
read_boundary          ``boundary.dat``)
load_motion_table +    ``LBMIB(case)`` -> ``lbmib_set_motion_table``
build_motion_step_table
evolution(mesh, n)     ``evolution(n_steps=1)`` -> ``lbmib_step``
erranalysis            ``erranalysis()`` -> ``lbmib_residual``
force.dat rows         ``history()`` -> ``lbmib_get_history``
output / saveflowfield ``fields()`` -> ``lbmib_download``
readflowfield          ``upload(f)`` -> ``lbmib_upload_state``
=====================  ==========================================================

:class:`Batch` groups cases that then advance together (``lbmib_batch_*``;
the reference has no counterpart: it runs one case per process).

Errors raise :class:`LBMIBError` carrying the library's code and message
(the reference stops the program instead).  The library is loaded from this
package directory; if it is missing the import of this module raises, there is
no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbmib.so")

PREC_FP64 = 0
PREC_FP32DEV = 1
GEOMETRY_BATCH = 16  # steps per geometry batch (LBMIB_GEOMETRY_BATCH)
SOLVER_AUTO = 0       # LBMIB_SOLVER_*: chain for FP64, separator form for FP32
SOLVER_CHAIN = 1
SOLVER_SEPARATOR = 2
NTIMINGS = 16  # LBMIB_NTIMINGS


class LBMIBError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("lbmib error %d: %s" % (code, msg))
        self.code = code


class Params(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32),
                ("ox", C.c_double), ("oy", C.c_double), ("scale", C.c_double), ("U", C.c_double),
                ("xc", C.c_double), ("yc", C.c_double), ("xp", C.c_double), ("yp", C.c_double),
                ("s", C.c_double * 9), ("smooth", C.c_int32), ("precision", C.c_int32),
                ("n_wings", C.c_int32), ("pts_per_wing", C.c_int32), ("device", C.c_int32),
                ("nan_guard", C.c_int32), ("ib_solver", C.c_int32), ("reserved", C.c_int32 * 5)]


class DynamicsParams(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("rk4", C.c_int32), ("closed_boundary", C.c_int32),
                ("reserved", C.c_int32), ("solid_den", C.c_double), ("area_airfoil", C.c_double),
                ("par1", C.c_double), ("par2", C.c_double), ("alphad", C.c_double)]


class LaunchShape(C.Structure):
    _fields_ = [("nodes_per_thread", C.c_int32), ("tile_order", C.c_int32), ("persistent", C.c_int32),
                ("ib_cus", C.c_int32), ("graph", C.c_int32), ("ib_collide_shared", C.c_int32),
                ("ib_rhs_reserved", C.c_int32), ("solve_unit", C.c_int32), ("fuse_steps", C.c_int32)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k in ("nodes_per_thread", "tile_order", "persistent", "ib_cus", "graph",
                                                    "ib_collide_shared", "ib_rhs_reserved", "solve_unit",
                                                    "fuse_steps")}


def make_shape(nodes_per_thread=1, tile_order=0, persistent=0, ib_cus=16, graph=0, ib_collide_shared=0,
               ib_rhs_reserved=1, solve_unit=0, fuse_steps=1) -> LaunchShape:
    s = LaunchShape()
    s.nodes_per_thread, s.tile_order, s.persistent = nodes_per_thread, tile_order, persistent
    s.ib_cus, s.graph, s.ib_collide_shared = ib_cus, graph, ib_collide_shared
    s.ib_rhs_reserved, s.solve_unit, s.fuse_steps = ib_rhs_reserved, solve_unit, fuse_steps
    return s


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lib = None

EXPORTS = ["lbmib_create", "lbmib_set_motion_table", "lbmib_step", "lbmib_steps_done",
           "lbmib_max_steps", "lbmib_get_forces", "lbmib_get_history", "lbmib_download",
           "lbmib_upload_state", "lbmib_set_step", "lbmib_markers", "lbmib_vorticity", "lbmib_surface_pressure",
           "lbmib_surface_pressure_noalloc", "lbmib_set_dynamics", "lbmib_get_body_state",
           "lbmib_set_body_state", "lbmib_get_motion_rows", "lbmib_get_edge_f0", "lbmib_set_edge_f0", "lbmib_residual", "lbmib_ib_nodes", "lbmib_lattice_window", "lbmib_set_profiling",
           "lbmib_get_timings", "lbmib_destroy", "lbmib_last_error", "lbmib_version",
           "lbmib_set_shape", "lbmib_get_shape", "lbmib_batch_create", "lbmib_batch_add",
           "lbmib_batch_remove", "lbmib_batch_size", "lbmib_batch_step", "lbmib_batch_set_shape",
           "lbmib_batch_get_shape", "lbmib_batch_set_profiling", "lbmib_batch_get_timings",
           "lbmib_batch_autotune", "lbmib_batch_destroy", "lbmib_device_bytes", "lbmib_mem_info"]


def load(path: str = LIB_PATH):
    """Load liblbmib.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError("liblbmib.so not built at %s (run __graft_entry__.build())" % path)
    L = C.CDLL(path)
    vp = C.c_void_p
    L.lbmib_create.argtypes = [C.POINTER(Params), _dp, _dp, C.POINTER(vp)]
    L.lbmib_set_motion_table.argtypes = [vp, _dp, C.c_int32, C.c_int32]
    L.lbmib_step.argtypes = [vp, C.c_int32]
    L.lbmib_steps_done.argtypes = [vp, _ip]
    L.lbmib_max_steps.argtypes = [vp, _ip]
    L.lbmib_get_forces.argtypes = [vp, _dp, _dp]
    L.lbmib_get_history.argtypes = [vp, _dp, C.c_int32, _ip]
    L.lbmib_download.argtypes = [vp, _dp, _dp, _dp, _dp]
    L.lbmib_upload_state.argtypes = [vp, _dp]
    L.lbmib_set_step.argtypes = [vp, C.c_int32]
    L.lbmib_markers.argtypes = [vp, _dp, _dp]
    L.lbmib_set_dynamics.argtypes = [vp, C.POINTER(DynamicsParams)]
    L.lbmib_get_body_state.argtypes = [vp, _dp, C.c_int32]
    L.lbmib_set_body_state.argtypes = [vp, _dp, C.c_int32]
    L.lbmib_get_motion_rows.argtypes = [vp, _dp, C.c_int32, C.POINTER(C.c_int32)]
    L.lbmib_get_edge_f0.argtypes = [vp, _dp, _dp]
    L.lbmib_set_edge_f0.argtypes = [vp, _dp, _dp]
    L.lbmib_vorticity.argtypes = [vp, C.c_int32, _dp, _dp]
    L.lbmib_surface_pressure.argtypes = [vp, _dp, C.c_int32]
    L.lbmib_surface_pressure_noalloc.argtypes = [vp, _dp, C.c_int32]
    L.lbmib_lattice_window.argtypes = [vp, _ip, _ip, _ip, _ip]
    L.lbmib_residual.argtypes = [vp, _dp]
    L.lbmib_ib_nodes.argtypes = [vp, _ip, _ip, _ip, _ip, _dp, _dp]
    L.lbmib_set_profiling.argtypes = [vp, C.c_int32]
    L.lbmib_get_timings.argtypes = [vp, _dp, C.c_int32]
    L.lbmib_destroy.argtypes = [vp]
    sp = C.POINTER(LaunchShape)
    L.lbmib_set_shape.argtypes = [vp, sp]
    L.lbmib_get_shape.argtypes = [vp, sp]
    L.lbmib_batch_create.argtypes = [C.c_int32, C.POINTER(vp)]
    L.lbmib_batch_add.argtypes = [vp, vp]
    L.lbmib_batch_remove.argtypes = [vp, vp]
    L.lbmib_batch_size.argtypes = [vp, _ip]
    L.lbmib_batch_step.argtypes = [vp, C.c_int32]
    L.lbmib_batch_set_shape.argtypes = [vp, sp]
    L.lbmib_batch_get_shape.argtypes = [vp, sp]
    L.lbmib_batch_set_profiling.argtypes = [vp, C.c_int32]
    L.lbmib_batch_get_timings.argtypes = [vp, _dp, C.c_int32]
    L.lbmib_batch_autotune.argtypes = [vp, C.c_int32, sp, C.c_char_p, C.c_int32]
    L.lbmib_batch_destroy.argtypes = [vp]
    L.lbmib_device_bytes.argtypes = [vp, C.POINTER(C.c_int64)]
    L.lbmib_mem_info.argtypes = [C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.lbmib_last_error.restype = C.c_char_p
    L.lbmib_version.restype = C.c_char_p
    _lib = L
    return L


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _check(rc: int):
    if rc != 0:
        raise LBMIBError(rc, load().lbmib_last_error().decode())


def make_params(p, precision: int = PREC_FP64, device: int = 0, nan_guard: int = 0,
                ib_solver: int = SOLVER_AUTO) -> Params:
    prm = Params()
    prm.nx, prm.ny = p.nx, p.ny
    prm.ox, prm.oy, prm.scale, prm.U = p.ox, p.oy, p.scale, p.U
    prm.xc, prm.yc, prm.xp, prm.yp = p.xc, p.yc, p.xp, p.yp
    for k in range(9):
        prm.s[k] = p.s[k]
    prm.smooth, prm.precision = p.smooth, precision
    prm.n_wings, prm.pts_per_wing = p.n_wings, p.pts_per_wing
    prm.device, prm.nan_guard, prm.ib_solver = device, nan_guard, ib_solver
    return prm


class LBMIB:
    """One solver case resident on one GPU."""

    def __init__(self, case, precision: int = PREC_FP64, device: int = 0, nan_guard: int = 0,
                 ib_solver: int = SOLVER_AUTO, dynamics=None):
        L = load()
        self.case = case
        self.p = case.params
        prm = make_params(self.p, precision, device, nan_guard, ib_solver)
        xb = np.ascontiguousarray(case.xb0, dtype=np.float64)
        yb = np.ascontiguousarray(case.yb0, dtype=np.float64)
        h = C.c_void_p()
        _check(L.lbmib_create(C.byref(prm), _d(xb), _d(yb), C.byref(h)))
        self.h = h
        if dynamics is not None:  # a cases.Dynamics
            d = DynamicsParams()
            d.enabled, d.rk4 = int(dynamics.enabled), int(dynamics.rk4)
            d.closed_boundary, d.reserved = int(dynamics.closed_boundary), 0
            d.solid_den, d.area_airfoil = dynamics.solid_den, dynamics.area_airfoil
            d.par1, d.par2, d.alphad = dynamics.par1, dynamics.par2, dynamics.alphad
            _check(L.lbmib_set_dynamics(self.h, C.byref(d)))
        mt = np.ascontiguousarray(case.motion, dtype=np.float64)
        _check(L.lbmib_set_motion_table(self.h, _d(mt), mt.shape[0], mt.shape[1]))
        self.N = self.p.nx * self.p.ny

    def close(self):
        if getattr(self, "h", None):
            load().lbmib_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # reference: evolution (solution.f90:269-410), n_steps times
    def evolution(self, n_steps: int = 1):
        _check(load().lbmib_step(self.h, int(n_steps)))

    step = evolution

    @property
    def steps_done(self) -> int:
        n = C.c_int32()
        _check(load().lbmib_steps_done(self.h, C.byref(n)))
        return n.value

    def max_steps(self) -> int:
        n = C.c_int32()
        _check(load().lbmib_max_steps(self.h, C.byref(n)))
        return n.value

    # reference: erranalysis (solution.f90:74-127)
    def erranalysis(self) -> float:
        e = C.c_double()
        _check(load().lbmib_residual(self.h, C.byref(e)))
        return e.value

    def forces(self):
        nw = self.p.n_wings
        wf = np.zeros(2 * nw)
        wm = np.zeros(nw)
        _check(load().lbmib_get_forces(self.h, _d(wf), _d(wm)))
        return wf.reshape(nw, 2), wm

    def history(self) -> np.ndarray:
        n = C.c_int32()
        _check(load().lbmib_get_history(self.h, None, 0, C.byref(n)))
        cols = 4 + 12 * self.p.n_wings
        out = np.zeros((n.value, cols))
        if n.value:
            _check(load().lbmib_get_history(self.h, _d(out), n.value, C.byref(n)))
        return out

    def fields(self, with_f: bool = True):
        nx, ny = self.p.nx, self.p.ny
        rho = np.zeros(self.N); u = np.zeros(self.N); v = np.zeros(self.N)
        f = np.zeros(9 * self.N) if with_f else None
        _check(load().lbmib_download(self.h, _d(rho), _d(u), _d(v), _d(f) if with_f else None))
        out = dict(rho=rho.reshape(ny, nx), u=u.reshape(ny, nx), v=v.reshape(ny, nx))
        if with_f:
            out["f"] = f.reshape(9, ny, nx)
        return out

    def upload(self, f: np.ndarray, step: Optional[int] = None):
        """Populations of the state after `step` steps (default: the
        current step count): readflowfield's restore of a saved run."""
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        if f.size != 9 * self.N:
            raise ValueError("f must have 9*nx*ny entries")
        if step is not None:
            _check(load().lbmib_set_step(self.h, int(step)))
        _check(load().lbmib_upload_state(self.h, _d(f)))

    def markers(self):
        """Marker polygon points (x, y) of every wing at the last step taken."""
        m = self.p.n_wings * self.p.pts_per_wing
        x = np.zeros(m); y = np.zeros(m)
        _check(load().lbmib_markers(self.h, _d(x), _d(y)))
        return x, y

    def body_state(self) -> np.ndarray:
        """11 per-wing arrays in the restart file's order, then movenum."""
        out = np.zeros(11 * self.p.n_wings + 1)
        _check(load().lbmib_get_body_state(self.h, _d(out), out.size))
        return out

    def set_body_state(self, state):
        st = np.ascontiguousarray(state, dtype=np.float64)
        _check(load().lbmib_set_body_state(self.h, _d(st), st.size))

    def edge_f0(self):
        """Stale f0 of the first and last columns, each (9, ny)."""
        a = np.zeros((9, self.p.ny)); b = np.zeros((9, self.p.ny))
        _check(load().lbmib_get_edge_f0(self.h, _d(a), _d(b)))
        return a, b

    def set_edge_f0(self, first, last):
        a = np.ascontiguousarray(first, dtype=np.float64)
        b = np.ascontiguousarray(last, dtype=np.float64)
        _check(load().lbmib_set_edge_f0(self.h, _d(a), _d(b)))

    def motion_rows(self) -> np.ndarray:
        """airfoil_motion.dat rows of the steps taken."""
        n = C.c_int32()
        L = load()
        _check(L.lbmib_get_motion_rows(self.h, None, 0, C.byref(n)))
        out = np.zeros((n.value, 1 + 3 * self.p.n_wings))
        if n.value:
            _check(L.lbmib_get_motion_rows(self.h, _d(out), n.value, C.byref(n)))
        return out

    def vorticity(self, rate: bool = False, copy: bool = True):
        """output's vorticity (ny, nx) of the last step; with ``rate`` also one
        vor_mom_rate update and its dvordt.  Returns (vor, dvordt or None);
        with copy=False nothing comes back to the host."""
        p = self.p
        vor = np.zeros((p.ny, p.nx)) if copy else None
        dv = np.zeros((p.ny, p.nx)) if (copy and rate) else None
        _check(load().lbmib_vorticity(self.h, 1 if rate else 0, _d(vor) if vor is not None else None,
                                      _d(dv) if dv is not None else None))
        return vor, dv

    def surface_pressure(self, noalloc: bool = False) -> np.ndarray:
        """cal_sur_cp: pressure at each marker midpoint of the last step.
        With ``noalloc`` through lbmib_surface_pressure_noalloc, which stages
        the values in the library's fixed per-device buffer."""
        n = self.p.n_wings * (self.p.pts_per_wing - 1)
        out = np.zeros(n)
        fn = load().lbmib_surface_pressure_noalloc if noalloc else load().lbmib_surface_pressure
        _check(fn(self.h, _d(out), n))
        return out

    def ib_nodes(self):
        n = C.c_int32()
        L = load()
        _check(L.lbmib_ib_nodes(self.h, C.byref(n), None, None, None, None, None))
        m = n.value
        pi = np.zeros(m, np.int32); pj = np.zeros(m, np.int32); w = np.zeros(m, np.int32)
        gx = np.zeros(m); gy = np.zeros(m)
        _check(L.lbmib_ib_nodes(self.h, C.byref(n), pi.ctypes.data_as(_ip), pj.ctypes.data_as(_ip),
                                w.ctypes.data_as(_ip), _d(gx), _d(gy)))
        return dict(i=pi, j=pj, wing=w, glx=gx, gly=gy)

    def lattice_window(self):
        """(i0, j0, width, height) of the lattice window holding every IB node
        of the run (0-based node indices)."""
        v = [C.c_int32() for _ in range(4)]
        _check(load().lbmib_lattice_window(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_profiling(self, every: int = 1):
        """Record phase timings on every ``every``-th step (0: off); resets them."""
        _check(load().lbmib_set_profiling(self.h, int(every)))

    def timings(self) -> dict:
        """Averages of the recorded phases (milliseconds) and their counts."""
        t = np.zeros(NTIMINGS)
        _check(load().lbmib_get_timings(self.h, _d(t), NTIMINGS))
        return _timings(t)

    def device_bytes(self) -> int:
        """Device memory this case holds."""
        b = C.c_int64()
        _check(load().lbmib_device_bytes(self.h, C.byref(b)))
        return b.value

    def set_shape(self, **kw):
        """Launch shape of this handle's steps (see :func:`make_shape`)."""
        _check(load().lbmib_set_shape(self.h, C.byref(make_shape(**kw))))

    def shape(self) -> dict:
        s = LaunchShape()
        _check(load().lbmib_get_shape(self.h, C.byref(s)))
        return s.as_dict()


def _timings(t) -> dict:
    ns, nb = max(t[2], 1.0), max(t[8], 1.0)
    return dict(step_ms=t[0] / max(t[1], 1.0), lattice_ms=t[3] / max(t[15], 1.0), ib_solve_ms=t[4] / ns,
                ib_collide_ms=t[5] / ns, geometry_batch_ms=t[6] / nb, factor_batch_ms=t[7] / nb,
                steps=int(t[1]), sampled_steps=int(t[2]), batches=int(t[8]),
                vorticity_ms=t[9] / max(t[10], 1.0), vorticity_calls=int(t[10]),
                far_pass_ms=t[11] / max(t[12], 1.0), far_passes=int(t[12]),
                window_step_ms=t[13] / max(t[14], 1.0), lattice_steps=int(t[15]))


class Batch:
    """Cases advanced together on one GPU (``lbmib_batch_*``).

    Members keep their :class:`LBMIB` handles for inputs and results; while
    in the batch they step only through :meth:`step`.
    """

    def __init__(self, device: int = 0):
        L = load()
        b = C.c_void_p()
        _check(L.lbmib_batch_create(int(device), C.byref(b)))
        self.b = b
        self.members = []

    def add(self, sim: "LBMIB"):
        _check(load().lbmib_batch_add(self.b, sim.h))
        self.members.append(sim)

    def remove(self, sim: "LBMIB"):
        _check(load().lbmib_batch_remove(self.b, sim.h))
        self.members.remove(sim)

    def __len__(self):
        n = C.c_int32()
        _check(load().lbmib_batch_size(self.b, C.byref(n)))
        return n.value

    def step(self, n_steps: int):
        _check(load().lbmib_batch_step(self.b, int(n_steps)))

    def set_shape(self, **kw):
        _check(load().lbmib_batch_set_shape(self.b, C.byref(make_shape(**kw))))

    def shape(self) -> dict:
        s = LaunchShape()
        _check(load().lbmib_batch_get_shape(self.b, C.byref(s)))
        return s.as_dict()

    def set_profiling(self, every: int = 1):
        _check(load().lbmib_batch_set_profiling(self.b, int(every)))

    def timings(self) -> dict:
        t = np.zeros(NTIMINGS)
        _check(load().lbmib_batch_get_timings(self.b, _d(t), NTIMINGS))
        return _timings(t)

    def autotune(self, trial_steps: int = 32) -> dict:
        """Pick the fastest launch shape on copies of the members; returns the
        trials and the choice (applied to this batch)."""
        import json
        best = LaunchShape()
        buf = C.create_string_buffer(1 << 16)
        _check(load().lbmib_batch_autotune(self.b, int(trial_steps), C.byref(best), buf, len(buf)))
        return json.loads(buf.value.decode())

    def close(self):
        if getattr(self, "b", None):
            load().lbmib_batch_destroy(self.b)
            self.b = None
            self.members = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mem_info(device: int = 0):
    """(free, total) bytes of device memory."""
    f, t = C.c_int64(), C.c_int64()
    _check(load().lbmib_mem_info(int(device), C.byref(f), C.byref(t)))
    return f.value, t.value


def version() -> str:
    return load().lbmib_version().decode()
