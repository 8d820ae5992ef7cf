"""Case inputs for the lattice-Boltzmann / immersed-boundary solver.

This module writes and reads the two input files the reference solver keeps at
its boundary and builds the parameter sets of the standard workloads:

* ``boundary.dat`` -- one closed marker polygon per wing in chord units, body
  frame, pivot at the origin, one ``(2f14.8)`` record per point
  (reference ``airfoil.f90:207-242``).  ``naca4_polygon`` generates such a
  polygon for a symmetric NACA 4-digit section with points evenly spaced by
  arclength, ordered like the reference's ``boundary.dat``: leading edge,
  upper surface to the trailing edge, lower surface back to the leading edge.
* ``motion_params.dat`` -- whitespace separated rows ``time, 9 columns per
  wing`` with the per-wing column order angle, omega, alpha-dot, sa, sau,
  saac, sb, sbv, sbac (reference ``mode.f90:68-174``).  Time is in chord/U
  units; positions in chords, velocities in units of U.

Lattice conventions follow ``ini_mesh.f90:14-67``: node (i, j) (1-based) sits
at physical position ``(ox + scale*(i-1), oy + scale*(j-1))`` with
``scale = 1/chord_lu``; viscosity ``nu = U*(1/scale)/Re``, ``tau = 3 nu + 1/2``
and ``s8 = s9 = 1/tau``.
"""
from __future__ import annotations

import dataclasses
import functools
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

# MRT rates of the standard workloads for the non-hydrodynamic moments
# (e, eps, q_x, q_y); s1 = s4 = s6 = 0 for the conserved moments.
STANDARD_S = (0.0, 1.64, 1.54, 0.0, 1.9, 0.0, 1.9)
# the reference's own rates (ini_mesh.f90:59-65)
REFERENCE_S = (0.0, 1.1, 1.05, 0.0, 1.2, 0.0, 1.2)


def naca_thickness(x: np.ndarray, t: float = 0.12, closed_te: bool = True) -> np.ndarray:
    """Half-thickness of a symmetric NACA 4-digit section at chordwise x in [0,1].

    ``closed_te`` uses the -0.1036 quartic coefficient so the trailing edge
    closes exactly (the reference's ``boundary.dat`` ends at y = 0).
    """
    a4 = -0.1036 if closed_te else -0.1015
    return 5.0 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x**2 + 0.2843 * x**3 + a4 * x**4)


def naca4_polygon(n_markers: int, t: float = 0.12, pivot: float = 0.25,
                  fine: int = 400_000) -> Tuple[np.ndarray, np.ndarray]:
    """Closed NACA 00xx polygon with ``n_markers`` segments of equal arclength
    (see :func:`_naca4_polygon`; computed once per argument set)."""
    px, py = _naca4_polygon(n_markers, t, pivot, fine)
    return np.array(px), np.array(py)


@functools.lru_cache(maxsize=64)
def _naca4_polygon(n_markers: int, t: float, pivot: float, fine: int):
    """Closed NACA 00xx polygon with ``n_markers`` segments of equal arclength.

    Returns ``n_markers + 1`` points (first == last) in chord units with the
    pivot at the origin, quantised to the ``(2f14.8)`` text format the
    reference reads (airfoil.f90:229).  Marker k is the midpoint of segment k.
    """
    if n_markers < 8 or n_markers % 2:
        raise ValueError("n_markers must be an even number >= 8")
    # cosine-clustered fine sampling of the upper surface, LE -> TE
    beta = np.linspace(0.0, math.pi, fine)
    xu = 0.5 * (1.0 - np.cos(beta))
    yu = naca_thickness(xu, t)
    arc = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(xu), np.diff(yu)))])
    half = n_markers // 2
    target = np.linspace(0.0, arc[-1], half + 1)
    ux = np.interp(target, arc, xu)
    uy = np.interp(target, arc, yu)
    ux[0], uy[0], ux[-1], uy[-1] = 0.0, 0.0, 1.0, 0.0
    ux, uy = quantise_f14_8(ux - pivot), quantise_f14_8(uy)
    # LE -> upper -> TE, then the mirrored lower surface back to the LE
    px = np.concatenate([ux, ux[::-1][1:]])
    py = np.concatenate([uy, -uy[::-1][1:]])
    return tuple(px.tolist()), tuple(py.tolist())


def quantise_f14_8(a: np.ndarray) -> np.ndarray:
    """Round-trip through the ``f14.8`` text format."""
    return np.array([float("%14.8f" % v) for v in np.asarray(a, dtype=np.float64)])


def write_boundary(path: str, x: Sequence[float], y: Sequence[float]) -> None:
    with open(path, "w") as fh:
        for xv, yv in zip(x, y):
            fh.write("%14.8f%14.8f\n" % (xv, yv))


def read_boundary(path: str, n_points: int) -> Tuple[np.ndarray, np.ndarray]:
    """Read ``n_points`` records of ``(2f14.8)`` (airfoil.f90:224-234)."""
    xs, ys = [], []
    with open(path) as fh:
        for line in fh:
            if len(xs) == n_points:
                break
            if len(line.rstrip("\n")) < 28:
                raise ValueError("boundary file record too short: %r" % line)
            xs.append(float(line[0:14]))
            ys.append(float(line[14:28]))
    if len(xs) != n_points:
        raise ValueError("boundary file has %d points, expected %d" % (len(xs), n_points))
    return np.array(xs), np.array(ys)


def write_motion(path: str, table: np.ndarray) -> None:
    """Write a motion table so that reading it back gives the same doubles."""
    with open(path, "w") as fh:
        for row in np.asarray(table, dtype=np.float64):
            fh.write(" ".join("%.17g" % v for v in row) + "\n")


def read_motion(path: str, n_wings: int) -> np.ndarray:
    """Parse ``motion_params.dat`` like load_motion_table (mode.f90:68-142)."""
    ncols = 1 + 9 * n_wings
    rows = []
    with open(path) as fh:
        for line in fh:
            if not line.strip():
                continue
            parts = line.split()
            if len(parts) < ncols:
                raise ValueError("motion table row has %d columns, expected %d" % (len(parts), ncols))
            rows.append([float(p) for p in parts[:ncols]])
    if not rows:
        raise ValueError("motion table is empty")
    t = np.array(rows, dtype=np.float64)
    if np.any(np.diff(t[:, 0]) < 0):
        raise ValueError("motion table time column must be non-decreasing")
    return t


def sinusoid_motion(n_steps: int, dt: float, heave: float, pitch_deg: float, freq: float,
                    phase_deg: float = 90.0, n_wings: int = 1) -> np.ndarray:
    """Per-step motion table for heave h(t) = heave*sin(2 pi f t) and pitch
    theta(t) = pitch*sin(2 pi f t + phase), sampled at t_n = (n-1)*dt so the
    reference's table lookup hits rows exactly (mode.f90:230-240).

    ``heave`` in chords, ``freq`` in 1/(chord/U), ``dt`` = scale*U.  Positive
    angle is nose-up (clockwise, airfoil.f90:261-263).
    """
    t = np.arange(n_steps, dtype=np.float64) * dt
    w = 2.0 * math.pi * freq
    th0 = math.radians(pitch_deg)
    ph = math.radians(phase_deg)
    cols = [t]
    for _ in range(n_wings):
        ang = th0 * np.sin(w * t + ph)
        om = th0 * w * np.cos(w * t + ph)
        oa = -th0 * w * w * np.sin(w * t + ph)
        z = np.zeros_like(t)
        h = heave * np.sin(w * t)
        hv = heave * w * np.cos(w * t)
        ha = -heave * w * w * np.sin(w * t)
        cols += [ang, om, oa, z, z, z, h, hv, ha]
    return np.stack(cols, axis=1)


@dataclasses.dataclass
class CaseParams:
    """One solver case.  Field names follow the reference's globals."""
    nx: int
    ny: int
    ox: float
    oy: float
    scale: float
    U: float
    s: Tuple[float, ...]
    smooth: int = 1
    n_wings: int = 1
    pts_per_wing: int = 129
    xc: float = 0.0
    yc: float = 0.0
    xp: float = 0.0
    yp: float = 0.0

    @property
    def dt(self) -> float:
        return self.scale * self.U


@dataclasses.dataclass
class Dynamics:
    """Rigid-body dynamics of wing 1 (solve_dynamic_equations, main.f90:101-129):
    x-translation driven by the IB force, optionally a pitch equation with
    par1/par2 (zero in the reference's build)."""
    enabled: bool = True
    rk4: bool = True
    closed_boundary: bool = True
    solid_den: float = 10.0
    area_airfoil: float = 6.807946476488794e-2
    par1: float = 0.0
    par2: float = 0.0
    alphad: float = 10.0 * math.pi / 180.0


def relaxation_rates(U: float, scale: float, Re: float,
                     fixed: Sequence[float] = STANDARD_S) -> Tuple[float, ...]:
    """s(1..9) with s8 = s9 = 1/tau, tau = 3 nu + 1/2, nu = U (1/scale)/Re
    (ini_mesh.f90:30-31, 59-67)."""
    niu = U * (1.0 / scale) / Re
    tau = 3.0 * niu + 0.5
    return tuple(fixed) + (1.0 / tau, 1.0 / tau)


@dataclasses.dataclass
class Case:
    params: CaseParams
    xb0: np.ndarray       # polygon points, all wings concatenated
    yb0: np.ndarray
    motion: np.ndarray    # rows x (1 + 9 n_wings)
    n_steps: int
    name: str = ""


def make_case(nx: int, ny: int, chord: float, n_markers: int, Re: float, n_steps: int,
              U: float = 0.05, le: Tuple[float, float] = None, heave: float = 0.5,
              pitch_deg: float = 0.0, phase_deg: float = 90.0, strouhal: float = 0.3,
              s_fixed: Sequence[float] = STANDARD_S, smooth: int = 1, name: str = "") -> Case:
    """A single-wing NACA0012 case.

    ``chord`` in lattice units; ``le`` the leading-edge lattice position
    (0-based node coordinates) at zero pitch, default (nx/4, ny/2).  The
    pivot is the quarter chord.  ``heave`` is the plunge amplitude in chords;
    the frequency follows from the Strouhal number on twice the heave
    amplitude (which equals the chord-based value for heave = c/2):
    f = St U / (2 h0) lattice, i.e. St/(2 heave) per chord/U time unit.
    """
    scale = 1.0 / chord
    if le is None:
        le = (nx / 4.0, ny / 2.0)
    # pivot at quarter chord: physical x of the LE is -0.25 (chord units)
    ox = -0.25 - le[0] * scale
    oy = -le[1] * scale
    s = relaxation_rates(U, scale, Re, s_fixed)
    xb, yb = naca4_polygon(n_markers)
    freq = strouhal / (2.0 * heave) if heave > 0 else strouhal
    motion = sinusoid_motion(n_steps, scale * U, heave, pitch_deg, freq, phase_deg)
    p = CaseParams(nx=nx, ny=ny, ox=ox, oy=oy, scale=scale, U=U, s=s, smooth=smooth,
                   n_wings=1, pts_per_wing=n_markers + 1)
    return Case(p, xb, yb, motion, n_steps, name)


def config(index: int, n_steps: Optional[int] = None, scale_down: int = 1) -> Case:
    """The standard workloads.

    0: 400x200, chord 40, 128 markers, Re 100, plunge 0.5c, St 0.3, 4000 steps.
    1: 4096x2048, chord 256, 1024 markers, Re 1000, plunge 0.5c + pitch 30 deg
       leading by 90 deg, St 0.3, 20000 steps.
    2: 8192x4096, chord 512, 2048 markers, Re 2000, the kinematics of 1, 50000 steps.
    U = 0.05 throughout.  ``scale_down`` divides lattice, chord and markers
    (keeping Re) for reduced test versions.
    """
    d = scale_down
    if index == 0:
        c = make_case(400 // d, 200 // d, 40.0 / d, 128 // d, 100.0, n_steps or 4000,
                      le=(100.0 / d, 100.0 / d), heave=0.5, pitch_deg=0.0, name="cfg0_400x200")
    elif index == 1:
        c = make_case(4096 // d, 2048 // d, 256.0 / d, 1024 // d, 1000.0, n_steps or 20000,
                      heave=0.5, pitch_deg=30.0, phase_deg=90.0, name="cfg1_4096x2048")
    elif index == 2:
        c = make_case(8192 // d, 4096 // d, 512.0 / d, 2048 // d, 2000.0, n_steps or 50000,
                      heave=0.5, pitch_deg=30.0, phase_deg=90.0, name="cfg2_8192x4096")
    else:
        raise ValueError("unknown config %d" % index)
    return c


def reference_boundary_case(boundary_xy: Tuple[np.ndarray, np.ndarray], nx: int = 241, ny: int = 121,
                            scale: float = 0.025, Re: float = 100.0, n_steps: int = 50,
                            U: float = 0.05) -> Case:
    """A case on the reference's own 205-point ``boundary.dat`` polygon."""
    xb, yb = boundary_xy
    ox = -(nx // 3) * scale
    oy = -(ny // 2) * scale
    s = relaxation_rates(U, scale, Re, REFERENCE_S)
    motion = sinusoid_motion(n_steps, scale * U, 0.5, 10.0, 0.3)
    p = CaseParams(nx=nx, ny=ny, ox=ox, oy=oy, scale=scale, U=U, s=s, n_wings=1,
                   pts_per_wing=len(xb))
    return Case(p, np.asarray(xb), np.asarray(yb), motion, n_steps, "reference_boundary")


def tandem_reference_case(boundary_xy: Tuple[np.ndarray, np.ndarray], nx: int = 321, ny: int = 121,
                          scale: float = 0.025, Re: float = 100.0, n_steps: int = 40, U: float = 0.05,
                          spacing: float = 1.6) -> Case:
    """The reference's two-wing setup (wing_count = 2, main.f90:18; its
    boundary.dat and boundary1.dat hold the same 205-point polygon): wing 2
    trails wing 1 by ``spacing`` chords (a constant in its sa column,
    mode.f90:161-172) and heaves in antiphase."""
    xb, yb = np.asarray(boundary_xy[0]), np.asarray(boundary_xy[1])
    ox = -(nx // 4) * scale
    oy = -(ny // 2) * scale
    s = relaxation_rates(U, scale, Re, REFERENCE_S)
    m = sinusoid_motion(n_steps, scale * U, 0.5, 10.0, 0.3, n_wings=2)
    w2 = 1 + 9
    m[:, w2 + 3] += spacing
    m[:, w2 + 6:w2 + 9] *= -1.0
    p = CaseParams(nx=nx, ny=ny, ox=ox, oy=oy, scale=scale, U=U, s=s, n_wings=2, pts_per_wing=len(xb))
    return Case(p, np.concatenate([xb, xb]), np.concatenate([yb, yb]), m, n_steps, "tandem_reference")


SWEEP_SEED = 20261203


# Stability rule of the drawn cases: a draw is accepted only when the BGK
# relaxation time tau = 3 nu + 1/2 is at least TAU_MIN and the fastest marker
# moves at no more than MACH_MAX times the lattice sound speed 1/sqrt(3).
TAU_MIN = 0.51
MACH_MAX = 0.2


def peak_marker_mach(xb: np.ndarray, yb: np.ndarray, U: float, heave: float, pitch_deg: float,
                     phase_deg: float, freq: float, samples: int = 720) -> float:
    """Largest speed of any polygon point over one period of the motion of
    :func:`sinusoid_motion` (heave ``heave`` chords, pitch ``pitch_deg``
    about the origin leading by ``phase_deg``, frequency ``freq`` per chord/U),
    in lattice units (U times the chord-unit velocity) over the lattice sound
    speed 1/sqrt(3)."""
    t = np.linspace(0.0, 1.0 / freq, samples, endpoint=False)[:, None]
    w, th0, ph = 2.0 * math.pi * freq, math.radians(pitch_deg), math.radians(phase_deg)
    th = th0 * np.sin(w * t + ph)
    thd = th0 * w * np.cos(w * t + ph)
    hd = heave * w * np.cos(w * t)
    x, y = np.asarray(xb)[None, :], np.asarray(yb)[None, :]
    vx = thd * (-np.sin(th) * x + np.cos(th) * y)
    vy = thd * (-np.cos(th) * x - np.sin(th) * y) + hd
    return float(U * np.sqrt(vx * vx + vy * vy).max() * math.sqrt(3.0))


def draw_case(rng: np.random.Generator, n_steps: int, nx_max: int = 2048, U: float = 0.05,
              name: str = "") -> Case:
    """One parameter-sweep case drawn from ``rng``:

    * nx uniform in {512, 640, ..., nx_max}, ny = nx/2, chord nx/16 lattice units
      with 4 markers per lattice unit of chord (4*chord markers);
    * Re uniform in [100, 1000] on the chord; U = 0.05 lattice units per step;
    * heave amplitude uniform in [0.25, 1.0] chords, pitch amplitude uniform in
      [0, 45] deg about the quarter chord, pitch leading heave by a phase
      uniform in [0, 180] deg;
    * Strouhal number uniform in [0.1, 0.5], defined on twice the heave
      amplitude: f = St U / (2 h0).

    Draws that break the stability rule (tau >= TAU_MIN and peak marker Mach
    number <= MACH_MAX) are discarded and drawn again, all six parameters
    together, so the accepted cases are the uniform draws conditioned on the
    rule.  Small chords at high Re (tau near 1/2) and high Strouhal numbers
    with large pitch (trailing-edge speeds near 0.25 c_s) are the draws it
    removes.
    """
    while True:
        nx = int(rng.choice(np.arange(512, nx_max + 1, 128)))
        Re = float(rng.uniform(100.0, 1000.0))
        heave = float(rng.uniform(0.25, 1.0))
        pitch = float(rng.uniform(0.0, 45.0))
        phase = float(rng.uniform(0.0, 180.0))
        st = float(rng.uniform(0.1, 0.5))
        chord = nx / 16.0
        tau = 3.0 * U * chord / Re + 0.5
        if tau < TAU_MIN:
            continue
        xb, yb = naca4_polygon(int(4 * chord))
        if peak_marker_mach(xb, yb, U, heave, pitch, phase, st / (2.0 * heave)) > MACH_MAX:
            continue
        break
    return make_case(nx, nx // 2, chord, int(4 * chord), Re, n_steps, U=U, heave=heave, pitch_deg=pitch,
                     phase_deg=phase, strouhal=st,
                     name=name or "sweep_%dx%d_Re%.0f_h%.2f_p%.0f_ph%.0f_St%.2f" % (nx, nx // 2, Re, heave, pitch,
                                                                                   phase, st))


def sweep_cases(n_cases: int = 96, n_steps: int = 10000, seed: int = SWEEP_SEED, nx_max: int = 2048) -> List[Case]:
    """The batched-form workload: ``n_cases`` independent cases drawn by
    :func:`draw_case` from ``numpy.random.default_rng(seed)``."""
    rng = np.random.default_rng(seed)
    return [draw_case(rng, n_steps, nx_max) for _ in range(n_cases)]


@dataclasses.dataclass
class Request:
    """One request of the serving workload: a case and its arrival time (s)."""
    case: Case
    arrival: float
    index: int


def serving_requests(n_requests: int = 500, rate_per_s: float = 1.0, seed: int = SWEEP_SEED + 1,
                     nx_max: int = 4096, steps_lo: int = 2000, steps_hi: int = 20000) -> List[Request]:
    """The serving workload: cases drawn as :func:`draw_case` with nx up to
    ``nx_max`` and a step count log-uniform in [steps_lo, steps_hi], arriving
    as a Poisson process of rate ``rate_per_s`` (exponential gaps)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n_requests):
        t += float(rng.exponential(1.0 / rate_per_s))
        steps = int(round(math.exp(rng.uniform(math.log(steps_lo), math.log(steps_hi)))))
        out.append(Request(draw_case(rng, steps, nx_max, name="req%03d" % k), t, k))
    return out
