"""The HIP FP64 path against the oracle's restatement of the reference's own
immersed-boundary solve.

The reference forms A = D D^T S and factors it with cuSOLVER's LU with
partial pivoting (boundary_condition.f90:1619-1640).  The GPU path solves the
same SPD system D D^T y = b by a banded L D L^T in zig-zag marker order and
is bit-identical to the oracle's solver 0 (tests/test_gpu_parity.py).  The
oracle's solver 1 restates the reference's call as a textbook unblocked LU
with partial pivoting of D D^T S.  Both are exact solvers of one system, so
they differ only in round-off, and that round-off is set by cond(D D^T):
about 1.0e8 at 128 markers on the 400x200 workload and 2.8e9 at 1024 markers
on the 4096x2048 workload (tests/test_ib_conditioning.py recomputes both).

Bars, written here:
  * 400x200 workload, its whole 4000-step run: the stated drop-in bars,
    relative L2 1e-12 on rho and u and 1e-10 on the C_D (C_x) and C_L (C_y)
    histories, all asserted.  v and C_M are printed.
  * 4096x2048 workload: at cond(D D^T) = 2.8e9 two correct LU codes already
    disagree: LAPACK's blocked dgetrf and the textbook unblocked LU give IB
    velocity corrections that differ by 4.7e-10 relative, the L D L^T solve
    and the textbook LU by 6.3e-10 (tests/test_ib_conditioning.py asserts
    that the second is within 4x the first).  After one step the fields
    already differ by 8e-12 (u) and 6e-10 (v); after 24 steps by 2.3e-12
    (rho), 1.5e-11 (u), 1.3e-10 (C_M).  The 1e-12 field bar is below the
    spread among valid implementations of the reference's own solve at this
    conditioning, so it cannot be met against cuSOLVER by any code that is
    not cuSOLVER.  The bars asserted here are 5e-12 (rho), 5e-11 (u) and
    5e-10 (C_D, C_L), about 3x the measured deviation.
  * 4096x2048 workload over 2000 steps, against the committed fixture
    tests/golden/cfg1_long_fp64.npz (both oracle solvers, run side by side by
    tests/golden/make_long_fp64.py): the GPU state after 2000 steps must hash
    to the L D L^T oracle's bit for bit; against the LU oracle the deviation
    is printed and held to 10x the deviation the two oracle solvers show
    against each other in the fixture.
"""
import hashlib
import os

import numpy as np
import pytest

from fortran_loadfile_mitiwing_amd import cases
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

LONG_FIXTURE = os.path.join(GOLDEN, "cfg1_long_fp64.npz")


def rel_l2(a, b):
    a = np.asarray(a).ravel(); b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def deviations(fg, hg, fo, ho):
    d = {k: rel_l2(fg[k], fo[k]) for k in ("rho", "u", "v")}
    for col, name in ((1, "C_D"), (2, "C_L"), (3, "C_M")):
        d[name] = rel_l2(hg[:, col], ho[:, col])
    print("  deviation from the LU oracle:", {k: "%.3g" % v for k, v in d.items()}, flush=True)
    return d


def test_config0_full_run_against_the_lu_solve(oracle_mod, lbmib_mod):
    case = cases.config(0)
    o = oracle_mod.Oracle(case, solver=1, dense_ib=1, n_threads=8)
    g = lbmib_mod.LBMIB(case)
    for _ in range(4):
        print("  1000 steps of %s" % case.name, flush=True)
        o.step(1000)
        g.evolution(1000)
    d = deviations(g.fields(with_f=False), g.history(), o.fields(), o.history())
    assert d["rho"] <= 1e-12 and d["u"] <= 1e-12
    assert d["C_D"] <= 1e-10 and d["C_L"] <= 1e-10


def test_config1_first_steps_against_the_lu_solve(oracle_mod, lbmib_mod):
    steps = 24
    case = cases.config(1, n_steps=steps + 8)
    # dense_ib = 0 evaluates the same products as the literal dense matrices
    # with the exact zeros skipped (bit-identical, tests/test_oracle.py)
    o = oracle_mod.Oracle(case, solver=1, dense_ib=0, n_threads=16)
    g = lbmib_mod.LBMIB(case)
    print("  %d steps of %s" % (steps, case.name), flush=True)
    o.step(steps)
    g.evolution(steps)
    d = deviations(g.fields(with_f=False), g.history(), o.fields(), o.history())
    assert d["rho"] <= 5e-12 and d["u"] <= 5e-11
    assert d["C_D"] <= 5e-10 and d["C_L"] <= 5e-10


@pytest.mark.skipif(not os.path.exists(LONG_FIXTURE), reason="long-run fixture not generated")
def test_config1_2000_steps_against_the_fixture(lbmib_mod):
    fx = np.load(LONG_FIXTURE, allow_pickle=False)
    steps = int(fx["steps"])
    case = cases.config(1, n_steps=steps)
    g = lbmib_mod.LBMIB(case)
    for _ in range(steps // 500):
        print("  500 steps of %s" % case.name, flush=True)
        g.evolution(500)
    g.evolution(steps % 500)
    fg = g.fields()
    h = hashlib.sha256()
    for k in ("rho", "u", "v", "f"):
        h.update(np.ascontiguousarray(fg[k], dtype=np.float64).tobytes())
    hg = g.history()[:, 1:4]
    assert np.array_equal(hg, fx["hist_ldl"]), "force history differs from the L D L^T oracle"
    assert h.hexdigest() == str(fx["sha_ldl"]), "fields differ from the L D L^T oracle"
    st = int(fx["sample_stride"])
    samp = np.stack([fg[k][::st, ::st] for k in ("rho", "u", "v")])
    d = {k: rel_l2(samp[i], fx["sample_lu"][i]) for i, k in enumerate(("rho", "u", "v"))}
    for c, name in enumerate(("C_D", "C_L", "C_M")):
        d[name] = rel_l2(hg[:, c], fx["hist_lu"][:, c])
    print("  deviation from the LU oracle after %d steps:" % steps, {k: "%.3g" % v for k, v in d.items()},
          flush=True)
    last = fx["drift"][-1]  # step, rho, u, v, C_x, C_y, C_M between the two oracle solvers
    for i, k in enumerate(("C_D", "C_L", "C_M")):
        assert d[k] <= 10.0 * last[4 + i] + 1e-12, k
