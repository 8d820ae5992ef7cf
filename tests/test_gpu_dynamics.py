"""GPU rigid-body dynamics against the oracle: wing 1 driven by the IB force
(solve_dynamic_equations; boundary_condition.f90:1676-1719, rk4 3168-3202)
and the domain shift that follows it (main.f90:370-388, move_mesh.f90).

The FP64 path evaluates the same operations in the same order as the
oracle, so the bar is bit-for-bit equality of fields, force rows, IB nodes,
body state and airfoil_motion rows, across a domain shift.
"""
import math

import numpy as np
import pytest

from fortran_loadfile_mitiwing_amd import cases

pytestmark = pytest.mark.gpu
DEN = 200.0  # the 200x100 plunge case drifts one node by step 497 with this body density


def case0(n):
    return cases.config(0, n_steps=n, scale_down=2)


def assert_same(o, g, what):
    fo, fg = o.fields(), g.fields()
    for k in ("rho", "u", "v", "f"):
        assert np.array_equal(fg[k], fo[k]), "%s: %s differs (max abs %g)" % (what, k, np.max(np.abs(fg[k] - fo[k])))
    assert np.array_equal(g.history(), o.history()), what
    ib_o, ib_g = o.ib_nodes(), g.ib_nodes()
    for k in ("i", "j", "wing", "glx", "gly"):
        assert np.array_equal(ib_o[k], ib_g[k]), "%s: IB %s" % (what, k)


def test_dynamics_across_a_domain_shift(oracle_mod, lbmib_mod):
    case = case0(520)
    dyn = cases.Dynamics(solid_den=DEN)
    o = oracle_mod.Oracle(case, n_threads=8)
    o.set_dynamics(dyn)
    g = lbmib_mod.LBMIB(case, dynamics=dyn)
    for k in (100, 396, 1, 23):  # the shift comes after step 497
        print("  dynamics: %d steps" % k, flush=True)
        o.step(k)
        g.evolution(k)
        assert_same(o, g, "step %d" % o.steps_done)
        assert np.array_equal(g.body_state(), o.body_state()), "body state at step %d" % o.steps_done
        if o.steps_done == 497:
            # right after the shift the residual compares the shifted velocity with itself
            assert g.body_state()[-1] == 1.0 and g.erranalysis() == 0.0
    assert g.body_state()[-1] == 1.0  # one shift taken
    assert np.array_equal(g.motion_rows(), o.motion_rows())
    mk_o = o.markers()
    xb1, _ = g.markers()
    ppw = case.params.pts_per_wing
    assert np.allclose(0.5 * (xb1[1:ppw] + xb1[:ppw - 1]), mk_o["xb2"], rtol=0, atol=1e-12)


def test_explicit_update_matches_the_oracle(oracle_mod, lbmib_mod):
    case = case0(60)
    dyn = cases.Dynamics(solid_den=DEN, rk4=False)
    o = oracle_mod.Oracle(case, n_threads=8)
    o.set_dynamics(dyn)
    g = lbmib_mod.LBMIB(case, dynamics=dyn)
    o.step(60)
    g.evolution(60)
    assert_same(o, g, "explicit")
    assert np.array_equal(g.body_state(), o.body_state())


def test_two_wings_move_together(oracle_mod, lbmib_mod):
    """wing_count = 2: the total x force of both wings drives the body, and
    update_body_extra_from_rel gives wing 2 the running sum of the offsets
    (mode.f90:285-307), so both wings move with wing 1."""
    import os
    from conftest import GOLDEN
    xy = cases.read_boundary(os.path.join(GOLDEN, "boundary.dat"), 205)
    case = cases.tandem_reference_case(xy, n_steps=60)
    dyn = cases.Dynamics(solid_den=DEN)
    o = oracle_mod.Oracle(case, n_threads=8)
    o.set_dynamics(dyn)
    g = lbmib_mod.LBMIB(case, dynamics=dyn)
    o.step(60)
    g.evolution(60)
    assert_same(o, g, "two wings")
    st = g.body_state()
    assert np.array_equal(st, o.body_state())
    m = g.motion_rows()
    assert np.array_equal(m, o.motion_rows())
    assert np.array_equal(m[:, 2], m[:, 5]) and np.array_equal(m[:, 3], m[:, 6]) and np.abs(m[:, 3]).max() > 0.0


def test_infinite_mass_is_the_prescribed_run(lbmib_mod):
    case = case0(40)
    a = lbmib_mod.LBMIB(case)
    b = lbmib_mod.LBMIB(case, dynamics=cases.Dynamics(solid_den=math.inf))
    a.evolution(40)
    b.evolution(40)
    fa, fb = a.fields(), b.fields()
    for k in ("rho", "u", "v", "f"):
        assert np.array_equal(fa[k], fb[k]), k
    assert np.array_equal(a.history(), b.history())
    assert np.all(b.body_state() == 0.0)
    assert np.all(a.motion_rows()[:, 1:] == 0.0) and np.array_equal(a.motion_rows(), b.motion_rows())


def test_resume_with_the_body_state(lbmib_mod):
    """A run resumed at step 300 from the populations and the body state (the
    restart file's contents) continues bit-identically through the shift."""
    case = case0(520)
    dyn = cases.Dynamics(solid_den=DEN)
    a = lbmib_mod.LBMIB(case, dynamics=dyn)
    a.evolution(300)
    f = a.fields()["f"]
    state = a.body_state()
    a.evolution(220)
    b = lbmib_mod.LBMIB(case, dynamics=dyn)
    b.upload(f, step=300)
    b.set_body_state(state)
    b.evolution(220)
    fa, fb = a.fields(), b.fields()
    for k in ("rho", "u", "v", "f"):
        assert np.array_equal(fa[k], fb[k]), k
    assert np.array_equal(a.history()[300:], b.history()[300:])
    assert np.array_equal(a.body_state(), b.body_state())
    assert np.array_equal(a.motion_rows()[300:], b.motion_rows()[300:])


def test_fp32_form_follows_the_fp64_body(lbmib_mod):
    case = case0(520)
    dyn = cases.Dynamics(solid_den=DEN)
    g64 = lbmib_mod.LBMIB(case, dynamics=dyn)
    g32 = lbmib_mod.LBMIB(case, precision=lbmib_mod.PREC_FP32DEV, dynamics=dyn)
    g64.evolution(520)
    g32.evolution(520)
    s64, s32 = g64.body_state(), g32.body_state()
    # body velocity and drift within 1% (the FP32 lattice's round-off on a 520-step drift)
    assert abs(s32[2] - s64[2]) <= 0.01 * abs(s64[2])
    assert abs(s32[4] - s64[4]) <= 0.01 * abs(s64[4])
    assert s32[-1] == s64[-1] == 1.0


def test_dynamics_argument_and_state_checks(lbmib_mod):
    L = lbmib_mod.load()
    case = case0(40)
    g = lbmib_mod.LBMIB(case, dynamics=cases.Dynamics(solid_den=DEN))
    b = lbmib_mod.Batch()
    with pytest.raises(lbmib_mod.LBMIBError, match="dynamics"):
        b.add(g)
    d = lbmib_mod.DynamicsParams()
    d.enabled, d.solid_den, d.area_airfoil = 1, DEN, 0.07
    assert L.lbmib_set_dynamics(g.h, d) < 0  # after the motion table
    bad = cases.Dynamics(solid_den=-1.0)
    with pytest.raises(lbmib_mod.LBMIBError, match="solid_den"):
        lbmib_mod.LBMIB(case, dynamics=bad)
    with pytest.raises(lbmib_mod.LBMIBError):
        g.set_body_state(np.zeros(5))
    b.close()
