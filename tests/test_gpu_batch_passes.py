"""Two-step passes in a batch: every member whose lattice window clears the
outer ring by 4 nodes takes its far field two steps per pass, each launch
(far pass, window, ring, IB chain) covering all members.  Every node keeps
the single-step arithmetic, so the bar is bit-for-bit equality with the same
batch stepped one step at a time (fuse_steps = 0) and with each case alone:
members of different sizes and precisions, calls of odd length, a member
whose motion table ends inside a call (passes stop there and the rest are
single steps), a case that joins with its state in the other plane, and
members that leave.
"""
import numpy as np
import pytest

from fortran_loadfile_mitiwing_amd import cases

pytestmark = pytest.mark.gpu


def members(n_steps):
    """Three cases of different sizes whose windows clear the ring."""
    return [
        cases.config(1, n_steps=n_steps, scale_down=4),   # 1024x512, 256 markers, plunge + pitch
        cases.config(0, n_steps=n_steps),                 # 400x200, 128 markers, plunge
        cases.make_case(512, 256, 32.0, 128, 500.0, n_steps, heave=0.5, pitch_deg=25.0, phase_deg=90.0,
                        strouhal=0.3, name="mid_512x256"),
    ]


def same(a, b, what=""):
    fa, fb = a.fields(), b.fields()
    for k in ("rho", "u", "v", "f"):
        assert np.array_equal(fa[k], fb[k]), (what, k)
    assert np.array_equal(a.history(), b.history()), what


def batch_of(lbmib_mod, cs, prec, fuse):
    sims = [lbmib_mod.LBMIB(c, precision=prec) for c in cs]
    b = lbmib_mod.Batch()
    b.set_shape(fuse_steps=fuse)
    for s in sims:
        b.add(s)
    return b, sims


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_passes_equal_single_steps(lbmib_mod, prec):
    cs = members(80)
    a, sa = batch_of(lbmib_mod, cs, prec, 1)
    b, sb = batch_of(lbmib_mod, cs, prec, 0)
    a.set_profiling(1)
    a.step(21)
    t = a.timings()
    print("  timings:", {k: round(v, 4) if isinstance(v, float) else v for k, v in t.items()}, flush=True)
    assert t["far_passes"] == 10 and t["lattice_steps"] == 1
    b.step(21)
    for chunk in (16, 5, 30):
        a.step(chunk)
        b.step(chunk)
    for c, x, y in zip(cs, sa, sb):
        same(x, y, c.name)
    for c, x in zip(cs, sa):
        g = lbmib_mod.LBMIB(c, precision=prec)
        g.evolution(72)
        same(x, g, c.name + " alone")
        g.close()


def test_member_ending_inside_a_call(lbmib_mod):
    cs = members(64)[:2] + [cases.make_case(512, 256, 32.0, 128, 500.0, 23, heave=0.4, name="short_512")]
    a, sa = batch_of(lbmib_mod, cs, 1, 1)
    a.set_profiling(1)
    a.step(32)
    t = a.timings()
    # passes while all three step (offsets 0..21), then single steps
    assert t["far_passes"] == 11 and t["lattice_steps"] == 10
    assert [s.steps_done for s in sa] == [32, 32, 23]
    a.step(16)
    for c, x in zip(cs, sa):
        g = lbmib_mod.LBMIB(c, precision=1)
        g.evolution(x.steps_done)
        same(x, g, c.name)
        g.close()


def test_join_with_state_in_the_other_plane_and_leave(lbmib_mod):
    """B runs 17 steps alone (passes leave its state in the other plane), then
    joins A's batch at step 16 + 16; A leaves and steps on alone."""
    ca, cb = members(96)[:2]
    prec = 0
    a = lbmib_mod.LBMIB(ca, precision=prec)
    bb = lbmib_mod.LBMIB(cb, precision=prec)
    bb.evolution(15)
    bb.evolution(1)
    b = lbmib_mod.Batch()
    b.add(a)
    b.step(16)
    b.add(bb)
    b.step(13)
    b.step(19)
    b.remove(a)
    a.evolution(9)
    b.step(10)
    b.remove(bb)
    assert a.steps_done == 57 and bb.steps_done == 58
    for c, x in ((ca, a), (cb, bb)):
        g = lbmib_mod.LBMIB(c, precision=prec)
        g.evolution(x.steps_done)
        same(x, g, c.name)
        g.close()


def test_batch_passes_match_the_oracle(oracle_mod, lbmib_mod):
    cs = members(34)[1:]
    a, sa = batch_of(lbmib_mod, cs, 0, 1)
    a.step(17)
    a.step(17)
    for c, x in zip(cs, sa):
        o = oracle_mod.Oracle(c, n_threads=8)
        o.step(34)
        fo, fg = o.fields(), x.fields()
        for k in ("rho", "u", "v", "f"):
            assert np.array_equal(fg[k], fo[k]), (c.name, k)
        assert np.array_equal(x.history(), o.history()), c.name
