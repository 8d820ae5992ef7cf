"""Two-step passes of an engine of one case (launch shape fuse_steps = 1, the
default): the nodes outside the body's lattice window take steps n and n+1
in one pass over the far field, the window one step at a time with the IB
chain between.  Every node keeps the single-step arithmetic, so the bar is
bit-for-bit equality with the same case stepped one step at a time
(fuse_steps = 0) -- populations, rho/u/v, force history, residual -- for
call lengths that start passes at every position (passes straddle the
geometry-batch boundaries; an odd chunk leaves one single step), in both
precisions, across a restart, and for a case that runs
passes alone and then joins a batch (its state then lies in the other
plane).  Parity with the oracle is covered by the single-case parity tests,
which run with the default shape.
"""
import numpy as np
import pytest

from fortran_loadfile_mitiwing_amd import cases

pytestmark = pytest.mark.gpu

CHUNKS = [5, 16, 3, 1, 27, 12]


def same(a, b):
    fa, fb = a.fields(), b.fields()
    for k in ("rho", "u", "v", "f"):
        assert np.array_equal(fa[k], fb[k]), k
    assert np.array_equal(a.history(), b.history())


def run(lbmib_mod, case, prec, fuse, chunks):
    g = lbmib_mod.LBMIB(case, precision=prec)
    g.set_shape(fuse_steps=fuse)
    for c in chunks:
        g.evolution(c)
    return g


@pytest.mark.parametrize("prec", [0, 1])
def test_passes_equal_single_steps(lbmib_mod, prec):
    case = cases.config(1, n_steps=80, scale_down=4)  # 1024x512, 256 markers, plunge + pitch
    print("  %s fuse on / off, chunks %s" % (case.name, CHUNKS), flush=True)
    a = run(lbmib_mod, case, prec, 1, CHUNKS)
    b = run(lbmib_mod, case, prec, 0, CHUNKS)
    same(a, b)
    assert a.erranalysis() == b.erranalysis()
    a.close()
    b.close()


def test_passes_are_taken(lbmib_mod):
    case = cases.config(1, n_steps=40, scale_down=4)
    g = lbmib_mod.LBMIB(case)
    g.set_profiling(1)
    g.evolution(32)
    t = g.timings()
    print("  timings:", {k: round(v, 4) if isinstance(v, float) else v for k, v in t.items()}, flush=True)
    assert t["far_passes"] == 16 and t["lattice_steps"] == 0
    g.set_shape(fuse_steps=0)
    g.set_profiling(1)
    g.evolution(8)
    t = g.timings()
    assert t["far_passes"] == 0 and t["lattice_steps"] == 8


def test_passes_start_at_any_position(lbmib_mod):
    # a window that starts at an odd position inside the geometry batch takes
    # passes from its first step; the pass whose second step opens the next
    # batch straddles the boundary
    case = cases.config(1, n_steps=60, scale_down=4)
    g = lbmib_mod.LBMIB(case)
    g.evolution(5)
    g.set_profiling(1)
    g.evolution(20)  # steps 6..25: the pass (16, 17) straddles a boundary
    t = g.timings()
    assert t["far_passes"] == 10 and t["lattice_steps"] == 0
    g.set_profiling(1)
    g.evolution(15)  # steps 26..40: seven passes and one single step
    t = g.timings()
    assert t["far_passes"] == 7 and t["lattice_steps"] == 1
    ref = run(lbmib_mod, case, 0, 0, [5, 20, 15])
    same(g, ref)
    g.close()
    ref.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_odd_lattice_widths(lbmib_mod, prec):
    # odd nx: the FP32 pass's last strip has an odd width and its last pair
    # holds a node past the lattice; strips of different widths
    for nx, ny in ((417, 203), (1031, 517)):
        case = cases.make_case(nx, ny, 28.0, 112, 400.0, 40, heave=0.45, pitch_deg=20.0, phase_deg=80.0,
                               strouhal=0.3, name="odd_%dx%d" % (nx, ny))
        print("  %s fuse on / off" % case.name, flush=True)
        a = run(lbmib_mod, case, prec, 1, [7, 16, 9])
        b = run(lbmib_mod, case, prec, 0, [7, 16, 9])
        a.set_profiling(1)
        a.evolution(2)
        assert a.timings()["far_passes"] == 1
        b.evolution(2)
        same(a, b)
        a.close()
        b.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_persistent_far_grid(lbmib_mod, prec):
    # persistent=1: one workgroup per CU (256 on MI355X) loops over the far
    # tiles (about 510 FP64 and 306 FP32 tiles at 1024x512, several thousand
    # at 4096x2048), against one workgroup per tile
    for case, chunks in ((cases.config(1, n_steps=60, scale_down=4), [5, 16, 11]),
                         (cases.config(1, n_steps=20), [3, 10])):
        if prec == 1 and case.params.nx > 1024:
            continue
        print("  %s persistent 1 / 0, chunks %s" % (case.name, chunks), flush=True)
        a = lbmib_mod.LBMIB(case, precision=prec)
        a.set_shape(persistent=1)
        a.set_profiling(1)
        for c in chunks:
            a.evolution(c)
        assert a.timings()["far_passes"] > 0
        b = run(lbmib_mod, case, prec, 1, chunks)
        same(a, b)
        a.close()
        b.close()


def test_full_size_first_steps_equal_single_steps(lbmib_mod):
    case = cases.config(1, n_steps=40)  # 4096x2048, 1024 markers
    a = run(lbmib_mod, case, 0, 1, [7, 18])
    b = run(lbmib_mod, case, 0, 0, [7, 18])
    same(a, b)


def test_restart_after_passes(lbmib_mod):
    case = cases.config(0, n_steps=60, scale_down=2)
    whole = run(lbmib_mod, case, 0, 1, [21, 30])
    first = run(lbmib_mod, case, 0, 1, [21])
    f = first.fields()["f"]
    resumed = lbmib_mod.LBMIB(case)
    resumed.upload(f, step=21)
    resumed.evolution(30)
    fw, fr = whole.fields(), resumed.fields()
    for k in ("rho", "u", "v", "f"):
        assert np.array_equal(fw[k], fr[k]), k


def test_case_after_passes_joins_a_batch(lbmib_mod):
    """Two cases take 14 steps alone (7 passes each, so their states lie in
    the other plane), then 18 together in a batch (one step at a time); the
    first then takes 9 more alone.  Every result equals the same case stepped
    alone one step at a time."""
    ca = cases.config(1, n_steps=64, scale_down=8)
    cb = cases.config(0, n_steps=64, scale_down=2)
    prec = 1
    a = lbmib_mod.LBMIB(ca, precision=prec)
    a.evolution(14)
    bb = lbmib_mod.LBMIB(cb, precision=prec)
    bb.evolution(14)
    b = lbmib_mod.Batch()
    b.add(a)
    b.add(bb)
    b.step(18)
    b.remove(a)
    b.remove(bb)
    a.evolution(9)
    same(a, run(lbmib_mod, ca, prec, 0, [41]))
    same(bb, run(lbmib_mod, cb, prec, 0, [32]))
