"""Geometry left to the next call.  A geometry-batch boundary that falls
fewer than 12 steps before the end of an engine call does not build the
next batch's geometry (that call's steps never use it); the next call builds
it at its start.  The results must not depend on where calls end: a case
stepped in call lengths that put boundaries near the ends of calls (so
builds are left over and picked up at the next call's start, one of them by
a call that starts at the boundary of the batch so built) equals the same
case stepped in one call, bit for bit, alone and in a batch, in both
precisions.
"""
import numpy as np
import pytest

from fortran_loadfile_mitiwing_amd import cases

pytestmark = pytest.mark.gpu

# with 16-step batches: the boundary at step 17 falls 4 steps before the end
# of the first call (left over), the second call starts inside the batch and
# builds it; step 33's boundary falls in the 13-step call 2 steps before its
# end; the 1-step and 3-step calls carry leftover builds across
CHUNKS = [20, 12, 1, 3, 13, 2, 17]


def same(a, b):
    fa, fb = a.fields(), b.fields()
    for k in ("rho", "u", "v", "f"):
        assert np.array_equal(fa[k], fb[k]), k
    assert np.array_equal(a.history(), b.history())


@pytest.mark.parametrize("prec", [0, 1])
def test_leftover_geometry_alone(lbmib_mod, prec):
    case = cases.config(1, n_steps=80, scale_down=4)  # 1024x512, 256 markers
    print("  %s chunks %s" % (case.name, CHUNKS), flush=True)
    a = lbmib_mod.LBMIB(case, precision=prec)
    for c in CHUNKS:
        a.evolution(c)
    b = lbmib_mod.LBMIB(case, precision=prec)
    b.evolution(sum(CHUNKS))
    same(a, b)
    a.close()
    b.close()


def test_leftover_geometry_in_a_batch(lbmib_mod):
    cs = [cases.config(0, n_steps=80, scale_down=2), cases.config(1, n_steps=80, scale_down=16)]
    prec = 1
    sims = [lbmib_mod.LBMIB(c, precision=prec) for c in cs]
    bt = lbmib_mod.Batch()
    for s in sims:
        bt.add(s)
    for c in CHUNKS:
        bt.step(c)
    for c, s in zip(cs, sims):
        ref = lbmib_mod.LBMIB(c, precision=prec)
        ref.evolution(sum(CHUNKS))
        same(s, ref)
        ref.close()


def test_leftover_geometry_then_batch(lbmib_mod):
    # a case leaves a build over when stepped alone, then joins a batch whose
    # first call must build it
    ca, cb = cases.config(1, n_steps=80, scale_down=8), cases.config(0, n_steps=80, scale_down=2)
    prec = 0
    a = lbmib_mod.LBMIB(ca, precision=prec)
    a.evolution(20)  # boundary at 17 left over
    b = lbmib_mod.LBMIB(cb, precision=prec)
    b.evolution(4)
    b.evolution(16)  # boundary at 17 left over as well
    bt = lbmib_mod.Batch()
    bt.add(a)
    bt.add(b)
    bt.step(30)
    bt.remove(a)
    bt.remove(b)
    for s, c in ((a, ca), (b, cb)):
        ref = lbmib_mod.LBMIB(c, precision=prec)
        ref.evolution(50)
        same(s, ref)
        ref.close()
