"""Restart from a saved state (the reference's readflowfield path): a run
stopped after k steps, its populations downloaded and uploaded into a new
handle set to step k, continues bit for bit like the uninterrupted run --
at a geometry-batch start and inside a batch, in FP64 and in the FP32
deviation form (whose download/upload round trip is exact)."""
import numpy as np
import pytest

from fortran_loadfile_mitiwing_amd import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1])
def test_restart_continues_bit_identically(lbmib_mod, prec):
    n = 60
    case = cases.config(1, n_steps=n, scale_down=8)  # 512x256, 128 markers, pitch + plunge
    full = lbmib_mod.LBMIB(case, precision=prec)
    full.evolution(n)
    ref = full.fields()
    ref_h = full.history()
    full.close()
    for k in (lbmib_mod.GEOMETRY_BATCH, 23):
        print("  restart %s at step %d" % (case.name, k), flush=True)
        a = lbmib_mod.LBMIB(case, precision=prec)
        a.evolution(k)
        saved = a.fields()["f"]
        a.close()
        b = lbmib_mod.LBMIB(case, precision=prec)
        b.upload(saved, step=k)
        assert b.steps_done == k
        b.evolution(n - k)
        got = b.fields()
        for key in ("rho", "u", "v", "f"):
            assert np.array_equal(got[key], ref[key]), (k, key)
        h = b.history()
        assert np.array_equal(h[k:], ref_h[k:]), k
        b.close()


def test_set_step_checks(lbmib_mod):
    case = cases.config(0, n_steps=20, scale_down=4)
    g = lbmib_mod.LBMIB(case)
    L = lbmib_mod.load()
    assert L.lbmib_set_step(g.h, -1) == -1
    assert L.lbmib_set_step(g.h, g.max_steps() + 1) == -1
    b = lbmib_mod.Batch()
    b.add(g)
    assert L.lbmib_set_step(g.h, 0) == -6  # a batch member
    b.remove(g)
    assert L.lbmib_set_step(g.h, 4) == 0
    with pytest.raises(RuntimeError):
        g.fields()  # the fields of step 4 were never computed
    b.close()
    g.close()


def test_markers_are_the_polygon_whose_midpoints_the_oracle_moves(lbmib_mod, oracle_mod):
    case = cases.config(1, n_steps=20, scale_down=16)
    g = lbmib_mod.LBMIB(case)
    g.evolution(20)
    o = oracle_mod.Oracle(case)
    o.step(20)
    x, y = g.markers()
    mk = o.markers()
    assert np.array_equal(0.5 * (x[1:] + x[:-1]), mk["xb2"])
    assert np.array_equal(0.5 * (y[1:] + y[:-1]), mk["yb2"])
    g.close()
