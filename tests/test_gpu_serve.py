"""Serving loop on the GPU: requests admitted while others run (continuous
batching) give the same result as the same case run alone.

The per-case arithmetic of a batch is the single-case path's, so the stated
tolerance is exact equality of every force-history row (FP32 deviation form).
"""
import numpy as np
import pytest

from fortran_loadfile_mitiwing_amd import cases, lbmib
from fortran_loadfile_mitiwing_amd.serve import GPUBackend, Server

pytestmark = pytest.mark.gpu


def small_requests():
    """Five small cases of different sizes, kinematics and lengths."""
    return [
        cases.config(0, n_steps=72, scale_down=2),
        cases.config(1, n_steps=45, scale_down=16),
        cases.make_case(160, 80, 16.0, 48, 300.0, 60, heave=0.4, pitch_deg=20.0, phase_deg=60.0,
                        strouhal=0.35, name="mini_160x80"),
        cases.make_case(192, 96, 16.0, 64, 500.0, 33, heave=0.6, pitch_deg=10.0, phase_deg=120.0,
                        strouhal=0.25, name="mini_192x96"),
        cases.config(0, n_steps=40, scale_down=4),
    ]


def test_results_are_the_same_alone_and_in_the_mix(lbmib_mod):
    reqs = small_requests()
    alone = []
    for c in reqs:
        g = lbmib_mod.LBMIB(c, precision=lbmib_mod.PREC_FP32DEV)
        g.evolution(c.n_steps)
        alone.append(g.history())
        g.close()
    srv = Server(GPUBackend(precision=lbmib.PREC_FP32DEV), hbm_budget=8 << 30, window=16)
    # arrivals: two at once, then one per window while the others run
    arrive = {0: [0, 1], 1: [2], 2: [3], 4: [4]}
    w = 0
    while w < 40 and (not srv.idle() or w <= max(arrive)):
        for k in arrive.get(w, []):
            srv.submit(reqs[k], tag=k)
        srv.step_window()
        print("  window %d: running %s" % (w, [j.tag for j in srv.running]), flush=True)
        w += 1
    res = {r.tag: r for r in srv.results}
    assert sorted(res) == list(range(len(reqs)))
    st = srv.stats()
    assert st["peak_running_cases"] >= 3
    for k, c in enumerate(reqs):
        assert res[k].steps == c.n_steps
        assert np.array_equal(res[k].history, alone[k]), c.name
    srv.close()


def test_admission_waits_for_memory(lbmib_mod):
    reqs = small_requests()[:3]
    be = GPUBackend(precision=lbmib.PREC_FP32DEV)
    budget = max(be.estimate_bytes(c) for c in reqs) + 1  # one case at a time
    srv = Server(be, hbm_budget=budget, window=16)
    for k, c in enumerate(reqs):
        srv.submit(c, tag=k)
    peak = 0
    while not srv.idle():
        srv.step_window()
        peak = max(peak, len(srv.running))
    assert peak == 1
    assert [r.tag for r in srv.results] == [0, 1, 2]
    srv.close()
