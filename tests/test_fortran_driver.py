"""The Fortran side of the drop-in boundary: the iso_c_binding module and the
driver that keeps the reference's main loop, input files (boundary.dat,
motion_params.dat) and outputs (force.dat, '(1000f15.8)' rows)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, gpu_available
from fortran_loadfile_mitiwing_amd import cases

DRIVER = os.path.join(ROOT, "fortran_loadfile_mitiwing_amd", "lbmib_driver")


def write_inputs(d, case, chunk=7, report_every=10):
    p = case.params
    cases.write_boundary(os.path.join(d, "boundary.dat"), case.xb0, case.yb0)
    cases.write_motion(os.path.join(d, "motion_params.dat"), case.motion)
    s = ",".join("%.17g" % x for x in p.s)
    with open(os.path.join(d, "lbmib.nml"), "w") as fh:
        fh.write("&lbmib_case\n nx=%d, ny=%d, ox=%r, oy=%r, scale=%r, UU=%r,\n" % (p.nx, p.ny, p.ox, p.oy, p.scale, p.U))
        fh.write(" s=%s, smooth=%d, wing_count=%d, points_per_wing=%d,\n" % (s, p.smooth, p.n_wings, p.pts_per_wing))
        fh.write(" chunk=%d, report_every=%d, fields_file='fields.bin'\n/\n" % (chunk, report_every))


def driver():
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (no amdflang)")
    return DRIVER


@pytest.mark.skipif(gpu_available(), reason="checks the behaviour without a GPU")
def test_driver_stops_with_the_library_error_without_a_device(tmp_path):
    exe = driver()
    write_inputs(str(tmp_path), cases.config(0, n_steps=10, scale_down=4))
    r = subprocess.run([exe, "lbmib.nml"], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "lbmib_create failed (-2)" in r.stdout


def test_driver_rejects_a_missing_boundary_file(tmp_path):
    exe = driver()
    write_inputs(str(tmp_path), cases.config(0, n_steps=10, scale_down=4))
    os.remove(os.path.join(str(tmp_path), "boundary.dat"))
    r = subprocess.run([exe, "lbmib.nml"], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "cannot open boundary file" in r.stdout


@pytest.mark.gpu
def test_driver_force_rows_and_fields_match_the_oracle(tmp_path, oracle_mod):
    exe = driver()
    case = cases.config(0, n_steps=30, scale_down=2)
    d = str(tmp_path)
    write_inputs(d, case)
    print("fortran driver: cfg0/2, 30 steps", flush=True)
    r = subprocess.run([exe, "lbmib.nml"], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "step 30 err" in r.stdout
    # the inputs went through the files exactly as the reference reads them
    xr, yr = cases.read_boundary(os.path.join(d, "boundary.dat"), len(case.xb0))
    case.xb0, case.yb0 = xr, yr
    o = oracle_mod.Oracle(case)
    o.step(30)
    h = o.history()
    rows = np.loadtxt(os.path.join(d, "force.dat"))
    assert rows.shape == h.shape
    assert np.max(np.abs(rows - h)) <= 0.5e-8 + 1e-15
    p = case.params
    raw = np.fromfile(os.path.join(d, "fields.bin"), dtype=np.float64).reshape(3, p.ny, p.nx)
    fl = o.fields()
    for k, name in enumerate(("rho", "u", "v")):
        assert np.array_equal(raw[k], fl[name]), name
    motion = np.loadtxt(os.path.join(d, "airfoil_motion.dat"))
    assert motion.shape == (30, 4) and np.all(motion[:, 1:] == 0.0)
