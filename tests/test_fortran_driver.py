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


def write_inputs(d, case, chunk=7, report_every=10, nml="lbmib.nml", fields="fields.bin", **extra):
    p = case.params
    cases.write_boundary(os.path.join(d, "boundary.dat"), case.xb0, case.yb0)
    cases.write_motion(os.path.join(d, "motion_params.dat"), case.motion)
    s = ",".join("%.17g" % x for x in p.s)
    with open(os.path.join(d, nml), "w") as fh:
        fh.write("&lbmib_case\n nx=%d, ny=%d, ox=%r, oy=%r, scale=%r, UU=%r,\n" % (p.nx, p.ny, p.ox, p.oy, p.scale, p.U))
        fh.write(" s=%s, smooth=%d, wing_count=%d, points_per_wing=%d,\n" % (s, p.smooth, p.n_wings, p.pts_per_wing))
        for k, v in extra.items():
            fh.write(" %s=%s,\n" % (k, ("'%s'" % v) if isinstance(v, str) else v))
        fh.write(" chunk=%d, report_every=%d, fields_file='%s'\n/\n" % (chunk, report_every, fields))


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


def run_driver(exe, d, nml="lbmib.nml"):
    r = subprocess.run([exe, nml], cwd=d, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.gpu
def test_driver_restart_file_continues_the_run(tmp_path):
    """saveflowfield at step 20 of a 30-step run, then a second run that
    resumes from that file (readflowfield): its force rows 21..30 and final
    fields equal the uninterrupted run's, bit for bit."""
    exe = driver()
    case = cases.config(1, n_steps=30, scale_down=8)
    d = str(tmp_path)
    p = case.params
    write_inputs(d, case, nml="a.nml", restart_every=20, force_file="force_a.dat", fields="a.bin")
    print("fortran driver: 30 steps with a restart file at step 20", flush=True)
    run_driver(exe, d, "a.nml")
    raw = open(os.path.join(d, "saved_flow_feild.dat"), "rb").read()
    nw = p.n_wings
    assert len(raw) == 4 * 3 + 8 * 11 * nw + 8 * 11 * p.nx * p.ny + 8 * 36 * (p.ny - 2 + p.nx)
    assert np.frombuffer(raw[:8], dtype=np.int32).tolist() == [20, nw]
    write_inputs(d, case, nml="b.nml", restart_in="saved_flow_feild.dat", restart_out="unused.dat",
                 force_file="force_b.dat", fields="b.bin")
    out = run_driver(exe, d, "b.nml")
    assert "resumed after step 20" in out
    fa = np.fromfile(os.path.join(d, "a.bin"))
    fb = np.fromfile(os.path.join(d, "b.bin"))
    assert np.array_equal(fa, fb)
    ra = open(os.path.join(d, "force_a.dat")).read().splitlines()
    rb = open(os.path.join(d, "force_b.dat")).read().splitlines()
    assert len(ra) == 30 and rb == ra[20:]
    # saveflowfield's copy beside the primary (solution.f90:1677-1706); a
    # primary cut short gives way to it (solution.f90:1737-1802)
    assert open(os.path.join(d, "saved_flow_feild_copy.dat"), "rb").read() == raw
    with open(os.path.join(d, "saved_flow_feild.dat"), "wb") as fh:
        fh.write(raw[:len(raw) // 2])
    write_inputs(d, case, nml="c.nml", restart_in="saved_flow_feild.dat", restart_out="unused.dat",
                 force_file="force_c.dat", fields="c.bin")
    out = run_driver(exe, d, "c.nml")
    assert "reading its copy" in out and "resumed after step 20" in out
    assert np.array_equal(np.fromfile(os.path.join(d, "c.bin")), fa)


@pytest.mark.gpu
def test_driver_tecplot_flow_field_file(tmp_path, oracle_mod):
    """F<n>.dat: the point zone holds X, Y, U/UU, V/UU, rho/3, the vorticity,
    the IB-node mask and a zero vorticity rate on every second node; the
    triangle zone holds the moved polygon (whose segment midpoints are the
    oracle's markers) and the wing centres."""
    exe = driver()
    case = cases.config(0, n_steps=24, scale_down=2)
    d = str(tmp_path)
    write_inputs(d, case, output_every=24)
    run_driver(exe, d)
    xr, yr = cases.read_boundary(os.path.join(d, "boundary.dat"), len(case.xb0))
    case.xb0, case.yb0 = xr, yr
    o = oracle_mod.Oracle(case)
    o.step(24)
    fl = o.fields()
    p = case.params
    lines = open(os.path.join(d, "F24.dat")).read().splitlines()
    assert lines[0].startswith('VARIABLES = "X" "Y" "U" "V" "Presure" "Vorticity" "body" "dvordt"')
    assert 'ZONE T="Flow feild"' in lines[1] and "SOLUTIONTIME=" in lines[1]
    onx, ony = (p.nx - 1) // 2 + 1, (p.ny - 1) // 2 + 1
    data = np.array([[float(x) for x in ln.split()] for ln in lines[2:2 + onx * ony]]).reshape(ony, onx, 8)
    ii, jj = np.arange(0, p.nx, 2), np.arange(0, p.ny, 2)
    u, v, rho = fl["u"], fl["v"], fl["rho"]
    close = lambda a, b: np.allclose(a, b, rtol=1e-7, atol=1e-12)
    assert close(data[:, :, 0], (p.ox + ii * p.scale)[None, :].repeat(ony, 0))
    assert close(data[:, :, 1], (p.oy + jj * p.scale)[:, None].repeat(onx, 1))
    assert close(data[:, :, 2], u[::2, ::2] / p.U)
    assert close(data[:, :, 3], v[::2, ::2] / p.U)
    assert close(data[:, :, 4], rho[::2, ::2] / 3.0)
    # vorticity: central differences inside, one-sided on the edge rows/columns
    dvx = np.empty_like(v)
    dvx[:, 1:-1] = (v[:, 2:] - v[:, :-2]) / (p.U * 2.0 * p.scale)
    dvx[:, 0] = (-1.5 * v[:, 0] + 2.0 * v[:, 1] - 0.5 * v[:, 2]) / (p.U * p.scale)
    dvx[:, -1] = (1.5 * v[:, -1] - 2.0 * v[:, -2] + 0.5 * v[:, -3]) / (p.U * p.scale)
    duy = np.empty_like(u)
    duy[1:-1, :] = (u[2:, :] - u[:-2, :]) / (p.U * 2.0 * p.scale)
    duy[0, :] = (-1.5 * u[0, :] + 2.0 * u[1, :] - 0.5 * u[2, :]) / (p.U * p.scale)
    duy[-1, :] = (1.5 * u[-1, :] - 2.0 * u[-2, :] + 0.5 * u[-3, :]) / (p.U * p.scale)
    assert np.allclose(data[:, :, 5], (dvx - duy)[::2, ::2], rtol=1e-7, atol=1e-9)
    body = np.zeros((p.ny, p.nx))
    ib = o.ib_nodes()
    body[ib["j"], ib["i"]] = 1.0
    assert np.array_equal(data[:, :, 6], body[::2, ::2]) and body[::2, ::2].sum() > 0
    assert np.all(data[:, :, 7] == 0.0)
    # the airfoil zone
    rest = lines[2 + onx * ony:]
    assert 'ZONE T="airfoil"' in rest[0] and "ZONETYPE=FETRIANGLE" in rest[0]
    assert "NODES=%d" % (p.n_wings * (p.pts_per_wing + 1)) in rest[0]
    npts = p.n_wings * p.pts_per_wing
    pts = np.array([[float(x) for x in ln.split()] for ln in rest[1:1 + npts + p.n_wings]])
    assert np.all(pts[:, 2] == 255.0)
    xb1, yb1 = pts[:npts, 0], pts[:npts, 1]
    mk = o.markers()
    assert np.allclose(0.5 * (xb1[1:] + xb1[:-1]), mk["xb2"], atol=2e-8)
    assert np.allclose(0.5 * (yb1[1:] + yb1[:-1]), mk["yb2"], atol=2e-8)
    tri = [ln.split() for ln in rest[1 + npts + p.n_wings:]]
    assert len(tri) == p.n_wings * (p.pts_per_wing - 1)
    assert tri[0] == ["001", "002", "%03d" % (npts + 1)]


@pytest.mark.gpu
def test_driver_vorticity_rate_column(tmp_path, oracle_mod):
    """vorticity_rate = 1 calls vor_mom_rate after every step (the call the
    reference leaves commented out, main.f90:391); F<n>.dat's dvordt column
    then holds the rate of the oracle's history updated at every step."""
    exe = driver()
    case = cases.config(0, n_steps=12, scale_down=2)
    d = str(tmp_path)
    write_inputs(d, case, output_every=12, vorticity_rate=1, output_stride=1)
    run_driver(exe, d)
    xr, yr = cases.read_boundary(os.path.join(d, "boundary.dat"), len(case.xb0))
    case.xb0, case.yb0 = xr, yr
    p = case.params
    o = oracle_mod.Oracle(case)
    st = oracle_mod.VorticityRate(p.nx, p.ny, p.U, p.scale)
    for _ in range(12):
        o.step(1)
        fl = o.fields()
        rate = st.update(fl["u"], fl["v"])
    vor = oracle_mod.vorticity(fl["u"], fl["v"], p.U, p.scale)
    lines = open(os.path.join(d, "F12.dat")).read().splitlines()
    data = np.array([[float(x) for x in ln.split()] for ln in lines[2:2 + p.nx * p.ny]]).reshape(p.ny, p.nx, 8)
    assert np.allclose(data[:, :, 5], vor, rtol=1e-7, atol=1e-9)
    assert np.allclose(data[:, :, 7], rate, rtol=1e-7, atol=1e-9)
    assert np.abs(rate).max() > 0.0 and np.all(data[0, :, 7] == 0.0)


@pytest.mark.gpu
def test_driver_dynamics_and_restart_after_a_domain_shift(tmp_path, oracle_mod):
    """solve_dynamic_equations = 1: force.dat and airfoil_motion.dat rows
    against the oracle's; a restart file written after the domain shift
    (body state, movenum, refilled edge f0) resumes bit-identically."""
    exe = driver()
    case = cases.config(0, n_steps=520, scale_down=2)
    d = str(tmp_path)
    dyn = dict(solve_dynamic_equations=1, solid_den=200.0)
    write_inputs(d, case, nml="a.nml", restart_every=500, force_file="force_a.dat", fields="a.bin",
                 motion_out_file="motion_a.dat", report_every=0, **dyn)
    print("fortran driver: 520 steps with the dynamics, restart file at step 500", flush=True)
    run_driver(exe, d, "a.nml")
    raw = open(os.path.join(d, "saved_flow_feild.dat"), "rb").read()
    p = case.params
    assert np.frombuffer(raw[8 + 8 * 11 * p.n_wings:12 + 8 * 11 * p.n_wings], dtype=np.int32)[0] == 1  # movenum
    write_inputs(d, case, nml="b.nml", restart_in="saved_flow_feild.dat", restart_out="unused.dat",
                 force_file="force_b.dat", fields="b.bin", motion_out_file="motion_b.dat", report_every=0, **dyn)
    out = run_driver(exe, d, "b.nml")
    assert "resumed after step 500" in out
    assert np.array_equal(np.fromfile(os.path.join(d, "a.bin")), np.fromfile(os.path.join(d, "b.bin")))
    ra = open(os.path.join(d, "force_a.dat")).read().splitlines()
    rb = open(os.path.join(d, "force_b.dat")).read().splitlines()
    assert len(ra) == 520 and rb == ra[500:]
    ma = open(os.path.join(d, "motion_a.dat")).read().splitlines()
    mb = open(os.path.join(d, "motion_b.dat")).read().splitlines()
    assert mb == ma[500:]
    xr, yr = cases.read_boundary(os.path.join(d, "boundary.dat"), len(case.xb0))
    case.xb0, case.yb0 = xr, yr
    o = oracle_mod.Oracle(case, n_threads=8)
    o.set_dynamics(cases.Dynamics(solid_den=200.0))
    o.step(520)
    assert np.max(np.abs(np.loadtxt(os.path.join(d, "force_a.dat")) - o.history())) <= 0.5e-8 + 1e-15
    m = np.loadtxt(os.path.join(d, "motion_a.dat"))
    assert m.shape == (520, 4) and np.max(np.abs(m - o.motion_rows())) <= 0.5e-8 + 1e-15
    assert np.abs(m[:, 3]).max() > 0.01
