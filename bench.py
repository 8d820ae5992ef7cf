#!/usr/bin/env python3
"""Benchmark of the per-timestep hot path on one MI355X.

One step = one timestep of the FP64 drop-in path on the 4096x2048 workload
(NACA0012, chord 256 lattice units, 1024 markers, Re 1000, plunge 0.5c with
30 deg pitch leading by 90 deg, Strouhal 0.3, U = 0.05): marker kinematics,
immersed-boundary interpolate/solve/spread/force, fused stream + MRT collide,
outer boundaries.  The lattice starts resident in HBM; the timed region is
exactly K steps between two device synchronisations.

value = lattice updates per second of whole steps (MLUPS, nx*ny per step).
roofline = the dominant kernel.  Steps are taken in two-step passes (the
default launch shape): the far-field kernel k_far2 advances every node
outside the lattice window that holds the body (and off the outer ring,
which takes its steps in kernels of its own) by two updates per launch (the
layer next to the ring by one: its second step goes with the ring's), so
its algorithmic bytes are 144 per node update (9 FP64 populations read + 9
written) times the updates of a launch over its mean launch time, timed with HIP events on
the stream it runs on.  It reads and writes each population once per two
updates, so this algorithmic rate can pass the HBM peak; `traffic` is the
HBM bytes a launch moves (rocprofv3 FETCH_SIZE/WRITE_SIZE) and
`traffic_frac` that rate against the peak.  With single steps
(fuse_steps=0 in --shape) the dominant kernel is the fused stream+collide
kernel k_lattice, 144 bytes per interior node.
cpu_baseline = the CPU oracle (C restatement of the reference step, OpenMP
over the host cores) on the first steps of the same workload.

--form batch: the parameter sweep of 96 seeded cases of uneven lattice sizes
(512..2048 x half, 4*chord markers, random Re / heave / pitch / phase /
Strouhal), all advanced together as one FP32-deviation batch with the
autotuned launch shape, against the same cases run one after another on the
plain FP64 path.  value = cases per hour of the batch; the JSON also holds
the sequential rate, and per-case C_L / C_D errors of the batch against the
FP64 runs.

--form precision: the precision form's workload (8192x4096, chord 512 lu,
2048 markers, Re 2000, the step workload's plunge + pitch), run in the FP32 deviation form
and in FP64 from the same start.  value = FP32 MLUPS over the whole run; the
JSON holds the FP64 rate, their ratio, the relative L2 differences of
rho / u / v after the last step and the C_L / C_D errors of the FP32 run.

--form post: post-processing of a step of the step workload on the device:
output's vorticity with one vor_mom_rate update per call, and cal_sur_cp.
value = nodes per second of the vorticity kernel; roofline over its 64
algorithmic bytes per node (u, v and two history planes read, the vorticity,
two history planes and the rate written).

--form dynamics: the step workload with the rigid-body dynamics of wing 1 on
(the reference's body density and area, rk4): value = MLUPS of whole steps,
each a window of its own (its geometry follows the force of the step before).

--form serve: the serving loop.  Requests (cases drawn as for the sweep with
nx up to 4096 and 2000..20000 steps, log-uniform) arrive as a Poisson process
at the rate that offers 80% of the device's capacity (measured first on a
probe batch of the same distribution); they are admitted by HBM footprint into
one running FP32-deviation batch (continuous batching) and retire when done.
value = aggregate MLUPS; the JSON holds p50/p99 time-to-result and a sample of
results checked against FP64 runs of the same cases.
"""
import argparse
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_NODE_F64 = 144
BYTES_PER_NODE_F32 = 72
WORKLOADS = {
    0: "400x200 lattice, NACA0012 chord 40 lu, 128 markers, Re 100, plunge 0.5c, St 0.3, U 0.05",
    1: "4096x2048 lattice, NACA0012 chord 256 lu, 1024 markers, Re 1000, plunge 0.5c + pitch 30deg "
       "(90deg lead), St 0.3, U 0.05",
    2: "8192x4096 lattice, NACA0012 chord 512 lu, 2048 markers, Re 2000, plunge 0.5c + pitch 30deg "
       "(90deg lead), St 0.3, U 0.05",
}
PROFILE_EVERY = 17  # coprime with the 16-step geometry batch: samples every phase of it
# HBM bytes per k_lattice_f64 launch on the 4096x2048 workload, from
# rocprofv3 FETCH_SIZE (x2.000) + WRITE_SIZE (x1.000), both scaled by a 1 GiB
# device copy measured the same way (the lattice traffic summary in profiles/)
LATTICE_TRAFFIC_BYTES = {(4096, 2048): (311140 * 2 + 590815) * 1024}
# the same for a k_far2<double> launch (two updates of the far field):
# FETCH_SIZE (x2.000) + WRITE_SIZE (x1.000) of a pass that records nothing
FAR_TRAFFIC_BYTES = {(4096, 2048): (436410 * 2 + 613215) * 1024}


def far_updates(nx, ny, ww, wh):
    """Node updates of one far-pass launch: step n at every node outside the
    lattice window and off the outer ring, step n+1 at those nodes but the
    layer next to the ring (which the window + ring kernel advances)."""
    n_far = nx * ny - ww * wh - (2 * nx + 2 * ny - 4)
    return 2 * n_far - (2 * (ny - 2) + 2 * (nx - 4))


def cpu_baseline(case, seconds_budget=20.0, threads=None):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = threads or min(16, os.cpu_count() or 1)
    o = pyoracle.Oracle(case, n_threads=threads)
    cells = case.params.nx * case.params.ny
    t0 = time.time()
    o.step(1)
    t1 = time.time()
    per = t1 - t0
    n = max(1, min(int(seconds_budget / max(per, 1e-3)), 200))
    t2 = time.time()
    o.step(n)
    dt = time.time() - t2
    o.close()
    return {"value": round(cells * n / dt / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": "%d timesteps (after 1 untimed) of the same %dx%d workload from rest, "
                      "oracle/liblbmib_oracle.so, %d OpenMP threads" % (n, case.params.nx, case.params.ny, threads)}


def coefficient_errors(h32, h64, skip_frac=0.5):
    """Errors of an FP32 force history against the FP64 one over the run's
    last (1 - skip_frac).  Mean C_L is taken relative to the C_L r.m.s. (a
    symmetric section's mean C_L is near zero); mean C_D both relative to
    |mean C_D| and relative to the C_D r.m.s. (a flapping section that makes
    thrust has a mean C_D near zero, where the first is ill-conditioned);
    peak |C_L| relative."""
    k = int(len(h64) * skip_frac)
    cd32, cd64 = h32[k:, 1], h64[k:, 1]
    cl32, cl64 = h32[k:, 2], h64[k:, 2]
    rms_l = float(np.sqrt(np.mean(cl64 ** 2)))
    rms_d = float(np.sqrt(np.mean(cd64 ** 2)))
    dcd = abs(cd32.mean() - cd64.mean())
    return {"mean_CD_rel": float(dcd / abs(cd64.mean())),
            "mean_CD": float(dcd / rms_d),
            "mean_CL": float(abs(cl32.mean() - cl64.mean()) / rms_l),
            "peak_CL": float(abs(np.abs(cl32).max() - np.abs(cl64).max()) / np.abs(cl64).max()),
            "_mean_CD64": float(cd64.mean()), "_rms_CD64": rms_d}


def parse_shape(text):
    return {k: int(v) for k, v in (kv.split("=") for kv in text.split(","))} if text else {}


def run_batch(cs, prec, steps, tune, chunk=1000, label="", shape=None):
    """Advance all cases together (launch shape: autotuned, or the given
    one, or the library's); returns (seconds, histories, timings, shape)."""
    from fortran_loadfile_mitiwing_amd import lbmib
    sims = [lbmib.LBMIB(c, precision=prec) for c in cs]
    b = lbmib.Batch()
    for s_ in sims:
        b.add(s_)
    if shape:
        b.set_shape(**shape)
        tune = False
    shape = b.autotune(trial_steps=32)["best"] if tune else b.shape()
    print("[batch] %s shape %s" % (label, shape), file=sys.stderr, flush=True)
    b.set_profiling(PROFILE_EVERY)
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        b.step(n)
        done += n
        print("[batch] %s %d steps, %.2f s" % (label, done, time.perf_counter() - t0), file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    tm = b.timings()
    hist = [s_.history() for s_ in sims]
    # node updates of one far-pass launch over all cases (far_updates)
    tm["far_updates"] = 0
    for c, s_ in zip(cs, sims):
        _, _, ww, wh = s_.lattice_window()
        p = c.params
        tm["far_updates"] += far_updates(p.nx, p.ny, ww, wh)
    b.close()
    for s_ in sims:
        s_.close()
    return dt, hist, tm, shape


def bench_batch(args):
    from fortran_loadfile_mitiwing_amd import cases, lbmib
    steps = args.sweep_steps
    cs = cases.sweep_cases(args.cases, n_steps=steps)
    nodes = sum(c.params.nx * c.params.ny for c in cs)
    interior = sum((c.params.nx - 2) * (c.params.ny - 2) for c in cs)
    print("[batch] %d cases, %d lattice nodes, %d steps each" % (len(cs), nodes, steps), file=sys.stderr, flush=True)
    if args.only_batched:
        t_batch, h32, tm, shape = run_batch(cs, lbmib.PREC_FP32DEV, steps, not args.no_tune, label="fp32",
                                            shape=parse_shape(args.shape))
        print(json.dumps({"form": "batch (batched FP32 only)", "ms_per_step": round(t_batch / steps * 1e3, 4),
                          "phases_ms": tm, "launch_shape": shape}))
        return
    # sequential plain FP64 runs, each case alone on one reused engine (a
    # batch of one: the plain path's launches): setup + 20 untimed + 200 timed
    # steps per case, the run time taken as setup + steps x (ms per timed step)
    t_seq, seq_w, seq_k = 0.0, 20, 200
    seq = lbmib.Batch()
    for i, c in enumerate(cs):
        t0 = time.perf_counter()
        g = lbmib.LBMIB(c)
        seq.add(g)
        t_setup = time.perf_counter() - t0
        seq.step(seq_w)
        t0 = time.perf_counter()
        seq.step(seq_k)
        t_seq += t_setup + (time.perf_counter() - t0) / seq_k * steps
        seq.remove(g)
        g.close()
        if i % 8 == 7:
            print("[batch] sequential FP64 timing: %d cases, %.1f s projected" % (i + 1, t_seq),
                  file=sys.stderr, flush=True)
    seq.close()
    t_batch, h32, tm, shape = run_batch(cs, lbmib.PREC_FP32DEV, steps, not args.no_tune, label="fp32")
    # FP64 reference histories of the same cases (the batch's per-case
    # arithmetic is the single-case path's, bit for bit)
    t_b64, h64, _, _ = run_batch(cs, lbmib.PREC_FP64, steps, False, label="fp64")
    finite = [(bool(np.all(np.isfinite(a[:, 1:3]))), bool(np.all(np.isfinite(b_[:, 1:3])))) for a, b_ in zip(h32, h64)]
    diverged = [{"case": c.name, "fp32_finite": f[0], "fp64_finite": f[1]} for c, f in zip(cs, finite) if not all(f)]
    keep = [i for i, f in enumerate(finite) if all(f)]
    errs = [coefficient_errors(h32[i], h64[i]) for i in keep]
    cs_ok = [cs[i] for i in keep]
    keys = ("mean_CD", "mean_CL", "peak_CL")
    worst = {k: max(e[k] for e in errs) for k in keys + ("mean_CD_rel",)}
    order = sorted(range(len(errs)), key=lambda i: -errs[i]["mean_CD_rel"])[:5]
    worst_cases = [{"case": cs_ok[i].name, "mean_CD_fp64": round(errs[i]["_mean_CD64"], 5),
                    "rms_CD_fp64": round(errs[i]["_rms_CD64"], 5),
                    **{k: float("%.3g" % errs[i][k]) for k in keys + ("mean_CD_rel",)}} for i in order]
    if tm["far_passes"] > 0:
        # two-step passes: the dominant kernel is the far pass over all cases
        kern = "k_far2<float> over all cases"
        bytes_launch = tm["far_updates"] * BYTES_PER_NODE_F32
        lat_s = tm["far_pass_ms"] * 1e-3
    else:
        kern = "k_lattice<float,%d> over all cases" % shape["nodes_per_thread"]
        bytes_launch = interior * BYTES_PER_NODE_F32
        lat_s = tm["lattice_ms"] * 1e-3
    achieved = bytes_launch / lat_s / 1e9
    out = {
        "metric": "sweep cases/hr (batched FP32-deviation form, %d cases x %d steps)" % (len(cs), steps),
        "value": round(len(cs) / t_batch * 3600.0, 1),
        "unit": "cases/hr",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": round(t_batch / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 lattice, f64 immersed boundary",
        "data": "synthetic (seeded sweep of NACA0012 cases, seed %d)" % cases.SWEEP_SEED,
        "config": {"workload": "parameter sweep: %d cases, nx in 512..2048 step 128, ny = nx/2, chord nx/16, "
                               "4*chord markers, Re 100..1000, heave 0.25..1c, pitch 0..45 deg, phase 0..180 deg, "
                               "St 0.1..0.5, %d steps each" % (len(cs), steps),
                   "lattice_nodes": nodes},
        "sequential_plain_fp64_cases_per_hr": round(len(cs) / t_seq * 3600.0, 1),
        "sequential_plain_timing": "per case alone (FP64, one reused engine): setup + %d untimed + %d timed "
                                   "steps, scaled to %d steps" % (seq_w, seq_k, steps),
        "speedup_vs_sequential_plain": round(t_seq / t_batch, 3),
        "batched_fp64_cases_per_hr": round(len(cs) / t_b64 * 3600.0, 1),
        "aggregate_MLUPS": round(nodes * steps / t_batch / 1e6, 1),
        "accuracy_vs_fp64_worst_case": {k: float("%.3g" % v) for k, v in worst.items()},
        "accuracy_window": "last half of each run; mean C_D and mean C_L relative to their r.m.s., peak |C_L| "
                           "relative; mean_CD_rel relative to |mean C_D|",
        "cases_compared": len(errs),
        "cases_within_0.5pct": int(sum(all(e[k] <= 0.005 for k in keys) for e in errs)),
        "cases_within_0.5pct_incl_mean_CD_rel": int(sum(all(e[k] <= 0.005 for k in keys + ("mean_CD_rel",))
                                                        for e in errs)),
        "largest_mean_CD_rel_cases": worst_cases,
        "diverged_cases": diverged,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kern,
                     "bytes_per_launch": int(bytes_launch), "mean_launch_ms": round(lat_s * 1e3, 5)},
        "phases_ms": {k: round(v, 5) for k, v in tm.items()},
        "launch_shape": shape,
    }
    print(json.dumps(out))


def bench_serve(args):
    from fortran_loadfile_mitiwing_amd import cases, lbmib
    from fortran_loadfile_mitiwing_amd.serve import GPUBackend, Server
    n_req = args.requests
    # capacity: a saturated batch of 16 cases of the serving distribution
    probe = [r.case for r in cases.serving_requests(16, rate_per_s=1.0, seed=cases.SWEEP_SEED + 2)]
    sims = [lbmib.LBMIB(c, precision=lbmib.PREC_FP32DEV) for c in probe]
    b = lbmib.Batch()
    for s_ in sims:
        b.add(s_)
    shape = b.autotune(trial_steps=32)["best"] if not args.no_tune else b.shape()
    b.step(16)
    t0 = time.perf_counter()
    b.step(96)
    dt = time.perf_counter() - t0
    cap_mlups = sum(c.params.nx * c.params.ny for c in probe) * 96 / dt / 1e6
    b.close()
    for s_ in sims:
        s_.close()
    print("[serve] capacity %.0f MLUPS (16-case probe), shape %s" % (cap_mlups, shape), file=sys.stderr, flush=True)
    # requests: Poisson arrivals offering args.load of that capacity
    reqs = cases.serving_requests(n_req, rate_per_s=1.0)
    work = np.array([r.case.params.nx * r.case.params.ny * r.case.n_steps for r in reqs], dtype=np.float64)
    rate = args.load * cap_mlups * 1e6 / work.mean()
    for r in reqs:
        r.arrival = r.arrival / rate
    free, total = lbmib.mem_info(0)
    budget = int(min(free * 0.85, args.budget_gb * 1e9)) if args.budget_gb else int(free * 0.85)
    print("[serve] %d requests at %.3f /s (load %.2f), HBM budget %.1f GB" % (n_req, rate, args.load, budget / 1e9),
          file=sys.stderr, flush=True)
    srv = Server(GPUBackend(precision=lbmib.PREC_FP32DEV, shape=shape), hbm_budget=budget, window=args.window)
    last = [time.perf_counter()]

    def progress(s):
        if time.perf_counter() - last[0] > 5.0:
            last[0] = time.perf_counter()
            print("[serve] t=%.0fs done %d running %d queued %d" % (s.now(), len(s.results), len(s.running),
                                                                    len(s.queue)), file=sys.stderr, flush=True)

    res = srv.run(reqs, progress=progress)
    st = srv.stats()
    srv.close()
    # accuracy: a seeded sample of requests rerun in FP64 (one batch)
    rng = np.random.default_rng(7)
    pick = sorted(rng.choice(n_req, size=min(args.check, n_req), replace=False).tolist())
    byidx = {r.tag: r for r in res}
    sims = [lbmib.LBMIB(reqs[i].case) for i in pick]
    b = lbmib.Batch()
    for s_ in sims:
        b.add(s_)
    b.step(max(reqs[i].case.n_steps for i in pick))
    errs = [coefficient_errors(byidx[i].history, s_.history()) for i, s_ in zip(pick, sims)]
    b.close()
    for s_ in sims:
        s_.close()
    keys = ("mean_CD", "mean_CL", "peak_CL")
    out = {
        "metric": "serving aggregate MLUPS (continuous batching, FP32-deviation form, %d requests)" % n_req,
        "value": round(st["aggregate_MLUPS"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": int(sum(r.case.n_steps for r in reqs)),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 lattice, f64 immersed boundary",
        "data": "synthetic (seeded requests: NACA0012 cases, nx 512..4096, 2000..20000 steps log-uniform)",
        "config": {"workload": "serving: %d requests, Poisson arrivals at %.3f /s = %.0f%% of the %.0f MLUPS "
                               "capacity, admission by HBM footprint (budget %.1f GB), windows of %d steps"
                               % (n_req, rate, 100 * args.load, cap_mlups, budget / 1e9, args.window)},
        "p50_time_to_result_s": round(st["p50_time_to_result_s"], 3),
        "p99_time_to_result_s": round(st["p99_time_to_result_s"], 3),
        "stats": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()},
        "capacity_MLUPS_probe": round(cap_mlups, 1),
        "accuracy_vs_fp64_sample": {"requests": pick,
                                    "worst": {k: float("%.3g" % max(e[k] for e in errs)) for k in keys},
                                    "within_0.5pct": int(sum(all(e[k] <= 0.005 for k in keys) for e in errs))},
        "launch_shape": shape,
    }
    print(json.dumps(out))


def bench_precision(args):
    from fortran_loadfile_mitiwing_amd import cases, lbmib
    steps = args.precision_steps
    case = cases.config(2, n_steps=steps + 8)
    p = case.params
    runs = {}
    for name, prec in (("fp32", lbmib.PREC_FP32DEV), ("fp64", lbmib.PREC_FP64)):
        g = lbmib.LBMIB(case, device=0, precision=prec)
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            k = min(2500, steps - done)
            g.evolution(k)
            done += k
            print("[precision] %s %d steps, %.1f s" % (name, done, time.perf_counter() - t0), file=sys.stderr,
                  flush=True)
        dt = time.perf_counter() - t0
        runs[name] = {"s": dt, "mlups": p.nx * p.ny * steps / dt / 1e6, "hist": g.history(),
                      "fields": g.fields(with_f=False), "shape": g.shape()}
        g.close()
    f32, f64 = runs["fp32"], runs["fp64"]
    field_l2 = {}
    for k in ("rho", "u", "v"):
        a, b = f32["fields"][k], f64["fields"][k]
        field_l2[k] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    ce = coefficient_errors(f32["hist"], f64["hist"])
    out = {
        "metric": "MLUPS of the FP32-deviation form and its error against FP64 (precision workload)",
        "value": round(f32["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": round(f32["s"] / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 lattice, f64 immersed boundary",
        "data": "synthetic (generated NACA0012 markers and sinusoidal kinematics)",
        "config": {"workload": "%s: %s, FP32 deviation form against FP64, %d steps from the start" % (
            case.name, WORKLOADS[2], steps), "nx": p.nx, "ny": p.ny, "markers": p.pts_per_wing - 1},
        "fp64_MLUPS": round(f64["mlups"], 1),
        "fp32_over_fp64": round(f32["mlups"] / f64["mlups"], 3),
        "field_rel_L2_after_last_step": field_l2,
        "coefficient_errors_last_half": {k: v for k, v in ce.items() if not k.startswith("_")},
        "coefficient_error_notes": "mean C_L and mean C_D relative to their r.m.s. over the last half, "
                                   "mean_CD_rel relative to |mean C_D|, peak |C_L| relative",
        "launch_shape": f32["shape"],
    }
    print(json.dumps(out))


def bench_dynamics(args):
    """The step workload with the rigid-body dynamics of wing 1 on (the
    reference's body parameters): each step's geometry is built after the
    force of the step before, so the steps run one window each."""
    from fortran_loadfile_mitiwing_amd import cases, lbmib
    K, W = args.steps, args.warmup
    case = cases.config(1, n_steps=K + W + 8)
    p = case.params
    dyn = cases.Dynamics()
    g = lbmib.LBMIB(case, device=0, dynamics=dyn)
    if W:
        g.evolution(W)
    g.set_profiling(PROFILE_EVERY)
    t0 = time.perf_counter()
    g.evolution(K)
    dt = time.perf_counter() - t0
    tm = g.timings()
    cells = p.nx * p.ny
    interior = (p.nx - 2) * (p.ny - 2)
    lat_s = tm["lattice_ms"] * 1e-3
    achieved = interior * BYTES_PER_NODE_F64 / lat_s / 1e9
    st = g.body_state()
    out = {
        "metric": "MLUPS (whole FP64 timestep with the rigid-body dynamics of wing 1)",
        "value": round(cells * K / dt / 1e6, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(dt / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated NACA0012 markers and sinusoidal kinematics)",
        "config": {"workload": "%s: %s, rigid-body dynamics (solid_den %g, area %g, rk4)" % (
            case.name, WORKLOADS[1], dyn.solid_den, dyn.area_airfoil), "nx": p.nx, "ny": p.ny,
            "markers": p.pts_per_wing - 1},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_lattice<double>",
                     "bytes_per_launch": interior * BYTES_PER_NODE_F64, "mean_launch_ms": round(lat_s * 1e3, 5)},
        "phases_ms": {k: round(v, 5) for k, v in tm.items()},
        "body": {"velocity": st[2], "drift_nodes": st[4], "domain_shifts": int(st[-1])},
    }
    g.close()
    print(json.dumps(out))


def bench_post(args):
    """Post-processing on the step workload: output's vorticity with one
    vor_mom_rate update per call (k_vorticity) and cal_sur_cp
    (k_surface_pressure), on the device from the field of the last step."""
    from fortran_loadfile_mitiwing_amd import cases, lbmib
    K, W = args.steps, args.warmup
    case = cases.config(1, n_steps=W + 8)
    p = case.params
    g = lbmib.LBMIB(case, device=0)
    g.evolution(max(W, 2))
    for _ in range(3):
        g.vorticity(rate=True, copy=False)
    g.set_profiling(1)  # events around each vorticity kernel
    t0 = time.perf_counter()
    for _ in range(K):
        g.vorticity(rate=True, copy=False)
    dt = time.perf_counter() - t0
    tm = g.timings()
    n = p.nx * p.ny
    kern_s = tm["vorticity_ms"] * 1e-3
    bpn = 64  # u, v, two history planes read; vorticity, two history planes, rate written
    achieved = n * bpn / kern_s / 1e9
    t1 = time.perf_counter()
    for _ in range(K):
        g.surface_pressure()
    sp_ms = (time.perf_counter() - t1) / K * 1e3
    out = {
        "metric": "vorticity + vor_mom_rate nodes per second on the device (post-processing of one step)",
        "value": round(n / kern_s / 1e6, 1),
        "unit": "Mnodes/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(dt / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated NACA0012 markers and sinusoidal kinematics)",
        "config": {"workload": "%s: %s, field after %d steps" % (case.name, WORKLOADS[1], max(W, 2)),
                   "nx": p.nx, "ny": p.ny, "markers": p.pts_per_wing - 1},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_vorticity",
                     "bytes_per_launch": n * bpn, "mean_launch_ms": round(kern_s * 1e3, 5)},
        "surface_pressure_call_ms": round(sp_ms, 4),
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        fl = g.fields(with_f=False)
        st = pyoracle.VorticityRate(p.nx, p.ny, p.U, p.scale)
        reps, t2 = 0, time.perf_counter()
        while reps < 3 or time.perf_counter() - t2 < 5.0:
            pyoracle.vorticity(fl["u"], fl["v"], p.U, p.scale)
            st.update(fl["u"], fl["v"])
            reps += 1
        cpu_s = (time.perf_counter() - t2) / reps
        out["cpu_baseline"] = {"value": round(n / cpu_s / 1e6, 2), "unit": "Mnodes/s", "cores": 1, "kind": "port",
                               "sample": "%d calls of the oracle's vorticity + vor_mom_rate on the same field" % reps}
    g.close()
    print(json.dumps(out))


def precision_summary(steps=200):
    """FP32 deviation form against FP64 on the 8192x4096 workload over a fixed
    window from rest: MLUPS of both (whole steps, after one geometry batch of
    warm-up), their ratio, and the FP32 run's C_L / C_D and field errors."""
    from fortran_loadfile_mitiwing_amd import cases, lbmib
    warm = 16
    case = cases.config(2, n_steps=warm + steps + 8)
    p = case.params
    runs = {}
    for name, prec in (("fp32", lbmib.PREC_FP32DEV), ("fp64", lbmib.PREC_FP64)):
        g = lbmib.LBMIB(case, device=0, precision=prec)
        g.evolution(warm)
        t0 = time.perf_counter()
        g.evolution(steps)
        dt = time.perf_counter() - t0
        runs[name] = (p.nx * p.ny * steps / dt / 1e6, g.history(), g.fields(with_f=False))
        g.close()
        print("[summary] precision %s %.0f MLUPS" % (name, runs[name][0]), file=sys.stderr, flush=True)
    (m32, h32, f32), (m64, h64, f64) = runs["fp32"], runs["fp64"]
    ce = coefficient_errors(h32, h64)
    return {"workload": "%s, %d timed steps after %d" % (WORKLOADS[2], steps, warm),
            "fp32_MLUPS": round(m32, 1), "fp64_MLUPS": round(m64, 1), "fp32_over_fp64": round(m32 / m64, 3),
            "CL_CD_errors_last_half": {k: float("%.3g" % v) for k, v in ce.items() if not k.startswith("_")},
            "field_rel_L2": {k: float("%.3g" % (np.linalg.norm(f32[k] - f64[k]) / np.linalg.norm(f64[k])))
                             for k in ("rho", "u", "v")},
            "notes": "mean C_L / C_D relative to their r.m.s. over the last half of the run, mean_CD_rel to "
                     "|mean C_D|, peak |C_L| relative"}


def sweep_summary(n_cases=96, steps=400):
    """The seeded sweep's cases cut to `steps` steps: all of them as one
    batched FP32-deviation run, then each alone as a fresh plain FP64 handle
    (the sequential baseline, measured, not projected; it also gives the FP64
    histories the batch is checked against).  Times include creating the
    cases.  A case whose history is not finite counts as failed, not as
    throughput."""
    from fortran_loadfile_mitiwing_amd import cases, lbmib
    cs = cases.sweep_cases(n_cases, n_steps=steps)
    nodes = sum(c.params.nx * c.params.ny for c in cs)
    t0 = time.perf_counter()
    sims = [lbmib.LBMIB(c, precision=lbmib.PREC_FP32DEV) for c in cs]
    b = lbmib.Batch()
    b.set_shape(ib_collide_shared=1)
    for s_ in sims:
        b.add(s_)
    b.step(steps)
    t_batch = time.perf_counter() - t0
    h32 = [s_.history() for s_ in sims]
    b.close()
    for s_ in sims:
        s_.close()
    print("[summary] sweep batch %.2f s" % t_batch, file=sys.stderr, flush=True)
    t_seq, h64 = 0.0, []
    for i, c in enumerate(cs):
        t1 = time.perf_counter()
        g = lbmib.LBMIB(c)
        g.evolution(steps)
        h64.append(g.history())
        g.close()
        t_seq += time.perf_counter() - t1
        if i % 32 == 31:
            print("[summary] sweep sequential %d cases %.2f s" % (i + 1, t_seq), file=sys.stderr, flush=True)
    ok = [i for i in range(len(cs)) if np.all(np.isfinite(h32[i][:, 1:4])) and np.all(np.isfinite(h64[i][:, 1:4]))]
    errs = [coefficient_errors(h32[i], h64[i]) for i in ok]
    keys = ("mean_CD", "mean_CL", "peak_CL")
    within = [i for i, e in zip(ok, errs) if all(e[k] <= 0.005 for k in keys)]
    return {"workload": "%d seeded sweep cases (nx 512..2048, tau >= %.2f, marker Mach <= %.1f), %d steps each"
                        % (len(cs), cases.TAU_MIN, cases.MACH_MAX, steps),
            "lattice_nodes": nodes,
            "batched_fp32_cases_per_hr": round(len(within) / t_batch * 3600.0, 1),
            "sequential_fp64_cases_per_hr": round(len(cs) / t_seq * 3600.0, 1),
            "speedup": round(t_seq / t_batch * len(within) / len(cs), 3),
            "batched_aggregate_MLUPS": round(nodes * steps / t_batch / 1e6, 1),
            "cases_failed": len(cs) - len(ok), "cases_within_0.5pct": len(within),
            "worst_CL_CD_errors": {k: float("%.3g" % max(e[k] for e in errs)) for k in keys + ("mean_CD_rel",)},
            "notes": "batched cases/hr counts only cases that stay finite and meet the 0.5% bar against their "
                     "FP64 run; times include creating the cases; errors over the last half of each run"}


def serve_summary(n_req=48, load=0.8, window=32):
    """A short serving run: requests of the serving distribution with step
    counts cut tenfold (200..2000, log-uniform), Poisson arrivals offering
    `load` of the capacity a 16-case probe batch measures, continuous
    batching in the FP32 deviation form; p50/p99 time-to-result."""
    from fortran_loadfile_mitiwing_amd import cases, lbmib
    from fortran_loadfile_mitiwing_amd.serve import GPUBackend, Server
    probe = [r.case for r in cases.serving_requests(16, rate_per_s=1.0, seed=cases.SWEEP_SEED + 2,
                                                    steps_lo=200, steps_hi=2000)]
    sims = [lbmib.LBMIB(c, precision=lbmib.PREC_FP32DEV) for c in probe]
    b = lbmib.Batch()
    b.set_shape(ib_collide_shared=1)
    for s_ in sims:
        b.add(s_)
    b.step(16)
    t0 = time.perf_counter()
    b.step(64)
    cap = sum(c.params.nx * c.params.ny for c in probe) * 64 / (time.perf_counter() - t0) / 1e6
    b.close()
    for s_ in sims:
        s_.close()
    reqs = cases.serving_requests(n_req, rate_per_s=1.0, steps_lo=200, steps_hi=2000)
    work = np.array([r.case.params.nx * r.case.params.ny * r.case.n_steps for r in reqs], dtype=np.float64)
    rate = load * cap * 1e6 / work.mean()
    for r in reqs:
        r.arrival = r.arrival / rate
    free, _ = lbmib.mem_info(0)
    srv = Server(GPUBackend(precision=lbmib.PREC_FP32DEV, shape={"ib_collide_shared": 1}),
                 hbm_budget=int(free * 0.85), window=window)
    srv.run(reqs)
    st = srv.stats()
    srv.close()
    print("[summary] serve p50 %.3f s p99 %.3f s" % (st["p50_time_to_result_s"], st["p99_time_to_result_s"]),
          file=sys.stderr, flush=True)
    return {"workload": "%d requests (nx 512..4096, 200..2000 steps), Poisson arrivals at %.2f /s = %.0f%% of the "
                        "%.0f MLUPS capacity, windows of %d steps" % (n_req, rate, 100 * load, cap, window),
            "p50_time_to_result_s": round(st["p50_time_to_result_s"], 3),
            "p99_time_to_result_s": round(st["p99_time_to_result_s"], 3),
            "aggregate_MLUPS": round(st["aggregate_MLUPS"], 1),
            "mean_running_cases": round(st["mean_running_cases"], 2),
            "peak_running_cases": st["peak_running_cases"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-summaries", action="store_true",
                    help="step form: leave out the precision, sweep and serving summaries of the default line")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                    help="fp64: the drop-in path (default); fp32: the FP32 deviation form")
    ap.add_argument("--shape", default="",
                    help="launch shape of the step form (and of --form batch --only-batched, instead of the "
                         "autotuner), e.g. nodes_per_thread=2,tile_order=1,persistent=4,ib_cus=16,"
                         "ib_collide_shared=0,ib_rhs_reserved=1 (default: the library's)")
    ap.add_argument("--solver", choices=["auto", "chain", "separator"], default="auto",
                    help="IB solve of the step form (auto: one chain in FP64, separator form in FP32)")
    ap.add_argument("--form", choices=["step", "batch", "serve", "precision", "post", "dynamics"], default="step",
                    help="step: one case's timestep (default); batch: the 96-case parameter sweep; "
                         "serve: the continuous-batching service; precision: FP32 against FP64 on the "
                         "8192x4096 workload")
    ap.add_argument("--precision-steps", type=int, default=50000, help="--form precision: steps of each run")
    ap.add_argument("--requests", type=int, default=500, help="--form serve: requests")
    ap.add_argument("--load", type=float, default=0.8, help="--form serve: offered load / capacity")
    ap.add_argument("--window", type=int, default=32, help="--form serve: steps per scheduling window")
    ap.add_argument("--budget-gb", type=float, default=0.0, help="--form serve: HBM budget (0: 85%% of free)")
    ap.add_argument("--check", type=int, default=8, help="--form serve: requests rerun in FP64")
    ap.add_argument("--cases", type=int, default=96, help="--form batch: cases in the sweep")
    ap.add_argument("--sweep-steps", type=int, default=10000, help="--form batch: steps per case")
    ap.add_argument("--no-tune", action="store_true", help="--form batch: keep the default launch shape")
    ap.add_argument("--only-batched", action="store_true", help="--form batch: only the batched FP32 run (profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 1 GiB (the known-size kernel that scales PMC byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 8, "copy_bytes_written": y.numel() * 8}))
        return
    if args.gpus != 1:
        print("bench.py: this build runs one case on one GPU; --gpus %d is not supported "
              "(the path shards by case across GPUs, see the design notes)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.form == "batch":
        return bench_batch(args)
    if args.form == "serve":
        return bench_serve(args)
    if args.form == "precision":
        return bench_precision(args)
    if args.form == "post":
        return bench_post(args)
    if args.form == "dynamics":
        return bench_dynamics(args)
    from fortran_loadfile_mitiwing_amd import cases, lbmib
    K, W = args.steps, args.warmup
    fp32 = args.precision == "fp32"
    case = cases.config(args.config, n_steps=K + W + 8)
    p = case.params
    solver = {"auto": lbmib.SOLVER_AUTO, "chain": lbmib.SOLVER_CHAIN, "separator": lbmib.SOLVER_SEPARATOR}[args.solver]
    g = lbmib.LBMIB(case, device=0, precision=lbmib.PREC_FP32DEV if fp32 else lbmib.PREC_FP64, ib_solver=solver)
    if args.shape:
        g.set_shape(**parse_shape(args.shape))
    if W:
        g.evolution(W)
    # phase events inside the timed region: on every step of a run shorter
    # than 16 steps, every 4th step (every other two-step pass) up to 64
    # steps, else every PROFILE_EVERY-th step.  Each sampled step costs the
    # run about 1% of its time at 20 steps.
    g.set_profiling(1 if K < 16 else 4 if K <= 64 else PROFILE_EVERY)
    t0 = time.perf_counter()
    g.evolution(K)  # synchronises the device before returning
    dt = time.perf_counter() - t0
    tm = g.timings()
    cells = p.nx * p.ny
    interior = (p.nx - 2) * (p.ny - 2)
    bpn = BYTES_PER_NODE_F32 if fp32 else BYTES_PER_NODE_F64
    wi0, wj0, ww, wh = g.lattice_window()
    if tm["far_passes"] > 0:
        # two-step passes: the far-field kernel's node updates per launch
        kern = "k_far2<%s>" % ("float" if fp32 else "double")
        bytes_launch = far_updates(p.nx, p.ny, ww, wh) * bpn
        lat_s = tm["far_pass_ms"] * 1e-3
        traffic = None if fp32 else FAR_TRAFFIC_BYTES.get((p.nx, p.ny))
    else:
        kern = "k_lattice<%s>" % ("float" if fp32 else "double")
        bytes_launch = interior * bpn
        lat_s = tm["lattice_ms"] * 1e-3
        traffic = None if fp32 else LATTICE_TRAFFIC_BYTES.get((p.nx, p.ny))
    achieved = bytes_launch / lat_s / 1e9
    value = cells * K / dt / 1e6
    out = {
        "metric": "MLUPS (whole %s timestep incl. immersed boundary)" % ("FP32-deviation" if fp32 else "FP64"),
        "value": round(value, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(dt / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 lattice, f64 immersed boundary" if fp32 else "f64",
        "data": "synthetic (generated NACA0012 markers and sinusoidal kinematics)",
        "config": {"workload": "%s: %s, %s" % (case.name, WORKLOADS.get(args.config, ""),
                                                "FP32 deviation form" if fp32 else "FP64 drop-in"),
                   "nx": p.nx, "ny": p.ny, "markers": p.pts_per_wing - 1},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kern,
                     "bytes_per_launch": bytes_launch, "mean_launch_ms": round(lat_s * 1e3, 5),
                     "traffic_frac": round(traffic / lat_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None},
        "lattice_window": {"i0": wi0, "j0": wj0, "width": ww, "height": wh},
        "whole_step_hbm_frac": round(value * 1e6 * bpn / (HBM_PEAK_GBS * 1e9), 4),
        "phases_ms": {k: round(v, 5) for k, v in tm.items()},
        "launch_shape": g.shape(),
    }
    g.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case)
    if not args.no_summaries and not fp32 and args.config == 1:
        out["precision"] = precision_summary()
        out["sweep"] = sweep_summary()
        out["serve"] = serve_summary()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
