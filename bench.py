#!/usr/bin/env python3
"""Benchmark of the per-timestep hot path on one MI355X.

One step = one timestep of the FP64 drop-in path on the 4096x2048 workload
(NACA0012, chord 256 lattice units, 1024 markers, Re 1000, plunge 0.5c with
30 deg pitch leading by 90 deg, Strouhal 0.3, U = 0.05): marker kinematics,
immersed-boundary interpolate/solve/spread/force, fused stream + MRT collide,
outer boundaries.  The lattice starts resident in HBM; the timed region is
exactly K steps between two device synchronisations.

value = lattice updates per second of whole steps (MLUPS, nx*ny per step).
roofline = the fused stream+collide kernel (the dominant one): 144 algorithmic
bytes per interior node (9 FP64 populations read + 9 written) over its mean
launch time, timed with HIP events on the stream it runs on.
cpu_baseline = the CPU oracle (C restatement of the reference step, OpenMP
over the host cores) on the first steps of the same workload.
"""
import argparse
import json
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
    2: "8192x4096 lattice, NACA0012 chord 512 lu, 2048 markers, Re 2000, plunge 0.5c, St 0.3, U 0.05",
}
PROFILE_EVERY = 16
# HBM bytes per k_lattice_f64 launch on the 4096x2048 workload, from
# rocprofv3 FETCH_SIZE (x2.000) + WRITE_SIZE (x1.000), both scaled by a 1 GiB
# device copy measured the same way (the lattice traffic summary in profiles/)
LATTICE_TRAFFIC_BYTES = {(4096, 2048): (311140 * 2 + 590815) * 1024}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                    help="fp64: the drop-in path (default); fp32: the FP32 deviation form")
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
    from fortran_loadfile_mitiwing_amd import cases, lbmib
    K, W = args.steps, args.warmup
    fp32 = args.precision == "fp32"
    case = cases.config(args.config, n_steps=K + W + 8)
    p = case.params
    g = lbmib.LBMIB(case, device=0, precision=lbmib.PREC_FP32DEV if fp32 else lbmib.PREC_FP64)
    if W:
        g.evolution(W)
    # phase events on every PROFILE_EVERY-th timed step: they time the lattice
    # kernel live inside the timed region at well under 1% of its cost
    g.set_profiling(PROFILE_EVERY)
    t0 = time.perf_counter()
    g.evolution(K)  # synchronises the device before returning
    dt = time.perf_counter() - t0
    tm = g.timings()
    cells = p.nx * p.ny
    interior = (p.nx - 2) * (p.ny - 2)
    bpn = BYTES_PER_NODE_F32 if fp32 else BYTES_PER_NODE_F64
    lat_s = tm["lattice_ms"] * 1e-3
    achieved = interior * bpn / lat_s / 1e9
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
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if fp32 else LATTICE_TRAFFIC_BYTES.get((p.nx, p.ny)),
                     "kernel": "k_lattice<%s>" % ("float" if fp32 else "double"),
                     "bytes_per_launch": interior * bpn, "mean_launch_ms": round(lat_s * 1e3, 5)},
        "whole_step_hbm_frac": round(value * 1e6 * bpn / (HBM_PEAK_GBS * 1e9), 4),
        "phases_ms": {k: round(v, 5) for k, v in tm.items()},
    }
    g.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
