#!/usr/bin/env python3
"""bench.py — throughput of the fractional-step LES step on one MI355X.

Workload: wind-tunnel channel of 512x128x128 cells (dx = D/32) with one
actuator-line rotor (3 blades x 32 points, TSR 7), Upwind3 + Smagorinsky,
BiCGStab with a red-black SOR preconditioner at tolerance 1e-6 on
|r|/N.  The headline figure is MLUPS of the mixed-precision step (float
stencils and Krylov passes, double defect correction); the FP64 step runs
beside it on the same workload.  A step is one full time step (ALM, pseudo
velocity, face flux, Poisson RHS, Poisson solve, projection, boundary fill,
SGS, divergence monitor) with all fields resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS + steps/s @512x128x128 ALM: mixed vs FP64 GPU at div<1e-6; HBM GB/s"
PEAK_HBM_GBS = 8000.0


def run_gpu(precision, steps, warmup, n, time_family):
    """Every falm_case_step call returns after synchronising the case's
    stream, so the wall clock brackets completed device work."""
    from falm_amd import _abi as abi
    from falm_amd import cases, runtime

    setup = cases.windtunnel(n=n, precision=precision)
    c = runtime.Case(setup)
    for _ in range(warmup):
        c.step(1)
    runtime.timing_enable(time_family)
    its, outer = [], []
    t0 = time.perf_counter()
    info = None
    for _ in range(steps):
        info = c.step(1)
        its.append(info.solver_it)
        outer.append(info.outer_it)
    dt = time.perf_counter() - t0
    launches, kms, kbytes = runtime.timing_read()
    runtime.timing_enable(abi.TIME_NONE)
    u = c.get(abi.FIELD_U)
    c.close()
    cells = setup.n_interior
    return {
        "ms_per_step": 1e3 * dt / steps,
        "steps_per_s": steps / dt,
        "mlups": cells * steps / dt / 1e6,
        "solver_it_per_step": sum(its) / len(its),
        "defect_passes_per_step": sum(outer) / len(outer),
        "div": info.div,
        "solver_err": info.solver_err,
        "status": info.status,
        "kernel": {"launches": launches, "ms": kms, "bytes": kbytes},
    }, u, setup


def cpu_baseline(n, steps=2, warmup=1):
    """The oracle (FP64 C++ restatement of the reference scheme, OpenMP) on
    the host cores: a bounded sample of the same workload."""
    from falm_amd import cases
    from oracle import pyoracle as O

    setup = cases.windtunnel(n=n)
    c = O.Case(setup)
    for _ in range(warmup):
        c.step(1)
    t0 = time.perf_counter()
    for _ in range(steps):
        c.step(1)
    dt = time.perf_counter() - t0
    return {
        "value": setup.n_interior * steps / dt / 1e6,
        "unit": "MLUPS",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": f"{steps} steps after {warmup} warm-up step(s) of the same 512x128x128 + rotor workload, FP64, "
                  f"{dt / steps:.2f} s/step",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, nargs=3, default=[512, 128, 128])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              f"(asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from falm_amd import _abi as abi
    from falm_amd import runtime

    runtime.init(0)
    n = tuple(args.n)
    r64, u64, setup = run_gpu(abi.CASE_F64, args.steps, args.warmup, n, abi.TIME_SOR)
    rmx, umx, _ = run_gpu(abi.CASE_MIXED, args.steps, args.warmup, n, abi.TIME_SOR)
    nx, ny, nz = setup.shape
    mask = np.zeros((nz, ny, nx), dtype=bool)
    mask[2:-2, 2:-2, 2:-2] = True
    m3 = np.tile(mask.ravel(), 3)
    rel_l2 = float(np.linalg.norm(umx[m3] - u64[m3]) / np.linalg.norm(u64[m3]))

    k = rmx["kernel"]
    achieved = k["bytes"] / (k["ms"] * 1e-3) / 1e9 if k["ms"] > 0 else None
    roofline = {
        "bound": "hbm",
        "kernel": "k_sor<float,float> (red-black SOR half sweep, float Krylov pass)",
        "achieved": achieved,
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": achieved / PEAK_HBM_GBS if achieved else None,
        "traffic": None,
        "launches": k["launches"],
        "avg_launch_us": 1e3 * k["ms"] / max(k["launches"], 1),
        "algorithmic_bytes_per_launch": k["bytes"] / max(k["launches"], 1),
    }
    k64 = r64["kernel"]
    r64_achieved = k64["bytes"] / (k64["ms"] * 1e-3) / 1e9 if k64["ms"] > 0 else None
    tr_path = os.path.join(ROOT, "profiles", "traffic_sor.json")
    if os.path.exists(tr_path):
        with open(tr_path) as f:
            roofline["traffic"] = json.load(f).get("bytes_per_launch")

    cpu = None if args.no_cpu else cpu_baseline(n, steps=args.cpu_steps)

    out = {
        "metric": METRIC,
        "value": rmx["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rmx["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 stencils+Krylov / f64 defect correction",
        "data": "synthetic (uniform inflow + seeded inlet noise, synthetic blade polar)",
        "config": {
            "workload": f"windtunnel {n[0]}x{n[1]}x{n[2]} cells + 1 ALM rotor (3x32 points), dx=D/32, "
                        "Re 2.5e5, Cs 0.1, dt 0.3dx/U, BiCGStab+RB-SOR(2, w=1.5) tol 1e-6",
            "cells": setup.n_interior,
            "precision": "mixed",
        },
        "steps_per_s": rmx["steps_per_s"],
        "mixed": {k_: v for k_, v in rmx.items() if k_ != "kernel"},
        "fp64": {k_: v for k_, v in r64.items() if k_ != "kernel"},
        "fp64_sor_achieved_gbs": r64_achieved,
        "speedup_mixed_vs_fp64": r64["ms_per_step"] / rmx["ms_per_step"],
        "velocity_rel_l2_mixed_vs_fp64": rel_l2,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
