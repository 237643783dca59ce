#!/usr/bin/env python3
"""BiCGSTAB throughput on one MI355X: FP32 + FP64 refinement vs FP64, to 1e-8 relative residual.

Workload (default): steady advection-diffusion div(U T) - div(kappa grad T) = 0 on a 64^3
hexahedral unit cube with interior nodes jittered by +-0.2h (seed 1), prescribed solenoidal
U = (sin pi x cos pi y, -cos pi x sin pi y, 0.3), kappa = 1e-3, MUSCL convection,
over-relaxed corrected diffusion, T = 1 on x = 0, zero-gradient elsewhere.  The system is
the one the reference assembles on the first outer iteration: MUSCL on the initial uniform T,
whose limiter ratios are 0/0 and therefore take the upwind weights (convection.cpp:69-128).

One step = reset T to 0 on the device, then one BiCGSTAB::step of the assembled system to
||b - A x|| <= 1e-8 ||b|| (inputs resident in HBM).  value = cells x BiCGSTAB iterations /
time of the FP32+refinement solve, in Mcell-iterations/s.  The FP64 solve of the same
system is timed the same way and reported beside it.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BiCGSTAB Mcell-iters/s + HBM GB/s, FP32+refine vs FP64 at 1e-8 rel residual"
HBM_PEAK_GBS = 8000.0


def spmv_bytes(n, nnz, w):
    """Algorithmic bytes of one bicgstab_spmv_v launch: values + column indices + row
    pointers, the gathered vector once, the product written once, r0 read for (r0, v)."""
    return (w + 4) * nnz + 4 * (n + 1) + 3 * w * n


def iteration_bytes(n, nnz, w):
    """Per BiCGSTAB iteration: 2 SpMV (values, indices, row pointers, gathered vector once,
    result once) + 15 vector streams (the compulsory-traffic model used for HBM GB/s)."""
    spmv = w * nnz + 4 * nnz + 4 * (n + 1) + 2 * w * n
    return 2 * spmv + 15 * w * n


def build_case(be, P, cases, nx):
    data = P.box(nx, nx, nx, jitter=0.2, seed=1)
    c = cases.advection_diffusion(be, data, scheme="MUSCL", kappa=1e-3)
    eqn = c["eqn"]
    eqn.update_coeffs()  # first-outer-iteration system (device assembly)
    return c, data


def timed_solves(be, eqn, T, slv, steps, warmup):
    for _ in range(warmup):
        T.fill(0.0)
        slv.step_eqn(eqn)
    be.synchronize()
    iters = 0
    res = None
    t0 = time.perf_counter()
    for _ in range(steps):
        T.fill(0.0)
        res = slv.step_eqn(eqn)
        iters += res.krylov_iters
    be.synchronize()
    dt = time.perf_counter() - t0
    return dt, iters, res


def cpu_baseline(nx, budget_s=15.0):
    """FP64 CPU restatement (oracle, one thread): the same assembled MUSCL system, the same
    unpreconditioned BiCGSTAB iteration, a bounded number of iterations."""
    import numpy as np

    import oracle
    import prism_amd as P
    from prism_amd import cases

    be = oracle.backend()
    data = P.box(nx, nx, nx, jitter=0.2, seed=1)
    t0 = time.perf_counter()
    c = cases.advection_diffusion(be, data, scheme="MUSCL", kappa=1e-3)
    eqn = c["eqn"]
    eqn.update_coeffs()  # upwind-weighted (uniform T) system
    t_asm = time.perf_counter() - t0
    rp, col, val, n = eqn.matrix()
    b = eqn.rhs()
    A = P.Matrix(be, rp, col, val)
    # calibrate: time a few iterations, then run a bounded sample
    slv = P.solver.BiCGSTAB(tol=1e-30, max_iters=5)
    t0 = time.perf_counter()
    res = []
    slv.step(A, np.zeros(n), b, res)
    per_it = (time.perf_counter() - t0) / max(1, res[0].krylov_iters)
    k = int(max(5, min(2000, budget_s / max(per_it, 1e-9))))
    slv = P.solver.BiCGSTAB(tol=1e-30, max_iters=k)
    res = []
    t0 = time.perf_counter()
    slv.step(A, np.zeros(n), b, res)
    dt = time.perf_counter() - t0
    it = res[0].krylov_iters
    return {"value": round(n * it / dt / 1e6, 3), "unit": "Mcell-iters/s", "cores": 1, "kind": "port",
            "sample": f"{it} unpreconditioned FP64 BiCGSTAB iterations on the {nx}^3 system "
                      f"({n} cells, {len(val)} nnz) in {dt:.2f} s, single thread, -O2",
            "assembly_s": round(t_asm, 3), "cpu": platform.processor() or platform.machine(),
            "host_cores_available": os.cpu_count()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64, help="cells per axis")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-8)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this path has no multi-GPU mode (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import prism_amd as P
    from prism_amd import cases

    be = P.Backend.hip(0)
    c, data = build_case(be, P, cases, args.n)
    eqn, T = c["eqn"], c["T"]
    mesh = c["mesh"]
    n, nnz = mesh.n_cells, mesh.nnz

    s32 = P.solver.BiCGSTAB(tol=args.tol, precision="fp32")
    s64 = P.solver.BiCGSTAB(tol=args.tol, precision="fp64")
    t32, it32, r32 = timed_solves(be, eqn, T, s32, args.steps, args.warmup)
    t64, it64, r64 = timed_solves(be, eqn, T, s64, args.steps, args.warmup)

    # assembly throughput (device): gradient refresh + face + cell kernels per pass
    asm_reps = 20
    be.synchronize()
    t0 = time.perf_counter()
    for _ in range(asm_reps):
        T.fill(0.0)
        eqn.zero_out_coeffs()
        eqn.update_coeffs()
    be.synchronize()
    t_asm = (time.perf_counter() - t0) / asm_reps
    # restore the benchmark system (T = 0 again)
    T.fill(0.0)
    eqn.zero_out_coeffs()
    eqn.update_coeffs()

    # roofline pass: per-launch HIP events on the library stream
    be.set_timing(True)
    be.reset_timing()
    for _ in range(3):
        T.fill(0.0)
        s32.step_eqn(eqn)
    stats32 = be.kernel_stats()
    be.reset_timing()
    for _ in range(3):
        T.fill(0.0)
        s64.step_eqn(eqn)
    stats64 = be.kernel_stats()
    be.set_timing(False)
    k32 = stats32.get("bicgstab_spmv_v_f32", (1, float("nan")))
    k64 = stats64.get("bicgstab_spmv_v_f64", (1, float("nan")))
    avg32_ms = k32[1] / k32[0]
    avg64_ms = k64[1] / k64[0]
    b32 = spmv_bytes(n, nnz, 4)
    b64 = spmv_bytes(n, nnz, 8)
    ach32 = b32 / (avg32_ms * 1e-3) / 1e9
    ach64 = b64 / (avg64_ms * 1e-3) / 1e9

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_bicgstab_spmv_v_f32.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        if tj.get("n") == n:
            traffic = tj.get("bytes_per_launch")

    ms32 = t32 / args.steps * 1e3
    ms64 = t64 / args.steps * 1e3
    v32 = n * it32 / t32 / 1e6
    v64 = n * it64 / t64 / 1e6
    gbs32 = v32 * 1e6 * iteration_bytes(n, nnz, 4) / n / 1e9
    gbs64 = v64 * 1e6 * iteration_bytes(n, nnz, 8) / n / 1e9
    name, ncu = be.device_info()
    out = {
        "metric": METRIC,
        "value": round(v32, 2),
        "unit": "Mcell-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 residual/dots/solution)",
        "data": "synthetic",
        "config": {"workload": f"advection-diffusion {args.n}^3 jittered hex, MUSCL (first outer iteration), "
                               f"kappa 1e-3, BiCGSTAB to {args.tol:g}",
                   "cells": n, "nnz": nnz, "faces": mesh.n_faces, "tol": args.tol},
        "hbm_gbs_model": round(gbs32, 1),
        "iters_per_solve": it32 / args.steps,
        "refinements": r32.refinements,
        "rel_residual": r32.krylov_rel_residual,
        "fp64": {"value": round(v64, 2), "ms_per_step": round(ms64, 4), "iters_per_solve": it64 / args.steps,
                 "hbm_gbs_model": round(gbs64, 1), "rel_residual": r64.krylov_rel_residual,
                 "spmv_avg_us": round(avg64_ms * 1e3, 3), "spmv_achieved_gbs": round(ach64, 1)},
        "speedup_time_to_solution": round(t64 / t32, 3),
        "speedup_mcell_iters": round(v32 / v64, 3),
        "assembly": {"faces_per_s": round(mesh.n_faces / t_asm, 1), "ms_per_pass": round(t_asm * 1e3, 4)},
        "roofline": {"bound": "hbm", "kernel": "bicgstab_spmv_v_f32", "achieved": round(ach32, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach32 / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "avg_us": round(avg32_ms * 1e3, 3), "algorithmic_bytes": b32},
        "device": {"arch": name, "cus": ncu},
    }
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(args.n)
        except Exception as e:  # the CPU leg never blocks the GPU figure
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
