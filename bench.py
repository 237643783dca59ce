#!/usr/bin/env python3
"""BiCGSTAB throughput on one MI355X: FP32 + FP64 refinement vs FP64, to 1e-8 relative residual.

Workload (default, the line's `value`): the pressure-correction solve of the steady SIMPLE
lid-driven cubic cavity on a 128^3 hexahedral mesh (Re 100: nu 0.01, lid speed 1, SMART
momentum convection, under-relaxation 0.7 / 0.3).  Five FP64 SIMPLE outer iterations run
first; the sixth outer iteration's momentum predictor then runs and its pressure-correction
system is built exactly as the step builds it (D = V / a_P from the relaxed momentum
diagonals, Rhie-Chow face fluxes, p' with constrained face values, all-Neumann with the
automatic reference cell 0; SIMPLE.cpp:19-178).  One step = p' reset to 0, then the
step's ISolver::solve of that system (assembly, reference cell, BiCGSTAB to
||b - A x|| <= 1e-8 ||b||, scaled-residual check; inputs resident in HBM), in FP32 +
refinement; the FP64 solve of the same system is timed the same way beside it.
value = cells x BiCGSTAB iterations / time of the FP32+refinement solve (Mcell-iters/s);
speedup_time_to_solution = FP64 time / FP32+refinement time for the same solve.

Extra blocks (keys of the same JSON line): the same comparison on the pressure systems of
other outer iterations; the 64^3 jittered MUSCL advection-diffusion system (its working set
stays in the 256 MB last-level cache); SIMPLE outer iterations end to end; PISO on
tetrahedra; the discrete adjoint; export and k-epsilon kernels; the multicolour SGS
preconditioned solve of the 64^3 MUSCL systems against the unpreconditioned one; the CPU
baseline.

Other modes (no JSON line):
  --pmc-workload       one FP32 solve of the workload plus a calibration copy kernel
                       (the program profiled by the rocprofv3 --pmc passes);
  --traffic F.csv W.csv  turn the FETCH_SIZE and WRITE_SIZE pass outputs into
                       profiles/traffic_<kernel>.json (no GPU needed).
"""
import argparse
import csv
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BiCGSTAB Mcell-iters/s + HBM GB/s, FP32+refine vs FP64 at 1e-8 rel residual"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_bicgstab_spmv_v_f32.json")
CALIB_FLOATS = 64 * 2 ** 20  # 256 MiB calibration copy


DIA_MAX = 8


def assembly_bytes(n, faces, nnz):
    """Compulsory bytes of one assembly pass of one FP64 equation (whole pass, wall time):
    per face the owner/neighbour pair, S_f, x_f and the mass flux (64 B); per cell x_C, V,
    phi, grad phi, kappa and rhs (80 B); the CSR values written once (8 B per nonzero)."""
    return 64 * faces + 80 * n + 8 * nnz


def matrix_format(rp, col, val=None):
    """The library's storage choice for a CSR pattern: ("dia", D) when the entries take at
    most 8 distinct offsets col - row and fill at least 3/4 of the D x n slots (("dia_sym", D)
    when, in addition, the offsets mirror and `val` is exactly symmetric, for a solve without
    Jacobi), else ("sell16", 0) when sell16_fits, else ("sell32", 0)."""
    import numpy as np
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    off = col.astype(np.int64) - rows
    offs = np.unique(off)
    d = len(offs)
    if os.environ.get("PRISM_KRYLOV_NO_DIA", "0") in ("", "0") and d <= DIA_MAX and 4 * len(col) >= 3 * d * n:
        if val is not None and d % 2 == 1 and np.array_equal(offs, -offs[::-1]):
            import scipy.sparse as sp
            A = sp.csr_matrix((val, col, rp), shape=(n, n))
            if (A != A.T).nnz == 0:
                return "dia_sym", d
        return "dia", d
    return ("sell16", 0) if sell16_fits(rp, col) else ("sell32", 0)


def sell16_fits(rp, col):
    """Whether the library stores the sliced-ELL columns in 16 bits: in every slice of 64 rows,
    the offsets col - row of each slot (the k-th entry of its rows, padding = own row) span
    less than 2^16 (a base per slot, then 16-bit offsets from it)."""
    import numpy as np
    n = len(rp) - 1
    lens = np.diff(rp)
    wmax = int(lens.max())
    ns = (n + 63) // 64
    rows = np.repeat(np.arange(n), lens)
    k = np.arange(len(col)) - np.repeat(rp[:-1], lens)
    off = np.zeros((ns * 64, wmax), dtype=np.int64)
    off[rows, k] = col.astype(np.int64) - rows
    if ns * 64 > n:  # rows past the end are never read: repeat the last row's offsets
        off[n:] = off[n - 1]
    off = off.reshape(ns, 64, wmax)
    return bool((off.max(axis=1) - off.min(axis=1)).max() <= 65535)


def matrix_bytes(n, nnz, w, fmt, diag=True):
    """Compulsory bytes of one pass over the matrix: DIA values (D x n, offsets are kernel
    arguments; (D/2 + 1) x n for symmetric DIA), or sliced-ELL values, column offsets (2 B,
    with their slot bases) or indices (4 B) and slice offsets.  diag=False: the FP32 difference-form SpMV of a
    zero-row-sum matrix, which does not read the diagonal's values (one DIA array fewer)."""
    kind, d = fmt
    if kind == "dia":
        return w * (d - (0 if diag else 1)) * n
    if kind == "dia_sym":  # the diagonal and the upper diagonals
        return w * (d // 2 + (1 if diag else 0)) * n
    if kind == "sell16":  # 16-bit offsets and a 4-byte base per slice and slot
        return (w + 2) * nnz + 4 * (nnz // 64) + 4 * (n // 64 + 1)
    return (w + 4) * nnz + 4 * (n // 64 + 1)


def spmv_bytes(n, nnz, w, fmt):
    """Algorithmic bytes of one bicgstab_spmv_v launch: the matrix once, the gathered p once,
    r read, v and s = r - alpha v written."""
    return matrix_bytes(n, nnz, w, fmt) + 4 * w * n


def iteration_bytes(n, nnz, w, fmt, diag=True):
    """Compulsory bytes of one three-kernel iteration: two matrix passes and 16 vector streams
    (V: p, r, v, s; T: s, r0, t; X: x twice, p, s, t, v, rhat, r, p_next)."""
    return 2 * matrix_bytes(n, nnz, w, fmt, diag) + 16 * w * n


def build_case(be, P, cases, nx):
    data = P.box(nx, nx, nx, jitter=0.2, seed=1)
    c = cases.advection_diffusion(be, data, scheme="MUSCL", kappa=1e-3)
    c["eqn"].update_coeffs()  # first-outer-iteration system (device assembly)
    return c


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
    return time.perf_counter() - t0, iters, res


def kernel_time(be, fn, name, batches=1):
    """(summed event time in ms, launches) of kernel `name` during fn(), and fn's result; with
    batches > 1, fn runs that many times and the batch with the median time is reported (a
    single launch can run several times slower on a busy box)."""
    runs = []
    for _ in range(batches):
        be.set_timing(True)
        be.reset_timing()
        res = fn()
        st = be.kernel_stats()
        be.set_timing(False)
        k = st.get(name, (0, float("nan")))
        runs.append((k[1], k[0], res))
    runs.sort(key=lambda r: r[0])
    return runs[len(runs) // 2]


ITER_KERNELS = ("bicgstab_spmv_v", "bicgstab_spmv_t", "bicgstab_update_xrp")


def iteration_kernel_bytes(kernel, n, nnz, w, fmt, diag=True):
    """Algorithmic bytes of one working launch of an iteration kernel: V reads the matrix, the
    gathered p once and r, writes v and s; T reads the matrix, s and r0, writes t; X reads p,
    s, t, v, x and rhat and writes x, r and the next p (no preconditioner)."""
    M = matrix_bytes(n, nnz, w, fmt, diag)
    return {"bicgstab_spmv_v": M + 4 * w * n, "bicgstab_spmv_t": M + 3 * w * n,
            "bicgstab_update_xrp": 9 * w * n}[kernel]


def kernel_profile(be, fn):
    """fn() with kernel timing on: ({kernel: (launches, total ms)}, fn's result).  Each launch
    records its own start and end (the dispatch's timestamps, as a kernel trace reads them)."""
    be.set_timing(True)
    be.reset_timing()
    res = fn()
    st = be.kernel_stats()
    be.set_timing(False)
    return st, res


PROFILE_SOLVES = 3


def kernel_profile_median(be, fn, reps):
    """kernel_profile over `reps` calls of fn: per kernel, the (launches, total ms) of the call
    whose total is the median (one timed solve can run 10-20 % slow on a busy box), and the
    last call's result."""
    runs = [kernel_profile(be, fn) for _ in range(reps)]
    st = {}
    for k in runs[0][0]:
        vals = sorted((r[0].get(k, (0, 0.0)) for r in runs), key=lambda v: v[1])
        st[k] = vals[len(vals) // 2]
    return st, runs[-1][1]


def traffic_of(kernel, n):
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        tj = json.load(f)
    return tj.get("bytes_per_working_launch") if tj.get("n") == n else None


def pressure_system_case(be, P, cases, n, outer, tol):
    """The 128^3 (n^3) SMART cavity after `outer` FP64 SIMPLE outer iterations, then the next
    outer iteration's momentum predictor and its pressure-correction system (SIMPLE.cpp:19-178)."""
    c = cases.cavity(be, n=n)
    ctl = P.algo.SIMPLEControls(be, momentum_tol=tol, pressure_tol=tol, max_krylov_iters=MAX_KRYLOV)
    S = P.algo.IncompressibleSIMPLE(ctl)
    for _ in range(outer):
        S.step(c["eqns"], c["U"], c["mdot"], c["p"])
    pe, pp, ref, _ = S.pressure_system(c["eqns"], c["U"], c["mdot"], c["p"])
    return dict(case=c, controls=ctl, eqn=pe, pp=pp, ref=ref)


def pprime_solver(P, prec, tol):
    return P.solver.BiCGSTAB(tol=tol, precision=prec, max_iters=MAX_KRYLOV)


def solve_pprime(sys_, slv):
    """The step's p' solve from p' = 0: ISolver::solve with the step's settings (SIMPLE.cpp:171-178)."""
    sys_["pp"].fill(0.0)
    ctl = sys_["controls"]
    return slv.solve(sys_["eqn"], ctl.pressure_max_iter, ctl.pressure_residual, ref_cell=sys_["ref"])


def cavity_pressure_block(be, P, cases, n, outer, tol, steps, warmup):
    """FP32 + refinement vs FP64 on the pressure-correction solve of the n^3 SMART cavity's
    outer iteration outer+1: `steps` timed solves each (after `warmup`), per-kernel durations of
    the median of three more solves each, the FP32 solution against the FP64 one."""
    import numpy as np
    t0 = time.perf_counter()
    sys_ = pressure_system_case(be, P, cases, n, outer, tol)
    progress(f"pressure system of outer iteration {outer + 1} on the {n}^3 cavity built in "
             f"{time.perf_counter() - t0:.1f} s")
    eqn = sys_["eqn"]
    eqn.update_coeffs()
    rp, col, val, cells = eqn.matrix()
    eqn.zero_out_coeffs()
    fmt, nnz = matrix_format(rp, col, val), len(col)
    out = {"cells": cells, "nnz": nnz, "matrix_format": fmt[0], "reference_cell": sys_["ref"]}
    sol, t, res = {}, {}, {}
    for prec in ("fp32", "fp64"):
        slv = pprime_solver(P, prec, tol)
        for _ in range(warmup):
            solve_pprime(sys_, slv)
        be.synchronize()
        it = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            r = solve_pprime(sys_, slv)
            it += r.krylov_iters
        be.synchronize()
        dt = time.perf_counter() - t0
        t[prec], res[prec] = dt, r
        sol[prec] = sys_["pp"].values
        w = 8 if prec == "fp64" else 4
        # the solve deflates the reference-cell system to one with zero row sums, whose FP32
        # difference-form SpMVs read no diagonal values
        diag = prec == "fp64" or sys_["ref"] < 0
        st, rr = kernel_profile_median(be, lambda: solve_pprime(sys_, slv), PROFILE_SOLVES)
        sfx = "_f64" if prec == "fp64" else "_f32"
        kern = {}
        for k in ITER_KERNELS:
            launches, tot_ms = st.get(k + sfx, (0, 0.0))
            b = iteration_kernel_bytes(k, cells, nnz, w, fmt, diag)
            kern[k + sfx] = {"launches": launches, "working_launches": rr.krylov_iters, "total_ms": round(tot_ms, 4),
                             "avg_us_all_launches": round(tot_ms / max(1, launches) * 1e3, 3),
                             "algorithmic_bytes_per_working_launch": b,
                             "achieved_gbs": round(b * rr.krylov_iters / (tot_ms * 1e-3) / 1e9, 1) if tot_ms else None}
        out[prec] = {"ms_per_solve": round(dt / steps * 1e3, 3), "iters": it // steps,
                     "refinements": r.refinements, "restarts": r.restarts, "rel_residual": r.krylov_rel_residual,
                     "scaled_residual": r.final_residual, "converged": r.converged,
                     "mcell_iters_per_s": round(cells * it / dt / 1e6, 1),
                     "hbm_gbs_iteration_model": round(cells * it / dt * iteration_bytes(cells, nnz, w, fmt, diag) / cells / 1e9,
                                                      1),
                     "kernels": kern,
                     "iteration_kernels_ms": round(sum(v["total_ms"] for v in kern.values()), 3)}
    out["speedup_time_to_solution"] = round(t["fp64"] / t["fp32"], 3)
    d = sol["fp32"] - sol["fp64"]
    out["fp32_vs_fp64_solution"] = {"l2_rel": float(np.linalg.norm(d) / np.linalg.norm(sol["fp64"])),
                                    "linf_rel": float(np.abs(d).max() / np.abs(sol["fp64"]).max())}
    progress(f"{n}^3 p' solve: fp32 {out['fp32']['ms_per_solve']} ms ({out['fp32']['iters']} it), "
             f"fp64 {out['fp64']['ms_per_solve']} ms ({out['fp64']['iters']} it)")
    return out


def cavity_pressure_others(be, P, cases, n, outers, tol, reps=2):
    """The same FP32 / FP64 comparison on the pressure systems of other outer iterations (each
    from a fresh cavity): BiCGSTAB's iteration count moves by +-10-20 % between right-hand
    sides, so one system is a single draw."""
    runs = []
    for outer in outers:
        sys_ = pressure_system_case(be, P, cases, n, outer, tol)
        run = {"outer_iteration": outer + 1}
        for prec in ("fp32", "fp64"):
            slv = pprime_solver(P, prec, tol)
            solve_pprime(sys_, slv)
            be.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                r = solve_pprime(sys_, slv)
            be.synchronize()
            run[prec + "_ms"] = round((time.perf_counter() - t0) / reps * 1e3, 3)
            run[prec + "_iters"] = r.krylov_iters
            run[prec + "_rel_residual"] = r.krylov_rel_residual
        run["speedup"] = round(run["fp64_ms"] / run["fp32_ms"], 3)
        progress(f"{n}^3 p' solve of outer iteration {outer + 1}: speedup {run['speedup']} "
                 f"({run['fp32_iters']} / {run['fp64_iters']} iterations)")
        runs.append(run)
    return {"systems": runs,
            "speedup_all": round(sum(r["fp64_ms"] for r in runs) / sum(r["fp32_ms"] for r in runs), 3),
            "iters_ratio_all": round(sum(r["fp32_iters"] for r in runs) / max(1, sum(r["fp64_iters"] for r in runs)), 3)}


PISO_N_MAX = 96
MAX_KRYLOV = 10000  # iteration cap of every BiCGSTAB step in the SIMPLE / PISO blocks (a bound, not a target)


def progress(msg):
    """One line on stderr per finished step of a long block (the JSON line stays on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _field_errors(a, b):
    import numpy as np
    nb = np.linalg.norm(b)
    mb = np.abs(b).max()
    return {"l2_rel": float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)),
            "linf_rel": float(np.abs(a - b).max() / (mb if mb > 0 else 1.0))}


def _pair_run(be, P, make, step, steps, label, monitor=None):
    """The same `steps` outer steps on two instances of a case, one with the pressure
    correction in FP64 and one in FP32 + refinement, side by side from the same start; per
    step the time of each and the FP32 run's p and U against the FP64 run's.  monitor(c), if
    given, runs before each step (untimed) and its value is recorded per precision."""
    runs = {}
    for prec in ("fp64", "fp32"):
        t0 = time.perf_counter()
        runs[prec] = make(prec)
        progress(f"{label} case ({prec} pressure) built in {time.perf_counter() - t0:.1f} s")
    rows = []
    tot = {"fp64": 0.0, "fp32": 0.0}
    for k in range(steps):
        row = {label: k + 1}
        for prec, (c, ctl) in runs.items():
            if monitor is not None:
                row[prec + "_monitor"] = monitor(c)
            be.synchronize()
            t0 = time.perf_counter()
            rep = step(c, ctl)
            be.synchronize()
            dt = time.perf_counter() - t0
            tot[prec] += dt
            row[prec + "_ms"] = round(dt * 1e3, 2)
            row[prec + "_p_krylov_iters"] = sum(r["krylov_iters"] for r in rep if r["field_name"].startswith("P"))
            row[prec + "_max_krylov_iters"] = max((r["krylov_iters"] for r in rep), default=0)
            progress(f"{label} {k + 1} ({prec} pressure): {row[prec + '_ms']} ms, "
                     f"largest BiCGSTAB step {row[prec + '_max_krylov_iters']} iterations")
        c64, c32 = runs["fp64"][0], runs["fp32"][0]
        row["p"] = _field_errors(c32["p"].values, c64["p"].values)
        row["U"] = _field_errors(c32["U"].values, c64["U"].values)
        rows.append(row)
        progress(f"{label} {k + 1}: fp64 {row['fp64_ms']} ms, fp32 {row['fp32_ms']} ms, "
                 f"p diff {row['p']['l2_rel']:.2e}")
    return rows, tot


def simple_block(be, P, cases, n, outer, tol):
    """SIMPLE lid-driven cubic cavity (n^3, Re 100: nu 0.01, lid speed 1, SMART momentum,
    relaxation 0.7 / 0.3, automatic reference cell): `outer` outer iterations with the
    pressure correction in FP64 and in FP32 + refinement (Krylov tolerance `tol` in both,
    momentum in FP64), field error of the FP32 run against the FP64 run per outer iteration."""
    def make(prec):
        c = cases.cavity(be, n=n)
        ctl = P.algo.SIMPLEControls(be, momentum_tol=tol, pressure_tol=tol, max_krylov_iters=MAX_KRYLOV,
                                    pressure_precision=0 if prec == "fp64" else 1)
        return c, ctl

    def step(c, ctl):
        return P.algo.IncompressibleSIMPLE(ctl).step(c["eqns"], c["U"], c["mdot"], c["p"])

    rows, tot = _pair_run(be, P, make, step, outer, "outer")
    return {"workload": f"SIMPLE lid-driven cavity {n}^3, Re 100, SMART, urf 0.7/0.3, p' to {tol:g}",
            "cells": n ** 3, "outer_iterations": rows,
            "speedup_total": round(tot["fp64"] / tot["fp32"], 3)}


def piso_block(be, P, cases, n, steps, tol):
    """Transient PISO on the [0, 2 pi]^3 cube split into 6 tetrahedra per hex (n^3 x 6
    cells): Taylor-Green vortex plus seeded divergence-free noise (RMS 1e-3), nu 1e-3,
    backward Euler with dt putting the initial maximum Courant number at 0.5, LinearUpwind,
    2 pressure correctors; `steps` time steps with the pressure correction in FP64 and in
    FP32 + refinement side by side.  Before each step the maximum cell Courant number of
    mdot is recorded, as examples/PISOSolver/main.cpp:66 logs it."""
    if n > PISO_N_MAX:
        raise ValueError(f"--piso-n {n}: meshes above {PISO_N_MAX}^3 x 6 cells are not measured")

    def make(prec):
        c = cases.taylor_green(be, n=n, tets=True)
        ctl = P.algo.PISOControls(be, outer_iterations=1, pressure_correction_steps=2,
                                  simple=dict(momentum_tol=tol, pressure_tol=tol, max_krylov_iters=MAX_KRYLOV,
                                              pressure_precision=0 if prec == "fp64" else 1))
        return c, ctl

    def step(c, ctl):
        rep = P.algo.PISO(ctl).step(c["eqns"], c["U"], c["mdot"], c["p"])
        for comp in (c["U"].x, c["U"].y, c["U"].z):
            comp.update_prev_time_steps()
        return rep

    dts = []

    def courant(c):
        dts.append(c["dt"])
        return {"co_max": round(float(P.courant_number(c["mdot"], c["dt"]).max()), 6)}

    rows, tot = _pair_run(be, P, make, step, steps, "step", monitor=courant)
    return {"workload": f"PISO Taylor-Green on {n}^3 x 6 tetrahedra, 2 correctors, p' to {tol:g}",
            "cells": 6 * n ** 3, "dt": dts[0] if dts else None, "steps": rows,
            "speedup_total": round(tot["fp64"] / tot["fp32"], 3)}


def converge_primal(c, slv, corrected, rel=1e-12, max_outer=200):
    """The primal's converged state: three FP64 steps; with a Corrected diffusion term, then
    further solves (each assembling the correction from the latest T) until T moves by less
    than `rel` relative.  Returns (last result, outer solves)."""
    import numpy as np
    r = slv.solve(c["eqn"], 3, 1e-30)
    outer = 1
    while corrected and outer < max_outer:
        prev = c["T"].values
        r = slv.solve(c["eqn"], 3, 1e-30)
        outer += 1
        if np.linalg.norm(c["T"].values - prev) <= rel * np.linalg.norm(prev):
            break
    return r, outer


def adjoint_block(be, P, cases, n, ndir, tol=1e-13, corrector=None):
    """Discrete adjoint of steady UPWIND advection-diffusion on the n^3 jittered mesh
    (log-normal kappa, source s ~ U(0, 1)): J, dJ/ds and dJ/dkappa from one A^T solve (with a
    Corrected diffusion term: A^T solves in a defect correction for the correction's terms), checked
    against central differences of the FP64 primal along `ndir` seeded random positive
    directions (half in s, half in kappa, the latter relative to kappa).  J is quadratic in s
    (T is linear in s), so the central difference in s is exact for any step and a large one
    (1e-3) keeps the primal's solver noise (~1e-11 relative in J) out of it; in kappa the step
    is 1e-4 (truncation ~1e-8 relative)."""
    import numpy as np
    c = cases.adjoint_transport(be, P.box(n, n, n, jitter=0.2, seed=5), corrector=corrector)
    slv = P.solver.BiCGSTAB(tol=tol, max_iters=20000)
    # three FP64 steps: each restarts BiCGSTAB from the current solution (refinement), so the
    # primal is converged far enough for a 1e-6 central difference of J to be meaningful
    t0 = time.perf_counter()
    primal, outer = converge_primal(c, slv, corrector is not None)
    be.synchronize()
    t_primal = time.perf_counter() - t0
    t0 = time.perf_counter()
    J, ds, dk, res = P.adjoint.gradient(c["eqn"], c["s"], c["kappa"], 0, 0.7, solver=slv)
    t_adj = time.perf_counter() - t0
    rng = np.random.default_rng(8)
    worst = 0.0
    eps_s, eps_k = 1e-3, 1e-4
    errs = []
    for k in range(ndir):
        field, grad = (c["s"], ds) if k % 2 == 0 else (c["kappa"], dk)
        eps = eps_s if k % 2 == 0 else eps_k
        base = field.values
        # positive directions: a zero-mean one makes dJ.v a sum of 2M cancelling terms, and the
        # central difference of J then measures the primal's round-off, not the derivative
        v = rng.uniform(0.0, 1.0, len(base)) * (1.0 if k % 2 == 0 else base)
        Js = []
        for sgn in (1.0, -1.0):
            field.values = base + sgn * eps * v
            converge_primal(c, slv, corrector is not None)
            Js.append(P.adjoint.objective(c["eqn"], 0, 0.7))
        field.values = base
        fd = (Js[0] - Js[1]) / (2 * eps)
        e = abs(float(grad @ v) - fd) / abs(fd)
        errs.append(round(e, 12))
        worst = max(worst, e)
    diff = "non-corrected" if corrector is None else f"{corrector} Corrected (least-squares gradients)"
    return {"workload": f"adjoint of UPWIND advection-diffusion {n}^3 jittered, log-normal kappa, {diff} "
                        "diffusion, J = mean T^2 over x > 0.7", "cells": n ** 3, "J": J,
            "primal_ms": round(t_primal * 1e3, 3), "primal_outer_solves": outer,
            "adjoint_solve_ms": round(t_adj * 1e3, 3), "adjoint_krylov_iters": res.krylov_iters,
            "adjoint_transpose_solves": res.iteration,
            "adjoint_converged": bool(res.converged), "adjoint_rel_residual": res.krylov_rel_residual,
            "primal_rel_residual": primal.krylov_rel_residual, "fd_directions": ndir, "fd_eps_s": eps_s, "fd_eps_kappa_rel": eps_k,
            "max_rel_error_vs_fd": worst, "rel_errors_s_kappa_alternating": errs}


def export_bytes(n_cells, n_vertices, n_adjacency, ncomp):
    """Algorithmic bytes of one cell -> point launch: vertex offsets, the adjacency's cell ids
    (int32) and weights (FP64), every cell value read once, every point value written once."""
    return 4 * (n_vertices + 1) + 12 * n_adjacency + 8 * ncomp * (n_cells + n_vertices)


def export_block(be, P, n, reps=10):
    """Output export: the device interpolation of a vector field to the vertices of an n^3
    jittered hex mesh (k_cell_to_point<3>, kernel timing), and the wall time of one compressed
    VTU write of that field (host compression and file I/O included)."""
    import numpy as np
    from prism_amd import output as O
    progress(f"export {n}^3")
    data = P.box(n, n, n, jitter=0.2, seed=1)
    m = P.Mesh(be, data, {p: {"U": ("fixed", (0.0, 0.0, 0.0))} for p in data.patch_names}, [("U", "vector")])
    U = m.vector("U", values=np.random.default_rng(0).standard_normal((m.n_cells, 3)))
    O.cell_to_point(U)  # first call uploads the adjacency and its weights
    tot_ms, launches, _ = kernel_time(be, lambda: [O.cell_to_point(U) for _ in range(reps)], "k_cell_to_point<3>", 3)
    avg_ms = tot_ms / launches
    nv, nadj = len(data.vertices), len(data.cell_vtx)
    b = export_bytes(m.n_cells, nv, nadj, 3)
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"bench_export_{os.getpid()}.vtu")
    be.synchronize()
    t0 = time.perf_counter()
    O.to_vtu(U, path)
    t_vtu = time.perf_counter() - t0
    size = os.path.getsize(path)
    os.remove(path)
    return {"workload": f"vector cell -> point, {n}^3 jittered hex ({nv} vertices)", "kernel": "k_cell_to_point<3>",
            "avg_us": round(avg_ms * 1e3, 2), "launches": launches, "algorithmic_bytes": b,
            "achieved_gbs": round(b / (avg_ms * 1e-3) / 1e9, 1),
            "frac": round(b / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "vtu_write_s": round(t_vtu, 3), "vtu_bytes": size, "vtu_cells": m.n_cells}


def kepsilon_block(be, P, n, reps=10):
    """k-epsilon coefficient update (k_kepsilon: nu_t, Pk, effective viscosities, epsilon
    source and sink in one pass) on an n^3 jittered hex mesh with a sheared velocity field,
    after the three least-squares gradients of U (HIP events per launch); then the gradients
    of U by the update's one three-component pass and by three single-field passes."""
    import numpy as np
    from prism_amd.turbulence import IncompressibleKEpsilon
    progress(f"k-epsilon {n}^3")
    data = P.box(n, n, n, jitter=0.2, seed=1)
    bcs = {p: {"U": ("fixed", (0.0, 0.0, 0.0)), "k": "zero-gradient", "epsilon": "zero-gradient",
               "nu": "zero-gradient"} for p in data.patch_names}
    m = P.Mesh(be, data, bcs, [("U", "vector"), ("k", "scalar"), ("epsilon", "scalar"), ("nu", "scalar")])
    g = m.geometry()["cell_center"]
    U = m.vector("U", values=np.stack([g[:, 1] ** 2, np.sin(g[:, 0]), g[:, 0] * g[:, 2]], axis=1))
    model = IncompressibleKEpsilon(U, U, m.scalar("nu", 1e-3))
    model.k().fill(0.01)
    model.epsilon().fill(1e-3)
    model.update_fields()
    tot_ms, launches, _ = kernel_time(be, lambda: [model.update_fields() for _ in range(reps)], "k_kepsilon", 3)
    avg_ms = tot_ms / launches
    b = 144 * m.n_cells  # k, eps, nu, grad U (9) read; nu_t, Pk, nu_eff_k, nu_eff_eps, source, sink written
    # grad U after every write of U: one three-component pass (the update's) against one
    # single-field pass per component
    comps = (U.x, U.y, U.z)
    cols = [c.values for c in comps]

    def rewritten(fn):
        for _ in range(reps):
            for c, v in zip(comps, cols):
                c.values = v
            fn()

    g3_ms, g3_n, _ = kernel_time(be, lambda: rewritten(model.update_fields), "k_grad_ls3")
    g1_ms, g1_n, _ = kernel_time(be, lambda: rewritten(lambda: [c.grad() for c in comps]), "k_grad_ls")
    return {"workload": f"k-epsilon coefficient update, {n}^3 jittered hex", "kernel": "k_kepsilon",
            "avg_us": round(avg_ms * 1e3, 2), "launches": launches, "algorithmic_bytes": b,
            "achieved_gbs": round(b / (avg_ms * 1e-3) / 1e9, 1),
            "frac": round(b / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "grad_U_us": {"k_grad_ls3_one_pass": round(g3_ms / g3_n * 1e3, 2),
                          "k_grad_ls_three_passes": round(3 * g1_ms / g1_n * 1e3, 2),
                          "launches": [g3_n, g1_n]}}


def cpu_baseline(rp, col, val, b, label, budget_s=15.0):
    """FP64 CPU restatement (oracle, one thread): the same unpreconditioned BiCGSTAB iteration
    on the same system (the headline's deflated p' system, copied from the device), a bounded
    number of iterations."""
    import numpy as np

    import oracle
    import prism_amd as P

    be = oracle.backend()
    n = len(b)
    A = P.Matrix(be, rp, col, val)
    slv = P.solver.BiCGSTAB(tol=1e-30, max_iters=5)  # calibrate
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
    out = {"value": round(n * it / dt / 1e6, 3), "unit": "Mcell-iters/s", "cores": 1, "kind": "port",
           "sample": f"{it} unpreconditioned FP64 BiCGSTAB iterations on {label} "
                     f"({n} cells, {len(val)} nnz) in {dt:.2f} s, single thread, -O2",
           "cpu": platform.processor() or platform.machine(), "host_cores_available": os.cpu_count()}
    try:
        out["all_cores"] = cpu_baseline_omp(rp, col, val, b, label, budget_s / 2)
    except Exception as e:
        out["all_cores"] = {"value": None, "error": str(e)}
    return out


def cpu_baseline_omp(rp, col, val, b, label, budget_s):
    """The same iteration on every host core this process may use (OpenMP, static schedule,
    capped by OMP_NUM_THREADS), sized from a 200-iteration calibration run."""
    import oracle

    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    threads = max(1, threads)
    oracle.omp_bicgstab(rp, col, val, b, 50, threads)  # warm-up: thread pool, first touch
    _, it0, dt0 = oracle.omp_bicgstab(rp, col, val, b, 200, threads)  # calibration
    k = int(max(10, min(100000, budget_s * it0 / max(dt0, 1e-9))))
    _, it, dt = oracle.omp_bicgstab(rp, col, val, b, k, threads)
    n = len(b)
    return {"value": round(n * it / dt / 1e6, 3), "unit": "Mcell-iters/s", "cores": threads, "kind": "port",
            "sample": f"{it} unpreconditioned FP64 BiCGSTAB iterations on {label} in {dt:.2f} s, "
                      f"{threads} OpenMP threads, -O2"}


def deflated_system(eqn, ref):
    """(rp, col, val, b') of the assembled p' system with the reference-cell deflation the solver
    applies (A symmetric with zero row sums: b' = b - (1^T b) e_ref, a compatible system)."""
    eqn.update_coeffs()
    rp, col, val, _ = eqn.matrix()
    b = eqn.rhs()
    eqn.zero_out_coeffs()
    if ref >= 0:
        b[ref] -= b.sum()
    return rp, col, val, b


def pmc_workload(args):
    """The program profiled by the --pmc passes: one FP32 p' solve of the headline workload,
    then a 256 MiB device copy of known size for calibrating FETCH_SIZE / WRITE_SIZE."""
    import torch

    import prism_amd as P
    from prism_amd import cases

    be = P.Backend.hip(0)
    sys_ = pressure_system_case(be, P, cases, args.n, args.outer, args.tol)
    r = solve_pprime(sys_, pprime_solver(P, "fp32", args.tol))
    be.synchronize()
    a = torch.ones(CALIB_FLOATS, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"pmc_workload": "fp32 p' solve", "n": args.n ** 3, "krylov_iters": r.krylov_iters,
                      "calibration_copy_bytes": CALIB_FLOATS * 4}))


def pmc_piso(args):
    """The program profiled by the tetrahedral --pmc passes: the first pressure-correction
    system of the PISO Taylor-Green step on --piso-n^3 x 6 tetrahedra (the PISO block's
    settings), solved once in FP32 + refinement, then the calibration copy."""
    import torch

    import prism_amd as P
    from prism_amd import cases

    be = P.Backend.hip(0)
    c = cases.taylor_green(be, n=args.piso_n, tets=True)
    pc = P.algo.PISOControls(be, outer_iterations=1, pressure_correction_steps=2,
                             simple=dict(momentum_tol=args.tol, pressure_tol=args.tol, max_krylov_iters=MAX_KRYLOV,
                                         pressure_precision=1))
    pe, pp, ref, _ = P.algo.IncompressibleSIMPLE(pc.simple).pressure_system(c["eqns"], c["U"], c["mdot"], c["p"])
    slv = P.solver.BiCGSTAB(tol=args.tol, precision="fp32", max_iters=MAX_KRYLOV)
    r = slv.solve(pe, pc.simple.pressure_max_iter, pc.simple.pressure_residual, ref_cell=ref)
    be.synchronize()
    a = torch.ones(CALIB_FLOATS, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"pmc_workload": "fp32 p' solve, PISO tetrahedra", "n": 6 * args.piso_n ** 3,
                      "krylov_iters": r.krylov_iters, "calibration_copy_bytes": CALIB_FLOATS * 4}))


def pmc_assembly(args):
    """The program profiled by the assembly --pmc passes: one assembly pass of the 64^3 MUSCL
    advection-diffusion system (least-squares gradient of T, then the assembly kernels), then
    the 256 MiB calibration copy."""
    import torch

    import prism_amd as P
    from prism_amd import cases

    be = P.Backend.hip(0)
    c = build_case(be, P, cases, args.advection_n)
    be.synchronize()
    a = torch.ones(CALIB_FLOATS, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    m = c["mesh"]
    print(json.dumps({"pmc_assembly": f"{args.advection_n}^3 MUSCL", "cells": m.n_cells, "faces": m.n_faces,
                      "nnz": m.nnz, "calibration_copy_bytes": CALIB_FLOATS * 4}))


def read_pmc(path, counter):
    rows = list(csv.DictReader(open(path)))
    per = {}
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        per.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return per


def write_traffic(fetch_csv, write_csv, n, kernel, match, working):
    """profiles/traffic_<kernel>.json from the two --pmc passes: the kernel's FETCH_SIZE and
    WRITE_SIZE summed over its dispatches (KiB), each scaled by the same counter's reading of
    the calibration copy, divided by the working launches of the profiled solve."""
    f = read_pmc(fetch_csv, "FETCH_SIZE")
    w = read_pmc(write_csv, "WRITE_SIZE")

    def total(d):
        vals = [v for k, vs in d.items() if match in k for v in vs]
        return (sum(vals), len(vals)) if vals else (None, 0)
    kf, nf = total(f)
    kw, nw = total(w)
    cal_f = max((max(vs) for k, vs in f.items() if "elementwise" in k), default=None)
    cal_w = max((max(vs) for k, vs in w.items() if "elementwise" in k), default=None)
    known = CALIB_FLOATS * 4
    ff = known / (cal_f * 1024) if cal_f else None
    wf = known / (cal_w * 1024) if cal_w else None
    ok = kf is not None and kw is not None and ff and wf
    out = {"kernel": kernel, "match": match, "n": n, "dispatches": [nf, nw], "working_launches": working,
           "fetch_kib_total": kf, "write_kib_total": kw,
           "calibration_copy_bytes": known, "calibration_fetch_kib": cal_f, "calibration_write_kib": cal_w,
           "fetch_factor": ff, "write_factor": wf,
           "bytes_per_working_launch": (kf * 1024 * ff + kw * 1024 * wf) / working if ok else None,
           "sources": [os.path.basename(fetch_csv), os.path.basename(write_csv)]}
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


def advection_block(be, P, cases, nx, steps, warmup, tol):
    """The 64^3 jittered MUSCL advection-diffusion system of the first outer iteration (limiter
    ratios 0/0: upwind weights; convection.cpp:69-128), FP32 + refinement vs FP64.  Its working
    set (about 20 MB per iteration) stays in the 256 MB last-level cache (MALL), so the rates
    here are not HBM rates and the iteration is launch-latency bound."""
    import numpy as np
    c = build_case(be, P, cases, nx)
    eqn, T = c["eqn"], c["T"]
    mesh = c["mesh"]
    n, nnz = mesh.n_cells, mesh.nnz
    rp, col, _, _ = eqn.matrix()
    fmt = matrix_format(rp, col)
    s32 = P.solver.BiCGSTAB(tol=tol, precision="fp32")
    s64 = P.solver.BiCGSTAB(tol=tol, precision="fp64")
    t32, it32, r32 = timed_solves(be, eqn, T, s32, steps, warmup)
    x32 = T.values
    t64, it64, r64 = timed_solves(be, eqn, T, s64, steps, warmup)
    x64 = T.values
    err = {"l2_rel": float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64)),
           "linf_rel": float(np.abs(x32 - x64).max() / np.abs(x64).max())}
    asm_reps = 20
    be.synchronize()
    t0 = time.perf_counter()
    for _ in range(asm_reps):
        T.fill(0.0)
        eqn.zero_out_coeffs()
        eqn.update_coeffs()
    be.synchronize()
    t_asm = (time.perf_counter() - t0) / asm_reps
    T.fill(0.0)
    eqn.zero_out_coeffs()
    eqn.update_coeffs()
    ab = assembly_bytes(n, mesh.n_faces, nnz)
    return {"workload": f"advection-diffusion {nx}^3 jittered hex, MUSCL (first outer iteration), kappa 1e-3, "
                        f"BiCGSTAB to {tol:g}; MALL-resident (working set below the 256 MB last-level cache)",
            "cells": n, "nnz": nnz, "matrix_format": fmt[0],
            "fp32": {"mcell_iters_per_s": round(n * it32 / t32 / 1e6, 1), "ms_per_solve": round(t32 / steps * 1e3, 4),
                     "iters_per_solve": it32 / steps, "refinements": r32.refinements,
                     "rel_residual": r32.krylov_rel_residual},
            "fp64": {"mcell_iters_per_s": round(n * it64 / t64 / 1e6, 1), "ms_per_solve": round(t64 / steps * 1e3, 4),
                     "iters_per_solve": it64 / steps, "rel_residual": r64.krylov_rel_residual},
            "speedup_time_to_solution": round(t64 / t32, 3), "fp32_vs_fp64_solution": err,
            "assembly": {"faces_per_s": round(mesh.n_faces / t_asm, 1), "ms_per_pass": round(t_asm * 1e3, 4),
                         "bytes_per_pass_model": ab, "hbm_gbs_model": round(ab / t_asm / 1e9, 1),
                         "hbm_frac_model": round(ab / t_asm / 1e9 / HBM_PEAK_GBS, 4)}}


def sgs_block(be, P, cases, nx, tol, max_iters=20000):
    """Multicolour symmetric Gauss-Seidel preconditioned BiCGSTAB (FP64) against
    unpreconditioned FP64 BiCGSTAB on nx^3 jittered advection-diffusion systems at kappa 1e-3
    of the first outer iteration (T = 0): MUSCL (zero gradients: upwind weights) and QUICK.
    One timed solve each after a warm-up solve (which builds the colouring), and the
    per-kernel times of one SGS solve of the MUSCL system."""
    out = {"workload": f"advection-diffusion {nx}^3 jittered hex, kappa 1e-3, first outer iteration, "
                       f"BiCGSTAB to {tol:g}, at most {max_iters} iterations"}
    data = P.box(nx, nx, nx, jitter=0.2, seed=1)
    for scheme in ("MUSCL", "QUICK"):
        c = cases.advection_diffusion(be, data, scheme=scheme, kappa=1e-3)
        eqn, T = c["eqn"], c["T"]
        eqn.update_coeffs()
        out["cells"] = c["mesh"].n_cells
        res = {}
        for pc in (None, "sgs"):
            slv = P.solver.BiCGSTAB(tol=tol, precond=pc, max_iters=max_iters)
            T.fill(0.0)
            slv.step_eqn(eqn)
            be.synchronize()
            T.fill(0.0)
            t0 = time.perf_counter()
            r = slv.step_eqn(eqn)
            be.synchronize()
            dt = time.perf_counter() - t0
            res[pc or "none"] = {"ms_per_solve": round(dt * 1e3, 3), "iters": r.krylov_iters,
                                 "us_per_iter": round(dt * 1e6 / max(1, r.krylov_iters), 2),
                                 "rel_residual": r.krylov_rel_residual, "converged": bool(r.converged)}
        if scheme == "MUSCL":
            st, _ = kernel_profile(be, lambda: (T.fill(0.0), slv.step_eqn(eqn))[1])
            res["sgs_kernels"] = {k: {"launches": v[0], "total_ms": round(v[1], 4)}
                                  for k, v in st.items() if k.startswith("sgs_")}
        res["speedup_time_to_solution"] = round(res["none"]["ms_per_solve"] / res["sgs"]["ms_per_solve"], 3)
        out[scheme] = res
        progress(f"sgs {nx}^3 {scheme}: none {res['none']['iters']} it {res['none']['ms_per_solve']} ms, "
                 f"sgs {res['sgs']['iters']} it {res['sgs']['ms_per_solve']} ms")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128, help="cells per axis of the headline cavity")
    ap.add_argument("--outer", type=int, default=5, help="FP64 SIMPLE outer iterations before the measured p' system")
    ap.add_argument("--other-outers", type=str, default="1,3,8",
                    help="outer iterations before the further p' systems compared ('' to skip)")
    ap.add_argument("--advection-n", type=int, default=64, help="cells per axis of the advection system (0: skip)")
    ap.add_argument("--adjoint-n", type=int, default=128, help="cells per axis of the adjoint check (0: skip)")
    ap.add_argument("--adjoint-dirs", type=int, default=32)
    ap.add_argument("--simple-n", type=int, default=128, help="cells per axis of the SIMPLE cavity (0: skip)")
    ap.add_argument("--simple-outer", type=int, default=5)
    ap.add_argument("--piso-n", type=int, default=96, help="hexes per axis of the PISO tet mesh (0: skip; at most 96)")
    ap.add_argument("--piso-steps", type=int, default=2)
    ap.add_argument("--export-n", type=int, default=128, help="cells per axis of the export measurement (0: skip)")
    ap.add_argument("--kepsilon-n", type=int, default=128, help="cells per axis of the k-epsilon update (0: skip)")
    ap.add_argument("--sgs-n", type=int, default=64,
                    help="cells per axis of the SGS-preconditioned advection solves (0: skip)")
    ap.add_argument("--sgs-only", action="store_true", help="run only the SGS block and print its JSON")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--pmc-workload", action="store_true")
    ap.add_argument("--pmc-piso", action="store_true",
                    help="one FP32 solve of the first PISO pressure system on --piso-n^3 x 6 tetrahedra and a "
                         "calibration copy (for --pmc passes)")
    ap.add_argument("--cells", type=int, default=0, help="--traffic: rows of the profiled system (default n^3)")
    ap.add_argument("--pmc-assembly", action="store_true",
                    help="one assembly pass of the advection system and a calibration copy (for --pmc passes)")
    ap.add_argument("--traffic", nargs=5, metavar=("FETCH_CSV", "WRITE_CSV", "KERNEL_MATCH", "WORKING_LAUNCHES", "LABEL"),
                    help="e.g. fetch.csv write.csv 'k_iter_t3p<float' 606 bicgstab_spmv_t_f32")
    ap.add_argument("--headline-only", action="store_true",
                    help="run only the headline p' comparison (for profiling it) and print its JSON line")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the adjoint blocks (non-corrected and over-relaxed) and print their JSON")
    ap.add_argument("--piso-only", action="store_true",
                    help="run only the PISO block (for profiling it) and print its JSON")
    ap.add_argument("--simple-only", action="store_true",
                    help="run only the SIMPLE cavity block (for profiling it) and print its JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this path has no multi-GPU mode (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    if args.traffic:
        fetch, write, match, working, label = args.traffic
        write_traffic(fetch, write, args.cells or args.n ** 3, label, match, int(working))
        return
    if args.pmc_workload:
        pmc_workload(args)
        return
    if args.pmc_piso:
        pmc_piso(args)
        return
    if args.pmc_assembly:
        pmc_assembly(args)
        return

    import prism_amd as P
    from prism_amd import cases

    be = P.Backend.hip(0)
    if args.piso_only:
        print(json.dumps(piso_block(be, P, cases, args.piso_n, args.piso_steps, args.tol)))
        return
    if args.adjoint_only:
        res = {"adjoint": adjoint_block(be, P, cases, args.adjoint_n, args.adjoint_dirs),
               "adjoint_corrected": adjoint_block(be, P, cases, args.adjoint_n, 4, corrector="over-relaxed")}
        print(json.dumps(res))
        return
    if args.sgs_only:
        print(json.dumps(sgs_block(be, P, cases, args.sgs_n, args.tol)))
        return
    if args.simple_only:
        print(json.dumps(simple_block(be, P, cases, args.simple_n, args.simple_outer, args.tol)))
        return
    hp = cavity_pressure_block(be, P, cases, args.n, args.outer, args.tol, args.steps, args.warmup)
    cells, nnz = hp["cells"], hp["nnz"]
    f32, f64 = hp["fp32"], hp["fp64"]
    # roofline: the FP32 iteration kernel with the largest total time in one solve; achieved =
    # algorithmic bytes of its working launches / its total time over all launches (launches
    # issued past a device event return at once and are counted in the time)
    dom = max(f32["kernels"].items(), key=lambda kv: kv[1]["total_ms"])
    dk, dv = dom
    traffic = traffic_of(dk, cells)
    name, ncu = be.device_info()
    out = {
        "metric": METRIC,
        "value": f32["mcell_iters_per_s"],
        "unit": "Mcell-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f32["ms_per_solve"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 residual/dots/solution)",
        "data": "synthetic",
        "config": {"workload": f"SIMPLE pressure-correction solve, {args.n}^3 SMART lid-driven cavity (Re 100, "
                               f"urf 0.7/0.3), the p' system of outer iteration {args.outer + 1} after {args.outer} "
                               f"FP64 outer iterations; ISolver::solve with reference cell {hp['reference_cell']}, "
                               f"BiCGSTAB to {args.tol:g} relative residual",
                   "cells": cells, "nnz": nnz, "tol": args.tol, "matrix_format": hp["matrix_format"]},
        "speedup_time_to_solution": hp["speedup_time_to_solution"],
        "iters_per_solve": f32["iters"],
        "fp64_iters_per_solve": f64["iters"],
        "hbm_gbs_iteration_model": f32["hbm_gbs_iteration_model"],
        "rel_residual": f32["rel_residual"],
        "fp32_vs_fp64_solution": hp["fp32_vs_fp64_solution"],
        "fp64": {"value": f64["mcell_iters_per_s"], "ms_per_step": f64["ms_per_solve"], "iters_per_solve": f64["iters"],
                 "hbm_gbs_iteration_model": f64["hbm_gbs_iteration_model"], "rel_residual": f64["rel_residual"]},
        "roofline": {"bound": "hbm", "kernel": dk, "achieved": dv["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dv["achieved_gbs"] / HBM_PEAK_GBS, 4) if dv["achieved_gbs"] else None,
                     "traffic": traffic, "algorithmic_bytes_per_working_launch": dv["algorithmic_bytes_per_working_launch"],
                     "working_launches": dv["working_launches"], "launches": dv["launches"], "total_ms": dv["total_ms"],
                     "avg_us_all_launches": dv["avg_us_all_launches"]},
        "headline_detail": hp,
        "device": {"arch": name, "cus": ncu},
    }
    if args.headline_only:
        print(json.dumps(out))
        return
    if args.other_outers:
        out["pressure_other_outer_iterations"] = cavity_pressure_others(
            be, P, cases, args.n, [int(k) for k in args.other_outers.split(",")], args.tol)
    if args.advection_n > 0:
        progress(f"advection {args.advection_n}^3")
        out["advection_mall_resident"] = advection_block(be, P, cases, args.advection_n, args.steps, args.warmup,
                                                         args.tol)
    if args.sgs_n > 0:
        out["sgs_preconditioned"] = sgs_block(be, P, cases, args.sgs_n, args.tol)
    if args.adjoint_n > 0:
        out["adjoint"] = adjoint_block(be, P, cases, args.adjoint_n, args.adjoint_dirs)
        out["adjoint_corrected"] = adjoint_block(be, P, cases, args.adjoint_n, 4, corrector="over-relaxed")
    if args.simple_n > 0:
        out["simple"] = simple_block(be, P, cases, args.simple_n, args.simple_outer, args.tol)
    if args.piso_n > 0:
        out["piso"] = piso_block(be, P, cases, args.piso_n, args.piso_steps, args.tol)
    if args.export_n > 0:
        out["export"] = export_block(be, P, args.export_n)
    if args.kepsilon_n > 0:
        out["kepsilon"] = kepsilon_block(be, P, args.kepsilon_n)
    if not args.no_cpu:
        try:
            sys_ = pressure_system_case(be, P, cases, args.n, args.outer, args.tol)
            rp, col, val, b = deflated_system(sys_["eqn"], sys_["ref"])
            out["cpu_baseline"] = cpu_baseline(rp, col, val, b, f"the {args.n}^3 cavity's p' system")
        except Exception as e:  # the CPU leg never blocks the GPU figure
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
