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
system is timed the same way and reported beside it, and so is the pressure-correction
system of a SIMPLE step on the 128^3 cavity mesh (symmetric, all-Neumann, reference cell).

Other modes (no JSON line):
  --pmc-workload       one FP32 solve of the workload plus a calibration copy kernel
                       (the program profiled by the rocprofv3 --pmc passes);
  --traffic F.csv W.csv  turn the FETCH_SIZE and WRITE_SIZE pass outputs into
                       profiles/traffic_bicgstab_spmv_v_f32.json (no GPU needed).
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
    Jacobi), else ("sell16", 0) when every |col - row| < 2^15, else ("sell32", 0)."""
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
    return ("sell16", 0) if np.abs(off).max() < 32768 else ("sell32", 0)


def matrix_bytes(n, nnz, w, fmt):
    """Compulsory bytes of one pass over the matrix: DIA values (D x n, offsets are kernel
    arguments; (D/2 + 1) x n for symmetric DIA), or sliced-ELL values, column offsets (2 B)
    or indices (4 B) and slice offsets."""
    kind, d = fmt
    if kind == "dia":
        return w * d * n
    if kind == "dia_sym":  # the diagonal and the upper diagonals
        return w * (d // 2 + 1) * n
    return (w + (2 if kind == "sell16" else 4)) * nnz + 4 * (n // 64 + 1)


def spmv_bytes(n, nnz, w, fmt):
    """Algorithmic bytes of one bicgstab_spmv_v launch: the matrix once, the gathered p once,
    r read, v and s = r - alpha v written."""
    return matrix_bytes(n, nnz, w, fmt) + 4 * w * n


def iteration_bytes(n, nnz, w, fmt):
    """Compulsory bytes of one three-kernel iteration: two matrix passes and 16 vector streams
    (V: p, r, v, s; T: s, r0, t; X: x twice, p, s, t, v, rhat, r, p_next)."""
    return 2 * matrix_bytes(n, nnz, w, fmt) + 16 * w * n


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


def kernel_time(be, fn, name):
    """(summed event time in ms, launches) of kernel `name` during fn(), and fn's result."""
    be.set_timing(True)
    be.reset_timing()
    res = fn()
    st = be.kernel_stats()
    be.set_timing(False)
    k = st.get(name, (0, float("nan")))
    return k[1], k[0], res


def pressure_block(be, P, cases, n, tol, reps=2, extra_seeds=(6, 7, 8)):
    """FP64 vs FP32+refinement on the pressure-correction system of the n^3 cavity: one
    ISolver::solve step each (assembly, reference cell 0, BiCGSTAB to `tol`), p' reset to 0."""
    c = cases.pressure_correction(be, n=n)
    eqn, p = c["eqn"], c["p"]
    eqn.update_coeffs()
    rp, col, val, _ = eqn.matrix()
    eqn.zero_out_coeffs()
    fmt, nnz, cells = matrix_format(rp, col, val), len(col), n ** 3
    out = {"workload": f"pressure correction {n}^3 cavity mesh, all-Neumann, reference cell 0, "
                       f"one solve() step to {tol:g}", "cells": cells, "nnz": nnz, "matrix_format": fmt[0]}
    t, sol = {}, {}
    for prec in ("fp64", "fp32"):
        slv = P.solver.BiCGSTAB(tol=tol, precision=prec)
        def one():
            p.fill(0.0)
            return slv.solve(eqn, 1, 1e-30, ref_cell=0)
        one()
        be.synchronize()
        it = 0
        t0 = time.perf_counter()
        for _ in range(reps):
            r = one()
            it += r.krylov_iters
        be.synchronize()
        dt = time.perf_counter() - t0
        t[prec] = dt
        sol[prec] = p.values
        w = 8 if prec == "fp64" else 4
        label = "bicgstab_spmv_v_" + ("f64" if prec == "fp64" else "f32")
        ktot, kl, kr = kernel_time(be, one, label)
        avg = ktot / max(1, kr.krylov_iters) * 1e-3  # s per working launch
        kb = spmv_bytes(cells, nnz, w, fmt)
        gbs_iter = cells * it / dt * iteration_bytes(cells, nnz, w, fmt) / cells / 1e9
        out[prec] = {"ms_per_solve": round(dt / reps * 1e3, 3), "iters": it // reps, "refinements": r.refinements,
                     "rel_residual": r.krylov_rel_residual, "scaled_residual": r.final_residual,
                     "mcell_iters_per_s": round(cells * it / dt / 1e6, 1),
                     "hbm_gbs_model": round(gbs_iter, 1), "hbm_frac_model": round(gbs_iter / HBM_PEAK_GBS, 4),
                     "spmv_v": {"avg_us": round(avg * 1e6, 3), "algorithmic_bytes": kb,
                                "achieved_gbs": round(kb / avg / 1e9, 1),
                                "frac": round(kb / avg / 1e9 / HBM_PEAK_GBS, 4), "launches": kl}}
    out["speedup_time_to_solution"] = round(t["fp64"] / t["fp32"], 3)
    import numpy as np
    d = sol["fp32"] - sol["fp64"]
    out["fp32_vs_fp64_solution"] = {"l2_rel": float(np.linalg.norm(d) / np.linalg.norm(sol["fp64"])),
                                    "linf_rel": float(np.abs(d).max() / np.abs(sol["fp64"]).max())}
    # BiCGSTAB's iteration count on this system moves by +-10 % under perturbations far below
    # FP32 rounding, so the FP32 / FP64 comparison is repeated over further seeded
    # right-hand sides (fluxes) of the same matrix and summed
    g = c["mesh"].geometry()
    area = np.linalg.norm(g["face_area"], axis=1)
    runs = [{"seed": 5, "fp64_ms": out["fp64"]["ms_per_solve"], "fp32_ms": out["fp32"]["ms_per_solve"],
             "fp64_iters": out["fp64"]["iters"], "fp32_iters": out["fp32"]["iters"]}]
    for seed in extra_seeds:
        flux = area * np.random.default_rng(seed).standard_normal(len(area))
        flux[g["neighbor"] < 0] = 0.0
        c["mdot"].set_face_fluxes(flux)
        run = {"seed": seed}
        for prec in ("fp64", "fp32"):
            slv = P.solver.BiCGSTAB(tol=tol, precision=prec)
            p.fill(0.0)
            slv.solve(eqn, 1, 1e-30, ref_cell=0)
            be.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                p.fill(0.0)
                r = slv.solve(eqn, 1, 1e-30, ref_cell=0)
            be.synchronize()
            run[prec + "_ms"] = round((time.perf_counter() - t0) / reps * 1e3, 3)
            run[prec + "_iters"] = r.krylov_iters
            run[prec + "_rel_residual"] = r.krylov_rel_residual
        runs.append(run)
    for run in runs:
        run["speedup"] = round(run["fp64_ms"] / run["fp32_ms"], 3)
    out["seeds"] = runs
    out["speedup_time_to_solution_all_seeds"] = round(sum(r["fp64_ms"] for r in runs) /
                                                      sum(r["fp32_ms"] for r in runs), 3)
    return out


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


def _pair_run(be, P, make, step, steps, label):
    """The same `steps` outer steps on two instances of a case, one with the pressure
    correction in FP64 and one in FP32 + refinement, side by side from the same start; per
    step the time of each and the FP32 run's p and U against the FP64 run's."""
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
    cells): Taylor-Green vortex plus seeded N(0, 1e-3) noise, nu 1e-3, backward Euler at
    Courant 0.5, LinearUpwind, 2 pressure correctors; `steps` time steps with the pressure
    correction in FP64 and in FP32 + refinement side by side."""
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

    rows, tot = _pair_run(be, P, make, step, steps, "step")
    return {"workload": f"PISO Taylor-Green on {n}^3 x 6 tetrahedra, 2 correctors, p' to {tol:g}",
            "cells": 6 * n ** 3, "steps": rows, "speedup_total": round(tot["fp64"] / tot["fp32"], 3)}


def adjoint_block(be, P, cases, n, ndir, tol=1e-13):
    """Discrete adjoint of steady UPWIND advection-diffusion on the n^3 jittered mesh
    (log-normal kappa, source s ~ U(0, 1)): J, dJ/ds and dJ/dkappa from one A^T solve, checked
    against central differences of the FP64 primal along `ndir` seeded random positive
    directions (half in s, half in kappa, the latter relative to kappa).  J is quadratic in s
    (T is linear in s), so the central difference in s is exact for any step and a large one
    (1e-3) keeps the primal's solver noise (~1e-11 relative in J) out of it; in kappa the step
    is 1e-4 (truncation ~1e-8 relative)."""
    import numpy as np
    c = cases.adjoint_transport(be, P.box(n, n, n, jitter=0.2, seed=5))
    slv = P.solver.BiCGSTAB(tol=tol, max_iters=20000)
    # three FP64 steps: each restarts BiCGSTAB from the current solution (refinement), so the
    # primal is converged far enough for a 1e-6 central difference of J to be meaningful
    primal = slv.solve(c["eqn"], 3, 1e-30)
    be.synchronize()
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
            slv.solve(c["eqn"], 3, 1e-30)
            Js.append(P.adjoint.objective(c["eqn"], 0, 0.7))
        field.values = base
        fd = (Js[0] - Js[1]) / (2 * eps)
        e = abs(float(grad @ v) - fd) / abs(fd)
        errs.append(round(e, 12))
        worst = max(worst, e)
    return {"workload": f"adjoint of UPWIND advection-diffusion {n}^3 jittered, log-normal kappa, "
                        "J = mean T^2 over x > 0.7", "cells": n ** 3, "J": J,
            "adjoint_solve_ms": round(t_adj * 1e3, 3), "adjoint_krylov_iters": res.krylov_iters,
            "adjoint_converged": bool(res.converged), "adjoint_rel_residual": res.krylov_rel_residual,
            "primal_rel_residual": primal.krylov_rel_residual, "fd_directions": ndir, "fd_eps_s": eps_s, "fd_eps_kappa_rel": eps_k,
            "max_rel_error_vs_fd": worst, "rel_errors_s_kappa_alternating": errs}


def export_bytes(n_cells, n_vertices, n_adjacency, ncomp):
    """Algorithmic bytes of one cell -> point launch: vertex offsets, the adjacency's cell ids
    (int32) and weights (FP64), every cell value read once, every point value written once."""
    return 4 * (n_vertices + 1) + 12 * n_adjacency + 8 * ncomp * (n_cells + n_vertices)


def export_block(be, P, n, vtu_field, reps=10):
    """Output export: the device interpolation of a vector field to the vertices of an n^3
    jittered hex mesh (k_cell_to_point<3>, HIP events), and the wall time of one compressed
    VTU write of `vtu_field` (host compression and file I/O included)."""
    import numpy as np
    from prism_amd import output as O
    progress(f"export {n}^3")
    data = P.box(n, n, n, jitter=0.2, seed=1)
    m = P.Mesh(be, data, {p: {"U": ("fixed", (0.0, 0.0, 0.0))} for p in data.patch_names}, [("U", "vector")])
    U = m.vector("U", values=np.random.default_rng(0).standard_normal((m.n_cells, 3)))
    O.cell_to_point(U)  # first call uploads the adjacency and its weights
    tot_ms, launches, _ = kernel_time(be, lambda: [O.cell_to_point(U) for _ in range(reps)], "k_cell_to_point<3>")
    avg_ms = tot_ms / launches
    nv, nadj = len(data.vertices), len(data.cell_vtx)
    b = export_bytes(m.n_cells, nv, nadj, 3)
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"bench_export_{os.getpid()}.vtu")
    be.synchronize()
    t0 = time.perf_counter()
    O.to_vtu(vtu_field, path)
    t_vtu = time.perf_counter() - t0
    size = os.path.getsize(path)
    os.remove(path)
    return {"workload": f"vector cell -> point, {n}^3 jittered hex ({nv} vertices)", "kernel": "k_cell_to_point<3>",
            "avg_us": round(avg_ms * 1e3, 2), "launches": launches, "algorithmic_bytes": b,
            "achieved_gbs": round(b / (avg_ms * 1e-3) / 1e9, 1),
            "frac": round(b / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "vtu_write_s": round(t_vtu, 3), "vtu_bytes": size, "vtu_cells": vtu_field.mesh.n_cells}


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
    tot_ms, launches, _ = kernel_time(be, lambda: [model.update_fields() for _ in range(reps)], "k_kepsilon")
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


def cpu_baseline(nx, budget_s=15.0):
    """FP64 CPU restatement (oracle, one thread): the same assembled MUSCL system, the same
    unpreconditioned BiCGSTAB iteration, a bounded number of iterations."""
    import numpy as np

    import oracle
    import prism_amd as P
    from prism_amd import cases

    be = oracle.backend()
    t0 = time.perf_counter()
    c = build_case(be, P, cases, nx)
    t_asm = time.perf_counter() - t0
    eqn = c["eqn"]
    rp, col, val, n = eqn.matrix()
    b = eqn.rhs()
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
           "sample": f"{it} unpreconditioned FP64 BiCGSTAB iterations on the {nx}^3 system "
                     f"({n} cells, {len(val)} nnz) in {dt:.2f} s, single thread, -O2",
           "assembly_s": round(t_asm, 3), "cpu": platform.processor() or platform.machine(),
           "host_cores_available": os.cpu_count()}
    try:
        out["all_cores"] = cpu_baseline_omp(A, b, nx, budget_s / 2)
    except Exception as e:
        out["all_cores"] = {"value": None, "error": str(e)}
    return out


def cpu_baseline_omp(A, b, nx, budget_s):
    """The same iteration on every host core this process may use (OpenMP, static schedule,
    capped by OMP_NUM_THREADS), sized from a 200-iteration calibration run."""
    import oracle

    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    threads = max(1, threads)
    rp, col, val = A.indptr, A.indices, A.data
    oracle.omp_bicgstab(rp, col, val, b, 50, threads)  # warm-up: thread pool, first touch
    _, it0, dt0 = oracle.omp_bicgstab(rp, col, val, b, 200, threads)  # calibration
    k = int(max(10, min(100000, budget_s * it0 / max(dt0, 1e-9))))
    _, it, dt = oracle.omp_bicgstab(rp, col, val, b, k, threads)
    n = len(b)
    return {"value": round(n * it / dt / 1e6, 3), "unit": "Mcell-iters/s", "cores": threads, "kind": "port",
            "sample": f"{it} unpreconditioned FP64 BiCGSTAB iterations on the {nx}^3 system in {dt:.2f} s, "
                      f"{threads} OpenMP threads, -O2"}


def pmc_workload(args):
    """The program profiled by the --pmc passes: one FP32 solve of the workload, then a
    256 MiB device copy of known size for calibrating FETCH_SIZE / WRITE_SIZE."""
    import torch

    import prism_amd as P
    from prism_amd import cases

    be = P.Backend.hip(0)
    if args.pressure_only:
        print(json.dumps(pressure_block(be, P, cases, args.pressure_n, args.tol, extra_seeds=())))
        return
    c = build_case(be, P, cases, args.n)
    c["T"].fill(0.0)
    r = P.solver.BiCGSTAB(tol=args.tol, precision="fp32").step_eqn(c["eqn"])
    be.synchronize()
    a = torch.ones(CALIB_FLOATS, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 2.0, out=b)
    torch.cuda.synchronize()
    print(f"pmc workload done: {r.krylov_iters} FP32 iterations, calibration copy {CALIB_FLOATS * 4} bytes")


def read_pmc(path, counter):
    rows = list(csv.DictReader(open(path)))
    per = {}
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        per.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return per


def write_traffic(fetch_csv, write_csv, n):
    f = read_pmc(fetch_csv, "FETCH_SIZE")
    w = read_pmc(write_csv, "WRITE_SIZE")
    def pick(d, key):
        vals = [v for k, vs in d.items() if key in k for v in vs]
        return sum(vals) / len(vals) if vals else None
    spmv_f = pick(f, "k_iter_v3<float")
    spmv_w = pick(w, "k_iter_v3<float")
    cal_f = max((max(vs) for k, vs in f.items() if "elementwise" in k), default=None)
    cal_w = max((max(vs) for k, vs in w.items() if "elementwise" in k), default=None)
    known = CALIB_FLOATS * 4
    ff = known / (cal_f * 1024) if cal_f else None
    wf = known / (cal_w * 1024) if cal_w else None
    out = {"kernel": "bicgstab_spmv_v_f32", "n": n,
           "fetch_kib_per_launch": spmv_f, "write_kib_per_launch": spmv_w,
           "calibration_copy_bytes": known, "calibration_fetch_kib": cal_f, "calibration_write_kib": cal_w,
           "fetch_factor": ff, "write_factor": wf,
           "bytes_per_launch": (spmv_f * 1024 * ff + spmv_w * 1024 * wf) if (spmv_f and spmv_w and ff and wf) else None,
           "sources": [os.path.basename(fetch_csv), os.path.basename(write_csv)]}
    os.makedirs(os.path.dirname(TRAFFIC_FILE), exist_ok=True)
    with open(TRAFFIC_FILE, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64, help="cells per axis")
    ap.add_argument("--pressure-n", type=int, default=128, help="cells per axis of the pressure system (0: skip)")
    ap.add_argument("--adjoint-n", type=int, default=128, help="cells per axis of the adjoint check (0: skip)")
    ap.add_argument("--adjoint-dirs", type=int, default=32)
    ap.add_argument("--simple-n", type=int, default=128, help="cells per axis of the SIMPLE cavity (0: skip)")
    ap.add_argument("--simple-outer", type=int, default=5)
    ap.add_argument("--piso-n", type=int, default=32, help="hexes per axis of the PISO tet mesh (0: skip; at most 96)")
    ap.add_argument("--piso-steps", type=int, default=10)
    ap.add_argument("--export-n", type=int, default=128, help="cells per axis of the export measurement (0: skip)")
    ap.add_argument("--kepsilon-n", type=int, default=128, help="cells per axis of the k-epsilon update (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--pmc-workload", action="store_true")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"))
    ap.add_argument("--pressure-only", action="store_true",
                    help="run only the pressure comparison (for profiling it) and print its JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this path has no multi-GPU mode (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    if args.traffic:
        write_traffic(args.traffic[0], args.traffic[1], args.n ** 3)
        return
    if args.pmc_workload:
        pmc_workload(args)
        return

    import prism_amd as P
    from prism_amd import cases

    be = P.Backend.hip(0)
    if args.pressure_only:
        print(json.dumps(pressure_block(be, P, cases, args.pressure_n, args.tol, extra_seeds=())))
        return
    c = build_case(be, P, cases, args.n)
    eqn, T = c["eqn"], c["T"]
    mesh = c["mesh"]
    n, nnz = mesh.n_cells, mesh.nnz
    rp, col, _, _ = eqn.matrix()
    fmt = matrix_format(rp, col)

    s32 = P.solver.BiCGSTAB(tol=args.tol, precision="fp32")
    s64 = P.solver.BiCGSTAB(tol=args.tol, precision="fp64")
    t32, it32, r32 = timed_solves(be, eqn, T, s32, args.steps, args.warmup)
    x32 = T.values
    t64, it64, r64 = timed_solves(be, eqn, T, s64, args.steps, args.warmup)
    x64 = T.values
    # FP32+refinement solution against the FP64 GPU solution of the same system
    import numpy as np
    err = {"l2_rel": float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64)),
           "linf_rel": float(np.abs(x32 - x64).max() / np.abs(x64).max())}

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
    T.fill(0.0)
    eqn.zero_out_coeffs()
    eqn.update_coeffs()  # the benchmark system again

    # roofline: per-launch HIP events on the library stream, one solve each.  Launches issued
    # past the device's convergence flag return at once; their time is kept in the total but
    # they do no work, so the per-launch figure divides by the iterations actually run.
    def one(slv):
        T.fill(0.0)
        return slv.step_eqn(eqn)
    tot32_ms, k32, rr32 = kernel_time(be, lambda: one(s32), "bicgstab_spmv_v_f32")
    tot64_ms, k64, rr64 = kernel_time(be, lambda: one(s64), "bicgstab_spmv_v_f64")
    avg32_ms = tot32_ms / rr32.krylov_iters
    avg64_ms = tot64_ms / rr64.krylov_iters
    b32 = spmv_bytes(n, nnz, 4, fmt)
    b64 = spmv_bytes(n, nnz, 8, fmt)
    ach32 = b32 / (avg32_ms * 1e-3) / 1e9
    ach64 = b64 / (avg64_ms * 1e-3) / 1e9

    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tj = json.load(f)
        if tj.get("n") == n:
            traffic = tj.get("bytes_per_launch")

    ms32 = t32 / args.steps * 1e3
    ms64 = t64 / args.steps * 1e3
    v32 = n * it32 / t32 / 1e6
    v64 = n * it64 / t64 / 1e6
    gbs32 = v32 * 1e6 * iteration_bytes(n, nnz, 4, fmt) / n / 1e9
    gbs64 = v64 * 1e6 * iteration_bytes(n, nnz, 8, fmt) / n / 1e9
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
                   "cells": n, "nnz": nnz, "faces": mesh.n_faces, "tol": args.tol,
                   "matrix_format": fmt[0], "diagonals": fmt[1]},
        "hbm_gbs_model": round(gbs32, 1),
        "iters_per_solve": it32 / args.steps,
        "refinements": r32.refinements,
        "rel_residual": r32.krylov_rel_residual,
        "fp32_vs_fp64_solution": err,
        "fp64": {"value": round(v64, 2), "ms_per_step": round(ms64, 4), "iters_per_solve": it64 / args.steps,
                 "hbm_gbs_model": round(gbs64, 1), "rel_residual": r64.krylov_rel_residual,
                 "spmv_avg_us": round(avg64_ms * 1e3, 3), "spmv_achieved_gbs": round(ach64, 1)},
        "speedup_time_to_solution": round(t64 / t32, 3),
        "speedup_mcell_iters": round(v32 / v64, 3),
        "assembly": {"faces_per_s": round(mesh.n_faces / t_asm, 1), "ms_per_pass": round(t_asm * 1e3, 4),
                     "bytes_per_pass_model": assembly_bytes(n, mesh.n_faces, nnz),
                     "hbm_gbs_model": round(assembly_bytes(n, mesh.n_faces, nnz) / t_asm / 1e9, 1),
                     "hbm_frac_model": round(assembly_bytes(n, mesh.n_faces, nnz) / t_asm / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": {"bound": "hbm", "kernel": "bicgstab_spmv_v_f32", "achieved": round(ach32, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach32 / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "avg_us": round(avg32_ms * 1e3, 3), "launches": k32,
                     "iterations": rr32.krylov_iters, "avg_us_all_launches": round(tot32_ms / k32 * 1e3, 3),
                     "algorithmic_bytes": b32},
        "device": {"arch": name, "cus": ncu},
    }
    if args.pressure_n > 0:
        out["pressure"] = pressure_block(be, P, cases, args.pressure_n, args.tol)
    if args.adjoint_n > 0:
        out["adjoint"] = adjoint_block(be, P, cases, args.adjoint_n, args.adjoint_dirs)
    if args.simple_n > 0:
        out["simple"] = simple_block(be, P, cases, args.simple_n, args.simple_outer, args.tol)
    if args.piso_n > 0:
        out["piso"] = piso_block(be, P, cases, args.piso_n, args.piso_steps, args.tol)
    if args.export_n > 0:
        out["export"] = export_block(be, P, args.export_n, T)
    if args.kepsilon_n > 0:
        out["kepsilon"] = kepsilon_block(be, P, args.kepsilon_n)
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(args.n)
        except Exception as e:  # the CPU leg never blocks the GPU figure
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
