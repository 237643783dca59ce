"""GPU parity on the reference's own test scenarios: the HIP library (FP64 path, through the
C ABI) against the FP64 oracle on the same fixtures, field by field at <= 1e-10 relative
L2, and the reference's own tolerances on the GPU result."""
import json
import os

import numpy as np
import pytest

import prism_amd as P
import refcases as R

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def both(hip, ref, fn):
    return fn(hip), fn(ref)


@pytest.mark.parametrize("mesh_file,bound", [("mesh_hex.unv", 0.0004), ("mesh.unv", 0.078)])
def test_poisson(hip, ref, mesh_file, bound):
    def run(be):
        c = R.poisson(be, mesh_file)
        R.solve_with_bicgstab(c["eqn"])
        return c["P"].values, c["exact"], c["P"].grad()
    (ph, ex, gh), (pr, _, gr) = both(hip, ref, run)
    assert rel(ph, pr) <= TOL
    assert rel(gh, gr) <= 1e-9
    assert R.l2_rel(ph, ex) < bound


def test_advection_1d(hip, ref):
    def run(be):
        c = R.advection_1d(be)
        R.solve_with_bicgstab(c["eqn"], 5, 5, 1e-20)
        return c["T"].values, c["exact"]
    (th, ex), (tr, _) = both(hip, ref, run)
    assert rel(th, tr) <= TOL
    assert np.linalg.norm(th - ex) < 0.25
    assert np.all(np.diff(th) <= 0.0)


def test_transient_1d_adam_moulton(hip, ref):
    rh, rr = both(hip, ref, lambda be: R.transient_1d(be))
    for a, b in zip(rh["history"], rr["history"]):
        assert rel(a, b) <= TOL
    assert rh["final"] < 0.02
    assert rh["errors_40_60"][1] < rh["errors_40_60"][0] < 0.10


def test_implicit_source(hip, ref):
    def run(be):
        c = R.implicit_source(be)
        R.solve_with_bicgstab(c["eqn"])
        return c["y"].values, c["exact"]
    (yh, ex), (yr, _) = both(hip, ref, run)
    assert rel(yh, yr) <= TOL
    assert R.l2_rel(yh, ex) < 0.005


@pytest.mark.parametrize("case,expected", [("cavity_hex", True), ("ductSIMPLE", False)])
def test_needs_pressure_reference(hip, case, expected):
    mesh = R.load_mesh(hip, case)
    assert P.needs_pressure_reference(mesh.pressure("P")) is expected


def _simple(be, case, symmetry, n):
    f = R.pressure_velocity(be, case, symmetry=symmetry)
    ctl = P.algo.SIMPLEControls(be)
    for _ in range(n):
        P.algo.IncompressibleSIMPLE(ctl).step(f["eqns"], f["U"], f["mdot"], f["p"])
    return f


@pytest.mark.parametrize("case,symmetry,bounds", [
    ("coarsePipeHex", False, (0.005, 0.005, 0.09)),   # tests/test_simple_coarse_pipe.cpp:36-49
    ("ductSIMPLE", True, (0.025, 0.07, 0.007)),       # tests/test_simple_duct_symmetry.cpp:38-53
])
def test_simple_matches_oracle_and_openfoam(hip, ref, case, symmetry, bounds):
    fh, fr = both(hip, ref, lambda be: _simple(be, case, symmetry, 100))
    assert rel(fh["p"].values, fr["p"].values) <= TOL
    assert rel(fh["U"].values, fr["U"].values) <= TOL
    assert rel(fh["mdot"].face_fluxes(), fr["mdot"].face_fluxes()) <= TOL
    with open(os.path.join(R.GOLDEN, case, "foam_fields.json")) as f:
        d = json.load(f)
    p_ref, U_ref = np.array(d["p"]), np.array(d["U"])
    U = fh["U"].values
    assert R.l2_rel(fh["p"].values, p_ref) < bounds[0]
    assert R.l2_rel(U[:, 0], U_ref[:, 0]) < bounds[1]
    assert R.l2_rel(U[:, 1], U_ref[:, 1]) < bounds[2]
    imb, _, _ = R.imbalance(fh["U"], *(("Inlet", "Outlet") if case == "coarsePipeHex" else ("inlet", "outlet")))
    assert imb < 1e-4


def test_piso_cavity_backward_euler(hip, ref):
    def run(be):
        f = R.pressure_velocity(be, "cavity_hex", nu=0.01, transient_dt=0.005)
        ctl = P.algo.PISOControls(be, outer_iterations=3, pressure_correction_steps=1,
                                  simple=dict(non_ortho_correctors=0))
        for _ in range(20):
            P.algo.PISO(ctl).step(f["eqns"], f["U"], f["mdot"], f["p"])
            for k in range(2):
                f["U"].component(k).update_prev_time_steps()
        return f
    fh, fr = both(hip, ref, run)
    assert np.all(np.isfinite(fh["p"].values))
    assert rel(fh["p"].values, fr["p"].values) <= TOL
    assert rel(fh["U"].values, fr["U"].values) <= TOL


def test_piso_prime_pipe(hip, ref):
    def run(be):
        f = R.pressure_velocity(be, "coarsePipeHex")
        ctl = P.algo.PISOControls(be, outer_iterations=1, momentum_implicit_steps=1, pressure_correction_steps=2,
                                  simple=dict(non_ortho_correctors=2, momentum_urf=0.7, pressure_urf=0.3,
                                              momentum_residual=1e-7, pressure_residual=1e-7))
        for _ in range(50):
            P.algo.PISO(ctl).step(f["eqns"], f["U"], f["mdot"], f["p"])
        return f
    fh, fr = both(hip, ref, run)
    assert rel(fh["p"].values, fr["p"].values) <= TOL
    assert rel(fh["U"].values, fr["U"].values) <= TOL
    imb, _, _ = R.imbalance(fh["U"], "Inlet", "Outlet")
    assert imb < 1e-6


def test_piso_taylor_green_tets(hip, ref):
    """The transient PISO workload on tetrahedra (6 per hex, Taylor-Green vortex plus seeded
    noise, backward Euler, LinearUpwind, 2 pressure correctors with 2 non-orthogonal
    correctors), at 6^3 x 6 cells for 4 steps: GPU against the oracle, FP64, and the same
    steps with the pressure correction in FP32 + refinement against the FP64 GPU run."""
    from prism_amd import cases

    def run(be, pressure_precision=0):
        c = cases.taylor_green(be, n=6, tets=True)
        ctl = P.algo.PISOControls(be, outer_iterations=1, pressure_correction_steps=2,
                                  simple=dict(pressure_precision=pressure_precision, pressure_tol=1e-12,
                                              momentum_tol=1e-12))
        for _ in range(4):
            P.algo.PISO(ctl).step(c["eqns"], c["U"], c["mdot"], c["p"])
            for comp in (c["U"].x, c["U"].y, c["U"].z):
                comp.update_prev_time_steps()
        return c["p"].values, c["U"].values

    (ph, uh), (pr, ur) = both(hip, ref, run)
    assert np.all(np.isfinite(ph)) and np.all(np.isfinite(uh))
    assert rel(ph, pr) <= TOL
    assert rel(uh, ur) <= TOL
    p32, u32 = run(hip, pressure_precision=1)
    assert rel(p32, ph) <= 1e-8
    assert rel(u32, uh) <= 1e-8


def test_simple_cavity_smart(hip, ref):
    """The SIMPLE lid-driven cavity workload (SMART momentum, relaxation 0.7 / 0.3, automatic
    reference cell) at 8^3, one outer iteration: GPU against the oracle, FP64.  After the
    second iteration the two runs differ by 1.2e-2 in p (measured): SMART's face weights
    switch discretely (with the reference's integer division it is downwind or upwind, and
    a 0/0 ratio falls back to upwind), so round-off differences in the many near-zero
    velocities of the early cavity flow select different weights.  SMART's weights
    themselves are pinned by test_gpu_parity.py::test_advection_assembly[SMART]."""
    from prism_amd import cases

    def run(be):
        c = cases.cavity(be, n=8)
        ctl = P.algo.SIMPLEControls(be, momentum_tol=1e-14, pressure_tol=1e-14)
        for _ in range(1):
            P.algo.IncompressibleSIMPLE(ctl).step(c["eqns"], c["U"], c["mdot"], c["p"])
        return c["p"].values, c["U"].values

    (ph, uh), (pr, ur) = both(hip, ref, run)
    assert np.all(np.isfinite(ph)) and np.all(np.isfinite(uh))
    assert rel(ph, pr) <= TOL
    assert rel(uh, ur) <= TOL
