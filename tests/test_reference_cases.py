"""The reference's own physics tests (tests/test_*.cpp), run on the FP64 CPU oracle with the
reference's fixtures and tolerances.  They pin the oracle: each check is the reference's
REQUIRE, and where the reference's comment states the value it observed, that value too."""
import json
import math
import os

import numpy as np
import pytest

import prism_amd as P
import refcases as R


@pytest.fixture(scope="module")
def be(ref):
    return ref


def test_poisson_structured(be):
    # tests/test_poisson.cpp:57-61: rel L2 < 0.0004 on poisson/mesh_hex.unv
    c = R.poisson(be, "mesh_hex.unv")
    R.solve_with_bicgstab(c["eqn"])
    err = R.l2_rel(c["P"].values, c["exact"])
    assert err < 0.0004


def test_poisson_unstructured(be):
    # tests/test_poisson.cpp:51-55: rel L2 < 0.078 on poisson/mesh.unv
    c = R.poisson(be, "mesh.unv")
    R.solve_with_bicgstab(c["eqn"])
    assert R.l2_rel(c["P"].values, c["exact"]) < 0.078


def test_advection_1d(be):
    # tests/test_advection_1d.cpp:47-58: ||T - exact|| < 0.25 ("around 0.209"), monotone
    c = R.advection_1d(be)
    R.solve_with_bicgstab(c["eqn"], 5, 5, 1e-20)
    T = c["T"].values
    diff = float(np.linalg.norm(T - c["exact"]))
    assert diff < 0.25
    assert abs(diff - 0.209) < 5e-4
    assert np.all(np.diff(T) <= 0.0)


def test_transient_1d(be):
    # tests/test_transient_1d.cpp:81-88
    r = R.transient_1d(be)
    e = r["errors_40_60"]
    assert len(e) == 2 and e[1] < e[0]
    assert e[0] < 0.10 and e[1] < 0.10
    assert r["final"] < 0.02


def test_implicit_source(be):
    # tests/test_implicit_source.cpp:46-48
    c = R.implicit_source(be)
    R.solve_with_bicgstab(c["eqn"])
    assert R.l2_rel(c["y"].values, c["exact"]) < 0.005


@pytest.mark.parametrize("case,expected", [("cavity_hex", True), ("coarsePipeHexSymmetrical", False),
                                           ("ductSIMPLE", False)])
def test_needs_pressure_reference(be, case, expected):
    # tests/test_needs_pressure_ref.cpp:10-29
    mesh = R.load_mesh(be, case)
    assert P.needs_pressure_reference(mesh.pressure("P")) is expected


def _foam(case):
    with open(os.path.join(R.GOLDEN, case, "foam_fields.json")) as f:
        d = json.load(f)
    return np.array(d["p"]), np.array(d["U"])


def test_simple_coarse_pipe_matches_openfoam(be):
    # tests/test_simple_coarse_pipe.cpp:12-49
    f = R.pressure_velocity(be, "coarsePipeHex")
    ctl = P.algo.SIMPLEControls(be)
    for _ in range(100):
        P.algo.IncompressibleSIMPLE(ctl).step(f["eqns"], f["U"], f["mdot"], f["p"])
    p_ref, U_ref = _foam("coarsePipeHex")
    U = f["U"].values
    assert R.l2_rel(f["p"].values, p_ref) < 0.005
    assert R.l2_rel(U[:, 0], U_ref[:, 0]) < 0.005
    assert R.l2_rel(U[:, 1], U_ref[:, 1]) < 0.09


def test_simple_duct_symmetry_matches_openfoam(be):
    # tests/test_simple_duct_symmetry.cpp:13-53 (Symmetry<Transport> on both momentum equations)
    f = R.pressure_velocity(be, "ductSIMPLE", symmetry=True)
    ctl = P.algo.SIMPLEControls(be)
    for _ in range(100):
        P.algo.IncompressibleSIMPLE(ctl).step(f["eqns"], f["U"], f["mdot"], f["p"])
    p_ref, U_ref = _foam("ductSIMPLE")
    U = f["U"].values
    assert R.l2_rel(f["p"].values, p_ref) < 0.025
    assert R.l2_rel(U[:, 0], U_ref[:, 0]) < 0.07
    assert R.l2_rel(U[:, 1], U_ref[:, 1]) < 0.007


def test_simple_pipe_mass_conservation(be):
    # tests/test_algorithms_pipe.cpp:14-37
    f = R.pressure_velocity(be, "coarsePipeHex")
    ctl = P.algo.SIMPLEControls(be)
    for _ in range(200):
        P.algo.IncompressibleSIMPLE(ctl).step(f["eqns"], f["U"], f["mdot"], f["p"])
    imb, a, b = R.imbalance(f["U"], "Inlet", "Outlet")
    assert imb < 1e-6


def test_piso_cavity_auto_pref(be):
    # tests/test_algorithms_pipe.cpp:39-76
    f = R.pressure_velocity(be, "cavity_hex", nu=0.01, transient_dt=0.005)
    ctl = P.algo.PISOControls(be, outer_iterations=3, pressure_correction_steps=1,
                              simple=dict(non_ortho_correctors=0))
    for _ in range(20):
        P.algo.PISO(ctl).step(f["eqns"], f["U"], f["mdot"], f["p"])
        for k in range(2):
            f["U"].component(k).update_prev_time_steps()
    p = f["p"].values
    assert np.all(np.isfinite(p))


@pytest.mark.parametrize("steps,n", [(1, 200), (2, 50)])
def test_piso_pipe_mass_conservation(be, steps, n):
    # tests/test_algorithms_pipe.cpp:107-187 (PISO without and with PRIME correctors)
    f = R.pressure_velocity(be, "coarsePipeHex")
    ctl = P.algo.PISOControls(be, outer_iterations=1, momentum_implicit_steps=1, pressure_correction_steps=steps,
                              simple=dict(non_ortho_correctors=2, momentum_urf=0.7, pressure_urf=0.3,
                                          momentum_residual=1e-7, pressure_residual=1e-7))
    for _ in range(n):
        P.algo.PISO(ctl).step(f["eqns"], f["U"], f["mdot"], f["p"])
    imb, a, b = R.imbalance(f["U"], "Inlet", "Outlet")
    assert imb < 1e-6


@pytest.mark.slow
def test_simple_duct_mass_conservation(be):
    # tests/test_algorithms_pipe.cpp:78-105 (500 outer iterations, symmetry handlers)
    f = R.pressure_velocity(be, "ductSIMPLE", symmetry=True)
    ctl = P.algo.SIMPLEControls(be)
    for _ in range(500):
        P.algo.IncompressibleSIMPLE(ctl).step(f["eqns"], f["U"], f["mdot"], f["p"])
    imb, a, b = R.imbalance(f["U"], "inlet", "outlet")
    assert imb < 1e-6


def test_known_answer_solvers(be):
    # tests/test_solver.cpp:10-57: 20 steps of BiCGSTAB and of Jacobi, rel err < 1e-4
    n = 10
    A = 6 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    x_true = np.arange(1, n + 1, dtype=float)
    b = A @ x_true
    M = P.Matrix.from_dense(be, A)
    for slv in (P.solver.BiCGSTAB(), P.solver.Jacobi()):
        x = np.zeros(n)
        for _ in range(20):
            x = slv.step(M, x, b)
        assert R.l2_rel(x, x_true) < 1e-4
