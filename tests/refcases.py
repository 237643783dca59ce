"""The reference's own test scenarios (tests/test_*.cpp), built through the Python mirror on
any backend (the FP64 oracle or the HIP library), on the reference's case fixtures copied
into tests/golden/cases by tests/golden/make_fixtures.py.  Each builder follows the cited
test line by line; the checks themselves live in the test modules."""
from __future__ import annotations

import math
import os

import numpy as np

import prism_amd as P
from prism_amd import case as case_io

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cases")


def load_mesh(be, name, mesh_file="mesh.unv"):
    """loadMesh (tests/test_utils.h:23-26): the mesh and the fields.json beside it."""
    return case_io.load(os.path.join(GOLDEN, name), mesh_file).mesh(be)


def l2_rel(x, ref):
    """l2NormRel (tests/test_utils.h:35-37)."""
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def solve_with_bicgstab(eqn, n_ortho_correctors=5, max_iters=15, tol=1e-20, solver=None):
    """solveWithBiCGSTAB (tests/test_utils.h:138-146)."""
    s = solver or P.solver.BiCGSTAB()
    for _ in range(n_ortho_correctors):
        s.solve(eqn, max_iters, tol)


def flow_rate(U, patch):
    """flowRate (tests/test_utils.h:49-63): sum over the patch of U.valueAtFace . S."""
    mesh = U.mesh
    g = mesh.geometry()
    idx = mesh.data.patch_names.index(patch)
    sel = (g["neighbor"] < 0) & (g["face_patch"] == idx)
    uf = U.face_values()[sel]
    S = g["face_area"][sel]
    return float(np.sum((uf[:, 0] * S[:, 0] + uf[:, 1] * S[:, 1]) + uf[:, 2] * S[:, 2]))


def cell_x(mesh):
    return mesh.geometry()["cell_center"][:, 0]


# --------------------------------------------------------------- tests/test_poisson.cpp ----
def poisson(be, mesh_file):
    """testPoissonWithMesh (test_poisson.cpp:24-48): Corrected(c = I, P) = S with
    S = 2 pi^2 sin(pi x) cos(pi y); returns P after solveWithBiCGSTAB and the exact field."""
    mesh = load_mesh(be, "poisson", mesh_file)
    c = mesh.geometry()["cell_center"]
    x, y = c[:, 0], c[:, 1]
    Pf = mesh.scalar("P", 0.0)
    src = mesh.scalar("S", values=2 * math.pi ** 2 * np.sin(math.pi * x) * np.cos(math.pi * y))
    kappa = mesh.tensor("c", np.eye(3))
    eqn = P.Transport(P.diffusion.Corrected(kappa, Pf), P.source.ConstantScalar(src, "+"))
    return dict(mesh=mesh, P=Pf, eqn=eqn, keep=(src, kappa),
                exact=np.sin(math.pi * x) * np.cos(math.pi * y))


# ---------------------------------------------------------- tests/test_advection_1d.cpp ----
def advection_1d(be):
    """test_advection_1d.cpp:24-58: Upwind(U, T) + NonCorrected(kappa = 0.1, T), U uniform at
    the inlet velocity, solveWithBiCGSTAB(eqn, 5, 5, 1e-20)."""
    mesh = load_mesh(be, "versteeg_advection_1d")
    case = case_io.load(os.path.join(GOLDEN, "versteeg_advection_1d"))
    u_in = np.array(case.bcs["Inlet"]["U"][1], dtype=float)
    T = mesh.scalar("T", 0.0)
    U = mesh.vector("U", tuple(u_in))
    kappa = mesh.scalar("kappa", 0.1)
    eqn = P.Transport(P.convection.Upwind(U, T), P.diffusion.NonCorrected(kappa, T))
    x = cell_x(mesh)
    u = u_in[0]
    exact = -((np.exp(u * x / 0.1) - 1) / (np.exp(u / 0.1) - 1)) + 1
    return dict(mesh=mesh, T=T, U=U, kappa=kappa, eqn=eqn, exact=exact)


# ---------------------------------------------------------- tests/test_transient_1d.cpp ----
def transient_analytical(x, t, kappa):
    """transientAnalyticalSolution (test_transient_1d.cpp:14-30), 100 modes."""
    s = np.zeros_like(x)
    for n in range(100):
        lam = ((2 * n) - 1) * math.pi / 2
        mode = ((-1.0) ** (n + 1)) * math.exp(-kappa * t * lam ** 2) * np.cos(lam * x)
        s += mode / ((2 * n) - 1)
    return s * (4.0 / math.pi) * 100


def transient_1d(be, n_steps=200, solver=None):
    """test_transient_1d.cpp:32-80: AdamMoulton(T, dt = 2) + Corrected(kappa = 1e-3, T), T = 200
    with two history levels, solve(eqn, 10, 1e-20) per step; returns the relative L2 error
    at t = 40, t = 60 and after the last step."""
    mesh = load_mesh(be, "versteeg_trans_1d")
    T = mesh.scalar("T", 200.0)
    T.set_history_size(2)
    T.update_prev_time_steps()
    kappa = mesh.scalar("kappa", 1e-3)
    dt = 2
    eqn = P.Transport(P.temporal.AdamMoulton(T, dt), P.diffusion.Corrected(kappa, T))
    s = solver or P.solver.BiCGSTAB()
    x = cell_x(mesh)
    at, final, history = [], None, []
    for step in range(n_steps):
        s.solve(eqn, 10, 1e-20)
        t = dt * step
        exact = transient_analytical(x, t, 1e-3)
        err = float(np.linalg.norm(T.values - exact) / np.linalg.norm(exact))
        if t in (40, 60):
            at.append(err)
        if step == n_steps - 1:
            final = err
        history.append(T.values.copy())
        T.update_prev_time_steps()
    return dict(errors_40_60=at, final=final, history=history, T=T)


# -------------------------------------------------------- tests/test_implicit_source.cpp ----
def implicit_source(be):
    """test_implicit_source.cpp:31-49: NonCorrected(c = 1, y) = ImplicitField<Negative>(4, y);
    exact y = sinh(2x + 1)."""
    mesh = load_mesh(be, "channel1d_coarse")
    y = mesh.scalar("y", 0.0)
    c = mesh.scalar("c", 1.0)
    eqn = P.Transport(P.diffusion.NonCorrected(c, y), P.source.ImplicitField(4, y, "-"))
    return dict(mesh=mesh, y=y, c=c, eqn=eqn, exact=np.sinh(2 * cell_x(mesh) + 1))


# ------------------------------------------------ pressure-velocity (tests/test_utils.h) ----
def pressure_velocity(be, case, nu=1e-3, symmetry=False, transient_dt=None, div="LinearUpwind"):
    """makePressureVelocityFields + MomentumEquations<LinearUpwind, Corrected,
    Gradient<Negative>> (test_utils.h:161-215); with transient_dt, BackwardEuler(U_k, dt)
    leads each equation and the components keep two history levels
    (test_algorithms_pipe.cpp:41-61); with symmetry, the equations carry the
    Symmetry<Transport> handler (test_algorithms_pipe.cpp:88-89)."""
    mesh = load_mesh(be, case)
    U = mesh.vector("U", (0.0, 0.0, 0.0))
    p = mesh.pressure("P", 0.0)
    nuf = mesh.scalar("nu", nu)
    mdot = U.clone()
    conv = getattr(P.convection, div)
    eqns = []
    for k in range(2):
        comp = U.component(k)
        terms = []
        if transient_dt is not None:
            comp.set_history_size(2)
            comp.update_prev_time_steps()
            terms.append(P.temporal.BackwardEuler(comp, transient_dt))
        terms += [conv(mdot, comp), P.diffusion.Corrected(nuf, comp), P.source.Gradient(p, k, "-")]
        e = P.Transport(*terms)
        if symmetry:
            e.add_symmetry_handler()
        eqns.append(e)
    return dict(mesh=mesh, U=U, p=p, nu=nuf, mdot=mdot, eqns=eqns)


def imbalance(U, inlet, outlet):
    a, b = flow_rate(U, inlet), flow_rate(U, outlet)
    return abs(a + b) / abs(a), a, b
