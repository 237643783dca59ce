"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle on the same inputs."""
import numpy as np
import pytest

import prism_amd as P
from prism_amd import cases

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def assembled(be, builder, *args, **kw):
    c = builder(be, *args, **kw)
    c["eqn"].update_coeffs()
    rp, col, val, n = c["eqn"].matrix()
    return c, val, c["eqn"].rhs()


SCHEMES = ["Upwind", "CentralDifference", "LinearUpwind", "QUICK", "FROMM", "MINMOD", "MUSCL", "SMART"]


@pytest.mark.parametrize("scheme", SCHEMES)
def test_advection_assembly(hip, ref, scheme):
    """Assembly of every NVF scheme with a nonzero gradient field (kappa 1e-3, cell Peclet
    number about 100 on this mesh): the same smooth T on both sides, then A, b and the
    gradient against the oracle."""
    data = P.box(10, 9, 7, jitter=0.2, seed=11)
    out = {}
    for name, be in (("hip", hip), ("ref", ref)):
        c = cases.advection_diffusion(be, data, scheme=scheme)
        x = c["mesh"].geometry()["cell_center"]
        c["T"].values = np.sin(3 * x[:, 0]) * np.cos(2 * x[:, 1]) + x[:, 2] ** 2
        c["eqn"].update_coeffs()
        out[name] = (c["eqn"].matrix()[2], c["eqn"].rhs(), c["T"].grad())
    assert rel(out["hip"][0], out["ref"][0]) < 1e-13
    assert rel(out["hip"][1], out["ref"][1]) < 1e-13
    assert rel(out["hip"][2], out["ref"][2]) < 1e-12


@pytest.mark.parametrize("scheme", SCHEMES)
def test_advection_assembly_and_solve(hip, ref, scheme):
    """Two solve() passes (the second with NVF weights from the first pass's gradients) at
    kappa 0.05, where every scheme's system is solvable by unpreconditioned BiCGSTAB (at
    1e-3 the central-difference, MUSCL and SMART systems of this coarse mesh are not, for
    scipy's BiCGSTAB either; the oracle solves them exactly)."""
    data = P.box(10, 9, 7, jitter=0.2, seed=11)
    out = {}
    for name, be in (("hip", hip), ("ref", ref)):
        c = cases.advection_diffusion(be, data, scheme=scheme, kappa=0.05)
        slv = P.solver.BiCGSTAB(tol=1e-13)
        r = slv.solve(c["eqn"], n_iter=2, eps=1e-20, clear_coeffs=False)  # upwind-like first pass
        A1 = c["eqn"].matrix()[2]
        b1 = c["eqn"].rhs()
        c["eqn"].zero_out_coeffs()
        slv.solve(c["eqn"], n_iter=2, eps=1e-20)  # NVF weights from the new gradients
        out[name] = (A1, b1, c["T"].values, c["T"].grad(), r)
    A_h, b_h, T_h, g_h, r_h = out["hip"]
    A_r, b_r, T_r, g_r, r_r = out["ref"]
    assert rel(A_h, A_r) < 1e-13
    assert rel(b_h, b_r) < 1e-13
    assert rel(T_h, T_r) < 1e-10
    assert rel(g_h, g_r) < 1e-9
    assert abs(r_h.initial_residual - r_r.initial_residual) <= 1e-12 * max(1.0, abs(r_r.initial_residual))


def test_poisson_green_gauss(hip, ref):
    out = {}
    for name, be in (("hip", hip), ("ref", ref)):
        c = cases.poisson(be, n=24)
        slv = P.solver.BiCGSTAB(tol=1e-14)
        for _ in range(3):
            slv.solve(c["eqn"], n_iter=3, eps=1e-20)
        out[name] = (c["T"].values, c["T"].grad(), c["exact"])
    assert rel(out["hip"][0], out["ref"][0]) < 1e-10
    assert rel(out["hip"][1], out["ref"][1]) < 1e-9
    assert rel(out["hip"][0], out["hip"][2]) < 5e-3


def test_known_answer_bicgstab(hip):
    # tests/test_solver.cpp:10-57
    n = 10
    M = 6 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    x_true = np.arange(1, n + 1, dtype=float)
    b = M @ x_true
    A = P.Matrix.from_dense(hip, M)
    x = np.zeros(n)
    s = P.solver.BiCGSTAB()
    for _ in range(20):
        x = s.step(A, x, b)
    assert rel(x, x_true) < 1e-4
    xj = np.zeros(n)
    for _ in range(20):
        xj = P.solver.Jacobi().step(A, xj, b)
    assert rel(xj, x_true) < 1e-4
