"""GPU parity on polyhedral cells (two jittered hexahedra merged into one 10-face cell, more
faces than the gradient kernels stage per cell): least-squares gradients, the MUSCL
advection-diffusion assembly and solve, and the k-epsilon update with its one-pass velocity
gradient, against the oracle."""
import numpy as np
import pytest

import prism_amd as P
from prism_amd import cases
from prism_amd.turbulence import IncompressibleKEpsilon
from polymesh import merged_box
from test_turbulence import _mesh

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_advection_diffusion_on_polyhedra(hip, ref):
    print("case polyhedra advection_diffusion MUSCL", flush=True)
    data = merged_box(P, 10, 9, 7, jitter=0.2, seed=11)
    assert np.diff(data.cell_face_off).max() == 10
    out = {}
    for name, be in (("hip", hip), ("ref", ref)):
        c = cases.advection_diffusion(be, data, scheme="MUSCL", kappa=0.05)
        x = c["mesh"].geometry()["cell_center"]
        c["T"].values = np.sin(3 * x[:, 0]) * np.cos(2 * x[:, 1]) + x[:, 2] ** 2
        grad = c["T"].grad()
        c["eqn"].update_coeffs()
        A, b = c["eqn"].matrix()[2], c["eqn"].rhs()
        c["eqn"].zero_out_coeffs()
        P.solver.BiCGSTAB(tol=1e-13).solve(c["eqn"], n_iter=2, eps=1e-20)
        out[name] = (grad, A, b, c["T"].values)
    h, r = out["hip"], out["ref"]
    assert rel(h[0], r[0]) < 1e-12
    assert rel(h[1], r[1]) < 1e-13 and rel(h[2], r[2]) < 1e-13
    assert rel(h[3], r[3]) < 1e-10


def test_kepsilon_on_polyhedra(hip, ref):
    print("case polyhedra kepsilon", flush=True)
    data = merged_box(P, 9, 7, 6, jitter=0.25, seed=3)
    out, per_comp = [], None
    for be in (hip, ref):
        m = _mesh(be, data)
        g = m.geometry()["cell_center"]
        U = m.vector("U", values=np.stack([np.sin(3 * g[:, 1]), g[:, 0] * g[:, 2], np.cos(2 * g[:, 0])], axis=1))
        model = IncompressibleKEpsilon(U, U, m.scalar("nu", 1e-3))
        rng = np.random.default_rng(6)
        model.k().values = rng.uniform(0.01, 0.1, m.n_cells)
        model.epsilon().values = rng.uniform(0.001, 0.02, m.n_cells)
        model.update_fields()
        grads = [c.grad() for c in (U.x, U.y, U.z)]
        out.append((model.production().values, model.turbulent_viscosity().values, grads))
        if be is hip:  # the same gradients by one single-field launch per component
            for c, v in zip((U.x, U.y, U.z), [c.values for c in (U.x, U.y, U.z)]):
                c.values = v
            per_comp = [c.grad() for c in (U.x, U.y, U.z)]
    for a, b in zip(out[0][2], per_comp):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(out[0][2], out[1][2]):
        assert rel(a, b) < 1e-12
    assert rel(out[0][0], out[1][0]) < 1e-12
    assert rel(out[0][1], out[1][1]) < 1e-12


@pytest.mark.parametrize("scheme", ["Upwind", "CentralDifference"])
def test_cells_sharing_two_faces(hip, ref, scheme):
    """Two merged cells side by side touch across two faces: one matrix entry per neighbouring
    pair holding the sum of both faces' coefficients (the reference sums duplicate triplets,
    linear.cpp:17-29); the pattern, A, b and the solution against the oracle."""
    from polymesh import merge_cells

    print(f"case polyhedra shared faces {scheme}", flush=True)
    data = merge_cells(P.box(4, 4, 2, jitter=0.1, seed=3), [(0, 1), (4, 5), (18, 19), (22, 23)])
    out = {}
    for name, be in (("hip", hip), ("ref", ref)):
        c = cases.advection_diffusion(be, data, scheme=scheme, kappa=0.05)
        c["eqn"].update_coeffs()
        rp, col, val, n = c["eqn"].matrix()
        b = c["eqn"].rhs()
        c["eqn"].zero_out_coeffs()
        P.solver.BiCGSTAB(tol=1e-13).solve(c["eqn"], 2, 1e-20)
        out[name] = (rp, col, val, b, c["T"].values)
    (rh, ch, vh, bh, th), (rr, cr, vr, br, tr) = out["hip"], out["ref"]
    np.testing.assert_array_equal(rh, rr)
    np.testing.assert_array_equal(ch, cr)
    assert rel(vh, vr) < 1e-13
    assert rel(bh, br) < 1e-13
    assert rel(th, tr) < 1e-10

