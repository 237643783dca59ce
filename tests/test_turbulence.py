"""k-epsilon (KEpsilon.cpp, k_production.cpp) on the CPU oracle.  The per-cell fields are pinned
by what the mathematics fixes on linear velocity fields, whose least-squares gradients are
exact: simple shear gives Pk = nu_t gamma^2, a rotation gives Pk = 0, uniaxial stretching
2 a^2 nu_t - (2/3) k a (the only case where the k I term contributes), and a general
gradient nu_t sum_ij (G_ij + G_ji) G_ij - (2/3) k tr G; nu_t = c_mu k^2 / eps and the epsilon
source and sink follow from them.  The reference ships no turbulence test; its two
equations are built from schemes pinned elsewhere, and a short transient run checks that
the model steps and stays finite."""
import numpy as np
import pytest

import oracle
import prism_amd as P
from prism_amd.turbulence import IncompressibleKEpsilon

C_MU, C1, C2, SK, SE = 0.09, 1.44, 1.92, 1.0, 1.3


def _mesh(be, data):
    bcs = {}
    for p in data.patch_names:
        bcs[p] = {"U": ("fixed", (0.0, 0.0, 0.0)), "k": "zero-gradient", "epsilon": "zero-gradient",
                  "nu": "zero-gradient"}
    bcs["left"].update({"U": ("fixed", (1.0, 0.0, 0.0)), "k": ("fixed", 0.01), "epsilon": ("fixed", 1e-3)})
    bcs["right"].update({"U": "outlet"})
    return P.Mesh(be, data, bcs, [("U", "vector"), ("k", "scalar"), ("epsilon", "scalar"), ("nu", "scalar")])


def _linear_U(m, G):
    """U_j(x) = sum_i G[i, j] x_i, so that grad U (as gradAtCell) is G; exact face values on
    every face make the least-squares gradient exact."""
    g = m.geometry()
    U = m.vector("U", values=g["cell_center"] @ G)
    for j, comp in enumerate((U.x, U.y, U.z)):
        comp.set_face_values(g["face_center"] @ G[:, j])
    return U


@pytest.fixture(scope="module")
def be():
    return oracle.backend()


@pytest.fixture(scope="module")
def mesh(be):
    return _mesh(be, P.box(6, 5, 4, jitter=0.25, seed=8, tets=True))


def _model(m, G, seed=0):
    U = _linear_U(m, G)
    nu = m.scalar("nu", 1e-3)
    model = IncompressibleKEpsilon(U, U, nu)
    rng = np.random.default_rng(seed)
    k = rng.uniform(0.01, 0.1, m.n_cells)
    eps = rng.uniform(0.001, 0.02, m.n_cells)
    model.k().values = k
    model.epsilon().values = eps
    model.update_fields()
    return model, k, eps


def test_nu_t_and_shear_production(mesh):
    gamma = 2.5
    G = np.zeros((3, 3))
    G[1, 0] = gamma  # dUx/dy
    model, k, eps = _model(mesh, G)
    nut = C_MU * (k * k / eps)
    np.testing.assert_allclose(model.turbulent_viscosity().values, nut, rtol=1e-15)
    np.testing.assert_allclose(model.production().values, nut * gamma ** 2, rtol=1e-11)


def test_rotation_produces_nothing(mesh):
    w = 1.7
    G = np.zeros((3, 3))
    G[1, 0], G[0, 1] = -w, w  # U = (-w y, w x, 0)
    model, k, eps = _model(mesh, G, seed=1)
    nut = C_MU * (k * k / eps)
    assert np.max(np.abs(model.production().values)) <= 1e-11 * np.max(nut) * w * w


def test_uniaxial_stretch_uses_the_k_term(mesh):
    a = 0.8
    G = np.zeros((3, 3))
    G[0, 0] = a  # U = (a x, 0, 0)
    model, k, eps = _model(mesh, G, seed=2)
    nut = C_MU * (k * k / eps)
    np.testing.assert_allclose(model.production().values, 2 * a * a * nut - (2.0 / 3.0) * k * a, rtol=1e-10,
                               atol=1e-13)


def test_general_gradient_and_derived_fields(mesh):
    G = np.random.default_rng(4).standard_normal((3, 3))
    model, k, eps = _model(mesh, G, seed=3)
    nut = C_MU * (k * k / eps)
    pk = nut * np.einsum("ij,ij->", G + G.T, G) - (2.0 / 3.0) * k * np.trace(G)
    got = model.production().values
    np.testing.assert_allclose(got, pk, rtol=1e-10, atol=1e-12 * np.abs(pk).max())
    np.testing.assert_allclose(model._nu_eff_k.values, 1e-3 + nut / SK, rtol=1e-15)
    np.testing.assert_allclose(model._nu_eff_eps.values, 1e-3 + nut / SE, rtol=1e-15)
    np.testing.assert_allclose(model._eps_source.values, C1 * (eps / k) * got, rtol=1e-14)
    np.testing.assert_allclose(model._eps_sink.values, C2 * eps * eps / k, rtol=1e-14)


def test_update_rejects_aliases_and_foreign_meshes(be, mesh):
    model, _, _ = _model(mesh, np.eye(3))
    import ctypes as C
    from prism_amd.core import _dp
    outs = [model._nu_t, model._pk, model._nu_eff_k, model._nu_eff_eps, model._eps_source, model.k()]
    arr = (C.c_void_p * 6)(*[f.h.value for f in outs])
    consts = np.array([C_MU, C1, C2, SK, SE])
    with pytest.raises(ValueError, match="also an input"):
        be("kepsilon_update", model.k().h, model.epsilon().h, model._nu.h, model._U.h, _dp(consts), arr)
    other = _mesh(be, P.box(6, 5, 4, jitter=0.25, seed=8, tets=True))
    nu2 = other.scalar("nu", 1e-3)
    with pytest.raises(P.InvalidMeshError):
        IncompressibleKEpsilon(model._U, model._U, nu2).update_fields()


def run_channel(be, steps=4, dt=0.05, n=(8, 4, 4)):
    """A transient k-epsilon run in a box channel: U from a sheared inflow profile, k and
    epsilon fixed at the inlet; each step rotates both histories, then solves the k and
    the epsilon equation (BiCGSTAB, FP64).  Returns the model and the per-step
    solver results."""
    m = _mesh(be, P.box(*n, x=(0.0, 2.0), jitter=0.15, seed=2))
    g = m.geometry()["cell_center"]
    U = m.vector("U", values=np.stack([0.5 + g[:, 1], 0.1 * np.sin(np.pi * g[:, 0]), 0.0 * g[:, 2]], axis=1))
    nu = m.scalar("nu", 1e-3)
    model = IncompressibleKEpsilon(U, U, nu, dt=dt)
    model.k().fill(0.01)
    model.epsilon().fill(1e-3)
    slv = P.solver.BiCGSTAB(tol=1e-13)
    reports = []
    for _ in range(steps):
        model.k().update_prev_time_steps()
        model.epsilon().update_prev_time_steps()
        ke, ee = model.equations()
        reports.append(slv.solve(ke, n_iter=2, eps=1e-20))
        reports.append(slv.solve(ee, n_iter=2, eps=1e-20))
    return model, reports


def test_transient_channel_steps(be):
    model, reports = run_channel(be)
    k, e, nut = model.k().values, model.epsilon().values, model.turbulent_viscosity().values
    assert np.all(np.isfinite(k)) and np.all(np.isfinite(e)) and np.all(np.isfinite(nut))
    assert np.all(k > 0) and np.all(e > 0)
    # the inflow values are held at the inlet and the fields moved away from their start
    assert not np.allclose(k, 0.01) and not np.allclose(e, 1e-3)
    assert all(r.converged or r.final_residual < 1e-10 for r in reports)
