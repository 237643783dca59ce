"""More pins of the FP64 oracle that the mathematics fixes, for the parts whose earlier pins
left a plausible mistake open:

* Green-Gauss with the option-2 skewness correction (green_gauss.cpp:11-24) is a stateful
  sequential sweep (green_gauss.cpp:34-76): cell c uses the gradients already updated in
  this pass for lower-numbered neighbours and the previous pass's for the rest.  On a
  skewed mesh with planar faces and exact linear boundary values, the exact gradient is a
  fixed point of that sweep; the first pass from the zero cache is reproduced by a
  brute-force replay of the sweep, and it changes when the cells are renumbered while the
  fixed point does not.  An all-old (Jacobi) sweep, a sign slip in the skew vector or a
  missing half fails one of these.
* The minimum and orthogonal non-orthogonal correctors (nonortho.cpp:13-21): the
  off-diagonal coefficient of each decomposition in closed form from the geometry, and the
  exact balance of a linear field that any E + T = S decomposition gives.
* The fixed-gradient diffusion boundary (diffusion_boundary.h:114-150): the one-dimensional
  closed form of the reference's sign convention.
* The Courant number (courant.cpp:7-24): uniform flow gives dt u / h in every cell.
"""
import numpy as np
import pytest

import prism_amd as P

SIDES = ("left", "right", "bottom", "top", "back", "front")
SHEAR = np.array([[1.0, 0.35, -0.2], [0.1, 0.9, 0.25], [0.0, -0.3, 1.2]])
GRAD = np.array([1.3, -2.1, 0.4])


def linear(p):
    return 0.7 + p @ GRAD


def skewed_box(nx, ny, nz, tets=False):
    """A box whose grid lines are stretched per axis (x + 0.6 x^2, ...) and then sheared:
    every face stays planar (a parallelogram, or a triangle with tets=True), and neighbouring
    cells differ in size, so the face centre is not the midpoint of the two cell centres (a
    nonzero skew vector)."""
    d = P.box(nx, ny, nz, tets=tets)
    v = d.vertices
    v = v + np.array([0.6, -0.4, 0.5]) * v ** 2
    d.vertices = v @ SHEAR.T
    return d


def renumbered(d, perm):
    """The same mesh with cell perm[i] renamed i.  Faces keep their order; a face whose owner
    gets the higher number swaps owner and neighbour and reverses its vertex loop, so its
    area vector still points from the owner to the neighbour."""
    n = d.n_cells
    new_of = np.empty(n, np.int64)
    new_of[perm] = np.arange(n)
    own = new_of[d.face_owner]
    nbr = np.where(d.face_neighbor >= 0, new_of[np.maximum(d.face_neighbor, 0)], -1)
    fv = d.face_vtx.copy()
    swap = (nbr >= 0) & (nbr < own)
    for f in np.nonzero(swap)[0]:
        a, b = d.face_vtx_off[f], d.face_vtx_off[f + 1]
        fv[a:b] = d.face_vtx[a:b][::-1]
    own2 = np.where(swap, nbr, own)
    nbr2 = np.where(swap, own, nbr)
    lists = [d.cell_faces[d.cell_face_off[c]:d.cell_face_off[c + 1]] for c in perm]
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(x) for x in lists])
    return P.MeshData(vertices=d.vertices.copy(), face_vtx_off=d.face_vtx_off.copy(), face_vtx=fv,
                      face_owner=own2, face_neighbor=nbr2, face_patch=d.face_patch.copy(),
                      cell_face_off=off, cell_faces=np.concatenate(lists), patch_names=list(d.patch_names))


def gg_mesh(be, data):
    bcs = {s: {"T": ("fixed", 0.0)} for s in data.patch_names}
    return P.Mesh(be, data, bcs, [("T", "scalar", "greenGauss")])


def linear_field(m):
    g = m.geometry()
    T = m.scalar("T", values=linear(g["cell_center"]))
    T.set_face_values(linear(g["face_center"]))
    return T


def gg_passes(m, npass):
    """Gradients after each of `npass` passes; every pass is triggered by rewriting the same
    cell values (a new event), so it starts from the previous pass's cache."""
    T = linear_field(m)
    vals = T.values
    out = []
    for _ in range(npass):
        T.values = vals
        out.append(T.grad())
    return out


def gg_replay(g, phi, phib, old=False):
    """Brute-force first pass of the sweep from a zero cache, cell by cell in index order:
    face value 1/2 (phi_C + phi_N) + 1/2 (G_C + G_N) . (x_f - 1/2 (x_C + x_N)) on interior
    faces, the boundary value on boundary faces.  old=True reads only the previous pass's
    (zero) gradients: a Jacobi sweep, which the reference's is not."""
    n = len(g["cell_volume"])
    c, xf, S = g["cell_center"], g["face_center"], g["face_area"]
    own, nbr = g["owner"], g["neighbor"]
    faces = [[] for _ in range(n)]
    for f in range(len(own)):
        faces[own[f]].append(f)
        if nbr[f] >= 0:
            faces[nbr[f]].append(f)
    G = np.zeros((n, 3))
    prev = np.zeros((n, 3))
    for i in range(n):
        acc = np.zeros(3)
        for f in faces[i]:
            if nbr[f] < 0:
                acc += phib[f] * S[f]
                continue
            j = nbr[f] if own[f] == i else own[f]
            Sf = S[f] if own[f] == i else -S[f]
            src = prev if old else G
            val = 0.5 * (phi[i] + phi[j]) + 0.5 * (src[i] + src[j]) @ (xf[f] - 0.5 * (c[i] + c[j]))
            acc += Sf * val
        G[i] = acc / g["cell_volume"][i]
    return G


@pytest.mark.parametrize("tets,npass", [(False, 16), (True, 40)])
def test_green_gauss_converges_to_exact_gradient(ref, tets, npass):
    """Passes contract towards the exact gradient on these meshes (on strongly jittered tets
    the sweep's iteration is not contractive, so they are not used here)."""
    m = gg_mesh(ref, skewed_box(4, 3, 3, tets=tets))
    passes = gg_passes(m, npass)
    err = [np.abs(G - GRAD).max() for G in passes]
    assert err[0] > 1e-2          # the first pass (zero cache) is not exact on this mesh
    assert err[-1] < 1e-12        # the exact gradient is the sweep's fixed point
    assert all(b < 0.5 * a for a, b in zip(err[:-8:4], err[4:-4:4]))


def test_green_gauss_first_pass_is_the_sequential_sweep(ref):
    data = skewed_box(3, 3, 2)
    m = gg_mesh(ref, data)
    g = m.geometry()
    first = gg_passes(m, 1)[0]
    phi = linear(g["cell_center"])
    phib = linear(g["face_center"])
    seq = gg_replay(g, phi, phib)
    jac = gg_replay(g, phi, phib, old=True)
    assert np.abs(first - seq).max() < 1e-12
    assert np.abs(first - jac).max() > 1e-3  # the sweep order matters on this mesh
    # cell 0 sees only the zero cache: the uncorrected Green-Gauss sum, written out directly
    own, nbr = g["owner"], g["neighbor"]
    acc = np.zeros(3)
    for f in range(len(own)):
        if own[f] != 0 and nbr[f] != 0:
            continue
        sgn = 1.0 if own[f] == 0 else -1.0
        val = phib[f] if nbr[f] < 0 else 0.5 * (phi[own[f]] + phi[nbr[f]])
        acc += sgn * val * g["face_area"][f]
    assert np.abs(first[0] - acc / g["cell_volume"][0]).max() < 1e-12


def test_green_gauss_renumbering(ref):
    """Renumbering the cells changes the first pass (the sweep is order dependent) but not
    the fixed point."""
    data = skewed_box(3, 3, 2)
    perm = np.random.default_rng(2).permutation(data.n_cells)
    m1 = gg_mesh(ref, data)
    m2 = gg_mesh(ref, renumbered(data, perm))
    g1, g2 = m1.geometry(), m2.geometry()
    assert np.abs(g2["cell_center"] - g1["cell_center"][perm]).max() < 1e-14
    assert np.abs(g2["cell_volume"] - g1["cell_volume"][perm]).max() < 1e-15
    p1, p2 = gg_passes(m1, 16), gg_passes(m2, 16)
    assert np.abs(p2[0] - p1[0][perm]).max() > 1e-3
    assert np.abs(p2[-1] - p1[-1][perm]).max() < 1e-12
    # the renumbered first pass is the sequential sweep in the new order
    seq2 = gg_replay(g2, linear(g2["cell_center"]), linear(g2["face_center"]))
    assert np.abs(p2[0] - seq2).max() < 1e-12


# ------------------------------------------------------------------------ correctors ----
def corrector_eqn(be, data, corrector, kappa=0.3):
    bcs = {s: {"T": ("fixed", 0.0)} for s in data.patch_names}
    m = P.Mesh(be, data, bcs, [("T", "scalar", "leastSquares")])
    g = m.geometry()
    T = m.scalar("T", values=linear(g["cell_center"]))
    T.set_face_values(linear(g["face_center"]))
    k = m.scalar("k", kappa)
    return m, T, P.Transport(P.diffusion.Corrected(k, T, corrector=corrector))


@pytest.mark.parametrize("corrector", ["over-relaxed", "minimum", "orthogonal"])
def test_corrector_coefficients_closed_form(ref, corrector):
    """Off-diagonal -kappa |E|/|d| with E = (e.S) e (minimum), |S| e (orthogonal) or
    |S|^2/(e.S) e (over-relaxed), e = d/|d| (Moukalled et al. 2016, sec. 8.6)."""
    kappa = 0.3
    m, T, eqn = corrector_eqn(ref, skewed_box(4, 3, 3), corrector, kappa)
    eqn.update_coeffs()
    A = eqn.dense()
    g = m.geometry()
    own, nbr, S, c = g["owner"], g["neighbor"], g["face_area"], g["cell_center"]
    inner = np.nonzero(nbr >= 0)[0]
    d = c[nbr[inner]] - c[own[inner]]
    dn = np.linalg.norm(d, axis=1)
    e = d / dn[:, None]
    Sn = np.linalg.norm(S[inner], axis=1)
    eS = np.einsum("ij,ij->i", e, S[inner])
    E = {"minimum": eS, "orthogonal": Sn, "over-relaxed": Sn ** 2 / eS}[corrector]
    expect = -kappa * E / dn
    got = A[own[inner], nbr[inner]]
    assert np.abs(got - expect).max() < 1e-13 * np.abs(expect).max()
    assert np.abs(A - A.T).max() < 1e-14
    # the three decompositions differ on this mesh (|E| strictly ordered)
    assert np.all(eS < Sn * (1 - 1e-6))


@pytest.mark.parametrize("corrector", ["minimum", "orthogonal"])
@pytest.mark.parametrize("kind", ["skewed", "tets"])
def test_corrector_flux_exact_for_linear_field(ref, corrector, kind):
    """Any E + T = S split with E || d gives the flux S . grad(phi) for linear phi and exact
    gradients, so every cell's balance of div(kappa grad phi) vanishes."""
    data = skewed_box(4, 3, 3) if kind == "skewed" else P.box(4, 4, 3, jitter=0.25, seed=2, tets=True)
    m, T, eqn = corrector_eqn(ref, data, corrector)
    eqn.update_coeffs()
    r = eqn.dense() @ T.values - eqn.rhs()
    assert np.abs(r).max() < 1e-12


# -------------------------------------------------------------------- fixed gradient ----
def fixed_gradient_case(be, diff, gx=0.8):
    data = P.box(16, 1, 1)
    bcs = {"left": {"T": ("fixed", 1.0)}, "right": {"T": ("fixed-gradient", (gx, 0.0, 0.0))}}
    bcs.update({s: {"T": "empty"} for s in SIDES[2:]})
    m = P.Mesh(be, data, bcs, [("T", "scalar", None)])
    T = m.scalar("T", 0.0)
    k = m.scalar("k", 0.7)
    return m, T, P.Transport(getattr(P.diffusion, diff)(k, T))


def test_fixed_gradient_boundary_closed_form(ref):
    """-kappa T'' = 0 on [0, 1] with T = 1 at x = 0 and a fixed-gradient patch at x = 1:
    the reference adds -g . (kappa S) to the right-hand side (diffusion_boundary.h:114-150),
    so the discrete solution is the line T = 1 - g_x x, exact at the cell centres."""
    gx = 0.8
    m, T, eqn = fixed_gradient_case(ref, "NonCorrected", gx)
    P.solver.BiCGSTAB().solve(eqn, 3, 1e-20)
    x = m.geometry()["cell_center"][:, 0]
    assert np.abs(T.values - (1.0 - gx * x)).max() < 1e-12


def test_fixed_gradient_corrected_needs_a_face_value(ref):
    """Corrected diffusion evaluates the field on boundary faces for its gradient, and a
    Scalar has no fixed-gradient face-value handler (scalar_boundary.h:53-77), so the
    reference throws NonImplementedBoundaryCondition there."""
    m, T, eqn = fixed_gradient_case(ref, "Corrected")
    with pytest.raises(P.BoundaryConditionError):
        P.solver.BiCGSTAB().solve(eqn, 3, 1e-20)


# --------------------------------------------------------------------------- Courant ----
def test_courant_uniform_flow(ref):
    """Face fluxes u |S| on the x-faces, zero elsewhere: each cell has two faces of flux
    u h^2 (its boundary face counts too), so Co = 0.5 dt 2 u h^2 / h^3 = dt u / h."""
    n, u, dt = 6, 1.7, 0.013
    data = P.box(n, n, n)
    bcs = {s: {"U": ("fixed", (u, 0.0, 0.0))} for s in SIDES}
    m = P.Mesh(ref, data, bcs, [("U", "velocity", None)])
    g = m.geometry()
    U = m.vector("U", (u, 0.0, 0.0))
    U.set_face_values(np.tile([u, 0.0, 0.0], (m.n_faces, 1)))
    U.set_face_fluxes(u * g["face_area"][:, 0])
    co = P.courant_number(U, dt)
    assert np.abs(co - dt * u * n).max() < 1e-13


def test_courant_skips_empty_faces(ref):
    """A 2-D mesh (front/back empty) with fluxes on every face: the empty faces' fluxes are
    not counted (courant.cpp:14-16); the others are summed in absolute value."""
    nx, ny, dt = 5, 4, 0.02
    data = P.box(nx, ny, 1, z=(0, 0.25))
    bcs = {s: {"U": ("fixed", (0.0, 0.0, 0.0))} for s in SIDES[:4]}
    bcs.update({s: {"U": "empty"} for s in SIDES[4:]})
    m = P.Mesh(ref, data, bcs, [("U", "velocity", None)])
    g = m.geometry()
    F = np.random.default_rng(3).normal(size=m.n_faces)
    U = m.vector("U", (0.0, 0.0, 0.0))
    U.set_face_values(np.zeros((m.n_faces, 3)))
    U.set_face_fluxes(F)
    co = P.courant_number(U, dt)
    names = data.patch_names
    empty = np.isin(g["face_patch"], [names.index("back"), names.index("front")]) & (g["neighbor"] < 0)
    acc = np.zeros(m.n_cells)
    keep = ~empty
    np.add.at(acc, g["owner"][keep], np.abs(F[keep]))
    inner = keep & (g["neighbor"] >= 0)
    np.add.at(acc, g["neighbor"][inner], np.abs(F[inner]))
    assert np.abs(co - 0.5 * dt * acc / g["cell_volume"]).max() < 1e-13
