"""Problem set-ups of the benchmark workloads, built through the reference-style API.

Each builder takes a backend and a MeshData (or sizes), creates the fields, boundary
conditions and equations exactly as a Prism user would (fields.json-equivalent BCs,
eqn::Transport of convection/diffusion/source schemes), and returns them in a dict.
"""
from __future__ import annotations

import numpy as np

from . import core as P

SIDES = ("left", "right", "bottom", "top", "back", "front")


def _face_eval(mesh, fn):
    g = mesh.geometry()
    return fn(g["face_center"][:, 0], g["face_center"][:, 1], g["face_center"][:, 2])


def poisson(be, n=128, kappa_tensor=True):
    """Steady Poisson -div(kappa grad T) = f on the unit square, n x n x 1 hexahedra one cell
    thick (front/back empty), kappa = I, f = 2 pi^2 sin(pi x) cos(pi y), fixed T = sin(pi x)
    cos(pi y) on the four sides (per-face values), diffusion::Corrected, Green-Gauss gradient."""
    data = P.box(n, n, 1, z=(0.0, 1.0 / n))
    bcs = {s: {"T": ("fixed", 0.0), "f": ("fixed", 0.0)} for s in SIDES[:4]}
    bcs.update({s: {"T": "empty", "f": "empty"} for s in SIDES[4:]})
    mesh = P.Mesh(be, data, bcs, [("T", "scalar", "greenGauss"), ("f", "scalar", None)])
    g = mesh.geometry()
    x, y = g["cell_center"][:, 0], g["cell_center"][:, 1]
    T = mesh.scalar("T", 0.0)
    T.set_face_values(_face_eval(mesh, lambda fx, fy, fz: np.sin(np.pi * fx) * np.cos(np.pi * fy)))
    f = mesh.scalar("f", values=2 * np.pi ** 2 * np.sin(np.pi * x) * np.cos(np.pi * y))
    kappa = mesh.tensor("kappa", np.eye(3)) if kappa_tensor else mesh.scalar("kappa", 1.0)
    eqn = P.Transport(P.diffusion.Corrected(kappa, T), P.source.ConstantScalar(f))
    exact = np.sin(np.pi * x) * np.cos(np.pi * y)
    return dict(mesh=mesh, T=T, f=f, kappa=kappa, eqn=eqn, exact=exact)


def velocity_field(x, y, z):
    """Prescribed solenoidal U = (sin pi x cos pi y, -cos pi x sin pi y, 0.3)."""
    return np.stack([np.sin(np.pi * x) * np.cos(np.pi * y), -np.cos(np.pi * x) * np.sin(np.pi * y),
                     np.full_like(x, 0.3)], axis=1)


def prescribed_velocity(mesh, name="U"):
    g = mesh.geometry()
    c = g["cell_center"]
    U = mesh.vector(name, values=velocity_field(c[:, 0], c[:, 1], c[:, 2]))
    fc = g["face_center"]
    uf = velocity_field(fc[:, 0], fc[:, 1], fc[:, 2])
    S = g["face_area"]
    U.set_face_values(uf)
    U.set_face_fluxes((uf[:, 0] * S[:, 0] + uf[:, 1] * S[:, 1]) + uf[:, 2] * S[:, 2])
    return U


def advection_diffusion(be, data, scheme="MUSCL", kappa=1e-3, diffusion="Corrected", grad=None):
    """Steady div(U T) - div(kappa grad T) = 0 with the prescribed U, T = 1 on x = 0 and
    zero-gradient elsewhere, NVF convection `scheme`, over-relaxed corrected diffusion."""
    bcs = {s: {"T": ("zero-gradient", None), "U": ("fixed", (0.0, 0.0, 0.0))} for s in SIDES}
    bcs["left"]["T"] = ("fixed", 1.0)
    mesh = P.Mesh(be, data, bcs, [("T", "scalar", grad), ("U", "velocity", None)])
    U = prescribed_velocity(mesh)
    T = mesh.scalar("T", 0.0)
    k = mesh.scalar("kappa", kappa) if np.ndim(kappa) == 0 else mesh.scalar("kappa", values=kappa)
    conv = getattr(P.convection, scheme)(U, T)
    diff = getattr(P.diffusion, diffusion)(k, T)
    eqn = P.Transport(conv, diff)
    return dict(mesh=mesh, T=T, U=U, kappa=k, eqn=eqn)


def cavity(be, n=16, nu=0.01, scheme="SMART", lid=1.0, data=None):
    """Lid-driven cubic cavity: lid (top, y = 1) moving at (lid, 0, 0), no-slip elsewhere,
    zero-gradient pressure everywhere (all-Neumann: reference cell 0), three momentum
    equations div(mdot u) - div(nu grad u) = -dp/dx_i."""
    data = data if data is not None else P.box(n, n, n)
    bcs = {s: {"U": ("no-slip", (0.0, 0.0, 0.0)), "P": ("zero-gradient", None)} for s in SIDES}
    bcs["top"]["U"] = ("fixed", (lid, 0.0, 0.0))
    mesh = P.Mesh(be, data, bcs, [("U", "velocity", None), ("P", "pressure", None)])
    U = mesh.vector("U", (0.0, 0.0, 0.0))
    p = mesh.pressure("P", 0.0)
    nuf = mesh.scalar("nu", nu)
    mdot = U.clone()
    conv = getattr(P.convection, scheme)
    eqns = [P.Transport(conv(mdot, c), P.diffusion.Corrected(nuf, c), P.source.Gradient(p, k, "-"))
            for k, c in enumerate((U.x, U.y, U.z))]
    return dict(mesh=mesh, U=U, p=p, nu=nuf, mdot=mdot, eqns=eqns)


def pressure_correction(be, n=128, nu=0.01, data=None, seed=5):
    """The pressure-correction system of a SIMPLE step on the n^3 cubic cavity mesh
    (SIMPLE.cpp:124-196): Corrected(D, p') = Divergence<Negative>(mdot) with zero-gradient p'
    on every wall (all-Neumann: reference cell 0).  D = diag(V / a_P) with the diagonal
    a_P = 6 nu h of the viscous momentum operator on the uniform mesh.  mdot carries the
    mass imbalance that p' corrects: seeded random interior face fluxes (|S| N(0, 1)) and no
    flux through the walls, so the right-hand side is compatible and rough, as a SIMPLE
    iterate's is (a smooth imbalance would be an eigenvector of the discrete Laplacian)."""
    data = data if data is not None else P.box(n, n, n)
    bcs = {s: {"P": ("zero-gradient", None), "U": ("no-slip", (0.0, 0.0, 0.0))} for s in SIDES}
    mesh = P.Mesh(be, data, bcs, [("P", "pressure", None), ("U", "velocity", None)])
    g = mesh.geometry()
    h = 1.0 / n
    d = g["cell_volume"] / (6 * nu * h)
    Dt = np.zeros((mesh.n_cells, 3, 3))
    Dt[:, 0, 0] = Dt[:, 1, 1] = Dt[:, 2, 2] = d
    D = mesh.tensor("D", values=Dt)
    mdot = mesh.vector("U", (0.0, 0.0, 0.0))
    area = np.linalg.norm(g["face_area"], axis=1)
    flux = area * np.random.default_rng(seed).standard_normal(len(area))
    flux[g["neighbor"] < 0] = 0.0
    mdot.set_face_values(np.zeros((len(area), 3)))
    mdot.set_face_fluxes(flux)
    pp = mesh.pressure("P", 0.0)
    eqn = P.Transport(P.diffusion.Corrected(D, pp), P.source.Divergence(mdot, "-"))
    return dict(mesh=mesh, p=pp, D=D, mdot=mdot, eqn=eqn)


def solenoidal_noise(x, y, z, sigma, seed, kmax=2):
    """Seeded Gaussian noise projected divergence-free on the 2 pi-periodic box, sampled at the
    points (x, y, z).  Each Fourier mode k with integer components |k_i| <= kmax (one of each
    pair +-k) gets a complex N(0, 1) coefficient vector a_k, projected onto the plane normal to
    k (a_k - (k . a_k) k / |k|^2, the Leray projection); the field is
    u = sum_k Re(a_k exp(i k . x)), so div u = 0 exactly at every point, scaled by one factor
    so that its RMS per component over the box (sum_k |a_k|^2 / 6) is `sigma`."""
    rng = np.random.default_rng(seed)
    r = range(-kmax, kmax + 1)
    ks = [(a, b, c) for a in r for b in r for c in r if (a, b, c) > (0, 0, 0)]  # half of the +-k pairs
    coef = rng.standard_normal((len(ks), 3)) + 1j * rng.standard_normal((len(ks), 3))
    ex = {k: np.exp(1j * k * x) for k in r}
    ey = {k: np.exp(1j * k * y) for k in r}
    ez = {k: np.exp(1j * k * z) for k in r}
    u = np.zeros((len(x), 3))
    power = 0.0
    for (a, b, c), v in zip(ks, coef):
        kv = np.array([a, b, c], dtype=float)
        v = v - (kv @ v) * kv / (kv @ kv)
        power += float(np.sum(np.abs(v) ** 2)) / 2.0
        e = ex[a] * ey[b] * ez[c]
        for i in range(3):
            u[:, i] += (v[i] * e).real
    return u * (sigma / np.sqrt(power / 3.0))


def courant_dt(mdot, co_max=0.5):
    """The time step at which the largest cell Courant number of monitor::courantNumber
    (courant.cpp:7-24, linear in dt) equals co_max."""
    return co_max / float(P.courant_number(mdot, 1.0).max())


def taylor_green(be, n=8, nu=1e-3, noise=1e-3, seed=4, tets=True, dt=None, co_max=0.5):
    """Transient Taylor-Green vortex on [0, 2 pi]^3 split into tetrahedra: symmetry planes for
    U (the Taylor-Green field satisfies them), zero-gradient pressure, backward Euler.  The
    initial U adds seeded divergence-free noise of RMS `noise` (solenoidal_noise); dt (unless
    given) puts the largest cell Courant number of the initial face fluxes at co_max."""
    L = 2 * np.pi
    data = P.box(n, n, n, x=(0, L), y=(0, L), z=(0, L), tets=tets)
    bcs = {s: {"U": ("symmetry", None), "P": ("zero-gradient", None)} for s in SIDES}
    mesh = P.Mesh(be, data, bcs, [("U", "velocity", None), ("P", "pressure", None)])
    c = mesh.geometry()["cell_center"]
    x, y, z = c[:, 0], c[:, 1], c[:, 2]
    u0 = np.stack([np.sin(x) * np.cos(y) * np.cos(z), -np.cos(x) * np.sin(y) * np.cos(z), np.zeros_like(x)], 1)
    if noise:
        u0 = u0 + solenoidal_noise(x, y, z, noise, seed)
    U = mesh.vector("U", values=u0)
    p = mesh.pressure("P", 0.0)
    nuf = mesh.scalar("nu", nu)
    mdot = U.clone()
    if dt is None:
        dt = courant_dt(mdot, co_max)
    for comp in (U.x, U.y, U.z):
        comp.set_history_size(1)
        comp.update_prev_time_steps()
    eqns = [P.Transport(P.temporal.BackwardEuler(cmp, dt), P.convection.LinearUpwind(mdot, cmp),
                        P.diffusion.Corrected(nuf, cmp), P.source.Gradient(p, k, "-"))
            for k, cmp in enumerate((U.x, U.y, U.z))]
    return dict(mesh=mesh, U=U, p=p, nu=nuf, mdot=mdot, eqns=eqns, dt=dt)


def adjoint_transport(be, data, kappa_seed=6, source_seed=7, corrector=None):
    """Steady UPWIND advection-diffusion with the prescribed U, log-normal kappa per cell
    (ln kappa ~ N(ln 1e-2, 0.5^2)), source s ~ U(0, 1) per cell, T = 1 on x = 0,
    zero-gradient elsewhere; non-corrected diffusion, or Corrected with `corrector`
    ("over-relaxed", "minimum", "orthogonal") and least-squares gradients."""
    bcs = {s: {"T": ("zero-gradient", None), "U": ("fixed", (0.0, 0.0, 0.0)),
               "s": ("zero-gradient", None), "kappa": ("zero-gradient", None)} for s in SIDES}
    bcs["left"]["T"] = ("fixed", 1.0)
    mesh = P.Mesh(be, data, bcs, [("T", "scalar", None)])
    n = mesh.n_cells
    kap = np.exp(np.log(1e-2) + 0.5 * np.random.default_rng(kappa_seed).standard_normal(n))
    src = np.random.default_rng(source_seed).uniform(0.0, 1.0, n)
    U = prescribed_velocity(mesh)
    T = mesh.scalar("T", 0.0)
    k = mesh.scalar("kappa", values=kap)
    s = mesh.scalar("s", values=src)
    diff = P.diffusion.NonCorrected(k, T) if corrector is None else P.diffusion.Corrected(k, T, corrector=corrector)
    eqn = P.Transport(P.convection.Upwind(U, T), diff, P.source.ConstantScalar(s))
    return dict(mesh=mesh, T=T, U=U, kappa=k, s=s, eqn=eqn)
