"""The pressure-correction system of a SIMPLE step, handed to the caller before its solve
(prism_simple_pressure_system; SIMPLE.cpp:19-178).  On the oracle (CPU): solving the handed
system and applying the step's pressure update p + 0.3 p' reproduces what a full
IncompressibleSIMPLE step leaves in p, so the system is the one the step solves; the
all-Neumann structure (symmetric, zero row sums, compatible right-hand side) holds.  On the
GPU: the HIP library's system equals the oracle's."""
import numpy as np
import pytest
import scipy.sparse as sp

import prism_amd as P
from prism_amd import cases


def simple_ctl(be):
    return P.algo.SIMPLEControls(be, momentum_tol=1e-13, pressure_tol=1e-13, max_krylov_iters=20000)


def captured(be, n, outer, scheme="SMART"):
    c = cases.cavity(be, n=n, scheme=scheme)
    S = P.algo.IncompressibleSIMPLE(simple_ctl(be))
    for _ in range(outer):
        S.step(c["eqns"], c["U"], c["mdot"], c["p"])
    p_before = c["p"].values
    pe, pp, ref, rep = S.pressure_system(c["eqns"], c["U"], c["mdot"], c["p"])
    return c, S, pe, pp, ref, rep, p_before


def test_pressure_system_is_the_steps_system(ref):
    n, outer = 6, 2
    c, S, pe, pp, refc, rep, p_before = captured(ref, n, outer)
    assert refc == 0  # all-Neumann cavity: automatic reference cell 0 (SIMPLE.cpp:110-122)
    assert [r["field_name"] for r in rep] == ["U_x", "U_y", "U_z"]
    ctl = simple_ctl(ref)
    slv = P.solver.BiCGSTAB(tol=1e-13, max_iters=20000)
    for _ in range(ctl.non_ortho_correctors + 1):
        slv.solve(pe, ctl.pressure_max_iter, ctl.pressure_residual, ref_cell=refc)
    p_new = p_before + ctl.pressure_urf * pp.values
    # the same outer iteration as a full step on a second instance
    c2 = cases.cavity(ref, n=n)
    S2 = P.algo.IncompressibleSIMPLE(simple_ctl(ref))
    for _ in range(outer + 1):
        S2.step(c2["eqns"], c2["U"], c2["mdot"], c2["p"])
    assert np.abs(p_new - c2["p"].values).max() <= 1e-12 * max(1.0, np.abs(p_new).max())


def test_pressure_system_structure(ref):
    c, S, pe, pp, refc, rep, _ = captured(ref, 7, 3)
    pe.update_coeffs()
    rp, col, val, n = pe.matrix()
    b = pe.rhs()
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    d = abs(A.diagonal()).max()
    assert abs(A - A.T).max() <= 1e-14 * d  # uniform orthogonal mesh: symmetric
    assert np.abs(np.asarray(A.sum(axis=1))).max() <= 1e-14 * d  # zero-gradient everywhere
    assert abs(b.sum()) <= 1e-12 * np.abs(b).sum()  # closed cavity: compatible


@pytest.mark.gpu
@pytest.mark.parametrize("scheme,outer", [("SMART", 0), ("Upwind", 3)])
def test_gpu_pressure_system_against_oracle(hip, ref, scheme, outer):
    """SMART switches a face between pure upwind and pure downwind on the sign of a normalised
    value (convection.cpp:113-125), so two runs that differ by round-off part after a few outer
    iterations; it is compared on the first outer iteration, Upwind (continuous in the face
    flux) after three."""
    print(f"case pressure system 10^3 {scheme} after {outer} outer iterations", flush=True)
    out = {}
    for name, be in (("hip", hip), ("ref", ref)):
        c, S, pe, pp, refc, rep, _ = captured(be, 10, outer, scheme)
        pe.update_coeffs()
        rp, col, val, n = pe.matrix()
        out[name] = (val, pe.rhs(), refc, c["mdot"].face_fluxes())
    (vh, bh, rh, fh), (vr, br, rr, fr) = out["hip"], out["ref"]
    assert rh == rr == 0
    assert np.abs(vh - vr).max() <= 1e-12 * np.abs(vr).max()
    assert np.abs(bh - br).max() <= 1e-10 * np.abs(br).max()
    assert np.abs(fh - fr).max() <= 1e-10 * np.abs(fr).max()
