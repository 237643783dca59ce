"""The allocation-free BiCGSTAB step (prism_eqn_step_noalloc) against prism_eqn_step on the
same systems: two calls in a row, the second on a smaller system (the static scratch array
keeps the first call's contents), FP64 bitwise and FP32 + refinement to its tolerance, plus
the size limit.  Runs through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import prism_amd as P
from prism_amd import _capi as A
from prism_amd import cases

pytestmark = pytest.mark.gpu


def step_both(be, eqn, T, slv):
    """(x from prism_eqn_step, x from prism_eqn_step_noalloc, both results) from T = 0."""
    out = []
    for fn in ("eqn_step", "eqn_step_noalloc"):
        T.fill(0.0)
        p = slv.params()
        r = A.SolveResult()
        be(fn, eqn.h, C.byref(p), C.byref(r))
        out.append((T.values, r))
    return out


def system(be, n, precision_scheme):
    c = cases.advection_diffusion(be, P.box(n, n - 2, n - 4, jitter=0.2, seed=3), scheme=precision_scheme)
    c["eqn"].update_coeffs()
    return c


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_matches_step_twice(hip, precision):
    slv = P.solver.BiCGSTAB(tol=1e-12 if precision == "fp64" else 1e-10, precision=precision)
    for n in (24, 12):  # second call smaller: stale scratch contents must not leak into it
        print(f"case noalloc {precision} n={n}", flush=True)
        c = system(hip, n, "Upwind")
        (x0, r0), (x1, r1) = step_both(hip, c["eqn"], c["T"], slv)
        assert r1.converged and r0.converged
        assert r1.krylov_iters == r0.krylov_iters
        if precision == "fp64":
            assert np.array_equal(x0, x1)
        else:
            assert np.linalg.norm(x1 - x0) / np.linalg.norm(x0) <= 1e-9


def test_noalloc_pressure_dia_symmetric(hip):
    """Symmetric DIA storage (the pressure system without reference cell: a plain step)."""
    print("case noalloc pressure", flush=True)
    c = cases.pressure_correction(hip, n=20)
    c["eqn"].update_coeffs()
    slv = P.solver.BiCGSTAB(tol=1e-10, max_iters=3000)
    (x0, r0), (x1, r1) = step_both(hip, c["eqn"], c["p"], slv)
    assert r1.krylov_iters == r0.krylov_iters
    assert np.array_equal(x0, x1)


def test_noalloc_rejects_large_systems(hip):
    """A 64^3 system does not fit the 24 MiB array: refused with PRISM_ERR_INVALID."""
    print("case noalloc too large", flush=True)
    c = cases.advection_diffusion(hip, P.box(64, 64, 64), scheme="Upwind")
    c["eqn"].update_coeffs()
    p = P.solver.BiCGSTAB(tol=1e-8).params()
    r = A.SolveResult()
    with pytest.raises(ValueError, match="scratch"):  # PRISM_ERR_INVALID
        hip("eqn_step_noalloc", c["eqn"].h, C.byref(p), C.byref(r))
