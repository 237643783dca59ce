"""Packed-row BiCGSTAB kernels (four FP32 or two FP64 consecutive rows per thread, DIA
gathers of whole aligned packs) against the FP64 oracle and against the row-per-thread
kernels, on box meshes whose offsets qualify (n and nx multiples of 4).  PRISM_KRYLOV_PACK=2
forces the packed form at these small sizes, 0 disables it."""
import os

import numpy as np
import pytest

import prism_amd as P
from prism_amd import cases

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


class packed:
    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        self.old = os.environ.get("PRISM_KRYLOV_PACK")
        os.environ["PRISM_KRYLOV_PACK"] = str(self.mode)

    def __exit__(self, *exc):
        if self.old is None:
            os.environ.pop("PRISM_KRYLOV_PACK", None)
        else:
            os.environ["PRISM_KRYLOV_PACK"] = self.old


@pytest.mark.parametrize("precision,bound", [("fp64", 1e-10), ("fp32", 1e-7)])
@pytest.mark.parametrize("n", [16, 20])
def test_packed_pressure_against_oracle(hip, ref, precision, bound, n):
    """Symmetric DIA (the all-Neumann pressure system, reference cell 0): packed kernels,
    row-per-thread kernels and the oracle's exact solve agree."""
    out = {}
    for name, be, mode in (("packed", hip, 2), ("rows", hip, 0), ("ref", ref, 0)):
        c = cases.pressure_correction(be, n=n)
        slv = P.solver.BiCGSTAB(tol=1e-12 if precision == "fp64" else 1e-9, precision=precision)
        with packed(mode):
            r = slv.solve(c["eqn"], 2, 1e-30, ref_cell=0)
        out[name] = (c["p"].values, r)
    assert rel(out["packed"][0], out["ref"][0]) <= bound
    assert rel(out["rows"][0], out["ref"][0]) <= bound
    assert out["packed"][1].krylov_rel_residual <= (1e-12 if precision == "fp64" else 1e-9)


@pytest.mark.parametrize("precision,bound", [("fp64", 1e-10), ("fp32", 1e-8)])
def test_packed_advection_against_oracle(hip, ref, precision, bound):
    """Nonsymmetric DIA (jittered 16 x 12 x 8 box, Upwind, offsets +-1, +-16, +-192)."""
    out = {}
    for name, be, mode in (("packed", hip, 2), ("ref", ref, 0)):
        c = cases.advection_diffusion(be, P.box(16, 12, 8, jitter=0.2, seed=2), scheme="Upwind")
        slv = P.solver.BiCGSTAB(tol=1e-12 if precision == "fp64" else 1e-10, precision=precision)
        with packed(mode):
            slv.solve(c["eqn"], 2, 1e-20)
        out[name] = c["T"].values
    assert rel(out["packed"], out["ref"]) <= bound


def test_packed_odd_width_uses_rows(hip, ref):
    """nx = 10: the +-nx offsets are not multiples of four, so the SpMV kernels keep the
    row-per-thread form (X may still be packed when n is a multiple of the width)."""
    out = {}
    for name, be in (("hip", hip), ("ref", ref)):
        c = cases.advection_diffusion(be, P.box(10, 8, 6, jitter=0.2, seed=2), scheme="Upwind")
        slv = P.solver.BiCGSTAB(tol=1e-12)
        with packed(2):
            slv.solve(c["eqn"], 2, 1e-20)
        out[name] = c["T"].values
    assert rel(out["hip"], out["ref"]) <= 1e-10
