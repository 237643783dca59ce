"""Sliced-ELL with 16-bit column offsets from a per-slot base, on a tetrahedral box mesh whose
z neighbours lie 6 nx ny = 34,656 rows away (more than 16-bit offsets from the row reach, so
such meshes used 32-bit indices before the bases): the format is chosen, its solves equal the
32-bit-index solves bit for bit (PRISM_KRYLOV_SELL32=1 forces those; both encodings give the
same columns), and the FP64 solution matches the oracle's."""
import os

import numpy as np
import pytest

import prism_amd as P
from prism_amd import cases

pytestmark = pytest.mark.gpu

NX, NY, NZ = 76, 76, 2


class env:
    def __init__(self, name, value):
        self.name, self.value = name, value

    def __enter__(self):
        self.old = os.environ.get(self.name)
        os.environ[self.name] = str(self.value)

    def __exit__(self, *exc):
        if self.old is None:
            os.environ.pop(self.name, None)
        else:
            os.environ[self.name] = self.old


@pytest.fixture(scope="module")
def mesh_data():
    return P.box(NX, NY, NZ, jitter=0.1, seed=5, tets=True)


def _solve(be, data, precision, tol, sell32):
    c = cases.advection_diffusion(be, data, scheme="Upwind", kappa=0.05)
    slv = P.solver.BiCGSTAB(tol=tol, precision=precision)
    with env("PRISM_KRYLOV_SELL32", 1 if sell32 else 0):
        r = slv.solve(c["eqn"], 2, 1e-20)
    return c, r


def test_tet_box_takes_the_16_bit_format(hip, mesh_data):
    import bench
    c = cases.advection_diffusion(hip, mesh_data, scheme="Upwind", kappa=0.05)
    c["eqn"].update_coeffs()
    rp, col, val, n = c["eqn"].matrix()
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert np.abs(col.astype(np.int64) - rows).max() > 32767  # out of reach of offsets from the row
    assert bench.matrix_format(rp, col, val)[0] == "sell16"


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-10)])
def test_16_bit_offsets_equal_32_bit_indices(hip, mesh_data, precision, tol):
    print(f"case sell16 vs sell32 {precision}", flush=True)
    c16, r16 = _solve(hip, mesh_data, precision, tol, False)
    c32, r32 = _solve(hip, mesh_data, precision, tol, True)
    assert r16.krylov_iters == r32.krylov_iters
    assert np.array_equal(c16["T"].values, c32["T"].values)


def test_16_bit_offsets_against_oracle(hip, ref, mesh_data):
    print("case sell16 vs oracle", flush=True)
    # both solves to 1e-13 relative residual: the solutions then agree to about 1e-11 (at
    # 1e-12 the conditioning of this system leaves 3e-10)
    ch, _ = _solve(hip, mesh_data, "fp64", 1e-13, False)
    cr = cases.advection_diffusion(ref, mesh_data, scheme="Upwind", kappa=0.05)
    P.solver.BiCGSTAB(tol=1e-13).solve(cr["eqn"], 2, 1e-20)
    d = ch["T"].values - cr["T"].values
    assert np.linalg.norm(d) / np.linalg.norm(cr["T"].values) <= 1e-10
