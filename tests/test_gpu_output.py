"""Output export on the GPU: the device cell -> point interpolation (k_cell_to_point) bit-exact
against the oracle's cellToPoint (export.cpp:79-112) on hexahedra, tetrahedra and the
reference's wedge / mixed case meshes; the VTU files of both libraries read back to the same
arrays; the compressed appended mode decoding to the ASCII mode's values; TimeWriter on the
device; and at 128^3 cells the properties the weighting fixes (constants, cell range)."""
import os

import numpy as np
import pytest

import prism_amd as P
from prism_amd import output as O

import vtu_io
from refcases import load_mesh

pytestmark = pytest.mark.gpu


def _mesh(be, data):
    bcs = {p: {"T": "zero-gradient", "U": ("fixed", (0.0, 0.0, 0.0))} for p in data.patch_names}
    return P.Mesh(be, data, bcs, [("T", "scalar"), ("U", "vector")])


def _fields(m, seed):
    rng = np.random.default_rng(seed)
    return (m.scalar("T", values=rng.standard_normal(m.n_cells)),
            m.vector("U", values=rng.standard_normal((m.n_cells, 3))),
            m.tensor("K", values=rng.standard_normal((m.n_cells, 3, 3))))


MESHES = [("box_hex", lambda: P.box(12, 10, 8, jitter=0.25, seed=3)),
          ("box_tet", lambda: P.box(6, 5, 4, jitter=0.2, seed=4, tets=True)),
          ("ductSIMPLE", None), ("poisson", None)]


@pytest.mark.parametrize("name,make", MESHES, ids=[n for n, _ in MESHES])
def test_points_and_vtu_match_oracle(hip, ref, name, make, tmp_path):
    print(f"case {name}", flush=True)
    if make is None:
        mh, mr = load_mesh(hip, name), load_mesh(ref, name)
    else:
        data = make()
        mh, mr = _mesh(hip, data), _mesh(ref, data)
    fh, fr = _fields(mh, 7), _fields(mr, 7)
    for a, b in zip(fh, fr):
        np.testing.assert_array_equal(O.cell_to_point(a), O.cell_to_point(b))
    O.to_vtu([fh[0]], [fh[1]], [fh[2]], str(tmp_path / "hip.vtu"), O.WriteMode.Ascii)
    O.to_vtu([fr[0]], [fr[1]], [fr[2]], str(tmp_path / "ref.vtu"), O.WriteMode.Ascii)
    O.to_vtu([fh[0]], [fh[1]], [fh[2]], str(tmp_path / "hipz.vtu"), O.WriteMode.RawBinaryCompressed)
    a, b, z = (vtu_io.read_vtu(str(tmp_path / f)) for f in ("hip.vtu", "ref.vtu", "hipz.vtu"))
    assert a["order"] == b["order"] == z["order"]
    assert z["header"]["compressor"] == "vtkZLibDataCompressor"
    for key in ("points", "connectivity", "offsets", "types"):
        np.testing.assert_array_equal(a[key], b[key])
        np.testing.assert_array_equal(z[key], a[key])
        assert z[key].dtype == a[key].dtype
    for sec in ("point_data", "cell_data"):
        for nm in ("T", "U", "K"):
            np.testing.assert_array_equal(a[sec][nm], b[sec][nm])
            np.testing.assert_array_equal(z[sec][nm], a[sec][nm])


def test_time_writer_on_device(hip, ref, tmp_path):
    print("case time_writer", flush=True)
    data = P.box(8, 6, 5, jitter=0.2, seed=9)
    m = _mesh(hip, data)
    T, U, _ = _fields(m, 1)
    prefix = str(tmp_path / "run")
    w = O.TimeWriter(prefix, 1)
    w.add(T)
    w.add(U)
    for k in range(3):
        T.fill(float(k))
        w.write_and_advance(0.25)
    assert vtu_io.read_pvd(prefix + ".pvd") == [("0.25", prefix + "_0001.vtu"), ("0.5", prefix + "_0002.vtu"),
                                                ("0.75", prefix + "_0003.vtu")]
    r = vtu_io.read_vtu(prefix + "_0003.vtu")
    np.testing.assert_array_equal(r["cell_data"]["T"], 2.0)
    np.testing.assert_array_equal(r["cell_data"]["U"], U.values)
    mr = _mesh(ref, data)
    Ur = mr.vector("U", values=U.values)
    np.testing.assert_array_equal(r["point_data"]["U"], O.cell_to_point(Ur))


def test_points_at_128_cubed(hip):
    print("case points_128", flush=True)
    data = P.box(128, 128, 128, jitter=0.2, seed=5)
    m = _mesh(hip, data)
    c = m.scalar("T", value=-1.625)
    pts = O.cell_to_point(c)
    assert pts.shape == (129 ** 3,)
    assert np.max(np.abs(pts + 1.625)) <= 4 * np.finfo(np.float64).eps * 1.625
    vals = np.random.default_rng(2).uniform(-3, 7, m.n_cells)
    T = m.scalar("T", values=vals)
    pts = O.cell_to_point(T)
    nv = len(data.vertices)
    lo, hi = np.full(nv, np.inf), np.full(nv, -np.inf)
    cell_of = np.repeat(np.arange(m.n_cells), np.diff(data.cell_vtx_off))
    np.minimum.at(lo, data.cell_vtx, vals[cell_of])
    np.maximum.at(hi, data.cell_vtx, vals[cell_of])
    assert np.all(pts >= lo - 1e-12) and np.all(pts <= hi + 1e-12)
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"points128_{os.getpid()}.vtu")
    try:
        O.to_vtu(T, path)
        r = vtu_io.read_vtu(path)
        np.testing.assert_array_equal(r["point_data"]["T"], pts)
        np.testing.assert_array_equal(r["cell_data"]["T"], vals)
    finally:
        if os.path.exists(path):
            os.remove(path)
