"""Output export (export.cpp) on the CPU: the oracle's cellToPoint against an independent
numpy evaluation of the inverse-distance weighting and against what the mathematics fixes
(constants reproduced, point values inside the range of their cells, equal weights on a
uniform grid), the VTU files it writes read back with tests/vtu_io.py, the UNV cells'
vertex lists against the file's own element records, the ParaView collection text and
TimeWriter's naming, interval and time stamps.  The reference ships no export test and its
VTU writer (the vtu11 submodule) is not vendored, so the file layout is checked against
the VTK XML format, not against vtu11's bytes."""
import gzip
import os

import numpy as np
import pytest

import oracle
import prism_amd as P
from prism_amd import output as O

import vtu_io
from refcases import GOLDEN, load_mesh

EPS = np.finfo(np.float64).eps


def _mesh(be, data, fields=(("T", "scalar"), ("U", "vector"))):
    bcs = {p: {"T": "zero-gradient", "U": ("fixed", (0.0, 0.0, 0.0))} for p in data.patch_names}
    return P.Mesh(be, data, bcs, list(fields))


def _adjacency(data):
    nv = len(data.vertices)
    adj = [[] for _ in range(nv)]
    off, vtx = data.cell_vtx_off, data.cell_vtx
    for c in range(len(off) - 1):
        for v in vtx[off[c]:off[c + 1]]:
            adj[v].append(c)
    return adj


def idw_numpy(data, centres, values):
    """sum_c f_c / d_c / sum_c 1 / d_c with d_c = max(|x_v - x_c|, eps), vectorised per vertex."""
    values = values.reshape(len(centres), -1)
    out = np.zeros((len(data.vertices), values.shape[1]))
    for v, cells in enumerate(_adjacency(data)):
        d = np.maximum(np.linalg.norm(data.vertices[v] - centres[cells], axis=1), EPS)
        w = 1.0 / d
        out[v] = (w[:, None] * values[cells]).sum(0) / w.sum()
    return out


@pytest.fixture(scope="module")
def be():
    return oracle.backend()


@pytest.mark.parametrize("tets", [False, True])
def test_cell_to_point_matches_numpy_idw(be, tets):
    data = P.box(5, 4, 3, jitter=0.25, seed=11, tets=tets)
    m = _mesh(be, data)
    rng = np.random.default_rng(5)
    T = m.scalar("T", values=rng.standard_normal(m.n_cells))
    U = m.vector("U", values=rng.standard_normal((m.n_cells, 3)))
    K = m.tensor("K", values=rng.standard_normal((m.n_cells, 3, 3)))
    centres = m.geometry()["cell_center"]
    for f, vals in ((T, T.values), (U, U.values), (K, K.values)):
        got = O.cell_to_point(f).reshape(len(data.vertices), -1)
        want = idw_numpy(data, centres, vals)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


def test_cell_to_point_against_what_the_weighting_fixes(be):
    # constants are reproduced and every point value lies within its cells' range
    data = P.box(6, 5, 4, jitter=0.3, seed=2)
    m = _mesh(be, data)
    c = m.scalar("T", value=3.75)
    np.testing.assert_allclose(O.cell_to_point(c), 3.75, rtol=4 * EPS)
    rng = np.random.default_rng(9)
    vals = rng.uniform(-2, 5, m.n_cells)
    T = m.scalar("T", values=vals)
    pts = O.cell_to_point(T)
    for v, cells in enumerate(_adjacency(data)):
        assert vals[cells].min() - 1e-12 <= pts[v] <= vals[cells].max() + 1e-12
    # on a uniform grid an interior vertex is equidistant from its 8 cells: their mean
    data = P.box(4, 4, 4)
    m = _mesh(be, data)
    T = m.scalar("T", values=rng.standard_normal(m.n_cells))
    pts = O.cell_to_point(T)
    adj = _adjacency(data)
    interior = [v for v in range(len(adj)) if len(adj[v]) == 8]
    assert len(interior) == 27
    for v in interior:
        np.testing.assert_allclose(pts[v], T.values[adj[v]].mean(), rtol=1e-14, atol=1e-15)
    # a corner vertex belongs to one cell and takes its value exactly
    corners = [v for v in range(len(adj)) if len(adj[v]) == 1]
    assert len(corners) == 8
    for v in corners:
        assert pts[v] == T.values[adj[v][0]]


def test_cell_to_point_without_cell_vertex_lists(be):
    # a descriptor without cell vertex lists: each cell names the distinct vertices of its faces
    data = P.box(4, 3, 3, jitter=0.2, seed=4)
    bare = P.MeshData(**{k: getattr(data, k) for k in ("vertices", "face_vtx_off", "face_vtx", "face_owner",
                                                       "face_neighbor", "face_patch", "cell_face_off",
                                                       "cell_faces", "patch_names")})
    a, b = _mesh(be, data), _mesh(be, bare)
    vals = np.random.default_rng(1).standard_normal(a.n_cells)
    np.testing.assert_array_equal(O.cell_to_point(a.scalar("T", values=vals)),
                                  O.cell_to_point(b.scalar("T", values=vals)))
    with pytest.raises(P.InvalidMeshError):
        O.to_vtu(b.scalar("T", values=vals), "unused.vtu", O.WriteMode.Ascii)


def test_vertex_without_cells_is_invalid_mesh(be):
    data = P.box(3, 3, 3)
    data.vertices = np.vstack([data.vertices, [[5.0, 5.0, 5.0]]])
    m = _mesh(be, data)
    with pytest.raises(P.InvalidMeshError, match="Vertex 64 has no adjacent cells"):
        O.cell_to_point(m.scalar("T", value=1.0))


def _unv_elements(path):
    """Volume elements of an Ideas-UNV file (dataset 2412, FE ids 111/112/115) as 0-based
    vertex index lists, vertices numbered in dataset-2411 order."""
    lines = gzip.open(path, "rt").read().split("\n")
    label_of, cells, i, ds = {}, [], 0, None
    while i < len(lines):
        s = lines[i].strip()
        if s == "-1":
            if ds is None and i + 1 < len(lines) and lines[i + 1].strip() in ("2411", "2412"):
                ds = lines[i + 1].strip()
                i += 2
                continue
            ds = None
            i += 1
            continue
        if ds == "2411":
            label_of[int(s.split()[0])] = len(label_of)
            i += 2
        elif ds == "2412":
            rec = s.split()
            fe, nn = int(rec[1]), int(rec[5])
            if fe in (11, 21, 22, 23, 24):
                i += 3
                continue
            nodes = []
            i += 1
            while len(nodes) < nn:
                nodes += [int(x) for x in lines[i].split()]
                i += 1
            if fe in (111, 112, 115):
                cells.append([label_of[x] for x in nodes])
        else:
            i += 1
    return cells


@pytest.mark.parametrize("case", ["ductSIMPLE", "poisson", "coarsePipeHex"])
def test_vtu_of_reference_case_meshes(be, case, tmp_path):
    m = load_mesh(be, case)
    cells = _unv_elements(os.path.join(GOLDEN, case, "mesh.unv.gz"))
    data = m.data
    assert len(cells) == m.n_cells
    np.testing.assert_array_equal(data.cell_vtx, np.concatenate(cells))
    rng = np.random.default_rng(3)
    T = m.scalar("T", values=rng.standard_normal(m.n_cells))
    U = m.vector("U", values=rng.standard_normal((m.n_cells, 3)))
    K = m.tensor("K", values=rng.standard_normal((m.n_cells, 3, 3)))
    path = str(tmp_path / "out.vtu")
    O.to_vtu([T], [U], [K], path, O.WriteMode.Ascii)
    r = vtu_io.read_vtu(path)
    assert (r["n_points"], r["n_cells"]) == (len(data.vertices), m.n_cells)
    np.testing.assert_array_equal(r["points"], data.vertices)
    np.testing.assert_array_equal(r["connectivity"], np.concatenate(cells))
    np.testing.assert_array_equal(r["offsets"], np.cumsum([len(c) for c in cells]))
    np.testing.assert_array_equal(r["types"], [{8: 12, 4: 10, 6: 13}[len(c)] for c in cells])
    np.testing.assert_array_equal(r["cell_data"]["T"], T.values)
    np.testing.assert_array_equal(r["cell_data"]["U"], U.values)
    np.testing.assert_array_equal(r["cell_data"]["K"], K.values.reshape(-1, 9))
    np.testing.assert_array_equal(r["point_data"]["T"], O.cell_to_point(T))
    np.testing.assert_array_equal(r["point_data"]["U"], O.cell_to_point(U))
    np.testing.assert_array_equal(r["point_data"]["K"], O.cell_to_point(K).reshape(-1, 9))
    assert r["order"][:6] == [("PointData", "T"), ("PointData", "U"), ("PointData", "K"),
                              ("CellData", "T"), ("CellData", "U"), ("CellData", "K")]
    np.testing.assert_allclose(r["point_data"]["U"], idw_numpy(data, m.geometry()["cell_center"], U.values),
                               rtol=1e-12, atol=1e-13)


def test_export_errors(be, tmp_path):
    data = P.box(3, 2, 2)
    m1, m2 = _mesh(be, data), _mesh(be, data)
    T, U = m1.scalar("T", value=1.0), m2.vector("U", value=(1.0, 0.0, 0.0))
    with pytest.raises(P.InvalidMeshError, match="Field 'U' does not share the mesh"):
        O.to_vtu([T], [U], [], str(tmp_path / "a.vtu"), O.WriteMode.Ascii)
    O.to_vtu([], [], [], str(tmp_path / "none.vtu"))
    assert not os.path.exists(tmp_path / "none.vtu")
    bad = P.box(3, 2, 2)
    bad.cell_vtx_off = bad.cell_vtx_off.copy()
    bad.cell_vtx = np.delete(bad.cell_vtx, 1)  # vertex 1 stays adjacent to cell 1
    bad.cell_vtx_off[1:] -= 1
    with pytest.raises(P.PrismError, match="Unsupported cell type"):
        O.to_vtu(_mesh(be, bad).scalar("T", value=1.0), str(tmp_path / "b.vtu"), O.WriteMode.Ascii)


def test_names_are_escaped(be, tmp_path):
    name = 'a<b&"c">'
    data = P.box(2, 2, 2)
    m = P.Mesh(be, data, {p: {name: "zero-gradient"} for p in data.patch_names}, [(name, "scalar")])
    O.to_vtu(m.scalar(name, value=1.0), str(tmp_path / "e.vtu"), O.WriteMode.Ascii)
    assert list(vtu_io.read_vtu(str(tmp_path / "e.vtu"))["point_data"]) == [name]
    O.write_pvd(be, str(tmp_path / "e.pvd"), [(1.0, 'x&"y".vtu')])
    assert vtu_io.read_pvd(str(tmp_path / "e.pvd")) == [("1", 'x&"y".vtu')]


def test_pvd_text(be, tmp_path):
    path = str(tmp_path / "c.pvd")
    O.write_pvd(be, path, [(0.1, "r_0001.vtu"), (1e-05, "r_0002.vtu"), (12345678.0, "r_0003.vtu"), (2.0, "x.vtu")])
    assert open(path).read() == (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="Collection" version="0.1">\n'
        '  <Collection>\n'
        '    <DataSet timestep="0.1" file="r_0001.vtu"/>\n'
        '    <DataSet timestep="1e-05" file="r_0002.vtu"/>\n'
        '    <DataSet timestep="1.23457e+07" file="r_0003.vtu"/>\n'
        '    <DataSet timestep="2" file="x.vtu"/>\n'
        '  </Collection>\n'
        '</VTKFile>\n')


def test_time_writer(be, tmp_path):
    with pytest.raises(ValueError, match="write_interval must be positive, got 0"):
        O.TimeWriter("r", 0)
    m = _mesh(be, P.box(3, 2, 2))
    T = m.scalar("T", value=0.0)
    prefix = str(tmp_path / "result")
    w = O.TimeWriter(prefix, 2, O.WriteMode.Ascii)
    w.write()                       # nothing registered yet
    w.add(T)
    w.write()                       # step 0 is a multiple of the interval
    for k in range(5):
        T.fill(float(k + 1))
        w.write_and_advance(0.1)
    assert w.current_step == 5 and abs(w.current_time - 0.5) < 1e-15
    written = sorted(f for f in os.listdir(tmp_path) if f.endswith(".vtu"))
    assert written == ["result_0000.vtu", "result_0002.vtu", "result_0004.vtu"]
    assert vtu_io.read_pvd(prefix + ".pvd") == [("0", prefix + "_0000.vtu"), ("0.2", prefix + "_0002.vtu"),
                                                ("0.4", prefix + "_0004.vtu")]
    r = vtu_io.read_vtu(prefix + "_0004.vtu")
    np.testing.assert_array_equal(r["cell_data"]["T"], 4.0)
    q = O.TimeWriter(str(tmp_path / "nopvd"), 1, O.WriteMode.Ascii, write_pvd=False)
    q.add(T)
    q.write_and_advance(1.0)
    assert os.path.exists(tmp_path / "nopvd_0001.vtu") and not os.path.exists(tmp_path / "nopvd.pvd")
