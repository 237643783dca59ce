"""Malformed input is refused on the host with an error code, before any device work: mesh
descriptors with out-of-range ids or bad offsets (PRISM_ERR_MESH), UNV files whose elements
have too few nodes (PRISM_ERR_MESH), and null arguments (PRISM_ERR_INVALID).  These calls
never reach the GPU (the descriptor is validated before the context is used), so they run
here through libprism_hip.so with a null context."""
import ctypes as C
import gzip
import os

import numpy as np
import pytest

import prism_amd as P
from prism_amd import _capi as A
from prism_amd.core import mesh_descriptor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    return P.hip_lib()


def create(lib, data, ctx=None):
    bcs = {s: {"T": ("fixed", 0.0)} for s in data.patch_names}
    d, keep = mesh_descriptor(data, bcs, [("T", "scalar", None)])
    h = A.vp()
    return lib.raw("mesh_create")(ctx, C.byref(d), C.byref(h))


def broken(mutate):
    data = P.box(3, 2, 2)
    mutate(data)
    return data


CASES = {
    "face vertex id past the end": lambda d: d.face_vtx.__setitem__(5, len(d.vertices)),
    "negative face vertex id": lambda d: d.face_vtx.__setitem__(0, -1),
    "decreasing face offsets": lambda d: d.face_vtx_off.__setitem__(3, d.face_vtx_off[2] - 1),
    "cell face id past the end": lambda d: d.cell_faces.__setitem__(4, len(d.face_owner)),
    "negative cell face id": lambda d: d.cell_faces.__setitem__(0, -3),
    "cell with no faces": lambda d: d.cell_face_off.__setitem__(1, 0),
    "owner out of range": lambda d: d.face_owner.__setitem__(2, d.n_cells),
    "neighbour out of range": lambda d: d.face_neighbor.__setitem__(int(np.argmax(d.face_neighbor >= 0)), d.n_cells),
    "owner is its own neighbour": lambda d: d.face_neighbor.__setitem__(
        int(np.argmax(d.face_neighbor >= 0)), d.face_owner[int(np.argmax(d.face_neighbor >= 0))]),
    "face offsets not from zero": lambda d: d.face_vtx_off.__setitem__(0, 1),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_malformed_descriptor_is_a_mesh_error(lib, case):
    rc = create(lib, broken(CASES[case]))
    assert rc == A.PRISM_ERR_MESH, lib.raw("last_error")().decode()


def test_valid_descriptor_then_needs_a_context(lib):
    rc = create(lib, P.box(3, 2, 2))
    assert rc == A.PRISM_ERR_INVALID
    assert b"context" in lib.raw("last_error")()


def test_null_arguments(lib):
    rc = lib.raw("courant")(None, C.c_double(0.1), None)
    assert rc == A.PRISM_ERR_INVALID
    rc = lib.raw("adjoint_objective")(None, None, None)
    assert rc == A.PRISM_ERR_INVALID
    rc = lib.raw("csr_bicgstab")(None, None, None, None, None, None)
    assert rc == A.PRISM_ERR_INVALID
    rc = lib.raw("mesh_create")(None, None, None)
    assert rc == A.PRISM_ERR_INVALID


def unv_text():
    path = os.path.join(ROOT, "tests", "golden", "cases", "poisson", "mesh_hex.unv.gz")
    with gzip.open(path, "rt") as f:
        return f.read()


def first_hex_record(lines):
    """Index of the first 2412 header line of an 8-node hexahedron (type 115)."""
    in2412 = False
    for i, ln in enumerate(lines):
        t = ln.split()
        if ln.strip() == "2412":
            in2412 = True
            continue
        if in2412 and len(t) == 6 and t[1] == "115":
            return i
    raise AssertionError("no hexahedron in the fixture")


def test_unv_element_with_too_few_nodes(lib, tmp_path):
    lines = unv_text().split("\n")
    i = first_hex_record(lines)
    t = lines[i].split()
    nodes = lines[i + 1].split()
    assert len(nodes) == 8
    # declare 6 nodes and give 6: a hexahedron that would index nodes 6 and 7
    lines[i] = " ".join(t[:5] + ["6"])
    lines[i + 1] = " ".join(nodes[:6])
    p = tmp_path / "short.unv"
    p.write_text("\n".join(lines))
    gen = A.vp()
    rc = lib.raw("meshgen_unv")(str(p).encode(), C.byref(gen))
    assert rc == A.PRISM_ERR_MESH
    assert b"115" in lib.raw("last_error")()


def test_unv_malformed_element_header(lib, tmp_path):
    lines = unv_text().split("\n")
    i = first_hex_record(lines)
    lines[i] = lines[i].split()[0] + " 115 x"
    p = tmp_path / "bad.unv"
    p.write_text("\n".join(lines))
    gen = A.vp()
    assert lib.raw("meshgen_unv")(str(p).encode(), C.byref(gen)) == A.PRISM_ERR_MESH


def test_unv_fixture_still_reads(lib, tmp_path):
    p = tmp_path / "ok.unv"
    p.write_text(unv_text())
    data = P.read_unv(str(p))
    assert data.n_cells > 0
