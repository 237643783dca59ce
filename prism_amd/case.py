"""Case directories in the reference's format: ``mesh.unv`` (or ``mesh.unv.gz``) beside a
``fields.json`` that lists the fields and one ``<field>.json`` per field with its patches.

Parsing follows mesh/boundary.cpp:94-319: a field entry needs ``name`` and ``type``
(``gradScheme`` optional); a field file needs a ``patches`` array whose entries carry
``name`` and ``type``; ``value`` is a number (scalar BC), an array of exactly three numbers
(vector BC) or absent (no value).  Every field file must exist beside ``fields.json``.
"""
from __future__ import annotations

import gzip
import json
import os
import shutil
import tempfile
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

from . import core


class CaseError(RuntimeError):
    """A malformed case directory (the reference raises std::runtime_error)."""


@dataclass
class Case:
    data: core.MeshData
    bcs: Dict[str, Dict[str, object]]
    fields: List[Tuple[str, str, Optional[str], Optional[List[float]]]]

    def mesh(self, backend: core.Backend) -> core.Mesh:
        return core.Mesh(backend, self.data, self.bcs, self.fields)


def parse_fields(path: str) -> List[Tuple[str, str, Optional[str], Optional[List[float]]]]:
    """fields.json -> [(name, type, gradScheme or None, units or None)] (boundary.cpp:94-121)."""
    with open(path) as f:
        try:
            doc = json.load(f)
        except json.JSONDecodeError as e:
            raise CaseError(f"prism::mesh::readFieldsFile(): Error parsing json file: {e}") from e
    out = []
    for fld in doc.get("fields", []):
        if "name" not in fld or "type" not in fld:
            raise CaseError("prism::mesh::parseField(): Field must contain name and type")
        units = fld.get("units")
        out.append((fld["name"], fld["type"], fld.get("gradScheme"),
                    [float(u) for u in units] if units is not None else None))
    return out


def parse_field_file(field_name: str, path: str) -> Dict[str, object]:
    """<field>.json -> {patch: kind or (kind, value)} (boundary.cpp:152-216)."""
    with open(path) as f:
        try:
            doc = json.load(f)
        except json.JSONDecodeError as e:
            raise CaseError(f"prism::mesh::readBoundaryFile(): Error parsing json file: {e}") from e
    patches = doc.get("patches") if isinstance(doc, dict) else None
    if not isinstance(patches, list):
        raise CaseError(f"prism::mesh::readBoundaryFile(): file `{field_name}` either doesn't contain "
                        "`patches` field of `patches` entry is not an array")
    out: Dict[str, object] = {}
    for p in patches:
        if "name" not in p or "type" not in p:
            raise CaseError(f"prism::mesh::readBoundaryFile(): file `{field_name}` either doesn't contain "
                            "`name` or `type` field")
        name, kind = p["name"], p["type"]
        if name in out:  # std::map::insert keeps the first entry
            continue
        if "value" in p:
            v = p["value"]
            if isinstance(v, bool) or not isinstance(v, (int, float, list)):
                continue  # neither number nor array: no entry (boundary.cpp:187-205)
            if isinstance(v, list):
                if len(v) != 3:
                    raise CaseError(f"prism::mesh::readBoundaryFile(): expected array of size 3 for field "
                                    f"`{field_name}` in patch `{name}`")
                out[name] = (kind, tuple(float(x) for x in v))
            else:
                out[name] = (kind, float(v))
        else:
            out[name] = kind
    return out


def read_boundary(fields_json: str):
    """fields.json and its field files -> (fields, bcs as patch -> field -> spec).  The patch
    set is the first field file's (boundary.cpp:295-319); a patch missing from a later field
    file is an error, as std::map::at raises there."""
    fields = parse_fields(fields_json)
    base = os.path.dirname(os.path.abspath(fields_json))
    files = []
    for name, *_ in fields:
        p = os.path.join(base, f"{name}.json")
        if not os.path.exists(p):
            raise CaseError(f"prism::mesh::readFieldsBoundaryFiles(): boundary file for field `{name}` "
                            f"does not exist in `{base}`")
        files.append((name, parse_field_file(name, p)))
    bcs: Dict[str, Dict[str, object]] = {}
    if files:
        for patch in sorted(files[0][1]):
            entry = {}
            for name, m in files:
                if patch not in m:
                    raise CaseError(f"prism::mesh::extractBoundaryPatches(): field `{name}` has no entry "
                                    f"for patch `{patch}`")
                entry[name] = m[patch]
            bcs[patch] = entry
    return fields, bcs


def read_mesh(path: str) -> core.MeshData:
    """UNV mesh; a ``.gz`` file is inflated to a temporary file first."""
    if not path.endswith(".gz"):
        return core.read_unv(path)
    with tempfile.TemporaryDirectory() as d:
        tmp = os.path.join(d, "mesh.unv")
        with gzip.open(path, "rb") as src, open(tmp, "wb") as dst:
            shutil.copyfileobj(src, dst)
        return core.read_unv(tmp)


def load(case_dir: str, mesh_file: str = "mesh.unv") -> Case:
    """Mesh and boundary conditions of a case directory (mesh.unv(.gz) + fields.json)."""
    mp = os.path.join(case_dir, mesh_file)
    if not os.path.exists(mp) and os.path.exists(mp + ".gz"):
        mp += ".gz"
    fields, bcs = read_boundary(os.path.join(case_dir, "fields.json"))
    data = read_mesh(mp)
    known = set(data.patch_names)
    bcs = {p: v for p, v in bcs.items() if p in known}
    return Case(data=data, bcs=bcs, fields=fields)
