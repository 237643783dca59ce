"""A small reader of VTK XML unstructured-grid files (ASCII DataArrays, or appended raw data
with or without vtkZLibDataCompressor blocks behind a UInt64 header) and of .pvd
collections, written from the VTK file-format description for the output tests."""
from __future__ import annotations

import html
import re
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np

_DT = {"Float64": np.float64, "Float32": np.float32, "Int64": np.int64, "Int32": np.int32, "Int8": np.int8,
       "UInt8": np.uint8, "UInt64": np.uint64}


def _attrs(tag: str) -> dict:
    return {k: html.unescape(v) for k, v in re.findall(r'(\w+)="([^"]*)"', tag)}


def read_vtu(path: str) -> dict:
    """{'header': attrs of VTKFile, 'n_points', 'n_cells', 'points' (n, 3), 'connectivity',
    'offsets', 'types', 'point_data': {name: array}, 'cell_data': {name: array},
    'order': [(section, name)]} with multi-component arrays shaped (n, k)."""
    raw = open(path, "rb").read()
    cut = raw.find(b"<AppendedData")
    xml_part = raw if cut < 0 else raw[:cut]
    appended = b""
    if cut >= 0:
        start = raw.index(b"_", raw.index(b">", cut)) + 1
        end = raw.rindex(b"</AppendedData>")
        appended = raw[start:end]
    text = xml_part.decode("ascii")
    head = _attrs(re.search(r"<VTKFile[^>]*>", text).group(0))
    compressed = head.get("compressor") == "vtkZLibDataCompressor"
    assert head.get("header_type", "UInt32") == "UInt64"
    piece = _attrs(re.search(r"<Piece[^>]*>", text).group(0))
    out = {"header": head, "n_points": int(piece["NumberOfPoints"]), "n_cells": int(piece["NumberOfCells"]),
           "point_data": {}, "cell_data": {}, "order": []}
    section = None
    for m in re.finditer(r"<(/?)(PointData|CellData|Points|Cells)>|<DataArray([^>]*?)(/?)>", text):
        if m.group(2):
            section = None if m.group(1) else m.group(2)
            continue
        a = _attrs(m.group(3))
        dt = _DT[a["type"]]
        if a["format"] == "ascii":
            body_end = text.index("</DataArray>", m.end())
            vals = np.array(text[m.end():body_end].split(), dtype=np.float64 if dt == np.float64 else np.int64)
            arr = vals.astype(dt)
        else:
            assert a["format"] == "appended"
            off = int(a["offset"])
            if compressed:
                nb, bs, last = struct.unpack_from("<3Q", appended, off)
                sizes = struct.unpack_from(f"<{nb}Q", appended, off + 24)
                at = off + 24 + 8 * nb
                chunks = []
                for b, cs in enumerate(sizes):
                    blk = zlib.decompress(appended[at:at + cs])
                    want = last if (b == nb - 1 and last) else bs
                    assert len(blk) == want
                    chunks.append(blk)
                    at += cs
                data = b"".join(chunks)
            else:
                (nbytes,) = struct.unpack_from("<Q", appended, off)
                data = appended[off + 8:off + 8 + nbytes]
            arr = np.frombuffer(data, dtype=dt).copy()
        k = int(a.get("NumberOfComponents", "1"))
        if k > 1:
            arr = arr.reshape(-1, k)
        name = a.get("Name", "")
        out["order"].append((section, name))
        if section == "PointData":
            out["point_data"][name] = arr
        elif section == "CellData":
            out["cell_data"][name] = arr
        elif section == "Points":
            out["points"] = arr
        elif section == "Cells":
            out[name] = arr
    return out


def read_pvd(path: str):
    """[(timestep string, file)] of a ParaView collection."""
    root = ET.parse(path).getroot()
    assert root.tag == "VTKFile" and root.get("type") == "Collection"
    return [(d.get("timestep"), d.get("file")) for d in root.find("Collection").findall("DataSet")]
