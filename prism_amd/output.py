"""prism::output (export.h, export.cpp) on the C ABI: toVtu, FieldRegistry, TimeWriter.

Point data is the inverse-distance-weighted cell -> vertex interpolation of cellToPoint
(export.cpp:79-112), computed on the device of the fields' backend; the file is written by
the library.  ``to_vtu(field, file_name, mode)`` and ``to_vtu(scalars, vectors, tensors,
file_name, mode)`` mirror the reference's overloads (export.h:14-35).
"""
from __future__ import annotations

import ctypes as C
import enum
from typing import List, Sequence, Tuple, Union

import numpy as np

from . import _capi as A
from .core import Scalar, Tensor, Vector, _dp

Field = Union[Scalar, Vector, Tensor]


class WriteMode(enum.IntEnum):
    """output::WriteMode (export.h:13)."""
    Ascii = A.WRITE_ASCII
    RawBinaryCompressed = A.WRITE_RAW_BINARY_COMPRESSED


def _kind(f: Field) -> int:
    if isinstance(f, Scalar):
        return A.EXPORT_SCALAR
    if isinstance(f, Vector):
        return A.EXPORT_VECTOR
    if isinstance(f, Tensor):
        return A.EXPORT_TENSOR
    raise TypeError(f"cannot export {type(f).__name__}")


def n_vertices(mesh) -> int:
    nv = C.c_int64()
    mesh.backend("mesh_n_vertices", mesh.h, C.byref(nv))
    return nv.value


def cell_to_point(f: Field) -> np.ndarray:
    """The point values export writes for `f`: shape (n_vertices,), (n_vertices, 3) or
    (n_vertices, 3, 3) (row-major tensors)."""
    k = _kind(f)
    nv = n_vertices(f.mesh)
    out = np.zeros(nv * k)
    fn = {A.EXPORT_SCALAR: "field_to_points", A.EXPORT_VECTOR: "vfield_to_points",
          A.EXPORT_TENSOR: "tfield_to_points"}[k]
    f.mesh.backend(fn, f.h, _dp(out))
    return out.reshape((nv,) if k == 1 else (nv, 3) if k == 3 else (nv, 3, 3))


def _export(fields: Sequence[Field], file_name: str, mode: WriteMode):
    if not fields:
        return  # export.cpp:273-275: nothing to write
    items = (A.ExportItem * len(fields))(*[A.ExportItem(_kind(f), f.h) for f in fields])
    fields[0].mesh.backend("export_vtu", items, len(fields), str(file_name).encode(), int(WriteMode(mode)))


def to_vtu(*args, mode: WriteMode = WriteMode.RawBinaryCompressed):
    """toVtu(field, file_name[, mode]) or toVtu(scalars, vectors, tensors, file_name[, mode])
    (export.cpp:248-300): each field's CellData and PointData, scalars first, then vectors,
    then tensors; all fields must share one mesh (InvalidMeshError otherwise)."""
    if len(args) in (2, 3) and not isinstance(args[0], (list, tuple)):
        if len(args) == 3:
            mode = args[2]
        return _export([args[0]], args[1], mode)
    if len(args) in (4, 5):
        scalars, vectors, tensors, file_name = args[:4]
        if len(args) == 5:
            mode = args[4]
        return _export(list(scalars) + list(vectors) + list(tensors), file_name, mode)
    raise TypeError("to_vtu(field, file_name[, mode]) or to_vtu(scalars, vectors, tensors, file_name[, mode])")


class FieldRegistry:
    """output::FieldRegistry (export.cpp:302-336)."""

    def __init__(self):
        self._scalars: List[Scalar] = []
        self._vectors: List[Vector] = []
        self._tensors: List[Tensor] = []

    def add(self, f: Field):
        {A.EXPORT_SCALAR: self._scalars, A.EXPORT_VECTOR: self._vectors,
         A.EXPORT_TENSOR: self._tensors}[_kind(f)].append(f)

    def export_all(self, file_name: str, mode: WriteMode = WriteMode.RawBinaryCompressed):
        to_vtu(self._scalars, self._vectors, self._tensors, file_name, mode)

    @property
    def scalars(self) -> List[Scalar]:
        return self._scalars

    @property
    def vectors(self) -> List[Vector]:
        return self._vectors

    @property
    def tensors(self) -> List[Tensor]:
        return self._tensors

    def _any(self) -> Field:
        return (self._scalars + self._vectors + self._tensors)[0]


def write_pvd(backend, file_name: str, entries: Sequence[Tuple[float, str]]):
    """A ParaView collection of (time, file) entries (TimeWriter::flushPvd, export.cpp:393-409)."""
    times = np.ascontiguousarray([t for t, _ in entries], dtype=np.float64)
    names = (C.c_char_p * max(1, len(entries)))(*[str(n).encode() for _, n in entries])
    backend("write_pvd", str(file_name).encode(), len(entries), _dp(times), names)


class TimeWriter:
    """output::TimeWriter (export.cpp:338-391): writes `{prefix}_{step:04d}.vtu` every
    `write_interval` steps and, with `write_pvd`, rewrites `{prefix}.pvd` after each file."""

    def __init__(self, prefix: str = "result", write_interval: int = 1,
                 mode: WriteMode = WriteMode.RawBinaryCompressed, write_pvd: bool = True):
        if write_interval == 0:
            raise ValueError("prism::output::TimeWriter(prefix, write_interval, ...): write_interval must be "
                             f"positive, got {write_interval}")
        self._prefix = str(prefix)
        self._interval = int(write_interval)
        self._mode = WriteMode(mode)
        self._write_pvd = bool(write_pvd)
        self._registry = FieldRegistry()
        self._step = 0
        self._time = 0.0
        self._pvd_entries: List[Tuple[float, str]] = []

    def add(self, f: Field):
        self._registry.add(f)

    def advance(self, dt: float = 0.0):
        self._step += 1
        self._time += dt

    def write(self):
        if self._step % self._interval != 0:
            return
        r = self._registry
        if not (r.scalars or r.vectors or r.tensors):
            return
        file_name = f"{self._prefix}_{self._step:04d}.vtu"
        r.export_all(file_name, self._mode)
        if self._write_pvd:
            self._pvd_entries.append((self._time, file_name))
            write_pvd(r._any().mesh.backend, f"{self._prefix}.pvd", self._pvd_entries)

    def write_and_advance(self, dt: float = 0.0):
        self.advance(dt)
        self.write()

    @property
    def current_time(self) -> float:
        return self._time

    @property
    def current_step(self) -> int:
        return self._step
