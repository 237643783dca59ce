"""Python mirror of the reference's public API on top of the C ABI (include/prism_hip.h).

Names follow the reference (eigemx/prism): ``eqn.Transport(convection.MUSCL(U, T),
diffusion.Corrected(kappa, T))``, ``solver.BiCGSTAB().solve(eqn, n_iter, eps, ...)``,
``algo.IncompressibleSIMPLE(controls).step(momentum, U, mdot, p)``.  Every object lives on
the device of its backend; host arrays are copies.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass, field as dfield
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _capi as A
from ._capi import Lib, PrismError, vp

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB_PATH = os.path.join(_HERE, "libprism_hip.so")

_hip_lib: Optional[Lib] = None


def hip_lib() -> Lib:
    """The in-tree HIP library; raises if it has not been built (no fallback exists)."""
    global _hip_lib
    if _hip_lib is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise RuntimeError(f"libprism_hip.so is missing at {HIP_LIB_PATH}: run __graft_entry__.build()")
        _hip_lib = Lib(HIP_LIB_PATH, "prism_")
    return _hip_lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(A.dp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------- backend ----
class Backend:
    """A library (HIP, or anything exposing the same ABI) plus one device context."""

    def __init__(self, lib: Lib, device: int = 0):
        self.lib = lib
        ctx = vp()
        lib("ctx_create", device, C.byref(ctx))
        self.ctx = ctx
        self._alive = True

    @classmethod
    def hip(cls, device: int = 0) -> "Backend":
        return cls(hip_lib(), device)

    def __call__(self, name, *args):
        return self.lib(name, *args)

    def close(self):
        if self._alive:
            self.lib("ctx_destroy", self.ctx)
            self._alive = False

    # kernel timing (HIP library only)
    def set_timing(self, on: bool):
        self.lib("ctx_set_timing", self.ctx, 1 if on else 0)

    def reset_timing(self):
        self.lib("ctx_reset_timing", self.ctx)

    def kernel_stats(self) -> Dict[str, Tuple[int, float]]:
        out = {}
        i = 0
        while True:
            name = C.create_string_buffer(128)
            n = C.c_int64()
            ms = C.c_double()
            rc = self.lib.raw("ctx_kernel_stats")(self.ctx, i, name, 128, C.byref(n), C.byref(ms))
            if rc != 0:
                break
            out[name.value.decode()] = (n.value, ms.value)
            i += 1
        return out

    def synchronize(self):
        self.lib("ctx_synchronize", self.ctx)

    def device_info(self):
        name = C.create_string_buffer(64)
        ncu = C.c_int32()
        self.lib("ctx_info", self.ctx, name, 64, C.byref(ncu))
        return name.value.decode(), ncu.value


# ------------------------------------------------------------------------ mesh data ----
@dataclass
class MeshData:
    """Topology in descriptor form (vertex coordinates, faces with owner-outward vertex
    order, cell face lists, patches)."""
    vertices: np.ndarray
    face_vtx_off: np.ndarray
    face_vtx: np.ndarray
    face_owner: np.ndarray
    face_neighbor: np.ndarray
    face_patch: np.ndarray
    cell_face_off: np.ndarray
    cell_faces: np.ndarray
    patch_names: List[str]
    # each cell's own vertex list (Cell::verticesIds, mesh/cell.h:25), used by output export;
    # None: a cell's vertices are the distinct vertices of its faces
    cell_vtx_off: Optional[np.ndarray] = None
    cell_vtx: Optional[np.ndarray] = None

    @property
    def n_cells(self):
        return len(self.cell_face_off) - 1


def _from_gen(lib: Lib, gen) -> MeshData:
    d = A.MeshDesc()
    lib("meshgen_desc", gen, C.byref(d))
    nv, nf, nc = d.n_vertices, d.n_faces, d.n_cells
    arr = lambda p, n, dt: np.ctypeslib.as_array(p, shape=(n,)).astype(dt, copy=True)
    nfv = d.face_vtx_off[nf]
    ncf = d.cell_face_off[nc]
    names = [d.patch_names[i].decode() for i in range(d.n_patches)]
    md = MeshData(
        vertices=arr(d.vertices, 3 * nv, np.float64).reshape(nv, 3),
        face_vtx_off=arr(d.face_vtx_off, nf + 1, np.int64),
        face_vtx=arr(d.face_vtx, nfv, np.int64),
        face_owner=arr(d.face_owner, nf, np.int64),
        face_neighbor=arr(d.face_neighbor, nf, np.int64),
        face_patch=arr(d.face_patch, nf, np.int32),
        cell_face_off=arr(d.cell_face_off, nc + 1, np.int64),
        cell_faces=arr(d.cell_faces, ncf, np.int64),
        patch_names=names,
    )
    if d.cell_vtx_off:
        md.cell_vtx_off = arr(d.cell_vtx_off, nc + 1, np.int64)
        md.cell_vtx = arr(d.cell_vtx, md.cell_vtx_off[-1], np.int64)
    lib("meshgen_destroy", gen)
    return md


def box(nx, ny, nz, x=(0.0, 1.0), y=(0.0, 1.0), z=(0.0, 1.0), jitter=0.0, seed=0, tets=False) -> MeshData:
    """Synthetic box mesh (patches left/right/bottom/top/back/front)."""
    lib = hip_lib()
    p = A.BoxParams(nx, ny, nz, x[0], x[1], y[0], y[1], z[0], z[1], jitter, seed, 1 if tets else 0)
    gen = vp()
    lib("meshgen_box", C.byref(p), C.byref(gen))
    return _from_gen(lib, gen)


def read_unv(path: str) -> MeshData:
    """Ideas-UNV mesh (the reference's tests/cases format)."""
    lib = hip_lib()
    gen = vp()
    lib("meshgen_unv", path.encode(), C.byref(gen))
    return _from_gen(lib, gen)


BCSpec = Dict[str, Dict[str, object]]  # patch -> field -> kind or (kind, value)


def mesh_descriptor(data: MeshData, bcs: BCSpec, fields: Sequence = ()):
    """The prism_mesh_desc of a mesh (borrowed arrays) and the objects that keep them alive."""
    keep = []
    names = (C.c_char_p * len(data.patch_names))(*[n.encode() for n in data.patch_names])
    keep.append(names)
    bl = []
    pidx = {n: i for i, n in enumerate(data.patch_names)}
    for pname, fmap in bcs.items():
        if pname not in pidx:
            raise KeyError(f"unknown patch {pname}")
        for fname, spec in fmap.items():
            kind, val = (spec, None) if isinstance(spec, str) else spec
            b = A.BCDesc()
            b.patch = pidx[pname]
            b.field = fname.encode()
            b.kind = kind.encode()
            if val is None:
                b.value_kind = 0
            elif np.ndim(val) == 0:
                b.value_kind = 1
                b.value[0] = float(val)
            else:
                b.value_kind = 2
                for i in range(3):
                    b.value[i] = float(val[i])
            bl.append(b)
    bcs_arr = (A.BCDesc * max(1, len(bl)))(*bl)
    fl = []
    for f in fields:
        name, typ, grad = (tuple(f) + (None, None))[:3]
        fl.append(A.FieldInfo(name.encode(), (typ or "").encode(), grad.encode() if grad else None))
    f_arr = (A.FieldInfo * max(1, len(fl)))(*fl)
    d = A.MeshDesc()
    v = _f64(data.vertices).reshape(-1)
    d.n_vertices = len(data.vertices)
    d.vertices = _dp(v)
    d.n_faces = len(data.face_owner)
    i64 = lambda a: np.ascontiguousarray(a, dtype=np.int64)
    arrs = [i64(data.face_vtx_off), i64(data.face_vtx), i64(data.face_owner),
            i64(data.face_neighbor), np.ascontiguousarray(data.face_patch, dtype=np.int32),
            i64(data.cell_face_off), i64(data.cell_faces)]
    a = arrs
    d.face_vtx_off = a[0].ctypes.data_as(A.i64p)
    d.face_vtx = a[1].ctypes.data_as(A.i64p)
    d.face_owner = a[2].ctypes.data_as(A.i64p)
    d.face_neighbor = a[3].ctypes.data_as(A.i64p)
    d.face_patch = a[4].ctypes.data_as(A.i32p)
    d.n_cells = data.n_cells
    d.cell_face_off = a[5].ctypes.data_as(A.i64p)
    d.cell_faces = a[6].ctypes.data_as(A.i64p)
    d.n_patches = len(data.patch_names)
    d.patch_names = names
    d.n_bcs = len(bl)
    d.bcs = bcs_arr
    d.n_fields = len(fl)
    d.fields = f_arr
    if data.cell_vtx_off is not None:
        arrs += [i64(data.cell_vtx_off), i64(data.cell_vtx)]
        d.cell_vtx_off = a[-2].ctypes.data_as(A.i64p)
        d.cell_vtx = a[-1].ctypes.data_as(A.i64p)
    return d, (keep, v, arrs, bcs_arr, f_arr)


class Mesh:
    """mesh::PMesh on a backend.  `bcs` maps patch name -> field name -> kind string or
    (kind, value); `fields` lists (name, type, gradScheme or None) as in fields.json."""

    def __init__(self, backend: Backend, data: MeshData, bcs: BCSpec, fields: Sequence = ()):
        self.backend = backend
        self.data = data
        d, self._keep = mesh_descriptor(data, bcs, fields)
        h = vp()
        backend("mesh_create", backend.ctx, C.byref(d), C.byref(h))
        self.h = h
        info = A.MeshInfo()
        backend("mesh_get_info", h, C.byref(info))
        self.n_cells = info.n_cells
        self.n_faces = info.n_faces
        self.n_boundary_faces = info.n_boundary_faces
        self.n_nonempty_boundary_faces = info.n_nonempty_boundary_faces
        self.nnz = info.nnz
        self.n_gg_levels = info.n_gg_levels
        self._geo = None
        weakref.finalize(self, _destroy, backend, "mesh_destroy", h)

    def geometry(self) -> Dict[str, np.ndarray]:
        if self._geo is None:
            n, F = self.n_cells, self.n_faces
            g = dict(cell_center=np.zeros((n, 3)), cell_volume=np.zeros(n), face_center=np.zeros((F, 3)),
                     face_area=np.zeros((F, 3)), owner=np.zeros(F, np.int64), neighbor=np.zeros(F, np.int64),
                     face_patch=np.zeros(F, np.int32))
            self.backend("mesh_geometry", self.h, _dp(g["cell_center"]), _dp(g["cell_volume"]),
                         _dp(g["face_center"]), _dp(g["face_area"]), g["owner"].ctypes.data_as(A.i64p),
                         g["neighbor"].ctypes.data_as(A.i64p), g["face_patch"].ctypes.data_as(A.i32p))
            self._geo = g
        return self._geo

    def face_permutation(self) -> np.ndarray:
        o = np.zeros(len(self.data.face_owner), np.int64)
        self.backend("mesh_face_permutation", self.h, o.ctypes.data_as(A.i64p))
        return o

    def scalar(self, name, value=0.0, values=None, pressure=False) -> "Scalar":
        return Scalar._create(self, name, value, values, 1 if pressure else 0)

    def pressure(self, name, value=0.0, values=None) -> "Scalar":
        return Scalar._create(self, name, value, values, 1)

    def vector(self, name, value=(0.0, 0.0, 0.0), values=None) -> "Vector":
        return Vector._create(self, name, value, values)

    def tensor(self, name, value=None, values=None) -> "Tensor":
        return Tensor._create(self, name, value, values)


def _destroy(backend, fn, h):
    try:
        if backend._alive:
            backend(fn, h)
    except Exception:
        pass


# -------------------------------------------------------------------------- fields ----
class Scalar:
    """field::Scalar / field::Pressure."""

    def __init__(self, mesh: Mesh, h, name: str, owner=None):
        self.mesh = mesh
        self.h = h
        self.name = name
        self._owner = owner  # vector that owns a component handle

    @classmethod
    def _create(cls, mesh, name, value, values, kind):
        h = vp()
        v = None if values is None else _f64(values)
        if v is not None and v.shape != (mesh.n_cells,):
            raise ValueError("values must have one entry per cell")
        mesh.backend("scalar_create", mesh.h, name.encode(), kind, None if v is None else _dp(v), float(value), C.byref(h))
        s = cls(mesh, h, name)
        weakref.finalize(s, _destroy, mesh.backend, "field_destroy", h)
        return s

    @property
    def values(self) -> np.ndarray:
        out = np.zeros(self.mesh.n_cells)
        self.mesh.backend("scalar_get", self.h, _dp(out))
        return out

    @values.setter
    def values(self, v):
        v = _f64(v)
        if v.shape != (self.mesh.n_cells,):
            raise ValueError("values must have one entry per cell")
        self.mesh.backend("scalar_set", self.h, _dp(v))

    def fill(self, value: float):
        self.mesh.backend("scalar_fill", self.h, float(value))

    def set_face_values(self, fv):
        fv = _f64(fv)
        if fv.shape != (self.mesh.n_faces,):
            raise ValueError("face values must have one entry per face")
        self.mesh.backend("scalar_set_face_values", self.h, _dp(fv))

    def clear_face_values(self):
        self.mesh.backend("scalar_clear_face_values", self.h)

    def grad(self) -> np.ndarray:
        out = np.zeros((self.mesh.n_cells, 3))
        self.mesh.backend("scalar_grad", self.h, _dp(out))
        return out

    def face_values(self) -> np.ndarray:
        out = np.zeros(self.mesh.n_faces)
        self.mesh.backend("scalar_face_values", self.h, _dp(out))
        return out

    def set_history_size(self, n: int):
        self.mesh.backend("scalar_set_history", self.h, n)

    def update_prev_time_steps(self):
        self.mesh.backend("scalar_update_prev", self.h)


class Vector:
    """field::Vector / field::Velocity (three component Scalars + optional face data)."""

    def __init__(self, mesh: Mesh, h, name: str):
        self.mesh = mesh
        self.h = h
        self.name = name
        comps = []
        for k, s in enumerate("xyz"):
            ch = vp()
            mesh.backend("vector_component", h, k, C.byref(ch))
            comps.append(Scalar(mesh, ch, f"{name}_{s}", owner=self))
        self.x, self.y, self.z = comps
        weakref.finalize(self, _destroy, mesh.backend, "vfield_destroy", h)

    @classmethod
    def _create(cls, mesh, name, value, values):
        h = vp()
        v = None if values is None else _f64(values).reshape(-1)
        u3 = _f64(value)
        mesh.backend("vector_create", mesh.h, name.encode(), None if v is None else _dp(v), _dp(u3), C.byref(h))
        return cls(mesh, h, name)

    def component(self, k):
        return (self.x, self.y, self.z)[k]

    def clone(self) -> "Vector":
        h = vp()
        self.mesh.backend("vector_clone", self.h, C.byref(h))
        return Vector(self.mesh, h, self.name)

    @property
    def values(self) -> np.ndarray:
        return np.stack([self.x.values, self.y.values, self.z.values], axis=1)

    def set_face_values(self, fv):
        fv = _f64(fv).reshape(-1)
        self.mesh.backend("vector_set_face_values", self.h, _dp(fv))

    def set_face_fluxes(self, ff):
        ff = _f64(ff)
        self.mesh.backend("vector_set_face_fluxes", self.h, _dp(ff))

    def face_fluxes(self) -> np.ndarray:
        out = np.zeros(self.mesh.n_faces)
        self.mesh.backend("vector_face_fluxes", self.h, _dp(out))
        return out

    def face_values(self) -> np.ndarray:
        out = np.zeros((self.mesh.n_faces, 3))
        self.mesh.backend("vector_face_values", self.h, _dp(out))
        return out


Velocity = Vector


class Tensor:
    """field::Tensor (per-cell 3x3)."""

    def __init__(self, mesh, h, name):
        self.mesh = mesh
        self.h = h
        self.name = name
        weakref.finalize(self, _destroy, mesh.backend, "tfield_destroy", h)

    @classmethod
    def _create(cls, mesh, name, value, values):
        h = vp()
        v = None if values is None else _f64(values).reshape(-1)
        u9 = _f64(np.eye(3) if value is None else (np.eye(3) * value if np.ndim(value) == 0 else value)).reshape(-1)
        mesh.backend("tensor_create", mesh.h, name.encode(), None if v is None else _dp(v), _dp(u9), C.byref(h))
        return cls(mesh, h, name)

    @property
    def values(self) -> np.ndarray:
        out = np.zeros((self.mesh.n_cells, 3, 3))
        self.mesh.backend("tensor_get", self.h, _dp(out))
        return out


# ------------------------------------------------------------------------- schemes ----
class _Scheme:
    kind = 0
    variant = 0

    def __init__(self):
        self.refs = []

    def desc(self) -> A.SchemeDesc:
        raise NotImplementedError


class _Convection(_Scheme):
    nvf = "Upwind"

    def __init__(self, U: Vector, phi: Scalar):
        super().__init__()
        self.U, self.phi = U, phi
        self.refs = [U, phi]

    def desc(self):
        return A.SchemeDesc(A.SCHEME_CONVECTION, A.NVF[self.nvf], self.phi.h, None, None, self.U.h, 0, 0.0)


class convection:
    """scheme::convection::* (convection.h:25-150)."""

    class Upwind(_Convection):
        nvf = "Upwind"

    class CentralDifference(_Convection):
        nvf = "CentralDifference"

    class LinearUpwind(_Convection):
        nvf = "LinearUpwind"

    class QUICK(_Convection):
        nvf = "QUICK"

    class FROMM(_Convection):
        nvf = "FROMM"

    class MINMOD(_Convection):
        nvf = "MINMOD"

    class MUSCL(_Convection):
        nvf = "MUSCL"

    class SMART(_Convection):
        nvf = "SMART"


class _Diffusion(_Scheme):
    corrector = A.DIFF_NONCORRECTED

    def __init__(self, kappa, phi: Scalar, corrector: Optional[str] = None):
        super().__init__()
        self.kappa, self.phi = kappa, phi
        self.refs = [kappa, phi]
        if corrector is not None:
            self.corrector = {"over-relaxed": A.DIFF_OVERRELAXED, "minimum": A.DIFF_MINIMUM,
                              "orthogonal": A.DIFF_ORTHOGONAL}[corrector]

    def desc(self):
        ks = self.kappa.h if isinstance(self.kappa, Scalar) else None
        kt = self.kappa.h if isinstance(self.kappa, Tensor) else None
        return A.SchemeDesc(A.SCHEME_DIFFUSION, self.corrector, self.phi.h, ks, kt, None, 0, 0.0)


class diffusion:
    """scheme::diffusion::NonCorrected / Corrected (diffusion.h:40-205)."""

    class NonCorrected(_Diffusion):
        corrector = A.DIFF_NONCORRECTED

        def __init__(self, kappa, phi):
            super().__init__(kappa, phi)

    class Corrected(_Diffusion):
        corrector = A.DIFF_OVERRELAXED


class source:
    """scheme::source::* (scheme/source/*.h); sign is "+" or "-"."""

    class ConstantScalar(_Scheme):
        def __init__(self, phi: Scalar, sign="+"):
            super().__init__()
            self.phi, self.sign = phi, sign
            self.refs = [phi]

        def desc(self):
            return A.SchemeDesc(A.SCHEME_SRC_CONSTANT, int(self.sign == "-"), self.phi.h, None, None, None, 0, 0.0)

    class Gradient(_Scheme):
        def __init__(self, phi: Scalar, coord: int, sign="+"):
            super().__init__()
            self.phi, self.coord, self.sign = phi, coord, sign
            self.refs = [phi]

        def desc(self):
            return A.SchemeDesc(A.SCHEME_SRC_GRADIENT, int(self.sign == "-"), self.phi.h, None, None, None,
                                self.coord, 0.0)

    class Divergence(_Scheme):
        def __init__(self, U: Vector, sign="+"):
            super().__init__()
            self.U, self.sign = U, sign
            self.refs = [U]

        def desc(self):
            return A.SchemeDesc(A.SCHEME_SRC_DIVERGENCE, int(self.sign == "-"), None, None, None, self.U.h, 0, 0.0)

    class ImplicitField(_Scheme):
        def __init__(self, coeff: float, phi: Scalar, sign="+"):
            super().__init__()
            self.coeff, self.phi, self.sign = coeff, phi, sign
            self.refs = [phi]

        def desc(self):
            return A.SchemeDesc(A.SCHEME_SRC_IMPLICIT, int(self.sign == "-"), self.phi.h, None, None, None, 0,
                                float(self.coeff))


class temporal:
    """scheme::temporal::BackwardEuler (backward_euler.cpp), AdamMoulton (adam_moulton.cpp)."""

    class BackwardEuler(_Scheme):
        def __init__(self, phi: Scalar, dt: float):
            super().__init__()
            if not dt > 0:
                raise ValueError("prism::scheme::temporal::BackwardEuler(phi, dt): dt must be positive")
            self.phi, self.dt = phi, dt
            self.refs = [phi]

        def desc(self):
            return A.SchemeDesc(A.SCHEME_BACKWARD_EULER, 0, self.phi.h, None, None, None, 0, float(self.dt))

    class AdamMoulton(_Scheme):
        """Second-order implicit (BDF2) scheme of adam_moulton.cpp:55-86; Backward-Euler
        while the field holds only one previous level."""

        def __init__(self, phi: Scalar, dt: float):
            super().__init__()
            if not dt > 0:
                raise ValueError("prism::scheme::temporal::AdamMoulton(phi, dt): dt must be positive")
            self.phi, self.dt = phi, dt
            self.refs = [phi]

        def desc(self):
            return A.SchemeDesc(A.SCHEME_ADAM_MOULTON, 0, self.phi.h, None, None, None, 0, float(self.dt))


# ------------------------------------------------------------------------ equation ----
class Transport:
    """eqn::Transport (transport.h:27-81)."""

    def __init__(self, *schemes: _Scheme):
        if not schemes:
            raise ValueError("Transport needs at least one scheme")
        first = schemes[0]
        self.mesh = first.phi.mesh if hasattr(first, "phi") else first.U.mesh
        self.backend = self.mesh.backend
        self.schemes = list(schemes)
        descs = (A.SchemeDesc * len(schemes))(*[s.desc() for s in schemes])
        h = vp()
        self.backend("eqn_create", descs, len(schemes), C.byref(h))
        self.h = h
        self._phi = first.phi
        weakref.finalize(self, _destroy, self.backend, "eqn_destroy", h)

    @classmethod
    def _adopt(cls, backend: Backend, mesh: Mesh, h, keep=()):
        """An equation the library built (it owns its field); `keep` holds what it reads."""
        e = cls.__new__(cls)
        e.mesh = mesh
        e.backend = backend
        e.schemes = []
        e.h = h
        e._keep = tuple(keep)
        fh = backend.lib.raw("eqn_field")(h)
        e._phi = Scalar(mesh, vp(fh), "p'", owner=e)
        weakref.finalize(e, _destroy, backend, "eqn_destroy", h)
        return e

    def field(self) -> Scalar:
        return self._phi

    def update_coeffs(self):
        self.backend("eqn_update_coeffs", self.h)

    def zero_out_coeffs(self):
        self.backend("eqn_zero_coeffs", self.h)

    def set_under_relax_factor(self, f: float):
        self.backend("eqn_set_under_relax", self.h, float(f))

    def relax(self):
        self.backend("eqn_relax", self.h)

    def add_symmetry_handler(self):
        """boundaryHandlersManager().addHandler<eqn::boundary::Symmetry<Transport>>()
        (symmetry.cpp:54-82); the field must be a vector component."""
        self.backend("eqn_add_symmetry_handler", self.h)

    def matrix(self):
        """(indptr, indices, data, n) of matrix() in CSR with sorted columns."""
        nnz = C.c_int64()
        n = C.c_int64()
        self.backend("eqn_nnz", self.h, C.byref(nnz), C.byref(n))
        rp = np.zeros(n.value + 1, np.int64)
        col = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value)
        self.backend("eqn_matrix", self.h, rp.ctypes.data_as(A.i64p), col.ctypes.data_as(A.i32p), _dp(val))
        return rp, col, val, n.value

    def dense(self) -> np.ndarray:
        rp, col, val, n = self.matrix()
        M = np.zeros((n, n))
        for i in range(n):
            M[i, col[rp[i]:rp[i + 1]]] = val[rp[i]:rp[i + 1]]
        return M

    def rhs(self) -> np.ndarray:
        out = np.zeros(self.mesh.n_cells)
        self.backend("eqn_rhs", self.h, _dp(out))
        return out

    def diagonal(self) -> np.ndarray:
        out = np.zeros(self.mesh.n_cells)
        self.backend("eqn_diag", self.h, _dp(out))
        return out


Momentum = Transport


# -------------------------------------------------------------------------- solver ----
@dataclass
class SolverResult:
    iteration: int
    initial_residual: float
    final_residual: float
    converged: bool
    krylov_iters: int = 0
    refinements: int = 0
    krylov_rel_residual: float = 0.0
    restarts: int = 0

    @classmethod
    def of(cls, r: A.SolveResult):
        return cls(r.iteration, r.initial_residual, r.final_residual, bool(r.converged), r.krylov_iters,
                   r.refinements, r.krylov_rel_residual, r.restarts)


class Matrix:
    """A bare CSR matrix for ISolver::step(A, x, b)."""

    def __init__(self, backend: Backend, indptr, indices, data):
        self.backend = backend
        self.indptr = np.ascontiguousarray(indptr, np.int64)
        self.indices = np.ascontiguousarray(indices, np.int32)
        self.data = _f64(data)
        self.n = len(self.indptr) - 1
        h = vp()
        backend("csr_create", backend.ctx, self.n, self.indptr.ctypes.data_as(A.i64p),
                self.indices.ctypes.data_as(A.i32p), _dp(self.data), C.byref(h))
        self.h = h
        weakref.finalize(self, _destroy, backend, "csr_destroy", h)

    @classmethod
    def from_dense(cls, backend, M):
        M = np.asarray(M, dtype=np.float64)
        indptr = [0]
        indices, data = [], []
        for i in range(M.shape[0]):
            nz = np.nonzero(M[i])[0]
            indices.extend(nz.tolist())
            data.extend(M[i, nz].tolist())
            indptr.append(len(indices))
        return cls(backend, indptr, indices, data)

    def spmv(self, x):
        x = _f64(x)
        y = np.zeros(self.n)
        self.backend("csr_spmv", self.h, _dp(x), _dp(y))
        return y


class solver:
    """solver::BiCGSTAB / solver::Jacobi (numerics/)."""

    class BiCGSTAB:
        def __init__(self, tol=1e-14, precision="fp64", precond=None, max_iters=0, inner_tol=0.0):
            self.tol = tol
            self.precision = {"fp64": A.PRECISION_FP64, "fp32": A.PRECISION_FP32_REFINE,
                              "fp32_refine": A.PRECISION_FP32_REFINE}[precision]
            # None, "jacobi" or "sgs" (multicolour symmetric Gauss-Seidel, FP64 iteration)
            self.precond = {None: A.PRECOND_NONE, "none": A.PRECOND_NONE, "jacobi": A.PRECOND_JACOBI,
                            "sgs": A.PRECOND_SGS}[precond]
            self.max_iters = max_iters
            self.inner_tol = inner_tol

        def params(self, n_iter=10, eps=1e-7, ref_cell=None, ref_value=0.0, clear_coeffs=True) -> A.SolveParams:
            return A.SolveParams(n_iter, eps, -1 if ref_cell is None else int(ref_cell), float(ref_value),
                                 1 if clear_coeffs else 0, self.precision, self.precond, self.tol,
                                 self.max_iters, self.inner_tol)

        def solve(self, eqn: Transport, n_iter=10, eps=1e-7, ref_cell=None, ref_value=0.0,
                  clear_coeffs=True) -> SolverResult:
            p = self.params(n_iter, eps, ref_cell, ref_value, clear_coeffs)
            r = A.SolveResult()
            eqn.backend("solve", eqn.h, C.byref(p), C.byref(r))
            return SolverResult.of(r)

        def step_eqn(self, eqn: Transport) -> SolverResult:
            """BiCGSTAB::step on the equation's assembled system, in place on the device."""
            p = self.params()
            r = A.SolveResult()
            eqn.backend("eqn_step", eqn.h, C.byref(p), C.byref(r))
            return SolverResult.of(r)

        def step(self, Amat: Matrix, x, b, result=None) -> np.ndarray:
            x = _f64(x)
            b = _f64(b)
            out = np.zeros(Amat.n)
            p = self.params()
            r = A.SolveResult()
            Amat.backend("csr_bicgstab", Amat.h, _dp(b), _dp(x), _dp(out), C.byref(p), C.byref(r))
            if result is not None:
                result.append(SolverResult.of(r))
            return out

    class Jacobi:
        def step(self, Amat: Matrix, x, b) -> np.ndarray:
            x = _f64(x)
            b = _f64(b)
            out = np.zeros(Amat.n)
            Amat.backend("csr_jacobi_step", Amat.h, _dp(b), _dp(x), _dp(out))
            return out


# ---------------------------------------------------------------- SIMPLE / PISO ----
def _reports(buf, n):
    return [dict(field_name=buf[i].field_name.decode(), n_iterations=buf[i].n_iterations,
                 initial_residual=buf[i].initial_residual, final_residual=buf[i].final_residual,
                 converged=bool(buf[i].converged), krylov_iters=buf[i].krylov_iters) for i in range(n)]


class algo:
    """algo::IncompressibleSIMPLE / PRIME / PISO (algorithm/)."""

    @staticmethod
    def SIMPLEControls(backend: Backend, **kw) -> A.SimpleControls:
        c = A.SimpleControls()
        backend("simple_controls_default", C.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    @staticmethod
    def PISOControls(backend: Backend, simple: Optional[dict] = None, **kw) -> A.PisoControls:
        c = A.PisoControls()
        backend("piso_controls_default", C.byref(c))
        for k, v in kw.items():
            setattr(c, k, v)
        for k, v in (simple or {}).items():
            setattr(c.simple, k, v)
        return c

    class _Step:
        fn = ""

        def __init__(self, controls):
            self.controls = controls

        def step(self, momentum: Sequence[Transport], U: Vector, mdot: Vector, p: Scalar):
            be = U.mesh.backend
            arr = (vp * len(momentum))(*[m.h for m in momentum])
            buf = (A.Report * 256)()
            n = C.c_int32()
            be(self.fn, arr, len(momentum), U.h, mdot.h, p.h, C.byref(self.controls), buf, 256, C.byref(n))
            return _reports(buf, n.value)

    class IncompressibleSIMPLE(_Step):
        fn = "simple_step"

        def pressure_system(self, momentum: Sequence[Transport], U: Vector, mdot: Vector, p: Scalar):
            """The step up to its pressure solve: the momentum predictor, then the
            pressure-correction equation as the step would solve it (SIMPLE.cpp:19-178).
            Returns (equation, p' field, reference cell or -1, momentum reports); the step
            does not continue."""
            be = U.mesh.backend
            arr = (vp * len(momentum))(*[m.h for m in momentum])
            buf = (A.Report * 16)()
            n = C.c_int32()
            h = vp()
            ref = C.c_int64()
            be("simple_pressure_system", arr, len(momentum), U.h, mdot.h, p.h, C.byref(self.controls), buf, 16,
               C.byref(n), C.byref(h), C.byref(ref))
            eqn = Transport._adopt(be, U.mesh, h, keep=(mdot,))
            return eqn, eqn.field(), ref.value, _reports(buf, n.value)

    class PRIME(_Step):
        fn = "prime_step"

    class PISO(_Step):
        fn = "piso_step"


class adjoint:
    """Discrete adjoint of a converged steady scalar transport equation (Upwind convection,
    NonCorrected diffusion with a scalar kappa field, ConstantScalar source).  Objective:
    J = sum_R V T^2 / sum_R V over the cells with x[axis] > threshold."""

    @staticmethod
    def objective(eqn: Transport, axis: int = 0, threshold: float = 0.7) -> float:
        J = C.c_double()
        o = A.Objective(axis, threshold)
        eqn.backend("adjoint_objective", eqn.h, C.byref(o), C.byref(J))
        return J.value

    @staticmethod
    def gradient(eqn: Transport, source: Optional[Scalar] = None, kappa: Optional[Scalar] = None, axis: int = 0,
                 threshold: float = 0.7, solver: Optional["solver.BiCGSTAB"] = None):
        """(J, dJ/ds per cell or None, dJ/dkappa per cell or None, SolverResult of the A^T solve)."""
        n = eqn.mesh.n_cells
        slv = solver or globals()["solver"].BiCGSTAB()
        p = slv.params()
        r = A.SolveResult()
        J = C.c_double()
        ds = np.zeros(n) if source is not None else None
        dk = np.zeros(n) if kappa is not None else None
        o = A.Objective(axis, threshold)
        eqn.backend("adjoint_gradient", eqn.h, source.h if source is not None else None,
                    kappa.h if kappa is not None else None, C.byref(o), C.byref(p), C.byref(J),
                    _dp(ds) if ds is not None else None, _dp(dk) if dk is not None else None, C.byref(r))
        return J.value, ds, dk, SolverResult.of(r)


def needs_pressure_reference(p: Scalar) -> bool:
    """algo::needsPressureReference (SIMPLE.cpp:110-122)."""
    out = C.c_int32()
    p.mesh.backend("needs_pressure_reference", p.h, C.byref(out))
    return bool(out.value)


def courant_number(mdot: Vector, dt: float) -> np.ndarray:
    out = np.zeros(mdot.mesh.n_cells)
    mdot.mesh.backend("courant", mdot.h, float(dt), _dp(out))
    return out
