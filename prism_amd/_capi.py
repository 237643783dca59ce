"""ctypes declarations of the C ABI in include/prism_hip.h.

`bind(lib, prefix)` attaches typed prototypes for every entry point to a loaded library.
The HIP library exports them as ``prism_*``; any other library exposing the same
signatures under another prefix can be bound the same way.
"""
from __future__ import annotations

import ctypes as C
import os

PRISM_OK = 0
PRISM_ERR_INVALID = 1
PRISM_ERR_RUNTIME = 2
PRISM_ERR_BC = 3
PRISM_ERR_MESH = 4
PRISM_ERR_DEVICE = 5
PRISM_ERR_NOMEM = 6

SCHEME_CONVECTION = 1
SCHEME_DIFFUSION = 2
SCHEME_SRC_CONSTANT = 3
SCHEME_SRC_GRADIENT = 4
SCHEME_SRC_DIVERGENCE = 5
SCHEME_SRC_IMPLICIT = 6
SCHEME_BACKWARD_EULER = 7
SCHEME_ADAM_MOULTON = 8

NVF = {"Upwind": 0, "CentralDifference": 1, "LinearUpwind": 2, "QUICK": 3, "FROMM": 4,
       "MINMOD": 5, "MUSCL": 6, "SMART": 7}
DIFF_NONCORRECTED = 0
DIFF_OVERRELAXED = 1
DIFF_MINIMUM = 2
DIFF_ORTHOGONAL = 3
SIGN_POSITIVE = 0
SIGN_NEGATIVE = 1
PRECISION_FP64 = 0
PRECISION_FP32_REFINE = 1
PRECOND_NONE = 0
PRECOND_JACOBI = 1
PRECOND_SGS = 2

vp = C.c_void_p
dp = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int32)


class BCDesc(C.Structure):
    _fields_ = [("patch", C.c_int32), ("field", C.c_char_p), ("kind", C.c_char_p),
                ("value_kind", C.c_int32), ("value", C.c_double * 3)]


class FieldInfo(C.Structure):
    _fields_ = [("name", C.c_char_p), ("type", C.c_char_p), ("grad_scheme", C.c_char_p)]


class MeshDesc(C.Structure):
    _fields_ = [("n_vertices", C.c_int64), ("vertices", dp),
                ("n_faces", C.c_int64), ("face_vtx_off", i64p), ("face_vtx", i64p),
                ("face_owner", i64p), ("face_neighbor", i64p), ("face_patch", i32p),
                ("n_cells", C.c_int64), ("cell_face_off", i64p), ("cell_faces", i64p),
                ("n_patches", C.c_int32), ("patch_names", C.POINTER(C.c_char_p)),
                ("n_bcs", C.c_int32), ("bcs", C.POINTER(BCDesc)),
                ("n_fields", C.c_int32), ("fields", C.POINTER(FieldInfo)),
                ("cell_vtx_off", i64p), ("cell_vtx", i64p)]


class ExportItem(C.Structure):
    _fields_ = [("kind", C.c_int32), ("field", vp)]


EXPORT_SCALAR, EXPORT_VECTOR, EXPORT_TENSOR = 1, 3, 9
WRITE_ASCII, WRITE_RAW_BINARY_COMPRESSED = 0, 1


class MeshInfo(C.Structure):
    _fields_ = [("n_cells", C.c_int64), ("n_faces", C.c_int64), ("n_boundary_faces", C.c_int64),
                ("n_nonempty_boundary_faces", C.c_int64), ("nnz", C.c_int64),
                ("n_gg_levels", C.c_int64)]


class BoxParams(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
                ("x0", C.c_double), ("x1", C.c_double), ("y0", C.c_double), ("y1", C.c_double),
                ("z0", C.c_double), ("z1", C.c_double), ("jitter", C.c_double),
                ("seed", C.c_uint64), ("tets", C.c_int32)]


class SchemeDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("variant", C.c_int32), ("phi", vp), ("kappa", vp),
                ("kappa_t", vp), ("vec", vp), ("coord", C.c_int32), ("coeff", C.c_double)]


class SolveParams(C.Structure):
    _fields_ = [("n_iter", C.c_int64), ("eps", C.c_double), ("ref_cell", C.c_int64),
                ("ref_value", C.c_double), ("clear_coeffs", C.c_int32), ("precision", C.c_int32),
                ("precond", C.c_int32), ("tol", C.c_double), ("max_iters", C.c_int64),
                ("inner_tol", C.c_double)]


class SolveResult(C.Structure):
    _fields_ = [("iteration", C.c_int64), ("initial_residual", C.c_double),
                ("final_residual", C.c_double), ("converged", C.c_int32),
                ("krylov_iters", C.c_int64), ("refinements", C.c_int64),
                ("krylov_rel_residual", C.c_double), ("restarts", C.c_int32)]


class SimpleControls(C.Structure):
    _fields_ = [("momentum_urf", C.c_double), ("momentum_max_iter", C.c_int64),
                ("momentum_residual", C.c_double), ("non_ortho_correctors", C.c_int64),
                ("pressure_urf", C.c_double), ("pressure_max_iter", C.c_int64),
                ("pressure_residual", C.c_double), ("p_ref_cell", C.c_int64),
                ("p_ref_value", C.c_double), ("momentum_precision", C.c_int32),
                ("pressure_precision", C.c_int32), ("precond", C.c_int32),
                ("momentum_tol", C.c_double), ("pressure_tol", C.c_double),
                ("max_krylov_iters", C.c_int64)]


class PisoControls(C.Structure):
    _fields_ = [("outer_iterations", C.c_int64), ("momentum_implicit_steps", C.c_int64),
                ("pressure_correction_steps", C.c_int64), ("simple", SimpleControls)]


class Report(C.Structure):
    _fields_ = [("field_name", C.c_char * 64), ("n_iterations", C.c_int64),
                ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("converged", C.c_int32), ("krylov_iters", C.c_int64)]


class Objective(C.Structure):
    _fields_ = [("axis", C.c_int32), ("threshold", C.c_double)]


I = C.c_int
PP = C.POINTER(vp)

# name -> (restype, argtypes)
PROTOS = {
    "last_error": (C.c_char_p, []),
    "ctx_create": (I, [C.c_int, PP]),
    "ctx_destroy": (I, [vp]),
    "ctx_info": (I, [vp, C.c_char_p, C.c_int32, i32p]),
    "ctx_set_timing": (I, [vp, C.c_int32]),
    "ctx_reset_timing": (I, [vp]),
    "ctx_kernel_stats": (I, [vp, C.c_int32, C.c_char_p, C.c_int32, i64p, dp]),
    "ctx_synchronize": (I, [vp]),
    "mesh_create": (I, [vp, C.POINTER(MeshDesc), PP]),
    "mesh_destroy": (I, [vp]),
    "mesh_get_info": (I, [vp, C.POINTER(MeshInfo)]),
    "mesh_geometry": (I, [vp, dp, dp, dp, dp, i64p, i64p, i32p]),
    "mesh_face_permutation": (I, [vp, i64p]),
    "meshgen_box": (I, [C.POINTER(BoxParams), PP]),
    "meshgen_unv": (I, [C.c_char_p, PP]),
    "meshgen_desc": (I, [vp, C.POINTER(MeshDesc)]),
    "meshgen_patch_name": (I, [vp, C.c_int32, C.c_char_p, C.c_int32]),
    "meshgen_destroy": (I, [vp]),
    "scalar_create": (I, [vp, C.c_char_p, C.c_int32, dp, C.c_double, PP]),
    "field_destroy": (I, [vp]),
    "scalar_set": (I, [vp, dp]),
    "scalar_get": (I, [vp, dp]),
    "scalar_fill": (I, [vp, C.c_double]),
    "scalar_set_face_values": (I, [vp, dp]),
    "scalar_clear_face_values": (I, [vp]),
    "scalar_grad": (I, [vp, dp]),
    "scalar_face_values": (I, [vp, dp]),
    "scalar_set_history": (I, [vp, C.c_int32]),
    "scalar_update_prev": (I, [vp]),
    "vector_create": (I, [vp, C.c_char_p, dp, dp, PP]),
    "vector_clone": (I, [vp, PP]),
    "vfield_destroy": (I, [vp]),
    "vector_component": (I, [vp, C.c_int32, PP]),
    "vector_set_face_values": (I, [vp, dp]),
    "vector_set_face_fluxes": (I, [vp, dp]),
    "vector_face_fluxes": (I, [vp, dp]),
    "vector_face_values": (I, [vp, dp]),
    "tensor_create": (I, [vp, C.c_char_p, dp, dp, PP]),
    "tfield_destroy": (I, [vp]),
    "tensor_get": (I, [vp, dp]),
    "eqn_create": (I, [C.POINTER(SchemeDesc), C.c_int32, PP]),
    "eqn_destroy": (I, [vp]),
    "eqn_update_coeffs": (I, [vp]),
    "eqn_zero_coeffs": (I, [vp]),
    "eqn_set_under_relax": (I, [vp, C.c_double]),
    "eqn_relax": (I, [vp]),
    "eqn_field": (vp, [vp]),
    "eqn_add_symmetry_handler": (I, [vp]),
    "needs_pressure_reference": (I, [vp, i32p]),
    "eqn_nnz": (I, [vp, i64p, i64p]),
    "eqn_matrix": (I, [vp, i64p, i32p, dp]),
    "eqn_rhs": (I, [vp, dp]),
    "eqn_diag": (I, [vp, dp]),
    "solve_params_default": (None, [C.POINTER(SolveParams)]),
    "solve": (I, [vp, C.POINTER(SolveParams), C.POINTER(SolveResult)]),
    "eqn_step": (I, [vp, C.POINTER(SolveParams), C.POINTER(SolveResult)]),
    "eqn_step_noalloc": (I, [vp, C.POINTER(SolveParams), C.POINTER(SolveResult)]),
    "csr_create": (I, [vp, C.c_int64, i64p, i32p, dp, PP]),
    "csr_destroy": (I, [vp]),
    "csr_bicgstab": (I, [vp, dp, dp, dp, C.POINTER(SolveParams), C.POINTER(SolveResult)]),
    "csr_jacobi_step": (I, [vp, dp, dp, dp]),
    "csr_spmv": (I, [vp, dp, dp]),
    "simple_controls_default": (None, [C.POINTER(SimpleControls)]),
    "piso_controls_default": (None, [C.POINTER(PisoControls)]),
    "simple_step": (I, [PP, C.c_int32, vp, vp, vp, C.POINTER(SimpleControls), C.POINTER(Report), C.c_int32, i32p]),
    "simple_pressure_system": (I, [PP, C.c_int32, vp, vp, vp, C.POINTER(SimpleControls), C.POINTER(Report), C.c_int32,
                                   i32p, PP, i64p]),
    "prime_step": (I, [PP, C.c_int32, vp, vp, vp, C.POINTER(SimpleControls), C.POINTER(Report), C.c_int32, i32p]),
    "piso_step": (I, [PP, C.c_int32, vp, vp, vp, C.POINTER(PisoControls), C.POINTER(Report), C.c_int32, i32p]),
    "courant": (I, [vp, C.c_double, dp]),
    "kepsilon_update": (I, [vp, vp, vp, vp, dp, PP]),
    "mesh_n_vertices": (I, [vp, i64p]),
    "field_to_points": (I, [vp, dp]),
    "vfield_to_points": (I, [vp, dp]),
    "tfield_to_points": (I, [vp, dp]),
    "export_vtu": (I, [C.POINTER(ExportItem), C.c_int32, C.c_char_p, C.c_int32]),
    "write_pvd": (I, [C.c_char_p, C.c_int32, dp, C.POINTER(C.c_char_p)]),
    "adjoint_objective": (I, [vp, C.POINTER(Objective), dp]),
    "adjoint_gradient": (I, [vp, vp, vp, C.POINTER(Objective), C.POINTER(SolveParams), dp, dp, dp,
                             C.POINTER(SolveResult)]),
}

# Entry points that only the HIP library provides (host-side mesh generation, device contexts).
HIP_ONLY = {"eqn_step_noalloc", "ctx_info", "ctx_set_timing", "ctx_reset_timing", "ctx_kernel_stats", "ctx_synchronize",
            "meshgen_box", "meshgen_unv", "meshgen_desc", "meshgen_patch_name", "meshgen_destroy"}


class PrismError(RuntimeError):
    """Raised for a non-zero status; `code` holds the PRISM_ERR_* value."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class BoundaryConditionError(PrismError):
    """error::NonImplementedBoundaryCondition."""


class InvalidMeshError(PrismError):
    """error::InvalidMesh."""


class Lib:
    """A loaded C-ABI library with typed entry points; calls raise PrismError on failure."""

    def __init__(self, path: str, prefix: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built")
        self.path = path
        self.prefix = prefix
        self.dll = C.CDLL(path)
        self.fns = {}
        for name, (res, args) in PROTOS.items():
            sym = prefix + name
            if not hasattr(self.dll, sym):
                continue
            f = getattr(self.dll, sym)
            f.restype = res
            f.argtypes = args
            self.fns[name] = f

    def has(self, name):
        return name in self.fns

    def raw(self, name):
        return self.fns[name]

    def __call__(self, name, *args):
        f = self.fns.get(name)
        if f is None:
            raise PrismError(PRISM_ERR_RUNTIME, f"{self.prefix}{name} is not provided by {self.path}")
        rc = f(*args)
        if f.restype is I and rc != PRISM_OK:
            msg = self.fns["last_error"]().decode(errors="replace")
            if rc == PRISM_ERR_BC:
                raise BoundaryConditionError(rc, msg)
            if rc == PRISM_ERR_MESH:
                raise InvalidMeshError(rc, msg)
            if rc == PRISM_ERR_INVALID:
                raise ValueError(msg)
            raise PrismError(rc, msg)
        return rc
