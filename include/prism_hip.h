/*
 * prism_hip.h -- C ABI of the MI355X finite-volume assembly + BiCGSTAB library.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and returns an int
 * status (PRISM_OK == 0).  No exception crosses the ABI; the message of the last failure
 * on the calling thread is returned by prism_last_error().  Host arrays passed in are
 * borrowed for the duration of the call only.  All device work of one context runs in
 * stream order on one HIP stream, and every call is synchronous at return (its results
 * are ready) unless it says otherwise.  Results are deterministic: no kernel uses a
 * floating-point atomic.
 *
 * The reference (eigemx/prism) has no FFI: its "plugin API" is the C++ class hierarchy.
 * Each group below names the C++ interface it replaces (file:line in the reference).
 */
#ifndef PRISM_HIP_H
#define PRISM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status codes ---- */
/* Mapping to the reference's exceptions (exceptions.h:9-27, solver.h:8-23):            */
#define PRISM_OK 0
#define PRISM_ERR_INVALID 1  /* std::invalid_argument (e.g. dt <= 0, temporal.cpp:9-12) */
#define PRISM_ERR_RUNTIME 2  /* std::runtime_error (bad sizes, linear.cpp:22-26)         */
#define PRISM_ERR_BC 3       /* error::NonImplementedBoundaryCondition (boundary.h:109)  */
#define PRISM_ERR_MESH 4     /* error::InvalidMesh (unv.cpp:66-69)                       */
#define PRISM_ERR_DEVICE 5   /* HIP runtime failure                                      */
#define PRISM_ERR_NOMEM 6    /* device allocation failure                                */

/* Message of the last failing call on this thread ("" if none). */
const char* prism_last_error(void);

/* ---------------------------------------------------------------- opaque handles ---- */
typedef struct prism_ctx prism_ctx;         /* one device + one HIP stream             */
typedef struct prism_mesh prism_mesh;       /* mesh::PMesh (pmesh.h:14-175)            */
typedef struct prism_field prism_field;     /* field::Scalar / Pressure (scalar.h:14)  */
typedef struct prism_vfield prism_vfield;   /* field::Vector / Velocity (vector.h:17)  */
typedef struct prism_tfield prism_tfield;   /* field::Tensor (tensor.h:7-38)           */
typedef struct prism_eqn prism_eqn;         /* eqn::Transport (transport.h:27-81)      */
typedef struct prism_csr prism_csr;         /* a bare SparseMatrix for ISolver::step   */
typedef struct prism_meshgen prism_meshgen; /* synthetic mesh builder (host memory)    */

/* ------------------------------------------------------------------------ context ---- */
int prism_ctx_create(int device, prism_ctx** out);
int prism_ctx_destroy(prism_ctx* ctx);
/* Device name and compute-unit count, for reports. */
int prism_ctx_info(prism_ctx* ctx, char* name, int32_t name_len, int32_t* n_cu);
/* Kernel timing on the context stream.  When enabled, every launch of every kernel records
 * its own start and end (hipExtLaunchKernelGGL's event pair: the dispatch's timestamps, the
 * durations a kernel trace reports); prism_ctx_kernel_stats() returns, per kernel name,
 * the launch count and the summed duration in ms.  Off by default. */
int prism_ctx_set_timing(prism_ctx* ctx, int32_t enabled);
int prism_ctx_reset_timing(prism_ctx* ctx);
int prism_ctx_kernel_stats(prism_ctx* ctx, int32_t idx, char* name, int32_t name_len,
                           int64_t* launches, double* total_ms); /* PRISM_ERR_INVALID past end */
int prism_ctx_synchronize(prism_ctx* ctx);

/* --------------------------------------------------------------------------- mesh ---- */
/* Replaces mesh::PMesh's constructor (pmesh.cpp:10-49) fed by UnvToPMeshConverter
 * (unv.cpp:61-110) and mesh::MeshBoundary (mesh/boundary.cpp:291-319).
 * Face vertex lists are in owner-outward order, boundary faces may appear anywhere: the
 * library applies the reference's [non-empty boundary | empty boundary | interior]
 * permutation (pmesh.cpp:51-119), keeping the relative order inside each group.          */
typedef struct prism_bc_desc {
    int32_t patch;         /* index into patch_names                                   */
    const char* field;     /* field name, e.g. "T", "U", "P"                            */
    const char* kind;      /* "fixed" "no-slip" "symmetry" "outlet" "zero-gradient"
                              "fixed-gradient" "empty"  (mesh/boundary.cpp:154-202)     */
    int32_t value_kind;    /* 0 none, 1 scalar (value[0]), 2 vector (value[0..2])       */
    double value[3];
} prism_bc_desc;

typedef struct prism_field_info {
    const char* name;        /* fields.json "name"                                     */
    const char* type;        /* fields.json "type" (informational)                     */
    const char* grad_scheme; /* "greenGauss"/"green-gauss" or NULL = least squares      */
} prism_field_info;

typedef struct prism_mesh_desc {
    int64_t n_vertices;
    const double* vertices;        /* [n_vertices][3]                                  */
    int64_t n_faces;
    const int64_t* face_vtx_off;   /* [n_faces+1]                                      */
    const int64_t* face_vtx;       /* vertex ids, owner-outward order                  */
    const int64_t* face_owner;     /* [n_faces]                                        */
    const int64_t* face_neighbor;  /* [n_faces], -1 on boundary faces                  */
    const int32_t* face_patch;     /* [n_faces], -1 on interior faces                  */
    int64_t n_cells;
    const int64_t* cell_face_off;  /* [n_cells+1]                                      */
    const int64_t* cell_faces;     /* face ids, in the cell's local face order         */
    int32_t n_patches;
    const char* const* patch_names;
    int32_t n_bcs;
    const prism_bc_desc* bcs;
    int32_t n_fields;
    const prism_field_info* fields;
    /* Optional (may be NULL): each cell's vertex ids in the element's own order (UNV, or
     * VTK's for generated boxes), as Cell::verticesIds() (mesh/cell.h); output export uses
     * them for the VTU connectivity and the vertex -> cell adjacency. */
    const int64_t* cell_vtx_off;   /* [n_cells+1]                                      */
    const int64_t* cell_vtx;
} prism_mesh_desc;

typedef struct prism_mesh_info {
    int64_t n_cells, n_faces, n_boundary_faces, n_nonempty_boundary_faces, nnz;
    int64_t n_gg_levels;             /* Green-Gauss sweep levels (wavefronts)          */
} prism_mesh_info;

int prism_mesh_create(prism_ctx* ctx, const prism_mesh_desc* desc, prism_mesh** out);
int prism_mesh_destroy(prism_mesh* mesh);
int prism_mesh_get_info(prism_mesh* mesh, prism_mesh_info* info);
/* Geometry in the permuted face order (host copies; any pointer may be NULL):
 * cell centers [n][3], volumes [n], face centers [F][3], area vectors [F][3],
 * owner [F], neighbor [F] (-1 boundary), face_patch [F]. */
int prism_mesh_geometry(prism_mesh* mesh, double* cell_center, double* cell_volume,
                        double* face_center, double* face_area_vec, int64_t* owner,
                        int64_t* neighbor, int32_t* face_patch);
/* Old (input) face index -> permuted index, [n_faces]. */
int prism_mesh_face_permutation(prism_mesh* mesh, int64_t* old_to_new);

/* Synthetic meshes: an nx*ny*nz box of hexahedra on [x0,x1]x[y0,y1]x[z0,z1], interior
 * vertices jittered uniformly by +-jitter*h per axis (counter-based RNG keyed on seed and
 * vertex id), optionally each hex split into 6 tetrahedra sharing the main diagonal.
 * Cells are numbered x-fastest; hex local faces follow the reference's UNV hex table
 * (unv.cpp:19-31), tets the tetra table (unv.cpp:34-44).  Patches: 0 "left" (x0),
 * 1 "right" (x1), 2 "bottom" (y0), 3 "top" (y1), 4 "back" (z0), 5 "front" (z1).  The
 * descriptor returned by prism_meshgen_desc points into the generator's own arrays and
 * has no BCs/fields; the caller fills those in before prism_mesh_create. */
typedef struct prism_box_params {
    int64_t nx, ny, nz;
    double x0, x1, y0, y1, z0, z1;
    double jitter;     /* fraction of the local spacing, 0 = regular                  */
    uint64_t seed;
    int32_t tets;      /* 0 hexahedra, 1 six tetrahedra per hexahedron                */
} prism_box_params;
int prism_meshgen_box(const prism_box_params* params, prism_meshgen** out);
int prism_meshgen_desc(prism_meshgen* gen, prism_mesh_desc* desc);
int prism_meshgen_destroy(prism_meshgen* gen);

/* Ideas-UNV reader (datasets 2411 vertices, 2412 elements, 2467 groups), following
 * UnvToPMeshConverter (unv.cpp:61-372): boundary quads/triangles define patch faces,
 * hex/tetra/wedge cells define interior faces by first-owner/second-neighbour. */
int prism_meshgen_unv(const char* path, prism_meshgen** out);
int prism_meshgen_patch_name(prism_meshgen* gen, int32_t idx, char* name, int32_t len);

/* -------------------------------------------------------------------------- fields ---- */
/* field::Scalar (scalar.h:14-127).  kind: 0 Scalar, 1 Pressure.  field::Pressure is a
 * Scalar with the same default boundary handlers (pressure.h:6-9, scalar.cpp:239-248): its
 * "symmetry" faces take the owner value like any scalar's, because the extrapolating
 * Symmetry<Pressure> of pressure_boundary.cpp is never registered by the reference.
 * values may be NULL (uniform `value`).  The name selects the boundary conditions and the
 * gradient scheme (scalar.cpp:260-304). */
int prism_scalar_create(prism_mesh* mesh, const char* name, int32_t kind, const double* values,
                        double value, prism_field** out);
int prism_field_destroy(prism_field* f);
int prism_scalar_set(prism_field* f, const double* values);     /* values() = ...          */
int prism_scalar_get(prism_field* f, double* values);           /* const values()          */
int prism_scalar_fill(prism_field* f, double value);            /* values().setConstant()  */
int prism_scalar_set_face_values(prism_field* f, const double* face_values); /* [F]       */
int prism_scalar_clear_face_values(prism_field* f);
/* gradAtCell for every cell ([n][3]); lazy recompute as gradient.h:26-31. */
int prism_scalar_grad(prism_field* f, double* grad);
/* valueAtFace for every face ([F]); boundary faces through the BC handlers. */
int prism_scalar_face_values(prism_field* f, double* out);
/* History ring (history.cpp:30-61) for temporal schemes. */
int prism_scalar_set_history(prism_field* f, int32_t n_steps);
int prism_scalar_update_prev(prism_field* f);

/* field::Vector / field::Velocity (vector.h:17-101, velocity.h:18-23); values [n][3]. */
int prism_vector_create(prism_mesh* mesh, const char* name, const double* values,
                        const double* uniform3, prism_vfield** out);
int prism_vector_clone(prism_vfield* src, prism_vfield** out);   /* GeneralVector::clone */
int prism_vfield_destroy(prism_vfield* v);
int prism_vector_component(prism_vfield* v, int32_t coord, prism_field** out); /* borrowed */
int prism_vector_set_face_values(prism_vfield* v, const double* face_values);  /* [F][3] */
int prism_vector_set_face_fluxes(prism_vfield* v, const double* face_fluxes);  /* [F]    */
int prism_vector_face_fluxes(prism_vfield* v, double* out);    /* fluxAtFace, [F]        */
int prism_vector_face_values(prism_vfield* v, double* out);    /* valueAtFace, [F][3]    */

/* field::Tensor (tensor.h:7-38); values [n][3][3] row-major, or NULL with uniform [3][3]. */
int prism_tensor_create(prism_mesh* mesh, const char* name, const double* values,
                        const double* uniform9, prism_tfield** out);
int prism_tfield_destroy(prism_tfield* t);
int prism_tensor_get(prism_tfield* t, double* values);

/* ---------------------------------------------------------------------- equations ---- */
/* Scheme descriptors for eqn::Transport(Scheme&&...) (transport.h:88-139). */
#define PRISM_SCHEME_CONVECTION 1     /* convection::* (convection.h:25-150)              */
#define PRISM_SCHEME_DIFFUSION 2      /* diffusion::NonCorrected / Corrected              */
#define PRISM_SCHEME_SRC_CONSTANT 3   /* source::ConstantScalar<Sign> (constant_scalar.h) */
#define PRISM_SCHEME_SRC_GRADIENT 4   /* source::Gradient<Sign> (source/gradient.h)       */
#define PRISM_SCHEME_SRC_DIVERGENCE 5 /* source::Divergence<Sign> (divergence.h)          */
#define PRISM_SCHEME_SRC_IMPLICIT 6   /* source::ImplicitField<Sign> (implicit_field.h)   */
#define PRISM_SCHEME_BACKWARD_EULER 7 /* temporal::BackwardEuler (backward_euler.cpp)     */
#define PRISM_SCHEME_ADAM_MOULTON 8   /* temporal::AdamMoulton (adam_moulton.cpp:55-86)   */

/* NVF weight tables of convection.cpp:69-128 */
#define PRISM_NVF_UPWIND 0
#define PRISM_NVF_CD 1
#define PRISM_NVF_LUD 2
#define PRISM_NVF_QUICK 3
#define PRISM_NVF_FROMM 4
#define PRISM_NVF_MINMOD 5
#define PRISM_NVF_MUSCL 6
#define PRISM_NVF_SMART 7

/* non-orthogonal correctors of nonortho.cpp:7-21 (diffusion::Corrected) */
#define PRISM_DIFF_NONCORRECTED 0
#define PRISM_DIFF_OVERRELAXED 1
#define PRISM_DIFF_MINIMUM 2
#define PRISM_DIFF_ORTHOGONAL 3

#define PRISM_SIGN_POSITIVE 0
#define PRISM_SIGN_NEGATIVE 1

typedef struct prism_scheme_desc {
    int32_t kind;            /* PRISM_SCHEME_*                                         */
    int32_t variant;         /* NVF id / PRISM_DIFF_* / PRISM_SIGN_*                   */
    prism_field* phi;        /* transported field, or the source field                 */
    prism_field* kappa;      /* scalar diffusion coefficient, or NULL                  */
    prism_tfield* kappa_t;   /* tensor diffusion coefficient, or NULL                  */
    prism_vfield* vec;       /* U for convection, the divergence argument              */
    int32_t coord;           /* 0,1,2 for source::Gradient                             */
    double coeff;            /* ImplicitField coefficient; BackwardEuler/AdamMoulton dt*/
} prism_scheme_desc;

int prism_eqn_create(const prism_scheme_desc* schemes, int32_t n, prism_eqn** out);
int prism_eqn_destroy(prism_eqn* e);
int prism_eqn_update_coeffs(prism_eqn* e);                 /* transport.cpp:4-21  */
int prism_eqn_zero_coeffs(prism_eqn* e);                   /* transport.cpp:23-36 */
int prism_eqn_set_under_relax(prism_eqn* e, double factor);/* transport.cpp:42-52 */
int prism_eqn_relax(prism_eqn* e);                         /* transport.cpp:54-64 */
/* eqn.boundaryHandlersManager().addHandler<eqn::boundary::Symmetry<Transport>>()
   (equation/boundary.h:64, symmetry.cpp:54-82): at the end of every update_coeffs, each
   face of a patch whose condition for the equation's field is "symmetry" adds
   g n_k^2 to its owner's diagonal and -g (U_C . n - U_k n_k) n_k to its rhs, with
   g = 2 |kappa_C^T S_f| / ((x_f - x_C) . n), U the (last) convection scheme's velocity and
   kappa the (last) diffusion scheme's coefficient.  The field must be a vector component. */
int prism_eqn_add_symmetry_handler(prism_eqn* e);
prism_field* prism_eqn_field(prism_eqn* e);
/* matrix()/rhs() mirrors: CSR with sorted columns, rhs [n], diagonal [n]. */
int prism_eqn_nnz(prism_eqn* e, int64_t* nnz, int64_t* n);
int prism_eqn_matrix(prism_eqn* e, int64_t* row_ptr, int32_t* col, double* val);
int prism_eqn_rhs(prism_eqn* e, double* rhs);
int prism_eqn_diag(prism_eqn* e, double* diag);

/* -------------------------------------------------------------------------- solver ---- */
/* ISolver::solve (solver.h:43-48, solver.cpp:27-82) with BiCGSTAB::step (bicgstab.cpp:7-9).
 * The outer loop (assemble, relax, reference cell, n_iter steps with the scaled residual
 * max|A x - b| / (max|diag*x| + eps)) is the reference's.  One step is a BiCGSTAB solve
 * from the current x to ||b - A x|| <= tol * ||b||, either in FP64, or (PRECISION_FP32)
 * as FP32 BiCGSTAB inside FP64 iterative refinement: FP32 matrix/vectors, FP64 residual
 * b - A x, FP64 dot products and FP64 solution accumulation. */
#define PRISM_PRECISION_FP64 0
#define PRISM_PRECISION_FP32_REFINE 1
#define PRISM_PRECOND_NONE 0
#define PRISM_PRECOND_JACOBI 1
/* Multicolour symmetric Gauss-Seidel, M = (D + L) D^-1 (D + U) with the rows ordered colour
 * by colour (greedy colouring of the pattern; red-black on structured hexahedra), applied
 * right as Eigen applies a preconditioner (BiCGSTAB.h).  Stands in for the reference's
 * IncompleteLUT (bicgstab.h:12) where unpreconditioned BiCGSTAB does not converge
 * (convection-dominated central-difference and limited-scheme systems); the converged x is
 * that of A x = b.  The iteration runs in FP64 in every precision mode; needs every
 * diagonal entry nonzero; not available in prism_eqn_step_noalloc. */
#define PRISM_PRECOND_SGS 2

typedef struct prism_solve_params {
    int64_t n_iter;        /* outer steps (reference default 10)                        */
    double eps;            /* scaled-residual stop (reference default 1e-7)             */
    int64_t ref_cell;      /* -1: none                                                  */
    double ref_value;
    int32_t clear_coeffs;  /* reference default 1                                       */
    int32_t precision;     /* PRISM_PRECISION_*                                         */
    int32_t precond;       /* PRISM_PRECOND_*                                           */
    double tol;            /* Krylov relative residual target per step                  */
    int64_t max_iters;     /* Krylov iteration cap per step (<=0: 2*n like Eigen)        */
    double inner_tol;      /* FP32 inner relative tolerance per refinement (<=0: auto)  */
} prism_solve_params;

typedef struct prism_solve_result {
    int64_t iteration;          /* SolverResult::iteration()                            */
    double initial_residual;    /* SolverResult::initialResidual()                      */
    double final_residual;      /* SolverResult::finalResidual()                        */
    int32_t converged;          /* SolverResult::hasConverged()                         */
    int64_t krylov_iters;       /* total BiCGSTAB iterations over all steps             */
    int64_t refinements;        /* FP64 residual evaluations (FP32_REFINE)              */
    double krylov_rel_residual; /* ||b - A x|| / ||b|| after the last step (FP64)       */
    int32_t restarts;           /* rho breakdown restarts                               */
} prism_solve_result;

void prism_solve_params_default(prism_solve_params* p);
int prism_solve(prism_eqn* e, const prism_solve_params* p, prism_solve_result* r);
/* One BiCGSTAB::step (bicgstab.cpp:7-9) on the equation's current (already assembled) matrix
 * and rhs, from and into the equation field's device values: no assembly, no host copies. */
int prism_eqn_step(prism_eqn* e, const prism_solve_params* p, prism_solve_result* r);

/* prism_eqn_step without device allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Every Krylov vector, partial-sum buffer and value copy
 * of the step lives in one static device array of 24 MiB per device (held from library load
 * until the process ends), which bounds the system: about 65,000 rows for an FP64 step on a
 * 7-diagonal box matrix, fewer with FP32 refinement or Jacobi; larger systems are refused
 * with PRISM_ERR_INVALID and the needed size in prism_last_error().  The caller's matrix,
 * right-hand side and solution stay where they are (the equation's device buffers).  All
 * calls on one device share that array: calls that overlap on one device (other threads,
 * other contexts) race and must be serialised by the caller. */
int prism_eqn_step_noalloc(prism_eqn* e, const prism_solve_params* p, prism_solve_result* r);

/* A bare matrix for ISolver::step(A, x, b) (solver.h:49) and Jacobi::step (jacobi.cpp:5-9). */
int prism_csr_create(prism_ctx* ctx, int64_t n, const int64_t* row_ptr, const int32_t* col,
                     const double* val, prism_csr** out);
int prism_csr_destroy(prism_csr* a);
/* x_out = BiCGSTAB(A, b, x_in) to p->tol; x_in and x_out may alias. */
int prism_csr_bicgstab(prism_csr* a, const double* b, const double* x_in, double* x_out,
                       const prism_solve_params* p, prism_solve_result* r);
int prism_csr_jacobi_step(prism_csr* a, const double* b, const double* x_in, double* x_out);
int prism_csr_spmv(prism_csr* a, const double* x, double* y);

/* ----------------------------------------------------------------- SIMPLE / PISO ---- */
/* algo::IncompressibleSIMPLE (SIMPLE.h:14-55), algo::PRIME (PRIME.h:13-42),
 * algo::PISO (PISO.h:11-61).  Momentum equations are passed as an array of 2 or 3. */
typedef struct prism_simple_controls {
    double momentum_urf;           /* 0.7  */
    int64_t momentum_max_iter;     /* 3    */
    double momentum_residual;      /* 1e-7 */
    int64_t non_ortho_correctors;  /* 2    */
    double pressure_urf;           /* 0.3  */
    int64_t pressure_max_iter;     /* 5    */
    double pressure_residual;      /* 1e-7 */
    int64_t p_ref_cell;            /* -1: automatic (cell 0 when all-Neumann)            */
    double p_ref_value;
    /* Krylov settings of the inner steps (no reference counterpart: Eigen defaults). */
    int32_t momentum_precision, pressure_precision;
    int32_t precond;
    double momentum_tol, pressure_tol;
    int64_t max_krylov_iters;
} prism_simple_controls;

typedef struct prism_piso_controls {
    int64_t outer_iterations;          /* 1 */
    int64_t momentum_implicit_steps;   /* 1 */
    int64_t pressure_correction_steps; /* 2 */
    prism_simple_controls simple;      /* urf 1.0, residuals 1e-4 by default for PISO  */
} prism_piso_controls;

typedef struct prism_report {
    char field_name[64];
    int64_t n_iterations;
    double initial_residual, final_residual;
    int32_t converged;
    int64_t krylov_iters;
} prism_report;

void prism_simple_controls_default(prism_simple_controls* c);
void prism_piso_controls_default(prism_piso_controls* c);
int prism_simple_step(prism_eqn* const* momentum, int32_t n_momentum, prism_vfield* U,
                      prism_vfield* mdot, prism_field* p, const prism_simple_controls* c,
                      prism_report* reports, int32_t max_reports, int32_t* n_reports);
/* The first half of prism_simple_step, up to the pressure solve: the momentum predictor
 * (SIMPLE.cpp:19-41, its reports in `reports`), then the pressure-correction equation as
 * solvePressureEquation hands it to solver.solve() (SIMPLE.cpp:124-178): D = V / a_P from
 * the relaxed momentum diagonals (SIMPLE.cpp:77-108; the predictors' coefficients are
 * zeroed), Rhie-Chow face fluxes written to mdot (rhie_chow.cpp:5-25), p' with constrained
 * face values (SIMPLE.cpp:43-75) and pEqn = Corrected(D, p') + Divergence<Negative>(mdot).
 * *out is a new equation that owns p' (prism_eqn_field) and D and reads mdot, which must
 * outlive it; *ref_cell receives the reference cell the step's solve would use (the
 * automatic choice of SIMPLE.cpp:110-122 resolved; -1: none).  Nothing of the step after
 * that point runs (no p' solve, no field correction). */
int prism_simple_pressure_system(prism_eqn* const* momentum, int32_t n_momentum, prism_vfield* U,
                                 prism_vfield* mdot, prism_field* p, const prism_simple_controls* c,
                                 prism_report* reports, int32_t max_reports, int32_t* n_reports,
                                 prism_eqn** out, int64_t* ref_cell);
int prism_prime_step(prism_eqn* const* momentum, int32_t n_momentum, prism_vfield* U,
                     prism_vfield* mdot, prism_field* p, const prism_simple_controls* c,
                     prism_report* reports, int32_t max_reports, int32_t* n_reports);
int prism_piso_step(prism_eqn* const* momentum, int32_t n_momentum, prism_vfield* U,
                    prism_vfield* mdot, prism_field* p, const prism_piso_controls* c,
                    prism_report* reports, int32_t max_reports, int32_t* n_reports);
/* monitor::courantNumber (courant.cpp:7-24): per-cell Courant number [n]. */
int prism_courant(prism_vfield* mdot, double dt, double* co);
/* algo::needsPressureReference (SIMPLE.cpp:110-122): *out = 1 when no non-empty patch gives
   the field a Dirichlet condition (fixed, no-slip), else 0. */
int prism_needs_pressure_reference(prism_field* p, int32_t* out);

/* ---------------------------------------------------------------- turbulence ---- */
/* The per-cell fields turbulence::IncompressibleKEpsilon::equations() builds before its two
 * transport equations (KEpsilon.cpp:27-77, kProductionIncompressible k_production.cpp:8-25),
 * from k, epsilon, the viscosity nu and grad U (each component's gradient scheme, as
 * U->gradAtCell), in one device pass:
 *   nu_t = c_mu k^2 / eps;   Pk = (nu_t (G + G^T) - 2/3 k I) : G,  G_ij = dU_j / dx_i;
 *   nu_eff_k = nu + nu_t / sigma_k;   nu_eff_eps = nu + nu_t / sigma_eps;
 *   eps_source = c_eps1 (eps / k) Pk;   eps_sink = c_eps2 eps (eps / k).
 * out = {nu_t, Pk, nu_eff_k, nu_eff_eps, eps_source, eps_sink} (fields of the same mesh,
 * overwritten); consts = {c_mu, c_eps1, c_eps2, sigma_k, sigma_eps}, or NULL for the
 * reference's 0.09, 1.44, 1.92, 1.0, 1.3 (KEpsilon.h:29-33).  A maintainer's KEpsilon.cpp
 * calls this in place of setTurbulentViscosity, kProductionIncompressible and the field
 * arithmetic of buildKEqn / buildEpsilonEqn, then builds the same two Transports. */
int prism_kepsilon_update(prism_field* k, prism_field* eps, prism_field* nu, prism_vfield* U, const double* consts,
                          prism_field* const* out);

/* -------------------------------------------------------------------- output ---- */
/* output::toVtu and its helpers (export.cpp).  Point data is the inverse-distance-weighted
 * interpolation of cellToPoint (export.cpp:79-112): for vertex v with adjacent cells c
 * (ascending), w_c = 1 / max(|x_v - x_c|, EPSILON) and p_v = sum_c (w_c / sum w) f_c,
 * computed on the device; a vertex with no adjacent cell is error::InvalidMesh.  The
 * adjacency comes from the cells' vertex lists (prism_mesh_desc.cell_vtx), else from
 * their faces. */
int prism_mesh_n_vertices(prism_mesh* mesh, int64_t* n_vertices);
int prism_field_to_points(prism_field* f, double* out);    /* [n_vertices]    */
int prism_vfield_to_points(prism_vfield* v, double* out);  /* [n_vertices][3] */
int prism_tfield_to_points(prism_tfield* t, double* out);  /* [n_vertices][9] row-major */
#define PRISM_EXPORT_SCALAR 1
#define PRISM_EXPORT_VECTOR 3
#define PRISM_EXPORT_TENSOR 9
#define PRISM_WRITE_ASCII 0
#define PRISM_WRITE_RAW_BINARY_COMPRESSED 1
typedef struct prism_export_item {
    int32_t kind;       /* PRISM_EXPORT_SCALAR / VECTOR / TENSOR */
    const void* field;  /* prism_field* / prism_vfield* / prism_tfield* */
} prism_export_item;
/* output::toVtu(scalars, vectors, tensors, file_name, mode) (export.cpp:262-300): every item's
 * CellData then PointData, in item order, on one unstructured grid of the mesh's vertices
 * and cells (hexahedron 12, tetrahedron 10, prism 13; any other vertex count is
 * PRISM_ERR_RUNTIME).  Items must share one mesh (else PRISM_ERR_MESH); n_items == 0
 * writes nothing.  Needs the cells' vertex lists.  mode: ASCII or raw binary, zlib-compressed
 * (VTK XML appended data). */
int prism_export_vtu(const prism_export_item* items, int32_t n_items, const char* file_name, int32_t mode);
/* TimeWriter::flushPvd (export.cpp:393-409): the ParaView collection of (time, file) pairs. */
int prism_write_pvd(const char* file_name, int32_t n, const double* times, const char* const* files);

/* ------------------------------------------------------------------- adjoint ---- */
/* No reference counterpart.  Discrete adjoint of a converged steady scalar transport
 * equation R(T; s, kappa) = A(kappa) T - b(T; s, kappa) = 0 built from convection::Upwind,
 * one diffusion term with a per-cell scalar kappa (no face values) -- NonCorrected, or
 * Corrected with any corrector when the field has least-squares gradients, no face values
 * and fixed / no-slip / zero-gradient / outlet / symmetry boundaries -- and
 * source::ConstantScalar(s); any other term is PRISM_ERR_INVALID.  prism_adjoint_gradient
 * re-assembles A at the current field and solves (dR/dT)^T lambda = (dJ/dT)^T with BiCGSTAB
 * on the transposed matrix (same CSR pattern, owner/neighbour entries swapped) under the
 * Krylov settings of p; with a Corrected term dR/dT = A - M (M: the explicit correction's
 * derivative through the gradient) and the solve is a defect correction A^T lambda' = g +
 * M^T lambda (result->iteration: A^T solves; converged = 0 when 200 did not settle lambda to
 * 10 tol).  Returns J, dJ/ds and dJ/dkappa per cell (either output may be NULL,
 * as may the matching field); it leaves the equation's coefficients zeroed, as a
 * clearing solve() does.  Objective: J = sum_{c: x_c[axis] > threshold} V_c T_c^2 /
 * sum_{same} V_c.  Only the converged state is used: no time history is stored. */
typedef struct prism_objective {
    int32_t axis;
    double threshold;
} prism_objective;
int prism_adjoint_objective(prism_eqn* e, const prism_objective* obj, double* J);
int prism_adjoint_gradient(prism_eqn* e, prism_field* source, prism_field* kappa,
                           const prism_objective* obj, const prism_solve_params* p,
                           double* J, double* dJds, double* dJdkappa, prism_solve_result* r);

#ifdef __cplusplus
}
#endif
#endif /* PRISM_HIP_H */
