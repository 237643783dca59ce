// internal.h -- device-side views and kernel launchers shared by the HIP kernels and the
// host runtime of libprism_hip.  Plain structs of device pointers, passed by value.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace prism_hip {

constexpr double EPS = 2.220446049250313080847e-16;  // DBL_EPSILON (constants.h:12)
constexpr int MAXP = 64;                              // boundary patches per mesh

#if defined(__HIPCC__)
#define PRISM_HD __host__ __device__
#else
#define PRISM_HD
#endif

// Sliced-ELL storage position of slot k (the k-th entry, padding included) of lane `lane` in
// a slice of 64 rows at offset o and width w.  Slots are stored in pairs (2m, 2m+1)
// interleaved per lane, slot 2m + e of lane l at o + 128 m + 2 l + e, so one 8-byte (FP32
// values, 16-bit column pairs: 4 bytes) or 16-byte (FP64) load per lane fetches two slots and
// a wavefront's load of a pair is contiguous; an odd width's last slot follows the pairs,
// at o + 64 (w - 1) + l.
inline PRISM_HD int64_t sellPos(int64_t o, int w, int k, int lane) {
    return k < (w & ~1) ? o + 128 * (int64_t)(k >> 1) + 2 * lane + (k & 1) : o + 64 * (int64_t)k + lane;
}

// Boundary-condition kinds (the strings of mesh/boundary.cpp:154-202)
enum BCKind : int8_t {
    BC_MISSING = 0,
    BC_FIXED = 1,
    BC_NOSLIP = 2,
    BC_SYMMETRY = 3,
    BC_OUTLET = 4,
    BC_ZEROGRAD = 5,
    BC_FIXEDGRAD = 6,
    BC_EMPTY = 7,
    BC_OTHER = 8,
};

// ------------------------------------------------------------------------- errors ----
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void hipCheck(hipError_t e, const char* what);
// Raises PRISM_ERR_RUNTIME when a per-block partial writer's grid needs more than `cap` doubles.
void require_partials(const char* who, size_t need, size_t cap);
#define HIPCHECK(x) ::prism_hip::hipCheck((x), #x)

// ------------------------------------------------------------------- device views ----
// Mesh in the permuted face order: [0,nne) non-empty boundary, [nne,nb) empty boundary,
// [nb,F) interior.  All arrays are SoA.
struct DMesh {
    int32_t n, F, nb, nne;
    const double *cx, *cy, *cz, *vol;          // cell centers, volumes
    const double *sx, *sy, *sz;                // face area vectors (area * unit normal)
    const double *fx, *fy, *fz;                // face centers
    const double *nx, *ny, *nz;                // unit normals
    const double *dx, *dy, *dz;                // interior: x_N - x_C;  boundary: x_f - x_C
    const double* gc;                          // geometric weight, interior faces (index f)
    const double *kx, *ky, *kz;                // x_f - (x_C + x_N)/2, interior (Green-Gauss skew)
    const int32_t *owner, *nbr, *fpatch;
    const int32_t *cf_off, *cf;                // cell -> faces; entry = face << 1 | is_neighbour
    const int32_t *rp, *col, *diag;            // CSR pattern (sorted columns)
    const int32_t *slot_on, *slot_no;          // interior face -> CSR slot of (o,n) and (n,o), or an
                                               // extra slot >= nnz (a pair's further faces)
    const int32_t *dup_slot, *dup_ptr;         // extra slots, grouped by the CSR slot they add into
    int32_t nnz, n_dup_groups;
    const double* pinv;                        // least-squares pseudo-inverse, 9 x n SoA
};

// Scalar field: cell values, optional face values, cached gradient, BC table per patch.
struct DScalar {
    const double* v;
    const double* fv;               // nullptr when the field has no face values
    const double *gx, *gy, *gz;     // cell gradient cache
    const int8_t* bck;              // kind per patch
    const double* bcs;              // scalar value per patch (component for vector parts)
    const double* bcv;              // vector value per patch, [MAXP][3]
    int32_t pressure;               // 1: "symmetry" extrapolates (pressure_boundary.cpp)
};

// Vector field: three components + optional face data / face fluxes.
struct DVector {
    DScalar c[3];
    const double* fd;   // face values, 3 x F SoA, or nullptr
    const double* ff;   // face fluxes, F, or nullptr
    const int8_t* bck;  // vector BC kind per patch
    const double* bcv;  // [MAXP][3]
};

// Tensor field: 9 x n SoA, row-major components (k = 3 r + c).
struct DTensor {
    const double* t;
    int32_t diag_only;  // off-diagonal components are all zero
};

// Face-based schemes of an equation, in constructor order.
constexpr int MAX_FACE_SCHEMES = 4;
struct DFaceScheme {
    int32_t kind;        // 1 convection, 2 diffusion
    int32_t variant;     // NVF id / corrector (0 non-corrected, 1 over-relaxed, 2 min, 3 orth)
    const double* flux;  // convection: materialised U . S per face
    DVector U;           // convection: velocity for boundary values
    int32_t kappa_kind;  // diffusion: 0 scalar, 1 tensor
    DScalar kappa;
    DTensor kappa_t;
};

// Cell-based contributions (sources, implicit fields, temporal), in constructor order.
constexpr int MAX_CELL_TERMS = 8;
struct DCellTerm {
    int32_t kind;       // 3 constant, 4 gradient, 5 divergence, 6 implicit, 7 backward Euler,
                        // 8 Adam-Moulton (BDF2)
    int32_t sign;       // 0 positive, 1 negative
    int32_t coord;
    double coeff;       // implicit coefficient / dt
    const double* phi;  // source values / previous time level
    const double* phi2; // Adam-Moulton: the level before the previous one (NULL: first step)
    const double* g;    // gradient component array (source::Gradient)
    const double* flux; // divergence: materialised face fluxes
};

struct DEquation {
    DScalar phi;
    int32_t n_face, n_cell;
    DFaceScheme fs[MAX_FACE_SCHEMES];
    DCellTerm ct[MAX_CELL_TERMS];
    double* A;
    double* b;
    double* scr;          // 4 x n_interior: diag owner, diag neighbour, rhs owner, rhs neighbour
    int32_t accumulate;   // add to the existing A, b instead of overwriting
};

// ------------------------------------------------------------------------ launches ----
struct Launch {
    hipStream_t stream;
    // timing hook (null when timing is off): hands out the start / stop event pair of one
    // launch of the kernel `name`, recorded by the launch itself (see PRISM_LAUNCH)
    void (*events)(void* user, const char* name, hipEvent_t* start, hipEvent_t* stop);
    void* user;
};

// Launches `kern` on L.stream.  With timing on, hipExtLaunchKernelGGL stamps the kernel's own
// start and end into the event pair (the dispatch's begin and end timestamps, which a kernel
// trace reads too), so the measured duration holds no queue or dispatch gap.  Expanded only in
// the .hip sources (which include hip/hip_ext.h).
#define PRISM_LAUNCH(label, kern, grid, threads, ...)                                               \
    do {                                                                                            \
        hipEvent_t ev_start_ = nullptr, ev_stop_ = nullptr;                                         \
        if (L.events) L.events(L.user, label, &ev_start_, &ev_stop_);                               \
        hipExtLaunchKernelGGL(kern, dim3(grid), dim3(threads), 0, L.stream, ev_start_, ev_stop_, 0, \
                              __VA_ARGS__);                                                         \
        HIPCHECK(hipGetLastError());                                                                \
    } while (0)

// assembly / fields
void launch_fill(const Launch& L, double* p, int64_t n, double v);
// Cells sharing several faces: adds each group's extra slots into its CSR slot (face order) and
// clears them (no launch when the mesh has none).
void launch_dup_fixup(const Launch& L, const DMesh& m, double* A);
void launch_fill(const Launch& L, float* p, int64_t n, float v);
void launch_copy(const Launch& L, double* dst, const double* src, int64_t n);
void launch_axpby(const Launch& L, double* y, const double* x, double a, double b, int64_t n);  // y = a*x + b*y
void launch_vector_flux(const Launch& L, const DMesh& m, const DVector& U, double* flux);
void launch_vector_face_values(const Launch& L, const DMesh& m, const DVector& U, double* out3);
void launch_scalar_face_values(const Launch& L, const DMesh& m, const DScalar& s, double* out);
// Least-squares gradient into g = [gx | gy | gz] (3n doubles); the three-field form fills
// the gradients of a vector's components in one pass, each bitwise as the one-field form.
void launch_grad_ls(const Launch& L, const DMesh& m, const DScalar& s, double* g);
void launch_grad_ls3(const Launch& L, const DMesh& m, const DScalar (&s)[3], double* const (&g)[3]);
void launch_grad_gg_level(const Launch& L, const DMesh& m, const DScalar& s, const int32_t* cells,
                          int32_t count, double* gx, double* gy, double* gz);
void launch_assemble(const Launch& L, const DMesh& m, const DEquation& e, int32_t relax, double urf);
void launch_relax(const Launch& L, const DMesh& m, double* A, double* b, const double* phi, double urf);
void launch_ref_cell(const Launch& L, const DMesh& m, double* A, double* b, int32_t cell, double ref);
// SIMPLE helpers
void launch_pressure_D(const Launch& L, const DMesh& m, const double* const* Adiag_src, int32_t nmom,
                       double* D9);
void launch_rhie_chow(const Launch& L, const DMesh& m, const DVector& U, const DTensor& D,
                      const DScalar& p, double* fd, double* ff);
void launch_map_face_values(const Launch& L, const DMesh& m, const DVector& U, double* fd, double* ff);
void launch_correct_U(const Launch& L, const DMesh& m, double* ux, double* uy, double* uz,
                      const DTensor& D, const DScalar& pp);
void launch_correct_mdot(const Launch& L, const DMesh& m, double* fd, double* ff, const DTensor& D,
                         const DScalar& pp);
void launch_update_p(const Launch& L, double* p, const double* pp, double urf, int64_t n);
void launch_explicit_momentum(const Launch& L, const DMesh& m, const double* A, const double* b,
                              double* x, double* tmp);
void launch_courant(const Launch& L, const DMesh& m, const double* ff, double dt, double* co);
// Per-cell fields of the k-epsilon equations (KEpsilon.cpp:27-77, k_production.cpp:8-25).
// g[j]: gradient of U_j, 3 x n SoA (d/dx, d/dy, d/dz).
struct DKEps {
    int n;
    const double *k, *eps, *nu;
    const double* g[3];
    double c_mu, c_eps1, c_eps2, sigma_k, sigma_eps;
    double *nu_t, *pk, *nu_eff_k, *nu_eff_eps, *eps_source, *eps_sink;
};
void launch_kepsilon(const Launch& L, const DKEps& a);
// Cell -> point interpolation of 1, 3 or 9 component arrays (cell-indexed) with normalised
// weights over the vertex -> cell adjacency; out interleaved [nv][ncomp].  off: the CSR
// offsets of the adjacency (degree of v = off[v+1] - off[v]); sliced copy of it: entry j of
// vertex v at sbase[v / 64] + 64 j + v % 64 in cid (cell) and w (weight).
struct DPointSrc { const double* c[9]; };
void launch_cell_to_point(const Launch& L, int nv, int ncomp, const int32_t* off, const int32_t* sbase,
                          const int32_t* cid, const double* w, const DPointSrc& src, double* out);
// Symmetry-plane momentum constraint (symmetry.cpp:54-82) for one patch: the owners of the
// patch's faces (first-appearance order), each with its faces in patch order.
struct DSymGroup {
    const int32_t* cells;
    const int32_t* off;    // [n + 1] into faces
    const int32_t* faces;
    int32_t n;
};
void launch_symmetry(const Launch& L, const DMesh& m, const DSymGroup& g, int32_t coord, const DVector& U,
                     const DFaceScheme& diff, double* A, double* b);
// Discrete adjoint (adjoint.hip)
int adjoint_objective_blocks(const DMesh& m);
void launch_objective(const Launch& L, const DMesh& m, const double* T, int axis, double thr, double* part,
                      size_t part_cap);  // part: 2 doubles per block of adjoint_objective_blocks, checked
void launch_objective_grad(const Launch& L, const DMesh& m, const double* T, int axis, double thr, double VR,
                           double* g);
void launch_transpose(const Launch& L, const DMesh& m, const double* A, double* AT);
void launch_dJds(const Launch& L, const DMesh& m, const double* lam, int negative, double* out);
// Adjoint of Corrected diffusion with least-squares gradients: rhs = g + M^T lambda (M = db/dphi
// of the explicit non-orthogonal correction; q: 3 n scratch), and dJ/dkappa for any corrector
// (scr: 2 per interior face).  The field's gradient cache must hold the converged state's.
void launch_adj_corr_rhs(const Launch& L, const DMesh& m, const DScalar& phi, const DFaceScheme& s, const double* lam,
                         double* q, const double* g, double* rhs);
void launch_adj_dkappa(const Launch& L, const DMesh& m, const DScalar& phi, int variant, const double* lam,
                       double* scr, double* out);
// per-block partials of sum (a - b)^2 and sum b^2 (2 doubles per block of diff_norm_blocks)
int diff_norm_blocks(int64_t n);
void launch_diff_norms(const Launch& L, int64_t n, const double* a, const double* b, double* part, size_t part_cap);
void launch_constrain_pprime(const Launch& L, const DMesh& m, const DScalar& pp, double* fv);

// Krylov (krylov.hip)
struct KrylovScalars;  // device-resident BiCGSTAB scalars
template <typename T>
struct KrylovWork {
    int64_t n;
    int32_t nblk;
    T *r, *r0, *p, *v, *s, *t, *y, *z, *x, *dinv;
    double* partials;   // 3 * nblk
    unsigned int* counter;
    KrylovScalars* ks;  // device
};

}  // namespace prism_hip
