// adjoint.hip -- discrete adjoint of a steady scalar transport equation
//   R(T; s, kappa) = A(kappa) T - b(s, kappa) = 0
// assembled from convection::Upwind, diffusion::NonCorrected with a scalar per-cell kappa and
// source::ConstantScalar(s).  With these schemes R is linear in T (upwind weights carry no
// limiter, the non-corrected flux no gradient term), so dR/dT = A, and A and b are linear in
// kappa and s.  The kernels here: the objective J and dJ/dT, the transposed matrix A^T on the
// same CSR pattern (the face -> slot maps swap the owner/neighbour entries), and the parameter
// sensitivities dJ/dtheta = -lambda^T dR/dtheta gathered per cell in face order (no atomics).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include "internal.h"

namespace prism_hip {

namespace {

constexpr int BS = 256;

__device__ inline double cellAxis(const DMesh& m, int c, int axis) {
    return axis == 0 ? m.cx[c] : (axis == 1 ? m.cy[c] : m.cz[c]);
}

// J = sum_R V T^2 / sum_R V over R = {c : x_c[axis] > threshold}: per-block partials.
__global__ void __launch_bounds__(BS) k_objective(DMesh m, const double* T, int axis, double thr, double* part) {
    __shared__ double s0[BS], s1[BS];
    int c = blockIdx.x * BS + threadIdx.x;
    double a = 0.0, v = 0.0;
    if (c < m.n && cellAxis(m, c, axis) > thr) {
        a = m.vol[c] * (T[c] * T[c]);
        v = m.vol[c];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = v;
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}

// dJ/dT_c = 2 V_c T_c / V_R on R, 0 elsewhere.
__global__ void k_objective_grad(DMesh m, const double* T, int axis, double thr, double VR, double* g) {
    int c = blockIdx.x * BS + threadIdx.x;
    if (c >= m.n) return;
    g[c] = cellAxis(m, c, axis) > thr ? 2.0 * m.vol[c] * T[c] / VR : 0.0;
}

// A^T on A's pattern: the diagonal stays, each interior face swaps its two off-diagonals (the
// entries of a pair of cells that share several faces are swapped once, by its first face).
__global__ void k_transpose_faces(DMesh m, const double* A, double* AT) {
    int fi = blockIdx.x * BS + threadIdx.x;
    int nint = m.F - m.nb;
    if (fi >= nint) return;
    int f = m.nb + fi;
    int so = m.slot_on[f], sn = m.slot_no[f];
    if (so >= m.nnz || sn >= m.nnz) return;  // a pair's further face: its entries are the first face's
    AT[so] = A[sn];
    AT[sn] = A[so];
}
__global__ void k_transpose_diag(DMesh m, const double* A, double* AT) {
    int c = blockIdx.x * BS + threadIdx.x;
    if (c < m.n) AT[m.diag[c]] = A[m.diag[c]];
}

// dJ/ds_c = -lambda_c dR_c/ds_c; ConstantScalar adds +-s_c V_c to b_c.
__global__ void k_dJds(DMesh m, const double* lam, int negative, double* out) {
    int c = blockIdx.x * BS + threadIdx.x;
    if (c >= m.n) return;
    double v = lam[c] * m.vol[c];
    out[c] = negative ? -v : v;
}

// Interior face f (owner O, neighbour N): the non-corrected flux D_f (T_O - T_N) enters R_O
// and -D_f (T_O - T_N) enters R_N, with D_f = kappa_f (S . d) / (|d|^2 + eps) and
// kappa_f = w kappa_O + (1 - w) kappa_N (diffusion.h:116-149, scalar.cpp:209-219).  Hence
// lambda^T dR/dkappa_O gets w q_f and lambda^T dR/dkappa_N gets (1 - w) q_f, with
// q_f = (lambda_O - lambda_N) (T_O - T_N) (S . d) / (|d|^2 + eps).
__global__ void k_dkappa_faces(DMesh m, const double* T, const double* lam, double* scr) {
    int fi = blockIdx.x * BS + threadIdx.x;
    int nint = m.F - m.nb;
    if (fi >= nint) return;
    int f = m.nb + fi;
    int o = m.owner[f], n = m.nbr[f];
    double dx = m.dx[f], dy = m.dy[f], dz = m.dz[f];
    double dd = (dx * dx + dy * dy) + dz * dz;
    double G = ((m.sx[f] * dx + m.sy[f] * dy) + m.sz[f] * dz) / (dd + EPS);
    double q = (lam[o] - lam[n]) * (T[o] - T[n]) * G;
    double w = m.gc[f];
    scr[fi] = w * q;
    scr[nint + fi] = (1.0 - w) * q;
}

// Per cell, in face order: interior faces from the scratch; a fixed (Dirichlet) boundary
// face adds lambda_C (T_C - T_b) (S . d_Cf) / (|d_Cf|^2 + eps) (diffusion_boundary.h:113-160,
// owner kappa); a fixed-gradient face adds lambda_C (g_w . S).  dJ/dkappa_C = -sum.
__global__ void k_dkappa_cells(DMesh m, DScalar T, const double* lam, const double* scr, double* out) {
    int c = blockIdx.x * BS + threadIdx.x;
    if (c >= m.n) return;
    int nint = m.F - m.nb;
    double acc = 0.0;
    for (int k = m.cf_off[c]; k < m.cf_off[c + 1]; ++k) {
        int e = m.cf[k];
        int f = e >> 1;
        if (f >= m.nb) {
            acc += (e & 1) ? scr[nint + (f - m.nb)] : scr[f - m.nb];
            continue;
        }
        if (f >= m.nne) continue;  // empty patch
        int p = m.fpatch[f];
        int kind = T.bck[p];
        if (kind == BC_FIXED || kind == BC_NOSLIP) {
            double tb = T.fv ? T.fv[f] : T.bcs[p];
            double dx = m.dx[f], dy = m.dy[f], dz = m.dz[f];
            double dd = (dx * dx + dy * dy) + dz * dz;
            double G = ((m.sx[f] * dx + m.sy[f] * dy) + m.sz[f] * dz) / (dd + EPS);
            acc += lam[c] * (T.v[c] - tb) * G;
        } else if (kind == BC_FIXEDGRAD) {
            double gS = (T.bcv[3 * p] * m.sx[f] + T.bcv[3 * p + 1] * m.sy[f]) + T.bcv[3 * p + 2] * m.sz[f];
            acc += lam[c] * gS;
        }
    }
    out[c] = -acc;
}

inline int nblk(int64_t n) { return (int)((n + BS - 1) / BS); }

}  // namespace

#define ALAUNCH(name, grid, ...)                          \
    do {                                                  \
        if ((grid) > 0) PRISM_LAUNCH(#name, name, grid, BS, __VA_ARGS__); \
    } while (0)

int adjoint_objective_blocks(const DMesh& m) { return nblk(m.n); }

void launch_objective(const Launch& L, const DMesh& m, const double* T, int axis, double thr, double* part,
                      size_t part_cap) {
    require_partials("adjoint objective", 2 * (size_t)nblk(m.n), part_cap);
    ALAUNCH(k_objective, nblk(m.n), m, T, axis, thr, part);
}
void launch_objective_grad(const Launch& L, const DMesh& m, const double* T, int axis, double thr, double VR,
                           double* g) {
    ALAUNCH(k_objective_grad, nblk(m.n), m, T, axis, thr, VR, g);
}
void launch_transpose(const Launch& L, const DMesh& m, const double* A, double* AT) {
    ALAUNCH(k_transpose_faces, nblk(m.F - m.nb), m, A, AT);
    ALAUNCH(k_transpose_diag, nblk(m.n), m, A, AT);
}
void launch_dJds(const Launch& L, const DMesh& m, const double* lam, int negative, double* out) {
    ALAUNCH(k_dJds, nblk(m.n), m, lam, negative, out);
}
void launch_dJdkappa(const Launch& L, const DMesh& m, const DScalar& T, const double* lam, double* scr, double* out) {
    ALAUNCH(k_dkappa_faces, nblk(m.F - m.nb), m, T.v, lam, scr);
    ALAUNCH(k_dkappa_cells, nblk(m.n), m, T, lam, scr, out);
}

}  // namespace prism_hip
