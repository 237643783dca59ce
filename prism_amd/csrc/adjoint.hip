// adjoint.hip -- discrete adjoint of a steady scalar transport equation
//   R(T; s, kappa) = A(kappa) T - b(T; s, kappa) = 0
// assembled from convection::Upwind, one diffusion term with a scalar per-cell kappa and
// source::ConstantScalar(s).  R is linear in T (upwind weights carry no limiter); with
// diffusion::NonCorrected dR/dT = A, with a Corrected form dR/dT = A - M, M the derivative of
// the explicit non-orthogonal correction through the least-squares gradient (kernels.hip,
// "adjoint of the non-orthogonal correction", with dJ/dkappa).  The kernels here: the objective
// J and dJ/dT, the transposed matrix A^T on the same CSR pattern (the face -> slot maps swap
// the owner/neighbour entries), and dJ/ds = -lambda^T dR/ds per cell.
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

}  // namespace prism_hip
