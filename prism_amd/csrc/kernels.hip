// kernels.hip -- finite-volume assembly, gradient and SIMPLE kernels for gfx950.
//
// Assembly is face-parallel with a deterministic owner/neighbour gather: one thread per
// interior face evaluates every face-based scheme of the equation, writes the two
// off-diagonal coefficients straight into their (unique) CSR slots and parks its four
// diagonal/rhs contributions in a face scratch array; one thread per cell then sums those
// in the cell's face order, adds boundary-face and cell-based terms, and writes the
// diagonal slot and rhs.  No floating-point atomics anywhere.
//
// Floating-point expressions keep the operation order of the reference (compiled with
// -ffp-contract=off) so branch decisions such as the NVF limiter intervals match it.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include "../../include/prism_hip.h"
#include "internal.h"

namespace prism_hip {

namespace {

constexpr int BS = 256;

struct V3d {
    double x, y, z;
};
__device__ inline V3d mk(double x, double y, double z) { return {x, y, z}; }
__device__ inline V3d add(V3d a, V3d b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3d sub(V3d a, V3d b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3d neg(V3d a) { return {-a.x, -a.y, -a.z}; }
__device__ inline V3d scl(double s, V3d a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ inline V3d divs(V3d a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ inline double dot(V3d a, V3d b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ inline double nrm(V3d a) { return sqrt(dot(a, a)); }
__device__ inline double smax(double a, double b) { return (a < b) ? b : a; }  // std::max

__device__ inline V3d faceS(const DMesh& m, int f) { return mk(m.sx[f], m.sy[f], m.sz[f]); }
__device__ inline V3d faceD(const DMesh& m, int f) { return mk(m.dx[f], m.dy[f], m.dz[f]); }
__device__ inline V3d faceN(const DMesh& m, int f) { return mk(m.nx[f], m.ny[f], m.nz[f]); }
__device__ inline V3d cellG(const DScalar& s, int c) { return mk(s.gx[c], s.gy[c], s.gz[c]); }

// Scalar::valueAtFace (scalar.cpp:189-237, scalar_boundary.h).  field::Pressure inherits
// Scalar's default handlers (pressure.h:6-9, scalar.cpp:239-248), so its symmetry faces take
// the owner value too: the extrapolating Symmetry<Pressure> of pressure_boundary.cpp is never
// registered by the reference.
__device__ double sFace(const DMesh& m, const DScalar& s, int f) {
    if (s.fv) return s.fv[f];
    int o = m.owner[f];
    if (f >= m.nb) {
        double g = m.gc[f];
        double val = g * s.v[o];
        val += (1 - g) * s.v[m.nbr[f]];
        return val;
    }
    int p = m.fpatch[f];
    switch (s.bck[p]) {
        case BC_FIXED:
        case BC_NOSLIP: return s.bcs[p];
        case BC_OUTLET:
        case BC_ZEROGRAD:
        case BC_SYMMETRY: return s.v[o];
        default: return __builtin_nan("");
    }
}

// gradient/boundary.cpp:23-34 (fixed / no-slip boundary face gradient)
__device__ V3d gradFixedFace(const DMesh& m, const DScalar& s, int f) {
    int o = m.owner[f];
    V3d gcl = cellG(s, o);
    V3d dCf = faceD(m, f);
    double dn = nrm(dCf);
    V3d e = divs(dCf, dn);
    double pb = sFace(m, s, f);
    double pc = s.v[o];
    V3d a = sub(gcl, scl(dot(gcl, e), e));
    return add(a, scl((pb - pc) / dn, e));
}

// GeneralVector::valueAtFace (vector.h:187-202, vector_boundary.h:55-115)
__device__ V3d vFace(const DMesh& m, const DVector& U, int f) {
    if (U.fd) return mk(U.fd[f], U.fd[m.F + f], U.fd[2 * m.F + f]);
    if (f >= m.nb) return mk(sFace(m, U.c[0], f), sFace(m, U.c[1], f), sFace(m, U.c[2], f));
    int p = m.fpatch[f];
    int o = m.owner[f];
    V3d uc = mk(U.c[0].v[o], U.c[1].v[o], U.c[2].v[o]);
    switch (U.bck[p]) {
        case BC_FIXED:
        case BC_NOSLIP: return mk(U.bcv[3 * p], U.bcv[3 * p + 1], U.bcv[3 * p + 2]);
        case BC_SYMMETRY: {
            V3d n = faceN(m, f);
            return sub(uc, scl(dot(uc, n), n));
        }
        case BC_OUTLET:
        case BC_ZEROGRAD: return uc;
        default: return mk(__builtin_nan(""), __builtin_nan(""), __builtin_nan(""));
    }
}

// GeneralVector::fluxAtFace (vector.h:216-257)
__device__ double vFlux(const DMesh& m, const DVector& U, int f) {
    if (U.ff) return U.ff[f];
    if (f < m.nb) {
        int k = U.bck[m.fpatch[f]];
        if (k == BC_NOSLIP || k == BC_SYMMETRY) return 0.0;
    }
    return dot(vFace(m, U, f), faceS(m, f));
}

struct NVF {
    double l, k;
};
// convection.cpp:69-128.  SMART's thresholds are C++ integer divisions (== 0).
__device__ inline NVF nvf(int scheme, double t) {
    switch (scheme) {
        case 1: return {0.5, 0.5};
        case 2: return {1.5, 0.0};
        case 3: return {0.75, 3.0 / 8.0};
        case 4: return {1.0, 0.25};
        case 5:
            if (t > 0.0 && t < 0.5) return {1.5, 0.0};
            if (t >= 0.5 && t < 1.0) return {0.5, 0.5};
            return {1.0, 0.0};
        case 6:
            if (t > 0.0 && t < 0.25) return {2.0, 0.0};
            if (t >= 0.25 && t < 0.75) return {1.0, 0.25};
            if (t >= 0.75 && t < 1.0) return {0.0, 1.0};
            return {1.0, 0.0};
        case 7:
            if (t >= 0.0 && t < 1.0) return {0.0, 1.0};
            return {1.0, 0.0};
        default: return {1.0, 0.0};
    }
}

struct Coef {
    double aC, aN, b;
};
// Upwind::interpolate (convection.cpp:46-67) with phiTilde / phiAtDummyUpwind
// (convection_hr.cpp:6-23).  dCN = x_N - x_C.
__device__ Coef interp(int scheme, double m, double pC, double pN, V3d gC, V3d gN, V3d dCN) {
    double pu = pN - (2 * dot(gC, dCN));
    double t = (pC - pu) / (pN - pu);
    NVF a = nvf(scheme, t);
    double aC = smax(m, 0.0) * a.l;
    double aN = smax(m, 0.0) * a.k;
    double b = -smax(m, 0.0) * (1 - a.l - a.k) * pu;
    V3d dNC = neg(dCN);
    double pu2 = pC - (2 * dot(gN, dNC));
    double t2 = (pN - pu2) / (pC - pu2);
    NVF q = nvf(scheme, t2);
    aC += -smax(-m, 0.0) * q.k;
    aN += -smax(-m, 0.0) * q.l;
    b += smax(-m, 0.0) * (1 - q.l - q.k) * pu2;
    return {aC, aN, b};
}

struct Dec {
    V3d E, T;
};
// nonortho.cpp:7-21
__device__ inline Dec decompose(int corr, V3d S, V3d e) {
    V3d E;
    if (corr == 1) E = scl((dot(S, S) / (dot(e, S) + EPS)), e);
    else if (corr == 2) E = scl(dot(e, S), e);
    else E = scl(nrm(S), e);
    return {E, sub(S, E)};
}

// diffusion_coeff.cpp:16-28: kappa_f^T S (tensor) or kappa_f S (scalar)
__device__ V3d kappaFace(const DMesh& m, const DFaceScheme& s, V3d S, int f) {
    if (s.kappa_kind == 0) return scl(sFace(m, s.kappa, f), S);
    int o = m.owner[f], n = m.nbr[f];
    double g = m.gc[f];
    const double* t = s.kappa_t.t;
    int N = m.n;
    double T[9];
    for (int k = 0; k < 9; ++k) T[k] = (g * t[k * N + o]) + ((1 - g) * t[k * N + n]);
    return mk((T[0] * S.x + T[3] * S.y) + T[6] * S.z, (T[1] * S.x + T[4] * S.y) + T[7] * S.z,
              (T[2] * S.x + T[5] * S.y) + T[8] * S.z);
}
__device__ V3d kappaCell(const DMesh& m, const DFaceScheme& s, V3d S, int c) {
    if (s.kappa_kind == 0) return scl(s.kappa.v[c], S);
    const double* t = s.kappa_t.t;
    int N = m.n;
    double T[9];
    for (int k = 0; k < 9; ++k) T[k] = t[k * N + c];
    return mk((T[0] * S.x + T[3] * S.y) + T[6] * S.z, (T[1] * S.x + T[4] * S.y) + T[7] * S.z,
              (T[2] * S.x + T[5] * S.y) + T[8] * S.z);
}

// Symmetry<Transport> (symmetry.cpp:12-82): per owner cell, the faces of one symmetry patch
// in patch order; g = 2 |kappa_C^T S_f| / ((x_f - x_C) . n).
__global__ void k_symmetry(DMesh m, DSymGroup g, int coord, DVector U, DFaceScheme diff, double* A, double* b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    const int c = g.cells[i];
    const V3d Uc = mk(U.c[0].v[c], U.c[1].v[c], U.c[2].v[c]);
    const V3d xc = mk(m.cx[c], m.cy[c], m.cz[c]);
    double dA = 0.0, dB = 0.0;
    for (int k = g.off[i]; k < g.off[i + 1]; ++k) {
        const int f = g.faces[k];
        const V3d n = faceN(m, f);
        const V3d dCb = sub(mk(m.fx[f], m.fy[f], m.fz[f]), xc);
        const double dn = dot(dCb, n);
        const V3d kS = kappaCell(m, diff, faceS(m, f), c);
        const double gg = 2 * sqrt(dot(kS, kS)) / dn;
        double ac, bb;
        if (coord == 0) { ac = n.x * n.x; bb = ((Uc.y * n.y) + (Uc.z * n.z)) * n.x; }
        else if (coord == 1) { ac = n.y * n.y; bb = ((Uc.x * n.x) + (Uc.z * n.z)) * n.y; }
        else { ac = n.z * n.z; bb = ((Uc.x * n.x) + (Uc.y * n.y)) * n.z; }
        dA = (k == g.off[i]) ? gg * ac : dA + gg * ac;
        dB += -gg * bb;
    }
    A[m.diag[c]] += dA;
    b[c] += dB;
}

// ------------------------------------------------------------------- utilities ----
__global__ void k_fill(double* p, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f32(float* p, int64_t n, float v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_copy(double* d, const double* s, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = s[i];
}
__global__ void k_axpby(double* y, const double* x, double a, double b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}

// ------------------------------------------------------------------ face values ----
__global__ void k_vector_flux(DMesh m, DVector U, double* flux) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    flux[f] = (f >= m.nne && f < m.nb) ? 0.0 : vFlux(m, U, f);
}
__global__ void k_vector_face_values(DMesh m, DVector U, double* out) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    V3d u = (f >= m.nne && f < m.nb && !U.fd) ? mk(0, 0, 0) : vFace(m, U, f);
    out[f] = u.x;
    out[m.F + f] = u.y;
    out[2 * m.F + f] = u.z;
}
__global__ void k_scalar_face_values(DMesh m, DScalar s, double* out) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    out[f] = (f >= m.nne && f < m.nb && !s.fv) ? 0.0 : sFace(m, s, f);
}

// --------------------------------------------------------------------- gradients ----
// LeastSquares::computeGradAtCell (least_squares.cpp:78-100); the weighted directions of
// least_squares.cpp:34-49 are rebuilt from the stored face vectors (bitwise identical).
// Cells with at most LS_MAXF faces (hexahedra, wedges, tetrahedra) issue the loads of all
// their faces before any arithmetic: face entries, then every face's d vector and neighbour
// id, then the neighbours' values, each stage as straight-line loads at valid indices, so
// the three dependent gathers of different faces overlap instead of running one face at a
// time.  The sum still runs over the faces in the cell's order, as before.
constexpr int LS_MAXF = 6;
// NC fields on one mesh (the components of a vector): the face entries, d vectors,
// neighbour ids and the pseudo-inverse are loaded once for all of them, and each field's
// sums run in the same order as for a single field, so every component's gradient is
// bitwise the one the single-field kernel gives.
template <int NC>
__device__ inline void gradLS(const DMesh& m, const DScalar (&s)[NC], double* const (&g)[NC]) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    double pc[NC];
    V3d b[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        pc[i] = s[i].v[c];
        b[i] = mk(0.0, 0.0, 0.0);
    }
    const int a = m.cf_off[c], nf = m.cf_off[c + 1] - a;
    if (nf <= LS_MAXF) {
        int e[LS_MAXF], o[LS_MAXF];
        V3d d[LS_MAXF];
        double oth[NC][LS_MAXF];
#pragma unroll
        for (int j = 0; j < LS_MAXF; ++j) e[j] = j < nf ? m.cf[a + j] : 0;
#pragma unroll
        for (int j = 0; j < LS_MAXF; ++j) {
            const int f = e[j] >> 1;
            d[j] = faceD(m, f);
            const int q = ((e[j] & 1) ? m.owner : m.nbr)[f];
            o[j] = f >= m.nb ? q : c;  // the other cell across an interior face
        }
#pragma unroll
        for (int i = 0; i < NC; ++i)
#pragma unroll
            for (int j = 0; j < LS_MAXF; ++j) oth[i][j] = s[i].v[o[j]];
#pragma unroll
        for (int j = 0; j < LS_MAXF; ++j) {
            if (j >= nf) break;
            const int f = e[j] >> 1;
            V3d r;
            double other[NC];
            if (f >= m.nb) {
                r = (e[j] & 1) ? neg(d[j]) : d[j];
#pragma unroll
                for (int i = 0; i < NC; ++i) other[i] = oth[i][j];
            } else if (f < m.nne) {
                r = d[j];
#pragma unroll
                for (int i = 0; i < NC; ++i) other[i] = sFace(m, s[i], f);
            } else {
                continue;
            }
            double wk = 1 / (nrm(r) + EPS);
            V3d w = scl(wk, r);
#pragma unroll
            for (int i = 0; i < NC; ++i) b[i] = add(b[i], scl(other[i] - pc[i], w));
        }
    } else for (int k = a; k < a + nf; ++k) {
        int e = m.cf[k];
        int f = e >> 1;
        double other[NC];
        V3d r;
        if (f >= m.nb) {
            int o = (e & 1) ? m.owner[f] : m.nbr[f];
            r = (e & 1) ? neg(faceD(m, f)) : faceD(m, f);
#pragma unroll
            for (int i = 0; i < NC; ++i) other[i] = s[i].v[o];
        } else if (f < m.nne) {
            r = faceD(m, f);
#pragma unroll
            for (int i = 0; i < NC; ++i) other[i] = sFace(m, s[i], f);
        } else {
            continue;
        }
        double wk = 1 / (nrm(r) + EPS);
        V3d w = scl(wk, r);
#pragma unroll
        for (int i = 0; i < NC; ++i) b[i] = add(b[i], scl(other[i] - pc[i], w));
    }
    const double* P = m.pinv;
    int N = m.n;
    double p[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) p[q] = P[q * N + c];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        double g0 = (p[0] * b[i].x + p[1] * b[i].y) + p[2] * b[i].z;
        double g1 = (p[3] * b[i].x + p[4] * b[i].y) + p[5] * b[i].z;
        double g2 = (p[6] * b[i].x + p[7] * b[i].y) + p[8] * b[i].z;
        g[i][c] = g0;
        g[i][N + c] = g1;
        g[i][2 * N + c] = g2;
    }
}

__global__ void k_grad_ls(DMesh m, DScalar s, double* g) {
    const DScalar ss[1] = {s};
    double* const gg[1] = {g};
    gradLS<1>(m, ss, gg);
}

struct DScalar3 {
    DScalar s[3];
    double* g[3];
};
__global__ void k_grad_ls3(DMesh m, DScalar3 a) { gradLS<3>(m, a.s, a.g); }

// GreenGauss::computeGradAtCell (green_gauss.cpp:49-76) for one wavefront level of the
// sequential sweep: cells of a level are never adjacent, lower-numbered neighbours sit on
// earlier levels (already updated this pass), higher-numbered ones on later levels.
__global__ void k_grad_gg_level(DMesh m, DScalar s, const int32_t* cells, int32_t count, double* gx,
                                double* gy, double* gz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    int c = cells[i];
    V3d grad = mk(0.0, 0.0, 0.0);
    for (int k = m.cf_off[c]; k < m.cf_off[c + 1]; ++k) {
        int e = m.cf[k];
        int f = e >> 1;
        if (f < m.nb) {
            if (f >= m.nne) continue;
            grad = add(grad, scl(sFace(m, s, f), faceS(m, f)));
            continue;
        }
        V3d S = (e & 1) ? neg(faceS(m, f)) : faceS(m, f);
        int nb = (e & 1) ? m.owner[f] : m.nbr[f];
        double fphi = 0.5 * (s.v[c] + s.v[nb]);
        V3d gs = mk(gx[c] + gx[nb], gy[c] + gy[nb], gz[c] + gz[nb]);
        V3d vec = mk(m.kx[f], m.ky[f], m.kz[f]);
        fphi += 0.5 * dot(gs, vec);
        grad = add(grad, scl(fphi, S));
    }
    grad = divs(grad, m.vol[c]);
    gx[c] = grad.x;
    gy[c] = grad.y;
    gz[c] = grad.z;
}

// ---------------------------------------------- adjoint of the non-orthogonal correction ----
// Corrected diffusion adds to b the explicit correction T_f . grad(phi)_f (interior faces, with
// the gradient interpolated, grad_f = g grad_O + (1 - g) grad_N, into b_O and out of b_N) and
// T_b . gradFixedFace on fixed faces.  With least-squares gradients, grad_c = P_c sum_k w_k
// r_k (phi_k - phi_c), those terms are linear in phi, db/dphi = M = K G + D_b, and the discrete
// adjoint solves (A - M)^T lambda = dJ/dphi.  The kernels below form M^T lambda: Q (per cell
// h_c = (K^T lambda)_c, stored as q_c = P_c^T h_c), then the transposed gradient gather G^T q
// and the direct D_b term, added to the right-hand side of A^T lambda' = g + M^T lambda.

// (I - e e^T) v
__device__ inline V3d tangential(V3d v, V3d e) { return sub(v, scl(dot(v, e), e)); }

__global__ void __launch_bounds__(BS) k_adj_corr_q(DMesh m, DScalar phi, DFaceScheme s, const double* lam,
                                                   double* q) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    V3d h = mk(0.0, 0.0, 0.0);
    for (int k = m.cf_off[c]; k < m.cf_off[c + 1]; ++k) {
        const int en = m.cf[k];
        const int f = en >> 1;
        if (f >= m.nb) {
            const int o = m.owner[f], n = m.nbr[f];
            const V3d d = faceD(m, f);
            const double dn = nrm(d);
            const Dec dc = decompose(s.variant, kappaFace(m, s, faceS(m, f), f), divs(d, dn));
            const double w = (en & 1) ? 1. - m.gc[f] : m.gc[f];
            h = add(h, scl(w * (lam[o] - lam[n]), dc.T));
        } else if (f < m.nne) {
            const int kind = phi.bck[m.fpatch[f]];
            if (kind != BC_FIXED && kind != BC_NOSLIP) continue;
            const V3d d = faceD(m, f);
            const double dn = nrm(d);
            const V3d ev = divs(d, dn);
            const Dec dc = decompose(s.variant, kappaCell(m, s, faceS(m, f), c), ev);
            h = add(h, scl(lam[c], tangential(dc.T, ev)));
        }
    }
    const double* P = m.pinv;
    const int N = m.n;
    double p[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) p[k] = P[k * N + c];
    q[c] = (p[0] * h.x + p[3] * h.y) + p[6] * h.z;
    q[N + c] = (p[1] * h.x + p[4] * h.y) + p[7] * h.z;
    q[2 * N + c] = (p[2] * h.x + p[5] * h.y) + p[8] * h.z;
}

// rhs_j = g_j + (G^T q)_j + (D_b^T lambda)_j: grad_c depends on phi_j through the weight
// w r of each face of c that reaches j, and on phi_c through minus the sum of the weights of
// its interior and fixed faces (zero-gradient, outlet and symmetry faces give phi_c - phi_c).
__global__ void __launch_bounds__(BS) k_adj_corr_rhs(DMesh m, DScalar phi, DFaceScheme s, const double* lam,
                                                     const double* q, const double* g, double* rhs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m.n) return;
    const int N = m.n;
    double acc = 0.0;
    V3d wsum = mk(0.0, 0.0, 0.0);
    for (int k = m.cf_off[j]; k < m.cf_off[j + 1]; ++k) {
        const int en = m.cf[k];
        const int f = en >> 1;
        if (f >= m.nb) {
            const int c = (en & 1) ? m.owner[f] : m.nbr[f];
            const V3d r = (en & 1) ? neg(faceD(m, f)) : faceD(m, f);  // from j to c
            const V3d w = scl(1 / (nrm(r) + EPS), r);
            acc -= dot(w, mk(q[c], q[N + c], q[2 * N + c]));  // c's weight towards j is -w
            wsum = add(wsum, w);
        } else if (f < m.nne) {
            const int kind = phi.bck[m.fpatch[f]];
            if (kind != BC_FIXED && kind != BC_NOSLIP) continue;
            const V3d r = faceD(m, f);
            const double dn = nrm(r);
            wsum = add(wsum, scl(1 / (dn + EPS), r));
            const V3d ev = divs(r, dn);
            const Dec dc = decompose(s.variant, kappaCell(m, s, faceS(m, f), j), ev);
            acc -= (dot(dc.T, ev) / dn) * lam[j];  // T_b . e (phi_b - phi_j) / |d|
        }
    }
    acc -= dot(wsum, mk(q[j], q[N + j], q[2 * N + j]));
    rhs[j] = g[j] + acc;
}

// dJ/dkappa of a diffusion term with scalar kappa, any corrector: per interior face
// Phi_f = (lambda_O - lambda_N) (a_f (phi_O - phi_N) - T_f . grad_f) for unit kappa (a_f, T_f
// from the decomposition of S), shared g Phi_f / (1 - g) Phi_f by the owner / neighbour
// (kappa_f = g kappa_O + (1 - g) kappa_N).  Non-corrected: a_f = S . d / (|d|^2 + eps), T_f = 0.
__global__ void __launch_bounds__(BS) k_adj_dkappa_faces(DMesh m, DScalar phi, int variant, const double* lam,
                                                         double* scr) {
    const int fi = blockIdx.x * blockDim.x + threadIdx.x;
    const int nint = m.F - m.nb;
    if (fi >= nint) return;
    const int f = m.nb + fi;
    const int o = m.owner[f], n = m.nbr[f];
    const V3d d = faceD(m, f);
    const V3d S = faceS(m, f);
    double a, t = 0.0;
    if (variant == 0) {
        a = dot(S, d) / (dot(d, d) + EPS);
    } else {
        const double dn = nrm(d);
        const Dec dc = decompose(variant, S, divs(d, dn));
        a = dot(dc.E, d) / (dn * dn + EPS);
        const double gw = m.gc[f];
        t = dot(dc.T, add(scl(gw, cellG(phi, o)), scl(1. - gw, cellG(phi, n))));
    }
    const double q = (lam[o] - lam[n]) * (a * (phi.v[o] - phi.v[n]) - t);
    const double w = m.gc[f];
    scr[fi] = w * q;
    scr[nint + fi] = (1.0 - w) * q;
}

// Per cell in face order: interior faces from the scratch; a fixed face adds
// lambda_C (a_b (phi_C - phi_b) - T_b . gradFixedFace) (owner kappa), a fixed-gradient face
// lambda_C (g_w . S).  dJ/dkappa_C = -sum.
__global__ void __launch_bounds__(BS) k_adj_dkappa_cells(DMesh m, DScalar phi, int variant, const double* lam,
                                                         const double* scr, double* out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    const int nint = m.F - m.nb;
    double acc = 0.0;
    for (int k = m.cf_off[c]; k < m.cf_off[c + 1]; ++k) {
        const int en = m.cf[k];
        const int f = en >> 1;
        if (f >= m.nb) {
            acc += (en & 1) ? scr[nint + (f - m.nb)] : scr[f - m.nb];
            continue;
        }
        if (f >= m.nne) continue;
        const int p = m.fpatch[f];
        const int kind = phi.bck[p];
        if (kind == BC_FIXED || kind == BC_NOSLIP) {
            const double pb = sFace(m, phi, f);
            const V3d d = faceD(m, f);
            const V3d S = faceS(m, f);
            if (variant == 0) {
                acc += lam[c] * (phi.v[c] - pb) * (dot(S, d) / (dot(d, d) + EPS));
            } else {
                const double dn = nrm(d);
                const Dec dc = decompose(variant, S, divs(d, dn));
                const double a = dot(dc.E, d) / (dn * dn + EPS);
                acc += lam[c] * (a * (phi.v[c] - pb) - dot(dc.T, gradFixedFace(m, phi, f)));
            }
        } else if (kind == BC_FIXEDGRAD) {
            const double gS = (phi.bcv[3 * p] * m.sx[f] + phi.bcv[3 * p + 1] * m.sy[f]) + phi.bcv[3 * p + 2] * m.sz[f];
            acc += lam[c] * gS;
        }
    }
    out[c] = -acc;
}

// per-block partials of sum (a - b)^2 and sum b^2
__global__ void __launch_bounds__(BS) k_diff_norms(int n, const double* a, const double* b, double* part) {
    __shared__ double s0[BS], s1[BS];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double x = 0.0, y = 0.0;
    if (i < n) {
        const double dd = a[i] - b[i];
        x = dd * dd;
        y = b[i] * b[i];
    }
    s0[threadIdx.x] = x;
    s1[threadIdx.x] = y;
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

// ---------------------------------------------------------------------- assembly ----
// The contributions of every face-based scheme of the equation on interior face f: owner and
// neighbour diagonal (Do, Dn), right-hand side (Bo, Bn) and the two off-diagonals
// (Aon in the owner's row, Ano in the neighbour's), in the schemes' constructor order.
struct FaceCoef {
    double Do, Dn, Bo, Bn, Aon, Ano;
};
__device__ inline FaceCoef faceCoef(const DMesh& m, const DEquation& e, int f) {
    int o = m.owner[f], n = m.nbr[f];
    double Do = 0, Dn = 0, Bo = 0, Bn = 0, Aon = 0, Ano = 0;
    V3d d = faceD(m, f);
    for (int k = 0; k < e.n_face; ++k) {
        const DFaceScheme& s = e.fs[k];
        if (s.kind == 1) {  // IConvection::applyInterior (convection.cpp:18-40)
            double md = s.flux[f];
            double pC = e.phi.v[o], pN = e.phi.v[n];
            V3d gC = cellG(e.phi, o), gN = cellG(e.phi, n);
            Coef a = interp(s.variant, md, pC, pN, gC, gN, d);
            Coef x = interp(s.variant, -md, pN, pC, gN, gC, neg(d));
            Do += a.aC;
            Aon += a.aN;
            Bo += a.b;
            Dn += x.aC;
            Ano += x.aN;
            Bn += x.b;
        } else {  // diffusion (diffusion.h:116-149, 171-205)
            double dn = nrm(d);
            V3d S = faceS(m, f);
            if (s.variant == 0) {
                V3d Sp = kappaFace(m, s, S, f);
                double g = dot(Sp, d) / (dn * dn + EPS);
                Do += g;
                Dn += g;
                Aon += -g;
                Ano += -g;
            } else {
                V3d ev = divs(d, dn);
                V3d Sp = kappaFace(m, s, S, f);
                Dec dc = decompose(s.variant, Sp, ev);
                double g = dot(dc.E, d) / (dn * dn + EPS);
                Do += g;
                Dn += g;
                Aon += -g;
                Ano += -g;
                double gw = m.gc[f];
                V3d gf = add(scl(gw, cellG(e.phi, o)), scl(1. - gw, cellG(e.phi, n)));
                double t = dot(dc.T, gf);
                Bo += t;
                Bn += -t;
            }
        }
    }
    return {Do, Dn, Bo, Bn, Aon, Ano};
}

// Cells sharing several faces: the CSR slot of such a pair holds its first face's coefficient;
// each further face wrote an extra slot.  One thread per slot adds them in face order (the
// reference sums duplicate triplets, linear.cpp:17-29) and clears the extras.
__global__ void k_dup_fixup(DMesh m, double* A) {
    const int g = blockIdx.x * BS + threadIdx.x;
    if (g >= m.n_dup_groups) return;
    const int s = m.dup_slot[g];
    double v = A[s];
    for (int k = m.dup_ptr[g]; k < m.dup_ptr[g + 1]; ++k) {
        v += A[m.nnz + k];
        A[m.nnz + k] = 0.0;
    }
    A[s] = v;
}

// Face-parallel form: one thread per interior face writes the two off-diagonals into their
// CSR slots and parks the four diagonal / rhs contributions in e.scr for k_assemble_cells.
__global__ void __launch_bounds__(BS) k_assemble_faces(DMesh m, DEquation e) {
    int fi = blockIdx.x * blockDim.x + threadIdx.x;
    int nint = m.F - m.nb;
    if (fi >= nint) return;
    int f = m.nb + fi;
    FaceCoef q = faceCoef(m, e, f);
    int so = m.slot_on[f], sn = m.slot_no[f];
    if (e.accumulate) {
        e.A[so] += q.Aon;
        e.A[sn] += q.Ano;
    } else {
        e.A[so] = q.Aon;
        e.A[sn] = q.Ano;
    }
    e.scr[fi] = q.Do;
    e.scr[nint + fi] = q.Dn;
    e.scr[2 * nint + fi] = q.Bo;
    e.scr[3 * nint + fi] = q.Bn;
}

// Boundary handlers of one face-based scheme on one boundary face owned by cell c.
__device__ void boundaryTerm(const DMesh& m, const DEquation& e, const DFaceScheme& s, int f, int c,
                             double& diag, double& rhs) {
    int p = m.fpatch[f];
    int kind = e.phi.bck[p];
    V3d S = faceS(m, f);
    if (s.kind == 1) {  // convection_boundary.cpp:4-61
        if (kind == BC_FIXED) {
            double md = s.flux[f];
            diag += smax(md, 0.0);
            rhs += smax(-md, 0.0) * e.phi.bcs[p];
        } else if (kind == BC_OUTLET || kind == BC_ZEROGRAD) {
            V3d uf = vFace(m, s.U, f);
            diag += dot(uf, S);
        }
        return;
    }
    // diffusion_boundary.h:113-234
    if (kind == BC_FIXED || kind == BC_NOSLIP) {
        double pw = sFace(m, e.phi, f);
        V3d dCf = faceD(m, f);
        double dn = nrm(dCf);
        V3d Sp = kappaCell(m, s, S, c);
        if (s.variant == 0) {
            double g = dot(Sp, dCf) / (dn * dn + EPS);
            diag += g;
            rhs += g * pw;
        } else {
            V3d ev = divs(dCf, dn);
            Dec dc = decompose(s.variant, Sp, ev);
            double g = dot(dc.E, dCf) / (dn * dn + EPS);
            diag += g;
            rhs += (g * pw) + dot(dc.T, gradFixedFace(m, e.phi, f));
        }
    } else if (kind == BC_FIXEDGRAD) {
        V3d wg = mk(e.phi.bcv[3 * p], e.phi.bcv[3 * p + 1], e.phi.bcv[3 * p + 2]);
        V3d Sp = kappaCell(m, s, S, c);
        rhs += -dot(wg, Sp);
    }
}

// Cell-based terms (sources, implicit fields, temporal) in constructor order, the accumulate
// and relax options, and the diagonal / rhs writes of cell c.
__device__ inline void cellTail(const DMesh& m, const DEquation& e, int c, double diag, double rhs, int32_t relax,
                                double urf) {
    double V = m.vol[c];
    for (int t = 0; t < e.n_cell; ++t) {
        const DCellTerm& ct = e.ct[t];
        switch (ct.kind) {
            case 3: {  // ConstantScalar (constant_scalar.h:27-37)
                double v = ct.phi[c] * V;
                rhs += ct.sign ? -v : v;
            } break;
            case 4: {  // source::Gradient (source/gradient.h:32-40)
                double v = ct.g[c] * V;
                rhs += ct.sign ? -v : v;
            } break;
            case 5: {  // source::Divergence (divergence.h:25-34, operations.h:144-169)
                double sum = 0.0;
                for (int k = m.cf_off[c]; k < m.cf_off[c + 1]; ++k) {
                    int en = m.cf[k];
                    int f = en >> 1;
                    if (f >= m.nne && f < m.nb) continue;
                    double fl = ct.flux[f];
                    if (en & 1) fl = -fl;
                    sum += fl;
                }
                double v = (sum / V) * V;
                rhs += ct.sign ? -v : v;
            } break;
            case 6: {  // ImplicitField (implicit_field.h:23-32)
                double v = ct.coeff * V;
                diag += ct.sign ? v : v * -1;
            } break;
            case 7: {  // BackwardEuler (backward_euler.cpp:62-83)
                diag += V / ct.coeff;
                rhs += V * ct.phi[c] / ct.coeff;
            } break;
            case 8: {  // AdamMoulton (adam_moulton.cpp:55-86): BDF2, Backward-Euler on the first step
                if (ct.phi2 == nullptr) {
                    diag += V / ct.coeff;
                    rhs += V * ct.phi[c] / ct.coeff;
                } else {
                    diag += (3.0 / 2.0) * V / ct.coeff;
                    double r = 2.0 * (V * ct.phi[c]) / ct.coeff;
                    r += -0.5 * (V * ct.phi2[c]) / ct.coeff;
                    rhs += r;
                }
            } break;
        }
    }
    int sd = m.diag[c];
    if (e.accumulate) {
        diag = e.A[sd] + diag;
        rhs = e.b[c] + rhs;
    }
    if (relax) {  // Transport::relax (transport.cpp:54-64)
        rhs += ((1.0 - urf) / urf) * (diag * e.phi.v[c]);
        diag /= urf;
    }
    e.A[sd] = diag;
    e.b[c] = rhs;
}

__global__ void __launch_bounds__(BS) k_assemble_cells(DMesh m, DEquation e, int32_t relax, double urf) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    int nint = m.F - m.nb;
    double diag = 0, rhs = 0;
    for (int k = m.cf_off[c]; k < m.cf_off[c + 1]; ++k) {
        int en = m.cf[k];
        int f = en >> 1;
        if (f >= m.nb) {
            int fi = f - m.nb;
            if (en & 1) {
                diag += e.scr[nint + fi];
                rhs += e.scr[3 * nint + fi];
            } else {
                diag += e.scr[fi];
                rhs += e.scr[2 * nint + fi];
            }
        } else if (f < m.nne) {
            for (int s = 0; s < e.n_face; ++s) boundaryTerm(m, e, e.fs[s], f, c, diag, rhs);
        }
    }
    cellTail(m, e, c, diag, rhs, relax, urf);
}

__global__ void k_relax(DMesh m, double* A, double* b, const double* phi, double urf) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    int sd = m.diag[c];
    double d = A[sd];
    b[c] += ((1.0 - urf) / urf) * (d * phi[c]);
    A[sd] = d / urf;
}

__global__ void k_ref_cell(DMesh m, double* A, double* b, int32_t cell, double ref) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // solver.cpp:50-57
        int sd = m.diag[cell];
        double d = A[sd];
        A[sd] = d + d;
        b[cell] += d * ref;
    }
}

// ------------------------------------------------------------------------ SIMPLE ----
// pressureEquationCoeffsTensor (SIMPLE.cpp:77-108): D = diag(V/(aP+eps)), 9 x n SoA.
__global__ void k_pressure_D(DMesh m, const double* a0, const double* a1, const double* a2, double* D) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    int N = m.n;
    int sd = m.diag[c];
    double V = m.vol[c];
    double du = V / (a0[sd] + EPS);
    double dv = V / (a1[sd] + EPS);
    double dw = a2 ? V / (a2[sd] + EPS) : 1.0;
    for (int k = 0; k < 9; ++k) D[k * N + c] = 0.0;
    D[0 * N + c] = du;
    D[4 * N + c] = dv;
    D[8 * N + c] = dw;
}

__device__ inline V3d dtMul(const DMesh& m, const double* T9, V3d v) {  // T v, row-major
    return mk((T9[0] * v.x + T9[1] * v.y) + T9[2] * v.z, (T9[3] * v.x + T9[4] * v.y) + T9[5] * v.z,
              (T9[6] * v.x + T9[7] * v.y) + T9[8] * v.z);
}
__device__ inline void tensorFace(const DMesh& m, const DTensor& D, int f, double* T) {
    int o = m.owner[f], n = m.nbr[f];
    double g = m.gc[f];
    int N = m.n;
    for (int k = 0; k < 9; ++k) T[k] = (g * D.t[k * N + o]) + ((1 - g) * D.t[k * N + n]);
}

// rhieChowCorrectFace (rhie_chow.h:47-63, rhie_chow.cpp:5-25) on interior faces.
__global__ void k_rhie_chow(DMesh m, DVector U, DTensor D, DScalar p, double* fd, double* ff) {
    int fi = blockIdx.x * blockDim.x + threadIdx.x;
    if (fi >= m.F - m.nb) return;
    int f = m.nb + fi;
    int o = m.owner[f], n = m.nbr[f];
    V3d Uf = mk(sFace(m, U.c[0], f), sFace(m, U.c[1], f), sFace(m, U.c[2], f));
    if (U.fd) Uf = mk(U.fd[f], U.fd[m.F + f], U.fd[2 * m.F + f]);
    double T[9];
    tensorFace(m, D, f, T);
    double gw = m.gc[f];
    V3d gavg = add(scl(gw, cellG(p, o)), scl(1. - gw, cellG(p, n)));
    V3d d = faceD(m, f);
    double dn = nrm(d);
    V3d ev = divs(d, dn);
    double gortho = (p.v[n] - p.v[o]) / dn;
    V3d gcalc = add(gavg, scl(gortho - dot(gavg, ev), ev));
    V3d r = sub(Uf, dtMul(m, T, sub(gcalc, gavg)));
    fd[f] = r.x;
    fd[m.F + f] = r.y;
    fd[2 * m.F + f] = r.z;
    ff[f] = dot(r, faceS(m, f));
}

// mapFaceValues(U.valueAtFace) (vector.h:296-316) on all non-empty faces.
__global__ void k_map_face_values(DMesh m, DVector U, double* fd, double* ff) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    if (f >= m.nne && f < m.nb) {
        fd[f] = fd[m.F + f] = fd[2 * m.F + f] = 0.0;
        ff[f] = 0.0;
        return;
    }
    V3d u = vFace(m, U, f);
    fd[f] = u.x;
    fd[m.F + f] = u.y;
    fd[2 * m.F + f] = u.z;
    ff[f] = dot(u, faceS(m, f));
}

// correctFields (SIMPLE.cpp:218-241)
__global__ void k_correct_U(DMesh m, double* ux, double* uy, double* uz, DTensor D, DScalar pp) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    int N = m.n;
    double T[9];
    for (int k = 0; k < 9; ++k) T[k] = D.t[k * N + c];
    V3d dg = dtMul(m, T, cellG(pp, c));
    ux[c] = ux[c] - dg.x;
    uy[c] = uy[c] - dg.y;
    uz[c] = uz[c] - dg.z;
}
__global__ void k_correct_mdot(DMesh m, double* fd, double* ff, DTensor D, DScalar pp) {
    int fi = blockIdx.x * blockDim.x + threadIdx.x;
    if (fi >= m.F - m.nb) return;
    int f = m.nb + fi;
    int o = m.owner[f], n = m.nbr[f];
    double T[9];
    tensorFace(m, D, f, T);
    double gw = m.gc[f];
    V3d gf = add(scl(gw, cellG(pp, o)), scl(1. - gw, cellG(pp, n)));
    V3d u = sub(mk(fd[f], fd[m.F + f], fd[2 * m.F + f]), dtMul(m, T, gf));
    fd[f] = u.x;
    fd[m.F + f] = u.y;
    fd[2 * m.F + f] = u.z;
    ff[f] = dot(u, faceS(m, f));
}
__global__ void k_update_p(double* p, const double* pp, double urf, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = p[i] + (urf * pp[i]);
}
// solveExplicitMomentum (PRIME.cpp:39-56): x = (b - (A x - aP x)) / aP, old x throughout.
__global__ void k_explicit_momentum(DMesh m, const double* A, const double* b, const double* x, double* out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    double s = 0.0;
    for (int k = m.rp[c]; k < m.rp[c + 1]; ++k) s += A[k] * x[m.col[k]];
    double ac = A[m.diag[c]];
    double H = s - ac * x[c];
    out[c] = (-H + b[c]) / ac;
}
// monitor::courantNumber (courant.cpp:7-24)
__global__ void k_courant(DMesh m, const double* ff, double dt, double* co) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    double s = 0.0;
    for (int k = m.cf_off[c]; k < m.cf_off[c + 1]; ++k) {
        int f = m.cf[k] >> 1;
        if (f >= m.nne && f < m.nb) continue;
        s += fabs(ff[f]);
    }
    co[c] = 0.5 * dt * s / m.vol[c];
}

// IncompressibleKEpsilon::equations (KEpsilon.cpp:27-77) in one pass per cell:
//   nu_t = c_mu (k k / eps)                                   (setTurbulentViscosity)
//   Pk = (nu_t (G + G^T) - (2/3) k I) : G, G(i,j) = dU_j/dx_i  (k_production.cpp:8-25)
//   nu_eff_k = nu + nu_t / sigma_k, nu_eff_eps = nu + nu_t / sigma_eps
//   eps_source = c_eps1 ((eps / k) Pk), eps_sink = c_eps2 (eps (eps / k))
// The double contraction sums the nine products as Eigen's 2-wide vectorised reduction of
// a column-major 3x3 (element e = i + 3 j): ((c0 + c2) + (c4 + c6)) + ((c1 + c3) + (c5 + c7)),
// then + c8.
__global__ void __launch_bounds__(BS) k_kepsilon(DKEps a) {
    const int c = blockIdx.x * BS + threadIdx.x;
    if (c >= a.n) return;
    const double k = a.k[c], e = a.eps[c];
    const double nut = a.c_mu * ((k * k) / e);
    double G[3][3];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) G[i][j] = a.g[j][(int64_t)i * a.n + c];
    const double s = (2.0 / 3.0) * k;
    double ce[9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double tau = nut * (G[i][j] + G[j][i]);
            tau = tau - s * (i == j ? 1.0 : 0.0);
            ce[i + 3 * j] = tau * G[i][j];
        }
    const double pk = (((ce[0] + ce[2]) + (ce[4] + ce[6])) + ((ce[1] + ce[3]) + (ce[5] + ce[7]))) + ce[8];
    const double nu = a.nu[c];
    const double eok = e / k;
    a.nu_t[c] = nut;
    a.pk[c] = pk;
    a.nu_eff_k[c] = nu + nut / a.sigma_k;
    a.nu_eff_eps[c] = nu + nut / a.sigma_eps;
    a.eps_source[c] = a.c_eps1 * (eok * pk);
    a.eps_sink[c] = a.c_eps2 * (e * eok);
}

// output::cellToPoint (export.cpp:79-112) with the normalised weights precomputed per mesh:
// point value of component k = sum over the vertex's cells, in ascending cell order, of
// w_i * f_k[cell], accumulated from 0.0 as the reference does.  One thread per vertex keeps
// the NC sums in registers; output interleaved [vertex][NC].  The adjacency is sliced
// (64 vertices per slice, slot-major inside a slice) so that each slot's cell ids and
// weights are one contiguous 256 / 512-byte load per wavefront; the blocks of one XCD own a
// contiguous range of vertices, so the cells they gather stay in that XCD's L2.
template <int NC>
__global__ void __launch_bounds__(BS) k_cell_to_point(int nv, const int32_t* __restrict__ off,
                                                      const int32_t* __restrict__ sbase,
                                                      const int32_t* __restrict__ cid, const double* __restrict__ w,
                                                      DPointSrc src, double* __restrict__ out) {
    const int nb = gridDim.x, q = nb >> 3, r = nb & 7, xcd = blockIdx.x & 7, idx = blockIdx.x >> 3;
    const int lb = (xcd < r) ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
    const int v = lb * BS + threadIdx.x;
    if (v >= nv) return;
    const int deg = off[v + 1] - off[v];
    const int base = sbase[v >> 6] + (v & 63);
    double acc[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) acc[k] = 0.0;
    for (int j = 0; j < deg; ++j) {
        const double wi = w[base + 64 * j];
        const int c = cid[base + 64 * j];
#pragma unroll
        for (int k = 0; k < NC; ++k) acc[k] += wi * src.c[k][c];
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) out[(int64_t)v * NC + k] = acc[k];
}
// constrainPPrime (SIMPLE.cpp:43-75) for a freshly zeroed p': every stored face value is 0
// (Dirichlet/symmetry faces by rule, the others take the owner value, which is 0).
__global__ void k_constrain_pprime(DMesh m, DScalar pp, double* fv) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    double v = 0.0;
    if (f < m.nne) {
        int k = pp.bck[m.fpatch[f]];
        if (k == BC_OUTLET || k == BC_ZEROGRAD) v = pp.v[m.owner[f]];
    }
    fv[f] = v;
}

inline int nblk(int64_t n) { return (int)((n + BS - 1) / BS); }

}  // namespace

#define LAUNCH(name, grid, ...)                          \
    do {                                                  \
        if ((grid) > 0) PRISM_LAUNCH(#name, name, grid, BS, __VA_ARGS__); \
    } while (0)

void launch_fill(const Launch& L, double* p, int64_t n, double v) { LAUNCH(k_fill, nblk(n), p, n, v); }
void launch_fill(const Launch& L, float* p, int64_t n, float v) { LAUNCH(k_fill_f32, nblk(n), p, n, v); }
void launch_copy(const Launch& L, double* d, const double* s, int64_t n) { LAUNCH(k_copy, nblk(n), d, s, n); }
void launch_axpby(const Launch& L, double* y, const double* x, double a, double b, int64_t n) {
    LAUNCH(k_axpby, nblk(n), y, x, a, b, n);
}
void launch_vector_flux(const Launch& L, const DMesh& m, const DVector& U, double* flux) {
    LAUNCH(k_vector_flux, nblk(m.F), m, U, flux);
}
void launch_vector_face_values(const Launch& L, const DMesh& m, const DVector& U, double* out3) {
    LAUNCH(k_vector_face_values, nblk(m.F), m, U, out3);
}
void launch_scalar_face_values(const Launch& L, const DMesh& m, const DScalar& s, double* out) {
    LAUNCH(k_scalar_face_values, nblk(m.F), m, s, out);
}
void launch_grad_ls(const Launch& L, const DMesh& m, const DScalar& s, double* g) {
    LAUNCH(k_grad_ls, nblk(m.n), m, s, g);
}
void launch_grad_ls3(const Launch& L, const DMesh& m, const DScalar (&s)[3], double* const (&g)[3]) {
    DScalar3 a;
    for (int i = 0; i < 3; ++i) {
        a.s[i] = s[i];
        a.g[i] = g[i];
    }
    LAUNCH(k_grad_ls3, nblk(m.n), m, a);
}
void launch_grad_gg_level(const Launch& L, const DMesh& m, const DScalar& s, const int32_t* cells,
                          int32_t count, double* gx, double* gy, double* gz) {
    LAUNCH(k_grad_gg_level, nblk(count), m, s, cells, count, gx, gy, gz);
}
void launch_assemble(const Launch& L, const DMesh& m, const DEquation& e, int32_t relax, double urf) {
    LAUNCH(k_assemble_faces, nblk(m.F - m.nb), m, e);
    launch_dup_fixup(L, m, e.A);
    LAUNCH(k_assemble_cells, nblk(m.n), m, e, relax, urf);
}
void launch_adj_corr_rhs(const Launch& L, const DMesh& m, const DScalar& phi, const DFaceScheme& s, const double* lam,
                         double* q, const double* g, double* rhs) {
    LAUNCH(k_adj_corr_q, nblk(m.n), m, phi, s, lam, q);
    LAUNCH(k_adj_corr_rhs, nblk(m.n), m, phi, s, lam, q, g, rhs);
}
void launch_adj_dkappa(const Launch& L, const DMesh& m, const DScalar& phi, int variant, const double* lam,
                       double* scr, double* out) {
    if (m.F > m.nb) LAUNCH(k_adj_dkappa_faces, nblk(m.F - m.nb), m, phi, variant, lam, scr);
    LAUNCH(k_adj_dkappa_cells, nblk(m.n), m, phi, variant, lam, scr, out);
}
int diff_norm_blocks(int64_t n) { return nblk(n); }
void launch_diff_norms(const Launch& L, int64_t n, const double* a, const double* b, double* part, size_t part_cap) {
    require_partials("difference norms", 2 * (size_t)nblk(n), part_cap);
    LAUNCH(k_diff_norms, nblk(n), (int)n, a, b, part);
}
void launch_dup_fixup(const Launch& L, const DMesh& m, double* A) {
    if (m.n_dup_groups > 0) LAUNCH(k_dup_fixup, nblk(m.n_dup_groups), m, A);
}
void launch_relax(const Launch& L, const DMesh& m, double* A, double* b, const double* phi, double urf) {
    LAUNCH(k_relax, nblk(m.n), m, A, b, phi, urf);
}
void launch_ref_cell(const Launch& L, const DMesh& m, double* A, double* b, int32_t cell, double ref) {
    LAUNCH(k_ref_cell, 1, m, A, b, cell, ref);
}
void launch_pressure_D(const Launch& L, const DMesh& m, const double* const* Ad, int32_t nmom, double* D9) {
    LAUNCH(k_pressure_D, nblk(m.n), m, Ad[0], Ad[1], nmom == 3 ? Ad[2] : nullptr, D9);
}
void launch_rhie_chow(const Launch& L, const DMesh& m, const DVector& U, const DTensor& D, const DScalar& p,
                      double* fd, double* ff) {
    LAUNCH(k_rhie_chow, nblk(m.F - m.nb), m, U, D, p, fd, ff);
}
void launch_map_face_values(const Launch& L, const DMesh& m, const DVector& U, double* fd, double* ff) {
    LAUNCH(k_map_face_values, nblk(m.F), m, U, fd, ff);
}
void launch_correct_U(const Launch& L, const DMesh& m, double* ux, double* uy, double* uz, const DTensor& D,
                      const DScalar& pp) {
    LAUNCH(k_correct_U, nblk(m.n), m, ux, uy, uz, D, pp);
}
void launch_correct_mdot(const Launch& L, const DMesh& m, double* fd, double* ff, const DTensor& D,
                         const DScalar& pp) {
    LAUNCH(k_correct_mdot, nblk(m.F - m.nb), m, fd, ff, D, pp);
}
void launch_update_p(const Launch& L, double* p, const double* pp, double urf, int64_t n) {
    LAUNCH(k_update_p, nblk(n), p, pp, urf, n);
}
void launch_explicit_momentum(const Launch& L, const DMesh& m, const double* A, const double* b, double* x,
                              double* tmp) {
    LAUNCH(k_explicit_momentum, nblk(m.n), m, A, b, x, tmp);
    LAUNCH(k_copy, nblk(m.n), x, tmp, (int64_t)m.n);
}
void launch_courant(const Launch& L, const DMesh& m, const double* ff, double dt, double* co) {
    LAUNCH(k_courant, nblk(m.n), m, ff, dt, co);
}
void launch_kepsilon(const Launch& L, const DKEps& a) { LAUNCH(k_kepsilon, nblk(a.n), a); }
void launch_cell_to_point(const Launch& L, int nv, int ncomp, const int32_t* off, const int32_t* sbase,
                          const int32_t* cid, const double* w, const DPointSrc& src, double* out) {
    if (ncomp == 1) LAUNCH(k_cell_to_point<1>, nblk(nv), nv, off, sbase, cid, w, src, out);
    else if (ncomp == 3) LAUNCH(k_cell_to_point<3>, nblk(nv), nv, off, sbase, cid, w, src, out);
    else if (ncomp == 9) LAUNCH(k_cell_to_point<9>, nblk(nv), nv, off, sbase, cid, w, src, out);
    else throw Error(PRISM_ERR_RUNTIME, "cell -> point interpolation of " + std::to_string(ncomp) + " components");
}
void launch_constrain_pprime(const Launch& L, const DMesh& m, const DScalar& pp, double* fv) {
    LAUNCH(k_constrain_pprime, nblk(m.F), m, pp, fv);
}
void launch_symmetry(const Launch& L, const DMesh& m, const DSymGroup& g, int32_t coord, const DVector& U,
                     const DFaceScheme& diff, double* A, double* b) {
    LAUNCH(k_symmetry, nblk(g.n), m, g, coord, U, diff, A, b);
}

}  // namespace prism_hip
