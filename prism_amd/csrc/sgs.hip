// sgs.hip -- kernels of the multicolour symmetric Gauss-Seidel preconditioned BiCGSTAB.
//
// One iteration, colours 0 .. nc-1 (positions of colour c: [cptr[c], cptr[c+1])):
//   P(c), c = 0..nc-1 : [c = 0: rho' = (r0, r) from the X partials, stopping test, beta]
//                       p = r + beta (p - omega v);  y = (p - L y) / d      on colour c
//   BY(c), c = nc-2..0: e = -(U y) / d; y += e                              on colour c
//   AV                : v = A y = p + L e; partials (r0, v)
//   S(c)              : [c = 0: alpha = rho' / (r0, v)]
//                       s = r - alpha v;  z = (s - L z) / d                 on colour c
//   BZ(c)             : e = -(U z) / d; z += e                              on colour c
//   AT                : t = A z = s + L e; partials (t, t), (t, s)
//   X                 : omega = (t, s) / (t, t); x += alpha y + omega z; r = s - omega t;
//                       partials (r0, r), (r, r)
// The products take Eisenstat's form: with f = (D + L)^-1 p the forward result and y = f + e
// after the backward sweeps, (D + U) y = D f gives A y = L y + D f = p + L (y - f) = p + L e,
// so AV and AT read only the lower-triangle entries (half the off-diagonals on two colours)
// and none of the diagonal; e of the last colour is zero and never read.
// 4 nc + 1 launches.  Every kernel that needs a scalar reduces the partial sums of the kernel
// that produced them in index order in each block (bit-identical in all blocks), or reads it
// from the parity slot the iteration's first kernel wrote.  Every kernel returns at once once
// the state holds done or flag.  Off-diagonals are stored as sliced ELL (slices of 64
// positions, column-major in a slice), so the loads of a wavefront are contiguous.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include "sgs.h"

namespace prism_hip {

namespace {

constexpr int BS = SGS_BS;
constexpr int NW = BS / 64;
constexpr double EPS = 2.220446049250313e-16;

__device__ inline double waveSum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-uniform: the state holds done or flag.
__device__ inline bool halted(const SgsState* st) {
    __shared__ int h;
    if (threadIdx.x == 0) h = (st->done != 0 || st->flag != 0);
    __syncthreads();
    return h != 0;
}

// Every block: the K totals of G partial sums per array (index order per thread, then the
// wave sums in order), identical in every block.
template <int K>
__device__ inline void reduceParts(const double* __restrict__ part, int G, double (&out)[K]) {
    __shared__ double sh[K][NW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int i = threadIdx.x; i < G; i += BS) s += part[k * G + i];
        s = waveSum(s);
        if (lane == 0) sh[k][w] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int i = 0; i < NW; ++i) s += sh[k][i];
        out[k] = s;
    }
    __syncthreads();
}

// This block's K partial sums into part[k * G + blockIdx.x].
template <int K>
__device__ inline void writeParts(double (&v)[K], double* __restrict__ part, int G) {
    __shared__ double sh[K][NW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = waveSum(v[k]);
        if (lane == 0) sh[k][w] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (int i = 0; i < NW; ++i) s += sh[k][i];
            part[k * G + blockIdx.x] = s;
        }
    }
}

// Off-diagonal product of position i with w: MODE 0 all entries, 1 those with j < lo
// (earlier colours), 2 those with j >= hi (later colours).  Padding (value 0, column i) is
// excluded by modes 1 and 2 and adds an exact zero in mode 0.
template <int MODE>
__device__ inline double offRow(const SgsView& V, int i, const double* __restrict__ w, int lo, int hi) {
    const int sl = i >> 6, lane = i & 63;
    const int o = V.soff[sl];
    const int wd = (V.soff[sl + 1] - o) >> 6;
    const int* __restrict__ cp = V.scol + o + lane;
    const double* __restrict__ vp = V.sval + o + lane;
    double sum = 0.0;
    if (wd <= 8) {
        int cj[8];
        double av[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            cj[k] = k < wd ? cp[k * 64] : i;
            av[k] = k < wd ? vp[k * 64] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const bool use = k < wd && (MODE == 0 || (MODE == 1 && cj[k] < lo) || (MODE == 2 && cj[k] >= hi));
            if (use) sum += av[k] * w[cj[k]];
        }
    } else {
        for (int k = 0; k < wd; ++k) {
            const int j = cp[k * 64];
            if (MODE == 0 || (MODE == 1 && j < lo) || (MODE == 2 && j >= hi)) sum += vp[k * 64] * w[j];
        }
    }
    return sum;
}

__global__ void k_sgs_values(int64_t slots, const int* __restrict__ emap, const double* __restrict__ val,
                             double* __restrict__ sval, int n, const int* __restrict__ dmap, double* __restrict__ d,
                             double* __restrict__ dinv, SgsState* st) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < slots; k += stride) {
        const int e = emap[k];
        sval[k] = e >= 0 ? val[e] : 0.0;
    }
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        const double a = val[dmap[i]];
        d[i] = a;
        dinv[i] = 1.0 / a;
        if (!(a != 0.0) || !(a - a == 0.0)) st->bad_diag = 1;
    }
}

__global__ void k_sgs_gather(int n, const int* __restrict__ perm, const double* __restrict__ src,
                             double* __restrict__ dst) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

__global__ void k_sgs_scatter(int n, const int* __restrict__ perm, const double* __restrict__ src,
                              double* __restrict__ dst) {
    const int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) dst[perm[i]] = src[i];
}

__global__ __launch_bounds__(BS) void k_sgs_seed(SgsView V, int start) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < V.n; i += V.G * BS) {
        const double ax = V.d[i] * V.x[i] + offRow<0>(V, i, V.x, 0, 0);
        const double ri = V.b[i] - ax;
        V.r[i] = ri;
        V.r0[i] = ri;
        if (start) {
            V.p[i] = 0.0;
            V.v[i] = 0.0;
        }
        acc[0] += ri * ri;
        acc[1] += ri * ri;
        acc[2] += V.b[i] * V.b[i];
    }
    writeParts<3>(acc, V.part_x, V.G);
}

__global__ __launch_bounds__(BS) void k_sgs_true_residual(SgsView V) {
    double acc[2] = {0.0, 0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < V.n; i += V.G * BS) {
        const double ri = V.b[i] - (V.d[i] * V.x[i] + offRow<0>(V, i, V.x, 0, 0));
        acc[0] += ri * ri;
        acc[1] += V.b[i] * V.b[i];
    }
    writeParts<2>(acc, V.part_t, V.G);
}

// P(c): search direction and forward sweep of y on colour c.
__global__ __launch_bounds__(BS) void k_sgs_fwd_p(SgsView V, int c, long long k, int fresh) {
    if (halted(V.st)) return;
    const int cur = (int)(k & 1), prev = cur ^ 1;
    double beta, omega_prev;
    if (c == 0) {
        double t[3];
        reduceParts<3>(V.part_x, V.G, t);
        const double rho = t[0], rr = t[1];
        const SgsSlot sp = V.st->s[prev];
        double tol2 = V.st->tol2;
        if (fresh == 1) tol2 = V.st->tol * V.st->tol * t[2];
        const double r0sq = fresh ? rr : sp.r0sq;
        beta = (rho / sp.rho) * (sp.alpha / sp.omega);
        omega_prev = sp.omega;
        int done = 0, flag = 0;
        if (fresh == 1 && t[2] == 0.0) done = 4;  // Eigen: b == 0 -> x = 0
        else if (!(rr == rr)) done = 3;
        else if (rr <= tol2) done = 1;
        else if (!fresh && fabs(rho) < EPS * EPS * r0sq) flag = 2;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            SgsState* st = V.st;
            st->rr = rr;
            if (fresh == 1) {
                st->bb = t[2];
                st->tol2 = tol2;
            }
            if (done || flag) {
                st->done = done;
                st->flag = flag;
                st->k_event = k;
            } else {
                SgsSlot& s = st->s[cur];
                s.rho = rho;
                s.r0sq = r0sq;
                s.beta = beta;
            }
        }
        if (done || flag) return;
    } else {
        beta = V.st->s[cur].beta;
        omega_prev = V.st->s[prev].omega;
    }
    const int lo = V.cptr[c], hi = V.cptr[c + 1];
    const int i = lo + blockIdx.x * BS + threadIdx.x;
    if (i >= hi) return;
    const double pi = V.r[i] + beta * (V.p[i] - omega_prev * V.v[i]);
    V.p[i] = pi;
    const double lw = c > 0 ? offRow<1>(V, i, V.y, lo, hi) : 0.0;
    V.y[i] = (pi - lw) * V.dinv[i];
}

// S(c): s = r - alpha v and the forward sweep of z on colour c.
__global__ __launch_bounds__(BS) void k_sgs_fwd_s(SgsView V, int c, long long k) {
    if (halted(V.st)) return;
    const int cur = (int)(k & 1);
    double alpha;
    if (c == 0) {
        double t[1];
        reduceParts<1>(V.part_v, V.G, t);
        alpha = V.st->s[cur].rho / t[0];
        if (blockIdx.x == 0 && threadIdx.x == 0) V.st->s[cur].alpha = alpha;
    } else {
        alpha = V.st->s[cur].alpha;
    }
    const int lo = V.cptr[c], hi = V.cptr[c + 1];
    const int i = lo + blockIdx.x * BS + threadIdx.x;
    if (i >= hi) return;
    const double si = V.r[i] - alpha * V.v[i];
    V.s[i] = si;
    const double lw = c > 0 ? offRow<1>(V, i, V.z, lo, hi) : 0.0;
    V.z[i] = (si - lw) * V.dinv[i];
}

// Backward sweep of w on colour c (w = y or z); the correction also goes to e.
__global__ __launch_bounds__(BS) void k_sgs_bwd(SgsView V, int c, double* __restrict__ w) {
    if (halted(V.st)) return;
    const int lo = V.cptr[c], hi = V.cptr[c + 1];
    const int i = lo + blockIdx.x * BS + threadIdx.x;
    if (i >= hi) return;
    const double ei = -offRow<2>(V, i, w, lo, hi) * V.dinv[i];
    w[i] += ei;
    V.e[i] = ei;
}

// First position of the colour that holds position i.
__device__ inline int colourStart(const SgsView& V, int i) {
    int lo = 0;
#pragma unroll
    for (int c = 1; c < SGS_MAXC; ++c)
        if (c < V.nc && i >= V.cptr[c]) lo = V.cptr[c];
    return lo;
}

// A w for w = (D + U)^-1 D (D + L)^-1 u: u + L e.
__device__ inline double eisenstat(const SgsView& V, int i, const double* __restrict__ u) {
    const int lo = colourStart(V, i);
    return u[i] + (lo > 0 ? offRow<1>(V, i, V.e, lo, 0) : 0.0);
}

// v = A y with partial (r0, v).
__global__ __launch_bounds__(BS) void k_sgs_av(SgsView V) {
    if (halted(V.st)) return;
    double acc[1] = {0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < V.n; i += V.G * BS) {
        const double vi = eisenstat(V, i, V.p);
        V.v[i] = vi;
        acc[0] += V.r0[i] * vi;
    }
    writeParts<1>(acc, V.part_v, V.G);
}

// t = A z with partials (t, t), (t, s).
__global__ __launch_bounds__(BS) void k_sgs_at(SgsView V) {
    if (halted(V.st)) return;
    double acc[2] = {0.0, 0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < V.n; i += V.G * BS) {
        const double ti = eisenstat(V, i, V.s);
        V.t[i] = ti;
        acc[0] += ti * ti;
        acc[1] += ti * V.s[i];
    }
    writeParts<2>(acc, V.part_t, V.G);
}

// omega, then x += alpha y + omega z; r = s - omega t; partials (r0, r), (r, r).
__global__ __launch_bounds__(BS) void k_sgs_x(SgsView V, long long k) {
    if (halted(V.st)) return;
    const int cur = (int)(k & 1);
    double t[2];
    reduceParts<2>(V.part_t, V.G, t);
    const double omega = t[0] > 0.0 ? t[1] / t[0] : 0.0;
    const double alpha = V.st->s[cur].alpha;
    if (blockIdx.x == 0 && threadIdx.x == 0) V.st->s[cur].omega = omega;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * BS + threadIdx.x; i < V.n; i += V.G * BS) {
        V.x[i] += alpha * V.y[i] + omega * V.z[i];
        const double ri = V.s[i] - omega * V.t[i];
        V.r[i] = ri;
        acc[0] += V.r0[i] * ri;
        acc[1] += ri * ri;
    }
    writeParts<3>(acc, V.part_x, V.G);
}

inline int blocksOf(int64_t n) { return (int)((n + BS - 1) / BS); }

}  // namespace

void launch_sgs_values(const Launch& L, int64_t slots, const int* emap, const double* val, double* sval, int n,
                       const int* dmap, double* d, double* dinv, SgsState* st) {
    const int64_t m = slots > n ? slots : n;
    const int g = (int)(m / BS + 1 < 4096 ? m / BS + 1 : 4096);
    PRISM_LAUNCH("sgs_values", k_sgs_values, g, BS, slots, emap, val, sval, n, dmap, d, dinv, st);
}

void launch_sgs_gather(const Launch& L, int n, const int* perm, const double* src, double* dst) {
    PRISM_LAUNCH("sgs_gather", k_sgs_gather, blocksOf(n), BS, n, perm, src, dst);
}

void launch_sgs_scatter(const Launch& L, int n, const int* perm, const double* src, double* dst) {
    PRISM_LAUNCH("sgs_scatter", k_sgs_scatter, blocksOf(n), BS, n, perm, src, dst);
}

void launch_sgs_seed(const Launch& L, const SgsView& V, bool start) {
    PRISM_LAUNCH("sgs_seed", k_sgs_seed, V.G, BS, V, start ? 1 : 0);
}

void launch_sgs_true_residual(const Launch& L, const SgsView& V) {
    PRISM_LAUNCH("sgs_true_residual", k_sgs_true_residual, V.G, BS, V);
}

void launch_sgs_iteration(const Launch& L, const SgsView& V, long long k, int fresh) {
    auto rows = [&](int c) { return blocksOf(V.cptr[c + 1] - V.cptr[c]); };
    for (int c = 0; c < V.nc; ++c) PRISM_LAUNCH("sgs_fwd_p", k_sgs_fwd_p, rows(c), BS, V, c, k, fresh);
    for (int c = V.nc - 2; c >= 0; --c) PRISM_LAUNCH("sgs_bwd", k_sgs_bwd, rows(c), BS, V, c, V.y);
    PRISM_LAUNCH("sgs_av", k_sgs_av, V.G, BS, V);
    for (int c = 0; c < V.nc; ++c) PRISM_LAUNCH("sgs_fwd_s", k_sgs_fwd_s, rows(c), BS, V, c, k);
    for (int c = V.nc - 2; c >= 0; --c) PRISM_LAUNCH("sgs_bwd", k_sgs_bwd, rows(c), BS, V, c, V.z);
    PRISM_LAUNCH("sgs_at", k_sgs_at, V.G, BS, V);
    PRISM_LAUNCH("sgs_x", k_sgs_x, V.G, BS, V, k);
}

}  // namespace prism_hip
