// krylov.hip -- BiCGSTAB for gfx950: three kernels per iteration, DIA or sliced-ELL SpMV that
// gathers one vector, FP64 dot products, deterministic fence-free reductions.
//
// Iteration (Eigen's BiCGSTAB with identity or right Jacobi preconditioner M = diag(A); with
// Jacobi the SpMV values are the column-scaled A M^-1, so the SpMVs are unchanged and only the
// x update applies M^-1).  rhat = (A M^-1)^T r0 is formed once per start or restart:
//   V: alpha = rho / (rhat, p) [= rho / (r0, v)], after the convergence / event decisions on
//      the X partials; v = A M^-1 p; s = r - alpha v                          (SpMV)
//   T: t = A M^-1 s; partials of (t,t), (t,s), (r0,s), (r0,t)                 (SpMV)
//   X: omega = (t,s)/(t,t); rho' = (r0,s) - omega (r0,t) [= (r0, r')]; beta;
//      x += alpha M^-1 p + omega M^-1 s; r' = s - omega t; p' = r' + beta (p - omega v);
//      partials of (rhat, p'), (r', r')                                       (stream)
// Three reductions, the fewest BiCGSTAB admits, each taken at a kernel boundary that the
// SpMVs need anyway.  After a seed (start, restart, reliable update) an entry kernel forms p
// from the seeded r.  The five-kernel form (p and s updates in kernels of their own, (r0, v)
// reduced directly) is selectable for comparison.
// Each SpMV gathers a single vector (one L2 line per neighbour).  V, which writes no partial
// sums, may run on more blocks than T and X (krylov_grid).  Reductions: each block writes its partial sums; the next consuming kernel's
// blocks each reduce them in index order (wave shuffles), so every block sees bit-identical
// scalars with no fence, atomic or extra launch.  Every kernel returns at once while the
// device has flagged an event (convergence, reliable update, restart).
// Matrix: DIA (diagonal-major values, offsets as kernel arguments) for patterns with at most
// 8 distinct offsets, else sliced ELL with 16-bit offsets from the row when every
// |col - row| < 2^15, else 32-bit absolute indices.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdlib>

#include "../../include/prism_hip.h"
#include "internal.h"
#include "krylov.h"

namespace prism_hip {

namespace {

constexpr int BS = 512;
constexpr int NW = BS / 64;
constexpr int LDS_CAP = 8 * BS;  // staged products per chunk of BS rows
constexpr int GMAX = 1024;       // partial-sum capacity of a reduction (blocks at most)
constexpr int GDEF = 512;        // default cap on the blocks of the iteration kernels
constexpr int GSELL = 768;       // the cap on sliced-ELL matrices

// blockIdx -> logical block, so that the blocks of one XCD (dispatched round-robin over the
// 8 XCDs) own one contiguous range of rows and share an L2 for the gathered vector.
__device__ inline int logicalBlock(int nblk) {
    int b = blockIdx.x;
    int q = nblk >> 3, r = nblk & 7;
    int xcd = b & 7, idx = b >> 3;
    return (xcd < r) ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

__device__ inline double waveSum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// max that keeps a NaN (fmax drops it): a NaN residual must not read as converged
__device__ inline double nanMax(double a, double b) { return (a != a || b != b) ? a + b : fmax(a, b); }
__device__ inline double waveMax(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = nanMax(v, __shfl_xor(v, off, 64));
    return v;
}

// Every block: the K totals of the previous kernel's partials (index order, wavefront 0).
template <int K>
__device__ inline void reducePartials(const double* __restrict__ part, int G, double (&out)[K]) {
    __shared__ double sh[K];
    if (threadIdx.x < 64) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (int i = threadIdx.x; i < G; i += 64) s += part[k * G + i];
            s = waveSum(s);
            if (threadIdx.x == 0) sh[k] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = sh[k];
}

// This block's K partial sums into part[k * G + lb].
template <int K>
__device__ inline void writePartials(double (&v)[K], double* __restrict__ part, int G, int lb) {
    __shared__ double sh[K][NW];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
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
            part[k * G + lb] = s;
        }
    }
}

template <typename T>
struct LdsBuf {
    __device__ static T* get() {
        __shared__ T buf[LDS_CAP];
        return buf;
    }
};

// One chunk of BS rows starting at row0: y_row = sum_k val[k] * op(col[k]) in column order.
// Products are formed with coalesced loads into LDS (CSR-stream), then each thread sums its
// row; a chunk holding more than LDS_CAP entries falls back to a per-row loop.
template <typename T, typename Op>
__device__ inline T spmvChunk(const int* __restrict__ rp, const int* __restrict__ col, const T* __restrict__ val,
                              int n, int row0, T* lds, Op op) {
    int rows = min(BS, n - row0);
    int k0 = rp[row0];
    int cnt = rp[row0 + rows] - k0;
    int row = row0 + threadIdx.x;
    T s = T(0);
    if (cnt <= LDS_CAP) {
        // all index/value loads first, then all gathers: up to 2 x SLOTS loads in flight
        constexpr int SLOTS = LDS_CAP / BS;
        int cj[SLOTS];
        T vv[SLOTS];
#pragma unroll
        for (int u = 0; u < SLOTS; ++u) {
            int k = threadIdx.x + u * BS;
            bool ok = k < cnt;
            cj[u] = ok ? col[k0 + k] : 0;
            vv[u] = ok ? val[k0 + k] : T(0);
        }
        T xv[SLOTS];
#pragma unroll
        for (int u = 0; u < SLOTS; ++u) xv[u] = (threadIdx.x + u * BS < cnt) ? op(cj[u]) : T(0);
#pragma unroll
        for (int u = 0; u < SLOTS; ++u) {
            int k = threadIdx.x + u * BS;
            if (k < cnt) lds[k] = vv[u] * xv[u];
        }
        __syncthreads();
        if (threadIdx.x < rows) {
            int a = rp[row] - k0, b = rp[row + 1] - k0;
            for (int k = a; k < b; ++k) s += lds[k];
        }
        __syncthreads();
    } else if (threadIdx.x < rows) {
        for (int k = rp[row]; k < rp[row + 1]; ++k) s += val[k] * op(col[k]);
    }
    return s;
}

// Prologue of the iteration kernels: thread 0 reads the event flags while every thread
// reduces its share of the previous kernel's K partial arrays (index order, wave shuffles,
// then the NW wave sums in order); one barrier.  Every block obtains bit-identical totals.
// The return value (halted) is block-uniform.
template <int K>
__device__ inline bool prologue(const KrylovScalars* ks, const double* __restrict__ part, int G, double (&out)[K]) {
    __shared__ double sh[K][NW];
    __shared__ int h;
    if (threadIdx.x == 0) h = (ks->done != 0 || ks->flag != 0);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
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
    return h != 0;
}

// The same with a caller's prefetch issued between the partial loads and their use: the
// partials (written by other XCDs, so served from beyond L2) and the caller's first rows
// are in flight together.  Each thread's partials are summed in index order, as above.
template <int K, typename Pre>
__device__ inline bool prologue(const KrylovScalars* ks, const double* __restrict__ part, int G, double (&out)[K],
                                Pre&& pre) {
    constexpr int PMAX = GMAX / BS;
    __shared__ double sh[K][NW];
    __shared__ int h;
    if (threadIdx.x == 0) h = (ks->done != 0 || ks->flag != 0);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double v[K][PMAX];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int j = 0; j < PMAX; ++j) {
            int i = threadIdx.x + j * BS;
            v[k][j] = i < G ? part[k * G + i] : 0.0;
        }
    pre();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = v[k][0];
#pragma unroll
        for (int j = 1; j < PMAX; ++j) s += v[k][j];
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
    return h != 0;
}

// Sliced-ELL row product: slices of 64 rows (one wavefront), rows padded to the slice width
// with zero values, slots paired per lane (sellPos); entries keep the CSR column order, so
// the sum is formed in the same order as a CSR row product.  REL: 16-bit column offsets from
// the row plus the base of the slot (one value per slice and slot, a scalar load).
//
// Widths 4 to 8 take a body of exactly w slots (a uniform branch: a wavefront is one slice):
// one load per lane for each pair of slots' values and one for their columns, all issued
// before the gathers, all nontemporal (the matrix is streamed once per product, and L2 is
// left to the gathered vector).
template <typename T>
using Pair = T __attribute__((ext_vector_type(2)));

template <int W, typename T, bool REL, int DIFF, typename Op>
__device__ inline T sellExact(const int* __restrict__ scol, const short* __restrict__ srel,
                              const int* __restrict__ sbase, const T* __restrict__ sval, int o, int lane, int row,
                              Op op) {
    int cj[W];
    T vv[W];
#pragma unroll
    for (int m = 0; m < W / 2; ++m) {
        const int p = o + 128 * m + 2 * lane;
        const Pair<T> a = __builtin_nontemporal_load(reinterpret_cast<const Pair<T>*>(sval + p));
        vv[2 * m] = a.x;
        vv[2 * m + 1] = a.y;
        if constexpr (REL) {
            const Pair<short> d = __builtin_nontemporal_load(reinterpret_cast<const Pair<short>*>(srel + p));
            cj[2 * m] = row + sbase[(o >> 6) + 2 * m] + (int)d.x;
            cj[2 * m + 1] = row + sbase[(o >> 6) + 2 * m + 1] + (int)d.y;
        } else {
            const Pair<int> c = __builtin_nontemporal_load(reinterpret_cast<const Pair<int>*>(scol + p));
            cj[2 * m] = c.x;
            cj[2 * m + 1] = c.y;
        }
    }
    if constexpr (W & 1) {
        const int p = o + 64 * (W - 1) + lane;
        vv[W - 1] = __builtin_nontemporal_load(sval + p);
        if constexpr (REL) cj[W - 1] = row + sbase[(o >> 6) + W - 1] + (int)__builtin_nontemporal_load(srel + p);
        else cj[W - 1] = __builtin_nontemporal_load(scol + p);
    }
    T xv[W];
#pragma unroll
    for (int k = 0; k < W; ++k) xv[k] = op(cj[k]);
    T sum = T(0);
    if constexpr (DIFF) {  // the row's own entry (and the padding) holds the row sum
        const T xi = op(row);
#pragma unroll
        for (int k = 0; k < W; ++k) {
            if (cj[k] != row) sum += vv[k] * (xv[k] - xi);
            else if (DIFF == 1) sum += vv[k] * xi;
        }
    } else {
#pragma unroll
        for (int k = 0; k < W; ++k) sum += vv[k] * xv[k];
    }
    return sum;
}

//
// Every caller maps the 64 rows of a slice to one wavefront (rows r0 + threadIdx.x + k BS with
// r0 and BS multiples of 64), so the slice index is wave-uniform: read from the first lane, it
// makes the slice offsets and the slot bases scalar loads instead of vector ones.
template <typename T, bool REL, int DIFF, typename Op>
__device__ inline T sellRow(const int* __restrict__ soff, const int* __restrict__ scol, const short* __restrict__ srel,
                            const int* __restrict__ sbase, const T* __restrict__ sval, int row, Op op) {
    const int s = __builtin_amdgcn_readfirstlane(row >> 6), lane = row & 63;
    const int o = soff[s];
    const int w = (soff[s + 1] - o) >> 6;  // >= 1: every row holds its diagonal
    switch (w) {
    case 4: return sellExact<4, T, REL, DIFF>(scol, srel, sbase, sval, o, lane, row, op);
    case 5: return sellExact<5, T, REL, DIFF>(scol, srel, sbase, sval, o, lane, row, op);
    case 6: return sellExact<6, T, REL, DIFF>(scol, srel, sbase, sval, o, lane, row, op);
    case 7: return sellExact<7, T, REL, DIFF>(scol, srel, sbase, sval, o, lane, row, op);
    case 8: return sellExact<8, T, REL, DIFF>(scol, srel, sbase, sval, o, lane, row, op);
    default: break;
    }
    T sum = T(0);
    const T xi = DIFF ? op(row) : T(0);
    for (int k = 0; k < w; ++k) {
        const int p = (int)sellPos(o, w, k, lane);
        int j;
        if constexpr (REL) j = row + sbase[(o >> 6) + k] + (int)srel[p];
        else j = scol[p];
        if (DIFF && j != row) sum += sval[p] * (op(j) - xi);
        else if (DIFF != 2) sum += sval[p] * op(j);
    }
    return sum;
}

// Diagonal (DIA) row product for patterns with few distinct offsets col - row (structured
// topologies): values of diagonal d at vals[d n + row], zero where the row has no such
// neighbour; the offsets are kernel arguments (scalar registers), so there are no index
// loads and the gathers of a wavefront are contiguous.  Offsets ascend, so the products are
// summed in CSR column order; the padding adds exact zeros.
template <typename V, int DIFF, typename Op>
__device__ inline V diaRow(const int (&off)[DIA_MAX], int D, const V* __restrict__ vals, int n, int row, Op op) {
    V vv[DIA_MAX], xv[DIA_MAX];
#pragma unroll
    for (int d = 0; d < DIA_MAX; ++d) vv[d] = (d < D && !(DIFF == 2 && off[d] == 0)) ? vals[(int64_t)d * n + row] : V(0);
#pragma unroll
    for (int d = 0; d < DIA_MAX; ++d) {
        int j = row + off[d];
        xv[d] = (d < D && (unsigned)j < (unsigned)n) ? op(j) : V(0);
    }
    V sum = V(0);
    const V xi = DIFF ? op(row) : V(0);
#pragma unroll
    for (int d = 0; d < DIA_MAX; ++d) {
        if (d >= D || (DIFF == 2 && off[d] == 0)) continue;
        sum += (DIFF && off[d] != 0) ? vv[d] * (xv[d] - xi) : vv[d] * xv[d];
    }
    return sum;
}

// Symmetric DIA (offsets -o_k .. 0 .. o_k, values exactly symmetric): the entry of row i on
// the lower diagonal -o is the upper diagonal's value at row i - o, so the lower diagonals'
// storage is never read; the same products in the same order as diaRow.
template <typename V, int DIFF, typename Op>
__device__ inline V diaRowSym(const int (&off)[DIA_MAX], int D, const V* __restrict__ vals, int n, int row, Op op) {
    V vv[DIA_MAX], xv[DIA_MAX];
    const int mid = D >> 1;
#pragma unroll
    for (int d = 0; d < DIA_MAX; ++d) {
        int j = row + off[d];
        bool ok = d < D && (unsigned)j < (unsigned)n;
        bool skip = DIFF == 2 && off[d] == 0;
        vv[d] = ok && !skip ? (d < mid ? vals[(int64_t)(D - 1 - d) * n + j] : vals[(int64_t)d * n + row]) : V(0);
        xv[d] = ok ? op(j) : V(0);
    }
    V sum = V(0);
    const V xi = DIFF ? op(row) : V(0);
#pragma unroll
    for (int d = 0; d < DIA_MAX; ++d) {
        if (d >= D || (DIFF == 2 && off[d] == 0)) continue;
        sum += (DIFF && off[d] != 0) ? vv[d] * (xv[d] - xi) : vv[d] * xv[d];
    }
    return sum;
}

// Row product in the matrix format of the solve (FMT_SELL32, FMT_SELL16, FMT_DIA, FMT_DIA_SYM).
// DIFF: the difference form of the same product, sum_i x_i + sum_{j != i} a_ij (x_j - x_i)
// with the row sum s_i = sum_j a_ij stored in the diagonal slot (algebraically A x).  In FP32
// it keeps the accuracy that the plain form loses to cancellation on smooth x: a Laplacian's
// A x of a smooth vector is far smaller than the terms a_ij x_j it sums, and each of those
// carries a rounding error of the terms' size, while the differences x_j - x_i of close values
// are exact.  The FP32 iteration of an all-Neumann pressure system applies its operator
// to the smooth modes that govern convergence to FP32 accuracy this way (and diverges on the
// tetrahedral systems without it).  DIFF = 2: a matrix with zero row sums (the deflated
// all-Neumann system, checked to 1e-12 of its diagonal), whose diagonal term is dropped and
// whose diagonal values are not read: one value array fewer per SpMV.
template <int FMT, int DIFF, typename V, typename KV, typename Op>
__device__ inline V matRow(const KV& kv, const V* __restrict__ vals, int row, Op op) {
    if constexpr (FMT == FMT_DIA) return diaRow<V, DIFF>(kv.doff, kv.ndia, vals, kv.n, row, op);
    else if constexpr (FMT == FMT_DIA_SYM) return diaRowSym<V, DIFF>(kv.doff, kv.ndia, vals, kv.n, row, op);
    else return sellRow<V, FMT == FMT_SELL16, DIFF>(kv.soff, kv.scol, kv.srel, kv.sbase, vals, row, op);
}

// 1.0 per block where some lower entry (i, i - o) differs from the upper entry (i - o, i).
__global__ void __launch_bounds__(BS) k_dia_symcheck(KView<double> kv, const double* vals, double* part) {
    int i = blockIdx.x * BS + threadIdx.x;
    int bad = 0;
    if (i < kv.n) {
        const int D = kv.ndia, mid = D >> 1;
        for (int d = 0; d < mid; ++d) {
            int j = i + kv.doff[d];
            if ((unsigned)j < (unsigned)kv.n && vals[(int64_t)d * kv.n + i] != vals[(int64_t)(D - 1 - d) * kv.n + j])
                bad = 1;
        }
    }
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (bad) any = 1;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = any ? 1.0 : 0.0;
}

// Event flag of the previous kernel, read by thread 0 and shared (block-uniform result).
__device__ inline bool halted(const KrylovScalars* ks) {
    __shared__ int h;
    if (threadIdx.x == 0) h = (ks->done != 0 || ks->flag != 0);
    __syncthreads();
    return h != 0;
}

struct Rows {  // chunk range of logical block lb (CSR helpers)
    int c0, c1;
    __device__ Rows(int n, int cpb, int lb) {
        int nchunks = (n + BS - 1) / BS;
        c0 = lb * cpb;
        c1 = min(c0 + cpb, nchunks);
    }
};

struct RowRange {  // contiguous rows of logical block lb: `spb` 64-row slices per block
    int r0, r1;
    __device__ RowRange(int n, int spb, int lb) {
        r0 = lb * spb * 64;
        r1 = min(r0 + spb * 64, n);
    }
};

// ---------------------------------------------------------------------- kernel P ----
template <typename T, bool OUT>
__global__ void __launch_bounds__(BS) k_iter_p(KView<T> kv, const double* part_in, double* part_out) {
    KrylovScalars* ks = kv.ks;
    const int lb = logicalBlock(kv.G);
    RowRange R(kv.n, kv.slices_per_block, lb);
    const T* __restrict__ r = kv.r;
    const T* __restrict__ po = kv.p_old;
    const T* __restrict__ vo = kv.v_old;
    const int row1 = R.r0 + threadIdx.x;  // first row of this thread, fetched with the partials
    T r1 = T(0), p1 = T(0), v1 = T(0);
    double red[2];
    if (prologue<2>(ks, part_in, kv.G, red, [&] {
            if (row1 < R.r1) {
                r1 = r[row1];
                p1 = po[row1];
                v1 = vo[row1];
            }
        }))
        return;
    const double rr = red[0], rho = red[1];
    int done = 0, flag = 0;
    if (!(rr == rr) || !(rho == rho)) done = 3;
    else if (kv.mixed) {
        if (rr <= ks->tol2 || rr <= ks->delta2 * ks->rr_ref || ks->iter >= ks->maxit) flag = 1;
    } else if (rr <= ks->tol2) done = 1;
    else if (ks->iter >= ks->maxit) done = 2;
    if (!done && !flag && fabs(rho) < ks->breakdown * ks->r0n) flag = 2;
    if (done || flag) {
        if (lb == 0 && threadIdx.x == 0) {
            ks->rr = rr;
            if (done) ks->done = done;
            else ks->flag = flag;
        }
        return;
    }
    const double beta = (rho / ks->rho[kv.par ^ 1]) * (ks->alpha / ks->omega);
    const T b = (T)beta, w = (T)ks->omega;
    T* __restrict__ pn = kv.p_new;
    if constexpr (OUT) {  // entry of the three-kernel iteration: (rhat, p), (r, r) partials
        const T* __restrict__ rh = kv.rhat;
        double acc[2] = {0.0, 0.0};
        for (int row = row1; row < R.r1; row += BS) {
            T ri = row == row1 ? r1 : r[row];
            T pi = ri + b * ((row == row1 ? p1 : po[row]) - w * (row == row1 ? v1 : vo[row]));
            pn[row] = pi;
            acc[0] += (double)rh[row] * (double)pi;
            acc[1] += (double)ri * (double)ri;
        }
        writePartials<2>(acc, part_out, kv.G, lb);
    } else {
        if (row1 < R.r1) pn[row1] = r1 + b * (p1 - w * v1);
#pragma unroll 4
        for (int row = row1 + BS; row < R.r1; row += BS) pn[row] = r[row] + b * (po[row] - w * vo[row]);
    }
    if (lb == 0 && threadIdx.x == 0) {
        ks->rho[kv.par] = rho;
        ks->rr = rr;
        ks->beta = beta;
    }
}

// ---------------------------------------------------------------------- kernel V ----
template <typename T, int FMT>
__global__ void __launch_bounds__(BS) k_iter_v(KView<T> kv, double* part_out) {
    if (halted(kv.ks)) return;
    const int lb = logicalBlock(kv.G);
    const T* __restrict__ pn = kv.p_new;
    const T* __restrict__ r0 = kv.r0;
    T* __restrict__ vn = kv.v_new;
    double acc = 0.0;
    RowRange R(kv.n, kv.slices_per_block, lb);
#pragma unroll 2
    for (int row = R.r0 + threadIdx.x; row < R.r1; row += BS) {
        T vi = matRow<FMT, 0>(kv, kv.sval, row, [&](int j) { return pn[j]; });
        vn[row] = vi;
        acc += (double)r0[row] * (double)vi;
    }
    double v[1] = {acc};
    writePartials<1>(v, part_out, kv.G, lb);
}

// ---------------------------------------------------------------------- kernel S ----
template <typename T>
__global__ void __launch_bounds__(BS) k_iter_s(KView<T> kv, const double* part_in) {
    KrylovScalars* ks = kv.ks;
    const int lb = logicalBlock(kv.G);
    RowRange R(kv.n, kv.slices_per_block, lb);
    const T* __restrict__ r = kv.r;
    const T* __restrict__ vn = kv.v_new;
    const int row1 = R.r0 + threadIdx.x;
    T r1 = T(0), v1 = T(0);
    double red[1];
    if (prologue<1>(ks, part_in, kv.G, red, [&] {
            if (row1 < R.r1) {
                r1 = r[row1];
                v1 = vn[row1];
            }
        }))
        return;
    const double alpha = ks->rho[kv.par] / red[0];
    const T a = (T)alpha;
    T* __restrict__ sv = kv.s;
    if (row1 < R.r1) sv[row1] = r1 - a * v1;
#pragma unroll 4
    for (int row = row1 + BS; row < R.r1; row += BS) sv[row] = r[row] - a * vn[row];
    if (lb == 0 && threadIdx.x == 0) {
        ks->alpha = alpha;
        ks->r0v = red[0];
    }
}

// ---------------------------------------------------------------------- kernel T ----
template <typename T, int FMT>
__global__ void __launch_bounds__(BS) k_iter_t(KView<T> kv, double* part_out) {
    if (halted(kv.ks)) return;
    const int lb = logicalBlock(kv.G);
    const T* __restrict__ sv = kv.s;
    T* __restrict__ tv = kv.t;
    double tt = 0.0, ts = 0.0;
    RowRange R(kv.n, kv.slices_per_block, lb);
#pragma unroll 2
    for (int row = R.r0 + threadIdx.x; row < R.r1; row += BS) {
        T ti = matRow<FMT, 0>(kv, kv.sval, row, [&](int j) { return sv[j]; });
        tv[row] = ti;
        double td = (double)ti;
        tt += td * td;
        ts += td * (double)sv[row];
    }
    double v[2] = {tt, ts};
    writePartials<2>(v, part_out, kv.G, lb);
}

// ---------------------------------------------------------------------- kernel X ----
template <typename T>
__global__ void __launch_bounds__(BS) k_iter_xr(KView<T> kv, const double* part_in, double* part_out) {
    KrylovScalars* ks = kv.ks;
    const int lb = logicalBlock(kv.G);
    RowRange R(kv.n, kv.slices_per_block, lb);
    const int pc = kv.precond;
    T* __restrict__ xv = kv.x;
    T* __restrict__ rv = kv.r;
    const T* __restrict__ pn = kv.p_new;
    const T* __restrict__ sv = kv.s;
    const T* __restrict__ tv = kv.t;
    const T* __restrict__ r0 = kv.r0;
    const T* __restrict__ dinv = kv.dinv;
    const int row1 = R.r0 + threadIdx.x;
    T p1 = T(0), s1 = T(0), t1 = T(0), x1 = T(0), z1 = T(0);
    double red[2];
    if (prologue<2>(ks, part_in, kv.G, red, [&] {
            if (row1 < R.r1) {
                p1 = pn[row1];
                s1 = sv[row1];
                t1 = tv[row1];
                x1 = xv[row1];
                z1 = r0[row1];
            }
        }))
        return;
    const double omega = red[0] > 0.0 ? red[1] / red[0] : 0.0;
    const T a = (T)ks->alpha, w = (T)omega;
    double rr = 0.0, r0r = 0.0;
    if (row1 < R.r1) {
        T yi = pc ? dinv[row1] * p1 : p1;
        T zi = pc ? dinv[row1] * s1 : s1;
        xv[row1] = x1 + (a * yi + w * zi);
        T ri = s1 - w * t1;
        rv[row1] = ri;
        double rd = (double)ri;
        rr += rd * rd;
        r0r += (double)z1 * rd;
    }
#pragma unroll 4
    for (int row = row1 + BS; row < R.r1; row += BS) {
        T pi = pn[row], si = sv[row];
        T yi = pc ? dinv[row] * pi : pi;
        T zi = pc ? dinv[row] * si : si;
        xv[row] += a * yi + w * zi;
        T ri = si - w * tv[row];
        rv[row] = ri;
        double rd = (double)ri;
        rr += rd * rd;
        r0r += (double)r0[row] * rd;
    }
    double v[2] = {rr, r0r};
    writePartials<2>(v, part_out, kv.G, lb);
    if (lb == 0 && threadIdx.x == 0) {
        ks->omega = omega;
        ks->tt = red[0];
        ks->ts = red[1];
        ks->iter += 1;
        ks->nit += 1;
    }
}

// ------------------------------------------------------------ three-kernel iteration ----
// Iteration k with p_k = p_new, v_k = v_new, rho_k = rho[par] and the (rhat, p_k), (r_k, r_k)
// partials in part_x (from X of iteration k-1, or from the entry kernel after a seed):
//   V: convergence / event decisions on (r,r) and rho_k (Eigen's order: convergence, then
//      rho breakdown); alpha = rho_k / (rhat, p_k), which equals rho_k / (r0, v_k) because
//      rhat = (A M^-1)^T r0; v = A M^-1 p; s = r - alpha v                 (SpMV)
//   T: t = A M^-1 s; partials of (t,t), (t,s), (r0,s), (r0,t)             (SpMV)
//   X: omega = (t,s)/(t,t); rho_k+1 = (r0,s) - omega (r0,t) = (r0, s - omega t);
//      beta = rho_k+1/rho_k * alpha/omega; x += alpha M^-1 p + omega M^-1 s; r = s - omega t;
//      p_k+1 = r + beta (p - omega v) into p_old; partials (rhat, p_k+1), (r, r)   (stream)
// Three reductions, the minimum for BiCGSTAB, hence three launches per iteration instead of
// five, and 16 vector passes instead of 19.  When V raises an event, p_k+1 is discarded: the
// host services the event, seeds, and the entry kernel forms p from the seeded r.
// V runs on its own grid (Gv blocks); it reduces the G partial sums of X (or of the entry
// kernel) and writes none.
template <typename T, int FMT, int DIFF>
__global__ void __launch_bounds__(BS) k_iter_v3(KView<T> kv, const double* part_in) {
    KrylovScalars* ks = kv.ks;
    const int lb = logicalBlock(kv.Gv);
    RowRange R(kv.n, kv.spb_v, lb);
    const T* __restrict__ pn = kv.p_new;
    const T* __restrict__ r = kv.r;
    auto pAt = [&](int j) { return pn[j]; };
    // the first row's product needs no scalar: it is formed while the partials arrive (not
    // after an event, when the launch only returns)
    const int halt = ks->done | ks->flag;
    const int row1 = R.r0 + threadIdx.x;
    T v1 = T(0), r1 = T(0);
    double red[2];
    if (prologue<2>(ks, part_in, kv.G, red, [&] {
            if (!halt && row1 < R.r1) {
                v1 = matRow<FMT, DIFF>(kv, kv.sval, row1, pAt);
                r1 = r[row1];
            }
        }))
        return;
    const double rp = red[0], rr = red[1], rho = ks->rho[kv.par];
    int done = 0, flag = 0;
    if (!(rr == rr)) done = 3;
    else if (kv.mixed) {
        if (rr <= ks->tol2 || rr <= ks->delta2 * ks->rr_ref || ks->iter >= ks->maxit) flag = 1;
    } else if (rr <= ks->tol2) done = 1;
    else if (ks->iter >= ks->maxit) done = 2;
    if (!done && !flag) {
        if (!(rp == rp) || !(rho == rho)) done = 3;
        else if (fabs(rho) < ks->breakdown * ks->r0n) flag = 2;
    }
    if (done || flag) {
        if (lb == 0 && threadIdx.x == 0) {
            ks->rr = rr;
            if (done) ks->done = done;
            else ks->flag = flag;
        }
        return;
    }
    const double alpha = rho / rp;
    const T a = (T)alpha;
    T* __restrict__ vn = kv.v_new;
    T* __restrict__ sv = kv.s;
    if (row1 < R.r1) {
        vn[row1] = v1;
        sv[row1] = r1 - a * v1;
    }
#pragma unroll 2
    for (int row = row1 + BS; row < R.r1; row += BS) {
        T vi = matRow<FMT, DIFF>(kv, kv.sval, row, pAt);
        vn[row] = vi;
        sv[row] = r[row] - a * vi;
    }
    if (lb == 0 && threadIdx.x == 0) {
        ks->alpha = alpha;
        ks->r0v = rp;
        ks->rr = rr;
    }
}

template <typename T, int FMT, int DIFF>
__global__ void __launch_bounds__(BS) k_iter_t3(KView<T> kv, double* part_out) {
    // Event flags of V, read by every thread (the same value: no kernel writes them while this
    // one runs), tested after the first row's product and before any store, so the read
    // overlaps the row's loads instead of a barrier waiting for it.
    const int halt = kv.ks->done | kv.ks->flag;
    const int lb = logicalBlock(kv.G);
    const T* __restrict__ sv = kv.s;
    const T* __restrict__ r0 = kv.r0;
    T* __restrict__ tv = kv.t;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    RowRange R(kv.n, kv.slices_per_block, lb);
    const int row1 = R.r0 + threadIdx.x;
    if (halt && row1 >= R.r1) return;
    auto sAt = [&](int j) { return sv[j]; };
    auto dots = [&](int row, T ti) {
        tv[row] = ti;
        double td = (double)ti, sd = (double)sv[row], zd = (double)r0[row];
        v[0] += td * td;
        v[1] += td * sd;
        v[2] += zd * sd;
        v[3] += zd * td;
    };
    if (row1 < R.r1) {
        T t1 = matRow<FMT, DIFF>(kv, kv.sval, row1, sAt);
        if (halt) return;  // block-uniform: no thread of the block reaches writePartials' barrier
        dots(row1, t1);
    }
#pragma unroll 2
    for (int row = row1 + BS; row < R.r1; row += BS) dots(row, matRow<FMT, DIFF>(kv, kv.sval, row, sAt));
    writePartials<4>(v, part_out, kv.G, lb);
}

template <typename T>
__global__ void __launch_bounds__(BS) k_iter_x3(KView<T> kv, const double* part_in, double* part_out) {
    KrylovScalars* ks = kv.ks;
    const int lb = logicalBlock(kv.G);
    RowRange R(kv.n, kv.slices_per_block, lb);
    const int pc = kv.precond;
    T* __restrict__ xv = kv.x;
    T* __restrict__ rv = kv.r;
    T* __restrict__ pnext = kv.p_old;
    const T* __restrict__ pn = kv.p_new;
    const T* __restrict__ vn = kv.v_new;
    const T* __restrict__ sv = kv.s;
    const T* __restrict__ tv = kv.t;
    const T* __restrict__ rh = kv.rhat;
    const T* __restrict__ dinv = kv.dinv;
    // the first row's operands are loaded while the partials arrive (not after an event)
    const int halt = ks->done | ks->flag;
    const int row1 = R.r0 + threadIdx.x;
    T p1 = T(0), s1 = T(0), t1 = T(0), v1 = T(0), x1 = T(0), h1 = T(0), d1 = T(1);
    double red[4];
    if (prologue<4>(ks, part_in, kv.G, red, [&] {
            if (!halt && row1 < R.r1) {
                p1 = pn[row1];
                s1 = sv[row1];
                t1 = tv[row1];
                v1 = vn[row1];
                x1 = xv[row1];
                h1 = rh[row1];
                if (pc) d1 = dinv[row1];
            }
        }))
        return;
    const double omega = red[0] > 0.0 ? red[1] / red[0] : 0.0;
    const double rho1 = red[2] - omega * red[3];
    const double alpha = ks->alpha;
    const double beta = (rho1 / ks->rho[kv.par]) * (alpha / omega);
    const T a = (T)alpha, w = (T)omega, b = (T)beta;
    double acc[2] = {0.0, 0.0};
    auto update = [&](int row, T pi, T si, T ti, T vi, T xi, T hi, T di) {
        T yi = pc ? di * pi : pi;
        T zi = pc ? di * si : si;
        xv[row] = xi + (a * yi + w * zi);
        T ri = si - w * ti;
        rv[row] = ri;
        T qi = ri + b * (pi - w * vi);
        pnext[row] = qi;
        double rd = (double)ri;
        acc[0] += (double)hi * (double)qi;
        acc[1] += rd * rd;
    };
    if (row1 < R.r1) update(row1, p1, s1, t1, v1, x1, h1, d1);
#pragma unroll 2
    for (int row = row1 + BS; row < R.r1; row += BS)
        update(row, pn[row], sv[row], tv[row], vn[row], xv[row], rh[row], pc ? dinv[row] : T(1));
    writePartials<2>(acc, part_out, kv.G, lb);
    if (lb == 0 && threadIdx.x == 0) {
        ks->omega = omega;
        ks->tt = red[0];
        ks->ts = red[1];
        ks->rho[kv.par ^ 1] = rho1;
        ks->beta = beta;
        ks->iter += 1;
        ks->nit += 1;
    }
}

// ------------------------------------------------ packed three-kernel iteration ----
// The same three kernels with W consecutive rows per thread (W = 4 in FP32, 2 in FP64): one
// 16-byte load or store per vector and pack, a quarter (half) of the memory instructions of
// the row-per-thread form.  A thread's first NPF packs (eight rows) are loaded and, in V, their
// products formed before the previous kernel's partial sums are reduced, so the reduction's
// latency overlaps the block's streaming instead of preceding it.  DIA SpMVs take this form
// when every offset is 0, +-1 or a multiple of W and n is a multiple of W (box meshes with
// nx a multiple of 4): a pack at row i then gathers whole aligned packs at i + o, and the
// offsets +-1 shift the own pack by one row (the row outside it is a scalar load).  Products
// are formed and summed in the same order as diaRow / diaRowSym.
template <typename T> struct VecOf;
template <> struct VecOf<float> { static constexpr int W = 4; using V = float4; };
template <> struct VecOf<double> { static constexpr int W = 2; using V = double2; };
template <typename T> struct Pk { T e[VecOf<T>::W]; };

template <typename T>
__device__ inline Pk<T> ldp(const T* __restrict__ a, int i) {  // i a multiple of W
    using V = typename VecOf<T>::V;
    V x = *reinterpret_cast<const V*>(a + i);
    Pk<T> r;
    __builtin_memcpy(&r, &x, sizeof(V));
    return r;
}
template <typename T>
__device__ inline void stp(T* __restrict__ a, int i, const Pk<T>& p) {
    using V = typename VecOf<T>::V;
    V x;
    __builtin_memcpy(&x, &p, sizeof(V));
    *reinterpret_cast<V*>(a + i) = x;
}
template <typename T>
__device__ inline Pk<T> zeroPk() {
    Pk<T> z;
#pragma unroll
    for (int w = 0; w < VecOf<T>::W; ++w) z.e[w] = T(0);
    return z;
}

// Rows i .. i+W-1 of A x (DIA, symmetric storage when SYM); x0 is the pack of x at i.  Each
// diagonal's product is added as soon as its two packs are loaded (the order of diaRow), which
// keeps few registers live; the compiler still issues the loads ahead of their use.
template <typename T, bool SYM, int DIFF, typename KV>
__device__ inline Pk<T> diaPack(const KV& kv, const T* __restrict__ vals, const T* __restrict__ x, int i,
                                const Pk<T>& x0) {
    constexpr int W = VecOf<T>::W;
    const int n = kv.n, D = kv.ndia, mid = D >> 1;
    Pk<T> y = zeroPk<T>();
#pragma unroll
    for (int d = 0; d < DIA_MAX; ++d) {
        if (d >= D) break;
        const int o = kv.doff[d];
        if (DIFF == 2 && o == 0) continue;  // zero row sums: no diagonal term, its values unread
        Pk<T> a, xv;
        if (SYM && d < mid) {  // lower diagonal: the upper entry of row i + w + o
            const T* up = vals + (int64_t)(D - 1 - d) * n;
            if (o == -1) {
                Pk<T> u = ldp(up, i);
                a.e[0] = i > 0 ? up[i - 1] : T(0);
#pragma unroll
                for (int w = 1; w < W; ++w) a.e[w] = u.e[w - 1];
            } else {
                a = i + o >= 0 ? ldp(up, i + o) : zeroPk<T>();
            }
        } else {
            a = ldp(vals + (int64_t)d * n, i);
        }
        if (o == 0) {
            xv = x0;
        } else if (o == 1) {
#pragma unroll
            for (int w = 0; w < W - 1; ++w) xv.e[w] = x0.e[w + 1];
            xv.e[W - 1] = i + W < n ? x[i + W] : T(0);
        } else if (o == -1) {
            xv.e[0] = i > 0 ? x[i - 1] : T(0);
#pragma unroll
            for (int w = 1; w < W; ++w) xv.e[w] = x0.e[w - 1];
        } else {
            const int j = i + o;
            xv = (j >= 0 && j < n) ? ldp(x, j) : zeroPk<T>();
        }
        if (DIFF && o != 0) {  // difference form (see matRow)
#pragma unroll
            for (int w = 0; w < W; ++w) y.e[w] += a.e[w] * (xv.e[w] - x0.e[w]);
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) y.e[w] += a.e[w] * xv.e[w];
        }
    }
    return y;
}

// 16-byte load at any 4-byte-aligned row (one dwordx4 load; the rows i .. i+W-1).
template <typename T>
__device__ inline Pk<T> ldu(const T* __restrict__ a, int i) {
    using V = typename VecOf<T>::V;
    V x;
    __builtin_memcpy(&x, a + i, sizeof(V));
    Pk<T> r;
    __builtin_memcpy(&r, &x, sizeof(V));
    return r;
}

// Rows i .. i+W-1 of A x for the seven-point box pattern, offsets -P, -NX, -1, 0, 1, NX, P
// (P = NX NY, NX and P multiples of W): the same products, in the same order, as diaPack,
// without its per-diagonal branches.  diaPack's offset tests and guarded loads compile to a
// branch and a wait on memory around every load, so a wave has one load in flight at a
// time; here every load is unconditional (addresses clamped into the vector, the
// coefficients of missing neighbours are zero, and the few that must read as zero are
// selected) and the +-1 neighbours are single unaligned 16-byte loads at i -+ 1.
template <typename T, bool SYM, int DIFF, typename KV>
__device__ inline Pk<T> stencilPack(const KV& kv, const T* __restrict__ vals, const T* __restrict__ x, int i,
                                    const Pk<T>& x0) {
    constexpr int W = VecOf<T>::W;
    const int n = kv.n, NX = kv.doff[5], P = kv.doff[6];
    const bool first = i == 0, last = i + W >= n;
    // neighbours of x (a clamped address only where the coefficient is zero)
    const Pk<T> xmP = ldp(x, i >= P ? i - P : 0), xmN = ldp(x, i >= NX ? i - NX : 0);
    Pk<T> xm1 = ldu(x, first ? 0 : i - 1), xp1 = ldu(x, last ? i : i + 1);
    const Pk<T> xpN = ldp(x, i + NX < n ? i + NX : i), xpP = ldp(x, i + P < n ? i + P : i);
    if (first) {  // row -1 does not exist: its coefficient is zero
        xm1.e[0] = T(0);
#pragma unroll
        for (int w = 1; w < W; ++w) xm1.e[w] = x0.e[w - 1];
    }
    if (last) {
#pragma unroll
        for (int w = 0; w < W - 1; ++w) xp1.e[w] = x0.e[w + 1];
        xp1.e[W - 1] = T(0);
    }
    // coefficients, diagonal-major: d = 0 .. 6 for the offsets above
    Pk<T> a[7];
    if (SYM) {  // lower entries are the mirrored upper ones
        const T* u1 = vals + (int64_t)4 * n;
        const T* uN = vals + (int64_t)5 * n;
        const T* uP = vals + (int64_t)6 * n;
        a[4] = ldp(u1, i);
        a[5] = ldp(uN, i);
        a[6] = ldp(uP, i);
        a[0] = ldp(uP, i >= P ? i - P : 0);
        a[1] = ldp(uN, i >= NX ? i - NX : 0);
        a[2] = ldu(u1, first ? 0 : i - 1);
        if (i < P) a[0] = zeroPk<T>();
        if (i < NX) a[1] = zeroPk<T>();
        if (first) {
            a[2].e[0] = T(0);
#pragma unroll
            for (int w = 1; w < W; ++w) a[2].e[w] = a[4].e[w - 1];
        }
    } else {
#pragma unroll
        for (int d = 0; d < 7; ++d)
            if (d != 3) a[d] = ldp(vals + (int64_t)d * n, i);
    }
    if (DIFF != 2) a[3] = ldp(vals + (int64_t)3 * n, i);
    const Pk<T>* xs[7] = {&xmP, &xmN, &xm1, &x0, &xp1, &xpN, &xpP};
    Pk<T> y = zeroPk<T>();
#pragma unroll
    for (int d = 0; d < 7; ++d) {
        if (DIFF == 2 && d == 3) continue;
        const Pk<T>& xv = *xs[d];
        if (DIFF && d != 3) {
#pragma unroll
            for (int w = 0; w < W; ++w) y.e[w] += a[d].e[w] * (xv.e[w] - x0.e[w]);
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w) y.e[w] += a[d].e[w] * xv.e[w];
        }
    }
    return y;
}

// The packed SpMV row product: the seven-point form when ST, else diaPack.
template <typename T, bool SYM, int DIFF, bool ST, typename KV>
__device__ inline Pk<T> packProduct(const KV& kv, const T* __restrict__ vals, const T* __restrict__ x, int i,
                                    const Pk<T>& x0) {
    if constexpr (ST) return stencilPack<T, SYM, DIFF>(kv, vals, x, i, x0);
    else return diaPack<T, SYM, DIFF>(kv, vals, x, i, x0);
}

template <typename T, bool SYM, int DIFF, bool ST>
__global__ void __launch_bounds__(BS, 4) k_iter_v3p(KView<T> kv, const double* part_in) {
    constexpr int W = VecOf<T>::W, NPF = 2;  // packs per thread loaded before the reduction
    KrylovScalars* ks = kv.ks;
    const int lb = logicalBlock(kv.Gv);
    RowRange R(kv.n, kv.spb_v, lb);
    const T* __restrict__ pn = kv.p_new;
    const T* __restrict__ r = kv.r;
    const int halt = ks->done | ks->flag;
    const int i0 = R.r0 + threadIdx.x * W;
    Pk<T> vv[NPF], rv[NPF];
    double red[2];
    if (prologue<2>(ks, part_in, kv.G, red, [&] {
            if (!halt) {
#pragma unroll
                for (int k = 0; k < NPF; ++k) {
                    const int i = i0 + k * BS * W;
                    if (i < R.r1) {
                        vv[k] = packProduct<T, SYM, DIFF, ST>(kv, kv.sval, pn, i, ldp(pn, i));
                        rv[k] = ldp(r, i);
                    }
                }
            }
        }))
        return;
    const double rp = red[0], rr = red[1], rho = ks->rho[kv.par];
    int done = 0, flag = 0;
    if (!(rr == rr)) done = 3;
    else if (kv.mixed) {
        if (rr <= ks->tol2 || rr <= ks->delta2 * ks->rr_ref || ks->iter >= ks->maxit) flag = 1;
    } else if (rr <= ks->tol2) done = 1;
    else if (ks->iter >= ks->maxit) done = 2;
    if (!done && !flag) {
        if (!(rp == rp) || !(rho == rho)) done = 3;
        else if (fabs(rho) < ks->breakdown * ks->r0n) flag = 2;
    }
    if (done || flag) {
        if (lb == 0 && threadIdx.x == 0) {
            ks->rr = rr;
            if (done) ks->done = done;
            else ks->flag = flag;
        }
        return;
    }
    const double alpha = rho / rp;
    const T a = (T)alpha;
    T* __restrict__ vn = kv.v_new;
    T* __restrict__ sv = kv.s;
    auto out = [&](int i, const Pk<T>& v, const Pk<T>& ri) {
        Pk<T> s;
#pragma unroll
        for (int w = 0; w < W; ++w) s.e[w] = ri.e[w] - a * v.e[w];
        stp(vn, i, v);
        stp(sv, i, s);
    };
#pragma unroll
    for (int k = 0; k < NPF; ++k) {
        const int i = i0 + k * BS * W;
        if (i < R.r1) out(i, vv[k], rv[k]);
    }
    for (int i = i0 + NPF * BS * W; i < R.r1; i += BS * W) out(i, packProduct<T, SYM, DIFF, ST>(kv, kv.sval, pn, i, ldp(pn, i)), ldp(r, i));
    if (lb == 0 && threadIdx.x == 0) {
        ks->alpha = alpha;
        ks->r0v = rp;
        ks->rr = rr;
    }
}

template <typename T, bool SYM, int DIFF, bool ST>
__global__ void __launch_bounds__(BS, 4) k_iter_t3p(KView<T> kv, double* part_out) {
    constexpr int W = VecOf<T>::W, NPF = 2;
    const int halt = kv.ks->done | kv.ks->flag;  // the same value in every thread (see k_iter_t3)
    const int lb = logicalBlock(kv.G);
    const T* __restrict__ sv = kv.s;
    const T* __restrict__ r0 = kv.r0;
    T* __restrict__ tv = kv.t;
    RowRange R(kv.n, kv.slices_per_block, lb);
    const int i0 = R.r0 + threadIdx.x * W;
    Pk<T> tk[NPF], sk[NPF], zk[NPF];
#pragma unroll
    for (int k = 0; k < NPF; ++k) {
        const int i = i0 + k * BS * W;
        if (i < R.r1) {
            sk[k] = ldp(sv, i);
            zk[k] = ldp(r0, i);
            tk[k] = packProduct<T, SYM, DIFF, ST>(kv, kv.sval, sv, i, sk[k]);
        }
    }
    if (halt) return;  // block-uniform
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    auto dots = [&](int i, const Pk<T>& t, const Pk<T>& s, const Pk<T>& z) {
        stp(tv, i, t);
#pragma unroll
        for (int w = 0; w < W; ++w) {
            double td = (double)t.e[w], sd = (double)s.e[w], zd = (double)z.e[w];
            v[0] += td * td;
            v[1] += td * sd;
            v[2] += zd * sd;
            v[3] += zd * td;
        }
    };
#pragma unroll
    for (int k = 0; k < NPF; ++k) {
        const int i = i0 + k * BS * W;
        if (i < R.r1) dots(i, tk[k], sk[k], zk[k]);
    }
    for (int i = i0 + NPF * BS * W; i < R.r1; i += BS * W) {
        Pk<T> s = ldp(sv, i);
        dots(i, packProduct<T, SYM, DIFF, ST>(kv, kv.sval, sv, i, s), s, ldp(r0, i));
    }
    writePartials<4>(v, part_out, kv.G, lb);
}

// X with packed rows: for any matrix format (no SpMV), n a multiple of W.
template <typename T>
__global__ void __launch_bounds__(BS, 4) k_iter_x3p(KView<T> kv, const double* part_in, double* part_out) {
    constexpr int W = VecOf<T>::W, NPF = 2;  // six vectors per pack: two packs keep 128 VGPRs
    KrylovScalars* ks = kv.ks;
    const int lb = logicalBlock(kv.G);
    RowRange R(kv.n, kv.slices_per_block, lb);
    const int pc = kv.precond;
    T* __restrict__ xv = kv.x;
    T* __restrict__ rv = kv.r;
    T* __restrict__ pnext = kv.p_old;
    const T* __restrict__ pn = kv.p_new;
    const T* __restrict__ vn = kv.v_new;
    const T* __restrict__ sv = kv.s;
    const T* __restrict__ tv = kv.t;
    const T* __restrict__ rh = kv.rhat;
    const T* __restrict__ dinv = kv.dinv;
    const int halt = ks->done | ks->flag;
    const int i0 = R.r0 + threadIdx.x * W;
    Pk<T> P[NPF], S[NPF], Tt[NPF], V[NPF], X[NPF], H[NPF];
    double red[4];
    if (prologue<4>(ks, part_in, kv.G, red, [&] {
            if (!halt) {
#pragma unroll
                for (int k = 0; k < NPF; ++k) {
                    const int i = i0 + k * BS * W;
                    if (i < R.r1) {
                        P[k] = ldp(pn, i);
                        S[k] = ldp(sv, i);
                        Tt[k] = ldp(tv, i);
                        V[k] = ldp(vn, i);
                        X[k] = ldp(xv, i);
                        H[k] = ldp(rh, i);
                    }
                }
            }
        }))
        return;
    const double omega = red[0] > 0.0 ? red[1] / red[0] : 0.0;
    const double rho1 = red[2] - omega * red[3];
    const double alpha = ks->alpha;
    const double beta = (rho1 / ks->rho[kv.par]) * (alpha / omega);
    const T a = (T)alpha, w = (T)omega, b = (T)beta;
    double acc[2] = {0.0, 0.0};
    auto update = [&](int i, const Pk<T>& p, const Pk<T>& s, const Pk<T>& t, const Pk<T>& v, const Pk<T>& x,
                      const Pk<T>& h) {
        Pk<T> xo, ro, qo;
        const Pk<T> dv = pc ? ldp(dinv, i) : Pk<T>{};
#pragma unroll
        for (int e = 0; e < W; ++e) {
            T yi = pc ? dv.e[e] * p.e[e] : p.e[e];
            T zi = pc ? dv.e[e] * s.e[e] : s.e[e];
            xo.e[e] = x.e[e] + (a * yi + w * zi);
            T ri = s.e[e] - w * t.e[e];
            ro.e[e] = ri;
            T qi = ri + b * (p.e[e] - w * v.e[e]);
            qo.e[e] = qi;
            double rd = (double)ri;
            acc[0] += (double)h.e[e] * (double)qi;
            acc[1] += rd * rd;
        }
        stp(xv, i, xo);
        stp(rv, i, ro);
        stp(pnext, i, qo);
    };
#pragma unroll
    for (int k = 0; k < NPF; ++k) {
        const int i = i0 + k * BS * W;
        if (i < R.r1) update(i, P[k], S[k], Tt[k], V[k], X[k], H[k]);
    }
    for (int i = i0 + NPF * BS * W; i < R.r1; i += BS * W)
        update(i, ldp(pn, i), ldp(sv, i), ldp(tv, i), ldp(vn, i), ldp(xv, i), ldp(rh, i));
    writePartials<2>(acc, part_out, kv.G, lb);
    if (lb == 0 && threadIdx.x == 0) {
        ks->omega = omega;
        ks->tt = red[0];
        ks->ts = red[1];
        ks->rho[kv.par ^ 1] = rho1;
        ks->beta = beta;
        ks->iter += 1;
        ks->nit += 1;
    }
}

// DIA: column j of Ahat holds, on diagonal d, the entry of row j - off_d, so rhat is a
// gather over the same value array the SpMV reads (the operator's own rounding and scaling).
// With the difference form the diagonal slot holds the row sum s_j, and the operator's own
// diagonal is s_j - sum_{d != 0} a_{j,d} (row j's off-diagonal values as stored).
template <typename T>
__global__ void __launch_bounds__(BS) k_rhat_dia(KView<T> kv, const double* r0) {
    int j = blockIdx.x * BS + threadIdx.x;
    if (j >= kv.n) return;
    double y = 0.0;
#pragma unroll
    for (int d = 0; d < DIA_MAX; ++d) {
        int i = j - kv.doff[d];
        if (d < kv.ndia && (unsigned)i < (unsigned)kv.n) {
            double a = (double)kv.sval[(int64_t)d * kv.n + i];
            if (kv.diff && kv.doff[d] == 0) {
                if (kv.diff == 2) a = 0.0;  // zero row sums
                for (int e = 0; e < kv.ndia; ++e)
                    if (kv.doff[e] != 0) a -= (double)kv.sval[(int64_t)e * kv.n + j];
            }
            y += a * (double)(T)r0[i];
        }
    }
    kv.rhat[j] = (T)y;
}

// rhat_j = sum_i Ahat_ij r0_i over column j (the transposed pattern), accumulated in FP64
// from exactly the operands of the iteration: the SpMV's values Ahat_ij = T(A_ij dinv_j)
// and the iteration's r0 = T(r0_64).  With any other rounding of Ahat, (rhat, p) and
// (r0, Ahat p) part along null-space components of p, which on a singular (deflated) system
// grow without bound in FP32.
// diff: the operator of the difference form (matRow): off-diagonals as stored, and on the
// diagonal the stored row sum T(sum_k A_jk dinv_k) less row j's stored off-diagonals.
template <typename T>
__global__ void k_rhat(int n, const int* cscp, const int* cscr, const int* cscm, const int* rp, const int* col,
                       const double* val, const double* r0, const double* dinv, int diff, T* rhat) {
    int j = blockIdx.x * BS + threadIdx.x;
    if (j >= n) return;
    const double dj = dinv ? dinv[j] : 1.0;
    // the operator's diagonal entry in the difference form (used where i == j)
    auto diagEntry = [&]() {
        double rs = 0.0;
        for (int q = rp[j]; q < rp[j + 1]; ++q) rs += dinv ? val[q] * dinv[col[q]] : val[q];
        double a = diff == 2 ? 0.0 : (double)(T)rs;  // 2: zero row sums
        for (int q = rp[j]; q < rp[j + 1]; ++q)
            if (col[q] != j) a -= (double)(T)(dinv ? val[q] * dinv[col[q]] : val[q]);
        return a;
    };
    const int k0 = cscp[j], k1 = cscp[j + 1];
    double y = 0.0;
    if (k1 > k0 && k1 - k0 <= 8) {
        // columns of 1 to 8 entries: every index, value and r0 load issued before the sums
        // (a slot past the column's end re-reads its last entry and is dropped); same order
        int ri[8], mi[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = k0 + q < k1 ? k0 + q : k1 - 1;
            ri[q] = cscr[k];
            mi[q] = cscm[k];
        }
        double av[8], xv[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            av[q] = (double)(T)(dinv ? val[mi[q]] * dj : val[mi[q]]);
            xv[q] = (double)(T)r0[ri[q]];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (k0 + q >= k1) break;
            const double a = (diff && ri[q] == j) ? diagEntry() : av[q];
            y += a * xv[q];
        }
    } else {
        for (int k = k0; k < k1; ++k) {
            const int i = cscr[k];
            double a = (double)(T)(dinv ? val[cscm[k]] * dj : val[cscm[k]]);
            if (diff && i == j) a = diagEntry();
            y += a * (double)(T)r0[i];
        }
    }
    rhat[j] = (T)y;
}

// Difference form: the row sums sum_k A_ik (dinv_k) in FP64, rounded into the diagonal slots
// of the iteration's value array (map: CSR position -> slot).
template <typename T>
__global__ void k_rowsum_diag(int n, const int* rp, const int* col, const int* diag, const int* map,
                              const double* val, const double* dinv, T* sval) {
    int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    double rs = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) rs += dinv ? val[k] * dinv[col[k]] : val[k];
    sval[map[diag[i]]] = (T)rs;
}

// ------------------------------------------------------------------ start / events ----
// r64 = b - A64 x64 (FP64); partials of (r,r), (b,b).
template <typename T, int FMT>
__global__ void __launch_bounds__(BS) k_residual(KView<T> kv, const double* b, const double* x64, double* r64,
                                                 double* part_out) {
    const int lb = logicalBlock(kv.G);
    double rr = 0.0, bb = 0.0;
    RowRange R(kv.n, kv.slices_per_block, lb);
    for (int row = R.r0 + threadIdx.x; row < R.r1; row += BS) {
        double ax = matRow<FMT, 0>(kv, kv.sval64, row, [&](int j) { return x64[j]; });
        double ri = b[row] - ax;
        r64[row] = ri;
        rr += ri * ri;
        bb += b[row] * b[row];
    }
    double v[2] = {rr, bb};
    writePartials<2>(v, part_out, kv.G, lb);
}

__global__ void k_residual_finish(KrylovScalars* ks, const double* part, int G, int start) {
    double red[2];
    reducePartials<2>(part, G, red);
    if (threadIdx.x == 0) {
        ks->rr_true = red[0];
        if (start) {
            ks->bb = red[1];
            ks->tol2 = ks->tol * ks->tol * red[1];
        }
    }
}

// r = T(r64); start 1: fresh (r0 = r, p = v = 0); start 2: restart (r0 = r); start 0:
// residual replacement (r0 kept).  Mixed precision: x32 = 0.  Partials: (r,r), (r0,r), (r0,r0).
template <typename T>
__global__ void __launch_bounds__(BS) k_seed(KView<T> kv, const double* r64, int start, double* part_out) {
    const int lb = logicalBlock(kv.G);
    double v[3] = {0.0, 0.0, 0.0};
    RowRange R(kv.n, kv.slices_per_block, lb);
    for (int row = R.r0 + threadIdx.x; row < R.r1; row += BS) {
        T ri = (T)r64[row];
        kv.r[row] = ri;
        T r0i;
        if (start) {
            kv.r0[row] = ri;
            if (start == 1) {
                kv.p_old[row] = T(0);
                kv.v_old[row] = T(0);
            }
            r0i = ri;
        } else {
            r0i = kv.r0[row];
        }
        if (kv.mixed) kv.x[row] = T(0);
        double rd = (double)ri, r0d = (double)r0i;
        v[0] += rd * rd;
        v[1] += r0d * rd;
        v[2] += r0d * r0d;
    }
    writePartials<3>(v, part_out, kv.G, lb);
}

// Clears the event state before a seed: done (clear_done) and the iteration counter (clear_iter).
__global__ void k_reset_events(KrylovScalars* ks, int clear_done, int clear_iter) {
    if (threadIdx.x == 0) {
        if (clear_done) ks->done = 0;
        if (clear_iter) ks->iter = 0;
    }
}

__global__ void k_seed_finish(KrylovScalars* ks, const double* part, int G, int start) {
    double red[3];
    reducePartials<3>(part, G, red);
    if (threadIdx.x == 0) {
        ks->rr = red[0];
        if (start) ks->r0n = red[2];
        if (start == 1) {
            ks->rho[0] = ks->rho[1] = 1.0;
            ks->alpha = 1.0;
            ks->omega = 1.0;
        }
        ks->rr_ref = ks->rr_true;
        ks->flag = 0;
        if (ks->bb == 0.0) ks->done = 4;
        else if (ks->rr_true <= ks->tol2) ks->done = 1;
    }
}

// CSR values -> sliced-ELL values (padding slots stay zero)
template <typename T>
__global__ void k_to_sell(int64_t nnz, const int* csr2sell, const double* val, T* sval) {
    int64_t k = (int64_t)blockIdx.x * BS + threadIdx.x;
    if (k < nnz) sval[csr2sell[k]] = (T)val[k];
}

// CSR values -> column-scaled sliced-ELL values A M^-1 (right Jacobi preconditioning)
template <typename T>
__global__ void k_to_sell_scaled(int64_t nnz, const int* csr2sell, const int* col, const double* val,
                                 const double* dinv, T* sval) {
    int64_t k = (int64_t)blockIdx.x * BS + threadIdx.x;
    if (k < nnz) sval[csr2sell[k]] = (T)(val[k] * dinv[col[k]]);
}

__global__ void k_flush(int n, double* x64, float* x32) {
    int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) {
        x64[i] += (double)x32[i];
        x32[i] = 0.0f;
    }
}

// Scaled residual of solver.cpp:8-24: max|A x - b| and max|diag * x| (partials, then finish).
__global__ void __launch_bounds__(BS) k_scaled(int n, int G, int cpb, const int* rp, const int* col,
                                               const double* val, const int* diag, const double* x, const double* b,
                                               double* part) {
    const int lb = logicalBlock(G);
    double* lds = LdsBuf<double>::get();
    double r = 0.0, d = 0.0;
    Rows R(n, cpb, lb);
    for (int c = R.c0; c < R.c1; ++c) {
        int row0 = c * BS;
        double ax = spmvChunk<double>(rp, col, val, n, row0, lds, [&](int j) { return x[j]; });
        int row = row0 + threadIdx.x;
        if (row < n) {
            r = nanMax(r, fabs(ax - b[row]));
            d = nanMax(d, fabs(val[diag[row]] * x[row]));
        }
    }
    r = waveMax(r);
    d = waveMax(d);
    __shared__ double sh[2][NW];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][w] = r;
        sh[1][w] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, c = 0;
        for (int i = 0; i < NW; ++i) {
            a = nanMax(a, sh[0][i]);
            c = nanMax(c, sh[1][i]);
        }
        part[lb] = a;
        part[G + lb] = c;
    }
}
__global__ void k_scaled_finish(KrylovScalars* ks, const double* part, int G) {
    if (threadIdx.x < 64) {
        double a = 0, c = 0;
        for (int i = threadIdx.x; i < G; i += 64) {
            a = nanMax(a, part[i]);
            c = nanMax(c, part[G + i]);
        }
        a = waveMax(a);
        c = waveMax(c);
        if (threadIdx.x == 0) {
            ks->scaled[0] = a;
            ks->scaled[1] = c;
        }
    }
}

__global__ void __launch_bounds__(BS) k_spmv64(int n, int G, int cpb, const int* rp, const int* col,
                                               const double* val, const double* x, double* y) {
    const int lb = logicalBlock(G);
    double* lds = LdsBuf<double>::get();
    Rows R(n, cpb, lb);
    for (int c = R.c0; c < R.c1; ++c) {
        int row0 = c * BS;
        double s = spmvChunk<double>(rp, col, val, n, row0, lds, [&](int j) { return x[j]; });
        int row = row0 + threadIdx.x;
        if (row < n) y[row] = s;
    }
}

// Jacobi sweep (jacobi.cpp:5-9): y = x + D^-1 (b - A x), row by row.
__global__ void k_jacobi_step(int n, const int* rp, const int* col, const double* val, const int* diag,
                              const double* b, const double* x, double* y) {
    int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    double ax = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) ax += val[k] * x[col[k]];
    double dinv = 1.0 / val[diag[i]];
    y[i] = x[i] + dinv * (b[i] - ax);
}

// Structure test of the reference-cell system (CSR, FP64): per block, max |row sum|,
// max |diagonal| and max |A_ij - A_ji| of A with the reference diagonal reduced to `a`.
__global__ void __launch_bounds__(BS) k_ref_check(int n, const int* rp, const int* col, const double* val,
                                                  const int* diag, int ref, double a, double* part) {
    int i = blockIdx.x * BS + threadIdx.x;
    double rs = 0.0, dg = 0.0, as = 0.0;
    if (i < n) {
        double sum = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            int j = col[k];
            double v = (k == diag[i] && i == ref) ? a : val[k];
            sum += v;
            if (j == i) continue;
            int lo = rp[j], hi = rp[j + 1];  // A_ji by bisection in the sorted row j
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (col[mid] < i) lo = mid + 1;
                else hi = mid;
            }
            double t = (lo < rp[j + 1] && col[lo] == i) ? val[lo] : 0.0;
            as = nanMax(as, fabs(v - t));
        }
        rs = fabs(sum);
        dg = fabs(i == ref ? a : val[diag[i]]);
    }
    __shared__ double sh[3][NW];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    rs = waveMax(rs);
    dg = waveMax(dg);
    as = waveMax(as);
    if (lane == 0) {
        sh[0][w] = rs;
        sh[1][w] = dg;
        sh[2][w] = as;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a0 = 0, a1 = 0, a2 = 0;
        for (int k = 0; k < NW; ++k) {
            a0 = nanMax(a0, sh[0][k]);
            a1 = nanMax(a1, sh[1][k]);
            a2 = nanMax(a2, sh[2][k]);
        }
        part[3 * blockIdx.x] = a0;
        part[3 * blockIdx.x + 1] = a1;
        part[3 * blockIdx.x + 2] = a2;
    }
}

// Per-block sums of x (the host adds the block sums in order).
__global__ void __launch_bounds__(BS) k_block_sum(int n, const double* x, double* part) {
    int i = blockIdx.x * BS + threadIdx.x;
    double v = i < n ? x[i] : 0.0;
    __shared__ double sh[NW];
    v = waveSum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < NW; ++k) t += sh[k];
        part[blockIdx.x] = t;
    }
}

__global__ void k_add_const(int n, double* x, double c) {
    int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) x[i] += c;
}

template <typename T>
__global__ void k_convert(int64_t n, const double* src, T* dst) {
    int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
    if (i < n) dst[i] = (T)src[i];
}
template <typename T>
__global__ void k_dinv(int n, const int* diag, const double* val, T* dinv) {
    int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) dinv[i] = (T)(1.0 / val[diag[i]]);
}

inline int nb(int64_t n) { return (int)((n + BS - 1) / BS); }

}  // namespace

// Device scratch of the allocation-free solve entry point (prism_eqn_step_noalloc): one
// static array per device, which the code object's loader places when the library is loaded
// on that device and which lives until the process ends.  Host code never uses the array's
// name as an address: scratch_address() obtains the device's address with
// hipGetSymbolAddress, and callers pass that pointer to the kernels.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

void* scratch_address(int device) {
    int cur = -1;
    HIPCHECK(hipGetDevice(&cur));
    if (cur != device) throw Error(PRISM_ERR_RUNTIME, "scratch_address: the context's device is not current");
    void* ptr = nullptr;
    HIPCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(PRISM_ERR_RUNTIME, "scratch_address: no device address for the scratch array");
    return ptr;
}

#define KLAUNCH(label, kern, grid, threads, ...) PRISM_LAUNCH(label, kern, grid, threads, __VA_ARGS__)

// Blocks of the iteration kernels: one per MIN_SPB 64-row slices (one row per thread) up to
// GDEF (256 in the band below); each block owns a contiguous run of slices.  Every block of
// a consuming kernel reads all G partial sums, a cost that grows with G (measured: 1.8, 2.8
// and 6.1 us per kernel for four partial arrays at G = 256, 512, 1024), while more than 512
// blocks of 512 threads do not raise the SpMV bandwidth on 256 CUs.
// On sliced-ELL matrices the cap is 768: on the 96^3 x 6 tetrahedral pressure system
// (82,944 slices) 768 blocks cut the FP32 iteration by 6 % against 512 (136.2 against 145.2
// us; 1024: 138.9) and leave the FP64 one unchanged (231.0 against 230.7 us; 1024: 236.8).
// *spb receives the slices per block.
int krylov_grid(int64_t n, int* spb, bool band, bool sell) {
    static const char* knob = std::getenv("PRISM_KRYLOV_GMAX");  // tuning knob: the cap on G
    static const bool knobbed = knob && std::atoi(knob) >= 8;
    const int gmax = std::min(knobbed ? std::atoi(knob) : sell ? GSELL : GDEF, GMAX);
    constexpr int64_t MIN_SPB = BS / 64;
    int64_t slices = (n + 63) / 64;
    int64_t want = (slices + MIN_SPB - 1) / MIN_SPB;
    int G = (int)(want < gmax ? want : gmax);
    // One block per CU (256) when that gives each thread two to four rows: measured at
    // 4096 slices (64^3 cells), 256 blocks beat 512 by 6 % (FP32) and 2 % (FP64) per
    // iteration, while 192 or 320 blocks lose 8-10 % against 256.
    if (band && !knobbed && !sell && slices >= 2 * MIN_SPB * 256 && slices < 2 * MIN_SPB * GDEF) G = 256;
    if (G < 1) G = 1;
    int per = (int)((slices + G - 1) / G);
    G = (int)((slices + per - 1) / per);
    *spb = per;
    return G;
}

static int csrGrid(int64_t n, int* cpb) {  // 512-row chunks for the CSR helpers
    int chunks = nb(n);
    int G = chunks < GMAX ? chunks : GMAX;
    int per = (chunks + G - 1) / G;
    G = (chunks + per - 1) / per;
    *cpb = per;
    return G;
}

void launch_spmv64(const Launch& L, int n, const int* rp, const int* col, const double* val, const double* x,
                   double* y) {
    int cpb;
    int G = csrGrid(n, &cpb);
    KLAUNCH("spmv_f64", k_spmv64, G, BS, n, G, cpb, rp, col, val, x, y);
}

void launch_jacobi_step(const Launch& L, int n, const int* rp, const int* col, const double* val, const int* diag,
                        const double* b, const double* x, double* y) {
    KLAUNCH("jacobi_step", k_jacobi_step, nb(n), BS, n, rp, col, val, diag, b, x, y);
}

int ref_blocks(int n) { return nb(n); }
// A per-block partial writer's buffer must hold `need` doubles; checked before the launch.
void require_partials(const char* who, size_t need, size_t cap) {
    if (need > cap)
        throw Error(PRISM_ERR_RUNTIME, std::string(who) + ": partial-sum buffer of " + std::to_string(cap) +
                                           " doubles for " + std::to_string(need) + " partials");
}
void launch_ref_check(const Launch& L, int n, const int* rp, const int* col, const double* val, const int* diag,
                      int ref, double a, double* part, size_t part_cap) {
    require_partials("reference-cell check", 3 * (size_t)nb(n), part_cap);
    KLAUNCH("ref_cell_check", k_ref_check, nb(n), BS, n, rp, col, val, diag, ref, a, part);
}
void launch_block_sum(const Launch& L, int n, const double* x, double* part, size_t part_cap) {
    require_partials("block sum", (size_t)nb(n), part_cap);
    KLAUNCH("block_sum", k_block_sum, nb(n), BS, n, x, part);
}
void launch_add_const(const Launch& L, int n, double* x, double c) {
    KLAUNCH("add_constant", k_add_const, nb(n), BS, n, x, c);
}

size_t scaled_residual_partials(int64_t n) {
    int cpb;
    return 2 * (size_t)csrGrid(n, &cpb);
}
void launch_scaled_residual(const Launch& L, int n, const int* rp, const int* col, const double* val,
                            const int* diag, const double* x, const double* b, double* part, size_t part_cap,
                            KrylovScalars* ks) {
    int cpb;
    int G = csrGrid(n, &cpb);
    require_partials("scaled residual", 2 * (size_t)G, part_cap);
    KLAUNCH("scaled_residual", k_scaled, G, BS, n, G, cpb, rp, col, val, diag, x, b, part);
    KLAUNCH("scaled_residual_finish", k_scaled_finish, 1, 64, ks, part, G);
}

template <typename KV>
inline int formatOf(const KV& kv) {
    return kv.ndia ? (kv.dsym ? FMT_DIA_SYM : FMT_DIA) : (kv.srel ? FMT_SELL16 : FMT_SELL32);
}

template <typename T>
void launch_residual(const Launch& L, const KView<T>& kv, const double* b, const double* x64, double* r64,
                     int start) {
    switch (formatOf(kv)) {
    case FMT_DIA:
        KLAUNCH("bicgstab_residual_f64", (k_residual<T, FMT_DIA>), kv.G, BS, kv, b, x64, r64, kv.part_out);
        break;
    case FMT_DIA_SYM:
        KLAUNCH("bicgstab_residual_f64", (k_residual<T, FMT_DIA_SYM>), kv.G, BS, kv, b, x64, r64, kv.part_out);
        break;
    case FMT_SELL16:
        KLAUNCH("bicgstab_residual_f64", (k_residual<T, FMT_SELL16>), kv.G, BS, kv, b, x64, r64, kv.part_out);
        break;
    default:
        KLAUNCH("bicgstab_residual_f64", (k_residual<T, FMT_SELL32>), kv.G, BS, kv, b, x64, r64, kv.part_out);
    }
    KLAUNCH("bicgstab_residual_finish", k_residual_finish, 1, 64, kv.ks, kv.part_out, kv.G, start);
}
template <typename T>
void launch_seed(const Launch& L, const KView<T>& kv, const double* r64, int start) {
    KLAUNCH(sizeof(T) == 8 ? "bicgstab_seed_f64" : "bicgstab_seed_f32", k_seed<T>, kv.G, BS, kv, r64, start,
            kv.part_in);
    KLAUNCH("bicgstab_seed_finish", k_seed_finish, 1, 64, kv.ks, kv.part_in, kv.G, start);
}
// part_in: (r,r), (r0,r) of the previous iteration (written back by C); part_out: two
// scratch partial buffers of 3 x G.
template <typename T>
void launch_iteration(const Launch& L, const KView<T>& kv) {
    const bool d = sizeof(T) == 8;
    double* pv = kv.part_out;
    double* pt = kv.part_out + 3 * (size_t)kv.G;
    KLAUNCH(d ? "bicgstab_update_p_f64" : "bicgstab_update_p_f32", (k_iter_p<T, false>), kv.G, BS, kv, kv.part_in,
            nullptr);
    const char* lv = d ? "bicgstab_spmv_v_f64" : "bicgstab_spmv_v_f32";
    const char* lt = d ? "bicgstab_spmv_t_f64" : "bicgstab_spmv_t_f32";
    switch (formatOf(kv)) {
    case FMT_DIA: KLAUNCH(lv, (k_iter_v<T, FMT_DIA>), kv.G, BS, kv, pv); break;
    case FMT_DIA_SYM: KLAUNCH(lv, (k_iter_v<T, FMT_DIA_SYM>), kv.G, BS, kv, pv); break;
    case FMT_SELL16: KLAUNCH(lv, (k_iter_v<T, FMT_SELL16>), kv.G, BS, kv, pv); break;
    default: KLAUNCH(lv, (k_iter_v<T, FMT_SELL32>), kv.G, BS, kv, pv);
    }
    KLAUNCH(d ? "bicgstab_update_s_f64" : "bicgstab_update_s_f32", k_iter_s<T>, kv.G, BS, kv, pv);
    switch (formatOf(kv)) {
    case FMT_DIA: KLAUNCH(lt, (k_iter_t<T, FMT_DIA>), kv.G, BS, kv, pt); break;
    case FMT_DIA_SYM: KLAUNCH(lt, (k_iter_t<T, FMT_DIA_SYM>), kv.G, BS, kv, pt); break;
    case FMT_SELL16: KLAUNCH(lt, (k_iter_t<T, FMT_SELL16>), kv.G, BS, kv, pt); break;
    default: KLAUNCH(lt, (k_iter_t<T, FMT_SELL32>), kv.G, BS, kv, pt);
    }
    KLAUNCH(d ? "bicgstab_update_xr_f64" : "bicgstab_update_xr_f32", k_iter_xr<T>, kv.G, BS, kv, pt, kv.part_in);
}
template <typename T>
void launch_rhat(const Launch& L, const KView<T>& kv, const double* r0_64, const double* dinv64) {
    if (kv.ndia)
        KLAUNCH(sizeof(T) == 8 ? "bicgstab_rhat_f64" : "bicgstab_rhat_f32", k_rhat_dia<T>, nb(kv.n), BS, kv, r0_64);
    else
        KLAUNCH(sizeof(T) == 8 ? "bicgstab_rhat_f64" : "bicgstab_rhat_f32", k_rhat<T>, nb(kv.n), BS, kv.n, kv.cscp,
                kv.cscr, kv.cscm, kv.rp, kv.col, kv.val64, r0_64, dinv64, kv.diff, kv.rhat);
}
// part_in: the seed's (r,r), (r0,r) partials; writes part_x.
template <typename T>
void launch_entry(const Launch& L, const KView<T>& kv) {
    KLAUNCH(sizeof(T) == 8 ? "bicgstab_entry_f64" : "bicgstab_entry_f32", (k_iter_p<T, true>), kv.G, BS, kv,
            kv.part_in, kv.part_x);
}
template <typename T>
void krylov_pack(KView<T>& kv) {
    const char* e = std::getenv("PRISM_KRYLOV_PACK");  // read per solve (tests switch it)
    const int mode = e ? std::atoi(e) : 1;
    constexpr int W = VecOf<T>::W;
    kv.pack = 0;
    // measured on the 128^3 pressure system: packed rows take 7.5 % off an FP32 iteration
    // and add 10 % to an FP64 one, so FP64 keeps the row-per-thread kernels unless forced
    if (mode == 0 || kv.n % W != 0 || (mode == 1 && sizeof(T) == 8)) return;
    auto enough = [&](int spb) { return mode == 2 || (int64_t)spb * 64 * 2 >= (int64_t)BS * W; };
    if (enough(kv.slices_per_block)) kv.pack |= 1;
    if (kv.ndia && enough(kv.slices_per_block) && enough(kv.spb_v)) {
        bool ok = true;
        for (int d = 0; d < kv.ndia; ++d) {
            const int o = kv.doff[d];
            if (!(o == 0 || o == 1 || o == -1 || o % W == 0)) ok = false;
        }
        if (ok) kv.pack |= 2;
    }
    // bit 2: the seven-point box pattern (offsets -P, -NX, -1, 0, 1, NX, P with NX, P multiples
    // of W and NX < P) takes the branch-free product.  PRISM_KRYLOV_STENCIL=0 turns it off.
    const char* se = std::getenv("PRISM_KRYLOV_STENCIL");
    if ((kv.pack & 2) && kv.ndia == 7 && !(se && std::atoi(se) == 0)) {
        const int NX = kv.doff[5], P = kv.doff[6];
        if (kv.doff[0] == -P && kv.doff[1] == -NX && kv.doff[2] == -1 && kv.doff[3] == 0 && kv.doff[4] == 1 &&
            NX > 1 && P > NX && NX % W == 0 && P % W == 0)
            kv.pack |= 4;
    }
}

// V and T of the three-kernel iteration in the view's format (packed or row per thread,
// plain or difference-form products).
template <typename T, int DIFF>
void launch_vt3(const Launch& L, const KView<T>& kv, const char* lv, const char* lt) {
    if (kv.pack & 2) {
        const bool st = (kv.pack & 4) != 0;
        if (kv.dsym) {
            if (st) {
                KLAUNCH(lv, (k_iter_v3p<T, true, DIFF, true>), kv.Gv, BS, kv, kv.part_x);
                KLAUNCH(lt, (k_iter_t3p<T, true, DIFF, true>), kv.G, BS, kv, kv.part_out);
            } else {
                KLAUNCH(lv, (k_iter_v3p<T, true, DIFF, false>), kv.Gv, BS, kv, kv.part_x);
                KLAUNCH(lt, (k_iter_t3p<T, true, DIFF, false>), kv.G, BS, kv, kv.part_out);
            }
        } else if (st) {
            KLAUNCH(lv, (k_iter_v3p<T, false, DIFF, true>), kv.Gv, BS, kv, kv.part_x);
            KLAUNCH(lt, (k_iter_t3p<T, false, DIFF, true>), kv.G, BS, kv, kv.part_out);
        } else {
            KLAUNCH(lv, (k_iter_v3p<T, false, DIFF, false>), kv.Gv, BS, kv, kv.part_x);
            KLAUNCH(lt, (k_iter_t3p<T, false, DIFF, false>), kv.G, BS, kv, kv.part_out);
        }
        return;
    }
    switch (formatOf(kv)) {
    case FMT_DIA:
        KLAUNCH(lv, (k_iter_v3<T, FMT_DIA, DIFF>), kv.Gv, BS, kv, kv.part_x);
        KLAUNCH(lt, (k_iter_t3<T, FMT_DIA, DIFF>), kv.G, BS, kv, kv.part_out);
        break;
    case FMT_DIA_SYM:
        KLAUNCH(lv, (k_iter_v3<T, FMT_DIA_SYM, DIFF>), kv.Gv, BS, kv, kv.part_x);
        KLAUNCH(lt, (k_iter_t3<T, FMT_DIA_SYM, DIFF>), kv.G, BS, kv, kv.part_out);
        break;
    case FMT_SELL16:
        KLAUNCH(lv, (k_iter_v3<T, FMT_SELL16, DIFF>), kv.Gv, BS, kv, kv.part_x);
        KLAUNCH(lt, (k_iter_t3<T, FMT_SELL16, DIFF>), kv.G, BS, kv, kv.part_out);
        break;
    default:
        KLAUNCH(lv, (k_iter_v3<T, FMT_SELL32, DIFF>), kv.Gv, BS, kv, kv.part_x);
        KLAUNCH(lt, (k_iter_t3<T, FMT_SELL32, DIFF>), kv.G, BS, kv, kv.part_out);
    }
}

// part_x: (rhat,p), (r,r) -> V; part_out (4 x G): T's partials -> X -> part_x.
template <typename T>
void launch_iteration3(const Launch& L, const KView<T>& kv) {
    const bool d = sizeof(T) == 8;
    const char* lv = d ? "bicgstab_spmv_v_f64" : "bicgstab_spmv_v_f32";
    const char* lt = d ? "bicgstab_spmv_t_f64" : "bicgstab_spmv_t_f32";
    const char* lx = d ? "bicgstab_update_xrp_f64" : "bicgstab_update_xrp_f32";
    if constexpr (sizeof(T) == 4) {
        if (kv.diff == 2) launch_vt3<T, 2>(L, kv, lv, lt);
        else if (kv.diff) launch_vt3<T, 1>(L, kv, lv, lt);
        else launch_vt3<T, 0>(L, kv, lv, lt);
    } else {
        launch_vt3<T, 0>(L, kv, lv, lt);
    }
    if (kv.pack & 1) KLAUNCH(lx, k_iter_x3p<T>, kv.G, BS, kv, kv.part_out, kv.part_x);
    else KLAUNCH(lx, k_iter_x3<T>, kv.G, BS, kv, kv.part_out, kv.part_x);
}
void launch_dia_symcheck(const Launch& L, const KView<double>& kv, const double* vals, double* part,
                         size_t part_cap) {
    require_partials("DIA symmetry check", (size_t)nb(kv.n), part_cap);
    KLAUNCH("dia_symcheck", k_dia_symcheck, nb(kv.n), BS, kv, vals, part);
}
void launch_reset_events(const Launch& L, KrylovScalars* ks, bool clear_done, bool clear_iter) {
    KLAUNCH("bicgstab_reset_events", k_reset_events, 1, 64, ks, clear_done ? 1 : 0, clear_iter ? 1 : 0);
}
void launch_flush(const Launch& L, int n, double* x64, float* x32) {
    KLAUNCH("bicgstab_flush", k_flush, nb(n), BS, n, x64, x32);
}
template <typename T>
void launch_convert(const Launch& L, int64_t n, const double* src, T* dst) {
    KLAUNCH("convert_values", k_convert<T>, nb(n), BS, n, src, dst);
}
template <typename T>
void launch_dinv(const Launch& L, int n, const int* diag, const double* val, T* dinv) {
    KLAUNCH("jacobi_dinv", k_dinv<T>, nb(n), BS, n, diag, val, dinv);
}
template <typename T>
void launch_to_sell_scaled(const Launch& L, int64_t nnz, const int* csr2sell, const int* col, const double* val,
                           const double* dinv, T* sval) {
    KLAUNCH(sizeof(T) == 8 ? "to_sell_scaled_f64" : "to_sell_scaled_f32", k_to_sell_scaled<T>, nb(nnz), BS, nnz,
            csr2sell, col, val, dinv, sval);
}
template void launch_to_sell_scaled<float>(const Launch&, int64_t, const int*, const int*, const double*,
                                           const double*, float*);
template void launch_to_sell_scaled<double>(const Launch&, int64_t, const int*, const int*, const double*,
                                            const double*, double*);
template <typename T>
void launch_rowsum_diag(const Launch& L, int n, const int* rp, const int* col, const int* diag, const int* map,
                        const double* val, const double* dinv, T* sval) {
    KLAUNCH(sizeof(T) == 8 ? "rowsum_diag_f64" : "rowsum_diag_f32", k_rowsum_diag<T>, nb(n), BS, n, rp, col, diag,
            map, val, dinv, sval);
}
template void launch_rowsum_diag<float>(const Launch&, int, const int*, const int*, const int*, const int*,
                                        const double*, const double*, float*);
template <typename T>
void launch_to_sell(const Launch& L, int64_t nnz, const int* csr2sell, const double* val, T* sval) {
    KLAUNCH(sizeof(T) == 8 ? "to_sell_f64" : "to_sell_f32", k_to_sell<T>, nb(nnz), BS, nnz, csr2sell, val, sval);
}

template void launch_residual<double>(const Launch&, const KView<double>&, const double*, const double*, double*, int);
template void launch_residual<float>(const Launch&, const KView<float>&, const double*, const double*, double*, int);
template void launch_seed<double>(const Launch&, const KView<double>&, const double*, int);
template void launch_seed<float>(const Launch&, const KView<float>&, const double*, int);
template void launch_iteration<double>(const Launch&, const KView<double>&);
template void launch_iteration<float>(const Launch&, const KView<float>&);
template void launch_rhat<double>(const Launch&, const KView<double>&, const double*, const double*);
template void launch_rhat<float>(const Launch&, const KView<float>&, const double*, const double*);
template void launch_entry<double>(const Launch&, const KView<double>&);
template void launch_entry<float>(const Launch&, const KView<float>&);
template void launch_iteration3<double>(const Launch&, const KView<double>&);
template void krylov_pack<double>(KView<double>&);
template void krylov_pack<float>(KView<float>&);
template void launch_iteration3<float>(const Launch&, const KView<float>&);
template void launch_convert<float>(const Launch&, int64_t, const double*, float*);
template void launch_convert<double>(const Launch&, int64_t, const double*, double*);
template void launch_dinv<float>(const Launch&, int, const int*, const double*, float*);
template void launch_dinv<double>(const Launch&, int, const int*, const double*, double*);
template void launch_to_sell<float>(const Launch&, int64_t, const int*, const double*, float*);
template void launch_to_sell<double>(const Launch&, int64_t, const int*, const double*, double*);

}  // namespace prism_hip
