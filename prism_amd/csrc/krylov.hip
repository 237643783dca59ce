// krylov.hip -- BiCGSTAB for gfx950: CSR SpMV staged through LDS with fused dot-product
// epilogues, element-wise kernels with fused dots, and deterministic reductions.
//
// Reductions: every block reduces its values with wave64 cross-lane shuffles and LDS,
// writes one partial, and takes an integer ticket; the block that draws the last ticket
// sums the partials in index order and updates the solver scalars in device memory.  The
// result does not depend on block scheduling, and no floating-point atomic is used.
//
// The iteration is Eigen's BiCGSTAB (with the identity or Jacobi preconditioner):
//   rho = (r0,r); beta = rho/rho_old * alpha/omega; p = r + beta (p - omega v)
//   v = A M^-1 p; alpha = rho/(r0,v); s = r - alpha v; t = A M^-1 s
//   omega = (t,s)/(t,t); x += alpha M^-1 p + omega M^-1 s; r = s - omega t
// Solver scalars live on the device so iterations are enqueued without host round trips;
// kernels return immediately once the device flags convergence or a pending event.
#include <hip/hip_runtime.h>

#include "internal.h"
#include "krylov.h"

namespace prism_hip {

namespace {

constexpr int BS = 256;
constexpr int NW = BS / 64;
constexpr int LDS_CAP = 8 * BS;  // staged products per block

// blockIdx -> logical block: consecutive logical blocks stay on one XCD (blocks are
// dispatched round-robin over the 8 XCDs), so neighbouring rows share an L2.
__device__ inline int logicalBlock(int nblk) {
    int b = blockIdx.x;
    int q = nblk >> 3, r = nblk & 7;
    int xcd = b & 7, idx = b >> 3;
    // XCD x receives q+1 blocks if x < r, else q: a bijection onto [0, nblk)
    return (xcd < r) ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

__device__ inline double waveSum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline double waveMax(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Sum of up to 3 values over the block; result valid in thread 0.
template <int K>
__device__ inline void blockSum(double (&v)[K], double (*sh)[NW]) {
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
            v[k] = s;
        }
    }
    __syncthreads();
}

// Writes this block's K partials, returns true in every thread of the last block, which
// then holds in v[] (thread 0) the sums over all blocks in index order.
template <int K>
__device__ bool reduceLast(double (&v)[K], double* part, unsigned* counter, int nblk, int lb) {
    __shared__ double sh[3][NW];
    __shared__ int last;
    blockSum<K>(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) part[k * nblk + lb] = v[k];
        __threadfence();
        unsigned t = atomicAdd(counter, 1u);
        last = (t == (unsigned)(nblk - 1));
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        acc[k] = 0.0;
        for (int i = threadIdx.x; i < nblk; i += BS) acc[k] += __builtin_nontemporal_load(&part[k * nblk + i]);
    }
    blockSum<K>(acc, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = acc[k];
        *counter = 0u;
    }
    return true;
}

// y_row = sum_k val[k] * x[col[k]] for the BS rows of logical block lb.  Products are
// streamed into LDS with coalesced loads (CSR-stream), then each thread sums its row in
// column order.  Falls back to a per-row loop when the block holds too many entries.
template <typename T>
__device__ inline T spmvRow(const int* __restrict__ rp, const int* __restrict__ col,
                            const T* __restrict__ val, const T* __restrict__ x, int n, int lb, T* lds) {
    int row0 = lb * BS;
    int rows = min(BS, n - row0);
    int k0 = rp[row0];
    int k1 = rp[row0 + rows];
    int cnt = k1 - k0;
    int row = row0 + threadIdx.x;
    T s = T(0);
    if (cnt <= LDS_CAP) {
        for (int k = threadIdx.x; k < cnt; k += BS) lds[k] = val[k0 + k] * x[col[k0 + k]];
        __syncthreads();
        if (threadIdx.x < rows) {
            int a = rp[row] - k0, b = rp[row + 1] - k0;
            for (int k = a; k < b; ++k) s += lds[k];
        }
        __syncthreads();
    } else if (threadIdx.x < rows) {
        for (int k = rp[row]; k < rp[row + 1]; ++k) s += val[k] * x[col[k]];
    }
    return s;
}

template <typename T>
struct LdsBuf {
    __device__ static T* get() {
        __shared__ T buf[LDS_CAP];
        return buf;
    }
};

__device__ inline bool halted(const KrylovScalars* ks) { return ks->done || ks->flag; }

// ------------------------------------------------------------------------ SpMV ----
template <typename T>
__global__ void __launch_bounds__(BS) k_spmv(int n, const int* rp, const int* col, const T* val,
                                             const T* x, T* y, int nblk) {
    int lb = logicalBlock(nblk);
    T* lds = LdsBuf<T>::get();
    T s = spmvRow<T>(rp, col, val, x, n, lb, lds);
    int row = lb * BS + threadIdx.x;
    if (row < n) y[row] = s;
}

// r = b - A x (FP64); partial sums of (r,r) and (b,b).  Used by the BiCGSTAB start and by
// the FP64 residual of the mixed-precision path.  mode 0: also r0 = r, p = v = 0.
template <typename T>
__global__ void __launch_bounds__(BS) k_residual(KView<T> kv, const double* b, const double* x64,
                                                 double* r64, int mode) {
    int lb = logicalBlock(kv.nblk);
    double* lds = LdsBuf<double>::get();
    double ax = spmvRow<double>(kv.rp, kv.col, kv.val64, x64, kv.n, lb, lds);
    int row = lb * BS + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (row < kv.n) {
        double ri = b[row] - ax;
        r64[row] = ri;
        v[0] = ri * ri;
        v[1] = b[row] * b[row];
    }
    if (reduceLast<2>(v, kv.part, kv.counter, kv.nblk, lb) && threadIdx.x == 0) {
        KrylovScalars* ks = kv.ks;
        ks->rr_true = v[0];
        if (mode == 0) {
            ks->bb = v[1];
            ks->tol2 = ks->tol * ks->tol * v[1];
        }
    }
}

// Start / restart / reliable update: r = T(r64), r0 (= r on start), x32 = 0 (mixed),
// p = v = 0 on start; scalars from (r,r), (r0,r), (r0,r0).
template <typename T>
__global__ void __launch_bounds__(BS) k_seed(KView<T> kv, const double* r64, int start) {
    int lb = logicalBlock(kv.nblk);
    int row = lb * BS + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (row < kv.n) {
        T ri = (T)r64[row];
        kv.r[row] = ri;
        T r0i;
        if (start) {  // 1: fresh start; 2: Eigen's rho-breakdown restart (keeps p, v)
            kv.r0[row] = ri;
            if (start == 1) {
                kv.p[row] = T(0);
                kv.v[row] = T(0);
            }
            r0i = ri;
        } else {
            r0i = kv.r0[row];
        }
        if (kv.mixed) kv.x[row] = T(0);
        v[0] = (double)ri * (double)ri;
        v[1] = (double)r0i * (double)ri;
        v[2] = (double)r0i * (double)r0i;
    }
    if (reduceLast<3>(v, kv.part, kv.counter, kv.nblk, lb) && threadIdx.x == 0) {
        KrylovScalars* ks = kv.ks;
        ks->rr = v[0];
        ks->rho_next = v[1];
        if (start) ks->r0n = v[2];
        if (start == 1) {
            ks->rho = 1.0;
            ks->alpha = 1.0;
            ks->omega = 1.0;
        }
        ks->rr_ref = ks->rr_true;
        ks->flag = 0;
        if (ks->bb == 0.0) ks->done = 4;
        else if (ks->rr_true <= ks->tol2) ks->done = 1;
        else if (start != 2 && fabs(ks->rho_next) < ks->breakdown * ks->r0n) ks->flag = 2;  // restart
    }
}

// p = r + beta (p - omega v); y = M^-1 p
template <typename T>
__global__ void __launch_bounds__(BS) k_update_p(KView<T> kv) {
    const KrylovScalars* ks = kv.ks;
    if (halted(ks)) return;
    double beta = (ks->rho_next / ks->rho) * (ks->alpha / ks->omega);
    T b = (T)beta, w = (T)ks->omega;
    int i = logicalBlock(kv.nblk) * BS + threadIdx.x;
    if (i >= kv.n) return;
    T pi = kv.r[i] + b * (kv.p[i] - w * kv.v[i]);
    kv.p[i] = pi;
    if (kv.precond) kv.y[i] = kv.dinv[i] * pi;
}

// v = A y; alpha = rho / (r0, v)
template <typename T>
__global__ void __launch_bounds__(BS) k_spmv_v(KView<T> kv) {
    KrylovScalars* ks = kv.ks;
    if (halted(ks)) return;
    int lb = logicalBlock(kv.nblk);
    T* lds = LdsBuf<T>::get();
    T vi = spmvRow<T>(kv.rp, kv.col, kv.val, kv.y, kv.n, lb, lds);
    int row = lb * BS + threadIdx.x;
    double v[1] = {0.0};
    if (row < kv.n) {
        kv.v[row] = vi;
        v[0] = (double)kv.r0[row] * (double)vi;
    }
    if (reduceLast<1>(v, kv.part, kv.counter, kv.nblk, lb) && threadIdx.x == 0) {
        ks->rho = ks->rho_next;
        ks->r0v = v[0];
        ks->alpha = ks->rho / v[0];
    }
}

// s = r - alpha v; z = M^-1 s
template <typename T>
__global__ void __launch_bounds__(BS) k_update_s(KView<T> kv) {
    const KrylovScalars* ks = kv.ks;
    if (halted(ks)) return;
    T a = (T)ks->alpha;
    int i = logicalBlock(kv.nblk) * BS + threadIdx.x;
    if (i >= kv.n) return;
    T si = kv.r[i] - a * kv.v[i];
    kv.s[i] = si;
    if (kv.precond) kv.z[i] = kv.dinv[i] * si;
}

// t = A z; omega = (t,s)/(t,t)
template <typename T>
__global__ void __launch_bounds__(BS) k_spmv_t(KView<T> kv) {
    KrylovScalars* ks = kv.ks;
    if (halted(ks)) return;
    int lb = logicalBlock(kv.nblk);
    T* lds = LdsBuf<T>::get();
    T ti = spmvRow<T>(kv.rp, kv.col, kv.val, kv.z, kv.n, lb, lds);
    int row = lb * BS + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (row < kv.n) {
        kv.t[row] = ti;
        double td = (double)ti;
        v[0] = td * td;
        v[1] = td * (double)kv.s[row];
    }
    if (reduceLast<2>(v, kv.part, kv.counter, kv.nblk, lb) && threadIdx.x == 0) {
        ks->tt = v[0];
        ks->ts = v[1];
        ks->omega = v[0] > 0.0 ? v[1] / v[0] : 0.0;
    }
}

// x += alpha y + omega z; r = s - omega t; (r,r), (r0,r); convergence / event flags.
template <typename T>
__global__ void __launch_bounds__(BS) k_update_xr(KView<T> kv) {
    KrylovScalars* ks = kv.ks;
    if (halted(ks)) return;
    int lb = logicalBlock(kv.nblk);
    T a = (T)ks->alpha, w = (T)ks->omega;
    int row = lb * BS + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (row < kv.n) {
        kv.x[row] += a * kv.y[row] + w * kv.z[row];
        T ri = kv.s[row] - w * kv.t[row];
        kv.r[row] = ri;
        double rd = (double)ri;
        v[0] = rd * rd;
        v[1] = (double)kv.r0[row] * rd;
    }
    if (reduceLast<2>(v, kv.part, kv.counter, kv.nblk, lb) && threadIdx.x == 0) {
        ks->rr = v[0];
        ks->rho_next = v[1];
        ks->iter += 1;
        if (!(v[0] == v[0])) {
            ks->done = 3;  // NaN: breakdown
        } else if (kv.mixed) {
            // reliable update once the recursive residual fell by delta, or reached tol
            if (v[0] <= ks->tol2 || v[0] <= ks->delta2 * ks->rr_ref) ks->flag = 1;
            else if (ks->iter >= ks->maxit) ks->flag = 1;
        } else if (v[0] <= ks->tol2) {
            ks->done = 1;
        } else if (ks->iter >= ks->maxit) {
            ks->done = 2;
        }
        if (!ks->done && !ks->flag && fabs(v[1]) < ks->breakdown * ks->r0n) ks->flag = 2;
    }
}

// x64 += x32 (mixed precision flush before the FP64 residual)
__global__ void k_flush(int n, double* x64, const float* x32) {
    int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) x64[i] += (double)x32[i];
}

// Scaled residual of solver.cpp:8-24: max|A x - b| and max|diag * x|.
__global__ void __launch_bounds__(BS) k_scaled_residual(int n, const int* rp, const int* col, const double* val,
                                                        const int* diag, const double* x, const double* b,
                                                        double* part, unsigned* counter, double* out, int nblk) {
    int lb = logicalBlock(nblk);
    double* lds = LdsBuf<double>::get();
    double ax = spmvRow<double>(rp, col, val, x, n, lb, lds);
    int row = lb * BS + threadIdx.x;
    double r = 0.0, d = 0.0;
    if (row < n) {
        r = fabs(ax - b[row]);
        d = fabs(val[diag[row]] * x[row]);
    }
    r = waveMax(r);
    d = waveMax(d);
    __shared__ double sh[2][NW];
    __shared__ int last;
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { sh[0][w] = r; sh[1][w] = d; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, c = 0;
        for (int i = 0; i < NW; ++i) { a = fmax(a, sh[0][i]); c = fmax(c, sh[1][i]); }
        part[lb] = a;
        part[nblk + lb] = c;
        __threadfence();
        unsigned t = atomicAdd(counter, 1u);
        last = (t == (unsigned)(nblk - 1));
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double a = 0, c = 0;
    for (int i = threadIdx.x; i < nblk; i += BS) {
        a = fmax(a, __builtin_nontemporal_load(&part[i]));
        c = fmax(c, __builtin_nontemporal_load(&part[nblk + i]));
    }
    a = waveMax(a);
    c = waveMax(c);
    __syncthreads();
    if (lane == 0) { sh[0][w] = a; sh[1][w] = c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double A = 0, C = 0;
        for (int i = 0; i < NW; ++i) { A = fmax(A, sh[0][i]); C = fmax(C, sh[1][i]); }
        out[0] = A;
        out[1] = C;
        *counter = 0u;
    }
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
template <typename T>
__global__ void k_to_double(int n, const T* src, double* dst) {
    int i = blockIdx.x * BS + threadIdx.x;
    if (i < n) dst[i] = (double)src[i];
}

struct Scope {
    const Launch& L;
    const char* name;
    Scope(const Launch& l, const char* nm) : L(l), name(nm) { if (L.begin) L.begin(L.user, name); }
    ~Scope() { if (L.end) L.end(L.user, name); }
};

inline int nb(int64_t n) { return (int)((n + BS - 1) / BS); }

}  // namespace

#define KLAUNCH(label, kern, grid, ...)                                                 \
    do {                                                                                \
        if ((grid) > 0) {                                                               \
            Scope _s(L, label);                                                         \
            hipLaunchKernelGGL(kern, dim3(grid), dim3(BS), 0, L.stream, __VA_ARGS__);   \
            HIPCHECK(hipGetLastError());                                                \
        }                                                                               \
    } while (0)

int krylov_blocks(int64_t n) { return nb(n); }

void launch_spmv64(const Launch& L, int n, const int* rp, const int* col, const double* val, const double* x,
                   double* y) {
    KLAUNCH("spmv_f64", k_spmv<double>, nb(n), n, rp, col, val, x, y, nb(n));
}

void launch_scaled_residual(const Launch& L, int n, const int* rp, const int* col, const double* val,
                            const int* diag, const double* x, const double* b, double* part, unsigned* counter,
                            double* out) {
    KLAUNCH("scaled_residual", k_scaled_residual, nb(n), n, rp, col, val, diag, x, b, part, counter, out, nb(n));
}

template <typename T>
void launch_residual(const Launch& L, const KView<T>& kv, const double* b, const double* x64, double* r64,
                     int mode) {
    KLAUNCH("bicgstab_residual_f64", k_residual<T>, kv.nblk, kv, b, x64, r64, mode);
}
template <typename T>
void launch_seed(const Launch& L, const KView<T>& kv, const double* r64, int start) {
    KLAUNCH(sizeof(T) == 8 ? "bicgstab_seed_f64" : "bicgstab_seed_f32", k_seed<T>, kv.nblk, kv, r64, start);
}
template <typename T>
void launch_iteration(const Launch& L, const KView<T>& kv) {
    const bool d = sizeof(T) == 8;
    KLAUNCH(d ? "bicgstab_update_p_f64" : "bicgstab_update_p_f32", k_update_p<T>, kv.nblk, kv);
    KLAUNCH(d ? "bicgstab_spmv_v_f64" : "bicgstab_spmv_v_f32", k_spmv_v<T>, kv.nblk, kv);
    KLAUNCH(d ? "bicgstab_update_s_f64" : "bicgstab_update_s_f32", k_update_s<T>, kv.nblk, kv);
    KLAUNCH(d ? "bicgstab_spmv_t_f64" : "bicgstab_spmv_t_f32", k_spmv_t<T>, kv.nblk, kv);
    KLAUNCH(d ? "bicgstab_update_xr_f64" : "bicgstab_update_xr_f32", k_update_xr<T>, kv.nblk, kv);
}
void launch_flush(const Launch& L, int n, double* x64, const float* x32) {
    KLAUNCH("bicgstab_flush", k_flush, nb(n), n, x64, x32);
}
template <typename T>
void launch_convert(const Launch& L, int64_t n, const double* src, T* dst) {
    KLAUNCH("convert_values", k_convert<T>, nb(n), n, src, dst);
}
template <typename T>
void launch_dinv(const Launch& L, int n, const int* diag, const double* val, T* dinv) {
    KLAUNCH("jacobi_dinv", k_dinv<T>, nb(n), n, diag, val, dinv);
}
template <typename T>
void launch_to_double(const Launch& L, int n, const T* src, double* dst) {
    KLAUNCH("to_double", k_to_double<T>, nb(n), n, src, dst);
}

template void launch_residual<double>(const Launch&, const KView<double>&, const double*, const double*, double*, int);
template void launch_residual<float>(const Launch&, const KView<float>&, const double*, const double*, double*, int);
template void launch_seed<double>(const Launch&, const KView<double>&, const double*, int);
template void launch_seed<float>(const Launch&, const KView<float>&, const double*, int);
template void launch_iteration<double>(const Launch&, const KView<double>&);
template void launch_iteration<float>(const Launch&, const KView<float>&);
template void launch_convert<float>(const Launch&, int64_t, const double*, float*);
template void launch_convert<double>(const Launch&, int64_t, const double*, double*);
template void launch_dinv<float>(const Launch&, int, const int*, const double*, float*);
template void launch_dinv<double>(const Launch&, int, const int*, const double*, double*);
template void launch_to_double<float>(const Launch&, int, const float*, double*);

}  // namespace prism_hip
