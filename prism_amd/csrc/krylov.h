// krylov.h -- BiCGSTAB device state and launchers (krylov.hip).
#pragma once

#include <cstdint>

#include "internal.h"

namespace prism_hip {

// Device-resident solver scalars.  done: 1 converged, 2 iteration cap, 3 breakdown (NaN),
// 4 zero right-hand side.  flag: 1 reliable update due (mixed precision), 2 restart due.
struct KrylovScalars {
    double rho, rho_next, alpha, omega;
    double r0v, tt, ts;
    double rr, rr_true, rr_ref, r0n, bb;
    double tol, tol2, delta2, breakdown;
    long long iter, maxit;
    int done, flag;
    double scaled[2];  // scaled residual numerator / denominator
};

template <typename T>
struct KView {
    int n, nblk;
    const int* rp;
    const int* col;
    const T* val;         // matrix values in the iteration precision
    const double* val64;  // FP64 values (residuals)
    T *x, *r, *r0, *p, *v, *s, *t, *y, *z;
    const T* dinv;
    int precond, mixed;
    double* part;
    unsigned* counter;
    KrylovScalars* ks;
};

int krylov_blocks(int64_t n);
void launch_spmv64(const Launch& L, int n, const int* rp, const int* col, const double* val, const double* x,
                   double* y);
void launch_scaled_residual(const Launch& L, int n, const int* rp, const int* col, const double* val,
                            const int* diag, const double* x, const double* b, double* part, unsigned* counter,
                            double* out);
template <typename T>
void launch_residual(const Launch& L, const KView<T>& kv, const double* b, const double* x64, double* r64, int mode);
template <typename T>
void launch_seed(const Launch& L, const KView<T>& kv, const double* r64, int start);
template <typename T>
void launch_iteration(const Launch& L, const KView<T>& kv);
void launch_flush(const Launch& L, int n, double* x64, const float* x32);
template <typename T>
void launch_convert(const Launch& L, int64_t n, const double* src, T* dst);
template <typename T>
void launch_dinv(const Launch& L, int n, const int* diag, const double* val, T* dinv);
template <typename T>
void launch_to_double(const Launch& L, int n, const T* src, double* dst);

}  // namespace prism_hip
