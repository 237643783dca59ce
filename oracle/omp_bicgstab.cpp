// omp_bicgstab.cpp -- multi-threaded FP64 BiCGSTAB over a CSR matrix (test infrastructure:
// the all-cores variant of the CPU baseline that bench.py times beside the GPU).
//
// The iteration is Eigen's unpreconditioned BiCGSTAB (BiCGSTAB.h's recurrences, as the
// single-threaded oracle runs them), with each pass over the data a parallel loop: the SpMV
// fused with the dot product that follows it, and the x, r update fused with (r0, r).  When
// (r0, r) falls below eps^2 (r0, r0) the shadow residual restarts from the true residual, as
// in Eigen's loop.  The dot products are OpenMP reductions, so their rounding differs from
// the serial oracle's.
#include <omp.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <vector>

extern "C" __attribute__((visibility("default"))) int oref_omp_bicgstab(int64_t n, const int64_t* rp,
                                                                       const int32_t* col, const double* val,
                                                                       const double* b, double* x, int64_t maxit,
                                                                       int32_t threads, int64_t* iters,
                                                                       double* seconds) {
    if (n <= 0 || threads <= 0) return 1;
    omp_set_num_threads(threads);
    std::vector<double> r(n), r0(n), p(n, 0.0), v(n, 0.0), s(n), t(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double ax = 0.0;
        for (int64_t k = rp[i]; k < rp[i + 1]; ++k) ax += val[k] * x[col[k]];
        r[i] = b[i] - ax;
        r0[i] = r[i];
    }
    double rho = 1.0, alpha = 1.0, omega = 1.0, rho_new = 0.0;
#pragma omp parallel for reduction(+ : rho_new) schedule(static)
    for (int64_t i = 0; i < n; ++i) rho_new += r0[i] * r[i];
    double r0_sq = rho_new;
    const double eps2 = DBL_EPSILON * DBL_EPSILON;
    const double t0 = omp_get_wtime();
    int64_t k = 0;
    for (; k < maxit; ++k) {
        if (std::fabs(rho_new) < eps2 * r0_sq) {
            // r has become orthogonal to the shadow residual: restart it from the true residual
            double rr = 0.0;
#pragma omp parallel for reduction(+ : rr) schedule(static)
            for (int64_t i = 0; i < n; ++i) {
                double ax = 0.0;
                for (int64_t q = rp[i]; q < rp[i + 1]; ++q) ax += val[q] * x[col[q]];
                r[i] = b[i] - ax;
                r0[i] = r[i];
                rr += r[i] * r[i];
            }
            rho_new = r0_sq = rr;
        }
        const double beta = (rho_new / rho) * (alpha / omega);
        rho = rho_new;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        double r0v = 0.0;
#pragma omp parallel for reduction(+ : r0v) schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            double a = 0.0;
            for (int64_t q = rp[i]; q < rp[i + 1]; ++q) a += val[q] * p[col[q]];
            v[i] = a;
            r0v += r0[i] * a;
        }
        alpha = rho / r0v;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
        double tt = 0.0, ts = 0.0;
#pragma omp parallel for reduction(+ : tt, ts) schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            double a = 0.0;
            for (int64_t q = rp[i]; q < rp[i + 1]; ++q) a += val[q] * s[col[q]];
            t[i] = a;
            tt += a * a;
            ts += a * s[i];
        }
        omega = tt > 0.0 ? ts / tt : 0.0;
        rho_new = 0.0;
        double rr = 0.0;
#pragma omp parallel for reduction(+ : rho_new, rr) schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i] + omega * s[i];
            r[i] = s[i] - omega * t[i];
            rho_new += r0[i] * r[i];
            rr += r[i] * r[i];
        }
        if (!(rr == rr)) break;
    }
    *seconds = omp_get_wtime() - t0;
    *iters = k;
    return 0;
}
