// krylov.h -- BiCGSTAB device state and launchers (krylov.hip).
#pragma once

#include <cstdint>

#include "internal.h"

namespace prism_hip {

// Device-resident solver scalars.  done: 1 converged, 2 iteration cap, 3 breakdown (NaN),
// 4 zero right-hand side.  flag: 1 reliable update due (mixed precision), 2 restart due.
// rho is double-buffered by iteration parity: iteration k writes rho[k & 1] and reads the
// previous one from rho[(k & 1) ^ 1], so no kernel reads a slot that it also writes.
struct KrylovScalars {
    double rho[2], alpha, omega, beta;
    double r0v, tt, ts;
    double rr, rr_true, rr_ref, r0n, bb;
    double tol, tol2, delta2, breakdown;
    long long iter, nit, maxit;
    int done, flag;
    double scaled[2];  // scaled residual: max|A x - b|, max|diag * x|
};

// Matrix formats of the Krylov kernels: sliced ELL with 32-bit columns or 16-bit offsets
// col - row, or DIA (diagonals) for patterns with at most DIA_MAX distinct offsets.
enum { FMT_SELL32 = 0, FMT_SELL16 = 1, FMT_DIA = 2, FMT_DIA_SYM = 3 };
constexpr int DIA_MAX = 8;

// Arguments of the per-iteration kernels.  Partial sums are produced by one kernel
// (part_out, 3 x G doubles) and reduced by the next (part_in): kernel boundaries order
// the exchange, so no fence or atomic is needed.
template <typename T>
struct KView {
    int n, G, slices_per_block;  // rows, blocks, 64-row slices per block
    int Gv, spb_v;               // the same for the V kernel (it writes no partial sums)
    const int* rp;
    const int* col;
    const T* val;         // CSR values in the iteration precision (unused by the SELL kernels)
    const double* val64;  // CSR FP64 values
    const int* soff;      // sliced-ELL: slice offsets [n/64 + 1] (entries), column-major slices
    const int* scol;      // sliced-ELL column indices (padding: own row)
    const short* srel;    // sliced-ELL 16-bit offsets col - row - sbase[slot], or nullptr (then scol)
    const int* sbase;     // with srel: per slice s and slot k, sbase[soff[s] / 64 + k]
    int ndia;             // DIA: number of diagonals (0: sliced ELL), their offsets ascending
    int doff[DIA_MAX];
    int dsym;             // DIA of an exactly symmetric matrix: only the diagonal and the upper
                          // diagonals are read (entry (i, i-o) is the upper entry (i-o, i))
    const T* sval;        // sliced-ELL (or DIA) values in the iteration precision
    const double* sval64; // sliced-ELL (or DIA) FP64 values (residuals)
    T *x, *r, *r0, *p_old, *v_old, *p_new, *v_new, *s, *t;
    const T* dinv;        // M^-1 for the x update (Jacobi); the SpMV values are pre-scaled
    T* rhat;              // (A M^-1)^T r0: (r0, A M^-1 p) = (rhat, p)
    const int *cscp, *cscr, *cscm;  // transpose of the pattern (ptr, row, CSR position)
    int precond, mixed, par;
    int pack;             // packed rows (krylov_pack): bit 0 the X kernel, bit 1 the DIA SpMV kernels V, T,
                          // bit 2 their branch-free seven-point product (box patterns)
    int diff;             // FP32 three-kernel SpMVs in the difference form: row sums in the diagonal
                          // slots (1), or zero row sums, diagonal slots unread (2)
    double *part_in, *part_out;
    double* part_x;       // (rhat, p), (r, r) of the next iteration
    KrylovScalars* ks;
};

// Blocks of the iteration kernels for n rows (and the 64-row slices per block).  band: one
// block per CU in the band of sizes where that was measured faster for kernels whose partial
// sums the next kernel reduces; the V kernel, which writes none, takes band = false.
int krylov_grid(int64_t n, int* slices_per_block, bool band = true, bool sell = false);
// Size of the static device scratch array behind prism_eqn_step_noalloc (one per device), and
// its address on `device`, which must be current (hipGetSymbolAddress; called once per context).
constexpr size_t SCRATCH_BYTES = size_t(24) << 20;
void* scratch_address(int device);
// Jacobi sweep of jacobi.cpp:5-9: y = x + D^-1 (b - A x) (CSR, FP64).
void launch_jacobi_step(const Launch& L, int n, const int* rp, const int* col, const double* val, const int* diag,
                        const double* b, const double* x, double* y);
// Reference-cell deflation helpers: structure test (per-block max |row sum|, |diag|,
// |A - A^T|, 3 doubles per block of ref_blocks(n)), per-block sums, x += c.
// Every launcher that writes per-block partials takes its buffer's capacity (doubles) and
// raises PRISM_ERR_RUNTIME before launching when the grid would write past it.
int ref_blocks(int n);
void launch_ref_check(const Launch& L, int n, const int* rp, const int* col, const double* val, const int* diag,
                      int ref, double a, double* part, size_t part_cap);
void launch_block_sum(const Launch& L, int n, const double* x, double* part, size_t part_cap);
void launch_add_const(const Launch& L, int n, double* x, double c);
void launch_spmv64(const Launch& L, int n, const int* rp, const int* col, const double* val, const double* x,
                   double* y);
// Scaled residual max|A x - b| / max|diag * x| into ks->scaled; `part` must hold
// scaled_residual_partials(n) doubles (its capacity, part_cap, is checked: the kernel writes
// two partial maxima per block of its own 512-row-chunk grid, up to GMAX blocks).
size_t scaled_residual_partials(int64_t n);
void launch_scaled_residual(const Launch& L, int n, const int* rp, const int* col, const double* val,
                            const int* diag, const double* x, const double* b, double* part, size_t part_cap,
                            KrylovScalars* ks);
// r64 = b - A64 x64; rr_true (and on start bb, tol2) into ks.
template <typename T>
void launch_residual(const Launch& L, const KView<T>& kv, const double* b, const double* x64, double* r64, int start);
// r = T(r64); start 1: r0 = r, p = v = 0, rho/alpha/omega = 1; start 2: r0 = r (restart);
// mixed: x32 = 0.  Leaves the (r,r), (r0,r) partials for the first iteration.
template <typename T>
void launch_seed(const Launch& L, const KView<T>& kv, const double* r64, int start);
template <typename T>
void launch_iteration(const Launch& L, const KView<T>& kv);
// Three-kernel iteration (V: alpha, v, s; T: t and four dots; X: omega, rho, beta, x, r, p).
// Entry after every seed: rhat = (A M^-1)^T r0 when r0 changed (start 1, 2; r0_64 is the FP64
// r0, dinv64 M^-1 or nullptr), then the p update with its (rhat, p), (r, r) partials.
template <typename T>
void launch_rhat(const Launch& L, const KView<T>& kv, const double* r0_64, const double* dinv64);
template <typename T>
void launch_entry(const Launch& L, const KView<T>& kv);
template <typename T>
void launch_iteration3(const Launch& L, const KView<T>& kv);
// Which kernels of launch_iteration3 take the packed-row form for this view (sets kv.pack):
// FP32 (FP64 only with PRISM_KRYLOV_PACK=2), n a multiple of the pack width W (4 rows in
// FP32, 2 in FP64) and, unless PRISM_KRYLOV_PACK=2, at least BS * W / 2 rows per block; the
// SpMV kernels also need a DIA matrix whose offsets are 0, +-1 or multiples of W.
// PRISM_KRYLOV_PACK=0 keeps the row-per-thread kernels.
template <typename T>
void krylov_pack(KView<T>& kv);
void launch_flush(const Launch& L, int n, double* x64, float* x32);
// Exact symmetry of a DIA value array with symmetric offsets: per block, 1.0 when some lower
// entry differs from its mirrored upper entry (ref_blocks(n) doubles).
void launch_dia_symcheck(const Launch& L, const KView<double>& kv, const double* vals, double* part,
                         size_t part_cap);
// ks->done = 0 and / or ks->iter = 0 (on the stream, before a seed).
void launch_reset_events(const Launch& L, KrylovScalars* ks, bool clear_done, bool clear_iter);
template <typename T>
void launch_convert(const Launch& L, int64_t n, const double* src, T* dst);
template <typename T>
void launch_dinv(const Launch& L, int n, const int* diag, const double* val, T* dinv);
template <typename T>
void launch_to_sell(const Launch& L, int64_t nnz, const int* csr2sell, const double* val, T* sval);
// Difference form (kv.diff): the FP64 row sums of A (column-scaled by dinv when given),
// rounded into the diagonal slots of an iteration value array (map: CSR position -> slot).
template <typename T>
void launch_rowsum_diag(const Launch& L, int n, const int* rp, const int* col, const int* diag, const int* map,
                        const double* val, const double* dinv, T* sval);
// Column-scaled values A diag(dinv) for the right Jacobi preconditioner.
template <typename T>
void launch_to_sell_scaled(const Launch& L, int64_t nnz, const int* csr2sell, const int* col, const double* val,
                           const double* dinv, T* sval);

}  // namespace prism_hip
