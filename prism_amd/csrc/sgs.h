// sgs.h -- BiCGSTAB right-preconditioned by multicolour symmetric Gauss-Seidel (FP64).
//
// The rows are coloured so that no two rows of one colour are coupled (greedy, natural
// order; a structured hexahedral mesh gets the two red-black colours) and renumbered colour
// by colour: colour c owns the positions [cptr[c], cptr[c+1]).  In that numbering the strict
// lower triangle L holds exactly the couplings to earlier colours and U those to later ones,
// so M = (D + L) D^-1 (D + U) is applied by one forward sweep per colour,
//   y_i = (p_i - sum_{j < cptr[c]} a_ij y_j) / a_ii,
// and one backward sweep per colour, from the last but one down to the first,
//   y_i -= (sum_{j >= cptr[c+1]} a_ij y_j) / a_ii,
// each a fully parallel pass over its colour's rows.  The iteration is Eigen's BiCGSTAB with
// a preconditioner (BiCGSTAB.h; the reference's solver is BiCGSTAB<SparseMatrix, IncompleteLUT>,
// bicgstab.h:12): x, r and the stopping test ||r|| <= tol ||b|| are those of A x = b.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.h"

namespace prism_hip {

constexpr int SGS_MAXC = 16;  // colours at most
constexpr int SGS_BS = 256;   // threads per block
constexpr int SGS_GMAX = 1024;  // blocks of the grid-stride kernels (partial sums per array)

// Iteration scalars, one slot per iteration parity: iteration k reads slot (k-1)&1 and its
// first kernel writes slot k&1, so no kernel writes a slot that its own blocks read.
struct SgsSlot {
    double rho, alpha, omega, beta, r0sq;
};
struct SgsState {
    SgsSlot s[2];
    double rr;          // (r, r) reduced at the top of the last iteration that ran
    double bb;          // (b, b)
    double tol2;        // tol^2 (b, b)
    double tol;
    long long k_event;  // iteration whose first kernel raised done / flag
    int done;           // 1 converged, 3 NaN, 4 b == 0
    int flag;           // 2 rho breakdown: restart with r0 = r = b - A x
    int bad_diag;       // a zero or non-finite diagonal entry
    int pad;
};

struct SgsView {
    int n, nc, G;
    int cptr[SGS_MAXC + 1];
    const int* soff;     // slice offsets (slices of 64 positions)
    const int* scol;     // off-diagonal columns in position numbering; padding = own position
    const double* sval;  // off-diagonal values (padding 0)
    const double* d;     // diagonal
    const double* dinv;  // 1 / diagonal
    double *x, *b, *r, *r0, *p, *v, *y, *s, *z, *t;  // position numbering
    double* e;  // backward-sweep corrections (y - f, then z - f'), read by the products
    double* part_x;  // seed / X partials: (r0, r), (r, r), (b, b)   3 G
    double* part_v;  // (r0, v)                                        G
    double* part_t;  // (t, t), (t, s)                                 2 G
    SgsState* st;
};

// Per-solve values: sval[k] = val[emap[k]] (0 where emap < 0), d / dinv from dmap; flags a
// zero or non-finite diagonal in st->bad_diag.
void launch_sgs_values(const Launch& L, int64_t slots, const int* emap, const double* val, double* sval, int n,
                       const int* dmap, double* d, double* dinv, SgsState* st);
// dst[k] = src[perm[k]] (natural -> position numbering)
void launch_sgs_gather(const Launch& L, int n, const int* perm, const double* src, double* dst);
// dst[perm[k]] = src[k] (position -> natural numbering)
void launch_sgs_scatter(const Launch& L, int n, const int* perm, const double* src, double* dst);
// r = b - A x, r0 = r, partials (r0, r), (r, r), (b, b); start: also p = v = 0.
void launch_sgs_seed(const Launch& L, const SgsView& V, bool start);
// One iteration k; fresh: 1 right after the start seed, 2 right after a restart seed, else 0.
void launch_sgs_iteration(const Launch& L, const SgsView& V, long long k, int fresh);
// ||b - A x||^2 and ||b||^2 partials into part_t (2 G), without touching the iteration's vectors.
void launch_sgs_true_residual(const Launch& L, const SgsView& V);

}  // namespace prism_hip
