// CPU check of the multicolour SGS layout (buildSgsLayout, host_mesh.cpp) and of the sweep
// arithmetic the device kernels perform on it (sgs.hip: forward sweeps over the entries with
// column < cptr[c], backward sweeps over those with column >= cptr[c+1], sliced-ELL slots at
// soff[s] + lane + 64 m).  For each pattern: the colouring is proper, perm is a permutation
// grouped by colour, every off-diagonal entry sits in exactly one slot of its position, the
// padding is inert; the sweeps applied through the layout equal a dense solve of M y = p with
// M = (D + L) D^-1 (D + U) in the colour ordering; and the product in Eisenstat's form,
// p + L e (e: the backward sweeps' corrections), equals A y.  Prints "ok" and exits 0.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <vector>

#include "../../prism_amd/csrc/host_mesh.h"
#include "../../prism_amd/csrc/internal.h"

using namespace prism_hip;

namespace {

int failures = 0;
#define CHECK(c, ...)                                   \
    do {                                                \
        if (!(c)) {                                     \
            std::fprintf(stderr, "FAIL %s: ", #c);      \
            std::fprintf(stderr, __VA_ARGS__);          \
            std::fprintf(stderr, "\n");                 \
            if (++failures > 20) std::exit(1);          \
        }                                               \
    } while (0)

struct Csr {
    int n;
    std::vector<int32_t> rp, col, diag;
    std::vector<double> val;
};

Csr fromRows(const std::vector<std::set<int>>& rows, std::mt19937& rng) {
    Csr a;
    a.n = (int)rows.size();
    a.rp.push_back(0);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (int i = 0; i < a.n; ++i) {
        double off = 0.0;
        int dpos = -1;
        for (int j : rows[i]) {
            if (j == i) dpos = (int)a.col.size();
            a.col.push_back(j);
            double v = j == i ? 0.0 : u(rng);
            off += std::fabs(v);
            a.val.push_back(v);
        }
        a.val[dpos] = 1.5 * off + 0.5;  // nonsymmetric, diagonally dominant
        a.diag.push_back(dpos);
        a.rp.push_back((int)a.col.size());
    }
    return a;
}

Csr grid7(int nx, int ny, int nz, std::mt19937& rng) {
    std::vector<std::set<int>> rows(nx * ny * nz);
    auto id = [&](int i, int j, int k) { return i + nx * (j + ny * k); };
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                auto& r = rows[id(i, j, k)];
                r.insert(id(i, j, k));
                if (i > 0) r.insert(id(i - 1, j, k));
                if (i + 1 < nx) r.insert(id(i + 1, j, k));
                if (j > 0) r.insert(id(i, j - 1, k));
                if (j + 1 < ny) r.insert(id(i, j + 1, k));
                if (k > 0) r.insert(id(i, j, k - 1));
                if (k + 1 < nz) r.insert(id(i, j, k + 1));
            }
    return fromRows(rows, rng);
}

// random pattern, symmetric or not, up to `deg` off-diagonals per row
Csr randomPattern(int n, int deg, bool symmetric, std::mt19937& rng) {
    std::vector<std::set<int>> rows(n);
    std::uniform_int_distribution<int> pick(0, n - 1);
    for (int i = 0; i < n; ++i) {
        rows[i].insert(i);
        for (int d = 0; d < deg / 2; ++d) {
            int j = pick(rng);
            rows[i].insert(j);
            if (symmetric) rows[j].insert(i);
        }
    }
    return fromRows(rows, rng);
}

void checkPattern(const char* name, const Csr& a, std::mt19937& rng) {
    SgsLayout L;
    buildSgsLayout(a.n, a.rp.data(), a.col.data(), a.diag.data(), 16, L);
    const int n = a.n;
    std::vector<int> colour(n, -1), pos(n, -1);
    CHECK(L.cptr.size() == (size_t)L.nc + 1 && L.cptr[0] == 0 && L.cptr[L.nc] == n, "%s: cptr", name);
    for (int c = 0; c < L.nc; ++c) {
        CHECK(L.cptr[c + 1] > L.cptr[c], "%s: empty colour %d", name, c);
        for (int q = L.cptr[c]; q < L.cptr[c + 1]; ++q) {
            const int r = L.perm[q];
            CHECK(r >= 0 && r < n && pos[r] < 0, "%s: perm not a permutation at %d", name, q);
            pos[r] = q;
            colour[r] = c;
            if (q > L.cptr[c]) CHECK(L.perm[q - 1] < r, "%s: row order inside colour %d", name, c);
        }
    }
    for (int i = 0; i < n; ++i)
        for (int k = a.rp[i]; k < a.rp[i + 1]; ++k)
            if (a.col[k] != i) CHECK(colour[a.col[k]] != colour[i], "%s: rows %d, %d share colour", name, i, a.col[k]);
    const int ns = (n + 63) / 64;
    CHECK((int)L.soff.size() == ns + 1 && L.soff[0] == 0, "%s: soff", name);
    CHECK(L.scol.size() == (size_t)L.soff[ns] && L.emap.size() == L.scol.size(), "%s: slot arrays", name);
    for (int q = 0; q < n; ++q) {
        const int s = q >> 6, l = q & 63, w = (L.soff[s + 1] - L.soff[s]) >> 6;
        const int r = L.perm[q];
        CHECK(L.dmap[q] == a.diag[r], "%s: dmap at %d", name, q);
        std::multiset<int> want, got;
        for (int k = a.rp[r]; k < a.rp[r + 1]; ++k)
            if (a.col[k] != r) want.insert(k);
        for (int m = 0; m < w; ++m) {
            const int slot = L.soff[s] + l + 64 * m;
            const int e = L.emap[slot];
            if (e < 0) {
                CHECK(L.scol[slot] == q, "%s: padding column at position %d", name, q);
                continue;
            }
            got.insert(e);
            CHECK(e >= a.rp[r] && e < a.rp[r + 1] && L.scol[slot] == pos[a.col[e]], "%s: slot of position %d", name, q);
        }
        CHECK(want == got, "%s: entries of position %d", name, q);
    }
    // values in position numbering, as k_sgs_values forms them
    std::vector<double> sval(L.scol.size()), d(n), dinv(n);
    for (size_t k = 0; k < sval.size(); ++k) sval[k] = L.emap[k] >= 0 ? a.val[L.emap[k]] : 0.0;
    for (int q = 0; q < n; ++q) {
        d[q] = a.val[L.dmap[q]];
        dinv[q] = 1.0 / d[q];
    }
    auto offRow = [&](int mode, int i, const std::vector<double>& w, int lo, int hi) {
        const int s = i >> 6, l = i & 63, o = L.soff[s], wd = (L.soff[s + 1] - o) >> 6;
        double sum = 0.0;
        for (int m = 0; m < wd; ++m) {
            const int j = L.scol[o + l + 64 * m];
            if (mode == 0 || (mode == 1 && j < lo) || (mode == 2 && j >= hi)) sum += sval[o + l + 64 * m] * w[j];
        }
        return sum;
    };
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<double> p(n), y(n, 1e300);  // y starts as garbage: the sweeps must not read it
    for (auto& v : p) v = u(rng);
    for (int c = 0; c < L.nc; ++c)
        for (int i = L.cptr[c]; i < L.cptr[c + 1]; ++i)
            y[i] = (p[i] - (c > 0 ? offRow(1, i, y, L.cptr[c], L.cptr[c + 1]) : 0.0)) * dinv[i];
    std::vector<double> e(n, 1e300);  // corrections of the backward sweeps (last colour: never read)
    for (int c = L.nc - 2; c >= 0; --c)
        for (int i = L.cptr[c]; i < L.cptr[c + 1]; ++i) {
            e[i] = -offRow(2, i, y, L.cptr[c], L.cptr[c + 1]) * dinv[i];
            y[i] += e[i];
        }
    // dense M = (D + L) D^-1 (D + U) in position numbering; check M y = p
    std::vector<double> Ad((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = a.rp[i]; k < a.rp[i + 1]; ++k) Ad[(size_t)pos[i] * n + pos[a.col[k]]] = a.val[k];
    std::vector<double> u1(n), My(n);
    for (int i = 0; i < n; ++i) {  // u1 = (D + U) y
        double s = 0.0;
        for (int j = i; j < n; ++j) s += Ad[(size_t)i * n + j] * y[j];
        u1[i] = s / Ad[(size_t)i * n + i];
    }
    double err = 0.0, ref = 0.0;
    for (int i = 0; i < n; ++i) {  // My = (D + L) D^-1 u1
        double s = 0.0;
        for (int j = 0; j <= i; ++j) s += Ad[(size_t)i * n + j] * u1[j];
        My[i] = s;
        err = std::max(err, std::fabs(My[i] - p[i]));
        ref = std::max(ref, std::fabs(p[i]));
    }
    CHECK(err <= 1e-12 * ref, "%s: M y - p = %.3e", name, err);
    // padding adds exact zeros to the full product
    double perr = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = d[i] * y[i] + offRow(0, i, y, 0, 0), t = 0.0;
        for (int j = 0; j < n; ++j) t += Ad[(size_t)i * n + j] * y[j];
        perr = std::max(perr, std::fabs(s - t));
    }
    CHECK(perr <= 1e-12 * (1.0 + ref), "%s: A y through the layout, error %.3e", name, perr);
    // Eisenstat's form of the product (k_sgs_av / k_sgs_at): A y = p + L e, reading only the
    // entries of earlier colours and e of colours before the last
    double eerr = 0.0;
    for (int c = 0; c < L.nc; ++c)
        for (int i = L.cptr[c]; i < L.cptr[c + 1]; ++i) {
            const double v = p[i] + (c > 0 ? offRow(1, i, e, L.cptr[c], 0) : 0.0);
            double t = 0.0;
            for (int j = 0; j < n; ++j) t += Ad[(size_t)i * n + j] * y[j];
            eerr = std::max(eerr, std::fabs(v - t));
        }
    CHECK(eerr <= 1e-12 * (1.0 + ref), "%s: p + L e against A y, error %.3e", name, eerr);
    std::printf("%s: n %d, %d colours, %zu slots\n", name, n, L.nc, L.scol.size());
}

}  // namespace

int main() {
    std::mt19937 rng(7);
    checkPattern("grid 9x7x5", grid7(9, 7, 5, rng), rng);
    checkPattern("grid 1x1x1", grid7(1, 1, 1, rng), rng);
    checkPattern("grid 130x1x1", grid7(130, 1, 1, rng), rng);
    checkPattern("random symmetric", randomPattern(300, 8, true, rng), rng);
    checkPattern("random nonsymmetric", randomPattern(257, 6, false, rng), rng);
    // a pattern whose greedy colouring needs more colours than allowed is refused
    bool refused = false;
    try {
        std::vector<std::set<int>> rows(20);
        for (int i = 0; i < 20; ++i)
            for (int j = 0; j < 20; ++j) rows[i].insert(j);  // complete graph: 20 colours
        Csr a = fromRows(rows, rng);
        SgsLayout L;
        buildSgsLayout(a.n, a.rp.data(), a.col.data(), a.diag.data(), 16, L);
    } catch (const Error& e) {
        refused = e.code == PRISM_ERR_RUNTIME;
    }
    CHECK(refused, "complete graph of 20 rows not refused");
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
