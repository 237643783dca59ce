// CPU check of the sliced-ELL column encodings (buildSell, sellRelative in host_mesh.cpp):
// for every row below n and every slot of its slice, the 16-bit offset plus the slot's base
// plus the row gives back the 32-bit column that buildSell stored, on patterns whose column
// offsets span up to 2^16 within a slot; patterns with a slot spanning more are refused.
// Prints "ok" and exits 0.
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
#define CHECK(c, ...)                              \
    do {                                           \
        if (!(c)) {                                \
            std::fprintf(stderr, "FAIL %s: ", #c); \
            std::fprintf(stderr, __VA_ARGS__);     \
            std::fprintf(stderr, "\n");            \
            if (++failures > 20) std::exit(1);     \
        }                                          \
    } while (0)

struct Csr {
    int64_t n;
    std::vector<int32_t> rp, col;
};

Csr fromRows(const std::vector<std::set<int64_t>>& rows) {
    Csr a;
    a.n = (int64_t)rows.size();
    a.rp.push_back(0);
    for (const auto& r : rows) {
        for (int64_t j : r) a.col.push_back((int32_t)j);
        a.rp.push_back((int32_t)a.col.size());
    }
    return a;
}

// rows of a tetrahedral box mesh (6 tetrahedra per hex, numbered hex by hex): the cell, two
// neighbours in its own hex and two in the adjacent hexes of two different directions
Csr tetBox(int nx, int ny, int nz) {
    const int64_t n = 6LL * nx * ny * nz;
    std::vector<std::set<int64_t>> rows(n);
    auto hex = [&](int i, int j, int k) { return 6LL * (i + (int64_t)nx * (j + (int64_t)ny * k)); };
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                for (int t = 0; t < 6; ++t) {
                    const int64_t c = hex(i, j, k) + t;
                    auto& r = rows[c];
                    r.insert(c);
                    r.insert(hex(i, j, k) + (t + 1) % 6);
                    r.insert(hex(i, j, k) + (t + 5) % 6);
                    const int dir = t % 3, sgn = t < 3 ? -1 : 1;  // one face on a cube face in x, y or z
                    int ii = i, jj = j, kk = k;
                    (dir == 0 ? ii : dir == 1 ? jj : kk) += sgn;
                    if (ii >= 0 && ii < nx && jj >= 0 && jj < ny && kk >= 0 && kk < nz) r.insert(hex(ii, jj, kk) + (5 - t));
                    const int dir2 = (t + 1) % 3;
                    ii = i, jj = j, kk = k;
                    (dir2 == 0 ? ii : dir2 == 1 ? jj : kk) -= sgn;
                    if (ii >= 0 && ii < nx && jj >= 0 && jj < ny && kk >= 0 && kk < nz) r.insert(hex(ii, jj, kk) + (5 - t));
                }
    return fromRows(rows);
}

// random rows with neighbours at +-far and near offsets
Csr randomFar(int64_t n, int64_t far, std::mt19937& rng) {
    std::vector<std::set<int64_t>> rows(n);
    std::uniform_int_distribution<int> near(-40, 40), pick(0, 3);
    for (int64_t i = 0; i < n; ++i) {
        rows[i].insert(i);
        for (int q = 0; q < 3; ++q) {
            int64_t j = i + near(rng);
            if (j >= 0 && j < n) rows[i].insert(j);
        }
        if (pick(rng) == 0 && i + far < n) rows[i].insert(i + far);
        if (pick(rng) == 0 && i - far >= 0) rows[i].insert(i - far);
    }
    return fromRows(rows);
}

// returns whether the pattern was encoded; checks every stored slot when it was
bool check(const char* name, const Csr& a) {
    std::vector<int32_t> off, scol, c2s;
    buildSell(a.n, a.rp.data(), a.col.data(), off, scol, c2s);
    std::vector<int16_t> rel;
    std::vector<int32_t> base;
    const bool ok = sellRelative(a.n, off, scol, rel, base);
    if (!ok) {
        CHECK(rel.empty() && base.empty(), "%s: refused but arrays kept", name);
        std::printf("%s: n %lld refused\n", name, (long long)a.n);
        return false;
    }
    const int64_t ns = (int64_t)off.size() - 1;
    CHECK(rel.size() == scol.size() && (int64_t)base.size() == off[ns] / 64, "%s: sizes", name);
    for (int64_t s = 0; s < ns; ++s) {  // sellPos maps a slice's slots onto its storage one to one
        const int w = (off[s + 1] - off[s]) / 64;
        std::vector<char> seen(64 * (size_t)w, 0);
        for (int k = 0; k < w; ++k)
            for (int lane = 0; lane < 64; ++lane) {
                const int64_t q = sellPos(off[s], w, k, lane) - off[s];
                CHECK(q >= 0 && q < 64 * w && !seen[q], "%s: slice %lld slot %d lane %d", name, (long long)s, k, lane);
                if (q >= 0 && q < 64 * w) seen[q] = 1;
            }
    }
    int64_t checked = 0;
    for (int64_t s = 0; s < ns; ++s) {
        const int w = (off[s + 1] - off[s]) / 64;
        for (int lane = 0; lane < 64; ++lane) {
            const int64_t row = s * 64 + lane;
            if (row >= a.n) continue;
            for (int k = 0; k < w; ++k) {
                const int64_t pos = sellPos(off[s], w, k, lane);
                const int64_t col = row + base[off[s] / 64 + k] + rel[pos];
                CHECK(col == scol[pos], "%s: row %lld slot %d: %lld != %d", name, (long long)row, k, (long long)col,
                      scol[pos]);
                ++checked;
            }
        }
    }
    // every CSR entry is where csr2sell says
    for (int64_t i = 0; i < a.n; ++i)
        for (int32_t e = a.rp[i]; e < a.rp[i + 1]; ++e) CHECK(scol[c2s[e]] == a.col[e], "%s: csr2sell at %d", name, e);
    std::printf("%s: n %lld, %lld slots checked\n", name, (long long)a.n, (long long)checked);
    return true;
}

}  // namespace

int main() {
    std::mt19937 rng(11);
    CHECK(check("tet box 8x8x4", tetBox(8, 8, 4)), "small tet box refused");
    // z neighbours 6 * 80 * 80 = 38400 rows away: beyond 16-bit offsets from the row, within a slot's span
    CHECK(check("tet box 80x80x3", tetBox(80, 80, 3)), "tet box with z offset 38400 refused");
    CHECK(check("random far 40000", randomFar(200000, 40000, rng)), "rows with neighbours 40000 away refused");
    // +-70000 in the same slot column: a span above 2^16 must be refused
    std::vector<std::set<int64_t>> rows(150000);
    for (int64_t i = 0; i < (int64_t)rows.size(); ++i) {
        rows[i].insert(i);
        if (i % 2 == 0 && i + 70000 < (int64_t)rows.size()) rows[i].insert(i + 70000);
        if (i % 2 == 1 && i - 70000 >= 0) rows[i].insert(i - 70000);
    }
    CHECK(!check("alternating +-70000", fromRows(rows)), "a slot spanning 140000 was accepted");
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
