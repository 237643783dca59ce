// host_mesh.cpp -- geometry and connectivity preparation on the host (once per mesh).
// Geometry follows the reference's formulas and floating-point order exactly
// (mesh/face.cpp:5-41, mesh/cell.cpp:7-41, mesh/utilities.cpp:15-40), so the device sees
// the same face/cell quantities the reference computes.
#include "host_mesh.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <numeric>
#include <sstream>
#include <unordered_map>

#include "internal.h"

namespace prism_hip {

namespace {

struct P3 {
    double x, y, z;
};
inline P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 operator*(P3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline P3 operator*(double s, P3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline P3 operator/(P3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(P3 a, P3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline double norm(P3 a) { return std::sqrt(dot(a, a)); }
inline P3 cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct FaceGeo {
    double area;
    P3 normal, center;
};

// face.cpp:5-41 -- triangle fan about the vertex average
FaceGeo faceGeometry(const double* V, const int64_t* vid, int64_t nv) {
    P3 geo{0, 0, 0};
    for (int64_t i = 0; i < nv; ++i) geo = geo + P3{V[3 * vid[i]], V[3 * vid[i] + 1], V[3 * vid[i] + 2]};
    geo = geo / (double)nv;
    P3 nrm{0, 0, 0}, ctr{0, 0, 0};
    double area = 0;
    for (int64_t i = 0; i < nv; ++i) {
        const int64_t a = vid[i], b = vid[(i + 1) % nv];
        P3 v1{V[3 * a], V[3 * a + 1], V[3 * a + 2]};
        P3 v2{V[3 * b], V[3 * b + 1], V[3 * b + 2]};
        P3 sc = ((v1 + v2) + geo) / 3.0;
        P3 sn = cross(v2 - v1, geo - v1);
        double sa = norm(sn) / 2.0;
        nrm = nrm + sn;
        area += sa;
        ctr = ctr + sc * sa;
    }
    ctr = ctr / area;
    double nn = norm(nrm);
    return {area, nrm / nn, ctr};
}

int8_t dummy_;

[[noreturn]] void fail(int code, const std::string& m) { throw Error(code, m); }

}  // namespace

void buildSell(int64_t n, const int32_t* rp, const int32_t* col, std::vector<int32_t>& off,
               std::vector<int32_t>& scol, std::vector<int32_t>& csr2sell) {
    const int64_t ns = (n + 63) / 64;
    off.assign(ns + 1, 0);
    for (int64_t s = 0; s < ns; ++s) {
        int32_t w = 0;
        for (int64_t r = s * 64; r < std::min<int64_t>(n, s * 64 + 64); ++r) w = std::max(w, rp[r + 1] - rp[r]);
        int64_t next = (int64_t)off[s] + (int64_t)w * 64;
        if (next >= (1LL << 31)) fail(PRISM_ERR_RUNTIME, "matrix too large for 32-bit sliced-ELL offsets");
        off[s + 1] = (int32_t)next;
    }
    scol.assign(off[ns], 0);
    csr2sell.assign(rp[n], 0);
    for (int64_t s = 0; s < ns; ++s) {
        int32_t w = (off[s + 1] - off[s]) / 64;
        for (int lane = 0; lane < 64; ++lane) {
            int64_t row = s * 64 + lane;
            int32_t len = row < n ? rp[row + 1] - rp[row] : 0;
            for (int32_t k = 0; k < w; ++k) {
                int32_t pos = (int32_t)sellPos(off[s], w, k, lane);
                if (k < len) {
                    scol[pos] = col[rp[row] + k];
                    csr2sell[rp[row] + k] = pos;
                } else {
                    scol[pos] = row < n ? (int32_t)row : 0;
                }
            }
        }
    }
}

void buildCsc(int64_t n, const int32_t* rp, const int32_t* col, std::vector<int32_t>& ptr,
              std::vector<int32_t>& row, std::vector<int32_t>& pos) {
    const int64_t nnz = rp[n];
    ptr.assign(n + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) ++ptr[col[k] + 1];
    for (int64_t j = 0; j < n; ++j) ptr[j + 1] += ptr[j];
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    row.assign(nnz, 0);
    pos.assign(nnz, 0);
    for (int64_t i = 0; i < n; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            int32_t d = fill[col[k]]++;
            row[d] = (int32_t)i;
            pos[d] = k;
        }
}

void buildSgsLayout(int64_t n64, const int32_t* rp, const int32_t* col, const int32_t* dg, int max_colours,
                    SgsLayout& L) {
    if (n64 <= 0 || n64 >= (1LL << 31)) fail(PRISM_ERR_INVALID, "SGS preconditioner: matrix size");
    const int n = (int)n64;
    for (int i = 0; i < n; ++i)
        if (dg[i] < 0) fail(PRISM_ERR_INVALID, "SGS preconditioner: row " + std::to_string(i) + " has no diagonal entry");
    // transpose of the pattern, for the symmetrised adjacency
    std::vector<int32_t> tp(n + 1, 0), tr(rp[n]);
    for (int k = 0; k < rp[n]; ++k) ++tp[col[k] + 1];
    for (int i = 0; i < n; ++i) tp[i + 1] += tp[i];
    {
        std::vector<int32_t> fill(tp.begin(), tp.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int k = rp[i]; k < rp[i + 1]; ++k) tr[fill[col[k]]++] = i;
    }
    std::vector<int32_t> colour(n, -1), mark(max_colours + 1, -1);
    int nc = 0;
    for (int i = 0; i < n; ++i) {
        auto see = [&](int j) {
            if (j != i && colour[j] >= 0) mark[colour[j]] = i;
        };
        for (int k = rp[i]; k < rp[i + 1]; ++k) see(col[k]);
        for (int k = tp[i]; k < tp[i + 1]; ++k) see(tr[k]);
        int c = 0;
        while (c < max_colours && mark[c] == i) ++c;
        if (c == max_colours)
            fail(PRISM_ERR_RUNTIME, "SGS preconditioner: the pattern needs more than " + std::to_string(max_colours) +
                                        " colours");
        colour[i] = c;
        nc = std::max(nc, c + 1);
    }
    L.nc = nc;
    L.cptr.assign(nc + 1, 0);
    for (int i = 0; i < n; ++i) ++L.cptr[colour[i] + 1];
    for (int c = 0; c < nc; ++c) L.cptr[c + 1] += L.cptr[c];
    L.perm.assign(n, 0);
    std::vector<int32_t> pos(n);
    {
        std::vector<int32_t> fill(L.cptr.begin(), L.cptr.end() - 1);
        for (int i = 0; i < n; ++i) {
            pos[i] = fill[colour[i]]++;
            L.perm[pos[i]] = i;
        }
    }
    const int ns = (n + 63) / 64;
    L.soff.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
        int w = 0;
        for (int l = 0; l < 64 && 64 * s + l < n; ++l) {
            const int r = L.perm[64 * s + l];
            w = std::max(w, rp[r + 1] - rp[r] - 1);
        }
        if ((int64_t)L.soff[s] + 64 * (int64_t)w >= (1LL << 31))
            fail(PRISM_ERR_RUNTIME, "matrix too large for 32-bit sliced-ELL offsets");
        L.soff[s + 1] = L.soff[s] + 64 * w;
    }
    L.scol.assign(L.soff[ns], 0);
    L.emap.assign(L.soff[ns], -1);
    L.dmap.assign(n, 0);
    for (int s = 0; s < ns; ++s) {
        const int w = (L.soff[s + 1] - L.soff[s]) / 64;
        for (int l = 0; l < 64; ++l) {
            const int q = 64 * s + l;
            int m = 0;
            if (q < n) {
                const int r = L.perm[q];
                L.dmap[q] = dg[r];
                for (int k = rp[r]; k < rp[r + 1]; ++k) {
                    if (col[k] == r) continue;
                    L.scol[L.soff[s] + l + 64 * m] = pos[col[k]];
                    L.emap[L.soff[s] + l + 64 * m] = k;
                    ++m;
                }
            }
            for (; m < w; ++m) L.scol[L.soff[s] + l + 64 * m] = q < n ? q : 0;
        }
    }
}

bool buildDia(int64_t n, const int32_t* rp, const int32_t* col, int dmax, std::vector<int32_t>& off,
              std::vector<int32_t>& csr2dia) {
    off.clear();
    csr2dia.clear();
    const int64_t nnz = rp[n];
    std::vector<int32_t> seen;
    for (int64_t i = 0; i < n; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            int32_t d = col[k] - (int32_t)i;
            if (std::find(seen.begin(), seen.end(), d) == seen.end()) {
                if ((int)seen.size() == dmax) return false;
                seen.push_back(d);
            }
        }
    const int64_t slots = (int64_t)seen.size() * n;
    if (seen.empty() || slots >= (1LL << 31) || 4 * nnz < 3 * slots) return false;
    std::sort(seen.begin(), seen.end());
    csr2dia.resize(nnz);
    for (int64_t i = 0; i < n; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            int32_t d = (int32_t)(std::lower_bound(seen.begin(), seen.end(), col[k] - (int32_t)i) - seen.begin());
            csr2dia[k] = (int32_t)(d * n + i);
        }
    off = seen;
    return true;
}

bool sellRelative(int64_t n, const std::vector<int32_t>& off, const std::vector<int32_t>& scol,
                  std::vector<int16_t>& rel, std::vector<int32_t>& base) {
    const int64_t ns = (int64_t)off.size() - 1;
    rel.assign(scol.size(), 0);
    base.assign(ns > 0 ? off[ns] / 64 : 0, 0);
    for (int64_t s = 0; s < ns; ++s) {
        const int32_t w = (off[s + 1] - off[s]) / 64;
        const int64_t rows = std::min<int64_t>(64, n - s * 64);  // rows past the end are never read
        for (int32_t k = 0; k < w; ++k) {
            int64_t lo = std::numeric_limits<int64_t>::max(), hi = std::numeric_limits<int64_t>::min();
            for (int64_t lane = 0; lane < rows; ++lane) {
                const int64_t d = (int64_t)scol[sellPos(off[s], w, k, (int)lane)] - (s * 64 + lane);
                lo = std::min(lo, d);
                hi = std::max(hi, d);
            }
            if (rows <= 0) continue;
            if (hi - lo > 65535) {
                rel.clear();
                base.clear();
                return false;
            }
            const int64_t b = lo + 32768;  // lo .. hi maps into -32768 .. 32767
            base[off[s] / 64 + k] = (int32_t)b;
            for (int64_t lane = 0; lane < rows; ++lane) {
                const int64_t pos = sellPos(off[s], w, k, (int)lane);
                rel[pos] = (int16_t)((int64_t)scol[pos] - (s * 64 + lane) - b);
            }
        }
    }
    return true;
}

// Every index and offset array of a descriptor, before anything is read through them.
void validateDesc(const prism_mesh_desc& d) {
    const int64_t F = d.n_faces, n = d.n_cells, nv = d.n_vertices;
    if (!d.vertices || !d.face_vtx_off || !d.face_vtx || !d.face_owner || !d.face_neighbor || !d.face_patch ||
        !d.cell_face_off || !d.cell_faces)
        fail(PRISM_ERR_MESH, "mesh descriptor with a null array");
    if ((d.n_bcs > 0 && !d.bcs) || (d.n_fields > 0 && !d.fields) || (d.n_patches > 0 && !d.patch_names) ||
        d.n_bcs < 0 || d.n_fields < 0 || d.n_patches < 0)
        fail(PRISM_ERR_MESH, "mesh descriptor with a null or negative-length table");
    if (d.face_vtx_off[0] != 0) fail(PRISM_ERR_MESH, "face vertex offsets must start at 0");
    for (int64_t f = 0; f < F; ++f)
        if (d.face_vtx_off[f + 1] < d.face_vtx_off[f]) fail(PRISM_ERR_MESH, "face vertex offsets decrease");
    for (int64_t k = 0; k < d.face_vtx_off[F]; ++k)
        if (d.face_vtx[k] < 0 || d.face_vtx[k] >= nv) fail(PRISM_ERR_MESH, "face vertex id out of range");
    for (int64_t f = 0; f < F; ++f) {
        int64_t o = d.face_owner[f], q = d.face_neighbor[f];
        if (o < 0 || o >= n || q >= n) fail(PRISM_ERR_MESH, "face owner/neighbour out of range");
        if (q == o) fail(PRISM_ERR_MESH, "face whose owner is its own neighbour");
    }
    if (d.cell_face_off[0] != 0) fail(PRISM_ERR_MESH, "cell face offsets must start at 0");
    for (int64_t c = 0; c < n; ++c)
        if (d.cell_face_off[c + 1] <= d.cell_face_off[c]) fail(PRISM_ERR_MESH, "cell with no faces");
    for (int64_t k = 0; k < d.cell_face_off[n]; ++k)
        if (d.cell_faces[k] < 0 || d.cell_faces[k] >= F) fail(PRISM_ERR_MESH, "cell face id out of range");
}

void buildMesh(const prism_mesh_desc& d, HMesh& m) {
    if (d.n_cells <= 0 || d.n_faces <= 0 || d.n_vertices <= 0) fail(PRISM_ERR_MESH, "mesh has no cells/faces/vertices");
    if (d.n_patches > MAXP) fail(PRISM_ERR_RUNTIME, "more than 64 boundary patches");
    if (d.n_cells >= (1LL << 31) || d.n_faces >= (1LL << 30)) fail(PRISM_ERR_RUNTIME, "mesh too large for 32-bit indices");
    validateDesc(d);
    // patches + BCs (mesh/boundary.cpp:391-407): "empty" decided by the first field in
    // alphabetical order
    m.patches.assign(d.n_patches, HPatch{});
    for (int32_t p = 0; p < d.n_patches; ++p) m.patches[p].name = d.patch_names[p];
    for (int32_t k = 0; k < d.n_bcs; ++k) {
        const prism_bc_desc& b = d.bcs[k];
        if (b.patch < 0 || b.patch >= d.n_patches) fail(PRISM_ERR_RUNTIME, "boundary condition for unknown patch");
        BCEntry e;
        e.kind = b.kind ? b.kind : "";
        e.vk = b.value_kind;
        for (int i = 0; i < 3; ++i) e.v[i] = b.value[i];
        m.patches[b.patch].bcs[b.field] = e;
    }
    for (auto& p : m.patches) {
        if (p.bcs.empty())
            fail(PRISM_ERR_RUNTIME, "BoundaryPatch constructor for patch " + p.name +
                                        " was called with an empty field name to boundary condition map");
        p.empty = p.bcs.begin()->second.kind == "empty";
    }
    m.infos.clear();
    for (int32_t k = 0; k < d.n_fields; ++k) {
        HFieldInfo fi;
        fi.name = d.fields[k].name;
        fi.type = d.fields[k].type ? d.fields[k].type : "";
        if (d.fields[k].grad_scheme) { fi.grad = d.fields[k].grad_scheme; fi.has_grad = true; }
        m.infos.push_back(fi);
    }

    const int64_t F = d.n_faces, n = d.n_cells;
    std::vector<FaceGeo> geo(F);
    for (int64_t f = 0; f < F; ++f) {
        int64_t a = d.face_vtx_off[f], b = d.face_vtx_off[f + 1];
        if (b - a < 3) fail(PRISM_ERR_MESH, "face with fewer than 3 vertices");
        geo[f] = faceGeometry(d.vertices, d.face_vtx + a, b - a);
        int64_t o = d.face_owner[f], nb = d.face_neighbor[f];
        if (o < 0 || o >= n || nb >= n) fail(PRISM_ERR_MESH, "face owner/neighbour out of range");
        if (nb < 0 && (d.face_patch[f] < 0 || d.face_patch[f] >= d.n_patches))
            fail(PRISM_ERR_MESH, "boundary face that is not part of any boundary patch");
    }
    // cell.cpp:7-41 -- pyramids on the face-centre average
    m.n = n;
    m.cx.resize(n); m.cy.resize(n); m.cz.resize(n); m.vol.resize(n);
    for (int64_t c = 0; c < n; ++c) {
        int64_t a = d.cell_face_off[c], b = d.cell_face_off[c + 1];
        P3 g{0, 0, 0};
        for (int64_t k = a; k < b; ++k) g = g + geo[d.cell_faces[k]].center;
        g = g / (double)(b - a);
        double vol = 0;
        P3 ctr{0, 0, 0};
        for (int64_t k = a; k < b; ++k) {
            const FaceGeo& fg = geo[d.cell_faces[k]];
            double pv = std::abs(dot(fg.normal, fg.center - g) * fg.area) / 3;
            P3 pc = (fg.center * 0.75) + (g * 0.25);
            vol += pv;
            ctr = ctr + pv * pc;
        }
        ctr = ctr / vol;
        m.cx[c] = ctr.x; m.cy[c] = ctr.y; m.cz[c] = ctr.z; m.vol[c] = vol;
    }
    // permutation (pmesh.cpp:51-119), stable within each group
    m.old_to_new.assign(F, 0);
    int64_t nb = 0, nne = 0;
    for (int64_t f = 0; f < F; ++f)
        if (d.face_neighbor[f] < 0) { ++nb; if (!m.patches[d.face_patch[f]].empty) ++nne; }
    int64_t pne = 0, pe = nne, pi = nb;
    for (int64_t f = 0; f < F; ++f)
        if (d.face_neighbor[f] < 0) m.old_to_new[f] = m.patches[d.face_patch[f]].empty ? pe++ : pne++;
    for (int64_t f = 0; f < F; ++f)
        if (d.face_neighbor[f] >= 0) m.old_to_new[f] = pi++;
    m.F = F; m.nb = nb; m.nne = nne;
    auto rs = [&](std::vector<double>& v) { v.assign(F, 0.0); };
    rs(m.sx); rs(m.sy); rs(m.sz); rs(m.fx); rs(m.fy); rs(m.fz); rs(m.nx); rs(m.ny); rs(m.nz);
    rs(m.dx); rs(m.dy); rs(m.dz); rs(m.gc); rs(m.kx); rs(m.ky); rs(m.kz);
    m.owner.assign(F, -1); m.nbr.assign(F, -1); m.fpatch.assign(F, -1);
    for (int64_t f = 0; f < F; ++f) {
        int64_t g = m.old_to_new[f];
        const FaceGeo& fg = geo[f];
        P3 S{fg.area * fg.normal.x, fg.area * fg.normal.y, fg.area * fg.normal.z};  // face.h:18
        m.sx[g] = S.x; m.sy[g] = S.y; m.sz[g] = S.z;
        m.fx[g] = fg.center.x; m.fy[g] = fg.center.y; m.fz[g] = fg.center.z;
        m.nx[g] = fg.normal.x; m.ny[g] = fg.normal.y; m.nz[g] = fg.normal.z;
        int64_t o = d.face_owner[f], q = d.face_neighbor[f];
        m.owner[g] = (int32_t)o;
        m.nbr[g] = (int32_t)q;
        m.fpatch[g] = q < 0 ? d.face_patch[f] : -1;
        P3 xo{m.cx[o], m.cy[o], m.cz[o]};
        if (q >= 0) {
            P3 xn{m.cx[q], m.cy[q], m.cz[q]};
            P3 dd = xn - xo;
            m.dx[g] = dd.x; m.dy[g] = dd.y; m.dz[g] = dd.z;
            // geometricWeight(owner, neighbour, face) (utilities.cpp:28-40)
            P3 dCf = fg.center - xo, dfN = xn - fg.center;
            P3 ef = S / norm(S);
            m.gc[g] = dot(dCf, ef) / (dot(dCf, ef) + dot(dfN, ef));
            // Green-Gauss skew vector (green_gauss.cpp:20-23)
            P3 k = fg.center - (0.5 * (xo + xn));
            m.kx[g] = k.x; m.ky[g] = k.y; m.kz[g] = k.z;
        } else {
            P3 dd = fg.center - xo;
            m.dx[g] = dd.x; m.dy[g] = dd.y; m.dz[g] = dd.z;
        }
    }
    for (int64_t g = 0; g < nb; ++g) m.patches[m.fpatch[g]].faces.push_back((int32_t)g);
    // cell -> faces (local order), entry = face << 1 | (cell is the neighbour)
    m.cf_off.assign(n + 1, 0);
    for (int64_t c = 0; c < n; ++c) m.cf_off[c + 1] = m.cf_off[c] + (int32_t)(d.cell_face_off[c + 1] - d.cell_face_off[c]);
    m.cf.resize(m.cf_off[n]);
    for (int64_t c = 0; c < n; ++c)
        for (int64_t k = d.cell_face_off[c]; k < d.cell_face_off[c + 1]; ++k) {
            int64_t g = m.old_to_new[d.cell_faces[k]];
            bool isn = m.nbr[g] == c;
            if (!isn && m.owner[g] != c) fail(PRISM_ERR_MESH, "cell lists a face it does not own or neighbour");
            m.cf[m.cf_off[c] + (k - d.cell_face_off[c])] = (int32_t)((g << 1) | (isn ? 1 : 0));
        }
    // CSR pattern: diagonal + one entry per neighbouring cell, sorted columns (two cells that
    // share several faces have one entry, as the reference's summed triplets do)
    std::vector<int32_t> deg(n, 1);
    for (int64_t g = nb; g < F; ++g) { deg[m.owner[g]]++; deg[m.nbr[g]]++; }
    std::vector<int32_t> rp0(n + 1, 0);
    for (int64_t c = 0; c < n; ++c) rp0[c + 1] = rp0[c] + deg[c];
    if ((int64_t)rp0[n] != n + 2 * (F - nb)) fail(PRISM_ERR_MESH, "pattern size mismatch");
    std::vector<int32_t> col0(rp0[n], 0);
    {
        std::vector<int32_t> fill(rp0.begin(), rp0.end() - 1);
        for (int64_t c = 0; c < n; ++c) col0[fill[c]++] = (int32_t)c;
        for (int64_t g = nb; g < F; ++g) {
            col0[fill[m.owner[g]]++] = m.nbr[g];
            col0[fill[m.nbr[g]]++] = m.owner[g];
        }
    }
    m.rp.assign(n + 1, 0);
    m.col.clear();
    m.col.reserve(rp0[n]);
    for (int64_t c = 0; c < n; ++c) {
        std::sort(col0.begin() + rp0[c], col0.begin() + rp0[c + 1]);
        for (int32_t k = rp0[c]; k < rp0[c + 1]; ++k)
            if (k == rp0[c] || col0[k] != col0[k - 1]) m.col.push_back(col0[k]);
        m.rp[c + 1] = (int32_t)m.col.size();
    }
    m.diag.resize(n);
    for (int64_t c = 0; c < n; ++c)
        m.diag[c] = (int32_t)(std::lower_bound(m.col.begin() + m.rp[c], m.col.begin() + m.rp[c + 1], (int32_t)c) - m.col.begin());
    buildSell(n, m.rp.data(), m.col.data(), m.sell_off, m.sell_col, m.csr2sell);
    sellRelative(m.n, m.sell_off, m.sell_col, m.sell_rel, m.sell_base);
    buildCsc(n, m.rp.data(), m.col.data(), m.csc_ptr, m.csc_row, m.csc_pos);
    buildDia(n, m.rp.data(), m.col.data(), 8, m.dia_off, m.csr2dia);
    m.slot_on.assign(F, -1);
    m.slot_no.assign(F, -1);
    {
        const int32_t nnz = m.rp[n];
        std::vector<uint8_t> taken(nnz, 0);
        std::vector<std::pair<int32_t, int32_t>> extra;  // (entry slot, extra index), face order
        auto place = [&](int32_t slot) {
            if (!taken[slot]) {
                taken[slot] = 1;
                return slot;
            }
            const int32_t k = (int32_t)extra.size();
            extra.push_back({slot, k});
            return nnz + k;
        };
        for (int64_t g = nb; g < F; ++g) {
            int32_t o = m.owner[g], q = m.nbr[g];
            m.slot_on[g] = place((int32_t)(std::lower_bound(m.col.begin() + m.rp[o], m.col.begin() + m.rp[o + 1], q) -
                                           m.col.begin()));
            m.slot_no[g] = place((int32_t)(std::lower_bound(m.col.begin() + m.rp[q], m.col.begin() + m.rp[q + 1], o) -
                                           m.col.begin()));
        }
        m.n_extra = (int64_t)extra.size();
        if ((int64_t)nnz + m.n_extra >= (1LL << 31)) fail(PRISM_ERR_MESH, "too many coefficients");
        std::stable_sort(extra.begin(), extra.end(),
                         [](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& b) {
                             return a.first < b.first;
                         });
        // extras renumbered group by group, face order inside a group
        std::vector<int32_t> renum(extra.size());
        m.dup_slot.clear();
        m.dup_ptr.assign(1, 0);
        for (size_t k = 0; k < extra.size(); ++k) {
            if (k == 0 || extra[k].first != extra[k - 1].first) {
                if (k) m.dup_ptr.push_back((int32_t)k);
                m.dup_slot.push_back(extra[k].first);
            }
            renum[extra[k].second] = (int32_t)k;
        }
        if (!extra.empty()) m.dup_ptr.push_back((int32_t)extra.size());
        for (int64_t g = nb; g < F; ++g) {
            if (m.slot_on[g] >= nnz) m.slot_on[g] = nnz + renum[m.slot_on[g] - nnz];
            if (m.slot_no[g] >= nnz) m.slot_no[g] = nnz + renum[m.slot_no[g] - nnz];
        }
    }
    // least-squares pseudo-inverses (least_squares.cpp:12-61); rows follow Eigen's 3x3
    // cofactor inverse and 3x3 products with (a0b0 + a1b1) + a2b2.
    m.pinv.assign(9 * n, 0.0);
    for (int64_t c = 0; c < n; ++c) {
        double D[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        P3 xc{m.cx[c], m.cy[c], m.cz[c]};
        for (int32_t k = m.cf_off[c]; k < m.cf_off[c + 1]; ++k) {
            int32_t g = m.cf[k] >> 1;
            P3 r;
            if (g >= nb) {
                int32_t o = (m.cf[k] & 1) ? m.owner[g] : m.nbr[g];
                r = P3{m.cx[o], m.cy[o], m.cz[o]} - xc;
            } else {
                r = P3{m.fx[g], m.fy[g], m.fz[g]} - xc;
            }
            double wk = 1 / (norm(r) + EPS);
            P3 w = r * wk;
            double wv[3] = {w.x, w.y, w.z}, rv[3] = {r.x, r.y, r.z};
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) D[i][j] += wv[i] * rv[j];
        }
        double Dt[3][3], M[3][3], I[3][3], P[3][3];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Dt[i][j] = D[j][i];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) M[i][j] = (Dt[i][0] * D[0][j] + Dt[i][1] * D[1][j]) + Dt[i][2] * D[2][j];
        auto cof = [&](int i, int j) {
            int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            return M[i1][j1] * M[i2][j2] - M[i1][j2] * M[i2][j1];
        };
        double c0 = cof(0, 0), c1 = cof(1, 0), c2 = cof(2, 0);
        double det = (c0 * M[0][0] + c1 * M[1][0]) + c2 * M[2][0];
        double inv = 1.0 / det;
        I[1][2] = cof(2, 1) * inv; I[2][1] = cof(1, 2) * inv; I[2][2] = cof(2, 2) * inv;
        I[1][0] = cof(0, 1) * inv; I[1][1] = cof(1, 1) * inv; I[2][0] = cof(0, 2) * inv;
        I[0][0] = c0 * inv; I[0][1] = c1 * inv; I[0][2] = c2 * inv;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) P[i][j] = (I[i][0] * Dt[0][j] + I[i][1] * Dt[1][j]) + I[i][2] * Dt[2][j];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m.pinv[(3 * i + j) * n + c] = P[i][j];
    }
    // Green-Gauss sweep levels: level(c) = 1 + max level of lower-numbered neighbours
    std::vector<int32_t> lev(n, 0);
    int32_t L = 0;
    for (int64_t c = 0; c < n; ++c) {
        int32_t l = 0;
        for (int32_t k = m.cf_off[c]; k < m.cf_off[c + 1]; ++k) {
            int32_t g = m.cf[k] >> 1;
            if (g < nb) continue;
            int32_t o = (m.cf[k] & 1) ? m.owner[g] : m.nbr[g];
            if (o < c) l = std::max(l, lev[o] + 1);
        }
        lev[c] = l;
        L = std::max(L, l + 1);
    }
    m.gg_off.assign(L + 1, 0);
    for (int64_t c = 0; c < n; ++c) m.gg_off[lev[c] + 1]++;
    for (int32_t l = 0; l < L; ++l) m.gg_off[l + 1] += m.gg_off[l];
    m.gg_cells.resize(n);
    std::vector<int32_t> pos(m.gg_off.begin(), m.gg_off.end() - 1);
    for (int64_t c = 0; c < n; ++c) m.gg_cells[pos[lev[c]]++] = (int32_t)c;
    (void)dummy_;
    buildVertexCells(d, m);
}

// Vertices, the cells' own vertex lists and the vertex -> cell adjacency of ensureCache
// (export.cpp:45-69): cells visited in order, each pushing itself onto the list of every
// vertex it names, so every list is ascending.  Without the cells' vertex lists a cell names
// the distinct vertices of its faces.
void buildVertexCells(const prism_mesh_desc& d, HMesh& m) {
    const int64_t nv = d.n_vertices, n = d.n_cells;
    m.vtx.assign(d.vertices, d.vertices + 3 * nv);
    m.cell_vtx_off.clear();
    m.cell_vtx.clear();
    if (d.cell_vtx_off) {
        if (d.cell_vtx_off[0] != 0) fail(PRISM_ERR_MESH, "cell vertex offsets must start at 0");
        for (int64_t c = 0; c < n; ++c)
            if (d.cell_vtx_off[c + 1] < d.cell_vtx_off[c]) fail(PRISM_ERR_MESH, "cell vertex offsets decrease");
        m.cell_vtx_off.assign(d.cell_vtx_off, d.cell_vtx_off + n + 1);
        m.cell_vtx.assign(d.cell_vtx, d.cell_vtx + d.cell_vtx_off[n]);
        for (int64_t v : m.cell_vtx)
            if (v < 0 || v >= nv) fail(PRISM_ERR_MESH, "cell vertex id out of range");
    }
    std::vector<int64_t> ids;
    auto cellVertices = [&](int64_t c) {
        ids.clear();
        if (!m.cell_vtx_off.empty()) {
            ids.assign(m.cell_vtx.begin() + m.cell_vtx_off[c], m.cell_vtx.begin() + m.cell_vtx_off[c + 1]);
            return;
        }
        for (int64_t k = d.cell_face_off[c]; k < d.cell_face_off[c + 1]; ++k) {
            int64_t f = d.cell_faces[k];
            for (int64_t j = d.face_vtx_off[f]; j < d.face_vtx_off[f + 1]; ++j) {
                int64_t v = d.face_vtx[j];
                if (std::find(ids.begin(), ids.end(), v) == ids.end()) ids.push_back(v);
            }
        }
    };
    if (nv >= (1LL << 31)) fail(PRISM_ERR_RUNTIME, "too many vertices for 32-bit indices");
    std::vector<int64_t> cnt(nv + 1, 0);
    for (int64_t c = 0; c < n; ++c) {
        cellVertices(c);
        for (int64_t v : ids) cnt[v + 1]++;
    }
    for (int64_t v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
    if (cnt[nv] >= (1LL << 31)) fail(PRISM_ERR_RUNTIME, "vertex adjacency too large for 32-bit indices");
    m.v2c_off.assign(cnt.begin(), cnt.end());
    m.v2c.assign(m.v2c_off[nv], 0);
    std::vector<int32_t> at(m.v2c_off.begin(), m.v2c_off.end() - 1);
    for (int64_t c = 0; c < n; ++c) {
        cellVertices(c);
        for (int64_t v : ids) m.v2c[at[v]++] = (int32_t)c;
    }
}

// Normalised inverse-distance weights of cellToPoint (export.cpp:93-104), one per adjacency
// entry: w_i = 1 / max(|x_v - x_c|, eps), then w_i / sum_j w_j with the sum taken in list
// order; |d| = sqrt((dx dx + dy dy) + dz dz) as Eigen's norm() of a 3-vector.
void idwWeights(const HMesh& m, std::vector<double>& w) {
    const int64_t nv = (int64_t)m.v2c_off.size() - 1;
    const double eps = std::numeric_limits<double>::epsilon();
    w.assign(m.v2c.size(), 0.0);
    for (int64_t v = 0; v < nv; ++v) {
        int32_t a = m.v2c_off[v], b = m.v2c_off[v + 1];
        if (a == b)
            fail(PRISM_ERR_MESH, "Vertex " + std::to_string(v) + " has no adjacent cells; point data cannot be computed.");
        double sum = 0.0;
        for (int32_t i = a; i < b; ++i) {
            int32_t c = m.v2c[i];
            double dx = m.vtx[3 * v] - m.cx[c], dy = m.vtx[3 * v + 1] - m.cy[c], dz = m.vtx[3 * v + 2] - m.cz[c];
            double dist = std::sqrt((dx * dx + dy * dy) + dz * dz);
            w[i] = 1.0 / std::max(dist, eps);
            sum += w[i];
        }
        for (int32_t i = a; i < b; ++i) w[i] = w[i] / sum;
    }
}

// ----------------------------------------------------------------- mesh generators ----
void MeshGen::describe(prism_mesh_desc& d) {
    std::memset(&d, 0, sizeof(d));
    d.n_vertices = (int64_t)vertices.size() / 3;
    d.vertices = vertices.data();
    d.n_faces = (int64_t)face_owner.size();
    d.face_vtx_off = face_vtx_off.data();
    d.face_vtx = face_vtx.data();
    d.face_owner = face_owner.data();
    d.face_neighbor = face_neighbor.data();
    d.face_patch = face_patch.data();
    d.n_cells = n_cells;
    d.cell_face_off = cell_face_off.data();
    d.cell_faces = cell_faces.data();
    d.cell_vtx_off = cell_vtx_off.empty() ? nullptr : cell_vtx_off.data();
    d.cell_vtx = cell_vtx_off.empty() ? nullptr : cell_vtx.data();
    patch_name_ptrs.clear();
    for (auto& s : patch_names) patch_name_ptrs.push_back(s.c_str());
    d.n_patches = (int32_t)patch_names.size();
    d.patch_names = patch_name_ptrs.data();
}

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
inline double unit(uint64_t seed, uint64_t id, uint64_t axis) {  // uniform in [-1, 1)
    uint64_t h = splitmix64(seed * 0x100000001b3ULL ^ splitmix64(id * 4 + axis));
    return ((h >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
}

// Faces from cell-local face lists: first visitor owns (its local vertex order), second
// neighbours, as UnvToPMeshConverter::processFace (unv.cpp:208-261).  Faces are numbered
// in creation order; `where` gives the patch of a face that stays single-visited.
struct LocalFace {
    unsigned __int128 key;
    int64_t cell;
    int32_t local;
};

}  // namespace

void generateBox(const prism_box_params& p, MeshGen& g) {
    if (p.nx < 1 || p.ny < 1 || p.nz < 1) fail(PRISM_ERR_INVALID, "box dimensions must be >= 1");
    const int64_t nx = p.nx, ny = p.ny, nz = p.nz;
    const int64_t NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    g = MeshGen{};
    g.vertices.resize(3 * NX * NY * NZ);
    const double hx = (p.x1 - p.x0) / nx, hy = (p.y1 - p.y0) / ny, hz = (p.z1 - p.z0) / nz;
    for (int64_t k = 0; k < NZ; ++k)
        for (int64_t j = 0; j < NY; ++j)
            for (int64_t i = 0; i < NX; ++i) {
                int64_t id = i + NX * (j + NY * k);
                double x = p.x0 + (double)i * hx, y = p.y0 + (double)j * hy, z = p.z0 + (double)k * hz;
                if (i == nx) x = p.x1;
                if (j == ny) y = p.y1;
                if (k == nz) z = p.z1;
                if (p.jitter > 0) {
                    if (i > 0 && i < nx) x += p.jitter * hx * unit(p.seed, id, 0);
                    if (j > 0 && j < ny) y += p.jitter * hy * unit(p.seed, id, 1);
                    if (k > 0 && k < nz) z += p.jitter * hz * unit(p.seed, id, 2);
                }
                g.vertices[3 * id] = x;
                g.vertices[3 * id + 1] = y;
                g.vertices[3 * id + 2] = z;
            }
    auto vid = [&](int64_t i, int64_t j, int64_t k) { return i + NX * (j + NY * k); };
    const int64_t nhex = nx * ny * nz;
    const int per = p.tets ? 6 : 1;
    const int nlf = p.tets ? 4 : 6;
    const int64_t ncell = nhex * per;
    g.n_cells = ncell;
    // cell vertex lists
    std::vector<int64_t> cv((size_t)ncell * (p.tets ? 4 : 8));
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                int64_t h = i + nx * (j + ny * k);
                int64_t c[8] = {vid(i, j, k),         vid(i + 1, j, k),         vid(i + 1, j + 1, k),
                                vid(i, j + 1, k),     vid(i, j, k + 1),         vid(i + 1, j, k + 1),
                                vid(i + 1, j + 1, k + 1), vid(i, j + 1, k + 1)};
                if (!p.tets) {
                    for (int q = 0; q < 8; ++q) cv[8 * h + q] = c[q];
                    continue;
                }
                // Freudenthal split along c0-c6: paths 0 -> e_a -> e_a+e_b -> 6
                static const int T6[6][4] = {{0, 1, 2, 6}, {0, 1, 5, 6}, {0, 3, 2, 6},
                                             {0, 3, 7, 6}, {0, 4, 5, 6}, {0, 4, 7, 6}};
                static const double C[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                               {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
                for (int t = 0; t < 6; ++t) {
                    int a = T6[t][0], b = T6[t][1], cc = T6[t][2], dd = T6[t][3];
                    // positive orientation (outward tetra table faces, unv.cpp:34-44)
                    double u[3], v[3], w[3];
                    for (int q = 0; q < 3; ++q) {
                        u[q] = C[b][q] - C[a][q];
                        v[q] = C[cc][q] - C[a][q];
                        w[q] = C[dd][q] - C[a][q];
                    }
                    double det = u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) +
                                 u[2] * (v[0] * w[1] - v[1] * w[0]);
                    if (det < 0) std::swap(b, cc);
                    int64_t cid = 6 * h + t;
                    cv[4 * cid] = c[a];
                    cv[4 * cid + 1] = c[b];
                    cv[4 * cid + 2] = c[cc];
                    cv[4 * cid + 3] = c[dd];
                }
            }
    // local face tables (unv.cpp:19-44)
    static const int HEXF[6][4] = {{1, 2, 6, 5}, {0, 4, 7, 3}, {2, 3, 7, 6}, {0, 1, 5, 4}, {4, 5, 6, 7}, {0, 3, 2, 1}};
    static const int TETF[4][3] = {{0, 2, 1}, {1, 2, 3}, {0, 1, 3}, {0, 3, 2}};
    const int fv = p.tets ? 3 : 4;
    auto faceVerts = [&](int64_t c, int l, int64_t* out) {
        for (int q = 0; q < fv; ++q) out[q] = p.tets ? cv[4 * c + TETF[l][q]] : cv[8 * c + HEXF[l][q]];
    };
    std::vector<LocalFace> lf((size_t)ncell * nlf);
    for (int64_t c = 0; c < ncell; ++c)
        for (int l = 0; l < nlf; ++l) {
            int64_t v[4];
            faceVerts(c, l, v);
            std::sort(v, v + fv);
            unsigned __int128 key = ((unsigned __int128)v[0] << 84) | ((unsigned __int128)v[1] << 42) | (unsigned __int128)v[2];
            lf[(size_t)c * nlf + l] = {key, c, l};
        }
    std::vector<size_t> ord(lf.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
        if (lf[a].key != lf[b].key) return lf[a].key < lf[b].key;
        return a < b;
    });
    // pair: (owner slot, neighbour cell)
    struct FR {
        int64_t oslot;
        int64_t nb;
        int64_t nslot;
    };
    std::vector<FR> fr;
    fr.reserve(lf.size() / 2 + 1);
    for (size_t i = 0; i < ord.size();) {
        size_t j = i + 1;
        while (j < ord.size() && lf[ord[j]].key == lf[ord[i]].key) ++j;
        if (j - i > 2) fail(PRISM_ERR_MESH, "face shared by more than two cells");
        if (j - i == 2) fr.push_back({(int64_t)ord[i], lf[ord[i + 1]].cell, (int64_t)ord[i + 1]});
        else fr.push_back({(int64_t)ord[i], -1, -1});
        i = j;
    }
    // creation order: (owner cell, owner local face)
    std::sort(fr.begin(), fr.end(), [](const FR& a, const FR& b) { return a.oslot < b.oslot; });
    const int64_t F = (int64_t)fr.size();
    g.face_vtx_off.assign(F + 1, 0);
    g.face_vtx.resize((size_t)F * fv);
    g.face_owner.resize(F);
    g.face_neighbor.resize(F);
    g.face_patch.assign(F, -1);
    std::vector<int64_t> slot2face(lf.size(), -1);
    const double tol = 1e-9;
    for (int64_t f = 0; f < F; ++f) {
        int64_t c = fr[f].oslot / nlf;
        int l = (int)(fr[f].oslot % nlf);
        g.face_vtx_off[f + 1] = g.face_vtx_off[f] + fv;
        faceVerts(c, l, &g.face_vtx[(size_t)f * fv]);
        g.face_owner[f] = c;
        g.face_neighbor[f] = fr[f].nb;
        slot2face[fr[f].oslot] = f;
        if (fr[f].nb >= 0) {
            slot2face[fr[f].nslot] = f;
            continue;
        }
        // patch from the (unjittered) boundary plane the face lies on
        int64_t v0 = g.face_vtx[(size_t)f * fv];
        int64_t ii = v0 % NX, jj = (v0 / NX) % NY, kk = v0 / (NX * NY);
        bool allx0 = true, allx1 = true, ally0 = true, ally1 = true, allz0 = true, allz1 = true;
        for (int q = 0; q < fv; ++q) {
            int64_t vv = g.face_vtx[(size_t)f * fv + q];
            int64_t a = vv % NX, b = (vv / NX) % NY, cc = vv / (NX * NY);
            allx0 &= a == 0; allx1 &= a == nx; ally0 &= b == 0; ally1 &= b == ny; allz0 &= cc == 0; allz1 &= cc == nz;
        }
        (void)ii; (void)jj; (void)kk; (void)tol;
        int32_t pid = allx0 ? 0 : allx1 ? 1 : ally0 ? 2 : ally1 ? 3 : allz0 ? 4 : allz1 ? 5 : -1;
        if (pid < 0) fail(PRISM_ERR_MESH, "generator: boundary face off the box planes");
        g.face_patch[f] = pid;
    }
    g.cell_face_off.assign(ncell + 1, 0);
    g.cell_faces.resize((size_t)ncell * nlf);
    for (int64_t c = 0; c < ncell; ++c) {
        g.cell_face_off[c + 1] = g.cell_face_off[c] + nlf;
        for (int l = 0; l < nlf; ++l) g.cell_faces[(size_t)c * nlf + l] = slot2face[(size_t)c * nlf + l];
    }
    g.patch_names = {"left", "right", "bottom", "top", "back", "front"};
    const int nvc = p.tets ? 4 : 8;
    g.cell_vtx = cv;
    g.cell_vtx_off.resize(ncell + 1);
    for (int64_t c = 0; c <= ncell; ++c) g.cell_vtx_off[c] = c * nvc;
}

// ------------------------------------------------------------------- UNV reader ----
// Ideas-UNV datasets 2411 (nodes), 2412 (elements), 2467 (groups).  Node and element
// labels are mapped to file-order indices.  Boundary quads/triangles create faces in file
// order; cells visit their local faces (hex/tetra/wedge tables of unv.cpp:19-58): a face
// seen first is created and owned by the visiting cell, a pre-created boundary face is
// owned by its first visitor, a second visitor becomes the neighbour
// (UnvToPMeshConverter::processFace, unv.cpp:208-261).  Faces are emitted as the
// reference's PMesh sees them: boundary faces in owner-discovery order, then interior
// faces in creation order (pmesh.cpp:51-85 keeps both orders).
void readUnv(const std::string& path, MeshGen& g) {
    std::ifstream in(path);
    if (!in) fail(PRISM_ERR_MESH, "cannot open UNV file `" + path + "`");
    g = MeshGen{};
    std::string line;
    std::unordered_map<int64_t, int64_t> node_idx, elem_idx;
    struct Elem {
        int fe;
        std::vector<int64_t> nodes;
    };
    std::vector<Elem> elems;
    struct Group {
        std::string name;
        std::vector<int64_t> labels;
    };
    std::vector<Group> groups;
    auto num = [](std::string t) {
        for (auto& ch : t) if (ch == 'D' || ch == 'd') ch = 'E';
        return std::stod(t);
    };
    auto isDelim = [](const std::string& l) {
        std::string t = l;
        t.erase(0, t.find_first_not_of(" \t\r"));
        t.erase(t.find_last_not_of(" \t\r") + 1);
        return t == "-1";
    };
    while (std::getline(in, line)) {
        if (!isDelim(line)) continue;
        if (!std::getline(in, line)) break;
        int ds = std::atoi(line.c_str());
        if (ds == 2411) {
            while (std::getline(in, line) && !isDelim(line)) {
                std::istringstream h(line);
                int64_t label;
                h >> label;
                std::string c;
                if (!std::getline(in, c)) fail(PRISM_ERR_MESH, "truncated 2411 record");
                std::istringstream cs(c);
                std::string a, b, d;
                cs >> a >> b >> d;
                node_idx[label] = (int64_t)g.vertices.size() / 3;
                g.vertices.push_back(num(a));
                g.vertices.push_back(num(b));
                g.vertices.push_back(num(d));
            }
        } else if (ds == 2412) {
            while (std::getline(in, line) && !isDelim(line)) {
                std::istringstream h(line);
                int64_t label = 0, fe = 0, pp = 0, mp = 0, col = 0, nn = -1;
                h >> label >> fe >> pp >> mp >> col >> nn;
                if (!h || nn < 1 || nn > 64) fail(PRISM_ERR_MESH, "malformed 2412 element record");
                bool beam = fe == 11 || fe == 21 || fe == 22 || fe == 23 || fe == 24;
                if (beam && !std::getline(in, line)) fail(PRISM_ERR_MESH, "truncated 2412 record");
                Elem e;
                e.fe = (int)fe;
                while ((int64_t)e.nodes.size() < nn) {
                    if (!std::getline(in, line)) fail(PRISM_ERR_MESH, "truncated 2412 record");
                    std::istringstream ns(line);
                    int64_t x;
                    while (ns >> x) e.nodes.push_back(x);
                }
                elem_idx[label] = (int64_t)elems.size();
                elems.push_back(std::move(e));
            }
        } else if (ds == 2467) {
            while (std::getline(in, line) && !isDelim(line)) {
                std::istringstream h(line);
                int64_t v[8];
                for (auto& x : v) h >> x;
                Group gr;
                std::getline(in, gr.name);
                gr.name.erase(gr.name.find_last_not_of(" \t\r") + 1);
                gr.name.erase(0, gr.name.find_first_not_of(" \t"));
                int64_t cnt = v[7];
                while ((int64_t)gr.labels.size() < cnt) {
                    if (!std::getline(in, line)) fail(PRISM_ERR_MESH, "truncated 2467 record");
                    std::istringstream es(line);
                    int64_t type, tag, leaf, comp;
                    int64_t k = 0;
                    while (es >> type >> tag >> leaf >> comp) {
                        if (type == 8) gr.labels.push_back(tag);
                        else gr.labels.push_back(-1);
                        ++k;
                    }
                    if (k == 0) break;
                }
                bool any = false;
                for (auto l : gr.labels) any |= l >= 0;
                if (any) groups.push_back(std::move(gr));
            }
        } else {
            while (std::getline(in, line) && !isDelim(line)) {}
        }
    }
    if (g.vertices.empty()) fail(PRISM_ERR_MESH, "UnvToPMeshConverter(): file `" + path + "` has zero vertices");
    for (auto& e : elems)
        for (auto& nd : e.nodes) {
            auto it = node_idx.find(nd);
            if (it == node_idx.end()) fail(PRISM_ERR_MESH, "element references an unknown node");
            nd = it->second;
        }
    static const int HEXF[6][4] = {{1, 2, 6, 5}, {0, 4, 7, 3}, {2, 3, 7, 6}, {0, 1, 5, 4}, {4, 5, 6, 7}, {0, 3, 2, 1}};
    static const int TETF[4][3] = {{0, 2, 1}, {1, 2, 3}, {0, 1, 3}, {0, 3, 2}};
    static const int WEDF[5][4] = {{0, 2, 1, -1}, {3, 4, 5, -1}, {0, 1, 4, 3}, {0, 3, 5, 2}, {1, 2, 5, 4}};
    struct RF {
        std::vector<int64_t> v;
        int64_t owner{-1}, nbr{-1};
        bool boundary{false};
    };
    std::vector<RF> faces;
    std::map<std::vector<int64_t>, int64_t> lookup;
    std::vector<int64_t> bfaces, ifaces;
    std::map<int64_t, int64_t> elem2bface;  // element index -> face id
    std::vector<std::vector<int64_t>> cellFaces;
    auto keyOf = [](std::vector<int64_t> v) { std::sort(v.begin(), v.end()); return v; };
    for (size_t ei = 0; ei < elems.size(); ++ei) {
        const Elem& e = elems[ei];
        int fe = e.fe;
        bool tri = fe == 41 || fe == 51 || fe == 61 || fe == 74 || fe == 81 || fe == 91;
        bool quad = fe == 44 || fe == 54 || fe == 64 || fe == 71 || fe == 84 || fe == 94;
        // every face-table index below must name one of the element's own nodes
        const size_t need = tri ? 3 : quad ? 4 : fe == 115 ? 8 : fe == 111 ? 4 : fe == 112 ? 6 : 0;
        if (e.nodes.size() < need)
            fail(PRISM_ERR_MESH, "UNV element of type " + std::to_string(fe) + " with " +
                                     std::to_string(e.nodes.size()) + " nodes (needs " + std::to_string(need) + ")");
        if (tri || quad) {
            RF f;
            f.v = e.nodes;
            f.boundary = true;
            int64_t id = (int64_t)faces.size();
            lookup[keyOf(f.v)] = id;
            faces.push_back(std::move(f));
            elem2bface[(int64_t)ei] = id;
            continue;
        }
        std::vector<std::vector<int64_t>> lf;
        if (fe == 115) {
            for (auto& t : HEXF) lf.push_back({e.nodes[t[0]], e.nodes[t[1]], e.nodes[t[2]], e.nodes[t[3]]});
        } else if (fe == 111) {
            for (auto& t : TETF) lf.push_back({e.nodes[t[0]], e.nodes[t[1]], e.nodes[t[2]]});
        } else if (fe == 112) {
            for (auto& t : WEDF) {
                std::vector<int64_t> v;
                for (int q = 0; q < 4 && t[q] >= 0; ++q) v.push_back(e.nodes[t[q]]);
                lf.push_back(v);
            }
        } else if (fe == 11 || fe == 21 || fe == 22 || fe == 23 || fe == 24) {
            continue;  // lines
        } else {
            fail(PRISM_ERR_MESH, "UnvToPMeshConverter::process_cells(): Input UNV mesh contains an unsupported element type!");
        }
        int64_t cid = (int64_t)cellFaces.size();
        if (g.cell_vtx_off.empty()) g.cell_vtx_off.push_back(0);
        for (auto v : e.nodes) g.cell_vtx.push_back(v);
        g.cell_vtx_off.push_back((int64_t)g.cell_vtx.size());
        std::vector<int64_t> ids;
        for (auto& v : lf) {
            auto k = keyOf(v);
            auto it = lookup.find(k);
            if (it != lookup.end()) {
                RF& f = faces[it->second];
                if (f.owner >= 0) f.nbr = cid;
                else { f.owner = cid; bfaces.push_back(it->second); }
                ids.push_back(it->second);
            } else {
                RF f;
                f.v = v;
                f.owner = cid;
                int64_t id = (int64_t)faces.size();
                lookup[k] = id;
                faces.push_back(std::move(f));
                ifaces.push_back(id);
                ids.push_back(id);
            }
        }
        cellFaces.push_back(std::move(ids));
    }
    if (cellFaces.empty())
        fail(PRISM_ERR_MESH, "UnvToPMeshConverter::process_cells(): Input UNV mesh is either empty or flat two dimensional. Prism acceptes only 3D meshes");
    // patches from face groups
    std::vector<int32_t> patchOf(faces.size(), -1);
    for (auto& gr : groups) {
        int32_t pid = (int32_t)g.patch_names.size();
        bool used = false;
        for (auto l : gr.labels) {
            if (l < 0) continue;
            auto it = elem_idx.find(l);
            if (it == elem_idx.end()) continue;
            auto jt = elem2bface.find(it->second);
            if (jt == elem2bface.end()) continue;
            patchOf[jt->second] = pid;
            used = true;
        }
        if (used) g.patch_names.push_back(gr.name);
    }
    int64_t undefined = 0;
    for (auto& kv : elem2bface) if (patchOf[kv.second] < 0) ++undefined;
    if (undefined)
        fail(PRISM_ERR_MESH, "UnvToPMeshConverter::check_boundary_faces(): Input UNV mesh has " + std::to_string(undefined) +
                                 " boundary faces that are not part of any boundary patches. Please check your mesh.");
    // emit: boundary faces (owner-discovery order), then interior (creation order)
    std::vector<int64_t> order = bfaces;
    for (auto id : ifaces) {
        if (faces[id].nbr < 0) fail(PRISM_ERR_MESH, "cell face without neighbour or boundary element");
        order.push_back(id);
    }
    std::vector<int64_t> newid(faces.size(), -1);
    for (size_t k = 0; k < order.size(); ++k) newid[order[k]] = (int64_t)k;
    g.face_vtx_off.assign(1, 0);
    for (auto id : order) {
        const RF& f = faces[id];
        for (auto v : f.v) g.face_vtx.push_back(v);
        g.face_vtx_off.push_back((int64_t)g.face_vtx.size());
        g.face_owner.push_back(f.owner);
        g.face_neighbor.push_back(f.boundary ? -1 : f.nbr);
        g.face_patch.push_back(f.boundary ? patchOf[id] : -1);
    }
    g.n_cells = (int64_t)cellFaces.size();
    g.cell_face_off.assign(1, 0);
    for (auto& cf : cellFaces) {
        for (auto id : cf) {
            if (newid[id] < 0) fail(PRISM_ERR_MESH, "boundary element not attached to any cell");
            g.cell_faces.push_back(newid[id]);
        }
        g.cell_face_off.push_back((int64_t)g.cell_faces.size());
    }
}

}  // namespace prism_hip
