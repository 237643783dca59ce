// oracle_core.cpp -- FP64 CPU restatement of the reference's finite-volume path.
// TEST INFRASTRUCTURE ONLY (see oracle_core.h).  Citations: eigemx/prism src/prism/...
#include "oracle_core.h"

#include <algorithm>
#include <map>
#include <limits>
#include <numeric>

namespace oref {

// ------------------------------------------------------------------ 3x3 helpers ----
static f64 cof(const M3& a, int i, int j) {  // Eigen cofactor_3x3<i,j>
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return a.m[i1][j1] * a.m[i2][j2] - a.m[i1][j2] * a.m[i2][j1];
}
M3 inverse3(const M3& a) {
    f64 c0 = cof(a, 0, 0), c1 = cof(a, 1, 0), c2 = cof(a, 2, 0);
    f64 det = (c0 * a.m[0][0] + c1 * a.m[1][0]) + c2 * a.m[2][0];
    f64 inv = 1.0 / det;
    M3 r;
    r.m[1][2] = cof(a, 2, 1) * inv;
    r.m[2][1] = cof(a, 1, 2) * inv;
    r.m[2][2] = cof(a, 2, 2) * inv;
    r.m[1][0] = cof(a, 0, 1) * inv;
    r.m[1][1] = cof(a, 1, 1) * inv;
    r.m[2][0] = cof(a, 0, 2) * inv;
    r.m[0][0] = c0 * inv;
    r.m[0][1] = c1 * inv;
    r.m[0][2] = c2 * inv;
    return r;
}
M3 matmul(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.m[i][j] = (a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j]) + a.m[i][2] * b.m[2][j];
    return r;
}
M3 transpose(const M3& a) {
    M3 r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}

// --------------------------------------------------------------------- patches ----
static bool isComponent(const std::string& n) {  // mesh/boundary.cpp:334-346
    if (n.size() < 2) return false;
    char l = n.back();
    return (l == 'x' || l == 'y' || l == 'z') && n[n.size() - 2] == '_';
}
const BC& Patch::bc(const std::string& field) const {
    auto it = bcs.find(field);
    if (it == bcs.end() && isComponent(field)) it = bcs.find(field.substr(0, field.size() - 2));
    if (it == bcs.end())
        throw std::runtime_error("Boundary patch `" + name + "` does not have a field named `" + field + "`");
    return it->second;
}
V3 Patch::vectorBC(const std::string& field) const {
    auto it = bcs.find(field);
    if (it == bcs.end()) throw std::runtime_error("patch " + name + " lacks field " + field);
    if (it->second.vk != 2) throw std::runtime_error("patch " + name + " field " + field + " is not a vector");
    return it->second.v;
}
f64 Patch::scalarBC(const std::string& field) const {
    if (isComponent(field)) {
        V3 v = vectorBC(field.substr(0, field.size() - 2));
        char c = field.back();
        return c == 'x' ? v.x : (c == 'y' ? v.y : v.z);
    }
    auto it = bcs.find(field);
    if (it == bcs.end()) throw std::runtime_error("patch " + name + " lacks field " + field);
    if (it->second.vk != 1) throw std::runtime_error("patch " + name + " field " + field + " is not a scalar");
    return it->second.s;
}

// -------------------------------------------------------------------- geometry ----
void faceGeometry(const std::vector<V3>& pts, Face& f) {  // face.cpp:5-41
    size_t n = f.verts.size();
    V3 geo{};
    for (size_t i = 0; i < n; ++i) geo += pts[f.verts[i]];
    geo = geo / (f64)n;
    V3 nrm{}, ctr{};
    f64 area = 0;
    for (size_t i = 0; i < n; ++i) {
        const V3& v1 = pts[f.verts[i]];
        const V3& v2 = pts[f.verts[(i + 1) % n]];
        const V3& v3 = geo;
        V3 sc = ((v1 + v2) + v3) / 3.0;
        V3 sn = cross(v2 - v1, v3 - v1);
        f64 sa = norm(sn) / 2.0;
        nrm += sn;
        area += sa;
        ctr += sc * sa;
    }
    ctr = ctr / area;
    f64 nn = norm(nrm);
    f.normal = nrm / nn;
    f.area = area;
    f.center = ctr;
}

void cellGeometry(const std::vector<Face>& faces, Cell& c) {  // cell.cpp:7-41
    V3 geo{};
    for (auto fid : c.faces) geo += faces[fid].center;
    geo = geo / (f64)c.faces.size();
    f64 vol = 0;
    V3 ctr{};
    for (auto fid : c.faces) {
        const Face& f = faces[fid];
        f64 pv = std::abs(dot(f.normal, f.center - geo) * f.area) / 3;
        V3 pc = (f.center * 0.75) + (geo * 0.25);
        vol += pv;
        ctr += pv * pc;
    }
    c.center = ctr / vol;
    c.volume = vol;
}

V3 outwardS(const Face& f, int64_t cell) { return f.owner == cell ? f.S() : -f.S(); }

f64 geometricWeight(const Mesh& m, int64_t c, int64_t n, const Face& f) {  // utilities.cpp:28-40
    V3 dCf = f.center - m.cells[c].center;
    V3 dfN = m.cells[n].center - f.center;
    V3 Sf = outwardS(f, c);
    V3 ef = Sf / norm(Sf);
    return dot(dCf, ef) / (dot(dCf, ef) + dot(dfN, ef));
}

// ------------------------------------------------------------------- gradients ----
V3 Gradient::atCell(const Scalar& f, int64_t c) {  // green_gauss.cpp:26-32, least_squares.cpp:63-69
    if (f.event != computed) {
        computeAll(f);
        computed = f.event;
    }
    return g[c];
}

V3 Gradient::atFace(const Scalar& f, const Face& face, int64_t fid) {  // gradient.cpp:15-34
    if (face.interior()) {
        V3 go = atCell(f, face.owner);
        V3 gn = atCell(f, face.neighbor);
        f64 gc = geometricWeight(*f.mesh, face.owner, face.neighbor, face);
        return (gc * go) + ((1. - gc) * gn);
    }
    return atBoundaryFace(f, face, fid);
}

V3 Gradient::atBoundaryFace(const Scalar& f, const Face& face, int64_t fid) {  // gradient/boundary.cpp
    const Patch& p = f.mesh->patchOf(face);
    const std::string& k = p.bc(f.name).kind;
    if (k == "symmetry" || k == "outlet" || k == "zero-gradient") return atCell(f, face.owner);
    if (k == "fixed" || k == "no-slip") {
        const Cell& o = f.mesh->cells[face.owner];
        V3 gcl = atCell(f, face.owner);
        V3 dCf = face.center - o.center;
        V3 e = dCf / norm(dCf);
        f64 pb = f.faceValue(fid);
        f64 pc = f.cellValue(face.owner);
        return gcl - (dot(gcl, e)) * e + (pb - pc) / norm(dCf) * e;
    }
    throw BCError("IGradient::gradAtBoundaryFace(): patch `" + p.name + "` kind `" + k + "`");
}

void GreenGauss::computeAll(const Scalar& f) {  // green_gauss.cpp:34-76 (sequential sweep)
    const Mesh& m = *f.mesh;
    for (int64_t c = 0; c < m.nc(); ++c) {
        V3 grad{0, 0, 0};
        for (auto fid : m.cells[c].faces) {
            const Face& face = m.faces[fid];
            if (!face.interior()) {
                const Patch& p = m.patchOf(face);
                if (p.empty) continue;  // boundaryFaceIntegral: empty -> 0 (+0 changes nothing)
                (void)p.bc(f.name);
                grad += f.faceValue(fid) * face.S();
                continue;
            }
            V3 Sf = outwardS(face, c);
            int64_t nb = m.other(c, face);
            f64 fphi = 0.5 * (f.cellValue(c) + f.cellValue(nb));
            V3 gsum = g[c] + g[nb];
            V3 vec = face.center - (0.5 * (m.cells[c].center + m.cells[nb].center));
            fphi += 0.5 * dot(gsum, vec);
            grad += Sf * fphi;
        }
        grad = grad / m.cells[c].volume;
        g[c] = grad;
    }
}

LeastSquares::LeastSquares(const Mesh& m) {  // least_squares.cpp:7-61
    g.assign(m.nc(), V3{});
    pinv.resize(m.nc());
    cache.resize(m.nc());
    for (int64_t c = 0; c < m.nc(); ++c) {
        M3 d;
        for (auto fid : m.cells[c].faces) {
            const Face& face = m.faces[fid];
            V3 r;
            if (face.interior()) {
                int64_t nb = m.other(c, face);
                r = m.cells[nb].center - m.cells[c].center;
                f64 wk = 1 / (norm(r) + EPS);
                V3 w = r * wk;
                for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) d.m[i][j] += w[i] * r[j];
                cache[c].push_back({true, nb, w});
            } else {
                r = face.center - m.cells[c].center;
                f64 wk = 1 / (norm(r) + EPS);
                V3 w = r * wk;
                for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) d.m[i][j] += w[i] * r[j];
                if (!m.patchOf(face).empty) cache[c].push_back({false, fid, w});
            }
        }
        M3 dt = transpose(d);
        pinv[c] = matmul(inverse3(matmul(dt, d)), dt);
    }
}

void LeastSquares::computeAll(const Scalar& f) {  // least_squares.cpp:78-100
    for (int64_t c = 0; c < f.mesh->nc(); ++c) {
        V3 b{0, 0, 0};
        f64 pc = f.cellValue(c);
        for (const auto& gf : cache[c]) {
            f64 o = gf.interior ? f.cellValue(gf.other) : f.faceValue(gf.other);
            b += gf.w * (o - pc);
        }
        g[c] = mul(pinv[c], b);
    }
}

// --------------------------------------------------------------------- fields ----
Scalar::Scalar(std::string n, std::shared_ptr<Mesh> m, std::vector<f64> vals, int kind_, int coord_)
    : name(std::move(n)), mesh(std::move(m)), v(std::move(vals)), kind(kind_), coord(coord_) {
    if ((int64_t)v.size() != mesh->nc()) throw std::runtime_error("Scalar: size != cell count");
    // scalar.cpp:260-304: component fields look up their parent's gradScheme
    std::string lookup = name;
    if (coord >= 0) lookup = name.substr(0, name.size() - 2);
    bool gg = false;
    for (const auto& fi : mesh->infos)
        if (fi.name == lookup && fi.grad.has_value()) {
            gg = (*fi.grad == "green-gauss" || *fi.grad == "greenGauss");
            break;
        }
    if (gg) grad = std::make_shared<GreenGauss>(*mesh);
    else grad = std::make_shared<LeastSquares>(*mesh);
}

f64 Scalar::faceValue(int64_t fid) const {  // scalar.cpp:189-237 + scalar_boundary.h
    if (hasFaceValues()) return fv[fid];
    const Face& f = mesh->faces[fid];
    if (f.interior()) {
        f64 gc = geometricWeight(*mesh, f.owner, f.neighbor, f);
        f64 val = gc * v[f.owner];
        val += (1 - gc) * v[f.neighbor];
        return val;
    }
    const Patch& p = mesh->patchOf(f);
    const std::string& k = p.bc(name).kind;
    if (k == "fixed" || k == "no-slip") return p.scalarBC(name);
    // scalar_boundary.h:64-77; field::Pressure keeps these default handlers (pressure.h:6-9):
    // Symmetry<Pressure> (pressure_boundary.cpp) is never registered
    if (k == "outlet" || k == "zero-gradient" || k == "symmetry") return v[f.owner];
    throw BCError("Scalar::valueAtBoundaryFace() field `" + name + "` patch `" + p.name + "` kind `" + k + "`");
}

void Scalar::updatePrev() {  // history.cpp:30-48
    if (hist_max == 0) return;
    if (hist.size() == hist_max) hist.pop_back();
    hist.insert(hist.begin(), v);
}

V3 Vector::boundaryValue(int64_t fid) const {  // vector_boundary.h:55-115
    const Face& f = mesh->faces[fid];
    const Patch& p = mesh->patchOf(f);
    const std::string& k = p.bc(name).kind;
    if (k == "fixed" || k == "no-slip") return p.vectorBC(name);
    if (k == "symmetry") {
        V3 uc = cellValue(f.owner);
        return uc - dot(uc, f.normal) * f.normal;
    }
    if (k == "outlet" || k == "zero-gradient") return cellValue(f.owner);
    throw BCError("GeneralVector::valueAtBoundaryFace() field `" + name + "` patch `" + p.name + "` kind `" + k + "`");
}

V3 Vector::faceValue(int64_t fid) const {
    if (has_fd) return fd[fid];
    const Face& f = mesh->faces[fid];
    if (!f.interior()) return boundaryValue(fid);
    return {c[0]->faceValue(fid), c[1]->faceValue(fid), c[2]->faceValue(fid)};
}

f64 Vector::flux(int64_t fid) const {
    if (has_ff) return ff[fid];
    const Face& f = mesh->faces[fid];
    if (f.interior()) return dot(faceValue(fid), f.S());
    const Patch& p = mesh->patchOf(f);
    const std::string& k = p.bc(name).kind;
    if (k == "no-slip" || k == "symmetry") return 0.0;
    if (k == "fixed" || k == "outlet" || k == "zero-gradient") return dot(faceValue(fid), f.S());
    throw BCError("GeneralVector::fluxAtBoundaryFace() field `" + name + "` patch `" + p.name + "` kind `" + k + "`");
}

void Vector::initFaceData() {  // vector.h:265-280
    std::vector<V3> fvals(mesh->nf());
    std::vector<f64> fflux(mesh->nf(), 0.0);
    for (const auto& p : mesh->patches) {
        if (p.empty) continue;
        for (auto fid : p.faces) {
            fvals[fid] = faceValue(fid);
            fflux[fid] = flux(fid);
        }
    }
    fd = std::move(fvals);
    ff = std::move(fflux);
    has_fd = has_ff = true;
}

M3 Tensor::faceValue(int64_t fid) const {  // tensor.cpp:58-77
    const Face& f = mesh->faces[fid];
    if (!f.interior()) return v[f.owner];
    f64 gc = geometricWeight(*mesh, f.owner, f.neighbor, f);
    return add(scale(gc, v[f.owner]), scale(1 - gc, v[f.neighbor]));
}

// --------------------------------------------------------------------- pattern ----
int64_t CSR::slot(int64_t i, int64_t j) const {
    auto b = col.begin() + rp[i], e = col.begin() + rp[i + 1];
    auto it = std::lower_bound(b, e, (int32_t)j);
    if (it == e || *it != j) throw std::runtime_error("CSR slot missing");
    return it - col.begin();
}

std::shared_ptr<CSR> patternFor(const std::shared_ptr<Mesh>& m) {
    auto p = std::make_shared<CSR>();
    int64_t n = m->nc();
    p->n = n;
    std::vector<std::vector<int32_t>> rows(n);
    for (int64_t i = 0; i < n; ++i) rows[i].push_back((int32_t)i);
    for (const auto& f : m->faces)
        if (f.interior()) {
            rows[f.owner].push_back((int32_t)f.neighbor);
            rows[f.neighbor].push_back((int32_t)f.owner);
        }
    p->rp.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
        std::sort(rows[i].begin(), rows[i].end());
        rows[i].erase(std::unique(rows[i].begin(), rows[i].end()), rows[i].end());
        p->rp[i + 1] = p->rp[i] + (int64_t)rows[i].size();
    }
    p->col.resize(p->rp[n]);
    p->diag.resize(n);
    for (int64_t i = 0; i < n; ++i) {
        std::copy(rows[i].begin(), rows[i].end(), p->col.begin() + p->rp[i]);
        p->diag[i] = p->slot(i, i);
    }
    return p;
}

// --------------------------------------------------------------------- schemes ----
namespace {

inline f64 smax(f64 a, f64 b) { return (a < b) ? b : a; }  // std::max(a, b)

struct Inserter {  // LinearSystem::insert + collect: duplicates summed in insertion order
    const CSR& p;
    std::vector<f64>& A;
    void operator()(int64_t i, int64_t j, f64 v) { A[p.slot(i, j)] += v; }
};

struct NVF { f64 l, k; };
NVF nvf(int scheme, f64 t) {  // convection.cpp:69-128
    switch (scheme) {
        case 0: return {1.0, 0.0};
        case 1: return {0.5, 0.5};
        case 2: return {1.5, 0.0};
        case 3: return {0.75, 3.0 / 8.0};
        case 4: return {1.0, 0.25};
        case 5:
            if (t > 0.0 && t < 0.5) return {1.5, 0.0};
            if (t >= 0.5 && t < 1.0) return {0.5, 0.5};
            return {1.0, 0};
        case 6:
            if (t > 0.0 && t < 0.25) return {2, 0.0};
            if (t >= 0.25 && t < 0.75) return {1, 0.25};
            if (t >= 0.75 && t < 1.0) return {0, 1.0};
            return {1.0, 0};
        case 7:  // integer divisions 1/6, 5/6, 3/8 are 0 in the reference
            if (t > 0.0 && t < (f64)(1 / 6)) return {4, 0.0};
            if (t >= (f64)(1 / 6) && t < (f64)(5 / 6)) return {0.75, (f64)(3 / 8)};
            if (t >= (f64)(5 / 6) && t < 1.0) return {0, 1.0};
            return {1.0, 0};
    }
    throw std::runtime_error("unknown NVF scheme");
}

f64 dummyUpwind(const Scalar& phi, int64_t c, int64_t d) {  // convection_hr.cpp:6-13
    const Mesh& m = *phi.mesh;
    f64 pd = phi.cellValue(d);
    V3 dCD = m.cells[d].center - m.cells[c].center;
    return pd - (2 * dot(phi.gradCell(c), dCD));
}
f64 phiTilde(const Scalar& phi, int64_t c, int64_t d) {  // convection_hr.cpp:15-23
    f64 pu = dummyUpwind(phi, c, d);
    f64 a = phi.cellValue(c) - pu;
    f64 b = phi.cellValue(d) - pu;
    return a / b;
}

struct Coeffs { f64 aC, aN, b; };
Coeffs interpolate(int scheme, const Scalar& phi, f64 m, int64_t c, int64_t n) {  // convection.cpp:46-67
    f64 t = phiTilde(phi, c, n);
    f64 pu = dummyUpwind(phi, c, n);
    NVF p = nvf(scheme, t);
    f64 aC = smax(m, 0.0) * p.l;
    f64 aN = smax(m, 0.0) * p.k;
    f64 b = -smax(m, 0.0) * (1 - p.l - p.k) * pu;
    t = phiTilde(phi, n, c);
    pu = dummyUpwind(phi, n, c);
    NVF q = nvf(scheme, t);
    aC += -smax(-m, 0.0) * q.k;
    aN += -smax(-m, 0.0) * q.l;
    b += smax(-m, 0.0) * (1 - q.l - q.k) * pu;
    return {aC, aN, b};
}

V3 kappaFace(const Scheme& s, const V3& v, int64_t fid) {  // diffusion_coeff.cpp:16-28
    if (s.kappa) return s.kappa->faceValue(fid) * v;
    return mulT(s.kappa_t->faceValue(fid), v);
}
V3 kappaCell(const Scheme& s, const V3& v, int64_t c) {
    if (s.kappa) return s.kappa->cellValue(c) * v;
    return mulT(s.kappa_t->v[c], v);
}

struct Decomp { V3 E, T; };
Decomp decompose(int corr, const V3& S, const V3& e) {  // nonortho.cpp:7-21
    V3 E;
    if (corr == 1) E = ((dot(S, S) / (dot(e, S) + EPS))) * e;
    else if (corr == 2) E = (dot(e, S)) * e;
    else E = norm(S) * e;
    return {E, S - E};
}

void applyDiffusion(Scheme& s, const CSR& pat) {
    Mesh& m = *s.phi->mesh;
    Scalar& phi = *s.phi;
    Inserter ins{pat, s.sys.A};
    bool corrected = s.variant != 0;
    // applyBoundary (diffusion_boundary.h:113-234) -- patch order, face order
    for (const auto& p : m.patches) {
        if (p.empty) continue;
        const std::string& k = p.bc(phi.name).kind;
        if (k == "symmetry" || k == "outlet" || k == "zero-gradient") continue;
        if (k == "fixed-gradient") {
            for (auto fid : p.faces) {
                const Face& f = m.faces[fid];
                V3 wg = p.vectorBC(phi.name);
                V3 Sp = kappaCell(s, f.S(), f.owner);
                s.sys.b[f.owner] += -dot(wg, Sp);
            }
            continue;
        }
        if (k != "fixed" && k != "no-slip")
            throw BCError("applyBoundary() diffusion patch `" + p.name + "` kind `" + k + "`");
        for (auto fid : p.faces) {
            const Face& f = m.faces[fid];
            int64_t c = f.owner;
            f64 pw = phi.faceValue(fid);
            V3 dCf = f.center - m.cells[c].center;
            f64 dn = norm(dCf);
            V3 Sp = kappaCell(s, f.S(), c);
            if (!corrected) {
                f64 g = dot(Sp, dCf) / (dn * dn + EPS);
                ins(c, c, g);
                s.sys.b[c] += g * pw;
            } else {
                V3 e = dCf / dn;
                Decomp d = decompose(s.variant, Sp, e);
                f64 g = dot(d.E, dCf) / (dn * dn + EPS);
                ins(c, c, g);
                s.sys.b[c] += (g * pw) + dot(d.T, phi.gradFace(fid));
            }
        }
    }
    // applyInterior (diffusion.h:116-149, 171-205)
    for (int64_t fid = m.n_boundary; fid < m.nf(); ++fid) {
        const Face& f = m.faces[fid];
        int64_t o = f.owner, n = f.neighbor;
        V3 dCF = m.cells[n].center - m.cells[o].center;
        f64 dn = norm(dCF);
        if (!corrected) {
            V3 Sp = kappaFace(s, f.S(), fid);
            f64 g = dot(Sp, dCF) / (dn * dn + EPS);
            ins(o, o, g); ins(n, n, g); ins(o, n, -g); ins(n, o, -g);
        } else {
            V3 e = dCF / dn;
            V3 Sp = kappaFace(s, outwardS(f, o), fid);
            Decomp d = decompose(s.variant, Sp, e);
            f64 g = dot(d.E, dCF) / (dn * dn + EPS);
            ins(o, o, g); ins(n, n, g); ins(o, n, -g); ins(n, o, -g);
            V3 gf = phi.gradFace(fid);
            s.sys.b[o] += dot(d.T, gf);
            s.sys.b[n] += -dot(d.T, gf);
        }
    }
}

void applyConvection(Scheme& s, const CSR& pat) {
    Mesh& m = *s.phi->mesh;
    Scalar& phi = *s.phi;
    Vector& U = *s.vec;
    Inserter ins{pat, s.sys.A};
    // convection_boundary.cpp:4-61
    for (const auto& p : m.patches) {
        if (p.empty) continue;
        const std::string& k = p.bc(phi.name).kind;
        if (k == "no-slip" || k == "symmetry") continue;
        if (k == "fixed") {
            for (auto fid : p.faces) {
                const Face& f = m.faces[fid];
                f64 pw = p.scalarBC(phi.name);
                f64 md = U.flux(fid);
                ins(f.owner, f.owner, smax(md, 0.0));
                s.sys.b[f.owner] += smax(-md, 0.0) * pw;
            }
            continue;
        }
        if (k == "outlet" || k == "zero-gradient") {
            for (auto fid : p.faces) {
                const Face& f = m.faces[fid];
                V3 uf = U.faceValue(fid);
                ins(f.owner, f.owner, dot(uf, f.S()));
            }
            continue;
        }
        throw BCError("applyBoundary() convection patch `" + p.name + "` kind `" + k + "`");
    }
    for (int64_t fid = m.n_boundary; fid < m.nf(); ++fid) {  // convection.cpp:18-40
        const Face& f = m.faces[fid];
        int64_t o = f.owner, n = f.neighbor;
        f64 md = U.flux(fid);
        Coeffs a = interpolate(s.variant, phi, md, o, n);
        Coeffs x = interpolate(s.variant, phi, -md, n, o);
        ins(o, o, a.aC); ins(o, n, a.aN);
        ins(n, n, x.aC); ins(n, o, x.aN);
        s.sys.b[o] += a.b;
        s.sys.b[n] += x.b;
    }
}

f64 fluxSum(const Vector& U, int64_t c) {  // operations.h:144-169
    const Mesh& m = *U.mesh;
    f64 sum = 0.0;
    for (auto fid : m.cells[c].faces) {
        const Face& f = m.faces[fid];
        if (!f.interior() && m.patchOf(f).empty) continue;
        f64 fl = U.flux(fid);
        if (f.owner != c) fl = -fl;
        sum += fl;
    }
    return sum;
}

}  // namespace

void applyScheme(Scheme& s, const CSR& pat) {
    const Mesh& m = *((s.kind == SK::SrcDivergence) ? s.vec->mesh : s.phi->mesh);
    int64_t n = m.nc();
    switch (s.kind) {
        case SK::Diffusion:
            std::fill(s.sys.A.begin(), s.sys.A.end(), 0.0);
            applyDiffusion(s, pat);
            break;
        case SK::Convection:
            std::fill(s.sys.A.begin(), s.sys.A.end(), 0.0);
            applyConvection(s, pat);
            break;
        case SK::SrcConstant: {  // constant_scalar.h:27-37
            s.phi->bump();  // non-const values()
            for (int64_t i = 0; i < n; ++i) {
                s.sys.b[i] = s.phi->v[i] * m.vol[i];
                if (s.variant == 1) s.sys.b[i] = -s.sys.b[i];
            }
            break;
        }
        case SK::SrcGradient: {  // source/gradient.h:32-40
            for (int64_t i = 0; i < n; ++i) {
                f64 g = s.phi->gradCell(i)[s.coord];
                s.sys.b[i] = g * m.vol[i];
                if (s.variant == 1) s.sys.b[i] = -s.sys.b[i];
            }
            break;
        }
        case SK::SrcDivergence: {  // divergence.h:25-34
            for (int64_t i = 0; i < n; ++i) {
                f64 d = fluxSum(*s.vec, i) / m.cells[i].volume;
                f64 v = d * m.vol[i];
                s.sys.b[i] = s.variant == 0 ? v : -v;
            }
            break;
        }
        case SK::SrcImplicit: {  // implicit_field.h:23-32
            std::fill(s.sys.A.begin(), s.sys.A.end(), 0.0);
            for (int64_t i = 0; i < n; ++i) {
                f64 d = s.coeff * m.vol[i];
                s.sys.A[pat.diag[i]] = s.variant == 0 ? d * -1 : d;
            }
            break;
        }
        case SK::BackwardEuler: {  // backward_euler.cpp:62-83
            const std::vector<f64>* prev = s.phi->prev();
            if (!prev) throw std::runtime_error("BackwardEuler::apply(): field has no previous time step values");
            std::fill(s.sys.A.begin(), s.sys.A.end(), 0.0);
            for (int64_t i = 0; i < n; ++i) {
                s.sys.A[pat.diag[i]] = m.vol[i] / s.coeff;
                s.sys.b[i] = m.vol[i] * (*prev)[i] / s.coeff;
            }
            break;
        }
        case SK::AdamMoulton: {  // adam_moulton.cpp:55-86 (BDF2; Backward-Euler until two levels exist)
            const std::vector<f64>* prev = s.phi->prev();
            if (!prev) throw std::runtime_error("AdamMoulton::apply(): field has no previous time step values");
            const std::vector<f64>* pp = s.phi->prevPrev();
            std::fill(s.sys.A.begin(), s.sys.A.end(), 0.0);
            for (int64_t i = 0; i < n; ++i) {
                if (!pp) {
                    s.sys.A[pat.diag[i]] = m.vol[i] / s.coeff;
                    s.sys.b[i] = m.vol[i] * (*prev)[i] / s.coeff;
                } else {
                    s.sys.A[pat.diag[i]] = (3.0 / 2.0) * m.vol[i] / s.coeff;
                    f64 r = 2.0 * (m.vol[i] * (*prev)[i]) / s.coeff;
                    r += -0.5 * (m.vol[i] * (*pp)[i]) / s.coeff;
                    s.sys.b[i] = r;
                }
            }
            break;
        }
    }
}

// -------------------------------------------------------------------- equation ----
void Equation::updateCoeffs() {  // transport.cpp:4-21
    for (auto& s : schemes) {
        applyScheme(s, *pattern);
        for (size_t k = 0; k < sys.A.size(); ++k) sys.A[k] += s.sys.A[k];
        for (size_t k = 0; k < sys.b.size(); ++k) sys.b[k] += s.sys.b[k];
    }
    for (auto& s : sources) {
        applyScheme(s, *pattern);
        for (size_t k = 0; k < sys.b.size(); ++k) sys.b[k] += s.sys.b[k];
    }
    if (symmetry) applySymmetry();  // applyBoundaryIfExists (boundary.h:128-159)
}

// Symmetry-plane momentum constraint of symmetry.cpp:12-82, patch by patch in mesh order:
// g = 2 |kappa_C^T S_f| / ((x_f - x_C) . n), diagonal += g n_k^2, rhs -= g (U_C . n - U_k n_k) n_k.
void Equation::applySymmetry() {
    const Scheme* conv = nullptr;
    const Scheme* diff = nullptr;
    for (const auto& s : schemes) {  // Transport keeps the last one added (transport.h:103-116)
        if (s.kind == SK::Convection) conv = &s;
        if (s.kind == SK::Diffusion) diff = &s;
    }
    if (!conv || !diff) throw std::runtime_error("eqn::boundary::Symmetry<Transport>: equation needs a convection and a diffusion scheme");
    if (phi->coord < 0) throw std::runtime_error("eqn::boundary::Symmetry<Transport>: field `" + phi->name + "` is not a vector component");
    const Mesh& m = *phi->mesh;
    const int64_t n = (int64_t)m.nc();
    for (const auto& p : m.patches) {
        if (p.empty || p.bc(phi->name).kind != "symmetry") continue;
        std::vector<f64> dA(n, 0.0), dB(n, 0.0);
        std::vector<char> hit(n, 0);
        for (int64_t fid : p.faces) {
            const Face& f = m.faces[fid];
            const int64_t o = f.owner;
            const V3 nn = f.normal;
            const V3 dCb = f.center - m.cells[o].center;
            const f64 dn = dot(dCb, nn);
            const V3 Uc = conv->vec->cellValue(o);
            const f64 g = 2 * norm(kappaCell(*diff, f.S(), o)) / dn;
            f64 ac, b;
            if (phi->coord == 0) { ac = nn.x * nn.x; b = ((Uc.y * nn.y) + (Uc.z * nn.z)) * nn.x; }
            else if (phi->coord == 1) { ac = nn.y * nn.y; b = ((Uc.x * nn.x) + (Uc.z * nn.z)) * nn.y; }
            else { ac = nn.z * nn.z; b = ((Uc.x * nn.x) + (Uc.y * nn.y)) * nn.z; }
            dA[o] = hit[o] ? dA[o] + g * ac : g * ac;  // setFromTriplets sums in insertion order
            hit[o] = 1;
            dB[o] += -g * b;
        }
        for (int64_t c = 0; c < n; ++c) {
            if (hit[c]) sys.A[pattern->diag[c]] += dA[c];
            sys.b[c] += dB[c];
        }
    }
}

void Equation::zeroOut() {  // transport.cpp:23-36
    for (auto& s : schemes) { std::fill(s.sys.A.begin(), s.sys.A.end(), 0.0); std::fill(s.sys.b.begin(), s.sys.b.end(), 0.0); }
    for (auto& s : sources) std::fill(s.sys.b.begin(), s.sys.b.end(), 0.0);
    std::fill(sys.A.begin(), sys.A.end(), 0.0);
    std::fill(sys.b.begin(), sys.b.end(), 0.0);
}

void Equation::setUnderRelax(f64 f) {  // transport.cpp:42-52
    if (f < 0.0 || f > 1.0) return;
    urf = f;
}

void Equation::relax() {  // transport.cpp:54-64
    phi->bump();
    const auto& P = *pattern;
    for (int64_t i = 0; i < P.n; ++i) {
        f64 d = sys.A[P.diag[i]];
        sys.b[i] += ((1.0 - urf) / urf) * (d * phi->v[i]);
    }
    for (int64_t i = 0; i < P.n; ++i) sys.A[P.diag[i]] /= urf;
}

// ---------------------------------------------------------------------- solver ----
void spmv(const CSR& p, const std::vector<f64>& A, const f64* x, f64* y) {
    for (int64_t i = 0; i < p.n; ++i) {
        f64 s = 0.0;
        for (int64_t k = p.rp[i]; k < p.rp[i + 1]; ++k) s += A[k] * x[p.col[k]];
        y[i] = s;
    }
}

f64 scaledResidual(const CSR& p, const std::vector<f64>& A, const std::vector<f64>& x,
                   const std::vector<f64>& b) {  // solver.cpp:8-24
    std::vector<f64> ax(p.n);
    spmv(p, A, x.data(), ax.data());
    f64 mdx = 0.0;
    for (int64_t i = 0; i < p.n; ++i) mdx = std::max(mdx, std::abs(A[p.diag[i]] * x[i]));
    f64 den = mdx + EPS;
    f64 r = 0.0;
    for (int64_t i = 0; i < p.n; ++i) r = std::max(r, std::abs(ax[i] - b[i]) / den);
    return r;
}

namespace {
struct ILU0 {
    const CSR& p;
    std::vector<f64> lu;
    ILU0(const CSR& pat, const std::vector<f64>& A) : p(pat), lu(A) {
        for (int64_t i = 0; i < p.n; ++i) {
            for (int64_t k = p.rp[i]; k < p.rp[i + 1] && p.col[k] < i; ++k) {
                int64_t c = p.col[k];
                f64 piv = lu[p.diag[c]];
                lu[k] /= piv;
                f64 lik = lu[k];
                // row i -= lik * upper part of row c
                int64_t kk = k + 1;
                for (int64_t j = p.diag[c] + 1; j < p.rp[c + 1]; ++j) {
                    int32_t cj = p.col[j];
                    while (kk < p.rp[i + 1] && p.col[kk] < cj) ++kk;
                    if (kk < p.rp[i + 1] && p.col[kk] == cj) lu[kk] -= lik * lu[j];
                }
            }
        }
    }
    void apply(const f64* r, f64* z) const {
        for (int64_t i = 0; i < p.n; ++i) {
            f64 s = r[i];
            for (int64_t k = p.rp[i]; k < p.diag[i]; ++k) s -= lu[k] * z[p.col[k]];
            z[i] = s;
        }
        for (int64_t i = p.n - 1; i >= 0; --i) {
            f64 s = z[i];
            for (int64_t k = p.diag[i] + 1; k < p.rp[i + 1]; ++k) s -= lu[k] * z[p.col[k]];
            z[i] = s / lu[p.diag[i]];
        }
    }
};

f64 dotv(const std::vector<f64>& a, const std::vector<f64>& b) {
    long double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += (long double)a[i] * b[i];
    return (f64)s;
}

// ILU(0)-preconditioned BiCGSTAB; returns the iterate with the smallest true residual.
void pbicgstab(const CSR& p, const std::vector<f64>& A, const ILU0& M, const std::vector<f64>& b,
               std::vector<f64>& x, f64 tol, int64_t maxit) {
    int64_t n = p.n;
    std::vector<f64> r(n), r0, v(n, 0), pp(n, 0), y(n), s(n), z(n), t(n);
    spmv(p, A, x.data(), r.data());
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    r0 = r;
    f64 bn2 = dotv(b, b);
    if (bn2 == 0) { std::fill(x.begin(), x.end(), 0.0); return; }
    f64 tol2 = tol * tol * bn2;
    f64 rho = 1, alpha = 1, w = 1;
    f64 r0n = dotv(r0, r0);
    for (int64_t it = 0; it < maxit && dotv(r, r) > tol2; ++it) {
        f64 rho_old = rho;
        rho = dotv(r0, r);
        if (std::abs(rho) < 1e-30 * r0n) {
            spmv(p, A, x.data(), r.data());
            for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
            r0 = r; rho = r0n = dotv(r, r);
        }
        f64 beta = (rho / rho_old) * (alpha / w);
        for (int64_t i = 0; i < n; ++i) pp[i] = r[i] + beta * (pp[i] - w * v[i]);
        M.apply(pp.data(), y.data());
        spmv(p, A, y.data(), v.data());
        alpha = rho / dotv(r0, v);
        for (int64_t i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
        M.apply(s.data(), z.data());
        spmv(p, A, z.data(), t.data());
        f64 tt = dotv(t, t);
        w = tt > 0 ? dotv(t, s) / tt : 0.0;
        for (int64_t i = 0; i < n; ++i) { x[i] += alpha * y[i] + w * z[i]; r[i] = s[i] - w * t[i]; }
    }
}
}  // namespace

void exactSolve(const CSR& p, const std::vector<f64>& A, const std::vector<f64>& b,
                std::vector<f64>& x) {
    int64_t n = p.n;
    f64 bn = std::sqrt(dotv(b, b));
    if (bn == 0) { std::fill(x.begin(), x.end(), 0.0); return; }  // Eigen: rhs == 0 -> x = 0
    ILU0 M(p, A);
    pbicgstab(p, A, M, b, x, 1e-15, 20 * n + 1000);
    // iterative refinement with long-double residuals
    std::vector<f64> r(n), d(n);
    for (int rep = 0; rep < 4; ++rep) {
        for (int64_t i = 0; i < n; ++i) {
            long double s = 0;
            for (int64_t k = p.rp[i]; k < p.rp[i + 1]; ++k) s += (long double)A[k] * x[p.col[k]];
            r[i] = (f64)((long double)b[i] - s);
        }
        f64 rn = std::sqrt(dotv(r, r));
        if (rn <= 1e-17 * bn) break;
        std::fill(d.begin(), d.end(), 0.0);
        pbicgstab(p, A, M, r, d, 1e-13, 20 * n + 1000);
        for (int64_t i = 0; i < n; ++i) x[i] += d[i];
    }
}

SolveResult solve(Equation& e, const SolveParams& prm) {  // solver.cpp:27-82
    auto& phi = *e.phi;
    const CSR& P = *e.pattern;
    f64 init = 0, cur = 0;
    bool conv = false;
    int64_t it = 0;
    e.updateCoeffs();
    e.relax();
    if (prm.ref_cell >= 0 && prm.ref_cell < P.n) {
        f64 d = e.sys.A[P.diag[prm.ref_cell]];
        e.sys.A[P.diag[prm.ref_cell]] += d;
        e.sys.b[prm.ref_cell] += d * prm.ref_value;
    }
    for (; it < prm.n_iter; it++) {
        if (it == 0) { phi.bump(); init = scaledResidual(P, e.sys.A, phi.v, e.sys.b); }
        phi.bump();
        exactSolve(P, e.sys.A, e.sys.b, phi.v);
        phi.bump();
        cur = scaledResidual(P, e.sys.A, phi.v, e.sys.b);
        if (cur < prm.eps) { conv = true; it++; break; }
    }
    if (prm.clear) e.zeroOut();
    return {it, init, cur, conv};
}

int64_t plainBicgstab(int64_t n, const int64_t* rp, const int32_t* col, const f64* val,
                      const f64* b, f64* x, f64 tol, int64_t maxit, f64* rel_res) {
    // Eigen BiCGSTAB iteration with the identity preconditioner (BiCGSTAB.h), plain doubles.
    auto mv = [&](const std::vector<f64>& in, std::vector<f64>& out) {
        for (int64_t i = 0; i < n; ++i) {
            f64 s = 0.0;
            for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * in[col[k]];
            out[i] = s;
        }
    };
    auto dt = [&](const std::vector<f64>& a, const std::vector<f64>& c) {
        f64 s = 0;
        for (int64_t i = 0; i < n; ++i) s += a[i] * c[i];
        return s;
    };
    std::vector<f64> xv(x, x + n), r(n), r0, v(n, 0), p(n, 0), s(n), t(n), bv(b, b + n);
    mv(xv, r);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    r0 = r;
    f64 r0n = dt(r0, r0), bn = dt(bv, bv);
    if (bn == 0) { std::fill(x, x + n, 0.0); if (rel_res) *rel_res = 0; return 0; }
    f64 tol2 = tol * tol * bn, rho = 1, alpha = 1, w = 1;
    int64_t i = 0;
    f64 rr = dt(r, r);
    while (rr > tol2 && i < maxit) {
        f64 rho_old = rho;
        rho = dt(r0, r);
        if (std::abs(rho) < EPS * EPS * r0n) {
            mv(xv, r);
            for (int64_t k = 0; k < n; ++k) r[k] = b[k] - r[k];
            r0 = r; rho = r0n = dt(r, r);
        }
        f64 beta = (rho / rho_old) * (alpha / w);
        for (int64_t k = 0; k < n; ++k) p[k] = r[k] + beta * (p[k] - w * v[k]);
        mv(p, v);
        alpha = rho / dt(r0, v);
        for (int64_t k = 0; k < n; ++k) s[k] = r[k] - alpha * v[k];
        mv(s, t);
        f64 tt = dt(t, t);
        w = tt > 0 ? dt(t, s) / tt : 0;
        for (int64_t k = 0; k < n; ++k) { xv[k] += alpha * p[k] + w * s[k]; r[k] = s[k] - w * t[k]; }
        rr = dt(r, r);
        ++i;
    }
    std::copy(xv.begin(), xv.end(), x);
    if (rel_res) *rel_res = std::sqrt(rr / bn);
    return i;
}

// ------------------------------------------------------------------ SIMPLE/PISO ----
namespace {
std::shared_ptr<Tensor> pressureCoeffs(std::vector<Equation*>& mom, const std::shared_ptr<Scalar>& p) {
    // SIMPLE.cpp:77-108
    const Mesh& m = *p->mesh;
    int64_t n = m.nc();
    auto D = std::make_shared<Tensor>();
    D->name = "D";
    D->mesh = p->mesh;
    D->v.assign(n, M3{});
    const CSR& P = *mom[0]->pattern;
    for (int64_t i = 0; i < n; ++i) {
        f64 du = m.vol[i] / (mom[0]->sys.A[P.diag[i]] + EPS);
        f64 dv = m.vol[i] / (mom[1]->sys.A[P.diag[i]] + EPS);
        f64 dw = 1.0;
        if (mom.size() == 3) dw = m.vol[i] / (mom[2]->sys.A[P.diag[i]] + EPS);
        D->v[i].m[0][0] = du; D->v[i].m[1][1] = dv; D->v[i].m[2][2] = dw;
    }
    for (auto* e : mom) e->zeroOut();
    return D;
}


void constrainPPrime(Scalar& pp) {  // SIMPLE.cpp:43-75
    std::vector<f64> fvals(pp.mesh->nf(), 0.0);
    for (const auto& p : pp.mesh->patches) {
        if (p.empty) continue;
        const std::string& k = p.bc("P").kind;
        bool known = k == "fixed" || k == "no-slip" || k == "symmetry" || k == "outlet" || k == "zero-gradient";
        if (!known) throw BCError("constrainPPrime() patch `" + p.name + "` kind `" + k + "`");
        if (k == "fixed" || k == "no-slip" || k == "symmetry") continue;
        for (auto fid : p.faces) fvals[fid] = pp.faceValue(fid);
    }
    pp.fv = std::move(fvals);
}

V3 rhieChow(int64_t fid, const Vector& U, const Tensor& D, const Scalar& P) {  // rhie_chow.h:47-63
    const Mesh& m = *U.mesh;
    const Face& f = m.faces[fid];
    V3 Uf = U.faceValue(fid);
    M3 Df = D.faceValue(fid);
    V3 gavg = P.gradFace(fid);
    // rhie_chow.cpp:5-25
    int64_t o = f.owner, n = f.neighbor;
    f64 po = P.cellValue(o), pn = P.cellValue(n);
    V3 dCF = m.cells[n].center - m.cells[o].center;
    f64 dn = norm(dCF);
    V3 e = dCF / dn;
    f64 gortho = (pn - po) / dn;
    V3 gcalc = gavg + ((gortho - dot(gavg, e)) * e);
    return Uf - mul(Df, gcalc - gavg);
}

struct PResult { std::shared_ptr<Scalar> pp; std::shared_ptr<Tensor> D; std::vector<Report> rep; };
}  // namespace

PressureSystem buildPressureSystem(SimpleControls c, std::vector<Equation*>& mom, std::shared_ptr<Vector>& U,
                                   std::shared_ptr<Vector>& mdot, std::shared_ptr<Scalar>& p) {  // SIMPLE.cpp:124-170
    auto D = pressureCoeffs(mom, p);
    Mesh& m = *p->mesh;
    // mdot->mapFaceValues(U->valueAtFace) (vector.h:296-316)
    mdot->initFaceData();
    for (int64_t fid = m.n_boundary; fid < m.nf(); ++fid) {
        V3 u = U->faceValue(fid);
        mdot->fd[fid] = u;
        mdot->ff[fid] = dot(u, m.faces[fid].S());
    }
    for (const auto& pt : m.patches) {
        if (pt.empty) continue;
        for (auto fid : pt.faces) {
            V3 u = U->faceValue(fid);
            mdot->fd[fid] = u;
            mdot->ff[fid] = dot(u, m.faces[fid].S());
        }
    }
    // mdot->mapInteriorFaceValues(rhieChowCorrectFace)
    for (int64_t fid = m.n_boundary; fid < m.nf(); ++fid) {
        V3 u = rhieChow(fid, *U, *D, *p);
        mdot->fd[fid] = u;
        mdot->ff[fid] = dot(u, m.faces[fid].S());
    }
    auto pp = std::make_shared<Scalar>(p->name, p->mesh, std::vector<f64>(m.nc(), 0.0), 1);
    constrainPPrime(*pp);
    if (c.p_ref_cell < 0 && needsRef(*pp)) c.p_ref_cell = 0;
    PressureSystem ps;
    ps.pp = pp;
    ps.D = D;
    Equation& pe = ps.pe;
    pe.phi = pp;
    pe.pattern = mom[0]->pattern;
    Scheme lap;
    lap.kind = SK::Diffusion; lap.variant = 1; lap.phi = pp; lap.kappa_t = D;
    lap.sys.resize(pe.pattern->col.size(), m.nc());
    Scheme dv;
    dv.kind = SK::SrcDivergence; dv.variant = 1; dv.vec = mdot;
    dv.sys.resize(0, m.nc());
    pe.schemes.push_back(lap);
    pe.sources.push_back(dv);
    pe.sys.resize(pe.pattern->col.size(), m.nc());
    ps.sp.n_iter = c.pressure_max_iter; ps.sp.eps = c.pressure_residual;
    ps.sp.ref_cell = c.p_ref_cell; ps.sp.ref_value = c.p_ref_value; ps.sp.clear = true;
    return ps;
}

std::vector<Report> momentumPredictor(std::vector<Equation*>& mom, const SimpleControls& c) {  // SIMPLE.cpp:19-41
    std::vector<Report> rep;
    for (auto* e : mom) {
        e->setUnderRelax(c.momentum_urf);
        SolveParams sp;
        sp.n_iter = c.momentum_max_iter; sp.eps = c.momentum_residual; sp.clear = false;
        auto r = solve(*e, sp);
        rep.push_back({e->phi->name, r.iter, r.r0, r.r, r.converged});
    }
    return rep;
}

namespace {
PResult solvePressure(SimpleControls c, std::vector<Equation*>& mom, std::shared_ptr<Vector>& U,
                      std::shared_ptr<Vector>& mdot, std::shared_ptr<Scalar>& p) {  // SIMPLE.cpp:124-196
    PressureSystem ps = buildPressureSystem(c, mom, U, mdot, p);
    std::vector<Report> rep;
    for (int64_t i = 0; i <= c.non_ortho; ++i) {
        auto r = solve(ps.pe, ps.sp);
        rep.push_back({ps.pp->name + "_corr_" + std::to_string(i), r.iter, r.r0, r.r, r.converged});
    }
    ps.pp->fv.clear();
    return {ps.pp, ps.D, rep};
}

void correctFields(std::shared_ptr<Vector>& U, std::shared_ptr<Vector>& mdot, std::shared_ptr<Scalar>& p,
                   const std::shared_ptr<Tensor>& D, std::shared_ptr<Scalar>& pp, f64 urf) {
    Mesh& m = *p->mesh;
    for (int64_t c = 0; c < m.nc(); ++c) {  // SIMPLE.cpp:224-232
        V3 u = U->cellValue(c) - mul(D->v[c], pp->gradCell(c));
        for (int k = 0; k < 3; ++k) { U->c[k]->v[c] = u[k]; U->c[k]->bump(); }
    }
    if (!mdot->has_fd) mdot->initFaceData();
    for (int64_t fid = m.n_boundary; fid < m.nf(); ++fid) {  // SIMPLE.cpp:234-237
        V3 u = mdot->faceValue(fid) - mul(D->faceValue(fid), pp->gradFace(fid));
        mdot->fd[fid] = u;
        mdot->ff[fid] = dot(u, m.faces[fid].S());
    }
    // p->update(p->values() + urf * pp->values())  (scalar.cpp:311-320)
    p->bump();
    std::vector<f64> nv(m.nc());
    for (int64_t i = 0; i < m.nc(); ++i) nv[i] = p->v[i] + (urf * pp->v[i]);
    p->updatePrev();
    p->v = std::move(nv);
    p->bump();
    p->fv.clear();
}
}  // namespace

f64 adjointObjective(const Scalar& T, int axis, f64 thr, f64* VR) {
    const Mesh& m = *T.mesh;
    f64 a = 0.0, v = 0.0;
    for (int64_t c = 0; c < (int64_t)m.nc(); ++c) {
        if (m.cells[c].center[axis] > thr) {
            a += m.vol[c] * T.v[c] * T.v[c];
            v += m.vol[c];
        }
    }
    if (!(v > 0.0)) throw std::invalid_argument("objective region holds no cell");
    *VR = v;
    return a / v;
}

namespace {
// The adjoint of a Corrected diffusion term's explicit non-orthogonal correction, formed as an
// explicit matrix: with least-squares gradients (least_squares.cpp:78-100),
// grad_c = pinv_c sum_k w_k (phi_k - phi_c), the correction b-terms T_f . grad_f (interior,
// into b_O, out of b_N) and T_b . gradAtBoundaryFace (fixed faces) are linear in phi, so
// dR/dphi = A - M with M = db/dphi; the adjoint is (A - M)^T lambda = g, solved exactly.
void adjointCorrectedSolve(Equation& e, const Scheme& diff, const std::vector<f64>& g, std::vector<f64>& lam) {
    const Scalar& T = *e.phi;
    const Mesh& m = *T.mesh;
    const CSR& P = *e.pattern;
    const int64_t n = (int64_t)m.nc();
    auto* ls = dynamic_cast<LeastSquares*>(T.grad.get());
    if (!ls) throw std::invalid_argument("adjoint: Corrected diffusion needs least-squares gradients");
    if (T.hasFaceValues()) throw std::invalid_argument("adjoint: Corrected diffusion needs a field without face values");
    auto isFixed = [&](int64_t fid) {
        const std::string& k = m.patchOf(m.faces[fid]).bc(T.name).kind;
        if (k == "fixed" || k == "no-slip") return true;
        if (k == "zero-gradient" || k == "outlet" || k == "symmetry") return false;
        throw std::invalid_argument("adjoint: boundary condition `" + k + "` is not supported");
    };
    // d grad_c / d phi_j as (j, vector) lists
    std::vector<std::vector<std::pair<int64_t, V3>>> G(n);
    for (int64_t c = 0; c < n; ++c) {
        V3 self{0, 0, 0};
        for (const auto& k : ls->cache[c]) {
            if (k.interior) {
                G[c].push_back({k.other, mul(ls->pinv[c], k.w)});
                self += k.w;
            } else if (isFixed(k.other)) {
                self += k.w;
            }
        }
        G[c].push_back({c, -1.0 * mul(ls->pinv[c], self)});
    }
    std::vector<std::map<int64_t, f64>> M(n);
    for (int64_t fid = 0; fid < (int64_t)m.faces.size(); ++fid) {
        const Face& f = m.faces[fid];
        const int64_t o = f.owner;
        if (f.interior()) {
            const int64_t nb = f.neighbor;
            V3 d = m.cells[nb].center - m.cells[o].center;
            f64 dn = norm(d);
            Decomp dc = decompose(diff.variant, kappaFace(diff, f.S(), fid), d / dn);
            f64 gw = geometricWeight(m, o, nb, f);
            for (const auto& [j, cv] : G[o]) {
                f64 v = gw * dot(dc.T, cv);
                M[o][j] += v;
                M[nb][j] -= v;
            }
            for (const auto& [j, cv] : G[nb]) {
                f64 v = (1. - gw) * dot(dc.T, cv);
                M[o][j] += v;
                M[nb][j] -= v;
            }
            continue;
        }
        if (m.patchOf(f).empty || !isFixed(fid)) continue;
        V3 d = f.center - m.cells[o].center;
        f64 dn = norm(d);
        V3 ev = d / dn;
        Decomp dc = decompose(diff.variant, kappaCell(diff, f.S(), o), ev);
        V3 tt = dc.T - dot(dc.T, ev) * ev;
        for (const auto& [j, cv] : G[o]) M[o][j] += dot(tt, cv);
        M[o][o] += -dot(dc.T, ev) / dn;
    }
    // J^T = (A - M)^T as CSR with sorted columns
    std::vector<std::map<int64_t, f64>> JT(n);
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t k = P.rp[i]; k < P.rp[i + 1]; ++k) JT[P.col[k]][i] += e.sys.A[k];
        for (const auto& [j, v] : M[i]) JT[j][i] -= v;
    }
    CSR Q;
    Q.n = n;
    Q.rp.assign(n + 1, 0);
    Q.diag.assign(n, -1);
    std::vector<f64> vals;
    for (int64_t i = 0; i < n; ++i) {
        for (const auto& [j, v] : JT[i]) {
            if (j == i) Q.diag[i] = (int64_t)Q.col.size();
            Q.col.push_back((int32_t)j);
            vals.push_back(v);
        }
        Q.rp[i + 1] = (int64_t)Q.col.size();
    }
    exactSolve(Q, vals, g, lam);
}
}  // namespace

void adjointGradient(Equation& e, const Scalar* src, const Scalar* kappa, int axis, f64 thr, f64* J,
                     std::vector<f64>* dJds, std::vector<f64>* dJdk) {
    const Scheme* diff = nullptr;
    const Scheme* source = nullptr;
    for (const auto& s : e.schemes) {
        if (s.kind == SK::Convection && s.variant == 0) continue;
        if (s.kind == SK::Diffusion && s.kappa && !s.kappa_t) { diff = &s; continue; }
        throw std::invalid_argument("adjoint: unsupported scheme in the equation (kind " +
                                    std::to_string((int)s.kind) + ", variant " + std::to_string(s.variant) + ")");
    }
    for (const auto& s : e.sources) {
        if (s.kind != SK::SrcConstant) throw std::invalid_argument("adjoint: unsupported source in the equation");
        if (src && s.phi.get() == src) source = &s;
    }
    if (kappa && (!diff || diff->kappa.get() != kappa)) throw std::invalid_argument("adjoint: kappa is not the diffusion coefficient");
    if (src && !source) throw std::invalid_argument("adjoint: source is not a ConstantScalar term");
    const Scalar& T = *e.phi;
    const Mesh& m = *T.mesh;
    const CSR& P = *e.pattern;
    const int64_t n = (int64_t)m.nc();
    e.zeroOut();
    e.updateCoeffs();
    f64 VR;
    *J = adjointObjective(T, axis, thr, &VR);
    std::vector<f64> g(n, 0.0), lam(n, 0.0), AT(P.col.size());
    for (int64_t c = 0; c < n; ++c)
        if (m.cells[c].center[axis] > thr) g[c] = 2.0 * m.vol[c] * T.v[c] / VR;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = P.rp[i]; k < P.rp[i + 1]; ++k) {
            int64_t j = P.col[k];
            auto b = P.col.begin() + P.rp[j], en = P.col.begin() + P.rp[j + 1];
            auto it = std::lower_bound(b, en, (int32_t)i);
            AT[k] = e.sys.A[it - P.col.begin()];
        }
    const bool corrected = diff && diff->variant != 0;
    if (!corrected) {
        exactSolve(P, AT, g, lam);
    } else {
        adjointCorrectedSolve(e, *diff, g, lam);
    }
    if (dJds && source) {
        dJds->assign(n, 0.0);
        for (int64_t c = 0; c < n; ++c) (*dJds)[c] = (source->variant == 1 ? -1.0 : 1.0) * lam[c] * m.vol[c];
    }
    if (dJdk && kappa) {
        std::vector<f64> acc(n, 0.0);
        for (size_t fid = 0; fid < m.faces.size(); ++fid) {
            const Face& f = m.faces[fid];
            const int64_t o = f.owner;
            if (f.interior()) {
                const int64_t nb = f.neighbor;
                V3 d = m.cells[nb].center - m.cells[o].center;
                f64 q;
                if (!corrected) {
                    f64 G = dot(f.S(), d) / (dot(d, d) + EPS);
                    q = (lam[o] - lam[nb]) * (T.v[o] - T.v[nb]) * G;
                } else {  // unit-kappa decomposition: orthogonal part and correction T1 . grad_f
                    f64 dn = norm(d);
                    Decomp d1 = decompose(diff->variant, f.S(), d / dn);
                    f64 a1 = dot(d1.E, d) / (dn * dn + EPS);
                    q = (lam[o] - lam[nb]) * (a1 * (T.v[o] - T.v[nb]) - dot(d1.T, T.gradFace((int64_t)fid)));
                }
                f64 w = geometricWeight(m, o, nb, f);
                acc[o] += w * q;
                acc[nb] += (1.0 - w) * q;
                continue;
            }
            const Patch& p = m.patchOf(f);
            if (p.empty) continue;
            const std::string& k = p.bc(T.name).kind;
            if (k == "fixed" || k == "no-slip") {
                V3 d = f.center - m.cells[o].center;
                if (!corrected) {
                    f64 G = dot(f.S(), d) / (dot(d, d) + EPS);
                    acc[o] += lam[o] * (T.v[o] - T.faceValue((int64_t)fid)) * G;
                } else {
                    f64 dn = norm(d);
                    Decomp d1 = decompose(diff->variant, f.S(), d / dn);
                    f64 a1 = dot(d1.E, d) / (dn * dn + EPS);
                    acc[o] += lam[o] * (a1 * (T.v[o] - T.faceValue((int64_t)fid)) -
                                        dot(d1.T, T.gradFace((int64_t)fid)));
                }
            } else if (!(k == "zero-gradient" || k == "outlet" || k == "symmetry")) {
                throw std::invalid_argument("adjoint: boundary condition `" + k + "` is not supported");
            }
        }
        dJdk->assign(n, 0.0);
        for (int64_t c = 0; c < n; ++c) (*dJdk)[c] = -acc[c];
    }
    e.zeroOut();  // leave the coefficients as a clearing solve() does
}

bool needsRef(const Scalar& pp) {  // SIMPLE.cpp:110-122
    for (const auto& p : pp.mesh->patches) {
        if (p.empty) continue;
        const std::string& k = p.bc(pp.name).kind;
        if (k == "fixed" || k == "no-slip") return false;
    }
    return true;
}

std::vector<Report> simpleStep(std::vector<Equation*>& mom, std::shared_ptr<Vector>& U,
                               std::shared_ptr<Vector>& mdot, std::shared_ptr<Scalar>& p,
                               const SimpleControls& c) {  // SIMPLE.cpp:200-216
    if (mom.size() != 2 && mom.size() != 3) throw std::runtime_error("SIMPLE expects 2 or 3 momentum predictors");
    std::vector<Report> rep = momentumPredictor(mom, c);
    auto pr = solvePressure(c, mom, U, mdot, p);
    rep.insert(rep.end(), pr.rep.begin(), pr.rep.end());
    correctFields(U, mdot, p, pr.D, pr.pp, c.pressure_urf);
    return rep;
}

std::vector<Report> primeStep(std::vector<Equation*>& mom, std::shared_ptr<Vector>& U,
                              std::shared_ptr<Vector>& mdot, std::shared_ptr<Scalar>& p,
                              const SimpleControls& c) {  // PRIME.cpp:11-56
    if (mom.size() != 2 && mom.size() != 3) throw std::runtime_error("PRIME expects 2 or 3 momentum predictors");
    for (auto* e : mom) {
        e->updateCoeffs();
        auto& phi = *e->phi;
        const CSR& P = *e->pattern;
        std::vector<f64> Ax(P.n);
        phi.bump();
        spmv(P, e->sys.A, phi.v.data(), Ax.data());
        std::vector<f64> nv(P.n);
        for (int64_t i = 0; i < P.n; ++i) {
            f64 ac = e->sys.A[P.diag[i]];
            f64 H = Ax[i] - ac * phi.v[i];
            nv[i] = (-H + e->sys.b[i]) / ac;
        }
        phi.v = std::move(nv);
        phi.bump();
    }
    auto pr = solvePressure(c, mom, U, mdot, p);
    correctFields(U, mdot, p, pr.D, pr.pp, c.pressure_urf);
    return pr.rep;
}

std::vector<Report> pisoStep(std::vector<Equation*>& mom, std::shared_ptr<Vector>& U,
                             std::shared_ptr<Vector>& mdot, std::shared_ptr<Scalar>& p,
                             int64_t outer, int64_t implicit_steps, int64_t corr_steps,
                             const SimpleControls& c) {  // PISO.cpp:10-55
    std::vector<Report> rep;
    for (int64_t i = 0; i < outer; ++i) {
        for (int64_t j = 0; j < implicit_steps; ++j) {
            auto r = simpleStep(mom, U, mdot, p, c);
            rep.insert(rep.end(), r.begin(), r.end());
        }
        for (int64_t j = 1; j < corr_steps; ++j) {
            auto r = primeStep(mom, U, mdot, p, c);
            rep.insert(rep.end(), r.begin(), r.end());
        }
    }
    return rep;
}

// ---------------------------------------------------------------------- output ----
// export.cpp:45-69 (ensureCache): vertex -> cells in cell order; export.cpp:79-112: per
// vertex, raw weights 1 / max(|v - c|, EPSILON) summed in list order, then each cell's
// normalised weight times its data added to the zero-initialised result.
std::vector<f64> cellToPoint(const Mesh& m, const std::vector<f64>& cell_data, size_t ncomp) {
    const size_t nv = m.vertices.size();
    std::vector<std::vector<size_t>> vert_to_cells(nv);
    for (size_t ci = 0; ci < (size_t)m.nc(); ++ci)
        for (auto vi : m.cells[ci].verts) vert_to_cells[vi].push_back(ci);
    const f64 EPSILON = std::numeric_limits<f64>::epsilon();
    std::vector<f64> result(nv * ncomp, 0.0);
    for (size_t v = 0; v < nv; ++v) {
        const auto& adj = vert_to_cells[v];
        if (adj.empty())
            throw MeshError("Vertex " + std::to_string(v) + " has no adjacent cells; point data cannot be computed.");
        std::vector<f64> weights(adj.size());
        f64 sum_weights = 0.0;
        for (size_t i = 0; i < adj.size(); ++i) {
            f64 d = norm(m.vertices[v] - m.cells[adj[i]].center);
            weights[i] = 1.0 / std::max(d, EPSILON);
            sum_weights += weights[i];
        }
        for (size_t i = 0; i < adj.size(); ++i) {
            f64 wi = weights[i] / sum_weights;
            for (size_t k = 0; k < ncomp; ++k) result[v * ncomp + k] += wi * cell_data[adj[i] * ncomp + k];
        }
    }
    return result;
}

}  // namespace oref
