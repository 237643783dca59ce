// library.cpp -- host runtime of libprism_hip: device objects, the reference's solver and
// pressure-velocity orchestration, and the C ABI of include/prism_hip.h.
//
// All device work of a context is issued on its one HIP stream.  Host-side state mirrors
// the reference's lazily refreshed gradient caches (a field's event counter is bumped
// wherever the reference takes a non-const values() reference; a gradient is recomputed at
// its first use after such a bump), so stateful schemes such as the Green-Gauss sweep see
// the same sequence of passes as the reference.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <type_traits>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/prism_hip.h"
#include "host_mesh.h"
#include "internal.h"
#include "krylov.h"
#include "output.h"
#include "sgs.h"

using namespace prism_hip;

namespace prism_hip {
void hipCheck(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        int code = (e == hipErrorOutOfMemory) ? PRISM_ERR_NOMEM : PRISM_ERR_DEVICE;
        throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
    }
}
}  // namespace prism_hip

namespace {

thread_local std::string g_err;

template <class F>
int guard(F&& f) {
    try {
        f();
        return PRISM_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc& e) {
        g_err = "host allocation failed";
        return PRISM_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PRISM_ERR_RUNTIME;
    }
}
[[noreturn]] void fail(int code, const std::string& m) { throw Error(code, m); }

template <typename T>
struct DevBuf {
    T* p{nullptr};
    size_t n{0};
    // borrowed: the buffer is a slice [base, base + cap) of memory it does not own (the static
    // scratch array); alloc() then only sets the size, and fails beyond cap
    T* base{nullptr};
    size_t cap{0};
    bool borrowed{false};
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    void bind(T* b, size_t c) {
        reset();
        borrowed = true;
        base = b;
        cap = c;
    }
    void reset() {
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        if (borrowed) {
            if (count > cap)
                throw Error(PRISM_ERR_RUNTIME, "scratch slice of " + std::to_string(cap) + " elements for " +
                                                   std::to_string(count));
            p = count ? base : nullptr;
            n = count;
            return;
        }
        reset();
        if (count == 0) return;
        HIPCHECK(hipMalloc((void**)&p, count * sizeof(T)));
        n = count;
    }
    void upload(const T* src, size_t count, hipStream_t s) {
        alloc(count);
        if (count) HIPCHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
};

}  // namespace

// ------------------------------------------------------------------------ context ----
struct prism_ctx {
    int device{0};
    hipStream_t stream{nullptr};
    bool timing{false};
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<int64_t, double>> stats;
    std::vector<std::string> order;
    KrylovScalars* hks{nullptr};  // pinned mirror
    unsigned char* scratch{nullptr};  // this device's static scratch array (scratch_address)

    Launch launch() {
        Launch L;
        L.stream = stream;
        L.user = this;
        L.events = timing ? &eventsHook : nullptr;
        return L;
    }
    hipEvent_t ev() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        // timing-only events: no system-scope fence (cache writeback) around every kernel,
        // which would otherwise inflate each measured duration by microseconds
        hipEvent_t e;
        HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        return e;
    }
    // the start / stop events of one kernel launch (recorded by the launch itself)
    static void eventsHook(void* u, const char* name, hipEvent_t* a, hipEvent_t* b) {
        auto* c = (prism_ctx*)u;
        if (c->pending.size() >= 4096) c->drain();
        Pending p{name, c->ev(), c->ev()};
        c->pending.push_back(p);
        *a = p.a;
        *b = p.b;
    }
    void drain() {
        if (pending.empty()) return;
        HIPCHECK(hipStreamSynchronize(stream));
        for (auto& p : pending) {
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, p.a, p.b));
            auto it = stats.find(p.name);
            if (it == stats.end()) {
                order.push_back(p.name);
                stats[p.name] = {1, (double)ms};
            } else {
                it->second.first += 1;
                it->second.second += ms;
            }
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    void sync() { HIPCHECK(hipStreamSynchronize(stream)); }
};

// --------------------------------------------------- SGS preconditioner plan ----
// The colouring and colour-ordered sliced-ELL layout of a sparsity pattern for the
// multicolour symmetric Gauss-Seidel preconditioner; built on the first preconditioned solve
// of the pattern and kept with it.
struct SgsPlan {
    int n{0}, nc{0};
    std::vector<int32_t> cptr;          // colour c: positions [cptr[c], cptr[c+1])
    DevBuf<int32_t> perm;               // position -> row
    DevBuf<int32_t> soff, scol, emap;   // off-diagonals: slice offsets, position columns, CSR index
    DevBuf<int32_t> dmap;               // position -> CSR index of its diagonal entry
    int64_t slots{0};
};
struct SgsCache {
    std::unique_ptr<SgsPlan> plan;
};

// --------------------------------------------------------------------------- mesh ----
struct prism_mesh {
    prism_ctx* ctx{nullptr};
    HMesh h;
    DevBuf<double> cx, cy, cz, vol, sx, sy, sz, fx, fy, fz, nx, ny, nz, dx, dy, dz, gc, kx, ky, kz, pinv;
    DevBuf<int32_t> owner, nbr, fpatch, cf_off, cf, rp, col, diag, slot_on, slot_no, gg_cells, dup_slot, dup_ptr;
    DevBuf<int32_t> sell_off, sell_col, csr2sell, csc_ptr, csc_row, csc_pos, csr2dia, sell_base;
    DevBuf<int16_t> sell_rel;
    // vertex -> cell adjacency for point export (uploaded on first use): CSR offsets, and the
    // sliced cell ids and normalised inverse-distance weights (64 vertices per slice)
    DevBuf<int32_t> v2c_off, v2c_sbase, v2c_cell;
    DevBuf<double> idw;
    DevBuf<double> pts;  // point values of the last export (grows to the widest field exported)
    DMesh dm{};
    SgsCache sgs;
};

namespace {

int8_t kindCode(const std::string& k) {
    if (k == "fixed") return BC_FIXED;
    if (k == "no-slip") return BC_NOSLIP;
    if (k == "symmetry") return BC_SYMMETRY;
    if (k == "outlet") return BC_OUTLET;
    if (k == "zero-gradient") return BC_ZEROGRAD;
    if (k == "fixed-gradient") return BC_FIXEDGRAD;
    if (k == "empty") return BC_EMPTY;
    return BC_OTHER;
}
bool isComponent(const std::string& n) {  // mesh/boundary.cpp:334-346
    if (n.size() < 2) return false;
    char l = n.back();
    return (l == 'x' || l == 'y' || l == 'z') && n[n.size() - 2] == '_';
}

// Per-field boundary table (host mirror + device copy), from BoundaryPatch lookups
// (mesh/boundary.cpp:426-492).
struct BCTable {
    std::vector<int8_t> kind;
    std::vector<std::string> kind_str;
    std::vector<double> s;   // scalar value (NaN if unavailable)
    std::vector<double> v;   // vector value [3] (NaN if unavailable)
    std::vector<uint8_t> s_ok, v_ok;
    DevBuf<int8_t> dk;
    DevBuf<double> ds, dv;

    void build(const HMesh& m, const std::string& name, hipStream_t st) {
        size_t P = m.patches.size();
        kind.assign(MAXP, BC_MISSING);
        kind_str.assign(P, "");
        s.assign(MAXP, NAN);
        v.assign(3 * MAXP, NAN);
        s_ok.assign(MAXP, 0);
        v_ok.assign(MAXP, 0);
        for (size_t p = 0; p < P; ++p) {
            const auto& bcs = m.patches[p].bcs;
            auto it = bcs.find(name);
            if (it == bcs.end() && isComponent(name)) it = bcs.find(name.substr(0, name.size() - 2));
            if (it != bcs.end()) {
                kind[p] = kindCode(it->second.kind);
                kind_str[p] = it->second.kind;
            }
            // getScalarBoundaryCondition
            if (isComponent(name)) {
                auto jt = bcs.find(name.substr(0, name.size() - 2));
                if (jt != bcs.end() && jt->second.vk == 2) {
                    int c = name.back() == 'x' ? 0 : (name.back() == 'y' ? 1 : 2);
                    s[p] = jt->second.v[c];
                    s_ok[p] = 1;
                }
            } else {
                auto jt = bcs.find(name);
                if (jt != bcs.end() && jt->second.vk == 1) { s[p] = jt->second.v[0]; s_ok[p] = 1; }
            }
            auto vt = bcs.find(name);  // getVectorBoundaryCondition: exact name
            if (vt != bcs.end() && vt->second.vk == 2) {
                for (int i = 0; i < 3; ++i) v[3 * p + i] = vt->second.v[i];
                v_ok[p] = 1;
            }
        }
        dk.upload(kind, st);
        ds.upload(s, st);
        dv.upload(v, st);
    }
};

void requireKinds(const HMesh& m, const BCTable& t, const std::string& field, const char* who,
                  std::initializer_list<int8_t> allowed) {
    for (size_t p = 0; p < m.patches.size(); ++p) {
        if (m.patches[p].empty) continue;
        if (t.kind[p] == BC_MISSING)
            fail(PRISM_ERR_RUNTIME, "prism::mesh::BoundaryPatch::getBoundaryCondition(): Boundary patch `" +
                                        m.patches[p].name + "` does not have a field named `" + field + "`");
        bool ok = false;
        for (auto a : allowed) ok |= (t.kind[p] == a);
        if (!ok)
            fail(PRISM_ERR_BC, std::string(who) + ": boundary condition `" + t.kind_str[p] + "` on patch `" +
                                   m.patches[p].name + "` is not implemented (field `" + field + "`)");
    }
}

}  // namespace

// ------------------------------------------------------------------------- fields ----
struct prism_eqn;
struct prism_vfield;
struct prism_field {
    prism_mesh* mesh{nullptr};
    std::string name;
    int kind{0};   // 1 pressure
    int coord{-1};
    bool gg{false};
    DevBuf<double> v, fv, g;
    bool has_fv{false};
    uint64_t event{1}, gevent{0};
    BCTable bc;
    int hist_max{0};
    std::vector<std::unique_ptr<DevBuf<double>>> hist;
    // The pressure-correction equation the SIMPLE / PRIME steps build on this pressure field
    // (with mass flux pcache_mdot), kept between steps with its p', D and Krylov workspace:
    // each step resets p' and recomputes D instead of allocating them again.
    struct EqnDelete {
        void operator()(prism_eqn* e) const;
    };
    std::unique_ptr<prism_eqn, EqnDelete> pcache;
    const prism_vfield* pcache_mdot{nullptr};

    int64_t n() const { return mesh->h.n; }
    DScalar view() const {
        DScalar s{};
        s.v = v.p;
        s.fv = has_fv ? fv.p : nullptr;
        s.gx = g.p;
        s.gy = g.p + n();
        s.gz = g.p + 2 * n();
        s.bck = bc.dk.p;
        s.bcs = bc.ds.p;
        s.bcv = bc.dv.p;
        s.pressure = kind == 1;
        return s;
    }
    void bump() { ++event; }
};

struct prism_vfield {
    prism_mesh* mesh{nullptr};
    std::string name;
    std::unique_ptr<prism_field> comp[3];
    DevBuf<double> fd, ff;
    bool has_fd{false}, has_ff{false};
    BCTable bc;
    DVector view() const {
        DVector U{};
        for (int k = 0; k < 3; ++k) U.c[k] = comp[k]->view();
        U.fd = has_fd ? fd.p : nullptr;
        U.ff = has_ff ? ff.p : nullptr;
        U.bck = bc.dk.p;
        U.bcv = bc.dv.p;
        return U;
    }
};

struct prism_tfield {
    prism_mesh* mesh{nullptr};
    std::string name;
    DevBuf<double> t;  // 9 x n SoA
    bool diag_only{false};
    DTensor view() const { return DTensor{t.p, diag_only ? 1 : 0}; }
};

namespace {

std::unique_ptr<prism_field> makeScalar(prism_mesh* m, const std::string& name, int kind, int coord,
                                        const double* values, double value) {
    auto f = std::make_unique<prism_field>();
    f->mesh = m;
    f->name = name;
    f->kind = kind;
    f->coord = coord;
    int64_t n = m->h.n;
    Launch L = m->ctx->launch();
    f->v.alloc(n);
    if (values) HIPCHECK(hipMemcpyAsync(f->v.p, values, n * sizeof(double), hipMemcpyHostToDevice, L.stream));
    else launch_fill(L, f->v.p, n, value);
    f->g.alloc(3 * n);
    launch_fill(L, f->g.p, 3 * n, 0.0);  // Green-Gauss cache starts at zero (green_gauss.cpp:36-38)
    // gradient scheme (scalar.cpp:260-304)
    std::string lookup = coord >= 0 ? name.substr(0, name.size() - 2) : name;
    for (const auto& fi : m->h.infos)
        if (fi.name == lookup && fi.has_grad) {
            f->gg = (fi.grad == "green-gauss" || fi.grad == "greenGauss");
            break;
        }
    f->bc.build(m->h, name, L.stream);
    return f;
}

// Scalar face values are needed on every non-empty patch (gradients, BCs).
void requireFaceValueBCs(const prism_field& f, const char* who) {
    if (f.has_fv) return;
    const HMesh& h = f.mesh->h;
    requireKinds(h, f.bc, f.name, who, {BC_FIXED, BC_NOSLIP, BC_SYMMETRY, BC_OUTLET, BC_ZEROGRAD});
    for (size_t p = 0; p < h.patches.size(); ++p) {
        if (h.patches[p].empty) continue;
        int8_t k = f.bc.kind[p];
        if ((k == BC_FIXED || k == BC_NOSLIP) && !f.bc.s_ok[p])
            fail(PRISM_ERR_RUNTIME, "prism::mesh::BoundaryPatch::getScalarBoundaryCondition(): Boundary patch '" +
                                        h.patches[p].name + "' field '" + f.name + "' is not a scalar");
    }
}

// Lazily refreshed gradient (gradient.h:26-31): recompute on first use after a write.
void ensureGrad(prism_field& f) {
    if (f.gevent == f.event) return;
    prism_mesh* m = f.mesh;
    requireFaceValueBCs(f, "IGradient");
    if (f.gg) requireKinds(m->h, f.bc, f.name, "GreenGauss", {BC_FIXED, BC_NOSLIP, BC_SYMMETRY, BC_OUTLET, BC_ZEROGRAD, BC_FIXEDGRAD, BC_EMPTY, BC_OTHER});
    Launch L = m->ctx->launch();
    DScalar s = f.view();
    int64_t n = f.n();
    if (!f.gg) {
        launch_grad_ls(L, m->dm, s, f.g.p);
    } else {
        const auto& off = m->h.gg_off;
        for (size_t l = 0; l + 1 < off.size(); ++l)
            launch_grad_gg_level(L, m->dm, s, m->gg_cells.p + off[l], off[l + 1] - off[l], f.g.p, f.g.p + n,
                                 f.g.p + 2 * n);
    }
    f.gevent = f.event;
}

// The gradients of a vector's three components where a caller needs all three at once:
// one least-squares pass over the mesh when all three are stale least-squares fields,
// else each component's own lazy refresh.  Either way each gradient equals what
// ensureGrad gives.
void ensureGradVec(prism_vfield& U) {
    prism_field* c[3] = {U.comp[0].get(), U.comp[1].get(), U.comp[2].get()};
    bool fuse = true;
    for (prism_field* f : c) fuse = fuse && !f->gg && f->gevent != f->event;
    if (!fuse) {
        for (prism_field* f : c) ensureGrad(*f);
        return;
    }
    for (prism_field* f : c) requireFaceValueBCs(*f, "IGradient");
    const DScalar s[3] = {c[0]->view(), c[1]->view(), c[2]->view()};
    double* const g[3] = {c[0]->g.p, c[1]->g.p, c[2]->g.p};
    launch_grad_ls3(U.mesh->ctx->launch(), U.mesh->dm, s, g);
    for (prism_field* f : c) f->gevent = f->event;
}

}  // namespace

// ---------------------------------------------------------------------- equations ----
struct SchemeRec {
    prism_scheme_desc d;
};

struct Workspace {  // Krylov buffers of one equation, allocated on first solve
    int64_t n{0};
    int G{0}, cpb{0};
    int Gv{0}, cpbv{0};  // grid of the V kernel
    int Gs{0}, cpbs{0}, Gvs{0}, cpbvs{0};  // the same on sliced-ELL matrices
    DevBuf<double> r64;
    DevBuf<double> d_r, d_r0, d_p[2], d_v[2], d_s, d_t, d_dinv, d_rhat;
    DevBuf<float> f_x, f_r, f_r0, f_p[2], f_v[2], f_s, f_t, f_dinv, f_rhat;
    DevBuf<double> part_xr, part_scr, part_aux, part_x;
    DevBuf<double> sval64, svalp64, dinv64;  // FP64 values; column-scaled values and M^-1 (Jacobi)
    DevBuf<float> sval32;
    DevBuf<KrylovScalars> ks;
    DevBuf<double> aval, bt, part_ref;  // reference-cell deflation: matrix, rhs, partials
    DevBuf<double> xsave;                // FP32 refinement: x at the best reliable update so far
    // multicolour SGS-preconditioned iteration (position numbering)
    struct Sgs {
        DevBuf<double> sval, d, dinv, x, b, r, r0, p, v, y, s, z, t, e, part_x, part_v, part_t;
        DevBuf<SgsState> st;
    } sgs;
    // Captured pairs of iterations (even/odd start) per precision, keyed by the kernel
    // arguments they were captured with.
    struct Graph {
        hipGraphExec_t exec{nullptr};
        std::vector<unsigned char> key;
    };
    Graph graphs[2][2];
    Workspace() = default;
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;
    ~Workspace() {
        for (auto& row : graphs)
            for (auto& g : row)
                if (g.exec) (void)hipGraphExecDestroy(g.exec);
    }
    // Binds every buffer a krylovStep of an n-row matrix with `vtot` stored values (DIA or
    // sliced ELL) may use to slices of [base, base + bytes) instead of device allocations
    // (the allocation-free entry point).  Fails with PRISM_ERR_INVALID when they do not fit.
    void bindScratch(unsigned char* base, size_t bytes, int64_t nn, int64_t vtot, bool mixed, bool pre) {
        size_t off = 0;
        auto carve = [&](auto& buf, size_t count, bool need) {
            using T = std::remove_pointer_t<decltype(buf.p)>;
            if (!need) return;
            off = (off + 255) & ~size_t(255);
            if (base) buf.bind(reinterpret_cast<T*>(base + off), count);
            off += count * sizeof(T);
        };
        int spb = 0;
        const size_t g = (size_t)std::max(krylov_grid(nn, &spb), krylov_grid(nn, &spb, true, true));
        for (int pass = 0; pass < 2; ++pass) {  // pass 0 sizes, pass 1 binds
            unsigned char* const b0 = base;
            if (pass == 0) base = nullptr;
            off = 0;
            carve(r64, nn, true);
            for (auto* d : {&d_r, &d_r0, &d_p[0], &d_p[1], &d_v[0], &d_v[1], &d_s, &d_t, &d_rhat}) carve(*d, nn, true);
            carve(d_dinv, nn, pre);
            for (auto* f : {&f_x, &f_r, &f_r0, &f_p[0], &f_p[1], &f_v[0], &f_v[1], &f_s, &f_t, &f_rhat})
                carve(*f, nn, mixed);
            carve(f_dinv, nn, mixed && pre);
            carve(part_xr, 3 * g, true);
            carve(part_scr, 6 * g, true);
            carve(part_aux, std::max<size_t>(3 * g, scaled_residual_partials(nn)), true);
            carve(part_x, 2 * g, true);
            carve(ks, 1, true);
            carve(part_ref, 3 * (size_t)ref_blocks((int)nn), true);
            carve(sval64, vtot, true);
            carve(svalp64, vtot, pre);
            carve(dinv64, nn, pre);
            carve(sval32, vtot, mixed);
            carve(xsave, nn, mixed);
            base = b0;
            if (pass == 0 && off > bytes)
                throw Error(PRISM_ERR_INVALID, "allocation-free step: a " + std::to_string(nn) + "-row system with " +
                                                   std::to_string(vtot) + " stored values needs " + std::to_string(off) +
                                                   " bytes of scratch, the device array holds " + std::to_string(bytes));
        }
    }
    void init(int64_t nn, hipStream_t) {
        if (n == nn) return;
        n = nn;
        G = krylov_grid(nn, &cpb);
        Gv = krylov_grid(nn, &cpbv, false);
        Gs = krylov_grid(nn, &cpbs, true, true);
        Gvs = krylov_grid(nn, &cpbvs, false, true);
        const size_t g = (size_t)std::max(G, Gs);  // partial sums of either grid
        part_xr.alloc(3 * g);
        part_scr.alloc(6 * g);
        part_aux.alloc(std::max<size_t>(3 * g, scaled_residual_partials(nn)));
        part_x.alloc(2 * g);
        ks.alloc(1);
        r64.alloc(nn);
    }
};

struct prism_eqn {
    prism_mesh* mesh{nullptr};
    prism_field* phi{nullptr};
    std::vector<prism_scheme_desc> schemes;  // ctor order
    double urf{1.0};
    bool dirty{false};
    DevBuf<double> A, b, scr;
    std::vector<std::unique_ptr<DevBuf<double>>> fluxbufs;
    Workspace ws;
    struct SymPatch {  // eqn::boundary::Symmetry<Transport> groups of one symmetry patch
        DevBuf<int32_t> cells, off, faces;
        int32_t n{0};
        DSymGroup view() const { return {cells.p, off.p, faces.p, n}; }
    };
    std::vector<std::unique_ptr<SymPatch>> sym;  // patches in mesh order; empty: no handler
    bool sym_handler{false};
    // fields owned by the equation (a pressure-correction system handed to the caller)
    std::unique_ptr<prism_field> own_phi;
    std::unique_ptr<prism_tfield> own_D;
    DevBuf<double> tmp;  // PRIME's explicit momentum update
};
void prism_field::EqnDelete::operator()(prism_eqn* e) const { delete e; }

namespace {

void checkSchemeBCs(prism_eqn& e) {
    const HMesh& h = e.mesh->h;
    prism_field& phi = *e.phi;
    for (const auto& s : e.schemes) {
        if (s.kind == PRISM_SCHEME_CONVECTION) {
            requireKinds(h, phi.bc, phi.name, "prism::scheme::convection::IConvection",
                         {BC_FIXED, BC_NOSLIP, BC_SYMMETRY, BC_OUTLET, BC_ZEROGRAD});
            for (size_t p = 0; p < h.patches.size(); ++p)
                if (!h.patches[p].empty && phi.bc.kind[p] == BC_FIXED && !phi.bc.s_ok[p])
                    fail(PRISM_ERR_RUNTIME, "getScalarBoundaryCondition(): patch '" + h.patches[p].name +
                                                "' field '" + phi.name + "' is not a scalar");
            prism_vfield& U = *s.vec;
            if (!U.has_ff || !U.has_fd)
                requireKinds(h, U.bc, U.name, "prism::field::GeneralVector",
                             {BC_FIXED, BC_NOSLIP, BC_SYMMETRY, BC_OUTLET, BC_ZEROGRAD});
        } else if (s.kind == PRISM_SCHEME_DIFFUSION) {
            requireKinds(h, phi.bc, phi.name, "prism::scheme::diffusion",
                         {BC_FIXED, BC_NOSLIP, BC_SYMMETRY, BC_OUTLET, BC_ZEROGRAD, BC_FIXEDGRAD});
            for (size_t p = 0; p < h.patches.size(); ++p) {
                if (h.patches[p].empty) continue;
                int8_t k = phi.bc.kind[p];
                // the Dirichlet handlers read the field's value on their own patch only
                // (diffusion_boundary.h:144-159, 188-207), so only those patches need a value
                if ((k == BC_FIXED || k == BC_NOSLIP) && !phi.has_fv && !phi.bc.s_ok[p])
                    fail(PRISM_ERR_RUNTIME, "prism::mesh::BoundaryPatch::getScalarBoundaryCondition(): Boundary patch '" +
                                                h.patches[p].name + "' field '" + phi.name + "' is not a scalar");
                if (k == BC_FIXEDGRAD && !phi.bc.v_ok[p])
                    fail(PRISM_ERR_RUNTIME, "getVectorBoundaryCondition(): patch '" + h.patches[p].name +
                                                "' field '" + phi.name + "' is not a vector");
            }
        }
    }
}

// Transport::updateCoeffs (transport.cpp:4-21), optionally with relax() fused.
void updateCoeffs(prism_eqn& e, bool relax) {
    prism_mesh* m = e.mesh;
    const HMesh& h = m->h;
    Launch L = m->ctx->launch();
    checkSchemeBCs(e);
    DEquation de{};
    de.phi = e.phi->view();
    // gradients first (lazy refresh at first use)
    for (const auto& s : e.schemes) {
        if (s.kind == PRISM_SCHEME_CONVECTION || (s.kind == PRISM_SCHEME_DIFFUSION && s.variant != 0)) ensureGrad(*e.phi);
        if (s.kind == PRISM_SCHEME_SRC_GRADIENT) ensureGrad(*s.phi);
    }
    de.phi = e.phi->view();
    size_t fb = 0;
    auto fluxOf = [&](prism_vfield* U) -> const double* {
        if (U->has_ff) return U->ff.p;
        if (e.fluxbufs.size() <= fb) e.fluxbufs.push_back(std::make_unique<DevBuf<double>>());
        DevBuf<double>& buf = *e.fluxbufs[fb++];
        buf.alloc(h.F);
        requireKinds(h, U->bc, U->name, "prism::field::GeneralVector::fluxAtFace",
                     {BC_FIXED, BC_NOSLIP, BC_SYMMETRY, BC_OUTLET, BC_ZEROGRAD});
        for (int k = 0; k < 3; ++k) if (!U->has_fd) requireFaceValueBCs(*U->comp[k], "prism::field::GeneralVector");
        launch_vector_flux(L, m->dm, U->view(), buf.p);
        return buf.p;
    };
    for (const auto& s : e.schemes) {
        switch (s.kind) {
            case PRISM_SCHEME_CONVECTION:
            case PRISM_SCHEME_DIFFUSION: {
                if (de.n_face >= MAX_FACE_SCHEMES) fail(PRISM_ERR_RUNTIME, "too many face-based schemes");
                DFaceScheme& fs = de.fs[de.n_face++];
                fs.kind = s.kind == PRISM_SCHEME_CONVECTION ? 1 : 2;
                fs.variant = s.variant;
                if (s.kind == PRISM_SCHEME_CONVECTION) {
                    fs.flux = fluxOf(s.vec);
                    fs.U = s.vec->view();
                } else if (s.kappa) {
                    fs.kappa_kind = 0;  // interior faces interpolate, boundary faces use the owner cell
                    fs.kappa = s.kappa->view();
                } else {
                    fs.kappa_kind = 1;
                    fs.kappa_t = s.kappa_t->view();
                }
            } break;
            default: {
                if (de.n_cell >= MAX_CELL_TERMS) fail(PRISM_ERR_RUNTIME, "too many cell-based terms");
                DCellTerm& ct = de.ct[de.n_cell++];
                ct.kind = s.kind == PRISM_SCHEME_SRC_CONSTANT ? 3 : s.kind == PRISM_SCHEME_SRC_GRADIENT ? 4
                        : s.kind == PRISM_SCHEME_SRC_DIVERGENCE ? 5 : s.kind == PRISM_SCHEME_SRC_IMPLICIT ? 6
                        : s.kind == PRISM_SCHEME_BACKWARD_EULER ? 7 : 8;
                ct.sign = s.variant;
                ct.coord = s.coord;
                ct.coeff = s.coeff;
                if (ct.kind == 3) { ct.phi = s.phi->v.p; s.phi->bump(); }
                if (ct.kind == 4) ct.g = s.phi->g.p + (int64_t)s.coord * h.n;
                if (ct.kind == 5) ct.flux = fluxOf(s.vec);
                if (ct.kind == 7) {
                    if (s.phi->hist.empty())
                        fail(PRISM_ERR_RUNTIME, "prism::scheme::temporal::BackwardEuler::apply() was called for field `" +
                                                    s.phi->name + "`, but the field does not have any previous time step values stored.");
                    ct.phi = s.phi->hist.front()->p;
                }
                if (ct.kind == 8) {
                    if (s.phi->hist.empty())
                        fail(PRISM_ERR_RUNTIME, "prism::scheme::temporal::AdamMoulton::apply() was called for field `" +
                                                    s.phi->name + "`, but the field does not have previous time steps values stored.");
                    ct.phi = s.phi->hist.front()->p;
                    ct.phi2 = s.phi->hist.size() >= 2 ? s.phi->hist[1]->p : nullptr;
                }
            } break;
        }
    }
    int64_t nint = h.F - h.nb;
    e.scr.alloc(std::max<int64_t>(1, 4 * nint));
    de.A = e.A.p;
    de.b = e.b.p;
    de.scr = e.scr.p;
    de.accumulate = e.dirty ? 1 : 0;
    // equation boundary handlers run after every scheme and source (transport.cpp:20), so
    // relax() cannot be fused into the cell kernel when one is present
    const bool sym = e.sym_handler && !e.sym.empty();
    launch_assemble(L, m->dm, de, (relax && !sym) ? 1 : 0, e.urf);
    if (sym) {
        const DFaceScheme* conv = nullptr;
        const DFaceScheme* diff = nullptr;
        for (int k = 0; k < de.n_face; ++k) {  // the last of each kind (transport.h:103-116)
            if (de.fs[k].kind == 1) conv = &de.fs[k];
            if (de.fs[k].kind == 2) diff = &de.fs[k];
        }
        if (!conv || !diff)
            fail(PRISM_ERR_RUNTIME, "eqn::boundary::Symmetry<Transport>: equation needs a convection and a diffusion scheme");
        for (const auto& sp : e.sym) launch_symmetry(L, m->dm, sp->view(), e.phi->coord, conv->U, *diff, e.A.p, e.b.p);
        if (relax) launch_relax(L, m->dm, e.A.p, e.b.p, e.phi->v.p, e.urf);
    }
    e.dirty = true;
}

void zeroCoeffs(prism_eqn& e) {
    Launch L = e.mesh->ctx->launch();
    launch_fill(L, e.A.p, (int64_t)e.A.n, 0.0);
    launch_fill(L, e.b.p, (int64_t)e.b.n, 0.0);
    e.dirty = false;
}

}  // namespace

// --------------------------------------------------------------------- bare matrix ----
struct prism_csr {
    prism_ctx* ctx{nullptr};
    int64_t n{0};
    DevBuf<int32_t> rp, col, diag, sell_off, sell_col, csr2sell, csc_ptr, csc_row, csc_pos, csr2dia, sell_base;
    DevBuf<int16_t> sell_rel;
    std::vector<int32_t> dia_off;
    int64_t stotal{0};
    bool all_diag{true};  // every row stores its diagonal entry
    DevBuf<double> val;
    Workspace ws;
    SgsCache sgs;
};

// ------------------------------------------------------------------------- solver ----
namespace {

struct MatRef {
    int n;
    const int32_t *rp, *col, *diag;
    const double* val;
    int64_t nnz;
    const int32_t *soff, *scol, *csr2sell;  // sliced-ELL view of the pattern
    int64_t stotal;
    const int16_t* srel;  // 16-bit column offsets (from sbase), or nullptr
    const int32_t *cscp, *cscr, *cscm;  // transpose of the pattern (ptr, row, CSR position)
    const int32_t* csr2dia;             // DIA view: CSR entry -> d n + row, or nullptr
    int ndia;
    int32_t doff[DIA_MAX];
    bool all_diag{true};                // every row stores its diagonal entry
    SgsCache* sgs{nullptr};             // the pattern's SGS plan (built on first use)
    const int32_t* sbase{nullptr};      // with srel: the base of each slice's slots
    // sets the DIA fields from a pattern's offsets (empty: no DIA view)
    void setDia(const int32_t* map, const std::vector<int32_t>& off) {
        ndia = map && !off.empty() ? (int)off.size() : 0;
        csr2dia = ndia ? map : nullptr;
        for (int d = 0; d < DIA_MAX; ++d) doff[d] = d < ndia ? off[d] : 0;
    }
};

struct StepOut {
    int64_t iters{0};
    int32_t restarts{0};
    int64_t refinements{0};
    double rel{0};
    int done{0};
};

template <typename T>
void allocVecs(Workspace& w, int64_t n, bool precond);

template <>
void allocVecs<double>(Workspace& w, int64_t n, bool precond) {
    w.d_r.alloc(n); w.d_r0.alloc(n); w.d_p[0].alloc(n); w.d_p[1].alloc(n); w.d_v[0].alloc(n); w.d_v[1].alloc(n);
    w.d_s.alloc(n); w.d_t.alloc(n); w.d_rhat.alloc(n);
    if (precond) w.d_dinv.alloc(n);
}
template <>
void allocVecs<float>(Workspace& w, int64_t n, bool precond) {
    w.f_r.alloc(n); w.f_r0.alloc(n); w.f_p[0].alloc(n); w.f_p[1].alloc(n); w.f_v[0].alloc(n); w.f_v[1].alloc(n);
    w.f_s.alloc(n); w.f_t.alloc(n); w.f_x.alloc(n); w.f_rhat.alloc(n);
    if (precond) w.f_dinv.alloc(n);
}

// Iterations replayed from a captured graph of two (PRISM_KRYLOV_GRAPH=1): the kernel
// arguments of iteration k depend only on k's parity, so one graph per start parity
// serves a whole solve.
// Reliable updates of the FP32 iteration restart it (symmetric A) or replace its residual
// (nonsymmetric A); PRISM_KRYLOV_REFINE=restart or =replace forces one for every matrix.
bool refineRestart(bool symmetric) {
    static const int mode = [] {
        const char* e = std::getenv("PRISM_KRYLOV_REFINE");
        if (!e) return 0;
        if (std::strcmp(e, "restart") == 0) return 1;
        if (std::strcmp(e, "replace") == 0) return 2;
        return 0;
    }();
    return mode == 0 ? symmetric : mode == 1;
}

// PRISM_KRYLOV_SELL32=1: 32-bit sliced-ELL column indices also where the 16-bit offsets fit
// (read per solve; the two encodings give the same columns).
bool sell32() {
    const char* e = std::getenv("PRISM_KRYLOV_SELL32");
    return e && std::atoi(e) != 0;
}

// The FP32 reliable-update threshold when the caller leaves it automatic (see krylovStep);
// PRISM_KRYLOV_INNER_TOL overrides it (read per solve, for measurement).
double autoInnerTol(bool structured) {
    const char* e = std::getenv("PRISM_KRYLOV_INNER_TOL");
    if (e && std::atof(e) > 0) return std::atof(e);
    return structured ? 1e-2 : 1e-1;
}

// PRISM_KRYLOV_NO_DIA=1: sliced ELL also for patterns that qualify for DIA.
bool noDia() {
    static const bool on = [] {
        const char* e = std::getenv("PRISM_KRYLOV_NO_DIA");
        return e && std::atoi(e) != 0;
    }();
    return on;
}

// PRISM_KRYLOV_DIFF=0: plain FP32 products instead of the difference form (read per solve).
bool noDiff() {
    const char* e = std::getenv("PRISM_KRYLOV_DIFF");
    return e && std::atoi(e) == 0;
}

// PRISM_KRYLOV_DSYM=0: full DIA storage also for exactly symmetric matrices (read per solve).
bool noDiaSym() {
    const char* e = std::getenv("PRISM_KRYLOV_DSYM");
    return e && std::atoi(e) == 0;
}

// PRISM_KRYLOV_FIVE=1: the five-kernel iteration (p and s updates in kernels of their own,
// (r0, A p) reduced directly) instead of the three-kernel one.
bool fiveKernels() {
    static const bool on = [] {
        const char* e = std::getenv("PRISM_KRYLOV_FIVE");
        return e && std::atoi(e) != 0;
    }();
    return on;
}

// PRISM_KRYLOV_TRACE=1: one line on stderr per solver event (divergence fallback, stall exit,
// FP64 continuation, end of a step), for diagnosing slow or failing solves.
bool traceEnabled() {
    static const bool on = [] {
        const char* e = std::getenv("PRISM_KRYLOV_TRACE");
        return e && std::atoi(e) != 0;
    }();
    return on;
}

bool graphsEnabled() {
    static const bool on = [] {
        const char* e = std::getenv("PRISM_KRYLOV_GRAPH");
        return e && std::atoi(e) != 0;
    }();
    return on;
}

template <typename T, typename SetIter>
hipGraphExec_t iterationPair(Workspace& w, const Launch& L, KView<T>& kv, SetIter&& setIter, long long k) {
    setIter(k);
    std::vector<unsigned char> key(sizeof(kv));
    std::memcpy(key.data(), &kv, sizeof(kv));
    Workspace::Graph& g = w.graphs[sizeof(T) == 4][k & 1];
    if (g.exec && g.key == key) return g.exec;
    if (g.exec) {
        HIPCHECK(hipGraphExecDestroy(g.exec));
        g.exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    HIPCHECK(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal));
    launch_iteration<T>(L, kv);
    setIter(k + 1);
    launch_iteration<T>(L, kv);
    HIPCHECK(hipStreamEndCapture(L.stream, &graph));
    HIPCHECK(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
    HIPCHECK(hipGraphDestroy(graph));
    g.key = std::move(key);
    return g.exec;
}

KrylovScalars pollScalars(prism_ctx* ctx, Workspace& w) {
    HIPCHECK(hipMemcpyAsync(ctx->hks, w.ks.p, sizeof(KrylovScalars), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    return *ctx->hks;
}

// One BiCGSTAB step x <- BiCGSTAB(A, b, x) (bicgstab.cpp:7-9 semantics: a full solve from the
// current guess).  precision 0: FP64.  precision 1: FP32 iteration with FP64 reliable updates
// (x accumulated in FP64, true residual b - A x recomputed in FP64 whenever the recursive
// residual has dropped by `inner_tol`, and to confirm convergence).  At an update the FP32
// iteration continues with the replaced residual (nonsymmetric A), or restarts from it
// (symmetric A: measured to need fewer iterations there, and far more on nonsymmetric ones).
bool diaSymmetric(prism_ctx* ctx, Workspace& w, const MatRef& A, const double* vals);

// ------------------------------------------- multicolour SGS-preconditioned BiCGSTAB ----
// The pattern's colouring and colour-ordered sliced ELL (buildSgsLayout), on the device.
std::unique_ptr<SgsPlan> buildSgsPlan(prism_ctx* ctx, const MatRef& A) {
    const int n = A.n;
    std::vector<int32_t> rp(n + 1), dg(n);
    HIPCHECK(hipMemcpyAsync(rp.data(), A.rp, (n + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipMemcpyAsync(dg.data(), A.diag, n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    std::vector<int32_t> col(rp[n]);
    HIPCHECK(hipMemcpyAsync(col.data(), A.col, col.size() * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    SgsLayout lay;
    buildSgsLayout(n, rp.data(), col.data(), dg.data(), SGS_MAXC, lay);
    auto plan = std::make_unique<SgsPlan>();
    plan->n = n;
    plan->nc = lay.nc;
    plan->cptr = lay.cptr;
    plan->slots = (int64_t)lay.scol.size();
    plan->perm.upload(lay.perm, ctx->stream);
    plan->soff.upload(lay.soff, ctx->stream);
    plan->scol.upload(lay.scol, ctx->stream);
    plan->emap.upload(lay.emap, ctx->stream);
    plan->dmap.upload(lay.dmap, ctx->stream);
    ctx->sync();
    return plan;
}

// BiCGSTAB::step with the multicolour SGS preconditioner (FP64 in every precision mode).
StepOut sgsStep(prism_ctx* ctx, Workspace& w, const MatRef& A, const double* b, double* x,
                const prism_solve_params& p) {
    if (!A.sgs) fail(PRISM_ERR_INVALID, "SGS preconditioner: not available for this matrix");
    if (!A.sgs->plan || A.sgs->plan->n != A.n) A.sgs->plan = buildSgsPlan(ctx, A);
    const SgsPlan& P = *A.sgs->plan;
    hipStream_t st = ctx->stream;
    Launch L = ctx->launch();
    const int n = A.n;
    const int G = (int)std::min<int64_t>(SGS_GMAX, (n + SGS_BS - 1) / SGS_BS);
    auto& W = w.sgs;
    W.sval.alloc(P.slots);
    for (auto* v : {&W.d, &W.dinv, &W.x, &W.b, &W.r, &W.r0, &W.p, &W.v, &W.y, &W.s, &W.z, &W.t, &W.e}) v->alloc(n);
    W.part_x.alloc(3 * (size_t)G);
    W.part_v.alloc(G);
    W.part_t.alloc(2 * (size_t)G);
    W.st.alloc(1);
    SgsState h{};
    h.s[1] = SgsSlot{1.0, 1.0, 1.0, 0.0, 0.0};
    h.tol = p.tol > 0 ? p.tol : 1e-14;
    HIPCHECK(hipMemcpyAsync(W.st.p, &h, sizeof(h), hipMemcpyHostToDevice, st));
    launch_sgs_values(L, P.slots, P.emap.p, A.val, W.sval.p, n, P.dmap.p, W.d.p, W.dinv.p, W.st.p);
    launch_sgs_gather(L, n, P.perm.p, b, W.b.p);
    launch_sgs_gather(L, n, P.perm.p, x, W.x.p);
    SgsView V{};
    V.n = n;
    V.nc = P.nc;
    V.G = G;
    for (int c = 0; c <= P.nc; ++c) V.cptr[c] = P.cptr[c];
    V.soff = P.soff.p;
    V.scol = P.scol.p;
    V.sval = W.sval.p;
    V.d = W.d.p;
    V.dinv = W.dinv.p;
    V.x = W.x.p; V.b = W.b.p; V.r = W.r.p; V.r0 = W.r0.p; V.p = W.p.p; V.v = W.v.p;
    V.y = W.y.p; V.s = W.s.p; V.z = W.z.p; V.t = W.t.p; V.e = W.e.p;
    V.part_x = W.part_x.p;
    V.part_v = W.part_v.p;
    V.part_t = W.part_t.p;
    V.st = W.st.p;
    auto poll = [&] {
        SgsState s{};
        HIPCHECK(hipMemcpyAsync(&s, W.st.p, sizeof(s), hipMemcpyDeviceToHost, st));
        ctx->sync();
        return s;
    };
    if (poll().bad_diag) fail(PRISM_ERR_RUNTIME, "SGS preconditioner: a diagonal entry is zero or not finite");
    launch_sgs_seed(L, V, true);
    const long long maxit = p.max_iters > 0 ? p.max_iters : 2 * (long long)n;
    StepOut out;
    long long k = 0, base = 0;  // base: Eigen restarts its iteration count at the first restart
    int fresh = 1, chunk = 4;
    for (;;) {
        const long long todo = std::min<long long>(chunk, maxit - (k - base));
        if (todo <= 0) {
            out.done = 2;
            out.iters = k - base;
            break;
        }
        for (long long j = 0; j < todo; ++j) launch_sgs_iteration(L, V, k + j, j == 0 ? fresh : 0);
        fresh = 0;
        const SgsState s = poll();
        if (s.done) {
            out.done = s.done;
            out.iters = s.k_event - base;
            break;
        }
        if (s.flag == 2) {  // rho breakdown: Eigen restarts with r0 = r = b - A x (keeps p, v)
            k = s.k_event;
            if (out.restarts++ == 0) base = k;
            static const int zero = 0;
            HIPCHECK(hipMemcpyAsync(&W.st.p->flag, &zero, sizeof(int), hipMemcpyHostToDevice, st));
            launch_sgs_seed(L, V, false);
            fresh = 2;
            continue;
        }
        k += todo;
        chunk = std::min(chunk * 2, 64);
    }
    // true relative residual ||b - A x|| / ||b||
    launch_sgs_true_residual(L, V);
    std::vector<double> part(2 * (size_t)G);
    HIPCHECK(hipMemcpyAsync(part.data(), W.part_t.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    ctx->sync();
    double rr = 0.0, bb = 0.0;
    for (int i = 0; i < G; ++i) {
        rr += part[i];
        bb += part[G + i];
    }
    if (out.done == 4) {
        launch_fill(L, x, n, 0.0);  // Eigen: b == 0 -> x = 0
        out.rel = 0.0;
    } else {
        launch_sgs_scatter(L, n, P.perm.p, W.x.p, x);
        out.rel = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    }
    if (traceEnabled())
        std::fprintf(stderr, "krylov n=%d sgs (%d colours): %lld iterations, done %d, %d restarts, rel %.3e\n", n, P.nc,
                     (long long)out.iters, out.done, out.restarts, out.rel);
    return out;
}

// zero_rowsum: A is known to have zero row sums (the deflated reference-cell system).
StepOut krylovStep(prism_ctx* ctx, Workspace& w, const MatRef& A, const double* b, double* x,
                   const prism_solve_params& p, bool symmetric = false, bool zero_rowsum = false) {
    if (p.precond == PRISM_PRECOND_SGS) return sgsStep(ctx, w, A, b, x, p);
    hipStream_t st = ctx->stream;
    Launch L = ctx->launch();
    w.init(A.n, st);
    const bool mixed32 = p.precision == PRISM_PRECISION_FP32_REFINE;
    const bool pre = p.precond == PRISM_PRECOND_JACOBI;
    KrylovScalars h{};
    h.tol = p.tol > 0 ? p.tol : 1e-14;
    h.maxit = p.max_iters > 0 ? p.max_iters : 2 * (long long)A.n;
    // reliable update every 100x of residual drop on structured (DIA) matrices; every 10x on
    // unstructured ones, where a CPU restatement of the FP32 iteration needs 0.03-0.1 to match
    // the FP64 iteration count on tetrahedral pressure systems (64^3 x 6: 774 at 0.1, 809 at
    // 0.05, 995 at 0.01, 806 in FP64) and the six pressure solves of two 96^3 x 6 PISO steps
    // all converge in FP32 at 0.1 (at 0.05 the first diverges twice and finishes in FP64),
    // while the 128^3 box system wants 0.01 (632 against 735 at 0.05)
    const bool structured = A.ndia > 0 && !noDia();
    double delta = p.inner_tol > 0 ? p.inner_tol : autoInnerTol(structured);
    h.delta2 = delta * delta;
    h.breakdown = EPS * EPS;  // Eigen: |rho| < eps^2 * |r0|^2
    h.rho[0] = h.rho[1] = h.alpha = h.omega = 1.0;
    StepOut out;
    auto run = [&](auto tag, const bool mixed) {
        using T = decltype(tag);
        HIPCHECK(hipMemcpyAsync(w.ks.p, &h, sizeof(h), hipMemcpyHostToDevice, st));
        allocVecs<T>(w, A.n, pre);
        KView<T> kv{};
        kv.n = A.n;
        kv.G = structured ? w.G : w.Gs;
        kv.slices_per_block = structured ? w.cpb : w.cpbs;
        kv.Gv = structured ? w.Gv : w.Gvs;
        kv.spb_v = structured ? w.cpbv : w.cpbvs;
        kv.rp = A.rp;
        kv.col = A.col;
        kv.val64 = A.val;
        kv.precond = pre;
        kv.mixed = mixed;
        kv.part_in = w.part_xr.p;
        kv.part_out = w.part_scr.p;
        kv.part_x = w.part_x.p;
        kv.ks = w.ks.p;
        kv.cscp = A.cscp;
        kv.cscr = A.cscr;
        kv.cscm = A.cscm;
        T* P[2];
        T* V[2];
        // sliced-ELL (or DIA) values of this solve (padding slots are zero from allocation)
        const bool dia = A.ndia > 0 && !noDia();
        const int32_t* vmap = dia ? A.csr2dia : A.csr2sell;
        const int64_t vtot = dia ? (int64_t)A.ndia * A.n : A.stotal;
        if (w.sval64.n != (size_t)vtot) {
            w.sval64.alloc(vtot);
            launch_fill(L, w.sval64.p, vtot, 0.0);  // padding slots stay zero
        }
        launch_to_sell<double>(L, A.nnz, vmap, A.val, w.sval64.p);
        kv.soff = A.soff;
        kv.scol = A.scol;
        kv.srel = sell32() ? nullptr : A.srel;
        kv.sbase = kv.srel ? A.sbase : nullptr;
        kv.ndia = dia ? A.ndia : 0;
        for (int d = 0; d < DIA_MAX; ++d) kv.doff[d] = dia ? A.doff[d] : 0;
        kv.sval64 = w.sval64.p;
        kv.dsym = dia && !pre && !noDiaSym() && diaSymmetric(ctx, w, A, w.sval64.p) ? 1 : 0;
        if (pre) {  // right Jacobi: M^-1 in FP64, the SpMV values column-scaled once per step
            w.dinv64.alloc(A.n);
            launch_dinv<double>(L, A.n, A.diag, A.val, w.dinv64.p);
        }
        if constexpr (sizeof(T) == 8) {
            kv.val = A.val;
            kv.sval = w.sval64.p;
            if (pre) {
                if (w.svalp64.n != (size_t)vtot) {
                    w.svalp64.alloc(vtot);
                    launch_fill(L, w.svalp64.p, vtot, 0.0);
                }
                launch_to_sell_scaled<double>(L, A.nnz, vmap, A.col, A.val, w.dinv64.p, w.svalp64.p);
                kv.sval = w.svalp64.p;
            }
            kv.x = x;
            kv.r = w.d_r.p; kv.r0 = w.d_r0.p; kv.s = w.d_s.p; kv.t = w.d_t.p;
            P[0] = w.d_p[0].p; P[1] = w.d_p[1].p; V[0] = w.d_v[0].p; V[1] = w.d_v[1].p;
            kv.rhat = w.d_rhat.p;
            kv.dinv = pre ? w.dinv64.p : nullptr;
        } else {
            if (w.sval32.n != (size_t)vtot) {
                w.sval32.alloc(vtot);
                launch_fill(L, w.sval32.p, vtot, 0.0f);
            }
            if (pre) launch_to_sell_scaled<float>(L, A.nnz, vmap, A.col, A.val, w.dinv64.p, w.sval32.p);
            else launch_to_sell<float>(L, A.nnz, vmap, A.val, w.sval32.p);
            // FP32 three-kernel iteration: SpMVs in the difference form (matRow), row sums in
            // the diagonal slots; needs every row's diagonal entry
            kv.diff = (!fiveKernels() && A.cscp && A.all_diag && !noDiff()) ? (zero_rowsum ? 2 : 1) : 0;
            if (kv.diff)
                launch_rowsum_diag<float>(L, A.n, A.rp, A.col, A.diag, vmap, A.val, pre ? w.dinv64.p : nullptr,
                                          w.sval32.p);
            kv.val = nullptr;
            kv.sval = w.sval32.p;
            kv.x = w.f_x.p;
            kv.r = w.f_r.p; kv.r0 = w.f_r0.p; kv.s = w.f_s.p; kv.t = w.f_t.p;
            P[0] = w.f_p[0].p; P[1] = w.f_p[1].p; V[0] = w.f_v[0].p; V[1] = w.f_v[1].p;
            kv.rhat = w.f_rhat.p;
            kv.dinv = pre ? w.f_dinv.p : nullptr;
            if (pre) launch_convert<float>(L, A.n, w.dinv64.p, w.f_dinv.p);
        }
        krylov_pack<T>(kv);
        bool three = !fiveKernels() && A.cscp;
        auto setIter = [&](long long k) {  // buffers and rho slot of iteration k
            int q = (int)(k & 1);
            kv.par = q;
            if (three) {  // p_k, v_k in slot q; X writes p_k+1 over p_k-1
                kv.p_old = P[q ^ 1]; kv.p_new = P[q];
                kv.v_old = V[q ^ 1]; kv.v_new = V[q];
            } else {
                kv.p_old = P[q]; kv.p_new = P[q ^ 1];
                kv.v_old = V[q]; kv.v_new = V[q ^ 1];
            }
        };
        // after every seed: rhat when r0 changed, then p and the partials the next V reduces
        auto enter = [&](int start) {
            if (!three) return;
            if (start) launch_rhat<T>(L, kv, w.r64.p, pre ? w.dinv64.p : nullptr);
            launch_entry<T>(L, kv);
        };
        setIter(0);
        launch_residual<T>(L, kv, b, x, w.r64.p, 1);
        launch_seed<T>(L, kv, w.r64.p, 1);
        enter(1);
        // FP32 refinement guard.  The three-kernel form takes alpha from (rhat, p) and rho
        // from (r0, s) - omega (r0, t); in FP32 these part from the stored v and r by
        // rounding, which on ill-conditioned systems (tetrahedral pressure systems) makes the
        // iteration diverge.  The last x whose true residual was the best so far is kept; a
        // true residual 1e4x above that best, or a NaN, restores it and ends the FP32 phase
        // (the step then finishes in FP64 from that x).
        const bool have_xsave = mixed && three;
        if (have_xsave) {
            w.xsave.alloc(A.n);
            HIPCHECK(hipMemcpyAsync(w.xsave.p, x, A.n * sizeof(double), hipMemcpyDeviceToDevice, st));
        }
        // On the first divergence the FP32 iteration restarts from the best x (a fresh shadow
        // residual and search direction); on the second the step finishes in FP64 from it.
        int retries = 0;
        auto diverged = [&](long long nit) -> bool {  // true: continue in FP32
            if (traceEnabled()) std::fprintf(stderr, "krylov n=%d: FP32 divergence at iteration %lld\n", A.n, nit);
            HIPCHECK(hipMemcpyAsync(x, w.xsave.p, A.n * sizeof(double), hipMemcpyDeviceToDevice, st));
            if (retries++ == 0) {
                launch_residual<T>(L, kv, b, x, w.r64.p, 0);
                launch_reset_events(L, w.ks.p, true, false);
                setIter(nit);
                launch_seed<T>(L, kv, w.r64.p, 1);
                enter(1);
                ++out.refinements;
                return true;
            }
            out.done = 6;
            out.iters = nit;
            return false;
        };
        out.refinements = 1;
        int chunk = 4;
        long long last_iter = 0;
        double last_rr = 0.0, rate = 0.0;
        // stagnation guard: when the residual has not dropped 4x over `window` iterations,
        // force a reliable update (FP32) or, once the relative residual is below 1e-8 (the
        // round-off floor of FP64 BiCGSTAB on ill-conditioned systems lies there or lower),
        // stop (FP64).  Above that level a long transient is ordinary (advection-dominated
        // operators, adjoints with a local right-hand side) and the iteration runs on to
        // maxit as Eigen's does.
        const long long window = 400;
        const double floor2 = 1e-16;
        double best = INFINITY;
        long long best_iter = 0;
        double last_true = INFINITY;
        bool forced = false;  // the pending reliable update was forced by stagnation
        int stalls = 0;       // forced updates in a row that did not lower the FP64 residual
        for (;;) {
            KrylovScalars s = pollScalars(ctx, w);
            if (!s.done && !s.flag) {
                if (s.rr < 0.25 * best) { best = s.rr; best_iter = s.nit; }
                else if (s.nit - best_iter > window) {
                    if (!mixed) {
                        if (s.rr <= floor2 * s.bb) { out.done = 5; out.iters = s.nit; break; }
                    } else {
                        s.flag = 1;  // serviced below
                        forced = true;
                    }
                    best = INFINITY;
                    best_iter = s.nit;
                }
            }
            if (s.done == 3 && have_xsave) {  // NaN in the three-kernel FP32 iteration
                if (diverged(s.nit)) continue;
                break;
            }
            if (s.done) {
                out.done = s.done;
                out.iters = s.nit;
                if (s.done == 4) launch_fill(L, x, A.n, 0.0);  // Eigen: b == 0 -> x = 0
                break;
            }
            setIter(s.nit);
            if (s.flag == 2) {  // rho breakdown: Eigen restarts with r0 = r = b - A x (keeps p, v)
                if constexpr (sizeof(T) == 4) launch_flush(L, A.n, x, kv.x);
                launch_residual<T>(L, kv, b, x, w.r64.p, 0);
                launch_seed<T>(L, kv, w.r64.p, 2);
                if (out.restarts++ == 0) launch_reset_events(L, w.ks.p, false, true);
                enter(2);
                ++out.refinements;
                continue;
            }
            if (s.flag == 1) {  // mixed: reliable update of x and r in FP64
                if constexpr (sizeof(T) == 4) launch_flush(L, A.n, x, kv.x);
                launch_residual<T>(L, kv, b, x, w.r64.p, 0);
                if (have_xsave) {
                    KrylovScalars u = pollScalars(ctx, w);
                    if (!(u.rr_true == u.rr_true) || u.rr_true > 1e4 * last_true) {
                        if (diverged(s.nit)) continue;
                        break;
                    }
                    if (u.rr_true <= last_true)
                        HIPCHECK(hipMemcpyAsync(w.xsave.p, x, A.n * sizeof(double), hipMemcpyDeviceToDevice, st));
                }
                const int start = refineRestart(symmetric) ? 1 : 0;
                launch_seed<T>(L, kv, w.r64.p, start);
                enter(start);
                ++out.refinements;
                KrylovScalars t = pollScalars(ctx, w);
                if (t.done) { out.done = t.done; out.iters = t.nit; break; }
                if (t.iter >= t.maxit) { out.done = 2; out.iters = t.nit; break; }
                if (forced) {  // two stalled windows in a row that the update did not help:
                    stalls = t.rr_true < 0.9 * last_true ? 0 : stalls + 1;
                    if (stalls >= 2) {  // the FP32 iteration can no longer reduce the FP64 residual
                        out.done = 5;
                        out.iters = t.nit;
                        if (traceEnabled())
                            std::fprintf(stderr, "krylov n=%d: FP32 stall exit at iteration %lld, rel %.3e\n", A.n, t.nit,
                                         std::sqrt(t.rr_true / t.bb));
                        if (have_xsave)  // the x of the best reliable update
                            HIPCHECK(hipMemcpyAsync(x, w.xsave.p, A.n * sizeof(double), hipMemcpyDeviceToDevice, st));
                        break;
                    }
                }
                forced = false;
                last_true = std::min(last_true, t.rr_true);
                continue;
            }
            long long progressed = s.nit - last_iter;
            if (progressed >= chunk / 2) chunk = std::min(chunk * 2, 64);
            // Size the next chunk from the observed convergence rate (log-decrease of (r,r)
            // per iteration, smoothed over polls), so that few launches are issued past the
            // device's next event: they return at once, but each still costs a launch.
            if (progressed > 0 && last_rr > 0 && s.rr > 0 && s.rr < last_rr) {
                double r1 = std::log(last_rr / s.rr) / (double)progressed;
                rate = rate > 0 ? 0.5 * (rate + r1) : r1;
            }
            if (rate > 0 && s.rr > 0) {
                double target = mixed ? std::max(s.tol2, s.delta2 * s.rr_ref) : s.tol2;
                double need = std::log(s.rr / target) / rate;
                if (need > 0 && need < chunk) chunk = std::max(4, 2 * (int)std::ceil(need * 0.5 + 0.5));
            }
            last_iter = s.nit;
            last_rr = s.rr;
            if (three) {
                for (int k = 0; k < chunk; ++k) {
                    setIter(s.nit + k);
                    launch_iteration3<T>(L, kv);
                }
            } else if (graphsEnabled() && !L.events) {
                hipGraphExec_t ex = iterationPair<T>(w, L, kv, setIter, s.nit);
                for (int k = 0; k < chunk; k += 2) HIPCHECK(hipGraphLaunch(ex, st));
            } else {
                for (int k = 0; k < chunk; ++k) {
                    setIter(s.nit + k);
                    launch_iteration<T>(L, kv);
                }
            }
        }
        // true relative residual ||b - A x|| / ||b|| (FP64)
        launch_residual<T>(L, kv, b, x, w.r64.p, 0);
        KrylovScalars f = pollScalars(ctx, w);
        out.rel = f.bb > 0 ? std::sqrt(f.rr_true / f.bb) : 0.0;
        if (traceEnabled())
            std::fprintf(stderr, "krylov n=%d %s: %lld iterations, done %d, %lld refinements, rel %.3e, %s\n", A.n,
                         sizeof(T) == 4 ? "fp32" : "fp64", (long long)out.iters, out.done, (long long)out.refinements, out.rel,
                         three ? "three kernels" : "five kernels");
    };
    if (!mixed32) {
        run(double{}, false);
        return out;
    }
    run(float{}, true);
    if (out.done == 5 || out.done == 6) {
        // The FP32 iteration stalled above the tolerance (two forced reliable updates in a row
        // without lowering the FP64 residual) or diverged: finish in FP64 from its best x.
        const StepOut first = out;
        out = StepOut{};
        run(double{}, false);
        out.iters += first.iters;
        out.restarts += first.restarts;
        out.refinements += first.refinements;
    }
    return out;
}

double scaledResidual(prism_ctx* ctx, Workspace& w, const MatRef& A, const double* x, const double* b) {
    Launch L = ctx->launch();
    w.init(A.n, ctx->stream);
    launch_scaled_residual(L, A.n, A.rp, A.col, A.val, A.diag, x, b, w.part_aux.p, w.part_aux.n, w.ks.p);
    KrylovScalars s = pollScalars(ctx, w);
    return s.scaled[0] / (s.scaled[1] + EPS);  // solver.cpp:20-23
}

// ---------------------------------------------------------- reference-cell deflation ----
// The reference-cell fix (solver.cpp:50-57) turns an all-Neumann matrix A (symmetric, zero
// row sums, null space span{1}) into A' = A + a e_r e_r^T (a = A_rr) with b' = b + a ref e_r.
// A' has one eigenvalue near a/n, which makes Krylov iterations long and FP32 ones unstable.
// Deflating span{1} is exact for this structure: with sigma = 1^T b, the solution of
// A' x = b' is x = y + (sigma/a + ref - y_r) 1 for any y with A y = b - sigma e_r (a
// compatible singular system, whose Krylov iteration never sees the null space).  So when
// the test below holds, the step solves A y = b - sigma e_r instead and shifts the result.
struct RefCell {
    bool ok{false};
    int32_t r{-1};
    int32_t slot{-1};  // CSR slot of A_rr
    double a{0};       // A_rr before the fix (A'_rr = 2a)
};

// max |A - A^T| <= 1e-12 max |diag|, from the structure-test kernel without a reference cell.
// Symmetric DIA storage applies when the offsets mirror (-o_k .. 0 .. o_k) and every lower
// entry equals its mirrored upper entry bit for bit (checked on the device per solve).
bool diaSymmetric(prism_ctx* ctx, Workspace& w, const MatRef& A, const double* vals) {
    const int D = A.ndia;
    if (D < 3 || D % 2 == 0 || A.doff[D / 2] != 0) return false;
    for (int d = 0; d < D / 2; ++d)
        if (A.doff[d] != -A.doff[D - 1 - d]) return false;
    KView<double> kd{};
    kd.n = A.n;
    kd.ndia = D;
    for (int d = 0; d < DIA_MAX; ++d) kd.doff[d] = A.doff[d];
    int nbk = ref_blocks(A.n);
    w.part_ref.alloc(std::max<size_t>(w.part_ref.n, 3 * (size_t)nbk));
    launch_dia_symcheck(ctx->launch(), kd, vals, w.part_ref.p, w.part_ref.n);
    std::vector<double> h(nbk);
    HIPCHECK(hipMemcpyAsync(h.data(), w.part_ref.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    for (double v : h)
        if (v != 0.0) return false;
    return true;
}

bool isSymmetric(prism_ctx* ctx, Workspace& w, const MatRef& A) {
    int nbk = ref_blocks(A.n);
    w.part_ref.alloc(std::max<size_t>(w.part_ref.n, 3 * (size_t)nbk));
    launch_ref_check(ctx->launch(), A.n, A.rp, A.col, A.val, A.diag, -1, 0.0, w.part_ref.p, w.part_ref.n);
    std::vector<double> h(3 * (size_t)nbk);
    HIPCHECK(hipMemcpyAsync(h.data(), w.part_ref.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    double dg = 0, as = 0;
    for (int i = 0; i < nbk; ++i) {
        dg = std::max(dg, h[3 * i + 1]);
        as = std::max(as, h[3 * i + 2]);
    }
    return dg > 0.0 && as <= 1e-12 * dg;
}

RefCell refCellDeflation(prism_ctx* ctx, Workspace& w, const MatRef& A, int32_t r, int32_t slot) {
    RefCell rc;
    rc.r = r;
    rc.slot = slot;
    double arr = 0.0;
    HIPCHECK(hipMemcpyAsync(&arr, A.val + slot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    rc.a = 0.5 * arr;
    if (!(rc.a > 0.0)) return rc;
    int nbk = ref_blocks(A.n);
    w.part_ref.alloc(std::max<size_t>(w.part_ref.n, 3 * (size_t)nbk));
    launch_ref_check(ctx->launch(), A.n, A.rp, A.col, A.val, A.diag, r, rc.a, w.part_ref.p, w.part_ref.n);
    std::vector<double> h(3 * (size_t)nbk);
    HIPCHECK(hipMemcpyAsync(h.data(), w.part_ref.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    double rs = 0, dg = 0, as = 0;
    for (int i = 0; i < nbk; ++i) {
        rs = std::max(rs, h[3 * i]);
        dg = std::max(dg, h[3 * i + 1]);
        as = std::max(as, h[3 * i + 2]);
    }
    rc.ok = dg > 0.0 && rs <= 1e-12 * dg && as <= 1e-12 * dg;
    return rc;
}

double deviceSum(prism_ctx* ctx, Workspace& w, int n, const double* x) {
    int nbk = ref_blocks(n);
    w.part_ref.alloc(std::max<size_t>(w.part_ref.n, (size_t)nbk));
    launch_block_sum(ctx->launch(), n, x, w.part_ref.p, w.part_ref.n);
    std::vector<double> h(nbk);
    HIPCHECK(hipMemcpyAsync(h.data(), w.part_ref.p, nbk * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    double s = 0.0;
    for (double v : h) s += v;
    return s;
}

// One BiCGSTAB::step of A' x = b' through the deflated system (see above).
StepOut krylovStepRef(prism_ctx* ctx, Workspace& w, const MatRef& A, const double* b, double* x, const RefCell& rc,
                      double ref, const prism_solve_params& p) {
    hipStream_t st = ctx->stream;
    Launch L = ctx->launch();
    w.aval.alloc(A.nnz);
    w.bt.alloc(A.n);
    HIPCHECK(hipMemcpyAsync(w.aval.p, A.val, A.nnz * sizeof(double), hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(w.aval.p + rc.slot, &rc.a, sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(w.bt.p, b, A.n * sizeof(double), hipMemcpyDeviceToDevice, st));
    double br = 0.0;
    HIPCHECK(hipMemcpyAsync(&br, b + rc.r, sizeof(double), hipMemcpyDeviceToHost, st));
    ctx->sync();
    const double b_r = br - rc.a * ref;                       // b_r before the fix
    const double sigma = deviceSum(ctx, w, A.n, b) - rc.a * ref;  // 1^T b
    const double bt_r = b_r - sigma;
    HIPCHECK(hipMemcpyAsync(w.bt.p + rc.r, &bt_r, sizeof(double), hipMemcpyHostToDevice, st));
    ctx->sync();
    MatRef Am = A;
    Am.val = w.aval.p;
    StepOut so = krylovStep(ctx, w, Am, w.bt.p, x, p, true, true);
    double yr = 0.0;
    HIPCHECK(hipMemcpyAsync(&yr, x + rc.r, sizeof(double), hipMemcpyDeviceToHost, st));
    ctx->sync();
    launch_add_const(L, A.n, x, sigma / rc.a + ref - yr);
    return so;
}

MatRef matOf(prism_eqn& e) {
    prism_mesh* m = e.mesh;
    MatRef A{(int)m->h.n, m->rp.p, m->col.p, m->diag.p, e.A.p, (int64_t)m->h.rp[m->h.n],
             m->sell_off.p, m->sell_col.p, m->csr2sell.p, (int64_t)m->h.sell_col.size(),
             m->h.sell_rel.empty() ? nullptr : m->sell_rel.p, m->csc_ptr.p, m->csc_row.p, m->csc_pos.p,
             nullptr, 0, {}};
    A.setDia(m->csr2dia.p, m->h.dia_off);
    A.sgs = &m->sgs;
    A.sbase = A.srel ? m->sell_base.p : nullptr;
    return A;
}

// ISolver::solve (solver.cpp:27-82)
prism_solve_result solveEqn(prism_eqn& e, const prism_solve_params& p) {
    prism_mesh* m = e.mesh;
    prism_ctx* ctx = m->ctx;
    Launch L = ctx->launch();
    prism_field& phi = *e.phi;
    prism_solve_result r{};
    updateCoeffs(e, true);
    phi.bump();  // relax() takes a non-const values() reference
    const bool has_ref = p.ref_cell >= 0 && p.ref_cell < m->h.n;
    if (has_ref) launch_ref_cell(L, m->dm, e.A.p, e.b.p, (int32_t)p.ref_cell, p.ref_value);
    MatRef A = matOf(e);
    RefCell rc;
    if (has_ref) rc = refCellDeflation(ctx, e.ws, A, (int32_t)p.ref_cell, m->h.diag[p.ref_cell]);
    const bool sym = p.precision == PRISM_PRECISION_FP32_REFINE && !rc.ok && isSymmetric(ctx, e.ws, A);
    double init = 0, cur = 0;
    bool conv = false;
    int64_t it = 0;
    for (; it < p.n_iter; it++) {
        if (it == 0) {
            phi.bump();
            init = scaledResidual(ctx, e.ws, A, phi.v.p, e.b.p);
        }
        phi.bump();
        StepOut so = rc.ok ? krylovStepRef(ctx, e.ws, A, e.b.p, phi.v.p, rc, p.ref_value, p)
                           : krylovStep(ctx, e.ws, A, e.b.p, phi.v.p, p, sym);
        r.krylov_iters += so.iters;
        r.refinements += so.refinements;
        r.restarts += so.restarts;
        r.krylov_rel_residual = so.rel;
        phi.bump();
        cur = scaledResidual(ctx, e.ws, A, phi.v.p, e.b.p);
        if (so.done == 3) cur = std::numeric_limits<double>::quiet_NaN();  // breakdown on a NaN
        if (cur < p.eps) {
            conv = true;
            it++;
            break;
        }
    }
    if (p.clear_coeffs) zeroCoeffs(e);
    r.iteration = it;
    r.initial_residual = init;
    r.final_residual = cur;
    r.converged = conv;
    return r;
}

}  // namespace

// ---------------------------------------------------------------------- C ABI ----
#define API extern "C" __attribute__((visibility("default")))

API const char* prism_last_error(void) { return g_err.c_str(); }

API int prism_ctx_create(int device, prism_ctx** out) {
    return guard([&] {
        int count = 0;
        HIPCHECK(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) fail(PRISM_ERR_INVALID, "no such device");
        HIPCHECK(hipSetDevice(device));
        auto c = std::make_unique<prism_ctx>();
        c->device = device;
        HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIPCHECK(hipHostMalloc((void**)&c->hks, sizeof(KrylovScalars), hipHostMallocDefault));
        c->scratch = static_cast<unsigned char*>(scratch_address(device));
        *out = c.release();
    });
}
API int prism_ctx_destroy(prism_ctx* c) {
    return guard([&] {
        if (!c) return;
        (void)hipStreamSynchronize(c->stream);
        c->drain();
        for (auto e : c->pool) (void)hipEventDestroy(e);
        (void)hipHostFree(c->hks);
        (void)hipStreamDestroy(c->stream);
        delete c;
    });
}
API int prism_ctx_info(prism_ctx* c, char* name, int32_t len, int32_t* n_cu) {
    return guard([&] {
        hipDeviceProp_t pr;
        HIPCHECK(hipGetDeviceProperties(&pr, c->device));
        if (name && len > 0) { std::strncpy(name, pr.gcnArchName, len - 1); name[len - 1] = 0; }
        if (n_cu) *n_cu = pr.multiProcessorCount;
    });
}
API int prism_ctx_set_timing(prism_ctx* c, int32_t on) {
    return guard([&] { c->drain(); c->timing = on != 0; });
}
API int prism_ctx_reset_timing(prism_ctx* c) {
    return guard([&] { c->drain(); c->stats.clear(); c->order.clear(); });
}
API int prism_ctx_kernel_stats(prism_ctx* c, int32_t idx, char* name, int32_t len, int64_t* launches, double* ms) {
    return guard([&] {
        c->drain();
        if (idx < 0 || idx >= (int32_t)c->order.size()) fail(PRISM_ERR_INVALID, "index past end");
        const auto& nm = c->order[idx];
        if (name && len > 0) { std::strncpy(name, nm.c_str(), len - 1); name[len - 1] = 0; }
        *launches = c->stats[nm].first;
        *ms = c->stats[nm].second;
    });
}
API int prism_ctx_synchronize(prism_ctx* c) { return guard([&] { c->sync(); }); }

// ----------------------------------------------------------------------- mesh ----
API int prism_mesh_create(prism_ctx* ctx, const prism_mesh_desc* d, prism_mesh** out) {
    return guard([&] {
        if (!d || !out) fail(PRISM_ERR_INVALID, "mesh create: null argument");
        auto m = std::make_unique<prism_mesh>();
        m->ctx = ctx;
        buildMesh(*d, m->h);  // the descriptor is validated first (PRISM_ERR_MESH), then the context
        if (!ctx) fail(PRISM_ERR_INVALID, "mesh create: null context");
        HMesh& h = m->h;
        hipStream_t s = ctx->stream;
        m->cx.upload(h.cx, s); m->cy.upload(h.cy, s); m->cz.upload(h.cz, s); m->vol.upload(h.vol, s);
        m->sx.upload(h.sx, s); m->sy.upload(h.sy, s); m->sz.upload(h.sz, s);
        m->fx.upload(h.fx, s); m->fy.upload(h.fy, s); m->fz.upload(h.fz, s);
        m->nx.upload(h.nx, s); m->ny.upload(h.ny, s); m->nz.upload(h.nz, s);
        m->dx.upload(h.dx, s); m->dy.upload(h.dy, s); m->dz.upload(h.dz, s);
        m->gc.upload(h.gc, s); m->kx.upload(h.kx, s); m->ky.upload(h.ky, s); m->kz.upload(h.kz, s);
        m->pinv.upload(h.pinv, s);
        m->owner.upload(h.owner, s); m->nbr.upload(h.nbr, s); m->fpatch.upload(h.fpatch, s);
        m->cf_off.upload(h.cf_off, s); m->cf.upload(h.cf, s);
        m->rp.upload(h.rp, s); m->col.upload(h.col, s); m->diag.upload(h.diag, s);
        m->slot_on.upload(h.slot_on, s); m->slot_no.upload(h.slot_no, s);
        if (!h.dup_slot.empty()) {
            m->dup_slot.upload(h.dup_slot, s);
            m->dup_ptr.upload(h.dup_ptr, s);
        }
        m->gg_cells.upload(h.gg_cells, s);
        m->sell_off.upload(h.sell_off, s); m->sell_col.upload(h.sell_col, s); m->csr2sell.upload(h.csr2sell, s);
        if (!h.sell_rel.empty()) {
            m->sell_rel.upload(h.sell_rel, s);
            m->sell_base.upload(h.sell_base, s);
        }
        m->csc_ptr.upload(h.csc_ptr, s); m->csc_row.upload(h.csc_row, s); m->csc_pos.upload(h.csc_pos, s);
        if (!h.csr2dia.empty()) m->csr2dia.upload(h.csr2dia, s);
        DMesh& dm = m->dm;
        dm.n = (int32_t)h.n; dm.F = (int32_t)h.F; dm.nb = (int32_t)h.nb; dm.nne = (int32_t)h.nne;
        dm.cx = m->cx.p; dm.cy = m->cy.p; dm.cz = m->cz.p; dm.vol = m->vol.p;
        dm.sx = m->sx.p; dm.sy = m->sy.p; dm.sz = m->sz.p;
        dm.fx = m->fx.p; dm.fy = m->fy.p; dm.fz = m->fz.p;
        dm.nx = m->nx.p; dm.ny = m->ny.p; dm.nz = m->nz.p;
        dm.dx = m->dx.p; dm.dy = m->dy.p; dm.dz = m->dz.p;
        dm.gc = m->gc.p; dm.kx = m->kx.p; dm.ky = m->ky.p; dm.kz = m->kz.p;
        dm.owner = m->owner.p; dm.nbr = m->nbr.p; dm.fpatch = m->fpatch.p;
        dm.cf_off = m->cf_off.p; dm.cf = m->cf.p;
        dm.rp = m->rp.p; dm.col = m->col.p; dm.diag = m->diag.p;
        dm.slot_on = m->slot_on.p; dm.slot_no = m->slot_no.p;
        dm.dup_slot = m->dup_slot.p; dm.dup_ptr = m->dup_ptr.p;
        dm.nnz = h.rp[h.n];
        dm.n_dup_groups = (int32_t)h.dup_slot.size();
        dm.pinv = m->pinv.p;
        ctx->sync();
        *out = m.release();
    });
}
API int prism_mesh_destroy(prism_mesh* m) { return guard([&] { delete m; }); }
API int prism_mesh_get_info(prism_mesh* m, prism_mesh_info* i) {
    return guard([&] {
        i->n_cells = m->h.n; i->n_faces = m->h.F; i->n_boundary_faces = m->h.nb;
        i->n_nonempty_boundary_faces = m->h.nne; i->nnz = m->h.rp[m->h.n];
        i->n_gg_levels = (int64_t)m->h.gg_off.size() - 1;
    });
}
API int prism_mesh_geometry(prism_mesh* m, double* cc, double* cv, double* fc, double* fs, int64_t* own,
                            int64_t* nbr, int32_t* fp) {
    return guard([&] {
        const HMesh& h = m->h;
        for (int64_t c = 0; c < h.n; ++c) {
            if (cc) { cc[3 * c] = h.cx[c]; cc[3 * c + 1] = h.cy[c]; cc[3 * c + 2] = h.cz[c]; }
            if (cv) cv[c] = h.vol[c];
        }
        for (int64_t f = 0; f < h.F; ++f) {
            if (fc) { fc[3 * f] = h.fx[f]; fc[3 * f + 1] = h.fy[f]; fc[3 * f + 2] = h.fz[f]; }
            if (fs) { fs[3 * f] = h.sx[f]; fs[3 * f + 1] = h.sy[f]; fs[3 * f + 2] = h.sz[f]; }
            if (own) own[f] = h.owner[f];
            if (nbr) nbr[f] = h.nbr[f];
            if (fp) fp[f] = h.fpatch[f];
        }
    });
}
API int prism_mesh_face_permutation(prism_mesh* m, int64_t* o2n) {
    return guard([&] { std::copy(m->h.old_to_new.begin(), m->h.old_to_new.end(), o2n); });
}

struct prism_meshgen {
    MeshGen g;
};
API int prism_meshgen_box(const prism_box_params* p, prism_meshgen** out) {
    return guard([&] {
        auto g = std::make_unique<prism_meshgen>();
        generateBox(*p, g->g);
        *out = g.release();
    });
}
API int prism_meshgen_unv(const char* path, prism_meshgen** out) {
    return guard([&] {
        auto g = std::make_unique<prism_meshgen>();
        readUnv(path, g->g);
        *out = g.release();
    });
}
API int prism_meshgen_desc(prism_meshgen* g, prism_mesh_desc* d) { return guard([&] { g->g.describe(*d); }); }
API int prism_meshgen_patch_name(prism_meshgen* g, int32_t idx, char* name, int32_t len) {
    return guard([&] {
        if (idx < 0 || idx >= (int32_t)g->g.patch_names.size()) fail(PRISM_ERR_INVALID, "patch index");
        std::strncpy(name, g->g.patch_names[idx].c_str(), len - 1);
        name[len - 1] = 0;
    });
}
API int prism_meshgen_destroy(prism_meshgen* g) { return guard([&] { delete g; }); }

// --------------------------------------------------------------------- fields ----
API int prism_scalar_create(prism_mesh* m, const char* name, int32_t kind, const double* values, double value,
                            prism_field** out) {
    return guard([&] {
        auto f = makeScalar(m, name, kind, -1, values, value);
        m->ctx->sync();
        *out = f.release();
    });
}
API int prism_field_destroy(prism_field* f) { return guard([&] { delete f; }); }
API int prism_scalar_set(prism_field* f, const double* v) {
    return guard([&] {
        f->bump();
        HIPCHECK(hipMemcpyAsync(f->v.p, v, f->n() * sizeof(double), hipMemcpyHostToDevice, f->mesh->ctx->stream));
        f->mesh->ctx->sync();
    });
}
API int prism_scalar_get(prism_field* f, double* v) {
    return guard([&] {
        HIPCHECK(hipMemcpyAsync(v, f->v.p, f->n() * sizeof(double), hipMemcpyDeviceToHost, f->mesh->ctx->stream));
        f->mesh->ctx->sync();
    });
}
API int prism_scalar_fill(prism_field* f, double value) {
    return guard([&] {
        f->bump();
        launch_fill(f->mesh->ctx->launch(), f->v.p, f->n(), value);
        f->mesh->ctx->sync();
    });
}
API int prism_scalar_set_face_values(prism_field* f, const double* fv) {
    return guard([&] {
        f->fv.upload(fv, f->mesh->h.F, f->mesh->ctx->stream);
        f->has_fv = true;
        f->mesh->ctx->sync();
    });
}
API int prism_scalar_clear_face_values(prism_field* f) { return guard([&] { f->has_fv = false; }); }
API int prism_scalar_grad(prism_field* f, double* g) {
    return guard([&] {
        ensureGrad(*f);
        int64_t n = f->n();
        std::vector<double> tmp(3 * n);
        HIPCHECK(hipMemcpyAsync(tmp.data(), f->g.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, f->mesh->ctx->stream));
        f->mesh->ctx->sync();
        for (int64_t c = 0; c < n; ++c) { g[3 * c] = tmp[c]; g[3 * c + 1] = tmp[n + c]; g[3 * c + 2] = tmp[2 * n + c]; }
    });
}
API int prism_scalar_face_values(prism_field* f, double* out) {
    return guard([&] {
        requireFaceValueBCs(*f, "prism::field::Scalar::valueAtFace");
        DevBuf<double> tmp;
        tmp.alloc(f->mesh->h.F);
        launch_scalar_face_values(f->mesh->ctx->launch(), f->mesh->dm, f->view(), tmp.p);
        HIPCHECK(hipMemcpyAsync(out, tmp.p, f->mesh->h.F * sizeof(double), hipMemcpyDeviceToHost, f->mesh->ctx->stream));
        f->mesh->ctx->sync();
    });
}
API int prism_scalar_set_history(prism_field* f, int32_t n) {
    return guard([&] {
        f->hist_max = n;
        while ((int32_t)f->hist.size() > n) f->hist.pop_back();
    });
}
API int prism_scalar_update_prev(prism_field* f) {  // history.cpp:30-48
    return guard([&] {
        if (f->hist_max == 0) return;
        std::unique_ptr<DevBuf<double>> b;
        if ((int32_t)f->hist.size() == f->hist_max) { b = std::move(f->hist.back()); f->hist.pop_back(); }
        else { b = std::make_unique<DevBuf<double>>(); b->alloc(f->n()); }
        launch_copy(f->mesh->ctx->launch(), b->p, f->v.p, f->n());
        f->hist.insert(f->hist.begin(), std::move(b));
    });
}

static std::unique_ptr<prism_vfield> makeVector(prism_mesh* m, const std::string& name) {
    auto v = std::make_unique<prism_vfield>();
    v->mesh = m;
    v->name = name;
    v->bc.build(m->h, name, m->ctx->stream);
    return v;
}
API int prism_vector_create(prism_mesh* m, const char* name, const double* values, const double* u3,
                            prism_vfield** out) {
    return guard([&] {
        auto v = makeVector(m, name);
        int64_t n = m->h.n;
        const char* sfx[3] = {"_x", "_y", "_z"};
        std::vector<double> tmp;
        for (int k = 0; k < 3; ++k) {
            if (values) {
                tmp.resize(n);
                for (int64_t i = 0; i < n; ++i) tmp[i] = values[3 * i + k];
            }
            v->comp[k] = makeScalar(m, std::string(name) + sfx[k], 0, k, values ? tmp.data() : nullptr, u3 ? u3[k] : 0.0);
            m->ctx->sync();
        }
        *out = v.release();
    });
}
API int prism_vector_clone(prism_vfield* src, prism_vfield** out) {
    return guard([&] {
        prism_mesh* m = src->mesh;
        auto v = makeVector(m, src->name);
        Launch L = m->ctx->launch();
        for (int k = 0; k < 3; ++k) {
            const prism_field& c = *src->comp[k];
            v->comp[k] = makeScalar(m, c.name, c.kind, c.coord, nullptr, 0.0);
            launch_copy(L, v->comp[k]->v.p, c.v.p, c.n());
            if (c.has_fv) {
                v->comp[k]->fv.alloc(m->h.F);
                launch_copy(L, v->comp[k]->fv.p, c.fv.p, m->h.F);
                v->comp[k]->has_fv = true;
            }
        }
        if (src->has_fd) { v->fd.alloc(3 * m->h.F); launch_copy(L, v->fd.p, src->fd.p, 3 * m->h.F); v->has_fd = true; }
        if (src->has_ff) { v->ff.alloc(m->h.F); launch_copy(L, v->ff.p, src->ff.p, m->h.F); v->has_ff = true; }
        m->ctx->sync();
        *out = v.release();
    });
}
API int prism_vfield_destroy(prism_vfield* v) { return guard([&] { delete v; }); }
API int prism_vector_component(prism_vfield* v, int32_t k, prism_field** out) {
    return guard([&] {
        if (k < 0 || k > 2) fail(PRISM_ERR_INVALID, "component index");
        *out = v->comp[k].get();
    });
}
API int prism_vector_set_face_values(prism_vfield* v, const double* fv) {
    return guard([&] {
        int64_t F = v->mesh->h.F;
        std::vector<double> soa(3 * F);
        for (int64_t f = 0; f < F; ++f) for (int k = 0; k < 3; ++k) soa[k * F + f] = fv[3 * f + k];
        v->fd.upload(soa, v->mesh->ctx->stream);
        v->has_fd = true;
        v->mesh->ctx->sync();
    });
}
API int prism_vector_set_face_fluxes(prism_vfield* v, const double* ff) {
    return guard([&] {
        v->ff.upload(ff, v->mesh->h.F, v->mesh->ctx->stream);
        v->has_ff = true;
        v->mesh->ctx->sync();
    });
}
API int prism_vector_face_fluxes(prism_vfield* v, double* out) {
    return guard([&] {
        prism_mesh* m = v->mesh;
        if (!v->has_ff) {
            requireKinds(m->h, v->bc, v->name, "prism::field::GeneralVector::fluxAtFace",
                         {BC_FIXED, BC_NOSLIP, BC_SYMMETRY, BC_OUTLET, BC_ZEROGRAD});
            if (!v->has_fd) for (int k = 0; k < 3; ++k) requireFaceValueBCs(*v->comp[k], "prism::field::GeneralVector");
        }
        DevBuf<double> tmp;
        tmp.alloc(m->h.F);
        launch_vector_flux(m->ctx->launch(), m->dm, v->view(), tmp.p);
        HIPCHECK(hipMemcpyAsync(out, tmp.p, m->h.F * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
        m->ctx->sync();
    });
}
API int prism_vector_face_values(prism_vfield* v, double* out) {
    return guard([&] {
        prism_mesh* m = v->mesh;
        if (!v->has_fd) {
            requireKinds(m->h, v->bc, v->name, "prism::field::GeneralVector::valueAtFace",
                         {BC_FIXED, BC_NOSLIP, BC_SYMMETRY, BC_OUTLET, BC_ZEROGRAD});
            for (int k = 0; k < 3; ++k) requireFaceValueBCs(*v->comp[k], "prism::field::GeneralVector");
        }
        int64_t F = m->h.F;
        DevBuf<double> tmp;
        tmp.alloc(3 * F);
        launch_vector_face_values(m->ctx->launch(), m->dm, v->view(), tmp.p);
        std::vector<double> soa(3 * F);
        HIPCHECK(hipMemcpyAsync(soa.data(), tmp.p, 3 * F * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
        m->ctx->sync();
        for (int64_t f = 0; f < F; ++f) for (int k = 0; k < 3; ++k) out[3 * f + k] = soa[k * F + f];
    });
}

API int prism_tensor_create(prism_mesh* m, const char* name, const double* values, const double* u9,
                            prism_tfield** out) {
    return guard([&] {
        auto t = std::make_unique<prism_tfield>();
        t->mesh = m;
        t->name = name;
        int64_t n = m->h.n;
        std::vector<double> soa(9 * n);
        bool diag = true;
        for (int64_t i = 0; i < n; ++i)
            for (int k = 0; k < 9; ++k) {
                double x = values ? values[9 * i + k] : u9[k];
                soa[k * n + i] = x;
                if (k != 0 && k != 4 && k != 8 && x != 0.0) diag = false;
            }
        t->diag_only = diag;
        t->t.upload(soa, m->ctx->stream);
        m->ctx->sync();
        *out = t.release();
    });
}
API int prism_tfield_destroy(prism_tfield* t) { return guard([&] { delete t; }); }
API int prism_tensor_get(prism_tfield* t, double* out) {
    return guard([&] {
        int64_t n = t->mesh->h.n;
        std::vector<double> soa(9 * n);
        HIPCHECK(hipMemcpyAsync(soa.data(), t->t.p, 9 * n * sizeof(double), hipMemcpyDeviceToHost, t->mesh->ctx->stream));
        t->mesh->ctx->sync();
        for (int64_t i = 0; i < n; ++i) for (int k = 0; k < 9; ++k) out[9 * i + k] = soa[k * n + i];
    });
}

// ------------------------------------------------------------------ equations ----
API int prism_eqn_create(const prism_scheme_desc* sd, int32_t n, prism_eqn** out) {
    return guard([&] {
        if (n < 1) fail(PRISM_ERR_RUNTIME, "eqn::Transport needs at least one scheme");
        auto e = std::make_unique<prism_eqn>();
        for (int32_t k = 0; k < n; ++k) {
            const prism_scheme_desc& d = sd[k];
            if (d.kind < PRISM_SCHEME_CONVECTION || d.kind > PRISM_SCHEME_ADAM_MOULTON) fail(PRISM_ERR_INVALID, "unknown scheme kind");
            if (d.kind == PRISM_SCHEME_BACKWARD_EULER && !(d.coeff > 0.0))
                fail(PRISM_ERR_INVALID, "prism::scheme::temporal::BackwardEuler(phi, dt): dt must be positive");
            if (d.kind == PRISM_SCHEME_ADAM_MOULTON && !(d.coeff > 0.0))
                fail(PRISM_ERR_INVALID, "prism::scheme::temporal::AdamMoulton(phi, dt): dt must be positive");
            if (d.kind == PRISM_SCHEME_CONVECTION && (!d.vec || !d.phi)) fail(PRISM_ERR_INVALID, "convection needs U and phi");
            if (d.kind == PRISM_SCHEME_DIFFUSION && (!d.phi || (!d.kappa && !d.kappa_t))) fail(PRISM_ERR_INVALID, "diffusion needs kappa and phi");
            if (d.kind == PRISM_SCHEME_CONVECTION && (d.variant < 0 || d.variant > 7)) fail(PRISM_ERR_INVALID, "NVF id");
            if (d.kind == PRISM_SCHEME_DIFFUSION && (d.variant < 0 || d.variant > 3)) fail(PRISM_ERR_INVALID, "corrector id");
            if (d.kind == PRISM_SCHEME_SRC_DIVERGENCE && !d.vec) fail(PRISM_ERR_INVALID, "divergence needs a vector");
            if ((d.kind == PRISM_SCHEME_SRC_CONSTANT || d.kind == PRISM_SCHEME_SRC_GRADIENT) && !d.phi) fail(PRISM_ERR_INVALID, "source needs a field");
            if (d.kind == PRISM_SCHEME_SRC_GRADIENT && (d.coord < 0 || d.coord > 2)) fail(PRISM_ERR_INVALID, "coordinate");
            if (k == 0) {
                if (!d.phi) fail(PRISM_ERR_RUNTIME, "the first scheme carries no field");
                e->phi = d.phi;
                e->mesh = d.phi->mesh;
            }
            if (d.kind != PRISM_SCHEME_SRC_DIVERGENCE && d.kind != PRISM_SCHEME_SRC_CONSTANT &&
                d.kind != PRISM_SCHEME_SRC_GRADIENT && d.phi && d.phi != e->phi && d.kind != PRISM_SCHEME_SRC_IMPLICIT &&
                d.kind != PRISM_SCHEME_BACKWARD_EULER && d.kind != PRISM_SCHEME_ADAM_MOULTON)
                fail(PRISM_ERR_RUNTIME, "all full schemes of one equation must transport the same field");
            e->schemes.push_back(d);
        }
        // coefficient storage: the CSR entries, then the extra slots of pairs of cells that
        // share several faces
        int64_t nnz = e->mesh->h.rp[e->mesh->h.n] + e->mesh->h.n_extra;
        e->A.alloc(nnz);
        e->b.alloc(e->mesh->h.n);
        zeroCoeffs(*e);
        e->mesh->ctx->sync();
        *out = e.release();
    });
}
API int prism_eqn_destroy(prism_eqn* e) { return guard([&] { delete e; }); }
API int prism_eqn_update_coeffs(prism_eqn* e) { return guard([&] { updateCoeffs(*e, false); e->mesh->ctx->sync(); }); }
API int prism_eqn_zero_coeffs(prism_eqn* e) { return guard([&] { zeroCoeffs(*e); }); }
API int prism_eqn_set_under_relax(prism_eqn* e, double f) {  // transport.cpp:42-52
    return guard([&] { if (!(f < 0.0 || f > 1.0)) e->urf = f; });
}
API int prism_eqn_relax(prism_eqn* e) {
    return guard([&] {
        e->phi->bump();
        launch_relax(e->mesh->ctx->launch(), e->mesh->dm, e->A.p, e->b.p, e->phi->v.p, e->urf);
        e->mesh->ctx->sync();
    });
}
API int prism_eqn_add_symmetry_handler(prism_eqn* e) {  // symmetry.cpp:54-82
    return guard([&] {
        if (e->phi->coord < 0)
            fail(PRISM_ERR_RUNTIME, "eqn::boundary::Symmetry<Transport>: field `" + e->phi->name + "` is not a vector component");
        if (e->sym_handler) return;
        const HMesh& h = e->mesh->h;
        std::vector<std::unique_ptr<prism_eqn::SymPatch>> out;
        for (size_t q = 0; q < h.patches.size(); ++q) {
            if (h.patches[q].empty || e->phi->bc.kind[q] != BC_SYMMETRY) continue;
            std::vector<int32_t> cells, off{0}, faces;
            std::map<int32_t, std::vector<int32_t>> by;
            for (int32_t f : h.patches[q].faces) {
                int32_t o = h.owner[f];
                auto it = by.find(o);
                if (it == by.end()) { cells.push_back(o); by[o] = {f}; }
                else it->second.push_back(f);
            }
            for (int32_t c : cells) {
                for (int32_t f : by[c]) faces.push_back(f);
                off.push_back((int32_t)faces.size());
            }
            auto sp = std::make_unique<prism_eqn::SymPatch>();
            sp->n = (int32_t)cells.size();
            hipStream_t st = e->mesh->ctx->stream;
            sp->cells.upload(cells, st);
            sp->off.upload(off, st);
            sp->faces.upload(faces, st);
            e->mesh->ctx->sync();  // the host vectors end with this scope
            out.push_back(std::move(sp));
        }
        e->sym = std::move(out);
        e->sym_handler = true;
    });
}
API int prism_needs_pressure_reference(prism_field* p, int32_t* out) {  // SIMPLE.cpp:110-122
    return guard([&] {
        const HMesh& h = p->mesh->h;
        requireKinds(h, p->bc, p->name, "prism::algo::needsPressureReference()",
                     {BC_FIXED, BC_NOSLIP, BC_SYMMETRY, BC_OUTLET, BC_ZEROGRAD, BC_FIXEDGRAD, BC_EMPTY, BC_OTHER});
        int32_t r = 1;
        for (size_t q = 0; q < h.patches.size(); ++q)
            if (!h.patches[q].empty && (p->bc.kind[q] == BC_FIXED || p->bc.kind[q] == BC_NOSLIP)) r = 0;
        *out = r;
    });
}
API prism_field* prism_eqn_field(prism_eqn* e) { return e ? e->phi : nullptr; }
API int prism_eqn_nnz(prism_eqn* e, int64_t* nnz, int64_t* n) {
    return guard([&] { *nnz = e->mesh->h.rp[e->mesh->h.n]; *n = e->mesh->h.n; });
}
API int prism_eqn_matrix(prism_eqn* e, int64_t* rp, int32_t* col, double* val) {
    return guard([&] {
        const HMesh& h = e->mesh->h;
        if (rp) for (int64_t i = 0; i <= h.n; ++i) rp[i] = h.rp[i];
        if (col) std::copy(h.col.begin(), h.col.end(), col);
        if (val)  // the CSR entries (the storage's extra slots of repeated cell pairs are zero here)
            HIPCHECK(hipMemcpyAsync(val, e->A.p, (size_t)h.rp[h.n] * sizeof(double), hipMemcpyDeviceToHost,
                                    e->mesh->ctx->stream));
        e->mesh->ctx->sync();
    });
}
API int prism_eqn_rhs(prism_eqn* e, double* b) {
    return guard([&] {
        HIPCHECK(hipMemcpyAsync(b, e->b.p, e->b.n * sizeof(double), hipMemcpyDeviceToHost, e->mesh->ctx->stream));
        e->mesh->ctx->sync();
    });
}
API int prism_eqn_diag(prism_eqn* e, double* d) {
    return guard([&] {
        std::vector<double> A(e->A.n);
        HIPCHECK(hipMemcpyAsync(A.data(), e->A.p, e->A.n * sizeof(double), hipMemcpyDeviceToHost, e->mesh->ctx->stream));
        e->mesh->ctx->sync();
        for (int64_t i = 0; i < e->mesh->h.n; ++i) d[i] = A[e->mesh->h.diag[i]];
    });
}

API void prism_solve_params_default(prism_solve_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->n_iter = 10;
    p->eps = 1e-7;
    p->ref_cell = -1;
    p->clear_coeffs = 1;
    p->precision = PRISM_PRECISION_FP64;
    p->precond = PRISM_PRECOND_NONE;
    p->tol = 1e-14;
    p->max_iters = 0;
    p->inner_tol = 0;
}
API int prism_solve(prism_eqn* e, const prism_solve_params* p, prism_solve_result* r) {
    return guard([&] {
        *r = solveEqn(*e, *p);
        e->mesh->ctx->sync();
    });
}

API int prism_eqn_step(prism_eqn* e, const prism_solve_params* p, prism_solve_result* r) {
    return guard([&] {
        e->phi->bump();
        MatRef A = matOf(*e);
        StepOut so = krylovStep(e->mesh->ctx, e->ws, A, e->b.p, e->phi->v.p, *p);
        e->phi->bump();
        std::memset(r, 0, sizeof(*r));
        r->iteration = 1;
        r->krylov_iters = so.iters;
        r->refinements = so.refinements;
        r->restarts = so.restarts;
        r->krylov_rel_residual = so.rel;
        r->converged = so.done == 1 || so.done == 4;
        e->mesh->ctx->sync();
    });
}

API int prism_eqn_step_noalloc(prism_eqn* e, const prism_solve_params* p, prism_solve_result* r) {
    return guard([&] {
        if (!e || !p || !r) fail(PRISM_ERR_INVALID, "prism_eqn_step_noalloc: null argument");
        if (p->precond == PRISM_PRECOND_SGS)
            fail(PRISM_ERR_INVALID, "prism_eqn_step_noalloc: the SGS preconditioner allocates its plan and vectors");
        prism_ctx* ctx = e->mesh->ctx;
        MatRef A = matOf(*e);
        const bool dia = A.ndia > 0 && !noDia();
        const int64_t vtot = dia ? (int64_t)A.ndia * A.n : A.stotal;
        Workspace w;
        w.bindScratch(ctx->scratch, SCRATCH_BYTES, A.n, vtot, p->precision == PRISM_PRECISION_FP32_REFINE,
                      p->precond == PRISM_PRECOND_JACOBI);
        e->phi->bump();
        StepOut so = krylovStep(ctx, w, A, e->b.p, e->phi->v.p, *p);
        e->phi->bump();
        std::memset(r, 0, sizeof(*r));
        r->iteration = 1;
        r->krylov_iters = so.iters;
        r->refinements = so.refinements;
        r->restarts = so.restarts;
        r->krylov_rel_residual = so.rel;
        r->converged = so.done == 1 || so.done == 4;
        ctx->sync();
    });
}

// ----------------------------------------------------------------- bare matrix ----
API int prism_csr_create(prism_ctx* ctx, int64_t n, const int64_t* rp, const int32_t* col, const double* val,
                         prism_csr** out) {
    return guard([&] {
        if (n <= 0 || rp[n] >= (1LL << 31)) fail(PRISM_ERR_INVALID, "matrix size");
        auto a = std::make_unique<prism_csr>();
        a->ctx = ctx;
        a->n = n;
        std::vector<int32_t> r32(n + 1), dg(n, -1);
        for (int64_t i = 0; i <= n; ++i) r32[i] = (int32_t)rp[i];
        for (int64_t i = 0; i < n; ++i) {
            for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
                if (col[k] < 0 || col[k] >= n) fail(PRISM_ERR_INVALID, "column index out of range");
                if (col[k] == i) dg[i] = (int32_t)k;
            }
        }
        a->rp.upload(r32, ctx->stream);
        a->col.upload(col, rp[n], ctx->stream);
        a->val.upload(val, rp[n], ctx->stream);
        a->diag.upload(dg, ctx->stream);
        a->all_diag = std::find(dg.begin(), dg.end(), -1) == dg.end();
        std::vector<int32_t> so, sc, c2s;
        buildSell(n, r32.data(), col, so, sc, c2s);
        a->sell_off.upload(so, ctx->stream);
        a->sell_col.upload(sc, ctx->stream);
        a->csr2sell.upload(c2s, ctx->stream);
        a->stotal = (int64_t)sc.size();
        std::vector<int16_t> rel;
        std::vector<int32_t> sbase;
        if (sellRelative(n, so, sc, rel, sbase)) {
            a->sell_rel.upload(rel, ctx->stream);
            a->sell_base.upload(sbase, ctx->stream);
        }
        std::vector<int32_t> cp, cr, cm;
        buildCsc(n, r32.data(), col, cp, cr, cm);
        a->csc_ptr.upload(cp, ctx->stream);
        a->csc_row.upload(cr, ctx->stream);
        a->csc_pos.upload(cm, ctx->stream);
        std::vector<int32_t> c2d;
        if (buildDia(n, r32.data(), col, DIA_MAX, a->dia_off, c2d)) a->csr2dia.upload(c2d, ctx->stream);
        ctx->sync();
        *out = a.release();
    });
}
API int prism_csr_destroy(prism_csr* a) { return guard([&] { delete a; }); }
API int prism_csr_spmv(prism_csr* a, const double* x, double* y) {
    return guard([&] {
        DevBuf<double> dx, dy;
        dx.upload(x, a->n, a->ctx->stream);
        dy.alloc(a->n);
        launch_spmv64(a->ctx->launch(), (int)a->n, a->rp.p, a->col.p, a->val.p, dx.p, dy.p);
        HIPCHECK(hipMemcpyAsync(y, dy.p, a->n * sizeof(double), hipMemcpyDeviceToHost, a->ctx->stream));
        a->ctx->sync();
    });
}
API int prism_csr_jacobi_step(prism_csr* a, const double* b, const double* xin, double* xout) {
    return guard([&] {  // jacobi.cpp:5-9: x + D^-1 (b - A x), on the device
        DevBuf<double> db, dx, dy;
        db.upload(b, a->n, a->ctx->stream);
        dx.upload(xin, a->n, a->ctx->stream);
        dy.alloc(a->n);
        launch_jacobi_step(a->ctx->launch(), (int)a->n, a->rp.p, a->col.p, a->val.p, a->diag.p, db.p, dx.p, dy.p);
        HIPCHECK(hipMemcpyAsync(xout, dy.p, a->n * sizeof(double), hipMemcpyDeviceToHost, a->ctx->stream));
        a->ctx->sync();
    });
}
API int prism_csr_bicgstab(prism_csr* a, const double* b, const double* xin, double* xout,
                           const prism_solve_params* p, prism_solve_result* r) {
    return guard([&] {
        if (!a || !p || !b || !xin || !xout) fail(PRISM_ERR_INVALID, "BiCGSTAB::step: null argument");
        DevBuf<double> db, dx;
        db.upload(b, a->n, a->ctx->stream);
        dx.upload(xin, a->n, a->ctx->stream);
        MatRef A{(int)a->n, a->rp.p, a->col.p, a->diag.p, a->val.p, (int64_t)a->val.n,
                 a->sell_off.p, a->sell_col.p, a->csr2sell.p, a->stotal, a->sell_rel.n ? a->sell_rel.p : nullptr,
                 a->csc_ptr.p, a->csc_row.p, a->csc_pos.p, nullptr, 0, {}};
        A.setDia(a->csr2dia.p, a->dia_off);
        A.all_diag = a->all_diag;
        A.sbase = A.srel ? a->sell_base.p : nullptr;
        A.sgs = &a->sgs;
        StepOut so = krylovStep(a->ctx, a->ws, A, db.p, dx.p, *p);
        HIPCHECK(hipMemcpyAsync(xout, dx.p, a->n * sizeof(double), hipMemcpyDeviceToHost, a->ctx->stream));
        a->ctx->sync();
        if (r) {
            std::memset(r, 0, sizeof(*r));
            r->iteration = 1;
            r->krylov_iters = so.iters;
            r->refinements = so.refinements;
            r->restarts = so.restarts;
            r->krylov_rel_residual = so.rel;
            r->converged = so.done == 1 || so.done == 4;
        }
    });
}

// ------------------------------------------------------------- SIMPLE / PISO ----
API void prism_simple_controls_default(prism_simple_controls* c) {  // SIMPLE.h:14-43
    std::memset(c, 0, sizeof(*c));
    c->momentum_urf = 0.7;
    c->momentum_max_iter = 3;
    c->momentum_residual = 1e-7;
    c->non_ortho_correctors = 2;
    c->pressure_urf = 0.3;
    c->pressure_max_iter = 5;
    c->pressure_residual = 1e-7;
    c->p_ref_cell = -1;
    c->p_ref_value = 0.0;
    c->momentum_precision = PRISM_PRECISION_FP64;
    c->pressure_precision = PRISM_PRECISION_FP64;
    c->precond = PRISM_PRECOND_NONE;
    c->momentum_tol = 1e-14;
    c->pressure_tol = 1e-14;
    c->max_krylov_iters = 0;
}
API void prism_piso_controls_default(prism_piso_controls* c) {  // PISO.h:11-49
    std::memset(c, 0, sizeof(*c));
    c->outer_iterations = 1;
    c->momentum_implicit_steps = 1;
    c->pressure_correction_steps = 2;
    prism_simple_controls_default(&c->simple);
    c->simple.momentum_urf = 1.0;
    c->simple.momentum_residual = 1e-4;
    c->simple.pressure_urf = 1.0;
    c->simple.pressure_residual = 1e-4;
}

namespace {

void addReport(std::vector<prism_report>& reps, const std::string& name, const prism_solve_result& r) {
    prism_report x{};
    std::strncpy(x.field_name, name.c_str(), 63);
    x.n_iterations = r.iteration;
    x.initial_residual = r.initial_residual;
    x.final_residual = r.final_residual;
    x.converged = r.converged;
    x.krylov_iters = r.krylov_iters;
    reps.push_back(x);
}

prism_solve_params paramsFrom(const prism_simple_controls& c, bool pressure) {
    prism_solve_params p;
    prism_solve_params_default(&p);
    p.precision = pressure ? c.pressure_precision : c.momentum_precision;
    p.tol = pressure ? c.pressure_tol : c.momentum_tol;
    p.precond = c.precond;
    p.max_iters = c.max_krylov_iters;
    return p;
}

// The pressure-correction equation as solvePressureEquation hands it to solver.solve()
// (SIMPLE.cpp:124-178): D = V / aP from the momentum diagonals (SIMPLE.cpp:77-108, the
// predictors' coefficients are then zeroed), Rhie-Chow face fluxes in mdot, p' with the name
// of p and constrained face values (SIMPLE.cpp:43-75), pEqn = Corrected(D, p') +
// Divergence<Negative>(mdot), and the solve parameters with the automatic reference cell
// resolved (SIMPLE.cpp:110-122).  The equation owns p' and D.
struct PressureSystem {
    std::unique_ptr<prism_eqn> owned;  // a system handed to the caller
    prism_eqn* pe{nullptr};            // the system (owned, or the field's cached one)
    prism_solve_params sp;
};

// The pressure-correction system of one SIMPLE step (SIMPLE.cpp:124-196): D from the momentum
// diagonals, mdot from Rhie-Chow, p' = 0 with its constrained face values, and the equation
// laplacian(D, p') - div(mdot).  cached: the equation, p' and D live in p->pcache between
// steps (same mdot); p' and its gradient cache are reset to zero and D recomputed, which is
// what a fresh system holds, without allocating it again.
PressureSystem buildPressureSystem(prism_eqn* const* mom, int32_t nm, prism_vfield* U, prism_vfield* mdot,
                                   prism_field* p, prism_simple_controls c, bool cached) {
    prism_mesh* m = p->mesh;
    const HMesh& h = m->h;
    Launch L = m->ctx->launch();
    const bool reuse = cached && p->pcache && p->pcache_mdot == mdot;
    std::unique_ptr<prism_tfield> Dnew;
    std::unique_ptr<prism_field> ppnew;
    prism_tfield* D = nullptr;
    prism_field* pp = nullptr;
    if (reuse) {
        D = p->pcache->own_D.get();
        pp = p->pcache->own_phi.get();
    } else {
        Dnew = std::make_unique<prism_tfield>();
        D = Dnew.get();
        D->mesh = m;
        D->name = "D";
        D->t.alloc(9 * h.n);
        D->diag_only = true;
    }
    const double* Ad[3] = {mom[0]->A.p, mom[1]->A.p, nm == 3 ? mom[2]->A.p : nullptr};
    launch_pressure_D(L, m->dm, Ad, nm, D->t.p);
    for (int32_t k = 0; k < nm; ++k) zeroCoeffs(*mom[k]);
    // mdot->mapFaceValues(U.valueAtFace) then Rhie-Chow on interior faces
    requireKinds(h, U->bc, U->name, "prism::field::GeneralVector::valueAtFace",
                 {BC_FIXED, BC_NOSLIP, BC_SYMMETRY, BC_OUTLET, BC_ZEROGRAD});
    for (int k = 0; k < 3; ++k) requireFaceValueBCs(*U->comp[k], "prism::field::GeneralVector");
    mdot->fd.alloc(3 * h.F);
    mdot->ff.alloc(h.F);
    launch_map_face_values(L, m->dm, U->view(), mdot->fd.p, mdot->ff.p);
    mdot->has_fd = mdot->has_ff = true;
    ensureGrad(*p);
    launch_rhie_chow(L, m->dm, U->view(), D->view(), p->view(), mdot->fd.p, mdot->ff.p);
    if (reuse) {  // the state makeScalar gives a new p'
        launch_fill(L, pp->v.p, h.n, 0.0);
        launch_fill(L, pp->g.p, 3 * h.n, 0.0);
        pp->bump();
        pp->has_fv = false;
    } else {
        ppnew = makeScalar(m, p->name, 1, -1, nullptr, 0.0);
        pp = ppnew.get();
    }
    requireKinds(h, pp->bc, pp->name, "prism::algo::constrainPPrime()",
                 {BC_FIXED, BC_NOSLIP, BC_SYMMETRY, BC_OUTLET, BC_ZEROGRAD});
    pp->fv.alloc(h.F);
    launch_constrain_pprime(L, m->dm, pp->view(), pp->fv.p);
    pp->has_fv = true;
    if (c.p_ref_cell < 0) {  // needsPressureReference (SIMPLE.cpp:110-122)
        bool dir = false;
        for (size_t q = 0; q < h.patches.size(); ++q)
            if (!h.patches[q].empty && (pp->bc.kind[q] == BC_FIXED || pp->bc.kind[q] == BC_NOSLIP)) dir = true;
        if (!dir) c.p_ref_cell = 0;
    }
    PressureSystem ps;
    if (reuse) {
        ps.pe = p->pcache.get();
        zeroCoeffs(*ps.pe);
    } else {
        ps.owned = std::make_unique<prism_eqn>();
        prism_eqn& pe = *ps.owned;
        pe.mesh = m;
        pe.phi = pp;
        prism_scheme_desc lap{};
        lap.kind = PRISM_SCHEME_DIFFUSION;
        lap.variant = PRISM_DIFF_OVERRELAXED;
        lap.phi = pp;
        lap.kappa_t = D;
        prism_scheme_desc dv{};
        dv.kind = PRISM_SCHEME_SRC_DIVERGENCE;
        dv.variant = PRISM_SIGN_NEGATIVE;
        dv.vec = mdot;
        pe.schemes = {lap, dv};
        pe.A.alloc(h.rp[h.n] + h.n_extra);
        pe.b.alloc(h.n);
        zeroCoeffs(pe);
        pe.own_phi = std::move(ppnew);
        pe.own_D = std::move(Dnew);
        if (cached) {
            p->pcache.reset(ps.owned.release());
            p->pcache_mdot = mdot;
            ps.pe = p->pcache.get();
        } else {
            ps.pe = ps.owned.get();
        }
    }
    ps.sp = paramsFrom(c, true);
    ps.sp.n_iter = c.pressure_max_iter;
    ps.sp.eps = c.pressure_residual;
    ps.sp.ref_cell = c.p_ref_cell;
    ps.sp.ref_value = c.p_ref_value;
    ps.sp.clear_coeffs = 1;
    return ps;
}

// solvePressureEquation + correctFields (SIMPLE.cpp:124-196, 218-241)
void pressureCorrection(prism_eqn* const* mom, int32_t nm, prism_vfield* U, prism_vfield* mdot, prism_field* p,
                        prism_simple_controls c, std::vector<prism_report>& reps) {
    prism_mesh* m = p->mesh;
    Launch L = m->ctx->launch();
    PressureSystem ps = buildPressureSystem(mom, nm, U, mdot, p, c, true);
    prism_eqn& pe = *ps.pe;
    prism_field* pp = pe.phi;
    const prism_tfield& D = *pe.own_D;
    for (int64_t i = 0; i <= c.non_ortho_correctors; ++i) {
        prism_solve_result r = solveEqn(pe, ps.sp);
        addReport(reps, pp->name + "_corr_" + std::to_string(i), r);
    }
    pp->has_fv = false;  // clearFaceValues (SIMPLE.cpp:193)
    // correctFields
    ensureGrad(*pp);
    launch_correct_U(L, m->dm, U->comp[0]->v.p, U->comp[1]->v.p, U->comp[2]->v.p, D.view(), pp->view());
    for (int k = 0; k < 3; ++k) U->comp[k]->bump();
    launch_correct_mdot(L, m->dm, mdot->fd.p, mdot->ff.p, D.view(), pp->view());
    // p->update(p + urf p') (scalar.cpp:311-320)
    p->bump();
    if (p->hist_max > 0) {
        std::unique_ptr<DevBuf<double>> b;
        if ((int32_t)p->hist.size() == p->hist_max) { b = std::move(p->hist.back()); p->hist.pop_back(); }
        else { b = std::make_unique<DevBuf<double>>(); b->alloc(m->h.n); }
        launch_copy(L, b->p, p->v.p, m->h.n);
        p->hist.insert(p->hist.begin(), std::move(b));
    }
    launch_update_p(L, p->v.p, pp->v.p, c.pressure_urf, m->h.n);
    p->bump();
    p->has_fv = false;
    m->ctx->sync();
}

void checkMomentum(int32_t nm) {
    if (nm != 2 && nm != 3)
        fail(PRISM_ERR_RUNTIME, "prism::algo::IncompressibleSIMPLE::step() expects 2 or 3 momentum predictors, not " +
                                    std::to_string(nm));
}

// solveImplicitMomentum of every predictor (SIMPLE.cpp:19-41)
void momentumPredictor(prism_eqn* const* mom, int32_t nm, const prism_simple_controls& c,
                       std::vector<prism_report>& reps) {
    for (int32_t k = 0; k < nm; ++k) {
        prism_eqn* e = mom[k];
        if (!(c.momentum_urf < 0.0 || c.momentum_urf > 1.0)) e->urf = c.momentum_urf;
        prism_solve_params sp = paramsFrom(c, false);
        sp.n_iter = c.momentum_max_iter;
        sp.eps = c.momentum_residual;
        sp.clear_coeffs = 0;
        addReport(reps, e->phi->name, solveEqn(*e, sp));
    }
}

std::vector<prism_report> simpleStep(prism_eqn* const* mom, int32_t nm, prism_vfield* U, prism_vfield* mdot,
                                     prism_field* p, const prism_simple_controls& c) {
    checkMomentum(nm);
    std::vector<prism_report> reps;
    momentumPredictor(mom, nm, c, reps);
    pressureCorrection(mom, nm, U, mdot, p, c, reps);
    return reps;
}

std::vector<prism_report> primeStep(prism_eqn* const* mom, int32_t nm, prism_vfield* U, prism_vfield* mdot,
                                    prism_field* p, const prism_simple_controls& c) {
    checkMomentum(nm);
    std::vector<prism_report> reps;
    for (int32_t k = 0; k < nm; ++k) {  // solveExplicitMomentum (PRIME.cpp:39-56)
        prism_eqn* e = mom[k];
        updateCoeffs(*e, false);
        prism_mesh* m = e->mesh;
        e->tmp.alloc(m->h.n);
        e->phi->bump();
        launch_explicit_momentum(m->ctx->launch(), m->dm, e->A.p, e->b.p, e->phi->v.p, e->tmp.p);
        e->phi->bump();
        m->ctx->sync();
    }
    pressureCorrection(mom, nm, U, mdot, p, c, reps);
    return reps;
}

void copyReports(const std::vector<prism_report>& reps, prism_report* out, int32_t maxr, int32_t* nr) {
    int32_t k = 0;
    for (; k < (int32_t)reps.size() && k < maxr; ++k) out[k] = reps[k];
    if (nr) *nr = k;
}

}  // namespace

API int prism_simple_step(prism_eqn* const* mom, int32_t nm, prism_vfield* U, prism_vfield* mdot, prism_field* p,
                          const prism_simple_controls* c, prism_report* rep, int32_t maxr, int32_t* nr) {
    return guard([&] { copyReports(simpleStep(mom, nm, U, mdot, p, *c), rep, maxr, nr); });
}
API int prism_simple_pressure_system(prism_eqn* const* mom, int32_t nm, prism_vfield* U, prism_vfield* mdot,
                                    prism_field* p, const prism_simple_controls* c, prism_report* rep, int32_t maxr,
                                    int32_t* nr, prism_eqn** out, int64_t* ref_cell) {
    return guard([&] {
        if (!mom || !U || !mdot || !p || !c || !out) fail(PRISM_ERR_INVALID, "simple_pressure_system: null argument");
        checkMomentum(nm);
        std::vector<prism_report> reps;
        momentumPredictor(mom, nm, *c, reps);
        PressureSystem ps = buildPressureSystem(mom, nm, U, mdot, p, *c, false);
        p->mesh->ctx->sync();
        if (ref_cell) *ref_cell = ps.sp.ref_cell;
        *out = ps.owned.release();
        if (rep && nr) copyReports(reps, rep, maxr, nr);
    });
}
API int prism_prime_step(prism_eqn* const* mom, int32_t nm, prism_vfield* U, prism_vfield* mdot, prism_field* p,
                         const prism_simple_controls* c, prism_report* rep, int32_t maxr, int32_t* nr) {
    return guard([&] { copyReports(primeStep(mom, nm, U, mdot, p, *c), rep, maxr, nr); });
}
API int prism_piso_step(prism_eqn* const* mom, int32_t nm, prism_vfield* U, prism_vfield* mdot, prism_field* p,
                        const prism_piso_controls* c, prism_report* rep, int32_t maxr, int32_t* nr) {
    return guard([&] {  // PISO.cpp:10-55
        std::vector<prism_report> all;
        for (int64_t i = 0; i < c->outer_iterations; ++i) {
            for (int64_t j = 0; j < c->momentum_implicit_steps; ++j) {
                auto r = simpleStep(mom, nm, U, mdot, p, c->simple);
                all.insert(all.end(), r.begin(), r.end());
            }
            for (int64_t j = 1; j < c->pressure_correction_steps; ++j) {
                auto r = primeStep(mom, nm, U, mdot, p, c->simple);
                all.insert(all.end(), r.begin(), r.end());
            }
        }
        copyReports(all, rep, maxr, nr);
    });
}
// ----------------------------------------------------------------- turbulence ----
API int prism_kepsilon_update(prism_field* k, prism_field* eps, prism_field* nu, prism_vfield* U, const double* consts,
                              prism_field* const* out) {
    return guard([&] {
        if (!k || !eps || !nu || !U || !out) fail(PRISM_ERR_INVALID, "IncompressibleKEpsilon: null argument");
        prism_mesh* m = k->mesh;
        if (eps->mesh != m || nu->mesh != m || U->mesh != m)
            fail(PRISM_ERR_MESH, "IncompressibleKEpsilon: k, epsilon, nu and U must share one mesh");
        for (int i = 0; i < 6; ++i) {
            if (!out[i] || out[i]->mesh != m) fail(PRISM_ERR_MESH, "IncompressibleKEpsilon: output field on another mesh");
            if (out[i] == k || out[i] == eps || out[i] == nu || out[i] == U->comp[0].get() ||
                out[i] == U->comp[1].get() || out[i] == U->comp[2].get())
                fail(PRISM_ERR_INVALID, "IncompressibleKEpsilon: an output field is also an input");
            for (int j = 0; j < i; ++j)
                if (out[j] == out[i]) fail(PRISM_ERR_INVALID, "IncompressibleKEpsilon: an output field is given twice");
        }
        ensureGradVec(*U);
        DKEps a{};
        a.n = (int)m->h.n;
        a.k = k->v.p; a.eps = eps->v.p; a.nu = nu->v.p;
        for (int j = 0; j < 3; ++j) a.g[j] = U->comp[j]->g.p;
        const double dflt[5] = {0.09, 1.44, 1.92, 1.0, 1.3};
        const double* c = consts ? consts : dflt;
        a.c_mu = c[0]; a.c_eps1 = c[1]; a.c_eps2 = c[2]; a.sigma_k = c[3]; a.sigma_eps = c[4];
        double** dst[6] = {&a.nu_t, &a.pk, &a.nu_eff_k, &a.nu_eff_eps, &a.eps_source, &a.eps_sink};
        for (int i = 0; i < 6; ++i) {
            out[i]->bump();
            *dst[i] = out[i]->v.p;
        }
        launch_kepsilon(m->ctx->launch(), a);
        m->ctx->sync();
    });
}

// --------------------------------------------------------------------- output ----
namespace {

// Interleaved [n_vertices][ncomp] point values of ncomp cell arrays (export.cpp:79-112).
std::vector<double> cellToPoint(prism_mesh* m, int ncomp, const DPointSrc& src) {
    const HMesh& h = m->h;
    hipStream_t st = m->ctx->stream;
    const int64_t nv = (int64_t)h.v2c_off.size() - 1;
    if (!m->idw.p) {
        std::vector<double> w;
        idwWeights(h, w);
        const int64_t S = (nv + 63) / 64;
        std::vector<int32_t> sbase(S + 1, 0);
        for (int64_t s = 0; s < S; ++s) {
            int32_t width = 0;
            for (int64_t v = 64 * s; v < std::min(nv, 64 * s + 64); ++v)
                width = std::max(width, h.v2c_off[v + 1] - h.v2c_off[v]);
            if ((int64_t)sbase[s] + 64LL * width >= (1LL << 31))
                fail(PRISM_ERR_RUNTIME, "vertex adjacency too large for 32-bit indices");
            sbase[s + 1] = sbase[s] + 64 * width;
        }
        std::vector<int32_t> cell(sbase[S], 0);
        std::vector<double> ws(sbase[S], 0.0);
        for (int64_t v = 0; v < nv; ++v)
            for (int32_t i = h.v2c_off[v], j = 0; i < h.v2c_off[v + 1]; ++i, ++j) {
                int64_t at = sbase[v >> 6] + 64LL * j + (v & 63);
                cell[at] = h.v2c[i];
                ws[at] = w[i];
            }
        m->v2c_off.upload(h.v2c_off, st);
        m->v2c_sbase.upload(sbase, st);
        m->v2c_cell.upload(cell, st);
        m->idw.upload(ws, st);
        m->ctx->sync();
    }
    if (m->pts.n < (size_t)(nv * ncomp)) m->pts.alloc(nv * ncomp);
    launch_cell_to_point(m->ctx->launch(), (int)nv, ncomp, m->v2c_off.p, m->v2c_sbase.p, m->v2c_cell.p, m->idw.p,
                         src, m->pts.p);
    std::vector<double> out(nv * ncomp);
    HIPCHECK(hipMemcpyAsync(out.data(), m->pts.p, out.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    m->ctx->sync();
    return out;
}
DPointSrc srcOf(const prism_field* f) { DPointSrc s{}; s.c[0] = f->v.p; return s; }
DPointSrc srcOf(const prism_vfield* v) {
    DPointSrc s{};
    for (int k = 0; k < 3; ++k) s.c[k] = v->comp[k]->v.p;
    return s;
}
DPointSrc srcOf(const prism_tfield* t) {
    DPointSrc s{};
    for (int k = 0; k < 9; ++k) s.c[k] = t->t.p + k * t->mesh->h.n;
    return s;
}
// Cell values interleaved per cell (flattenScalar / flattenVector / flattenTensor, export.cpp:131-182).
std::vector<double> cellData(prism_mesh* m, int ncomp, const DPointSrc& src) {
    const int64_t n = m->h.n;
    std::vector<double> soa(ncomp * n), out(ncomp * n);
    for (int k = 0; k < ncomp; ++k)
        HIPCHECK(hipMemcpyAsync(soa.data() + k * n, src.c[k], n * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
    m->ctx->sync();
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < ncomp; ++k) out[i * ncomp + k] = soa[k * n + i];
    return out;
}

}  // namespace

API int prism_mesh_n_vertices(prism_mesh* m, int64_t* nv) {
    return guard([&] { *nv = (int64_t)m->h.vtx.size() / 3; });
}
API int prism_field_to_points(prism_field* f, double* out) {
    return guard([&] {
        if (!f || !out) fail(PRISM_ERR_INVALID, "cell -> point: null argument");
        auto p = cellToPoint(f->mesh, 1, srcOf(f));
        std::memcpy(out, p.data(), p.size() * sizeof(double));
    });
}
API int prism_vfield_to_points(prism_vfield* v, double* out) {
    return guard([&] {
        if (!v || !out) fail(PRISM_ERR_INVALID, "cell -> point: null argument");
        auto p = cellToPoint(v->mesh, 3, srcOf(v));
        std::memcpy(out, p.data(), p.size() * sizeof(double));
    });
}
API int prism_tfield_to_points(prism_tfield* t, double* out) {
    return guard([&] {
        if (!t || !out) fail(PRISM_ERR_INVALID, "cell -> point: null argument");
        auto p = cellToPoint(t->mesh, 9, srcOf(t));
        std::memcpy(out, p.data(), p.size() * sizeof(double));
    });
}
API int prism_export_vtu(const prism_export_item* items, int32_t n_items, const char* file_name, int32_t mode) {
    return guard([&] {
        if (n_items <= 0) return;  // toVtu with no fields skips the export (export.cpp:273-275)
        if (!items || !file_name) fail(PRISM_ERR_INVALID, "toVtu: null argument");
        if (mode != PRISM_WRITE_ASCII && mode != PRISM_WRITE_RAW_BINARY_COMPRESSED)
            fail(PRISM_ERR_INVALID, "unsupported WriteMode value " + std::to_string(mode));
        auto meshOf = [](const prism_export_item& it) -> prism_mesh* {
            switch (it.kind) {
                case PRISM_EXPORT_SCALAR: return static_cast<const prism_field*>(it.field)->mesh;
                case PRISM_EXPORT_VECTOR: return static_cast<const prism_vfield*>(it.field)->mesh;
                case PRISM_EXPORT_TENSOR: return static_cast<const prism_tfield*>(it.field)->mesh;
                default: fail(PRISM_ERR_INVALID, "export item kind must be 1, 3 or 9");
            }
        };
        for (int32_t i = 0; i < n_items; ++i)
            if (!items[i].field) fail(PRISM_ERR_INVALID, "export item without a field");
        prism_mesh* m = meshOf(items[0]);
        const HMesh& h = m->h;
        if (h.cell_vtx_off.empty())
            fail(PRISM_ERR_MESH, "VTU export needs the cells' vertex lists (prism_mesh_desc.cell_vtx)");
        std::vector<VtuArray> arrays;
        for (int32_t i = 0; i < n_items; ++i) {
            const auto& it = items[i];
            std::string name;
            DPointSrc src;
            if (it.kind == PRISM_EXPORT_SCALAR) {
                auto* f = static_cast<const prism_field*>(it.field);
                name = f->name; src = srcOf(f);
            } else if (it.kind == PRISM_EXPORT_VECTOR) {
                auto* v = static_cast<const prism_vfield*>(it.field);
                name = v->name; src = srcOf(v);
            } else {
                auto* t = static_cast<const prism_tfield*>(it.field);
                name = t->name; src = srcOf(t);
            }
            if (meshOf(it) != m)
                fail(PRISM_ERR_MESH, "Field '" + name + "' does not share the mesh of the other exported fields.");
            arrays.push_back({name, false, it.kind, cellData(m, it.kind, src)});
            arrays.push_back({name, true, it.kind, cellToPoint(m, it.kind, src)});
        }
        writeVtu(file_name, mode, h.vtx, h.cell_vtx_off, h.cell_vtx, arrays);
    });
}
API int prism_write_pvd(const char* file_name, int32_t n, const double* times, const char* const* files) {
    return guard([&] {
        if (!file_name || n < 0 || (n > 0 && (!times || !files))) fail(PRISM_ERR_INVALID, "flushPvd: null argument");
        writePvd(file_name, n, times, files);
    });
}

API int prism_courant(prism_vfield* mdot, double dt, double* co) {
    return guard([&] {
        if (!mdot || !co) fail(PRISM_ERR_INVALID, "courantNumber: null argument");
        prism_mesh* m = mdot->mesh;
        DevBuf<double> ff, out;
        const double* fp;
        if (mdot->has_ff) fp = mdot->ff.p;
        else {
            ff.alloc(m->h.F);
            launch_vector_flux(m->ctx->launch(), m->dm, mdot->view(), ff.p);
            fp = ff.p;
        }
        out.alloc(m->h.n);
        launch_courant(m->ctx->launch(), m->dm, fp, dt, out.p);
        HIPCHECK(hipMemcpyAsync(co, out.p, m->h.n * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
        m->ctx->sync();
    });
}

// ------------------------------------------------------------------- adjoint ----
// Discrete adjoint of R(T; s, kappa) = A T - b for Upwind + NonCorrected(scalar kappa) +
// ConstantScalar(s) equations (no reference counterpart).  The caller has solved the primal
// equation; this re-assembles A at the current T, solves A^T lambda = dJ/dT with BiCGSTAB on
// the transposed matrix and forms dJ/ds = -lambda^T dR/ds, dJ/dkappa = -lambda^T dR/dkappa.
namespace {

struct AdjointTerms {
    const prism_scheme_desc* diff{nullptr};
    const prism_scheme_desc* src{nullptr};
};

AdjointTerms adjointTerms(const prism_eqn& e, prism_field* source, prism_field* kappa) {
    AdjointTerms t;
    for (const auto& s : e.schemes) {
        if (s.kind == PRISM_SCHEME_CONVECTION && s.variant == PRISM_NVF_UPWIND) continue;
        if (s.kind == PRISM_SCHEME_DIFFUSION && s.kappa && !s.kappa_t) {
            if (t.diff) fail(PRISM_ERR_INVALID, "adjoint: more than one diffusion scheme");
            t.diff = &s;
            continue;
        }
        if (s.kind == PRISM_SCHEME_SRC_CONSTANT) {
            if (s.phi == source) {
                if (t.src) fail(PRISM_ERR_INVALID, "adjoint: the source field enters more than one source term");
                t.src = &s;
            }
            continue;
        }
        fail(PRISM_ERR_INVALID,
             "adjoint: the equation may hold only convection::Upwind, one diffusion term with a scalar kappa "
             "and source::ConstantScalar terms");
    }
    if (t.diff && t.diff->variant != PRISM_DIFF_NONCORRECTED) {  // the correction's gradient enters dR/dphi
        if (e.phi->gg) fail(PRISM_ERR_INVALID, "adjoint: Corrected diffusion needs least-squares gradients of the field");
        if (e.phi->has_fv) fail(PRISM_ERR_INVALID, "adjoint: Corrected diffusion needs a field without face values");
        if (t.diff->kappa->has_fv) fail(PRISM_ERR_INVALID, "adjoint: kappa must not carry face values");
        requireKinds(e.mesh->h, e.phi->bc, e.phi->name, "adjoint of Corrected diffusion",
                     {BC_FIXED, BC_NOSLIP, BC_SYMMETRY, BC_OUTLET, BC_ZEROGRAD, BC_EMPTY});
    }
    if (kappa && (!t.diff || t.diff->kappa != kappa))
        fail(PRISM_ERR_INVALID, "adjoint: kappa is not the diffusion coefficient of the equation");
    if (kappa && kappa->has_fv) fail(PRISM_ERR_INVALID, "adjoint: kappa must not carry face values");
    if (source && !t.src) fail(PRISM_ERR_INVALID, "adjoint: the source field is not a ConstantScalar term");
    return t;
}

double objectiveOf(prism_eqn& e, const prism_objective& obj, double* VR) {
    prism_mesh* m = e.mesh;
    if (obj.axis < 0 || obj.axis > 2) fail(PRISM_ERR_INVALID, "objective axis must be 0, 1 or 2");
    int nb = adjoint_objective_blocks(m->dm);
    DevBuf<double> part;
    part.alloc(2 * (size_t)nb);
    launch_objective(m->ctx->launch(), m->dm, e.phi->v.p, obj.axis, obj.threshold, part.p, part.n);
    std::vector<double> h(2 * (size_t)nb);
    HIPCHECK(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
    m->ctx->sync();
    double a = 0.0, v = 0.0;
    for (int i = 0; i < nb; ++i) {
        a += h[2 * i];
        v += h[2 * i + 1];
    }
    if (!(v > 0.0)) fail(PRISM_ERR_INVALID, "objective region holds no cell");
    *VR = v;
    return a / v;
}

}  // namespace

API int prism_adjoint_objective(prism_eqn* e, const prism_objective* obj, double* J) {
    return guard([&] {
        if (!e || !obj || !J) fail(PRISM_ERR_INVALID, "adjoint objective: null argument");
        double VR;
        *J = objectiveOf(*e, *obj, &VR);
    });
}

API int prism_adjoint_gradient(prism_eqn* e, prism_field* source, prism_field* kappa, const prism_objective* obj,
                               const prism_solve_params* p, double* J, double* dJds, double* dJdkappa,
                               prism_solve_result* r) {
    return guard([&] {
        if (!e || !obj || !p || !J || !r) fail(PRISM_ERR_INVALID, "adjoint gradient: null argument");
        prism_mesh* m = e->mesh;
        const HMesh& h = m->h;
        AdjointTerms t = adjointTerms(*e, source, kappa);
        // A at the converged state (the dirty flag decides accumulation: start clean)
        zeroCoeffs(*e);
        updateCoeffs(*e, false);
        double VR;
        *J = objectiveOf(*e, *obj, &VR);
        Launch L = m->ctx->launch();
        DevBuf<double> g, lam, AT, scr, out;
        g.alloc(h.n);
        lam.alloc(h.n);
        AT.alloc(h.rp[h.n]);
        launch_objective_grad(L, m->dm, e->phi->v.p, obj->axis, obj->threshold, VR, g.p);
        launch_fill(L, lam.p, h.n, 0.0);
        launch_transpose(L, m->dm, e->A.p, AT.p);
        MatRef A = matOf(*e);
        A.val = AT.p;
        Workspace ws;
        const bool corrected = t.diff && t.diff->variant != PRISM_DIFF_NONCORRECTED;
        const double tol = p->tol > 0 ? p->tol : 1e-14;
        // Corrected: every A^T solve is one step of the defect correction below, so it needs
        // only an accuracy in proportion to the correction still to come: 1e-2 of it, down
        // to tol (inexact defect correction; the fixed point, and the stopping test on the
        // corrections, are those of exact solves).
        prism_solve_params pk = *p;
        if (corrected) pk.tol = std::max(tol, 1e-2);
        StepOut so = krylovStep(m->ctx, ws, A, g.p, lam.p, pk);
        std::memset(r, 0, sizeof(*r));
        r->iteration = 1;
        r->krylov_iters = so.iters;
        r->refinements = so.refinements;
        r->restarts = so.restarts;
        r->krylov_rel_residual = so.rel;
        r->converged = so.done == 1 || so.done == 4;
        if (corrected) {
            // (A - M)^T lambda = g by defect correction, A^T lambda_{k+1} = g + M^T lambda_k, the
            // transpose of the primal's non-orthogonal correction loop (it converges where that
            // does, at the same rate); each solve starts from the previous lambda, to 1e-2 of
            // the last correction's relative size (at least tol).  Stops when lambda moves by
            // less than 10 tol relative in a solve to tol, or after 200 corrections.
            ensureGrad(*e->phi);
            DFaceScheme ds{};
            ds.kind = 2;
            ds.variant = t.diff->variant;
            ds.kappa_kind = 0;
            ds.kappa = t.diff->kappa->view();
            DevBuf<double> q, rhs, prev, part;
            q.alloc(3 * (size_t)h.n);
            rhs.alloc(h.n);
            prev.alloc(h.n);
            const int nbk = diff_norm_blocks(h.n);
            part.alloc(2 * (size_t)nbk);
            std::vector<double> hp(2 * (size_t)nbk);
            const double stop = 10.0 * tol;
            bool settled = false;
            int outer = 0;
            double change = 1.0;  // relative size of the last correction
            for (; outer < 200 && r->converged; ++outer) {
                launch_adj_corr_rhs(L, m->dm, e->phi->view(), ds, lam.p, q.p, g.p, rhs.p);
                launch_copy(L, prev.p, lam.p, h.n);
                pk.tol = std::max(tol, std::min(1e-2, 1e-2 * change));
                StepOut sk = krylovStep(m->ctx, ws, A, rhs.p, lam.p, pk);
                r->krylov_iters += sk.iters;
                r->refinements += sk.refinements;
                r->restarts += sk.restarts;
                r->krylov_rel_residual = sk.rel;
                r->converged = sk.done == 1 || sk.done == 4;
                launch_diff_norms(L, h.n, prev.p, lam.p, part.p, part.n);
                HIPCHECK(hipMemcpyAsync(hp.data(), part.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost,
                                        m->ctx->stream));
                m->ctx->sync();
                double dd = 0.0, ll = 0.0;
                for (int i = 0; i < nbk; ++i) {
                    dd += hp[2 * i];
                    ll += hp[2 * i + 1];
                }
                if (!(dd == dd)) break;
                if (dd <= stop * stop * ll && pk.tol <= tol) {
                    settled = true;
                    ++outer;
                    break;
                }
                change = ll > 0.0 ? std::sqrt(dd / ll) : 0.0;
            }
            r->iteration = 1 + outer;
            r->final_residual = 0.0;
            if (!settled) r->converged = 0;
        }
        out.alloc(h.n);
        if (dJds && source) {
            launch_dJds(L, m->dm, lam.p, t.src->variant == PRISM_SIGN_NEGATIVE ? 1 : 0, out.p);
            HIPCHECK(hipMemcpyAsync(dJds, out.p, h.n * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
            m->ctx->sync();
        }
        if (dJdkappa && kappa) {
            scr.alloc(std::max<int64_t>(1, 2 * (h.F - h.nb)));
            ensureGrad(*e->phi);
            launch_adj_dkappa(L, m->dm, e->phi->view(), t.diff->variant, lam.p, scr.p, out.p);
            HIPCHECK(hipMemcpyAsync(dJdkappa, out.p, h.n * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
        }
        zeroCoeffs(*e);  // leave the coefficients as a clearing solve() does
        m->ctx->sync();
    });
}
