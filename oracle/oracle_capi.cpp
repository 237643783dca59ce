// oracle_capi.cpp -- C ABI of the FP64 CPU oracle, with the same signatures as the HIP
// library's include/prism_hip.h under the prefix `pref_` (test infrastructure only).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <numeric>
#include <string>

#include "../include/prism_hip.h"
#include "oracle_core.h"

using namespace oref;

struct prism_ctx { int dummy; };
struct prism_mesh { std::shared_ptr<Mesh> m; std::shared_ptr<CSR> pat; };
struct prism_field { std::shared_ptr<Scalar> s; prism_mesh* mesh; };
struct prism_vfield { std::shared_ptr<Vector> v; prism_field comp[3]; prism_mesh* mesh; };
struct prism_tfield { std::shared_ptr<Tensor> t; };
struct prism_eqn { Equation e; prism_field phi_handle; prism_mesh* mesh; };
struct prism_csr { CSR p; std::vector<f64> A; };

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& f) {
    try {
        f();
        g_err.clear();
        return PRISM_OK;
    } catch (const BCError& e) {
        g_err = e.what();
        return PRISM_ERR_BC;
    } catch (const MeshError& e) {
        g_err = e.what();
        return PRISM_ERR_MESH;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return PRISM_ERR_INVALID;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PRISM_ERR_RUNTIME;
    }
}

}  // namespace

#define OAPI extern "C" __attribute__((visibility("default")))

OAPI const char* pref_last_error(void) { return g_err.c_str(); }

OAPI int pref_ctx_create(int, prism_ctx** out) { *out = new prism_ctx{0}; return PRISM_OK; }
OAPI int pref_ctx_destroy(prism_ctx* c) { delete c; return PRISM_OK; }
OAPI int pref_ctx_synchronize(prism_ctx*) { return PRISM_OK; }

// ------------------------------------------------------------------------ mesh ----
OAPI int pref_mesh_create(prism_ctx*, const prism_mesh_desc* d, prism_mesh** out) {
    return guard([&] {
        auto m = std::make_shared<Mesh>();
        m->vertices.resize(d->n_vertices);
        for (int64_t i = 0; i < d->n_vertices; ++i)
            m->vertices[i] = {d->vertices[3 * i], d->vertices[3 * i + 1], d->vertices[3 * i + 2]};
        // patches + BCs (mesh/boundary.cpp:291-319, 391-407)
        m->patches.resize(d->n_patches);
        for (int32_t p = 0; p < d->n_patches; ++p) m->patches[p].name = d->patch_names[p];
        for (int32_t k = 0; k < d->n_bcs; ++k) {
            const auto& b = d->bcs[k];
            if (b.patch < 0 || b.patch >= d->n_patches) throw std::runtime_error("bc patch index out of range");
            BC bc;
            bc.kind = b.kind;
            bc.vk = b.value_kind;
            bc.s = b.value[0];
            bc.v = {b.value[0], b.value[1], b.value[2]};
            m->patches[b.patch].bcs[b.field] = bc;
        }
        for (auto& p : m->patches) {
            if (p.bcs.empty()) throw std::runtime_error("BoundaryPatch " + p.name + " has no boundary conditions");
            p.empty = p.bcs.begin()->second.kind == "empty";
        }
        for (int32_t k = 0; k < d->n_fields; ++k) {
            FieldInfo fi;
            fi.name = d->fields[k].name;
            fi.type = d->fields[k].type ? d->fields[k].type : "";
            if (d->fields[k].grad_scheme) fi.grad = std::string(d->fields[k].grad_scheme);
            m->infos.push_back(fi);
        }
        // faces in input order, geometry
        int64_t F = d->n_faces;
        std::vector<Face> faces(F);
        for (int64_t f = 0; f < F; ++f) {
            faces[f].verts.assign(d->face_vtx + d->face_vtx_off[f], d->face_vtx + d->face_vtx_off[f + 1]);
            faces[f].owner = d->face_owner[f];
            faces[f].neighbor = d->face_neighbor[f];
            faces[f].patch = d->face_patch[f];
            if (faces[f].neighbor < 0 && (faces[f].patch < 0 || faces[f].patch >= d->n_patches))
                throw MeshError("boundary face without a patch");
            faceGeometry(m->vertices, faces[f]);
        }
        std::vector<Cell> cells(d->n_cells);
        for (int64_t c = 0; c < d->n_cells; ++c) {
            cells[c].faces.assign(d->cell_faces + d->cell_face_off[c], d->cell_faces + d->cell_face_off[c + 1]);
            cellGeometry(faces, cells[c]);
            if (d->cell_vtx_off) {
                cells[c].verts.assign(d->cell_vtx + d->cell_vtx_off[c], d->cell_vtx + d->cell_vtx_off[c + 1]);
                for (auto v : cells[c].verts)
                    if (v < 0 || v >= d->n_vertices) throw MeshError("cell vertex id out of range");
            } else {  // the distinct vertices of the cell's faces, in order of appearance
                for (auto fid : cells[c].faces)
                    for (auto v : faces[fid].verts)
                        if (std::find(cells[c].verts.begin(), cells[c].verts.end(), v) == cells[c].verts.end())
                            cells[c].verts.push_back(v);
            }
        }
        m->has_cell_vtx = d->cell_vtx_off != nullptr;
        // permutation [non-empty boundary | empty boundary | interior] (pmesh.cpp:51-85)
        std::vector<int64_t> o2n(F);
        int64_t nne = 0, nb = 0;
        for (int64_t f = 0; f < F; ++f)
            if (faces[f].neighbor < 0) { ++nb; if (!m->patches[faces[f].patch].empty) ++nne; }
        int64_t pne = 0, pe = nne, pi = nb;
        for (int64_t f = 0; f < F; ++f) {
            if (faces[f].neighbor >= 0) continue;
            o2n[f] = m->patches[faces[f].patch].empty ? pe++ : pne++;
        }
        for (int64_t f = 0; f < F; ++f) if (faces[f].neighbor >= 0) o2n[f] = pi++;
        m->faces.resize(F);
        for (int64_t f = 0; f < F; ++f) m->faces[o2n[f]] = std::move(faces[f]);
        for (auto& c : cells) for (auto& fid : c.faces) fid = o2n[fid];
        for (int64_t f = 0; f < F; ++f)
            if (!m->faces[f].interior()) m->patches[m->faces[f].patch].faces.push_back(f);
        m->cells = std::move(cells);
        m->n_boundary = nb;
        m->n_nonempty = nne;
        m->old_to_new = o2n;
        m->vol.resize(m->nc());
        for (int64_t c = 0; c < m->nc(); ++c) m->vol[c] = m->cells[c].volume;
        auto* h = new prism_mesh;
        h->m = m;
        h->pat = patternFor(m);
        *out = h;
    });
}
OAPI int pref_mesh_destroy(prism_mesh* m) { delete m; return PRISM_OK; }

OAPI int pref_mesh_get_info(prism_mesh* h, prism_mesh_info* info) {
    info->n_cells = h->m->nc();
    info->n_faces = h->m->nf();
    info->n_boundary_faces = h->m->n_boundary;
    info->n_nonempty_boundary_faces = h->m->n_nonempty;
    info->nnz = (int64_t)h->pat->col.size();
    info->n_gg_levels = 0;
    return PRISM_OK;
}

OAPI int pref_mesh_geometry(prism_mesh* h, double* cc, double* cv, double* fc, double* fs,
                            int64_t* own, int64_t* nbr, int32_t* fp) {
    const Mesh& m = *h->m;
    for (int64_t c = 0; c < m.nc(); ++c) {
        if (cc) { cc[3 * c] = m.cells[c].center.x; cc[3 * c + 1] = m.cells[c].center.y; cc[3 * c + 2] = m.cells[c].center.z; }
        if (cv) cv[c] = m.cells[c].volume;
    }
    for (int64_t f = 0; f < m.nf(); ++f) {
        const Face& F = m.faces[f];
        V3 S = F.S();
        if (fc) { fc[3 * f] = F.center.x; fc[3 * f + 1] = F.center.y; fc[3 * f + 2] = F.center.z; }
        if (fs) { fs[3 * f] = S.x; fs[3 * f + 1] = S.y; fs[3 * f + 2] = S.z; }
        if (own) own[f] = F.owner;
        if (nbr) nbr[f] = F.neighbor;
        if (fp) fp[f] = F.patch;
    }
    return PRISM_OK;
}

OAPI int pref_mesh_face_permutation(prism_mesh* h, int64_t* o2n) {
    std::copy(h->m->old_to_new.begin(), h->m->old_to_new.end(), o2n);
    return PRISM_OK;
}

// ---------------------------------------------------------------------- fields ----
OAPI int pref_scalar_create(prism_mesh* mh, const char* name, int32_t kind, const double* values,
                            double value, prism_field** out) {
    return guard([&] {
        std::vector<f64> v(mh->m->nc(), value);
        if (values) v.assign(values, values + mh->m->nc());
        auto* f = new prism_field;
        f->s = std::make_shared<Scalar>(name, mh->m, std::move(v), kind);
        f->mesh = mh;
        *out = f;
    });
}
OAPI int pref_field_destroy(prism_field* f) { delete f; return PRISM_OK; }
OAPI int pref_scalar_set(prism_field* f, const double* v) {
    f->s->bump();
    std::copy(v, v + f->s->v.size(), f->s->v.begin());
    f->s->bump();
    return PRISM_OK;
}
OAPI int pref_scalar_get(prism_field* f, double* v) { std::copy(f->s->v.begin(), f->s->v.end(), v); return PRISM_OK; }
OAPI int pref_scalar_fill(prism_field* f, double value) {
    f->s->bump();
    std::fill(f->s->v.begin(), f->s->v.end(), value);
    f->s->bump();
    return PRISM_OK;
}
OAPI int pref_scalar_set_face_values(prism_field* f, const double* fv) {
    f->s->fv.assign(fv, fv + f->s->mesh->nf());
    return PRISM_OK;
}
OAPI int pref_scalar_clear_face_values(prism_field* f) { f->s->fv.clear(); return PRISM_OK; }
OAPI int pref_scalar_grad(prism_field* f, double* g) {
    return guard([&] {
        for (int64_t c = 0; c < f->s->mesh->nc(); ++c) {
            V3 v = f->s->gradCell(c);
            g[3 * c] = v.x; g[3 * c + 1] = v.y; g[3 * c + 2] = v.z;
        }
    });
}
OAPI int pref_scalar_face_values(prism_field* f, double* out) {
    return guard([&] {
        const Mesh& m = *f->s->mesh;
        for (int64_t fid = 0; fid < m.nf(); ++fid) {
            const Face& F = m.faces[fid];
            if (!F.interior() && m.patchOf(F).empty && !f->s->hasFaceValues()) { out[fid] = 0.0; continue; }
            out[fid] = f->s->faceValue(fid);
        }
    });
}
OAPI int pref_scalar_set_history(prism_field* f, int32_t n) {
    f->s->hist_max = n;
    if ((size_t)n < f->s->hist.size()) f->s->hist.resize(n);
    return PRISM_OK;
}
OAPI int pref_scalar_update_prev(prism_field* f) { f->s->updatePrev(); return PRISM_OK; }

static prism_vfield* wrapVector(std::shared_ptr<Vector> v, prism_mesh* mh) {
    auto* h = new prism_vfield;
    h->v = v;
    h->mesh = mh;
    for (int k = 0; k < 3; ++k) { h->comp[k].s = v->c[k]; h->comp[k].mesh = mh; }
    return h;
}

OAPI int pref_vector_create(prism_mesh* mh, const char* name, const double* values, const double* u3,
                            prism_vfield** out) {
    return guard([&] {
        auto v = std::make_shared<Vector>();
        v->name = name;
        v->mesh = mh->m;
        int64_t n = mh->m->nc();
        const char* sfx[3] = {"_x", "_y", "_z"};
        for (int k = 0; k < 3; ++k) {
            std::vector<f64> c(n, u3 ? u3[k] : 0.0);
            if (values) for (int64_t i = 0; i < n; ++i) c[i] = values[3 * i + k];
            v->c[k] = std::make_shared<Scalar>(std::string(name) + sfx[k], mh->m, std::move(c), 0, k);
        }
        *out = wrapVector(v, mh);
    });
}
OAPI int pref_vector_clone(prism_vfield* src, prism_vfield** out) {
    return guard([&] {
        auto v = std::make_shared<Vector>(*src->v);
        for (int k = 0; k < 3; ++k) {
            auto c = std::make_shared<Scalar>(src->v->c[k]->name, src->mesh->m, src->v->c[k]->v, 0, k);
            c->fv = src->v->c[k]->fv;
            c->event = src->v->c[k]->event;
            v->c[k] = c;
        }
        *out = wrapVector(v, src->mesh);
    });
}
OAPI int pref_vfield_destroy(prism_vfield* v) { delete v; return PRISM_OK; }
OAPI int pref_vector_component(prism_vfield* v, int32_t k, prism_field** out) { *out = &v->comp[k]; return PRISM_OK; }
OAPI int pref_vector_set_face_values(prism_vfield* v, const double* fv) {
    int64_t F = v->mesh->m->nf();
    v->v->fd.resize(F);
    for (int64_t f = 0; f < F; ++f) v->v->fd[f] = {fv[3 * f], fv[3 * f + 1], fv[3 * f + 2]};
    v->v->has_fd = true;
    return PRISM_OK;
}
OAPI int pref_vector_set_face_fluxes(prism_vfield* v, const double* ff) {
    v->v->ff.assign(ff, ff + v->mesh->m->nf());
    v->v->has_ff = true;
    return PRISM_OK;
}
OAPI int pref_vector_face_fluxes(prism_vfield* v, double* out) {
    return guard([&] {
        const Mesh& m = *v->mesh->m;
        for (int64_t f = 0; f < m.nf(); ++f) {
            const Face& F = m.faces[f];
            if (!F.interior() && m.patchOf(F).empty && !v->v->has_ff) { out[f] = 0.0; continue; }
            out[f] = v->v->flux(f);
        }
    });
}
OAPI int pref_vector_face_values(prism_vfield* v, double* out) {
    return guard([&] {
        const Mesh& m = *v->mesh->m;
        for (int64_t f = 0; f < m.nf(); ++f) {
            const Face& F = m.faces[f];
            V3 u{};
            if (F.interior() || !m.patchOf(F).empty || v->v->has_fd) u = v->v->faceValue(f);
            out[3 * f] = u.x; out[3 * f + 1] = u.y; out[3 * f + 2] = u.z;
        }
    });
}

OAPI int pref_tensor_create(prism_mesh* mh, const char* name, const double* values, const double* u9,
                            prism_tfield** out) {
    return guard([&] {
        auto t = std::make_shared<Tensor>();
        t->name = name;
        t->mesh = mh->m;
        int64_t n = mh->m->nc();
        t->v.resize(n);
        for (int64_t i = 0; i < n; ++i)
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) t->v[i].m[r][c] = values ? values[9 * i + 3 * r + c] : u9[3 * r + c];
        *out = new prism_tfield{t};
    });
}
OAPI int pref_tfield_destroy(prism_tfield* t) { delete t; return PRISM_OK; }
OAPI int pref_tensor_get(prism_tfield* t, double* out) {
    for (size_t i = 0; i < t->t->v.size(); ++i)
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) out[9 * i + 3 * r + c] = t->t->v[i].m[r][c];
    return PRISM_OK;
}

// ------------------------------------------------------------------- equations ----
OAPI int pref_eqn_create(const prism_scheme_desc* sd, int32_t n, prism_eqn** out) {
    return guard([&] {
        if (n < 1) throw std::runtime_error("Transport needs at least one scheme");
        auto* h = new prism_eqn;
        prism_mesh* mh = nullptr;
        for (int32_t k = 0; k < n; ++k) {
            Scheme s;
            const auto& d = sd[k];
            switch (d.kind) {
                case PRISM_SCHEME_CONVECTION: s.kind = SK::Convection; break;
                case PRISM_SCHEME_DIFFUSION: s.kind = SK::Diffusion; break;
                case PRISM_SCHEME_SRC_CONSTANT: s.kind = SK::SrcConstant; break;
                case PRISM_SCHEME_SRC_GRADIENT: s.kind = SK::SrcGradient; break;
                case PRISM_SCHEME_SRC_DIVERGENCE: s.kind = SK::SrcDivergence; break;
                case PRISM_SCHEME_SRC_IMPLICIT: s.kind = SK::SrcImplicit; break;
                case PRISM_SCHEME_BACKWARD_EULER: s.kind = SK::BackwardEuler; break;
                case PRISM_SCHEME_ADAM_MOULTON: s.kind = SK::AdamMoulton; break;
                default: throw std::runtime_error("unknown scheme kind");
            }
            s.variant = d.variant;
            if (d.phi) { s.phi = d.phi->s; mh = d.phi->mesh; }
            if (d.kappa) s.kappa = d.kappa->s;
            if (d.kappa_t) s.kappa_t = d.kappa_t->t;
            if (d.vec) { s.vec = d.vec->v; if (!mh) mh = d.vec->mesh; }
            s.coord = d.coord;
            s.coeff = d.coeff;
            if (s.kind == SK::BackwardEuler && !(d.coeff > 0.0))
                throw std::invalid_argument("BackwardEuler(phi, dt): dt must be positive");
            if (s.kind == SK::AdamMoulton && !(d.coeff > 0.0))
                throw std::invalid_argument("AdamMoulton(phi, dt): dt must be positive");
            if (k == 0) {
                if (!d.phi) throw std::runtime_error("first scheme has no field");
                h->e.phi = d.phi->s;
                h->phi_handle = *d.phi;
            }
            if (s.kind == SK::Diffusion && !s.kappa && !s.kappa_t) throw std::runtime_error("diffusion without kappa");
            if (s.kind == SK::Convection && !s.vec) throw std::runtime_error("convection without U");
            if (s.full()) h->e.schemes.push_back(s);
            else h->e.sources.push_back(s);
        }
        h->mesh = mh;
        h->e.pattern = mh->pat;
        size_t nnz = mh->pat->col.size(), nc = mh->m->nc();
        for (auto& s : h->e.schemes) s.sys.resize(nnz, nc);
        for (auto& s : h->e.sources) s.sys.resize(0, nc);
        h->e.sys.resize(nnz, nc);
        *out = h;
    });
}
OAPI int pref_eqn_destroy(prism_eqn* e) { delete e; return PRISM_OK; }
OAPI int pref_eqn_update_coeffs(prism_eqn* e) { return guard([&] { e->e.updateCoeffs(); }); }
OAPI int pref_eqn_zero_coeffs(prism_eqn* e) { e->e.zeroOut(); return PRISM_OK; }
OAPI int pref_eqn_set_under_relax(prism_eqn* e, double f) { e->e.setUnderRelax(f); return PRISM_OK; }
OAPI int pref_eqn_relax(prism_eqn* e) { e->e.relax(); return PRISM_OK; }
OAPI int pref_eqn_add_symmetry_handler(prism_eqn* e) {
    return guard([&] {
        if (e->e.phi->coord < 0)
            throw std::runtime_error("eqn::boundary::Symmetry<Transport>: field `" + e->e.phi->name + "` is not a vector component");
        e->e.symmetry = true;
    });
}
OAPI int pref_adjoint_objective(prism_eqn* e, const prism_objective* obj, double* J) {
    return guard([&] {
        if (obj->axis < 0 || obj->axis > 2) throw std::invalid_argument("objective axis must be 0, 1 or 2");
        f64 VR;
        *J = adjointObjective(*e->e.phi, obj->axis, obj->threshold, &VR);
    });
}
OAPI int pref_adjoint_gradient(prism_eqn* e, prism_field* source, prism_field* kappa, const prism_objective* obj,
                               const prism_solve_params*, double* J, double* dJds, double* dJdk,
                               prism_solve_result* r) {
    return guard([&] {
        if (obj->axis < 0 || obj->axis > 2) throw std::invalid_argument("objective axis must be 0, 1 or 2");
        std::vector<f64> ds, dk;
        adjointGradient(e->e, source ? source->s.get() : nullptr, kappa ? kappa->s.get() : nullptr, obj->axis,
                        obj->threshold, J, dJds ? &ds : nullptr, dJdk ? &dk : nullptr);
        if (dJds && source) std::copy(ds.begin(), ds.end(), dJds);
        if (dJdk && kappa) std::copy(dk.begin(), dk.end(), dJdk);
        if (r) { std::memset(r, 0, sizeof(*r)); r->iteration = 1; r->converged = 1; }
    });
}
OAPI int pref_needs_pressure_reference(prism_field* p, int32_t* out) {
    return guard([&] { *out = needsRef(*p->s) ? 1 : 0; });
}
OAPI prism_field* pref_eqn_field(prism_eqn* e) { return &e->phi_handle; }
OAPI int pref_eqn_nnz(prism_eqn* e, int64_t* nnz, int64_t* n) {
    *nnz = (int64_t)e->e.pattern->col.size();
    *n = e->e.pattern->n;
    return PRISM_OK;
}
OAPI int pref_eqn_matrix(prism_eqn* e, int64_t* rp, int32_t* col, double* val) {
    const CSR& p = *e->e.pattern;
    if (rp) std::copy(p.rp.begin(), p.rp.end(), rp);
    if (col) std::copy(p.col.begin(), p.col.end(), col);
    if (val) std::copy(e->e.sys.A.begin(), e->e.sys.A.end(), val);
    return PRISM_OK;
}
OAPI int pref_eqn_rhs(prism_eqn* e, double* b) { std::copy(e->e.sys.b.begin(), e->e.sys.b.end(), b); return PRISM_OK; }
OAPI int pref_eqn_diag(prism_eqn* e, double* d) {
    const CSR& p = *e->e.pattern;
    for (int64_t i = 0; i < p.n; ++i) d[i] = e->e.sys.A[p.diag[i]];
    return PRISM_OK;
}

OAPI void pref_solve_params_default(prism_solve_params* p) {
    std::memset(p, 0, sizeof(*p));
    p->n_iter = 10; p->eps = 1e-7; p->ref_cell = -1; p->ref_value = 0; p->clear_coeffs = 1;
    p->tol = 1e-14; p->max_iters = 0;
}

// BiCGSTAB::step on the assembled system (bicgstab.cpp:7-9): the exact solution from x.
OAPI int pref_eqn_step(prism_eqn* e, const prism_solve_params*, prism_solve_result* r) {
    return guard([&] {
        Scalar& phi = *e->e.phi;
        phi.bump();
        std::vector<f64> x = phi.v;
        exactSolve(*e->e.pattern, e->e.sys.A, e->e.sys.b, x);
        phi.v = x;
        phi.bump();
        if (r) { std::memset(r, 0, sizeof(*r)); r->iteration = 1; r->converged = 1; }
    });
}
OAPI int pref_solve(prism_eqn* e, const prism_solve_params* p, prism_solve_result* r) {
    return guard([&] {
        SolveParams sp;
        sp.n_iter = p->n_iter; sp.eps = p->eps; sp.ref_cell = p->ref_cell; sp.ref_value = p->ref_value;
        sp.clear = p->clear_coeffs != 0;
        auto res = solve(e->e, sp);
        std::memset(r, 0, sizeof(*r));
        r->iteration = res.iter;
        r->initial_residual = res.r0;
        r->final_residual = res.r;
        r->converged = res.converged;
    });
}

// ------------------------------------------------------------------ bare matrix ----
OAPI int pref_csr_create(prism_ctx*, int64_t n, const int64_t* rp, const int32_t* col, const double* val,
                         prism_csr** out) {
    return guard([&] {
        auto* a = new prism_csr;
        a->p.n = n;
        a->p.rp.assign(rp, rp + n + 1);
        a->p.col.assign(col, col + rp[n]);
        a->A.assign(val, val + rp[n]);
        a->p.diag.assign(n, -1);
        for (int64_t i = 0; i < n; ++i)
            for (int64_t k = rp[i]; k < rp[i + 1]; ++k) if (col[k] == i) a->p.diag[i] = k;
        *out = a;
    });
}
OAPI int pref_csr_destroy(prism_csr* a) { delete a; return PRISM_OK; }
OAPI int pref_csr_spmv(prism_csr* a, const double* x, double* y) { spmv(a->p, a->A, x, y); return PRISM_OK; }
OAPI int pref_csr_jacobi_step(prism_csr* a, const double* b, const double* xin, double* xout) {
    // jacobi.cpp:5-9: x + D^-1 (b - A x)
    int64_t n = a->p.n;
    std::vector<f64> ax(n), x(xin, xin + n);
    spmv(a->p, a->A, x.data(), ax.data());
    for (int64_t i = 0; i < n; ++i) {
        f64 dinv = 1.0 / a->A[a->p.diag[i]];
        xout[i] = x[i] + dinv * (b[i] - ax[i]);
    }
    return PRISM_OK;
}
OAPI int pref_csr_bicgstab(prism_csr* a, const double* b, const double* xin, double* xout,
                           const prism_solve_params* p, prism_solve_result* r) {
    return guard([&] {
        int64_t n = a->p.n;
        std::vector<f64> x(xin, xin + n);
        f64 rel = 0;
        int64_t maxit = p->max_iters > 0 ? p->max_iters : 2 * n;
        int64_t it = plainBicgstab(n, a->p.rp.data(), a->p.col.data(), a->A.data(), b, x.data(), p->tol, maxit, &rel);
        std::copy(x.begin(), x.end(), xout);
        if (r) { std::memset(r, 0, sizeof(*r)); r->krylov_iters = it; r->krylov_rel_residual = rel; r->iteration = 1; }
    });
}
// exact solve of the bare matrix (used to pin known answers)
OAPI int pref_csr_exact(prism_csr* a, const double* b, double* x) {
    return guard([&] {
        std::vector<f64> bv(b, b + a->p.n), xv(a->p.n, 0.0);
        exactSolve(a->p, a->A, bv, xv);
        std::copy(xv.begin(), xv.end(), x);
    });
}

// ------------------------------------------------------------------ SIMPLE/PISO ----
OAPI void pref_simple_controls_default(prism_simple_controls* c) {
    std::memset(c, 0, sizeof(*c));
    c->momentum_urf = 0.7; c->momentum_max_iter = 3; c->momentum_residual = 1e-7;
    c->non_ortho_correctors = 2; c->pressure_urf = 0.3; c->pressure_max_iter = 5;
    c->pressure_residual = 1e-7; c->p_ref_cell = -1; c->p_ref_value = 0;
    c->momentum_tol = 1e-14; c->pressure_tol = 1e-14;
}
OAPI void pref_piso_controls_default(prism_piso_controls* c) {
    std::memset(c, 0, sizeof(*c));
    c->outer_iterations = 1; c->momentum_implicit_steps = 1; c->pressure_correction_steps = 2;
    pref_simple_controls_default(&c->simple);
    c->simple.momentum_urf = 1.0; c->simple.momentum_residual = 1e-4;
    c->simple.pressure_urf = 1.0; c->simple.pressure_residual = 1e-4;
}

static SimpleControls toCtl(const prism_simple_controls* c) {
    SimpleControls s;
    s.momentum_urf = c->momentum_urf; s.momentum_max_iter = c->momentum_max_iter;
    s.momentum_residual = c->momentum_residual; s.non_ortho = c->non_ortho_correctors;
    s.pressure_urf = c->pressure_urf; s.pressure_max_iter = c->pressure_max_iter;
    s.pressure_residual = c->pressure_residual; s.p_ref_cell = c->p_ref_cell; s.p_ref_value = c->p_ref_value;
    return s;
}
static void toReports(const std::vector<Report>& rep, prism_report* out, int32_t maxr, int32_t* nr) {
    int32_t k = 0;
    for (const auto& r : rep) {
        if (k >= maxr) break;
        std::memset(&out[k], 0, sizeof(prism_report));
        std::strncpy(out[k].field_name, r.name.c_str(), 63);
        out[k].n_iterations = r.iters; out[k].initial_residual = r.r0; out[k].final_residual = r.r;
        out[k].converged = r.converged;
        ++k;
    }
    *nr = k;
}

OAPI int pref_simple_step(prism_eqn* const* mom, int32_t nm, prism_vfield* U, prism_vfield* mdot, prism_field* p,
                          const prism_simple_controls* c, prism_report* rep, int32_t maxr, int32_t* nr) {
    return guard([&] {
        std::vector<Equation*> m;
        for (int32_t k = 0; k < nm; ++k) m.push_back(&mom[k]->e);
        auto r = simpleStep(m, U->v, mdot->v, p->s, toCtl(c));
        toReports(r, rep, maxr, nr);
    });
}
OAPI int pref_simple_pressure_system(prism_eqn* const* mom, int32_t nm, prism_vfield* U, prism_vfield* mdot,
                                     prism_field* p, const prism_simple_controls* c, prism_report* rep, int32_t maxr,
                                     int32_t* nr, prism_eqn** out, int64_t* ref_cell) {
    return guard([&] {
        if (!mom || !U || !mdot || !p || !c || !out) throw std::invalid_argument("simple_pressure_system: null argument");
        if (nm != 2 && nm != 3) throw std::runtime_error("SIMPLE expects 2 or 3 momentum predictors");
        std::vector<Equation*> m;
        for (int32_t k = 0; k < nm; ++k) m.push_back(&mom[k]->e);
        SimpleControls ctl = toCtl(c);
        auto r = momentumPredictor(m, ctl);
        PressureSystem ps = buildPressureSystem(ctl, m, U->v, mdot->v, p->s);
        auto* h = new prism_eqn;
        h->e = std::move(ps.pe);
        h->phi_handle.s = ps.pp;
        h->phi_handle.mesh = p->mesh;
        h->mesh = p->mesh;
        *out = h;
        if (ref_cell) *ref_cell = ps.sp.ref_cell;
        if (rep && nr) toReports(r, rep, maxr, nr);
    });
}
OAPI int pref_prime_step(prism_eqn* const* mom, int32_t nm, prism_vfield* U, prism_vfield* mdot, prism_field* p,
                         const prism_simple_controls* c, prism_report* rep, int32_t maxr, int32_t* nr) {
    return guard([&] {
        std::vector<Equation*> m;
        for (int32_t k = 0; k < nm; ++k) m.push_back(&mom[k]->e);
        auto r = primeStep(m, U->v, mdot->v, p->s, toCtl(c));
        toReports(r, rep, maxr, nr);
    });
}
OAPI int pref_piso_step(prism_eqn* const* mom, int32_t nm, prism_vfield* U, prism_vfield* mdot, prism_field* p,
                        const prism_piso_controls* c, prism_report* rep, int32_t maxr, int32_t* nr) {
    return guard([&] {
        std::vector<Equation*> m;
        for (int32_t k = 0; k < nm; ++k) m.push_back(&mom[k]->e);
        auto r = pisoStep(m, U->v, mdot->v, p->s, c->outer_iterations, c->momentum_implicit_steps,
                          c->pressure_correction_steps, toCtl(&c->simple));
        toReports(r, rep, maxr, nr);
    });
}
// ------------------------------------------------------------------ turbulence ----
// IncompressibleKEpsilon::setTurbulentViscosity, buildKEqn and buildEpsilonEqn field
// arithmetic (KEpsilon.cpp:27-77) and kProductionIncompressible (k_production.cpp:8-25), with
// Eigen's association: gradU column j = grad U_j (vector.h:395-401); twoSymm A + A^T;
// tau -= (2/3 k) I over all nine entries; A.cwiseProduct(B).sum() over the column-major
// elements as its 2-wide vectorised tree ((e0+e2)+(e4+e6)) + ((e1+e3)+(e5+e7)), then + e8.
OAPI int pref_kepsilon_update(prism_field* k, prism_field* eps, prism_field* nu, prism_vfield* U, const double* consts,
                              prism_field* const* out) {
    return guard([&] {
        if (!k || !eps || !nu || !U || !out) throw std::invalid_argument("IncompressibleKEpsilon: null argument");
        const Mesh* m = k->s->mesh.get();
        if (eps->s->mesh.get() != m || nu->s->mesh.get() != m || U->v->mesh.get() != m)
            throw MeshError("IncompressibleKEpsilon: k, epsilon, nu and U must share one mesh");
        for (int i = 0; i < 6; ++i) {
            if (!out[i] || out[i]->s->mesh.get() != m) throw MeshError("IncompressibleKEpsilon: output field on another mesh");
            for (auto* in : {k, eps, nu, &U->comp[0], &U->comp[1], &U->comp[2]})
                if (out[i]->s == in->s) throw std::invalid_argument("IncompressibleKEpsilon: an output field is also an input");
            for (int j = 0; j < i; ++j)
                if (out[j]->s == out[i]->s) throw std::invalid_argument("IncompressibleKEpsilon: an output field is given twice");
        }
        const f64 dflt[5] = {0.09, 1.44, 1.92, 1.0, 1.3};
        const f64* c = consts ? consts : dflt;
        const f64 c_mu = c[0], c_eps1 = c[1], c_eps2 = c[2], sigma_k = c[3], sigma_eps = c[4];
        const int64_t n = m->nc();
        std::vector<f64> nu_t(n), pk(n), nek(n), nee(n), src(n), snk(n);
        const auto &kv = k->s->v, &ev = eps->s->v, &nv = nu->s->v;
        for (int64_t i = 0; i < n; ++i) nu_t[i] = c_mu * ((kv[i] * kv[i]) / ev[i]);
        for (int64_t i = 0; i < n; ++i) {
            f64 G[3][3];  // G[r][col]
            for (int col = 0; col < 3; ++col) {
                V3 g = U->v->c[col]->gradCell(i);
                G[0][col] = g.x; G[1][col] = g.y; G[2][col] = g.z;
            }
            f64 tau[3][3];
            for (int r = 0; r < 3; ++r)
                for (int col = 0; col < 3; ++col) tau[r][col] = nu_t[i] * (G[r][col] + G[col][r]);
            const f64 s = (2.0 / 3.0) * kv[i];
            for (int r = 0; r < 3; ++r)
                for (int col = 0; col < 3; ++col) tau[r][col] -= s * (r == col ? 1.0 : 0.0);
            f64 e[9];
            for (int col = 0; col < 3; ++col)
                for (int r = 0; r < 3; ++r) e[r + 3 * col] = tau[r][col] * G[r][col];
            pk[i] = (((e[0] + e[2]) + (e[4] + e[6])) + ((e[1] + e[3]) + (e[5] + e[7]))) + e[8];
        }
        for (int64_t i = 0; i < n; ++i) {
            nek[i] = nv[i] + (nu_t[i] / sigma_k);
            nee[i] = nv[i] + (nu_t[i] / sigma_eps);
            const f64 eok = ev[i] / kv[i];
            src[i] = c_eps1 * (eok * pk[i]);
            snk[i] = c_eps2 * (ev[i] * eok);
        }
        std::vector<f64>* vals[6] = {&nu_t, &pk, &nek, &nee, &src, &snk};
        for (int j = 0; j < 6; ++j) {
            out[j]->s->v = *vals[j];
            out[j]->s->bump();
        }
    });
}

// ---------------------------------------------------------------------- output ----
namespace {
std::vector<f64> flatScalar(const Scalar& f) { return f.v; }  // flattenScalar (export.cpp:131-139)
std::vector<f64> flatVector(const Vector& v) {                // flattenVector (export.cpp:146-157)
    std::vector<f64> d;
    for (int64_t i = 0; i < v.mesh->nc(); ++i) {
        V3 x = v.cellValue(i);
        d.push_back(x.x); d.push_back(x.y); d.push_back(x.z);
    }
    return d;
}
std::vector<f64> flatTensor(const Tensor& t) {                // flattenTensor (export.cpp:164-182)
    std::vector<f64> d;
    for (const auto& x : t.v)
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) d.push_back(x.m[r][c]);
    return d;
}
std::string xmlEscape(const std::string& v) {
    std::string o;
    for (char c : v)
        o += c == '&' ? "&amp;" : c == '<' ? "&lt;" : c == '>' ? "&gt;" : c == '"' ? "&quot;" : std::string(1, c);
    return o;
}
void putArray(std::ostream& o, const std::string& attrs, const std::vector<f64>& a) {
    o << "<DataArray type=\"Float64\"" << attrs << " format=\"ascii\">\n";
    for (f64 x : a) o << x << ' ';
    o << "\n</DataArray>\n";
}
}  // namespace

OAPI int pref_mesh_n_vertices(prism_mesh* h, int64_t* nv) {
    *nv = (int64_t)h->m->vertices.size();
    return PRISM_OK;
}
OAPI int pref_field_to_points(prism_field* f, double* out) {
    return guard([&] {
        auto p = cellToPoint(*f->s->mesh, flatScalar(*f->s), 1);
        std::copy(p.begin(), p.end(), out);
    });
}
OAPI int pref_vfield_to_points(prism_vfield* v, double* out) {
    return guard([&] {
        auto p = cellToPoint(*v->v->mesh, flatVector(*v->v), 3);
        std::copy(p.begin(), p.end(), out);
    });
}
OAPI int pref_tfield_to_points(prism_tfield* t, double* out) {
    return guard([&] {
        auto p = cellToPoint(*t->t->mesh, flatTensor(*t->t), 9);
        std::copy(p.begin(), p.end(), out);
    });
}
// toVtu (export.cpp:262-300) and writeVtu (export.cpp:191-232), ASCII only: each item's
// CellData and PointData; point sets first, then cell sets, then Points and Cells.
OAPI int pref_export_vtu(const prism_export_item* items, int32_t n_items, const char* file_name, int32_t mode) {
    return guard([&] {
        if (n_items <= 0) return;
        if (!items || !file_name) throw std::invalid_argument("toVtu: null argument");
        if (mode != PRISM_WRITE_ASCII) throw std::invalid_argument("the oracle writes ASCII VTU files only");
        struct Set { std::string name; int nc; std::vector<f64> cell, point; };
        std::vector<Set> sets;
        const Mesh* mesh = nullptr;
        for (int32_t i = 0; i < n_items; ++i) {
            const Mesh* m;
            Set s;
            if (items[i].kind == PRISM_EXPORT_SCALAR) {
                auto* f = static_cast<const prism_field*>(items[i].field);
                m = f->s->mesh.get(); s = {f->s->name, 1, flatScalar(*f->s), {}};
            } else if (items[i].kind == PRISM_EXPORT_VECTOR) {
                auto* v = static_cast<const prism_vfield*>(items[i].field);
                m = v->v->mesh.get(); s = {v->v->name, 3, flatVector(*v->v), {}};
            } else if (items[i].kind == PRISM_EXPORT_TENSOR) {
                auto* t = static_cast<const prism_tfield*>(items[i].field);
                m = t->t->mesh.get(); s = {t->t->name, 9, flatTensor(*t->t), {}};
            } else {
                throw std::invalid_argument("export item kind must be 1, 3 or 9");
            }
            if (!mesh) mesh = m;
            if (m != mesh)
                throw MeshError("Field '" + s.name + "' does not share the mesh of the other exported fields.");
            s.point = cellToPoint(*m, s.cell, s.nc);
            sets.push_back(std::move(s));
        }
        if (!mesh->has_cell_vtx) throw MeshError("VTU export needs the cells' vertex lists");
        std::ofstream o(file_name);
        if (!o) throw std::runtime_error(std::string("cannot open ") + file_name);
        o << std::setprecision(17);
        o << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
             "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n<UnstructuredGrid>\n";
        o << "<Piece NumberOfPoints=\"" << mesh->vertices.size() << "\" NumberOfCells=\"" << mesh->nc() << "\">\n";
        auto attrs = [](const Set& s) {
            return " Name=\"" + xmlEscape(s.name) + "\" NumberOfComponents=\"" + std::to_string(s.nc) + "\"";
        };
        o << "<PointData>\n";
        for (auto& s : sets) putArray(o, attrs(s), s.point);
        o << "</PointData>\n<CellData>\n";
        for (auto& s : sets) putArray(o, attrs(s), s.cell);
        o << "</CellData>\n<Points>\n";
        std::vector<f64> pts;
        for (const auto& v : mesh->vertices) { pts.push_back(v.x); pts.push_back(v.y); pts.push_back(v.z); }
        putArray(o, " NumberOfComponents=\"3\"", pts);
        o << "</Points>\n<Cells>\n";
        o << "<DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n";
        for (const auto& c : mesh->cells) for (auto v : c.verts) o << v << ' ';
        o << "\n</DataArray>\n<DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n";
        int64_t at = 0;
        for (const auto& c : mesh->cells) { at += (int64_t)c.verts.size(); o << at << ' '; }
        o << "\n</DataArray>\n<DataArray type=\"Int8\" Name=\"types\" format=\"ascii\">\n";
        for (const auto& c : mesh->cells) {
            size_t k = c.verts.size();
            if (k != 8 && k != 4 && k != 6) throw std::runtime_error("Unsupported cell type for vtu export.");
            o << (k == 8 ? 12 : k == 4 ? 10 : 13) << ' ';
        }
        o << "\n</DataArray>\n</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
        if (!o) throw std::runtime_error(std::string("write to ") + file_name + " failed");
    });
}
// TimeWriter::flushPvd (export.cpp:393-409)
OAPI int pref_write_pvd(const char* file_name, int32_t n, const double* times, const char* const* files) {
    return guard([&] {
        if (!file_name || n < 0 || (n > 0 && (!times || !files))) throw std::invalid_argument("flushPvd: null argument");
        std::ofstream f(file_name);
        if (!f) throw std::runtime_error(std::string("cannot open ") + file_name);
        f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\">\n  <Collection>\n";
        for (int32_t i = 0; i < n; ++i) {
            char t[64];
            std::snprintf(t, sizeof(t), "%.6g", times[i]);
            f << "    <DataSet timestep=\"" << t << "\" file=\"" << xmlEscape(files[i]) << "\"/>\n";
        }
        f << "  </Collection>\n</VTKFile>\n";
    });
}

OAPI int pref_courant(prism_vfield* mdot, double dt, double* co) {  // courant.cpp:7-24
    return guard([&] {
        const Mesh& m = *mdot->mesh->m;
        for (int64_t c = 0; c < m.nc(); ++c) {
            f64 s = 0;
            for (auto fid : m.cells[c].faces) {
                const Face& F = m.faces[fid];
                if (!F.interior() && m.patchOf(F).empty) continue;
                s += std::abs(mdot->v->flux(fid));
            }
            co[c] = 0.5 * dt * s / m.cells[c].volume;
        }
    });
}
