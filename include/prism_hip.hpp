// prism_hip.hpp -- C++20 host mirror of the reference's public API (eigemx/prism) on top of
// the C ABI in prism_hip.h.  The class names, constructor arguments, defaults and exception
// types follow the reference so that code written against it reads the same:
//
//   auto mesh = mesh::UnvToPMeshConverter(unv, fields_json).toPMesh();
//   auto P = makeShared<field::Scalar>("P", mesh, 0.0);
//   auto eqn = eqn::Transport(diffusion::Corrected(c, P),
//                             source::ConstantScalar<Sign::Positive>(S));
//   solver::BiCGSTAB().solve(eqn, 15, 1e-20);
//
// Every object lives on the device of one context; host vectors are copies (values() returns
// a fresh std::vector).  Eigen types are replaced by std::vector<double> and two small value
// types (Vector3d, Matrix3d).  The entry points are reached through a function table, which
// is either the linked library (Api::linked(), link with -lprism_hip) or any library exposing
// the same ABI under a prefix (Api::load(path, "prism_")); mesh reading always uses the HIP
// library's host-side reader.  Header-only; needs <dlfcn.h> for Api::load.
#pragma once

#include <dlfcn.h>

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <optional>
#include <span>
#include <sstream>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

#include "prism_hip.h"

namespace prism {

using f64 = double;
template <class T>
using SharedPtr = std::shared_ptr<T>;
template <class T, class... A>
auto makeShared(A&&... a) -> SharedPtr<T> {
    return std::make_shared<T>(std::forward<A>(a)...);
}
template <class T>
using Optional = std::optional<T>;
inline constexpr auto NullOption = std::nullopt;
using VectorXd = std::vector<f64>;
inline constexpr f64 PI = 3.14159265358979323846;

struct Vector3d {
    f64 v[3]{0, 0, 0};
    Vector3d() = default;
    Vector3d(f64 x, f64 y, f64 z) : v{x, y, z} {}
    f64 x() const { return v[0]; }
    f64 y() const { return v[1]; }
    f64 z() const { return v[2]; }
    f64 operator[](int i) const { return v[i]; }
    f64 dot(const Vector3d& o) const { return (v[0] * o.v[0] + v[1] * o.v[1]) + v[2] * o.v[2]; }
};

struct Matrix3d {
    std::array<f64, 9> m{};  // row-major
    static Matrix3d Identity() {
        Matrix3d a;
        a.m[0] = a.m[4] = a.m[8] = 1.0;
        return a;
    }
};

enum class VectorCoord { X = 0, Y = 1, Z = 2 };

// --------------------------------------------------------------------- exceptions ----
namespace error {
struct NonImplementedBoundaryCondition : std::runtime_error {  // exceptions.h:9-18
    using std::runtime_error::runtime_error;
};
struct InvalidMesh : std::runtime_error {  // exceptions.h:20-27
    using std::runtime_error::runtime_error;
};
struct DeviceError : std::runtime_error {  // HIP runtime failure (no reference counterpart)
    using std::runtime_error::runtime_error;
};
}  // namespace error

// ---------------------------------------------------------------- function table ----
#define PRISM_HIP_FUNCS(X)                                                                                  \
    X(last_error) X(ctx_create) X(ctx_destroy) X(ctx_synchronize) X(mesh_create) X(mesh_destroy)            \
    X(mesh_get_info) X(mesh_geometry) X(scalar_create) X(field_destroy) X(scalar_set) X(scalar_get)          \
    X(scalar_fill) X(scalar_set_face_values) X(scalar_clear_face_values) X(scalar_grad)                      \
    X(scalar_face_values) X(scalar_set_history) X(scalar_update_prev) X(vector_create) X(vector_clone)       \
    X(vfield_destroy) X(vector_component) X(vector_set_face_values) X(vector_set_face_fluxes)                \
    X(vector_face_fluxes) X(vector_face_values) X(tensor_create) X(tfield_destroy) X(eqn_create)             \
    X(eqn_destroy) X(eqn_update_coeffs) X(eqn_zero_coeffs) X(eqn_set_under_relax) X(eqn_relax)              \
    X(eqn_add_symmetry_handler) X(eqn_nnz) X(eqn_matrix) X(eqn_rhs) X(eqn_diag) X(solve_params_default)     \
    X(solve) X(csr_create) X(csr_destroy) X(csr_bicgstab) X(csr_jacobi_step) X(simple_controls_default)     \
    X(piso_controls_default) X(simple_step) X(prime_step) X(piso_step) X(needs_pressure_reference)            \
    X(mesh_n_vertices) X(field_to_points) X(vfield_to_points) X(tfield_to_points) X(export_vtu) X(write_pvd)     \
    X(kepsilon_update)

#define PRISM_HIP_MESHGEN_FUNCS(X) X(meshgen_unv) X(meshgen_box) X(meshgen_desc) X(meshgen_destroy)

struct Api {
#define PRISM_X(n) decltype(&::prism_##n) n = nullptr;
    PRISM_HIP_FUNCS(PRISM_X)
    PRISM_HIP_MESHGEN_FUNCS(PRISM_X)
#undef PRISM_X

    // The library this program links (-lprism_hip).
    static Api linked() {
        Api a;
#define PRISM_X(n) a.n = &::prism_##n;
        PRISM_HIP_FUNCS(PRISM_X)
        PRISM_HIP_MESHGEN_FUNCS(PRISM_X)
#undef PRISM_X
        return a;
    }
    // Compute entry points from `path` under `prefix`; mesh reading from `meshgen_path`
    // (the HIP library, whose reader is host code).
    static Api load(const std::string& path, const std::string& prefix, const std::string& meshgen_path) {
        Api a;
        void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) throw std::runtime_error("prism: cannot load " + path + ": " + dlerror());
        void* g = dlopen(meshgen_path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!g) throw std::runtime_error("prism: cannot load " + meshgen_path + ": " + dlerror());
        auto sym = [](void* lib, const std::string& name) {
            void* s = dlsym(lib, name.c_str());
            if (!s) throw std::runtime_error("prism: missing symbol " + name);
            return s;
        };
#define PRISM_X(n) a.n = reinterpret_cast<decltype(a.n)>(sym(h, prefix + #n));
        PRISM_HIP_FUNCS(PRISM_X)
#undef PRISM_X
#define PRISM_X(n) a.n = reinterpret_cast<decltype(a.n)>(sym(g, std::string("prism_") + #n));
        PRISM_HIP_MESHGEN_FUNCS(PRISM_X)
#undef PRISM_X
        return a;
    }
};

namespace detail {
inline Api& api() {
    static Api a;
    return a;
}
inline void check(int rc) {
    if (rc == PRISM_OK) return;
    std::string msg = api().last_error ? api().last_error() : "";
    switch (rc) {
        case PRISM_ERR_INVALID: throw std::invalid_argument(msg);
        case PRISM_ERR_BC: throw error::NonImplementedBoundaryCondition(msg);
        case PRISM_ERR_MESH: throw error::InvalidMesh(msg);
        case PRISM_ERR_DEVICE:
        case PRISM_ERR_NOMEM: throw error::DeviceError(msg);
        default: throw std::runtime_error(msg);
    }
}
}  // namespace detail

// One device context per process (device 0 unless init() says otherwise).
class Context {
  public:
    static void init(const Api& api, int device = 0) {
        detail::api() = api;
        auto& c = instance();
        if (c.h_) detail::api().ctx_destroy(c.h_);
        detail::check(detail::api().ctx_create(device, &c.h_));
    }
    static prism_ctx* handle() {
        if (!instance().h_) throw std::runtime_error("prism: Context::init() was not called");
        return instance().h_;
    }

  private:
    prism_ctx* h_{nullptr};
    static Context& instance() {
        static Context c;
        return c;
    }
};

// ------------------------------------------------------------------ small JSON ----
// Enough of JSON for fields.json and the per-field boundary files (mesh/boundary.cpp).
namespace detail {
struct Json {
    std::variant<std::nullptr_t, bool, f64, std::string, std::vector<Json>, std::map<std::string, Json>> v;
    bool isNumber() const { return std::holds_alternative<f64>(v); }
    bool isArray() const { return std::holds_alternative<std::vector<Json>>(v); }
    bool isObject() const { return std::holds_alternative<std::map<std::string, Json>>(v); }
    bool contains(const std::string& k) const { return isObject() && obj().count(k); }
    const std::map<std::string, Json>& obj() const { return std::get<std::map<std::string, Json>>(v); }
    const std::vector<Json>& arr() const { return std::get<std::vector<Json>>(v); }
    const Json& operator[](const std::string& k) const { return obj().at(k); }
    const std::string& str() const { return std::get<std::string>(v); }
    f64 num() const { return std::get<f64>(v); }
};
struct JsonParser {
    const std::string& s;
    size_t i{0};
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    [[noreturn]] void bad() { throw std::runtime_error("prism: malformed JSON at offset " + std::to_string(i)); }
    Json parse() {
        ws();
        if (i >= s.size()) bad();
        char c = s[i];
        if (c == '{') {
            ++i;
            std::map<std::string, Json> m;
            ws();
            if (s[i] == '}') { ++i; return {m}; }
            for (;;) {
                ws();
                std::string k = std::get<std::string>(parse().v);
                ws();
                if (s[i++] != ':') bad();
                m[k] = parse();
                ws();
                if (s[i] == ',') { ++i; continue; }
                if (s[i++] == '}') return {m};
                bad();
            }
        }
        if (c == '[') {
            ++i;
            std::vector<Json> a;
            ws();
            if (s[i] == ']') { ++i; return {a}; }
            for (;;) {
                a.push_back(parse());
                ws();
                if (s[i] == ',') { ++i; continue; }
                if (s[i++] == ']') return {a};
                bad();
            }
        }
        if (c == '"') {
            ++i;
            std::string out;
            while (i < s.size() && s[i] != '"') {
                if (s[i] == '\\' && i + 1 < s.size()) ++i;
                out += s[i++];
            }
            ++i;
            return {out};
        }
        if (s.compare(i, 4, "true") == 0) { i += 4; return {true}; }
        if (s.compare(i, 5, "false") == 0) { i += 5; return {false}; }
        if (s.compare(i, 4, "null") == 0) { i += 4; return {nullptr}; }
        size_t end = i;
        f64 d = std::stod(s.substr(i), &end);
        i += end;
        return {d};
    }
};
inline Json readJson(const std::filesystem::path& p) {
    std::ifstream f(p);
    if (!f) throw std::runtime_error("prism: cannot open " + p.string());
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str();
    JsonParser jp{text};
    return jp.parse();
}
}  // namespace detail

// -------------------------------------------------------------------------- mesh ----
namespace mesh {

struct BoundaryCondition {  // mesh/boundary.h: kind string + Nil / Scalar / Vector value
    std::string kind;
    int value_kind{0};
    f64 value[3]{0, 0, 0};
};

struct FieldInfo {
    std::string name, type;
    Optional<std::string> grad_scheme;
};

class PMesh;

// A boundary patch's conditions (mesh/boundary.h: BoundaryPatch).
class BoundaryPatch {
  public:
    BoundaryPatch(std::string name, std::map<std::string, BoundaryCondition> bcs)
        : name_(std::move(name)), bcs_(std::move(bcs)) {}
    auto name() const -> const std::string& { return name_; }
    auto isEmpty() const -> bool { return !bcs_.empty() && bcs_.begin()->second.kind == "empty"; }
    auto getBoundaryCondition(const std::string& field) const -> const BoundaryCondition& {
        auto it = bcs_.find(field);
        if (it == bcs_.end())
            throw std::runtime_error("prism::mesh::BoundaryPatch: patch '" + name_ + "' has no field '" + field + "'");
        return it->second;
    }
    auto getVectorBoundaryCondition(const std::string& field) const -> Vector3d {  // boundary.cpp:471-489
        const auto& bc = getBoundaryCondition(field);
        if (bc.value_kind != 2)
            throw std::runtime_error("prism::mesh::BoundaryPatch::getVectorBoundaryCondition(): Boundary patch '" +
                                     name_ + "' field '" + field + "' is not a vector");
        return {bc.value[0], bc.value[1], bc.value[2]};
    }
    auto getScalarBoundaryCondition(const std::string& field) const -> f64 {
        const auto& bc = getBoundaryCondition(field);
        if (bc.value_kind != 1) throw std::runtime_error("prism::mesh::BoundaryPatch: field '" + field + "' is not a scalar");
        return bc.value[0];
    }
    auto facesIds() const -> const std::vector<size_t>& { return faces_; }

  private:
    friend class PMesh;
    std::string name_;
    std::map<std::string, BoundaryCondition> bcs_;
    std::vector<size_t> faces_;
};

class PMesh {
  public:
    PMesh(prism_meshgen* gen, std::map<std::string, std::map<std::string, BoundaryCondition>> bcs,
          std::vector<FieldInfo> fields)
        : bcs_(std::move(bcs)), fields_(std::move(fields)) {
        auto& A = detail::api();
        prism_mesh_desc d{};
        detail::check(A.meshgen_desc(gen, &d));
        std::vector<prism_bc_desc> bl;
        for (int32_t p = 0; p < d.n_patches; ++p) {
            auto it = bcs_.find(d.patch_names[p]);
            if (it == bcs_.end()) continue;
            for (auto& [field, bc] : it->second) {
                prism_bc_desc b{};
                b.patch = p;
                b.field = field.c_str();
                b.kind = bc.kind.c_str();
                b.value_kind = bc.value_kind;
                for (int k = 0; k < 3; ++k) b.value[k] = bc.value[k];
                bl.push_back(b);
            }
        }
        std::vector<prism_field_info> fl;
        for (auto& f : fields_)
            fl.push_back({f.name.c_str(), f.type.c_str(), f.grad_scheme ? f.grad_scheme->c_str() : nullptr});
        d.n_bcs = (int32_t)bl.size();
        d.bcs = bl.data();
        d.n_fields = (int32_t)fl.size();
        d.fields = fl.data();
        std::vector<std::string> names;
        for (int32_t p = 0; p < d.n_patches; ++p) names.emplace_back(d.patch_names[p]);
        detail::check(A.mesh_create(Context::handle(), &d, &h_));
        prism_mesh_info info{};
        detail::check(A.mesh_get_info(h_, &info));
        n_ = info.n_cells;
        F_ = info.n_faces;
        centers_.resize(3 * n_);
        vol_.resize(n_);
        area_.resize(3 * F_);
        nbr_.resize(F_);
        fpatch_.resize(F_);
        detail::check(A.mesh_geometry(h_, centers_.data(), vol_.data(), nullptr, area_.data(), nullptr, nbr_.data(),
                                      fpatch_.data()));
        for (size_t p = 0; p < names.size(); ++p) {
            auto it = bcs_.find(names[p]);
            patches_.emplace_back(names[p], it == bcs_.end() ? std::map<std::string, BoundaryCondition>{} : it->second);
        }
        for (int64_t f = 0; f < F_; ++f)
            if (nbr_[f] < 0 && fpatch_[f] >= 0) patches_[fpatch_[f]].faces_.push_back((size_t)f);
    }
    ~PMesh() {
        if (h_) detail::api().mesh_destroy(h_);
    }
    PMesh(const PMesh&) = delete;
    PMesh& operator=(const PMesh&) = delete;

    auto cellCount() const -> size_t { return (size_t)n_; }
    auto faceCount() const -> size_t { return (size_t)F_; }
    auto cellCenter(size_t c) const -> Vector3d { return {centers_[3 * c], centers_[3 * c + 1], centers_[3 * c + 2]}; }
    auto cellsVolumeVector() const -> const VectorXd& { return vol_; }
    // Face area vector S_f (owner-outward) in the library's face numbering.
    auto faceAreaVector(size_t f) const -> Vector3d { return {area_[3 * f], area_[3 * f + 1], area_[3 * f + 2]}; }
    auto boundaryPatches() const -> const std::vector<BoundaryPatch>& { return patches_; }
    prism_mesh* handle() const { return h_; }

  private:
    prism_mesh* h_{nullptr};
    int64_t n_{0}, F_{0};
    VectorXd centers_, vol_, area_;
    std::vector<int64_t> nbr_;
    std::vector<int32_t> fpatch_;
    std::vector<BoundaryPatch> patches_;
    std::map<std::string, std::map<std::string, BoundaryCondition>> bcs_;
    std::vector<FieldInfo> fields_;
};

// UnvToPMeshConverter(unv, fields.json).toPMesh() (unv.cpp:61-110, mesh/boundary.cpp:291-319)
class UnvToPMeshConverter {
  public:
    UnvToPMeshConverter(std::filesystem::path unv, std::filesystem::path fields_json)
        : unv_(std::move(unv)), fields_(std::move(fields_json)) {}

    auto toPMesh() -> SharedPtr<PMesh> {
        std::vector<FieldInfo> infos;
        std::map<std::string, std::map<std::string, BoundaryCondition>> bcs;
        auto doc = detail::readJson(fields_);
        for (const auto& f : doc["fields"].arr()) {
            if (!f.contains("name") || !f.contains("type"))
                throw std::runtime_error("prism::mesh::parseField(): Field must contain name and type");
            FieldInfo fi{f["name"].str(), f["type"].str(), NullOption};
            if (f.contains("gradScheme")) fi.grad_scheme = f["gradScheme"].str();
            infos.push_back(fi);
        }
        for (const auto& fi : infos) {
            auto file = fields_.parent_path() / (fi.name + ".json");
            if (!std::filesystem::exists(file))
                throw std::runtime_error("prism::mesh: boundary file of field `" + fi.name + "` is missing");
            auto bdoc = detail::readJson(file);
            if (!bdoc.contains("patches") || !bdoc["patches"].isArray())
                throw std::runtime_error("prism::mesh::readBoundaryFile(): file `" + fi.name +
                                         "` has no `patches` array");
            for (const auto& p : bdoc["patches"].arr()) {
                if (!p.contains("name") || !p.contains("type"))
                    throw std::runtime_error("prism::mesh::readBoundaryFile(): patch without name or type");
                BoundaryCondition bc{p["type"].str()};
                if (p.contains("value")) {
                    const auto& v = p["value"];
                    if (v.isNumber()) {
                        bc.value_kind = 1;
                        bc.value[0] = v.num();
                    } else if (v.isArray()) {
                        if (v.arr().size() != 3) throw std::runtime_error("prism::mesh: expected array of size 3");
                        bc.value_kind = 2;
                        for (int k = 0; k < 3; ++k) bc.value[k] = v.arr()[k].num();
                    } else {
                        continue;
                    }
                }
                bcs[p["name"].str()].emplace(fi.name, bc);
            }
        }
        prism_meshgen* gen = nullptr;
        detail::check(detail::api().meshgen_unv(unv_.string().c_str(), &gen));
        try {
            auto m = std::make_shared<PMesh>(gen, bcs, infos);
            detail::api().meshgen_destroy(gen);
            return m;
        } catch (...) {
            detail::api().meshgen_destroy(gen);
            throw;
        }
    }

  private:
    std::filesystem::path unv_, fields_;
};

}  // namespace mesh

// ------------------------------------------------------------------------ fields ----
namespace field {

class Scalar {  // field::Scalar (scalar.h:14-127)
  public:
    Scalar(std::string name, SharedPtr<mesh::PMesh> mesh, f64 value, int kind = 0)
        : name_(std::move(name)), mesh_(std::move(mesh)) {
        detail::check(detail::api().scalar_create(mesh_->handle(), name_.c_str(), kind, nullptr, value, &h_));
    }
    Scalar(std::string name, SharedPtr<mesh::PMesh> mesh, const VectorXd& values, int kind = 0)
        : name_(std::move(name)), mesh_(std::move(mesh)) {
        if (values.size() != mesh_->cellCount()) throw std::runtime_error("Scalar: values size != cell count");
        detail::check(detail::api().scalar_create(mesh_->handle(), name_.c_str(), kind, values.data(), 0.0, &h_));
    }
    Scalar(std::string name, SharedPtr<mesh::PMesh> mesh, prism_field* borrowed)
        : name_(std::move(name)), mesh_(std::move(mesh)), h_(borrowed), owned_(false) {}
    virtual ~Scalar() {
        if (h_ && owned_) detail::api().field_destroy(h_);
    }
    Scalar(const Scalar&) = delete;
    Scalar& operator=(const Scalar&) = delete;

    auto name() const -> const std::string& { return name_; }
    auto mesh() const -> const SharedPtr<mesh::PMesh>& { return mesh_; }
    auto values() const -> VectorXd {
        VectorXd v(mesh_->cellCount());
        detail::check(detail::api().scalar_get(h_, v.data()));
        return v;
    }
    void setValues(const VectorXd& v) { detail::check(detail::api().scalar_set(h_, v.data())); }
    auto valueAtCell(size_t c) const -> f64 { return values()[c]; }
    void setFaceValues(const VectorXd& fv) { detail::check(detail::api().scalar_set_face_values(h_, fv.data())); }
    void clearFaceValues() { detail::check(detail::api().scalar_clear_face_values(h_)); }
    void setHistorySize(size_t n) { detail::check(detail::api().scalar_set_history(h_, (int32_t)n)); }
    void updatePrevTimeSteps() { detail::check(detail::api().scalar_update_prev(h_)); }
    prism_field* handle() const { return h_; }

  private:
    std::string name_;
    SharedPtr<mesh::PMesh> mesh_;
    prism_field* h_{nullptr};
    bool owned_{true};
};

class Pressure : public Scalar {  // pressure.h:6-9
  public:
    Pressure(std::string name, SharedPtr<mesh::PMesh> mesh, f64 value) : Scalar(std::move(name), std::move(mesh), value, 1) {}
};

class Tensor {  // field::Tensor (tensor.h:7-38)
  public:
    Tensor(std::string name, SharedPtr<mesh::PMesh> mesh, const Matrix3d& value)
        : name_(std::move(name)), mesh_(std::move(mesh)) {
        detail::check(detail::api().tensor_create(mesh_->handle(), name_.c_str(), nullptr, value.m.data(), &h_));
    }
    ~Tensor() {
        if (h_) detail::api().tfield_destroy(h_);
    }
    Tensor(const Tensor&) = delete;
    auto name() const -> const std::string& { return name_; }
    auto mesh() const -> const SharedPtr<mesh::PMesh>& { return mesh_; }
    prism_tfield* handle() const { return h_; }

  private:
    std::string name_;
    SharedPtr<mesh::PMesh> mesh_;
    prism_tfield* h_{nullptr};
};

class Vector {  // field::Vector / GeneralVector (vector.h:17-101)
  public:
    Vector(std::string name, SharedPtr<mesh::PMesh> mesh, const Vector3d& value)
        : name_(std::move(name)), mesh_(std::move(mesh)) {
        detail::check(detail::api().vector_create(mesh_->handle(), name_.c_str(), nullptr, value.v, &h_));
        bindComponents();
    }
    ~Vector() {
        comp_ = {};
        if (h_) detail::api().vfield_destroy(h_);
    }
    Vector(const Vector&) = delete;
    auto clone() const -> Vector { return Vector(*this, 0); }
    auto x() const -> SharedPtr<Scalar> { return comp_[0]; }
    auto y() const -> SharedPtr<Scalar> { return comp_[1]; }
    auto z() const -> SharedPtr<Scalar> { return comp_[2]; }
    auto name() const -> const std::string& { return name_; }
    auto mesh() const -> const SharedPtr<mesh::PMesh>& { return mesh_; }
    auto faceValues() const -> std::vector<Vector3d> {
        VectorXd buf(3 * mesh_->faceCount());
        detail::check(detail::api().vector_face_values(h_, buf.data()));
        std::vector<Vector3d> out(mesh_->faceCount());
        for (size_t f = 0; f < out.size(); ++f) out[f] = {buf[3 * f], buf[3 * f + 1], buf[3 * f + 2]};
        return out;
    }
    prism_vfield* handle() const { return h_; }
    Vector(Vector&& o) noexcept : name_(std::move(o.name_)), mesh_(std::move(o.mesh_)), h_(o.h_), comp_(std::move(o.comp_)) {
        o.h_ = nullptr;
    }

  private:
    Vector(const Vector& src, int) : name_(src.name_), mesh_(src.mesh_) {
        detail::check(detail::api().vector_clone(src.h_, &h_));
        bindComponents();
    }
    void bindComponents() {
        static const char* sfx[3] = {"_x", "_y", "_z"};
        for (int k = 0; k < 3; ++k) {
            prism_field* c = nullptr;
            detail::check(detail::api().vector_component(h_, k, &c));
            comp_[k] = std::make_shared<Scalar>(name_ + sfx[k], mesh_, c);
        }
    }
    std::string name_;
    SharedPtr<mesh::PMesh> mesh_;
    prism_vfield* h_{nullptr};
    std::array<SharedPtr<Scalar>, 3> comp_;
};

using Velocity = Vector;  // velocity.h:18-23

}  // namespace field

// ----------------------------------------------------------------------- schemes ----
enum class Sign { Positive, Negative };

namespace scheme {
using prism::Sign;

struct Desc {  // what every scheme hands to eqn::Transport
    prism_scheme_desc d{};
    std::vector<std::shared_ptr<const void>> keep;  // fields the equation keeps alive
    SharedPtr<field::Scalar> phi;                   // the scheme's (transported or source) field
};

namespace convection {
template <int NVF>
class NVFScheme {
  public:
    NVFScheme(SharedPtr<field::Vector> U, SharedPtr<field::Scalar> phi) : U_(std::move(U)), phi_(std::move(phi)) {}
    Desc desc() const {
        Desc r;
        r.d.kind = PRISM_SCHEME_CONVECTION;
        r.d.variant = NVF;
        r.d.phi = phi_->handle();
        r.d.vec = U_->handle();
        r.keep = {U_, phi_};
        r.phi = phi_;
        return r;
    }

  private:
    SharedPtr<field::Vector> U_;
    SharedPtr<field::Scalar> phi_;
};
using Upwind = NVFScheme<PRISM_NVF_UPWIND>;
using CentralDifference = NVFScheme<PRISM_NVF_CD>;
using LinearUpwind = NVFScheme<PRISM_NVF_LUD>;
using QUICK = NVFScheme<PRISM_NVF_QUICK>;
using FROMM = NVFScheme<PRISM_NVF_FROMM>;
using MINMOD = NVFScheme<PRISM_NVF_MINMOD>;
using MUSCL = NVFScheme<PRISM_NVF_MUSCL>;
using SMART = NVFScheme<PRISM_NVF_SMART>;
}  // namespace convection

namespace diffusion {
template <int CORR>
class DiffusionScheme {
  public:
    DiffusionScheme(SharedPtr<field::Scalar> kappa, SharedPtr<field::Scalar> phi)
        : ks_(std::move(kappa)), phi_(std::move(phi)) {}
    DiffusionScheme(SharedPtr<field::Tensor> kappa, SharedPtr<field::Scalar> phi)
        : kt_(std::move(kappa)), phi_(std::move(phi)) {}
    Desc desc() const {
        Desc r;
        r.d.kind = PRISM_SCHEME_DIFFUSION;
        r.d.variant = CORR;
        r.d.phi = phi_->handle();
        r.d.kappa = ks_ ? ks_->handle() : nullptr;
        r.d.kappa_t = kt_ ? kt_->handle() : nullptr;
        r.keep = {phi_};
        if (ks_) r.keep.push_back(ks_);
        if (kt_) r.keep.push_back(kt_);
        r.phi = phi_;
        return r;
    }

  private:
    SharedPtr<field::Scalar> ks_;
    SharedPtr<field::Tensor> kt_;
    SharedPtr<field::Scalar> phi_;
};
using NonCorrected = DiffusionScheme<PRISM_DIFF_NONCORRECTED>;
using Corrected = DiffusionScheme<PRISM_DIFF_OVERRELAXED>;  // diffusion.h:153: OverRelaxed corrector
}  // namespace diffusion

namespace source {
template <Sign S>
class ConstantScalar {  // constant_scalar.h:27-37
  public:
    explicit ConstantScalar(SharedPtr<field::Scalar> phi) : phi_(std::move(phi)) {}
    Desc desc() const {
        Desc r;
        r.d.kind = PRISM_SCHEME_SRC_CONSTANT;
        r.d.variant = S == Sign::Positive ? PRISM_SIGN_POSITIVE : PRISM_SIGN_NEGATIVE;
        r.d.phi = phi_->handle();
        r.keep = {phi_};
        r.phi = phi_;
        return r;
    }

  private:
    SharedPtr<field::Scalar> phi_;
};
template <Sign S>
class Gradient {  // source/gradient.h:32-40
  public:
    Gradient(SharedPtr<field::Scalar> phi, VectorCoord coord) : phi_(std::move(phi)), coord_(coord) {}
    Desc desc() const {
        Desc r;
        r.d.kind = PRISM_SCHEME_SRC_GRADIENT;
        r.d.variant = S == Sign::Positive ? PRISM_SIGN_POSITIVE : PRISM_SIGN_NEGATIVE;
        r.d.phi = phi_->handle();
        r.d.coord = (int32_t)coord_;
        r.keep = {phi_};
        r.phi = phi_;
        return r;
    }

  private:
    SharedPtr<field::Scalar> phi_;
    VectorCoord coord_;
};
template <Sign S>
class Divergence {  // divergence.h:25-34
  public:
    explicit Divergence(SharedPtr<field::Vector> U) : U_(std::move(U)) {}
    Desc desc() const {
        Desc r;
        r.d.kind = PRISM_SCHEME_SRC_DIVERGENCE;
        r.d.variant = S == Sign::Positive ? PRISM_SIGN_POSITIVE : PRISM_SIGN_NEGATIVE;
        r.d.vec = U_->handle();
        r.keep = {U_};
        return r;
    }

  private:
    SharedPtr<field::Vector> U_;
};
template <Sign S>
class ImplicitField {  // implicit_field.h:23-32
  public:
    ImplicitField(f64 coeff, SharedPtr<field::Scalar> phi) : coeff_(coeff), phi_(std::move(phi)) {}
    Desc desc() const {
        Desc r;
        r.d.kind = PRISM_SCHEME_SRC_IMPLICIT;
        r.d.variant = S == Sign::Positive ? PRISM_SIGN_POSITIVE : PRISM_SIGN_NEGATIVE;
        r.d.phi = phi_->handle();
        r.d.coeff = coeff_;
        r.keep = {phi_};
        r.phi = phi_;
        return r;
    }

  private:
    f64 coeff_;
    SharedPtr<field::Scalar> phi_;
};
}  // namespace source

namespace temporal {
template <int KIND, const char* NAME>
class Temporal {
  public:
    Temporal(SharedPtr<field::Scalar> phi, f64 dt) : phi_(std::move(phi)), dt_(dt) {
        if (!(dt > 0.0)) throw std::invalid_argument(std::string("prism::scheme::temporal::") + NAME + "(phi, dt): dt must be positive");
    }
    Desc desc() const {
        Desc r;
        r.d.kind = KIND;
        r.d.phi = phi_->handle();
        r.d.coeff = dt_;
        r.keep = {phi_};
        r.phi = phi_;
        return r;
    }

  private:
    SharedPtr<field::Scalar> phi_;
    f64 dt_;
};
inline constexpr char kBackwardEuler[] = "BackwardEuler";
inline constexpr char kAdamMoulton[] = "AdamMoulton";
using BackwardEuler = Temporal<PRISM_SCHEME_BACKWARD_EULER, kBackwardEuler>;
using AdamMoulton = Temporal<PRISM_SCHEME_ADAM_MOULTON, kAdamMoulton>;
}  // namespace temporal
}  // namespace scheme

// ---------------------------------------------------------------------- equation ----
namespace eqn {
class Transport;
namespace boundary {
template <class Eqn>
struct Symmetry {};  // eqn::boundary::Symmetry<Transport> (equation/boundary.h:64)
}  // namespace boundary

class Transport {  // eqn::Transport (transport.h:27-81)
  public:
    template <class... Schemes>
    explicit Transport(Schemes&&... schemes) {
        std::vector<scheme::Desc> ds{schemes.desc()...};
        std::vector<prism_scheme_desc> raw;
        for (auto& d : ds) {
            raw.push_back(d.d);
            for (auto& k : d.keep) keep_.push_back(k);
        }
        detail::check(detail::api().eqn_create(raw.data(), (int32_t)raw.size(), &h_));
        phi_ = ds.front().phi;  // the first scheme's field is the equation's (transport.h:98-101)
    }
    ~Transport() {
        if (h_) detail::api().eqn_destroy(h_);
    }
    Transport(const Transport&) = delete;
    Transport(Transport&& o) noexcept : h_(o.h_), keep_(std::move(o.keep_)), phi_(std::move(o.phi_)) { o.h_ = nullptr; }

    void updateCoeffs() { detail::check(detail::api().eqn_update_coeffs(h_)); }
    void zeroOutCoeffs() { detail::check(detail::api().eqn_zero_coeffs(h_)); }
    void setUnderRelaxFactor(f64 f) { detail::check(detail::api().eqn_set_under_relax(h_, f)); }
    void relax() { detail::check(detail::api().eqn_relax(h_)); }
    auto field() const -> const SharedPtr<field::Scalar>& { return phi_; }
    auto rhs() const -> VectorXd {
        VectorXd b(phi_->mesh()->cellCount());
        detail::check(detail::api().eqn_rhs(h_, b.data()));
        return b;
    }
    auto diagonal() const -> VectorXd {
        VectorXd d(phi_->mesh()->cellCount());
        detail::check(detail::api().eqn_diag(h_, d.data()));
        return d;
    }
    // boundaryHandlersManager().addHandler<eqn::boundary::Symmetry<Transport>>()
    struct HandlerManager {
        Transport* e;
        template <class H>
        void addHandler() {
            static_assert(std::is_same_v<H, boundary::Symmetry<Transport>>, "only Symmetry<Transport> exists");
            detail::check(detail::api().eqn_add_symmetry_handler(e->h_));
        }
    };
    auto boundaryHandlersManager() -> HandlerManager { return {this}; }
    prism_eqn* handle() const { return h_; }

  private:
    prism_eqn* h_{nullptr};
    std::vector<std::shared_ptr<const void>> keep_;
    SharedPtr<field::Scalar> phi_;
};
using Momentum = Transport;  // transport.h:83-86
}  // namespace eqn

// ------------------------------------------------------------------------ solver ----
namespace solver {
class SolverResult {  // solver.h:8-23
  public:
    explicit SolverResult(const prism_solve_result& r) : r_(r) {}
    auto iteration() const -> size_t { return (size_t)r_.iteration; }
    auto initialResidual() const -> f64 { return r_.initial_residual; }
    auto finalResidual() const -> f64 { return r_.final_residual; }
    auto hasConverged() const -> bool { return r_.converged != 0; }
    auto krylovIterations() const -> int64_t { return r_.krylov_iters; }

  private:
    prism_solve_result r_;
};

class BiCGSTAB {  // bicgstab.h:12, solver.cpp:27-82
  public:
    BiCGSTAB() { detail::api().solve_params_default(&p_); }
    // Krylov settings (no reference counterpart): precision, tolerance, iteration cap
    auto& params() { return p_; }
    auto solve(eqn::Transport& eqn, size_t n_iter = 10, f64 eps = 1e-7, Optional<size_t> ref_cell = NullOption,
               f64 ref_value = 0.0, bool clear_coeffs = true) -> SolverResult {
        prism_solve_params p = p_;
        p.n_iter = (int64_t)n_iter;
        p.eps = eps;
        p.ref_cell = ref_cell ? (int64_t)*ref_cell : -1;
        p.ref_value = ref_value;
        p.clear_coeffs = clear_coeffs ? 1 : 0;
        prism_solve_result r{};
        detail::check(detail::api().solve(eqn.handle(), &p, &r));
        return SolverResult(r);
    }

  private:
    prism_solve_params p_{};
};
}  // namespace solver

// ------------------------------------------------------------------- algorithms ----
namespace algo {
struct SIMPLEControls {  // SIMPLE.h:14-43
    f64 momentum_urf = 0.7;
    size_t momentum_max_iter = 3;
    f64 momentum_residual = 1e-7;
    f64 momentum_residual_stop = 1e-5;  // carried, unused (as in the reference)
    size_t non_ortho_correctors = 2;
    f64 pressure_urf = 0.3;
    size_t pressure_max_iter = 5;
    f64 pressure_residual = 1e-7;
    Optional<size_t> p_ref_cell = NullOption;
    f64 p_ref_value = 0.0;

    prism_simple_controls raw() const {
        prism_simple_controls c{};
        prism::detail::api().simple_controls_default(&c);  // Krylov settings
        c.momentum_urf = momentum_urf;
        c.momentum_max_iter = (int64_t)momentum_max_iter;
        c.momentum_residual = momentum_residual;
        c.non_ortho_correctors = (int64_t)non_ortho_correctors;
        c.pressure_urf = pressure_urf;
        c.pressure_max_iter = (int64_t)pressure_max_iter;
        c.pressure_residual = pressure_residual;
        c.p_ref_cell = p_ref_cell ? (int64_t)*p_ref_cell : -1;
        c.p_ref_value = p_ref_value;
        return c;
    }
};
struct PISOControls {  // PISO.h:11-49
    size_t outer_iterations = 1;
    size_t momentum_implicit_steps = 1;
    size_t pressure_correction_steps = 2;
    f64 momentum_urf = 1.0;
    size_t momentum_max_iter = 3;
    f64 momentum_residual = 1e-4;
    f64 momentum_residual_stop = 1e-5;  // carried, unused (as in the reference)
    size_t non_ortho_correctors = 2;
    f64 pressure_urf = 1.0;
    size_t pressure_max_iter = 5;
    f64 pressure_residual = 1e-4;
    Optional<size_t> p_ref_cell = NullOption;
    f64 p_ref_value = 0.0;

    prism_piso_controls raw() const {
        prism_piso_controls c{};
        prism::detail::api().piso_controls_default(&c);
        c.outer_iterations = (int64_t)outer_iterations;
        c.momentum_implicit_steps = (int64_t)momentum_implicit_steps;
        c.pressure_correction_steps = (int64_t)pressure_correction_steps;
        c.simple.momentum_urf = momentum_urf;
        c.simple.momentum_max_iter = (int64_t)momentum_max_iter;
        c.simple.momentum_residual = momentum_residual;
        c.simple.non_ortho_correctors = (int64_t)non_ortho_correctors;
        c.simple.pressure_urf = pressure_urf;
        c.simple.pressure_max_iter = (int64_t)pressure_max_iter;
        c.simple.pressure_residual = pressure_residual;
        c.simple.p_ref_cell = p_ref_cell ? (int64_t)*p_ref_cell : -1;
        c.simple.p_ref_value = p_ref_value;
        return c;
    }
};

namespace detail_algo {
inline std::vector<prism_eqn*> handles(std::span<eqn::Momentum*> m) {
    std::vector<prism_eqn*> h;
    for (auto* e : m) h.push_back(e->handle());
    return h;
}
}  // namespace detail_algo

class IncompressibleSIMPLE {  // SIMPLE.h:45-55
  public:
    explicit IncompressibleSIMPLE(SIMPLEControls c) : c_(c) {}
    void step(std::span<eqn::Momentum*> momentum, SharedPtr<field::Velocity>& U, SharedPtr<field::Velocity>& mdot,
              SharedPtr<field::Pressure>& p) {
        auto h = detail_algo::handles(momentum);
        prism_simple_controls c = c_.raw();
        prism_report rep[64];
        int32_t n = 0;
        prism::detail::check(prism::detail::api().simple_step(h.data(), (int32_t)h.size(), U->handle(), mdot->handle(),
                                                              p->handle(), &c, rep, 64, &n));
    }

  private:
    SIMPLEControls c_;
};

class PISO {  // PISO.h:51-61
  public:
    explicit PISO(PISOControls c) : c_(c) {}
    void step(std::span<eqn::Momentum*> momentum, SharedPtr<field::Velocity>& U, SharedPtr<field::Velocity>& mdot,
              SharedPtr<field::Pressure>& p) {
        auto h = detail_algo::handles(momentum);
        prism_piso_controls c = c_.raw();
        prism_report rep[256];
        int32_t n = 0;
        prism::detail::check(prism::detail::api().piso_step(h.data(), (int32_t)h.size(), U->handle(), mdot->handle(),
                                                            p->handle(), &c, rep, 256, &n));
    }

  private:
    PISOControls c_;
};

inline auto needsPressureReference(const SharedPtr<field::Pressure>& p) -> bool {  // SIMPLE.cpp:110-122
    int32_t out = 0;
    prism::detail::check(prism::detail::api().needs_pressure_reference(p->handle(), &out));
    return out != 0;
}
}  // namespace algo

// -------------------------------------------------------------------- turbulence ----
// turbulence::IncompressibleKEpsilon (KEpsilon.h:11-40, KEpsilon.cpp): equations() refreshes
// nu_t, Pk and the coefficient fields in one device pass, then builds the k and epsilon
// Transports as KEpsilon.cpp:41-77 does (new ones on every call, as there).
namespace turbulence {
class IncompressibleKEpsilon {
  public:
    IncompressibleKEpsilon(const SharedPtr<field::Velocity>& mdot, const SharedPtr<field::Velocity>& U,
                           const SharedPtr<field::Scalar>& mu, Optional<f64> dt = NullOption)
        : mdot_(mdot), U_(U), nu_(mu) {
        const auto& m = mu->mesh();
        k_ = makeShared<field::Scalar>("k", m, 0.0);
        eps_ = makeShared<field::Scalar>("epsilon", m, 0.0);
        nu_t_ = makeShared<field::Scalar>("mu_t", m, 0.0);
        pk_ = makeShared<field::Scalar>("Pk", m, 0.0);
        nu_eff_k_ = makeShared<field::Scalar>("mu_eff_k", m, 0.0);
        nu_eff_eps_ = makeShared<field::Scalar>("mu_eff_eps", m, 0.0);
        eps_source_ = makeShared<field::Scalar>("eps_source", m, 0.0);
        eps_sink_ = makeShared<field::Scalar>("eps_sink", m, 0.0);
        if (dt.has_value() && dt.value() > 0.0) {
            transient_ = true;
            dt_ = dt.value();
            k_->setHistorySize(2);
            eps_->setHistorySize(2);
        }
    }
    auto k() -> const SharedPtr<field::Scalar>& { return k_; }
    auto epsilon() -> const SharedPtr<field::Scalar>& { return eps_; }
    auto turbulentViscosity() -> const SharedPtr<field::Scalar>& { return nu_t_; }

    auto equations() -> std::pair<eqn::Transport, eqn::Transport> {
        const f64 consts[5] = {c_mu_, c_eps1_, c_eps2_, sigma_k_, sigma_eps_};
        prism_field* out[6] = {nu_t_->handle(), pk_->handle(), nu_eff_k_->handle(), nu_eff_eps_->handle(),
                               eps_source_->handle(), eps_sink_->handle()};
        prism::detail::check(prism::detail::api().kepsilon_update(k_->handle(), eps_->handle(), nu_->handle(),
                                                                  U_->handle(), consts, out));
        using namespace scheme;
        if (transient_)
            return {eqn::Transport(convection::LinearUpwind(mdot_, k_), diffusion::NonCorrected(nu_eff_k_, k_),
                                   source::ConstantScalar<Sign::Positive>(pk_),
                                   source::ConstantScalar<Sign::Negative>(eps_), temporal::AdamMoulton(k_, dt_)),
                    eqn::Transport(convection::LinearUpwind(mdot_, eps_), diffusion::NonCorrected(nu_eff_eps_, eps_),
                                   source::ConstantScalar<Sign::Positive>(eps_source_),
                                   source::ConstantScalar<Sign::Negative>(eps_sink_), temporal::AdamMoulton(eps_, dt_))};
        return {eqn::Transport(convection::LinearUpwind(mdot_, k_), diffusion::NonCorrected(nu_eff_k_, k_),
                               source::ConstantScalar<Sign::Positive>(pk_), source::ConstantScalar<Sign::Negative>(eps_)),
                eqn::Transport(convection::LinearUpwind(mdot_, eps_), diffusion::NonCorrected(nu_eff_eps_, eps_),
                               source::ConstantScalar<Sign::Positive>(eps_source_),
                               source::ConstantScalar<Sign::Negative>(eps_sink_))};
    }

  private:
    f64 c_mu_ = 0.09, c_eps1_ = 1.44, c_eps2_ = 1.92, sigma_k_ = 1.0, sigma_eps_ = 1.3;  // KEpsilon.h:29-33
    SharedPtr<field::Velocity> mdot_, U_;
    SharedPtr<field::Scalar> nu_, k_, eps_, nu_t_, pk_, nu_eff_k_, nu_eff_eps_, eps_source_, eps_sink_;
    bool transient_{false};
    f64 dt_{0.0};
};
}  // namespace turbulence

// ------------------------------------------------------------------------ output ----
// prism::output (export.h): toVtu overloads, FieldRegistry and TimeWriter.  Point data is the
// device-side inverse-distance interpolation of cellToPoint; the library writes the file.
namespace output {

enum class WriteMode : std::uint8_t { Ascii, RawBinaryCompressed };  // export.h:13

namespace detail {
inline int32_t mode(WriteMode m) {
    return m == WriteMode::Ascii ? PRISM_WRITE_ASCII : PRISM_WRITE_RAW_BINARY_COMPRESSED;
}
}  // namespace detail

// toVtu(scalars, vectors, tensors, file_name, mode) (export.cpp:262-300)
inline void toVtu(std::span<const field::Scalar*> scalars, std::span<const field::Vector*> vectors,
                  std::span<const field::Tensor*> tensors, const std::string& file_name,
                  WriteMode mode = WriteMode::RawBinaryCompressed) {
    std::vector<prism_export_item> items;
    for (auto* f : scalars) items.push_back({PRISM_EXPORT_SCALAR, f->handle()});
    for (auto* f : vectors) items.push_back({PRISM_EXPORT_VECTOR, f->handle()});
    for (auto* f : tensors) items.push_back({PRISM_EXPORT_TENSOR, f->handle()});
    prism::detail::check(
        prism::detail::api().export_vtu(items.data(), (int32_t)items.size(), file_name.c_str(), detail::mode(mode)));
}
inline void toVtu(const field::Scalar& f, const std::string& file_name, WriteMode mode = WriteMode::RawBinaryCompressed) {
    std::array<const field::Scalar*, 1> s{&f};
    toVtu(s, {}, {}, file_name, mode);
}
inline void toVtu(const field::Vector& f, const std::string& file_name, WriteMode mode = WriteMode::RawBinaryCompressed) {
    std::array<const field::Vector*, 1> v{&f};
    toVtu({}, v, {}, file_name, mode);
}
inline void toVtu(const field::Tensor& f, const std::string& file_name, WriteMode mode = WriteMode::RawBinaryCompressed) {
    std::array<const field::Tensor*, 1> t{&f};
    toVtu({}, {}, t, file_name, mode);
}

// The point values toVtu writes: [n_vertices] / [n_vertices][3] / [n_vertices][9] row-major.
inline auto vertexCount(const mesh::PMesh& m) -> size_t {
    int64_t nv = 0;
    prism::detail::check(prism::detail::api().mesh_n_vertices(m.handle(), &nv));
    return (size_t)nv;
}
inline auto cellToPoint(const field::Scalar& f) -> VectorXd {
    VectorXd out(vertexCount(*f.mesh()));
    prism::detail::check(prism::detail::api().field_to_points(f.handle(), out.data()));
    return out;
}
inline auto cellToPoint(const field::Vector& f) -> VectorXd {
    VectorXd out(3 * vertexCount(*f.mesh()));
    prism::detail::check(prism::detail::api().vfield_to_points(f.handle(), out.data()));
    return out;
}
inline auto cellToPoint(const field::Tensor& f) -> VectorXd {
    VectorXd out(9 * vertexCount(*f.mesh()));
    prism::detail::check(prism::detail::api().tfield_to_points(f.handle(), out.data()));
    return out;
}

class FieldRegistry {  // export.cpp:302-336
  public:
    void add(SharedPtr<field::Scalar> f) { scalars_.push_back(std::move(f)); }
    void add(SharedPtr<field::Vector> f) { vectors_.push_back(std::move(f)); }
    void add(SharedPtr<field::Tensor> f) { tensors_.push_back(std::move(f)); }
    void exportAll(const std::string& file_name, WriteMode mode = WriteMode::RawBinaryCompressed) const {
        std::vector<const field::Scalar*> s;
        std::vector<const field::Vector*> v;
        std::vector<const field::Tensor*> t;
        for (auto& f : scalars_) s.push_back(f.get());
        for (auto& f : vectors_) v.push_back(f.get());
        for (auto& f : tensors_) t.push_back(f.get());
        toVtu(s, v, t, file_name, mode);
    }
    auto scalars() const noexcept -> const std::vector<SharedPtr<field::Scalar>>& { return scalars_; }
    auto vectors() const noexcept -> const std::vector<SharedPtr<field::Vector>>& { return vectors_; }
    auto tensors() const noexcept -> const std::vector<SharedPtr<field::Tensor>>& { return tensors_; }

  private:
    std::vector<SharedPtr<field::Scalar>> scalars_;
    std::vector<SharedPtr<field::Vector>> vectors_;
    std::vector<SharedPtr<field::Tensor>> tensors_;
};

class TimeWriter {  // export.cpp:338-409
  public:
    TimeWriter(std::string prefix = "result", size_t write_interval = 1,
               WriteMode mode = WriteMode::RawBinaryCompressed, bool write_pvd = true)
        : prefix_(std::move(prefix)), interval_(write_interval), mode_(mode), write_pvd_(write_pvd) {
        if (write_interval == 0)
            throw std::invalid_argument(
                "prism::output::TimeWriter(prefix, write_interval, ...): write_interval must be positive, got 0");
    }
    void add(SharedPtr<field::Scalar> f) { registry_.add(std::move(f)); }
    void add(SharedPtr<field::Vector> f) { registry_.add(std::move(f)); }
    void add(SharedPtr<field::Tensor> f) { registry_.add(std::move(f)); }
    void advance(f64 dt = 0.0) {
        ++step_;
        time_ += dt;
    }
    void write() {
        if (step_ % interval_ != 0) return;
        if (registry_.scalars().empty() && registry_.vectors().empty() && registry_.tensors().empty()) return;
        char num[32];
        std::snprintf(num, sizeof(num), "%04zu", step_);
        auto file_name = prefix_ + "_" + num + ".vtu";
        registry_.exportAll(file_name, mode_);
        if (write_pvd_) {
            entries_.emplace_back(time_, file_name);
            std::vector<f64> t;
            std::vector<const char*> names;
            for (auto& [tt, n] : entries_) {
                t.push_back(tt);
                names.push_back(n.c_str());
            }
            prism::detail::check(prism::detail::api().write_pvd((prefix_ + ".pvd").c_str(), (int32_t)t.size(), t.data(),
                                                                names.data()));
        }
    }
    void writeAndAdvance(f64 dt = 0.0) {
        advance(dt);
        write();
    }
    auto currentTime() const noexcept -> f64 { return time_; }
    auto currentStep() const noexcept -> size_t { return step_; }

  private:
    FieldRegistry registry_;
    size_t step_{0};
    f64 time_{0.0};
    std::string prefix_;
    size_t interval_;
    WriteMode mode_;
    bool write_pvd_;
    std::vector<std::pair<f64, std::string>> entries_;
};

}  // namespace output

}  // namespace prism
