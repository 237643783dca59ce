// oracle_core.h -- FP64 CPU restatement of the reference's finite-volume path.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP library is checked against; it is
// linked only by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
// It is never part of the product path.
//
// Every routine restates the reference (eigemx/prism, C++20 + Eigen) line by line,
// including the order of floating-point operations where the reference fixes it, so that
// a branchy decision (NVF limiter intervals, max(m,0)) takes the same side on both paths.
// File:line citations are relative to the reference repository root.
#pragma once

#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace oref {

using f64 = double;
constexpr f64 EPS = 2.220446049250313080847e-16;  // constants.h:12  numeric_limits::epsilon
constexpr f64 PI = 3.141592653589793238462643383279502884;

// -------------------------------------------------------------------- exceptions ----
struct BCError : std::runtime_error { using std::runtime_error::runtime_error; };
struct MeshError : std::runtime_error { using std::runtime_error::runtime_error; };

// ------------------------------------------------------------- small vector types ----
// Association follows Eigen's fixed-size reductions: dot = (x0y0 + x1y1) + x2y2.
struct V3 {
    f64 x{0}, y{0}, z{0};
    f64& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    f64 operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(const V3& a, const V3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(const V3& a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(const V3& a, f64 s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(f64 s, const V3& a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(const V3& a, f64 s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, const V3& b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline V3& operator-=(V3& a, const V3& b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
inline f64 dot(const V3& a, const V3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline f64 sqnorm(const V3& a) { return dot(a, a); }
inline f64 norm(const V3& a) { return std::sqrt(sqnorm(a)); }
inline V3 cross(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// 3x3, m[r][c]
struct M3 {
    f64 m[3][3]{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    static M3 identity() { M3 r; r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.0; return r; }
};
inline V3 mul(const M3& a, const V3& v) {
    V3 r;
    for (int i = 0; i < 3; ++i) r[i] = (a.m[i][0] * v.x + a.m[i][1] * v.y) + a.m[i][2] * v.z;
    return r;
}
inline V3 mulT(const M3& a, const V3& v) {  // a^T v
    V3 r;
    for (int i = 0; i < 3; ++i) r[i] = (a.m[0][i] * v.x + a.m[1][i] * v.y) + a.m[2][i] * v.z;
    return r;
}
inline M3 scale(f64 s, const M3& a) {
    M3 r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = s * a.m[i][j];
    return r;
}
inline M3 add(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[i][j] + b.m[i][j];
    return r;
}
M3 inverse3(const M3& a);               // cofactor inverse
M3 matmul(const M3& a, const M3& b);    // row . column, (a0b0 + a1b1) + a2b2
M3 transpose(const M3& a);

// ---------------------------------------------------------------------------- mesh ----
struct BC {
    std::string kind;
    int vk{0};     // 0 nil, 1 scalar, 2 vector
    f64 s{0};
    V3 v{};
};

struct Patch {
    std::string name;
    std::map<std::string, BC> bcs;  // std::map: alphabetical, as mesh/boundary.cpp:391
    std::vector<int64_t> faces;
    bool empty{false};
    const BC& bc(const std::string& field) const;          // mesh/boundary.cpp:426-444
    f64 scalarBC(const std::string& field) const;          // :446-470 (component fallback)
    V3 vectorBC(const std::string& field) const;           // :473-492
};

struct Face {
    std::vector<int64_t> verts;
    f64 area{0};
    V3 normal{}, center{};
    int64_t owner{-1}, neighbor{-1};
    int32_t patch{-1};
    V3 S() const { return area * normal; }   // face.h:18  _area * _normal
    bool interior() const { return neighbor >= 0; }
};

struct Cell {
    std::vector<int64_t> faces;
    std::vector<int64_t> verts;   // Cell::verticesIds() (cell.h:25)
    f64 volume{0};
    V3 center{};
};

struct FieldInfo {
    std::string name, type;
    std::optional<std::string> grad;
};

struct Mesh {
    std::vector<V3> vertices;
    std::vector<Face> faces;
    std::vector<Cell> cells;
    std::vector<Patch> patches;
    std::vector<FieldInfo> infos;
    std::vector<int64_t> old_to_new;
    int64_t n_boundary{0}, n_nonempty{0};
    std::vector<f64> vol;
    bool has_cell_vtx{false};   // the cells' own vertex lists were given

    int64_t nc() const { return (int64_t)cells.size(); }
    int64_t nf() const { return (int64_t)faces.size(); }
    const Patch& patchOf(const Face& f) const { return patches[f.patch]; }
    int64_t other(int64_t c, const Face& f) const { return f.owner == c ? f.neighbor : f.owner; }
};

// output::cellToPoint (export.cpp:45-112): inverse-distance-weighted point values,
// interleaved [n_vertices][ncomp], of interleaved cell data [n_cells][ncomp].
std::vector<f64> cellToPoint(const Mesh& m, const std::vector<f64>& cell_data, size_t ncomp);
void faceGeometry(const std::vector<V3>& pts, Face& f);                   // face.cpp:5-41
void cellGeometry(const std::vector<Face>& faces, Cell& c);               // cell.cpp:7-41
V3 outwardS(const Face& f, int64_t cell);                                 // utilities.cpp:15
f64 geometricWeight(const Mesh& m, int64_t c, int64_t n, const Face& f);  // utilities.cpp:28

// ------------------------------------------------------------------------ fields ----
struct Scalar;

struct Gradient {  // gradient.h:12-43; cache of cell gradients keyed on the field event no.
    virtual ~Gradient() = default;
    std::vector<V3> g;
    size_t computed{SIZE_MAX};
    V3 atCell(const Scalar& f, int64_t c);
    V3 atFace(const Scalar& f, const Face& face, int64_t face_id);   // gradient.cpp:15-34
    V3 atBoundaryFace(const Scalar& f, const Face& face, int64_t fid);
    virtual void computeAll(const Scalar& f) = 0;
};

struct GreenGauss : Gradient {   // green_gauss.cpp
    explicit GreenGauss(const Mesh& m) { g.assign(m.nc(), V3{}); }
    void computeAll(const Scalar& f) override;
};

struct LeastSquares : Gradient {  // least_squares.cpp
    struct Contrib { bool interior; int64_t other; V3 w; };
    std::vector<M3> pinv;
    std::vector<std::vector<Contrib>> cache;
    explicit LeastSquares(const Mesh& m);
    void computeAll(const Scalar& f) override;
};

struct Scalar {
    std::string name;
    std::shared_ptr<Mesh> mesh;
    std::vector<f64> v;
    std::vector<f64> fv;        // face values (empty = none), scalar.h:112
    int kind{0};                // 0 Scalar, 1 Pressure
    int coord{-1};
    size_t event{0};
    std::shared_ptr<Gradient> grad;
    size_t hist_max{0};
    std::vector<std::vector<f64>> hist;

    Scalar(std::string n, std::shared_ptr<Mesh> m, std::vector<f64> vals, int kind_ = 0,
           int coord_ = -1);
    bool hasFaceValues() const { return !fv.empty(); }
    f64 cellValue(int64_t c) const { return v[c]; }
    f64 faceValue(int64_t fid) const;                         // scalar.cpp:189-237
    V3 gradCell(int64_t c) const { return grad->atCell(*this, c); }
    V3 gradFace(int64_t fid) const { return grad->atFace(*this, mesh->faces[fid], fid); }
    void bump() { ++event; }                                  // non-const values() access
    void updatePrev();                                        // history.cpp:30-48
    const std::vector<f64>* prev() const { return hist.empty() ? nullptr : &hist.front(); }
    const std::vector<f64>* prevPrev() const { return hist.size() < 2 ? nullptr : &hist[1]; }  // history.cpp:17-22
    bool isDirichletBC(const std::string& kind) const { return kind == "fixed" || kind == "no-slip"; }
};

struct Vector {  // GeneralVector (vector.h)
    std::string name;
    std::shared_ptr<Mesh> mesh;
    std::shared_ptr<Scalar> c[3];
    std::vector<V3> fd;      // face data (empty = none)
    std::vector<f64> ff;     // face flux data
    bool has_fd{false}, has_ff{false};
    V3 cellValue(int64_t i) const { return {c[0]->v[i], c[1]->v[i], c[2]->v[i]}; }
    V3 faceValue(int64_t fid) const;       // vector.h:187-202
    V3 boundaryValue(int64_t fid) const;   // vector_boundary.h
    f64 flux(int64_t fid) const;           // vector.h:216-257
    void initFaceData();                   // vector.h:265-280
};

struct Tensor {
    std::string name;
    std::shared_ptr<Mesh> mesh;
    std::vector<M3> v;
    M3 faceValue(int64_t fid) const;  // tensor.cpp:58-77
};

// ------------------------------------------------------------------------ schemes ----
struct CSR {   // static pattern of one mesh: sorted columns per row
    int64_t n{0};
    std::vector<int64_t> rp;
    std::vector<int32_t> col;
    std::vector<int64_t> diag;
    int64_t slot(int64_t i, int64_t j) const;
};

struct System {   // LinearSystem (linear.h): matrix values on the mesh pattern + rhs
    std::vector<f64> A;
    std::vector<f64> b;
    void resize(size_t nnz, size_t n) { A.assign(nnz, 0.0); b.assign(n, 0.0); }
};

enum class SK { Convection, Diffusion, SrcConstant, SrcGradient, SrcDivergence, SrcImplicit, BackwardEuler, AdamMoulton };

struct Scheme {
    SK kind;
    int variant{0};
    std::shared_ptr<Scalar> phi;
    std::shared_ptr<Scalar> kappa;
    std::shared_ptr<Tensor> kappa_t;
    std::shared_ptr<Vector> vec;
    int coord{0};
    f64 coeff{1.0};
    System sys;  // full schemes: A (replaced by each apply) + rhs (accumulated); sources: rhs
    bool full() const { return kind != SK::SrcConstant && kind != SK::SrcGradient && kind != SK::SrcDivergence; }
};

struct Equation {  // eqn::Transport
    std::shared_ptr<Scalar> phi;
    std::vector<Scheme> schemes;   // full schemes, ctor order
    std::vector<Scheme> sources;   // explicit sources, ctor order
    std::shared_ptr<CSR> pattern;
    System sys;
    f64 urf{1.0};
    bool symmetry{false};          // eqn::boundary::Symmetry<Transport> handler added
    void updateCoeffs();
    void applySymmetry();          // symmetry.cpp:54-82
    void zeroOut();
    void relax();
    void setUnderRelax(f64 f);
};

std::shared_ptr<CSR> patternFor(const std::shared_ptr<Mesh>& m);
void applyScheme(Scheme& s, const CSR& p);

// ------------------------------------------------------------------------- solver ----
struct SolveParams {
    int64_t n_iter{10};
    f64 eps{1e-7};
    int64_t ref_cell{-1};
    f64 ref_value{0};
    bool clear{true};
};
struct SolveResult {
    int64_t iter{0};
    f64 r0{0}, r{0};
    bool converged{false};
};
SolveResult solve(Equation& e, const SolveParams& p);
f64 scaledResidual(const CSR& p, const std::vector<f64>& A, const std::vector<f64>& x,
                   const std::vector<f64>& b);
void spmv(const CSR& p, const std::vector<f64>& A, const f64* x, f64* y);
// "exact" solve of A x = b (ILU(0)-BiCGSTAB to round-off + refinement); x holds the guess.
void exactSolve(const CSR& p, const std::vector<f64>& A, const std::vector<f64>& b,
                std::vector<f64>& x);
// Unpreconditioned Eigen-style BiCGSTAB (same iteration as the GPU solver); returns iters.
int64_t plainBicgstab(int64_t n, const int64_t* rp, const int32_t* col, const f64* val,
                      const f64* b, f64* x, f64 tol, int64_t maxit, f64* rel_res);

// ------------------------------------------------------------------ SIMPLE / PISO ----
struct SimpleControls {
    f64 momentum_urf{0.7};
    int64_t momentum_max_iter{3};
    f64 momentum_residual{1e-7};
    int64_t non_ortho{2};
    f64 pressure_urf{0.3};
    int64_t pressure_max_iter{5};
    f64 pressure_residual{1e-7};
    int64_t p_ref_cell{-1};
    f64 p_ref_value{0};
};
struct Report {
    std::string name;
    int64_t iters;
    f64 r0, r;
    bool converged;
};
// Discrete adjoint of R(T) = A T - b for Upwind + NonCorrected(scalar kappa) +
// ConstantScalar(s) (no reference counterpart): J = sum_R V T^2 / sum_R V over
// R = {x_c[axis] > thr}; A^T lambda = dJ/dT; dJ/dtheta = -lambda^T dR/dtheta.
f64 adjointObjective(const Scalar& T, int axis, f64 thr, f64* VR);
void adjointGradient(Equation& e, const Scalar* src, const Scalar* kappa, int axis, f64 thr, f64* J,
                     std::vector<f64>* dJds, std::vector<f64>* dJdk);
bool needsRef(const Scalar& pp);  // algo::needsPressureReference (SIMPLE.cpp:110-122)
std::vector<Report> simpleStep(std::vector<Equation*>& mom, std::shared_ptr<Vector>& U,
                               std::shared_ptr<Vector>& mdot, std::shared_ptr<Scalar>& p,
                               const SimpleControls& c);
std::vector<Report> primeStep(std::vector<Equation*>& mom, std::shared_ptr<Vector>& U,
                              std::shared_ptr<Vector>& mdot, std::shared_ptr<Scalar>& p,
                              const SimpleControls& c);
// solveImplicitMomentum of every predictor (SIMPLE.cpp:19-41).
std::vector<Report> momentumPredictor(std::vector<Equation*>& mom, const SimpleControls& c);
// The pressure-correction equation as solvePressureEquation hands it to solver.solve()
// (SIMPLE.cpp:124-178): D from the momentum diagonals (the predictors' coefficients are
// zeroed), Rhie-Chow face fluxes in mdot, p' with constrained face values, and the solve
// parameters (automatic reference cell resolved).
struct PressureSystem {
    std::shared_ptr<Scalar> pp;
    std::shared_ptr<Tensor> D;
    Equation pe;
    SolveParams sp;
};
PressureSystem buildPressureSystem(SimpleControls c, std::vector<Equation*>& mom, std::shared_ptr<Vector>& U,
                                   std::shared_ptr<Vector>& mdot, std::shared_ptr<Scalar>& p);
std::vector<Report> pisoStep(std::vector<Equation*>& mom, std::shared_ptr<Vector>& U,
                             std::shared_ptr<Vector>& mdot, std::shared_ptr<Scalar>& p,
                             int64_t outer, int64_t implicit_steps, int64_t corr_steps,
                             const SimpleControls& c);

}  // namespace oref
