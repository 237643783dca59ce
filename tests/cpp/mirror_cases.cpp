// The reference's own tests written against the C++ mirror (include/prism_hip.hpp), as a
// user of the reference would write them: tests/test_poisson.cpp:24-61,
// tests/test_advection_1d.cpp:24-58 and tests/test_algorithms_pipe.cpp:14-37, the examples'
// output export, and a k-epsilon run.
//
// usage: mirror_cases <compute library> <symbol prefix> <libprism_hip.so> <cases dir>
// Prints one line per case and exits non-zero when a check of the reference fails.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "prism_hip.hpp"

using namespace prism;
using namespace prism::scheme;

namespace {

auto l2NormRel(const VectorXd& x, const VectorXd& ref) -> f64 {
    f64 a = 0, b = 0;
    for (size_t i = 0; i < x.size(); ++i) {
        a += (x[i] - ref[i]) * (x[i] - ref[i]);
        b += ref[i] * ref[i];
    }
    return std::sqrt(a) / std::sqrt(b);
}

auto loadMesh(const std::filesystem::path& unv) -> SharedPtr<mesh::PMesh> {
    return mesh::UnvToPMeshConverter(unv, unv.parent_path() / "fields.json").toPMesh();
}

template <class Eqn>
void solveWithBiCGSTAB(Eqn& eqn, size_t n = 5, size_t max_iters = 15, f64 tol = 1e-20) {
    solver::BiCGSTAB solver;
    for (size_t i = 0; i < n; ++i) solver.solve(eqn, max_iters, tol);
}

auto flowRate(const field::Velocity& U, const mesh::PMesh& mesh, const std::string& patch_name) -> f64 {
    auto uf = U.faceValues();
    for (const auto& patch : mesh.boundaryPatches()) {
        if (patch.name() != patch_name) continue;
        f64 rate = 0.0;
        for (auto fid : patch.facesIds()) rate += uf[fid].dot(mesh.faceAreaVector(fid));
        return rate;
    }
    return 0.0;
}

int failures = 0;
void check(bool ok, const char* what, f64 value) {
    std::printf("%-40s %.6e %s\n", what, value, ok ? "ok" : "FAILED");
    if (!ok) ++failures;
}

void poisson(const std::filesystem::path& dir) {
    auto mesh = loadMesh(dir / "poisson" / "mesh_hex.unv");
    auto P = makeShared<field::Scalar>("P", mesh, 0.0);
    VectorXd src(mesh->cellCount()), exact(mesh->cellCount());
    for (size_t c = 0; c < mesh->cellCount(); ++c) {
        auto x = mesh->cellCenter(c);
        src[c] = 2 * std::pow(PI, 2) * std::sin(PI * x.x()) * std::cos(PI * x.y());
        exact[c] = std::sin(PI * x.x()) * std::cos(PI * x.y());
    }
    auto cten = makeShared<field::Tensor>("c", mesh, Matrix3d::Identity());
    using laplacian = diffusion::Corrected;
    auto source = makeShared<field::Scalar>("S", mesh, src);
    auto eqn = eqn::Transport(laplacian(cten, P), source::ConstantScalar<Sign::Positive>(source));
    solveWithBiCGSTAB(eqn);
    check(l2NormRel(P->values(), exact) < 0.0004, "poisson structured rel L2", l2NormRel(P->values(), exact));
}

void advection(const std::filesystem::path& dir) {
    auto mesh = loadMesh(dir / "versteeg_advection_1d" / "mesh.unv");
    auto T = makeShared<field::Scalar>("T", mesh, 0.0);
    Vector3d inlet_velocity;
    for (const auto& patch : mesh->boundaryPatches())
        if (patch.name() == "Inlet") inlet_velocity = patch.getVectorBoundaryCondition("U");
    auto U = makeShared<field::Velocity>("U", mesh, inlet_velocity);
    auto kappa = makeShared<field::Scalar>("kappa", mesh, 0.1);
    using div = convection::Upwind;
    using laplacian = diffusion::NonCorrected;
    auto eqn = eqn::Transport(div(U, T), laplacian(kappa, T));
    solveWithBiCGSTAB(eqn, 5, 5, 1e-20);
    VectorXd v = T->values();
    f64 d = 0;
    bool sorted = true;
    for (size_t c = 0; c < v.size(); ++c) {
        f64 x = mesh->cellCenter(c).x(), u = inlet_velocity.x();
        f64 a = -((std::exp(u * x / 0.1) - 1) / (std::exp(u / 0.1) - 1)) + 1;
        d += (v[c] - a) * (v[c] - a);
        if (c > 0 && v[c] > v[c - 1]) sorted = false;
    }
    check(std::sqrt(d) < 0.25 && sorted, "advection 1D ||T - exact||", std::sqrt(d));
}

void simplePipe(const std::filesystem::path& dir) {
    auto mesh = loadMesh(dir / "coarsePipeHex" / "mesh.unv");
    auto U = makeShared<field::Velocity>("U", mesh, Vector3d{0, 0, 0});
    auto p = makeShared<field::Pressure>("P", mesh, 0.0);
    auto nu = makeShared<field::Scalar>("nu", mesh, 1e-3);
    auto mdot = makeShared<field::Velocity>(U->clone());
    using div = convection::LinearUpwind;
    using laplacian = diffusion::Corrected;
    using grad = source::Gradient<Sign::Negative>;
    auto u = eqn::Momentum(div(mdot, U->x()), laplacian(nu, U->x()), grad(p, VectorCoord::X));
    auto v = eqn::Momentum(div(mdot, U->y()), laplacian(nu, U->y()), grad(p, VectorCoord::Y));
    std::vector<eqn::Momentum*> eqns{&u, &v};
    algo::SIMPLEControls controls;
    for (int iter = 0; iter < 200; ++iter)
        algo::IncompressibleSIMPLE(controls).step(std::span<eqn::Momentum*>(eqns), U, mdot, p);
    auto inlet = flowRate(*U, *mesh, "Inlet");
    auto outlet = flowRate(*U, *mesh, "Outlet");
    f64 imbalance = std::abs(inlet + outlet) / std::abs(inlet);
    check(imbalance < 1e-6, "SIMPLE coarse pipe mass imbalance", imbalance);
}

// The export of the examples: toVtu of one field (examples/poissonSolver/main.cpp:86) and a
// TimeWriter writing every second step (examples/PISOSolver/main.cpp:60).
void exportFields(const std::filesystem::path& dir) {
    auto mesh = loadMesh(dir / "poisson" / "mesh_hex.unv");
    auto P = makeShared<field::Scalar>("P", mesh, 2.5);
    auto U = makeShared<field::Vector>("U", mesh, Vector3d{1, 2, 3});
    auto out = dir / "export";
    std::filesystem::create_directories(out);
    output::toVtu(*P, (out / "solution.vtu").string(), output::WriteMode::Ascii);
    output::TimeWriter writer((out / "result").string(), 2, output::WriteMode::Ascii);
    writer.add(P);
    writer.add(U);
    for (int s = 0; s < 4; ++s) writer.writeAndAdvance(0.5);
    f64 dev = 0;
    for (auto x : output::cellToPoint(*P)) dev = std::max(dev, std::abs(x - 2.5));
    namespace fs = std::filesystem;
    bool files = fs::exists(out / "solution.vtu") && fs::exists(out / "result_0002.vtu") &&
                 fs::exists(out / "result_0004.vtu") && fs::exists(out / "result.pvd") &&
                 !fs::exists(out / "result_0001.vtu") && writer.currentStep() == 4;
    check(files && dev < 1e-14, "toVtu / TimeWriter files, constant points", dev);
}

// A transient k-epsilon run (KEpsilon.cpp) on the coarse pipe: uniform inflow velocity,
// k and epsilon fixed at the inlet, three steps of both equations.
void kEpsilonPipe(const std::filesystem::path& dir) {
    auto mesh = loadMesh(dir / "kepsilon_pipe" / "mesh.unv");
    auto U = makeShared<field::Velocity>("U", mesh, Vector3d{0.01, 0.0, 0.0});
    auto nu = makeShared<field::Scalar>("nu", mesh, 1e-3);
    turbulence::IncompressibleKEpsilon model(U, U, nu, 0.05);
    model.k()->setValues(VectorXd(mesh->cellCount(), 1e-4));
    model.epsilon()->setValues(VectorXd(mesh->cellCount(), 1e-5));
    solver::BiCGSTAB solver;
    for (int step = 0; step < 3; ++step) {
        model.k()->updatePrevTimeSteps();
        model.epsilon()->updatePrevTimeSteps();
        auto [k_eqn, eps_eqn] = model.equations();
        solver.solve(k_eqn, 2, 1e-20);
        solver.solve(eps_eqn, 2, 1e-20);
    }
    f64 kmin = 1e300, emin = 1e300;
    bool finite = true;
    for (auto v : model.k()->values()) { kmin = std::min(kmin, v); finite = finite && std::isfinite(v); }
    for (auto v : model.epsilon()->values()) { emin = std::min(emin, v); finite = finite && std::isfinite(v); }
    check(finite && kmin > 0 && emin > 0, "k-epsilon pipe: min k > 0, min eps > 0", std::min(kmin, emin));
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s <compute library> <prefix> <libprism_hip.so> <cases dir>\n", argv[0]);
        return 2;
    }
    Context::init(Api::load(argv[1], argv[2], argv[3]));
    std::filesystem::path dir(argv[4]);
    poisson(dir);
    advection(dir);
    simplePipe(dir);
    exportFields(dir);
    kEpsilonPipe(dir);
    try {  // the reference's exception types cross the mirror
        auto mesh = loadMesh(dir / "poisson" / "mesh_hex.unv");
        auto T = makeShared<field::Scalar>("P", mesh, 0.0);
        scheme::temporal::BackwardEuler bad(T, -1.0);
        check(false, "BackwardEuler(dt < 0) throws", 0.0);
    } catch (const std::invalid_argument&) {
        check(true, "BackwardEuler(dt < 0) throws", 0.0);
    }
    return failures == 0 ? 0 : 1;
}
