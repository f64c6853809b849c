// facade_demo.cpp -- a C++ caller of the finite-volume path written against the
// header-only facade (include/fvhip.hpp), the way FoamAdapter's C++ calls NeoN:
// the explicit operators of test/test_operators.cpp:52-131, the heat-transfer
// time loop of examples/heatTransfer/heatTransfer.cpp:60-86 through
// dsl::solve, a pressure equation through PDESolver with setReference
// (include/FoamAdapter/datastructures/expression.hpp:23-179), and the
// neoIcoFoam time loop (examples/neoIcoFoam/neoIcoFoam.cpp:80-180) on a
// lid-driven cavity through PDESolver<Vec3> and the coupling helpers.  Test
// infrastructure: tests/test_gpu_cpp_facade.py runs it on the GPU and compares
// what it writes (raw FP64 arrays in the output directory) with the oracle.
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/fvhip.hpp"

namespace dsl = fvhip::dsl;
namespace fvcc = fvhip::fvcc;
namespace la = fvhip::la;
using fvhip::Vector;

static std::string g_dir;

static void write(const char* name, const Vector<double>& v) {
    const std::vector<double> h = v.copyToHost();
    FILE* f = std::fopen((g_dir + "/" + name + ".bin").c_str(), "wb");
    if (!f || std::fwrite(h.data(), sizeof(double), h.size(), f) != h.size()) {
        std::fprintf(stderr, "cannot write %s\n", name);
        std::exit(2);
    }
    std::fclose(f);
}

static void uniform(Vector<double>& v, double lo, double hi, uint64_t seed, uint64_t stream) {
    fvhip::check(fv_fill_uniform(v.exec().get(), FV_F64, v.data(), (int64_t)v.size(), lo, hi,
                                 seed, stream),
                 "fv_fill_uniform");
}

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: facade_demo <output directory>\n");
        return 2;
    }
    g_dir = argv[1];
    try {
        fvhip::Executor exec(0);
        // 12^3 hex block, interior vertices jittered by 0.2 h, cells numbered at random
        fvhip::HexBlock block(12, 12, 12, 0.2, 21, 22);
        auto mesh = fvhip::UnstructuredMesh::fromPolyMesh(exec, block.desc());
        const size_t N = mesh.nCells();
        std::printf("mesh: %zu cells, %d internal faces, %d boundary faces\n", N,
                    mesh.nInternalFaces(), mesh.nBoundaryFaces());

        // ---- explicit operators (test/test_operators.cpp) ----
        std::vector<int32_t> kinds;
        for (int p = 0; p < mesh.nPatches(); ++p) kinds.push_back(p % 3);  // fixedValue, zeroGradient, calculated
        auto bcs = std::make_shared<fvcc::BoundaryConditions>(mesh, kinds);
        fvcc::VolumeField T(mesh, "T", bcs);
        uniform(T.internalVector(), -1.0, 1.0, 23, 0);
        uniform(T.boundaryValues(), -1.0, 1.0, 23, 1);
        fvcc::SurfaceField phi(mesh, "phi"), gamma(mesh, "gamma"), Tf(mesh, "Tf"), sn(mesh, "sn");
        uniform(phi.internalVector(), -1.0, 1.0, 23, 2);
        uniform(gamma.internalVector(), 0.5, 2.0, 23, 3);
        Vector<double> gradT(exec, 3 * N), out(exec, N);
        fvcc::SurfaceInterpolation(exec, mesh, {"linear"}).interpolate(T, Tf);
        write("interp_linear", Tf.internalVector());
        fvcc::SurfaceInterpolation(exec, mesh, {"upwind"}).interpolate(T, Tf, &phi);
        write("interp_upwind", Tf.internalVector());
        fvcc::GaussGreenGrad(exec, mesh).grad(T, dsl::Coeff(), gradT);
        write("grad", gradT);
        fvcc::GaussGreenDiv(exec, mesh, {"Gauss", "linear"}).div(out, phi, T, dsl::Coeff(1.0));
        write("div_linear", out);
        fvcc::GaussGreenDiv(exec, mesh, {"Gauss", "upwind"}).div(out, phi, T, dsl::Coeff(1.0));
        write("div_upwind", out);
        fvcc::GaussGreenLaplacian(exec, mesh, {"Gauss", "linear", "uncorrected"})
            .laplacian(out, gamma, T, dsl::Coeff(1.0));
        write("laplacian", out);
        fvcc::FaceNormalGradient(exec, mesh, {"uncorrected"}).faceNormalGrad(T, sn);
        write("sngrad", sn.internalVector());

        // ---- heat transfer (examples/heatTransfer/heatTransfer.cpp:60-86) ----
        auto fixed = std::make_shared<fvcc::BoundaryConditions>(
            mesh, std::vector<int32_t>(mesh.nPatches(), FV_BC_FIXED_VALUE));
        fvcc::VolumeField Th(mesh, "T", fixed);
        uniform(Th.internalVector(), 300.0, 400.0, 1001, 0);
        {
            const std::vector<int32_t> po = mesh.get<int32_t>(FV_G_PATCH_OFFSETS, mesh.nPatches() + 1);
            std::vector<double> tb(mesh.nBoundaryFaces(), 350.0);
            for (int b = po[0]; b < po[1]; ++b) tb[b] = 300.0;
            for (int b = po[1]; b < po[2]; ++b) tb[b] = 400.0;
            Th.boundaryValues() = Vector<double>(exec, tb);
        }
        fvcc::SurfaceField kappa(mesh, "kappa", 1e-2);
        const fvhip::Dictionary fvSchemes{{"div", "Gauss linear"}};
        const fvhip::Dictionary solverT{{"solver", "PCG"}, {"relTol", "1e-10"}, {"maxIter", "1000"}};
        const double dt = 1e-3;
        for (int step = 0; step < 10; ++step) {
            const double t = step * dt;
            dsl::Expression eqn(dsl::imp::ddt(Th) - dsl::imp::laplacian(kappa, Th));
            const la::SolverStats st = dsl::solve(eqn, Th, t, dt, fvSchemes, solverT);
            std::printf("heat step %d: initial %.3e final %.3e iterations %d\n", step,
                        st.initResNorm, st.finalResNorm, st.numIter);
            if (!(st.finalResNorm <= 1e-10)) return 3;
        }
        write("heat_T", Th.internalVector());

        // ---- pressure equation through PDESolver with a reference value ----
        auto zg = std::make_shared<fvcc::BoundaryConditions>(
            mesh, std::vector<int32_t>(mesh.nPatches(), FV_BC_FIXED_GRADIENT));
        fvcc::VolumeField p(mesh, "p", zg);
        fvcc::SurfaceField one(mesh, "one", 1.0);
        fvhip::PDESolver pEqn(dsl::imp::laplacian(one, p) - dsl::exp::div(phi), p, 0.0, 1.0, {},
                              {{"solver", "PCG"}, {"relTol", "1e-10"}, {"maxIter", "5000"}});
        pEqn.setReference(0, 0.0);
        const la::SolverStats ps = pEqn.solve();
        std::printf("pressure: initial %.3e final %.3e iterations %d\n", ps.initResNorm,
                    ps.finalResNorm, ps.numIter);
        write("poisson_p", p.internalVector());

        // ---- neoIcoFoam (examples/neoIcoFoam/neoIcoFoam.cpp:80-180) on a lid-driven
        // cavity: 8^3 jittered hex block, lid (patch y1) U = (1, 0, 0), no-slip walls,
        // p zeroGradient with reference cell 0, nu = 0.01, dt = 0.05, upwind
        // convection, momentum predictor and two pressure correctors, four steps ----
        {
            fvhip::HexBlock cb(8, 8, 8, 0.15, 91, 92);
            auto cm = fvhip::UnstructuredMesh::fromPolyMesh(exec, cb.desc());
            const size_t Bc = cm.nBoundaryFaces();
            auto ubc = std::make_shared<fvcc::BoundaryConditions>(
                cm, std::vector<int32_t>(cm.nPatches(), FV_BC_FIXED_VALUE));
            auto pbc = std::make_shared<fvcc::BoundaryConditions>(
                cm, std::vector<int32_t>(cm.nPatches(), FV_BC_FIXED_GRADIENT));
            fvcc::VolumeVectorField U(cm, "U", ubc);
            fvcc::VolumeField pc(cm, "p", pbc);
            {
                const std::vector<int32_t> po = cm.get<int32_t>(FV_G_PATCH_OFFSETS, cm.nPatches() + 1);
                std::vector<double> ub(3 * Bc, 0.0);
                for (int b = po[3]; b < po[4]; ++b) ub[3 * (size_t)b] = 1.0;
                U.boundaryValues() = Vector<double>(exec, ub);
            }
            fvcc::SurfaceField phiC = fvhip::flux(U);
            fvcc::SurfaceField nu(cm, "nu", 0.01);
            const double dtc = 0.05;
            const fvhip::Dictionary schemes{{"div", "Gauss upwind"}};
            const fvhip::Dictionary solU{{"solver", "PBiCGStab"}, {"relTol", "1e-13"}, {"maxIter", "1000"}};
            const fvhip::Dictionary solP{{"solver", "PCG"}, {"relTol", "1e-13"}, {"maxIter", "2000"}};
            for (int step = 0; step < 4; ++step) {
                U.oldTime().copyFrom(U.internalVector());
                const double co = fvhip::computeCoNum(phiC, dtc);
                fvhip::PDESolver<fvhip::Vec3> UEqn(
                    dsl::imp::ddt(U) + dsl::imp::div(phiC, U) - dsl::imp::laplacian(nu, U), U,
                    (step + 1) * dtc, dtc, schemes, solU);
                UEqn.solve();
                int pIters = 0;
                for (int corr = 0; corr < 2; ++corr) {
                    auto [rAU, HbyA] = fvhip::computeRAUandHByA(UEqn);
                    fvhip::constrainHbyA(U, pc, HbyA);
                    fvcc::SurfaceField rAUf(cm, "rAUf");
                    fvcc::SurfaceInterpolation(exec, cm, {"linear"}).interpolate(rAU, rAUf);
                    fvcc::SurfaceField phiHbyA = fvhip::flux(HbyA);
                    fvhip::PDESolver pEqn(dsl::imp::laplacian(rAUf, pc) - dsl::exp::div(phiHbyA), pc,
                                          (step + 1) * dtc, dtc, schemes, solP);
                    pEqn.setReference(0, 0.0);
                    pIters += pEqn.solve().numIter;   // corrects p's boundary values
                    fvhip::updateFaceVelocity(phiHbyA, pEqn, phiC);
                    fvhip::updateVelocity(HbyA, rAU, pc, U);
                    U.correctBoundaryConditions();
                }
                std::printf("icoFoam step %d: Co %.3f, pressure iterations %d\n", step, co, pIters);
            }
            write("ico_U", U.internalVector());
            write("ico_p", pc.internalVector());
            write("ico_phi", phiC.internalVector());
        }
        exec.fence();
    } catch (const fvhip::Error& e) {
        std::fprintf(stderr, "fvhip error %d: %s\n", e.code, e.what());
        return 1;
    }
    std::printf("facade_demo ok\n");
    return 0;
}
