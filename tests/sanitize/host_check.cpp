// host_check.cpp -- runs the host C++ of the library (hex block meshes,
// polyMesh renumbering) and every routine of the CPU oracle under
// AddressSanitizer and UndefinedBehaviorSanitizer (built by the Makefile beside
// this file; tests/test_sanitizers.py runs it).  Test infrastructure only.
//
// For each mesh -- single cell, a row, a 2x2x2 block, an anisotropic jittered
// block, jittered and randomly numbered cubes -- it generates the polyMesh
// (fv_hexmesh_sizes / fv_hexmesh_generate), renumbers it both ways
// (fv_polymesh_renumber: reverse Cuthill-McKee, Z-order), computes the oracle
// geometry, sparsity, every explicit operator in FP64 and FP32, the multi-core
// operators, every assembly routine, the residual, both CGs, the coupling
// helpers and the Courant number, and checks a few invariants so that the code
// paths really ran (sum V = box volume, closed cells, CG converged).  The
// error paths of the host entry points are exercised with bad arguments.
#include "../../oracle/fv_oracle.cpp"

#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../include/fvhip.h"

// the library's error plumbing (defined in runtime.hip, which needs the HIP
// runtime): the host sources only set and clear the thread-local message
namespace fvhip {
static thread_local std::string g_msg;
void set_error(const std::string& m) { g_msg = m; }
void clear_error() { g_msg.clear(); }
}  // namespace fvhip

namespace {

int g_fail = 0;

void expect(bool ok, const char* what, const std::string& ctx) {
    if (!ok) {
        std::fprintf(stderr, "CHECK FAILED: %s (%s)\n", what, ctx.c_str());
        ++g_fail;
    }
}

double urand(uint64_t& s) {  // splitmix64 in [0, 1)
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
}

struct Poly {
    int nP = 0, nC = 0, nI = 0, nB = 0, nPatch = 0, nFP = 0;
    std::vector<double> pts;
    std::vector<int32_t> foff, fpts, own, nbr, pstart, psize, pempty;
    fv_polymesh_desc desc() const {
        fv_polymesh_desc d{};
        d.n_points = nP;
        d.n_faces = nI + nB;
        d.n_internal_faces = nI;
        d.n_cells = nC;
        d.n_patches = nPatch;
        d.h_points = pts.data();
        d.h_face_offsets = foff.data();
        d.h_face_points = fpts.data();
        d.h_owner = own.data();
        d.h_neighbour = nbr.data();
        d.h_patch_start = pstart.data();
        d.h_patch_size = psize.data();
        d.h_patch_empty = pempty.data();
        return d;
    }
};

Poly hex(int nx, int ny, int nz, double jitter, uint64_t jseed, uint64_t pseed) {
    int64_t s6[6];
    expect(fv_hexmesh_sizes(nx, ny, nz, s6) == FV_OK, "fv_hexmesh_sizes", "");
    Poly p;
    p.nP = (int)s6[0];
    p.nC = (int)s6[1];
    p.nI = (int)s6[2];
    p.nB = (int)s6[3];
    p.nPatch = (int)s6[4];
    p.nFP = (int)s6[5];
    p.pts.resize(3 * (size_t)p.nP);
    p.foff.resize(p.nI + p.nB + 1);
    p.fpts.resize(p.nFP);
    p.own.resize(p.nI + p.nB);
    p.nbr.resize(std::max(p.nI, 1));  // the entry point wants a buffer even for no face
    p.pstart.resize(p.nPatch);
    p.psize.resize(p.nPatch);
    p.pempty.assign(p.nPatch, 0);
    const int rc = fv_hexmesh_generate(nx, ny, nz, 1.0, 1.0, 1.0, jitter, jseed, pseed,
                                       p.pts.data(), p.foff.data(), p.fpts.data(), p.own.data(),
                                       p.nbr.data(), p.pstart.data(), p.psize.data());
    expect(rc == FV_OK, "fv_hexmesh_generate", fvhip::g_msg);
    return p;
}

void run_oracle(const Poly& p, const std::string& name) {
    const int nF = p.nI + p.nB, N = p.nC, F = p.nI, B = p.nB, P = p.nPatch;
    std::vector<double> Cf(3 * (size_t)nF), Sf(3 * (size_t)nF), C(3 * (size_t)N), V(N);
    orc_face_geometry(nF, p.foff.data(), p.fpts.data(), p.pts.data(), Cf.data(), Sf.data());
    orc_cell_geometry(N, nF, F, p.own.data(), p.nbr.data(), Cf.data(), Sf.data(), C.data(),
                      V.data());
    double vol = 0.0;
    for (double v : V) vol += v;
    expect(std::fabs(vol - 1.0) < 1e-12, "sum V = 1", name);
    std::vector<double> magSf(F), w(F), dc(F), nodc(F);
    orc_internal_coeffs(F, p.own.data(), p.nbr.data(), C.data(), Cf.data(), Sf.data(),
                        magSf.data(), w.data(), dc.data(), nodc.data());
    std::vector<int32_t> fc(B), poff(P + 1);
    std::vector<double> Sfb(3 * (size_t)B), Cfb(3 * (size_t)B);
    for (int b = 0; b < B; ++b) {
        fc[b] = p.own[F + b];
        for (int k = 0; k < 3; ++k) {
            Sfb[3 * b + k] = Sf[3 * (size_t)(F + b) + k];
            Cfb[3 * b + k] = Cf[3 * (size_t)(F + b) + k];
        }
    }
    for (int i = 0; i < P; ++i) poff[i + 1] = poff[i] + p.psize[i];
    std::vector<double> magSfb(B), nfb(3 * (size_t)B), deltab(3 * (size_t)B), wb(B), dcb(B),
        nodcb(B);
    orc_boundary_coeffs(B, fc.data(), C.data(), Cfb.data(), Sfb.data(), magSfb.data(),
                        nfb.data(), deltab.data(), wb.data(), dcb.data(), nodcb.data());
    // closed cells: sum of outward Sf per cell vanishes
    std::vector<double> cl(3 * (size_t)N, 0.0);
    for (int f = 0; f < F; ++f)
        for (int k = 0; k < 3; ++k) {
            cl[3 * (size_t)p.own[f] + k] += Sf[3 * (size_t)f + k];
            cl[3 * (size_t)p.nbr[f] + k] -= Sf[3 * (size_t)f + k];
        }
    for (int b = 0; b < B; ++b)
        for (int k = 0; k < 3; ++k) cl[3 * (size_t)fc[b] + k] += Sfb[3 * (size_t)b + k];
    double mx = 0.0;
    for (double v : cl) mx = std::max(mx, std::fabs(v));
    expect(mx < 1e-13, "closed cells", name);

    // fields
    uint64_t rs = 12345;
    std::vector<double> T(N), Tb(B), Tg(B), phi(F + B), gam(F + B), Tf(F + B), out3(3 * (size_t)N),
        out(N), sn(F + B);
    for (auto& v : T) v = urand(rs) * 2 - 1;
    for (auto& v : Tb) v = urand(rs) * 2 - 1;
    for (auto& v : Tg) v = urand(rs) * 2 - 1;
    for (auto& v : phi) v = urand(rs) * 2 - 1;
    for (auto& v : gam) v = 0.5 + 1.5 * urand(rs);
    std::vector<int32_t> kinds(P);
    for (int i = 0; i < P; ++i) kinds[i] = i % 3;
    for (int s = 0; s < 2; ++s) orc_interp_f64(s, F, B, p.own.data(), p.nbr.data(), w.data(),
                                               phi.data(), T.data(), Tb.data(), Tf.data());
    orc_grad_f64(N, F, B, p.own.data(), p.nbr.data(), fc.data(), w.data(), Sf.data(), Sfb.data(),
                 V.data(), T.data(), Tb.data(), out3.data());
    for (int s = 0; s < 2; ++s)
        orc_div_f64(s, N, F, B, p.own.data(), p.nbr.data(), fc.data(), w.data(), V.data(),
                    phi.data(), T.data(), Tb.data(), out.data());
    orc_sngrad_f64(F, B, P, poff.data(), kinds.data(), p.own.data(), p.nbr.data(), fc.data(),
                   nodc.data(), dcb.data(), T.data(), Tb.data(), Tg.data(), sn.data());
    orc_laplacian_f64(N, F, B, P, poff.data(), kinds.data(), p.own.data(), p.nbr.data(),
                      fc.data(), magSf.data(), nodc.data(), magSfb.data(), dcb.data(), V.data(),
                      gam.data(), T.data(), Tb.data(), Tg.data(), out.data());
    // FP32 forms on converted arrays
    auto f32 = [](const std::vector<double>& a) { return std::vector<float>(a.begin(), a.end()); };
    auto w32 = f32(w), Sf32 = f32(Sf), Sfb32 = f32(Sfb), V32 = f32(V), T32 = f32(T), Tb32 = f32(Tb),
         Tg32 = f32(Tg), phi32 = f32(phi), gam32 = f32(gam), magSf32 = f32(magSf),
         nodc32 = f32(nodc), magSfb32 = f32(magSfb), dcb32 = f32(dcb);
    std::vector<float> o32(3 * (size_t)N), f32o(F + B);
    orc_interp_f32(0, F, B, p.own.data(), p.nbr.data(), w32.data(), phi32.data(), T32.data(),
                   Tb32.data(), f32o.data());
    orc_grad_f32(N, F, B, p.own.data(), p.nbr.data(), fc.data(), w32.data(), Sf32.data(),
                 Sfb32.data(), V32.data(), T32.data(), Tb32.data(), o32.data());
    orc_div_f32(1, N, F, B, p.own.data(), p.nbr.data(), fc.data(), w32.data(), V32.data(),
                phi32.data(), T32.data(), Tb32.data(), o32.data());
    orc_sngrad_f32(F, B, P, poff.data(), kinds.data(), p.own.data(), p.nbr.data(), fc.data(),
                   nodc32.data(), dcb32.data(), T32.data(), Tb32.data(), Tg32.data(), f32o.data());
    orc_laplacian_f32(N, F, B, P, poff.data(), kinds.data(), p.own.data(), p.nbr.data(),
                      fc.data(), magSf32.data(), nodc32.data(), magSfb32.data(), dcb32.data(),
                      V32.data(), gam32.data(), T32.data(), Tb32.data(), Tg32.data(), o32.data());
    // multi-core forms
    void* st = orc_omp_stencil(N, F, B, p.own.data(), p.nbr.data(), fc.data());
    orc_omp_interp(F, B, p.own.data(), p.nbr.data(), w.data(), T.data(), Tb.data(), Tf.data());
    orc_omp_grad(st, p.own.data(), p.nbr.data(), w.data(), Sf.data(), Sfb.data(), V.data(),
                 T.data(), Tb.data(), out3.data());
    orc_omp_div(st, 0, p.own.data(), p.nbr.data(), w.data(), V.data(), phi.data(), T.data(),
                Tb.data(), out.data());
    orc_omp_laplacian(st, P, poff.data(), kinds.data(), p.own.data(), p.nbr.data(),
                      magSf.data(), nodc.data(), magSfb.data(), dcb.data(), V.data(), gam.data(),
                      T.data(), Tb.data(), out.data());
    orc_omp_sngrad(F, B, P, poff.data(), kinds.data(), p.own.data(), p.nbr.data(), fc.data(),
                   nodc.data(), dcb.data(), T.data(), Tb.data(), sn.data());
    orc_omp_free(st);
    // sparsity, assembly, residual, solvers
    std::vector<int32_t> ro(N + 1), ci(N + 2 * (size_t)F), dg(N), oo(F), no(F);
    orc_sparsity(N, F, p.own.data(), p.nbr.data(), ro.data(), ci.data(), dg.data(), oo.data(),
                 no.data());
    std::vector<double> vals(ci.size(), 0.0), rhs(N, 0.0), bm(B), br(B), res(N);
    std::vector<int32_t> fv(P, BC_FIXED_VALUE);
    orc_assemble_ddt(N, ro.data(), dg.data(), V.data(), T.data(), 1e-2, 1.0, vals.data(),
                     rhs.data());
    orc_assemble_laplacian(N, F, B, P, poff.data(), fv.data(), p.own.data(), p.nbr.data(),
                           fc.data(), magSf.data(), nodc.data(), magSfb.data(), dcb.data(),
                           gam.data(), Tb.data(), Tg.data(), ro.data(), dg.data(), oo.data(),
                           no.data(), -1.0, vals.data(), rhs.data());
    orc_laplacian_bcoeffs(F, B, P, poff.data(), fv.data(), magSfb.data(), dcb.data(), gam.data(),
                          Tb.data(), Tg.data(), -1.0, bm.data(), br.data());
    orc_assemble_div(1, N, F, B, P, poff.data(), kinds.data(), p.own.data(), p.nbr.data(),
                     fc.data(), w.data(), dcb.data(), phi.data(), Tb.data(), Tg.data(), ro.data(),
                     dg.data(), oo.data(), no.data(), 0.1, vals.data(), rhs.data());
    orc_assemble_source(N, ro.data(), dg.data(), V.data(), T.data(), 0.01, vals.data());
    orc_set_reference(ro.data(), dg.data(), 0, 0.5, vals.data(), rhs.data());
    orc_residual(N, ro.data(), ci.data(), vals.data(), rhs.data(), T.data(), res.data());
    // a symmetric positive definite system for the CGs: ddt - laplacian, fixedValue
    std::vector<double> sv(ci.size(), 0.0), sr(N, 0.0), x(N, 0.0), x2(N, 0.0), stats(3);
    orc_assemble_ddt(N, ro.data(), dg.data(), V.data(), T.data(), 1e-2, 1.0, sv.data(), sr.data());
    orc_assemble_laplacian(N, F, B, P, poff.data(), fv.data(), p.own.data(), p.nbr.data(),
                           fc.data(), magSf.data(), nodc.data(), magSfb.data(), dcb.data(),
                           gam.data(), Tb.data(), nullptr, ro.data(), dg.data(), oo.data(),
                           no.data(), -1.0, sv.data(), sr.data());
    expect(orc_cg(N, ro.data(), ci.data(), sv.data(), sr.data(), x.data(), 1e-10, 5000,
                  stats.data()) == 0,
           "serial CG converged", name);
    expect(orc_omp_cg(N, ro.data(), ci.data(), sv.data(), sr.data(), x2.data(), 1e-10, 5000,
                      stats.data()) == 0,
           "multi-core CG converged", name);
    // coupling helpers on the (asymmetric) momentum-like matrix
    std::vector<double> rau(N), raub(B), b3(3 * (size_t)N), U(3 * (size_t)N), h(3 * (size_t)N),
        hb(3 * (size_t)B), Ub(3 * (size_t)B), phi2(F + B), gp(3 * (size_t)N);
    for (auto& v : b3) v = urand(rs);
    for (auto& v : U) v = urand(rs);
    for (auto& v : Ub) v = urand(rs);
    orc_rau(N, B, ro.data(), dg.data(), vals.data(), V.data(), fc.data(), rau.data(), raub.data());
    orc_hbya(N, F, B, p.own.data(), p.nbr.data(), ro.data(), oo.data(), no.data(), vals.data(),
             b3.data(), U.data(), rau.data(), V.data(), fc.data(), h.data(), hb.data());
    orc_constrain_hbya(B, P, poff.data(), kinds.data(), Ub.data(), hb.data());
    orc_update_face_velocity(F, B, p.own.data(), p.nbr.data(), ro.data(), oo.data(), no.data(),
                             sv.data(), fc.data(), bm.data(), br.data(), phi.data(), x.data(),
                             phi2.data());
    orc_update_velocity(N, h.data(), rau.data(), out3.data(), gp.data());
    const double co = orc_courant(N, F, B, p.own.data(), p.nbr.data(), fc.data(), phi.data(),
                                  V.data(), 1e-3);
    expect(co > 0.0, "Courant number positive", name);
}

void run_mesh(int nx, int ny, int nz, double jitter, uint64_t seed) {
    const std::string name = std::to_string(nx) + "x" + std::to_string(ny) + "x" +
                             std::to_string(nz) + " jitter " + std::to_string(jitter);
    Poly p = hex(nx, ny, nz, jitter, seed, seed);
    run_oracle(p, name);
    for (int method = 0; method < 2; ++method) {
        const fv_polymesh_desc d = p.desc();
        Poly r = p;
        std::vector<int32_t> c2o(p.nC), f2o(p.nI + p.nB);
        const int rc = fv_polymesh_renumber(&d, method, r.foff.data(), r.fpts.data(),
                                            r.own.data(), r.nbr.data(), c2o.data(), f2o.data());
        expect(rc == FV_OK, "fv_polymesh_renumber", name);
        std::vector<char> seen(p.nC, 0);
        for (int32_t c : c2o) seen[c] = 1;
        for (char s : seen) expect(s == 1, "renumbering is a permutation", name);
        for (int f = 0; f < r.nI; ++f)
            expect(r.own[f] < r.nbr[f], "upper-triangular faces after renumbering", name);
        run_oracle(r, name + (method ? " z-order" : " rcm"));
    }
}

}  // namespace

int main() {
    run_mesh(1, 1, 1, 0.0, 0);
    run_mesh(3, 1, 1, 0.0, 0);
    run_mesh(2, 2, 2, 0.0, 7);
    run_mesh(7, 5, 3, 0.3, 5);
    run_mesh(12, 12, 12, 0.2, 1002);
    run_mesh(17, 9, 11, 0.25, 99);
    // error paths of the host entry points
    int64_t s6[6];
    expect(fv_hexmesh_sizes(0, 1, 1, s6) == FV_EINVAL, "zero-size mesh rejected", "");
    Poly p = hex(3, 3, 3, 0.0, 0, 0);
    fv_polymesh_desc d = p.desc();
    std::vector<int32_t> a(p.foff.size()), b(p.fpts.size()), o(p.own.size()), n(p.nbr.size()),
        c(p.nC), f(p.nI + p.nB);
    expect(fv_polymesh_renumber(&d, 7, a.data(), b.data(), o.data(), n.data(), c.data(),
                                f.data()) == FV_EINVAL,
           "unknown renumbering method rejected", "");
    expect(fv_hexmesh_generate(3, 3, 3, 1, 1, 1, 0.6, 0, 0, p.pts.data(), p.foff.data(),
                               p.fpts.data(), p.own.data(), p.nbr.data(), p.pstart.data(),
                               p.psize.data()) == FV_EINVAL,
           "jitter >= 0.5 rejected", "");
    std::printf("host_check: %s\n", g_fail ? "FAILED" : "ok");
    return g_fail ? 1 : 0;
}
