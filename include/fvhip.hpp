/*
 * fvhip.hpp -- header-only C++ facade over the C ABI (fvhip.h) with the names of
 * NeoN's finite-volume API as FoamAdapter's C++ code calls it, so that code
 * written against NeoN's fvcc / la / dsl classes and FoamAdapter's PDESolver
 * ports by changing the namespace:
 *
 *   fvhip::Executor exec(0);                                   // createExecutor("GPU")
 *   auto mesh = fvhip::UnstructuredMesh::fromPolyMesh(exec, desc);
 *   fvcc::GaussGreenGrad(exec, mesh).grad(T, dsl::Coeff(), gradT);   // test/test_operators.cpp:72
 *   fvcc::GaussGreenDiv(exec, mesh, {"Gauss", "linear"}).div(divT, phi, T, dsl::Coeff(1.0));
 *   dsl::Expression eqn(dsl::imp::ddt(T) - dsl::imp::laplacian(kappa, T));
 *   dsl::solve(eqn, T, t, dt, fvSchemes, solverDict);          // examples/heatTransfer/heatTransfer.cpp:67-76
 *   FoamAdapter-style PDESolver: fvhip::PDESolver(eqn, p).setReference(0, 0.0); .solve();
 *   PDESolver<Vec3> UEqn(dsl::imp::ddt(U) + dsl::imp::div(phi, U) - dsl::imp::laplacian(nu, U), U);
 *   auto [rAU, HbyA] = fvhip::computeRAUandHByA(UEqn);          // neoIcoFoam.cpp:94-171
 *
 * Every call goes through libfvhip's extern "C" entry points (hand-written HIP
 * kernels); failures throw fvhip::Error carrying the status code and
 * fv_last_error().  Device storage belongs to the Executor's context (HBM
 * budget accounting included).  Scalar fields are FP64 (NeoN::scalar with
 * WM_DP); a Vec3 result (the gradient) is an interleaved xyz Vector<double> of
 * 3N values.
 */
#ifndef FVHIP_HPP
#define FVHIP_HPP

#include <cstdint>
#include <initializer_list>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fvhip.h"

namespace fvhip {

// ---- errors ------------------------------------------------------------------

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const char* what) {
    if (rc != FV_OK) throw Error(rc, std::string(what) + ": " + fv_last_error());
}

// ---- executor and device vectors ----------------------------------------------

// NeoN::Executor on the GPU (createExecutor("GPU"), src/auxiliary/setup.cpp:51-80):
// one HIP stream and an HBM budget (0: unlimited).  Copies share the context.
class Executor {
   public:
    explicit Executor(int device = 0, size_t hbmBudget = 0) {
        fv_ctx* c = nullptr;
        check(fv_ctx_create(device, hbmBudget, &c), "fv_ctx_create");
        ctx_ = std::shared_ptr<fv_ctx>(c, [](fv_ctx* p) { fv_ctx_destroy(p); });
    }
    fv_ctx* get() const { return ctx_.get(); }
    void fence() const { check(fv_sync(ctx_.get()), "fv_sync"); }  // Kokkos::fence

   private:
    std::shared_ptr<fv_ctx> ctx_;
};

template <typename T>
struct PrecOf;
template <>
struct PrecOf<double> {
    static constexpr fv_prec value = FV_F64;
};
template <>
struct PrecOf<float> {
    static constexpr fv_prec value = FV_F32;
};

// NeoN::Vector<T> on the GPU executor: device storage owned by the vector.
template <typename T>
class Vector {
   public:
    Vector() = default;
    Vector(const Executor& exec, size_t n) : exec_(std::make_shared<Executor>(exec)), n_(n) {
        void* p = nullptr;
        if (n) check(fv_device_alloc(exec.get(), n * sizeof(T), &p), "fv_device_alloc");
        p_ = static_cast<T*>(p);
    }
    Vector(const Executor& exec, size_t n, T value) : Vector(exec, n) { fill(value); }
    Vector(const Executor& exec, const std::vector<T>& host) : Vector(exec, host.size()) {
        if (n_) check(fv_memcpy_h2d(exec.get(), p_, host.data(), n_ * sizeof(T)), "fv_memcpy_h2d");
    }
    Vector(const Vector&) = delete;
    Vector& operator=(const Vector&) = delete;
    Vector(Vector&& o) noexcept { swap(o); }
    Vector& operator=(Vector&& o) noexcept {
        if (this != &o) {
            release();
            swap(o);
        }
        return *this;
    }
    ~Vector() { release(); }

    T* data() { return p_; }
    const T* data() const { return p_; }
    size_t size() const { return n_; }
    const Executor& exec() const { return *exec_; }
    void fill(T v) {
        if (n_) check(fv_fill(exec_->get(), PrecOf<T>::value, p_, (int64_t)n_, (double)v), "fv_fill");
    }
    std::vector<T> copyToHost() const {
        std::vector<T> h(n_);
        if (n_) check(fv_memcpy_d2h(exec_->get(), h.data(), p_, n_ * sizeof(T)), "fv_memcpy_d2h");
        return h;
    }
    void copyFrom(const Vector& o) {
        if (o.n_ != n_) throw Error(FV_EINVAL, "Vector::copyFrom: size mismatch");
        if (n_) check(fv_memcpy_d2d(exec_->get(), p_, o.p_, n_ * sizeof(T)), "fv_memcpy_d2d");
    }

   private:
    void release() {
        if (p_ && exec_) fv_device_free(exec_->get(), p_);
        p_ = nullptr;
        n_ = 0;
    }
    void swap(Vector& o) noexcept {
        std::swap(exec_, o.exec_);
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
    }
    std::shared_ptr<Executor> exec_;
    T* p_ = nullptr;
    size_t n_ = 0;
};

// NeoN::scalar and NeoN::Vec3, the value types of the equations (a Vec3 field is
// stored xyz interleaved, 3 doubles per cell or face).
using scalar = double;
struct Vec3 {
    double x = 0, y = 0, z = 0;
};

// Component k of the interleaved [n][3] vector src into dst [n] (mode 0), or
// dst [n][3] component k = src [n] (mode 1).
inline void moveComponent(const Vector<double>& src, Vector<double>& dst, int k, size_t n, int mode) {
    if (n)
        check(fv_component(src.exec().get(), FV_F64, (int64_t)n, 3, k, mode, src.data(), dst.data()),
              "fv_component");
}

// ---- mesh ----------------------------------------------------------------------

// NeoN::UnstructuredMesh built by MeshAdapter from the polyMesh primitives
// (src/datastructures/meshAdapter.cpp:59-136); geometry is computed on the GPU.
class UnstructuredMesh {
   public:
    static UnstructuredMesh fromPolyMesh(const Executor& exec, const fv_polymesh_desc& d) {
        fv_mesh* m = nullptr;
        check(fv_mesh_create_poly(exec.get(), &d, &m), "fv_mesh_create_poly");
        return UnstructuredMesh(exec, m);
    }
    const Executor& exec() const { return exec_; }
    fv_mesh* get() const { return m_.get(); }
    int32_t nCells() const { return sizes_[0]; }
    int32_t nInternalFaces() const { return sizes_[1]; }
    int32_t nBoundaryFaces() const { return sizes_[2]; }
    int32_t nPatches() const { return sizes_[3]; }
    int32_t nFaces() const { return sizes_[1] + sizes_[2]; }
    // one geometry array on the host (fv_geom), e.g. FV_G_V, FV_G_PATCH_OFFSETS
    template <typename T>
    std::vector<T> get(fv_geom which, size_t n) const {
        std::vector<T> h(n);
        check(fv_mesh_get(m_.get(), which, h.data()), "fv_mesh_get");
        return h;
    }

   private:
    UnstructuredMesh(const Executor& exec, fv_mesh* m)
        : exec_(exec), m_(m, [](fv_mesh* p) { fv_mesh_destroy(p); }) {
        check(fv_mesh_sizes(m, sizes_), "fv_mesh_sizes");
    }
    Executor exec_;
    std::shared_ptr<fv_mesh> m_;
    int32_t sizes_[4] = {};
};

// Structured hex block in polyMesh form (the blockMesh cases of the reference),
// kept with the host arrays its descriptor points at.
struct HexBlock {
    int64_t sizes[6] = {};
    std::vector<double> points;
    std::vector<int32_t> faceOffsets, facePoints, owner, neighbour, patchStart, patchSize,
        patchEmpty;
    HexBlock(int nx, int ny, int nz, double jitter = 0.0, uint64_t jitterSeed = 0,
             uint64_t permSeed = 0, double lx = 1.0, double ly = 1.0, double lz = 1.0) {
        check(fv_hexmesh_sizes(nx, ny, nz, sizes), "fv_hexmesh_sizes");
        points.resize(3 * (size_t)sizes[0]);
        faceOffsets.resize(sizes[2] + sizes[3] + 1);
        facePoints.resize(sizes[5]);
        owner.resize(sizes[2] + sizes[3]);
        neighbour.resize(sizes[2] > 0 ? sizes[2] : 1);
        patchStart.resize(sizes[4]);
        patchSize.resize(sizes[4]);
        patchEmpty.assign(sizes[4], 0);
        check(fv_hexmesh_generate(nx, ny, nz, lx, ly, lz, jitter, jitterSeed, permSeed,
                                  points.data(), faceOffsets.data(), facePoints.data(),
                                  owner.data(), neighbour.data(), patchStart.data(),
                                  patchSize.data()),
              "fv_hexmesh_generate");
    }
    fv_polymesh_desc desc() const {
        fv_polymesh_desc d{};
        d.n_points = (int32_t)sizes[0];
        d.n_cells = (int32_t)sizes[1];
        d.n_internal_faces = (int32_t)sizes[2];
        d.n_faces = (int32_t)(sizes[2] + sizes[3]);
        d.n_patches = (int32_t)sizes[4];
        d.h_points = points.data();
        d.h_face_offsets = faceOffsets.data();
        d.h_face_points = facePoints.data();
        d.h_owner = owner.data();
        d.h_neighbour = neighbour.data();
        d.h_patch_start = patchStart.data();
        d.h_patch_size = patchSize.data();
        d.h_patch_empty = patchEmpty.data();
        return d;
    }
};

namespace la {
class LinearSystem;
}

namespace finiteVolume {
namespace cellCentred {

using TokenList = std::vector<std::string>;

// Boundary-condition kinds per patch after FoamAdapter's mapping
// (include/FoamAdapter/auxiliary/readers.hpp:43-95).
class BoundaryConditions {
   public:
    BoundaryConditions(const UnstructuredMesh& mesh, const std::vector<int32_t>& kinds)
        : kinds_(kinds) {
        fv_bc* b = nullptr;
        check(fv_bc_create(mesh.get(), kinds.data(), &b), "fv_bc_create");
        bc_ = std::shared_ptr<fv_bc>(b, [](fv_bc* p) { fv_bc_destroy(p); });
    }
    static int32_t kindOf(const std::string& name) {
        if (name == "fixedValue" || name == "noSlip") return FV_BC_FIXED_VALUE;
        if (name == "fixedGradient" || name == "zeroGradient") return FV_BC_FIXED_GRADIENT;
        if (name == "calculated" || name == "extrapolatedCalculated") return FV_BC_CALCULATED;
        throw Error(FV_EINVAL, "unsupported boundary condition " + name);
    }
    BoundaryConditions(const UnstructuredMesh& mesh, std::initializer_list<const char*> names)
        : BoundaryConditions(mesh, kindsOf(names)) {}
    const fv_bc* get() const { return bc_.get(); }
    const std::vector<int32_t>& kinds() const { return kinds_; }

   private:
    static std::vector<int32_t> kindsOf(std::initializer_list<const char*> names) {
        std::vector<int32_t> k;
        for (const char* n : names) k.push_back(kindOf(n));
        return k;
    }
    std::vector<int32_t> kinds_;
    std::shared_ptr<fv_bc> bc_;
};

// fvcc::VolumeField<scalar>: internalVector() [N], boundaryData().value() [B],
// the boundary conditions, an optional fixedGradient value per boundary face,
// and the old-time copy Euler ddt reads (fvcc::oldTime).
class VolumeField {
   public:
    VolumeField(const UnstructuredMesh& mesh, std::string name,
                std::shared_ptr<BoundaryConditions> bcs)
        : name(std::move(name)),
          mesh_(&mesh),
          bcs_(std::move(bcs)),
          internal_(mesh.exec(), (size_t)mesh.nCells(), 0.0),
          boundary_(mesh.exec(), (size_t)mesh.nBoundaryFaces(), 0.0) {}
    Vector<double>& internalVector() { return internal_; }
    const Vector<double>& internalVector() const { return internal_; }
    Vector<double>& boundaryValues() { return boundary_; }  // boundaryData().value()
    const Vector<double>& boundaryValues() const { return boundary_; }
    Vector<double>* boundaryGradient() { return gradient_.get(); }
    const Vector<double>* boundaryGradient() const { return gradient_.get(); }
    void setBoundaryGradient(Vector<double>&& g) { gradient_ = std::make_unique<Vector<double>>(std::move(g)); }
    const BoundaryConditions& bcs() const { return *bcs_; }
    const UnstructuredMesh& mesh() const { return *mesh_; }
    // VolumeField::correctBoundaryConditions (readers.hpp:123)
    void correctBoundaryConditions() {
        if (!mesh_->nBoundaryFaces()) return;
        check(fv_correct_boundary(mesh_->exec().get(), mesh_->get(), FV_F64, internal_.data(),
                                  bcs_->get(), gradient_ ? gradient_->data() : nullptr,
                                  boundary_.data()),
              "fv_correct_boundary");
    }
    // fvcc::oldTime(field): created on first use as a copy of the field
    Vector<double>& oldTime() {
        if (!old_) {
            old_ = std::make_unique<Vector<double>>(mesh_->exec(), internal_.size());
            old_->copyFrom(internal_);
        }
        return *old_;
    }
    std::shared_ptr<la::LinearSystem> linearSystemCache;  // dsl::solve's system on this mesh
    std::string name;

   private:
    const UnstructuredMesh* mesh_;
    std::shared_ptr<BoundaryConditions> bcs_;
    Vector<double> internal_, boundary_;
    std::unique_ptr<Vector<double>> gradient_, old_;
};

// fvcc::VolumeField<Vec3>: internal [3N] and boundary [3B] values xyz
// interleaved, one set of boundary conditions for the three components.  The
// equations of a vector field are solved component by component: component()
// copies one out into a scalar field (internal, boundary and old-time values),
// setComponent() copies one back.
class VolumeVectorField {
   public:
    VolumeVectorField(const UnstructuredMesh& mesh, std::string name,
                      std::shared_ptr<BoundaryConditions> bcs)
        : name(std::move(name)),
          mesh_(&mesh),
          bcs_(std::move(bcs)),
          internal_(mesh.exec(), 3 * (size_t)mesh.nCells(), 0.0),
          boundary_(mesh.exec(), 3 * (size_t)mesh.nBoundaryFaces(), 0.0) {}
    Vector<double>& internalVector() { return internal_; }
    const Vector<double>& internalVector() const { return internal_; }
    Vector<double>& boundaryValues() { return boundary_; }
    const Vector<double>& boundaryValues() const { return boundary_; }
    const BoundaryConditions& bcs() const { return *bcs_; }
    const std::shared_ptr<BoundaryConditions>& bcsPtr() const { return bcs_; }
    const UnstructuredMesh& mesh() const { return *mesh_; }
    bool hasOldTime() const { return (bool)old_; }
    // fvcc::oldTime(U): created on first use as a copy of the field
    Vector<double>& oldTime() {
        if (!old_) {
            old_ = std::make_unique<Vector<double>>(mesh_->exec(), internal_.size());
            old_->copyFrom(internal_);
        }
        return *old_;
    }
    void component(int k, VolumeField& out) {
        moveComponent(internal_, out.internalVector(), k, (size_t)mesh_->nCells(), 0);
        moveComponent(boundary_, out.boundaryValues(), k, (size_t)mesh_->nBoundaryFaces(), 0);
        if (old_) moveComponent(*old_, out.oldTime(), k, (size_t)mesh_->nCells(), 0);
    }
    void setComponent(int k, const VolumeField& in) {
        moveComponent(in.internalVector(), internal_, k, (size_t)mesh_->nCells(), 1);
        moveComponent(in.boundaryValues(), boundary_, k, (size_t)mesh_->nBoundaryFaces(), 1);
    }
    // each component's fixedGradient / zeroGradient faces from its cell values
    void correctBoundaryConditions() {
        if (!mesh_->nBoundaryFaces()) return;
        for (int k = 0; k < 3; ++k) {
            VolumeField c(*mesh_, name, bcs_);
            moveComponent(internal_, c.internalVector(), k, (size_t)mesh_->nCells(), 0);
            moveComponent(boundary_, c.boundaryValues(), k, (size_t)mesh_->nBoundaryFaces(), 0);
            c.correctBoundaryConditions();
            moveComponent(c.boundaryValues(), boundary_, k, (size_t)mesh_->nBoundaryFaces(), 1);
        }
    }
    std::string name;

   private:
    const UnstructuredMesh* mesh_;
    std::shared_ptr<BoundaryConditions> bcs_;
    Vector<double> internal_, boundary_;
    std::unique_ptr<Vector<double>> old_;
};

// fvcc::SurfaceField<scalar>: internal faces then boundary faces [F+B]
// (readers.hpp:175-217).
class SurfaceField {
   public:
    SurfaceField(const UnstructuredMesh& mesh, std::string name, double value = 0.0)
        : name(std::move(name)), mesh_(&mesh), values_(mesh.exec(), (size_t)mesh.nFaces(), value) {}
    SurfaceField(const UnstructuredMesh& mesh, std::string name, const std::vector<double>& host)
        : name(std::move(name)), mesh_(&mesh), values_(mesh.exec(), host) {
        if (host.size() != (size_t)mesh.nFaces()) throw Error(FV_EINVAL, "SurfaceField: size");
    }
    Vector<double>& internalVector() { return values_; }
    const Vector<double>& internalVector() const { return values_; }
    const UnstructuredMesh& mesh() const { return *mesh_; }
    std::string name;

   private:
    const UnstructuredMesh* mesh_;
    Vector<double> values_;
};

inline fv_scheme schemeOf(const TokenList& t) {
    for (const std::string& s : t) {
        if (s == "linear") return FV_LINEAR;
        if (s == "upwind") return FV_UPWIND;
    }
    for (const std::string& s : t)
        if (s != "Gauss" && s != "uncorrected")
            throw Error(FV_EINVAL, "unsupported interpolation scheme " + s);
    return FV_LINEAR;
}

}  // namespace cellCentred
}  // namespace finiteVolume

namespace dsl {
// dsl::Coeff: a scalar factor on an operator's result
struct Coeff {
    double value = 1.0;
    Coeff() = default;
    Coeff(double v) : value(v) {}  // NOLINT: implicit, as NeoN's
};
}  // namespace dsl

namespace finiteVolume {
namespace cellCentred {

// SurfaceInterpolation linear / upwind (test_operators.cpp:52-59)
class SurfaceInterpolation {
   public:
    SurfaceInterpolation(const Executor& exec, const UnstructuredMesh& mesh, const TokenList& s)
        : exec_(exec), mesh_(mesh), scheme_(schemeOf(s)) {}
    void interpolate(const VolumeField& T, SurfaceField& out, const SurfaceField* phi = nullptr) const {
        if (scheme_ == FV_UPWIND && !phi) throw Error(FV_EINVAL, "upwind needs the face flux");
        check(fv_interpolate(exec_.get(), mesh_.get(), scheme_, FV_F64,
                             phi ? phi->internalVector().data() : nullptr,
                             T.internalVector().data(), T.boundaryValues().data(),
                             out.internalVector().data()),
              "fv_interpolate");
    }

   private:
    Executor exec_;
    const UnstructuredMesh& mesh_;
    fv_scheme scheme_;
};

// GaussGreenGrad::grad(T, coeff, out) (test_operators.cpp:72): out [3N] xyz
class GaussGreenGrad {
   public:
    GaussGreenGrad(const Executor& exec, const UnstructuredMesh& mesh) : exec_(exec), mesh_(mesh) {}
    void grad(const VolumeField& T, dsl::Coeff c, Vector<double>& out) const {
        if (out.size() != 3 * (size_t)mesh_.nCells()) throw Error(FV_EINVAL, "grad: out is not 3N");
        check(fv_grad_gauss(exec_.get(), mesh_.get(), FV_F64, c.value, T.internalVector().data(),
                            T.boundaryValues().data(), out.data()),
              "fv_grad_gauss");
    }

   private:
    Executor exec_;
    const UnstructuredMesh& mesh_;
};

// GaussGreenDiv<scalar>::div(out, phi, T, coeff) (test_operators.cpp:109-110)
class GaussGreenDiv {
   public:
    GaussGreenDiv(const Executor& exec, const UnstructuredMesh& mesh, const TokenList& s)
        : exec_(exec), mesh_(mesh), scheme_(schemeOf(s)) {}
    void div(Vector<double>& out, const SurfaceField& phi, const VolumeField& T, dsl::Coeff c) const {
        check(fv_div_gauss(exec_.get(), mesh_.get(), scheme_, FV_F64, c.value,
                           phi.internalVector().data(), T.internalVector().data(),
                           T.boundaryValues().data(), out.data()),
              "fv_div_gauss");
    }
    void div(VolumeField& out, const SurfaceField& phi, const VolumeField& T, dsl::Coeff c) const {
        div(out.internalVector(), phi, T, c);
    }
    // single-argument form: surfaceIntegrate(phi) (neoIcoFoam.cpp:145)
    void div(Vector<double>& out, const SurfaceField& phi, dsl::Coeff c) const {
        check(fv_div_gauss(exec_.get(), mesh_.get(), scheme_, FV_F64, c.value,
                           phi.internalVector().data(), nullptr, nullptr, out.data()),
              "fv_div_gauss");
    }

   private:
    Executor exec_;
    const UnstructuredMesh& mesh_;
    fv_scheme scheme_;
};

// GaussGreenLaplacian "linear uncorrected" (bench_explicitOperators.cpp:212)
class GaussGreenLaplacian {
   public:
    GaussGreenLaplacian(const Executor& exec, const UnstructuredMesh& mesh, const TokenList& = {})
        : exec_(exec), mesh_(mesh) {}
    void laplacian(Vector<double>& out, const SurfaceField& gamma, const VolumeField& T,
                   dsl::Coeff c) const {
        const Vector<double>* g = T.boundaryGradient();
        check(fv_laplacian_uncorrected(exec_.get(), mesh_.get(), FV_F64, c.value,
                                       gamma.internalVector().data(), T.internalVector().data(),
                                       T.bcs().get(), T.boundaryValues().data(),
                                       g ? g->data() : nullptr, out.data()),
              "fv_laplacian_uncorrected");
    }

   private:
    Executor exec_;
    const UnstructuredMesh& mesh_;
};

// FaceNormalGradient "uncorrected" (bench_explicitOperators.cpp:478)
class FaceNormalGradient {
   public:
    FaceNormalGradient(const Executor& exec, const UnstructuredMesh& mesh, const TokenList& = {})
        : exec_(exec), mesh_(mesh) {}
    void faceNormalGrad(const VolumeField& T, SurfaceField& out) const {
        const Vector<double>* g = T.boundaryGradient();
        check(fv_sngrad_uncorrected(exec_.get(), mesh_.get(), FV_F64, T.internalVector().data(),
                                    T.bcs().get(), T.boundaryValues().data(),
                                    g ? g->data() : nullptr, out.internalVector().data()),
              "fv_sngrad_uncorrected");
    }

   private:
    Executor exec_;
    const UnstructuredMesh& mesh_;
};

}  // namespace cellCentred
}  // namespace finiteVolume

namespace fvcc = finiteVolume::cellCentred;

// ---- linear algebra ----------------------------------------------------------------

namespace la {

// la::SparsityPattern::readOrCreate(mesh) (expression.hpp:35)
class SparsityPattern {
   public:
    explicit SparsityPattern(const UnstructuredMesh& mesh) {
        fv_sparsity* s = nullptr;
        check(fv_sparsity_create(mesh.get(), &s), "fv_sparsity_create");
        sp_ = std::shared_ptr<fv_sparsity>(s, [](fv_sparsity* p) { fv_sparsity_destroy(p); });
    }
    fv_sparsity* get() const { return sp_.get(); }

   private:
    std::shared_ptr<fv_sparsity> sp_;
};

// la::SolverStats (expression.hpp:155-158) with the refinement count and the
// explicit residual
struct SolverStats {
    double initResNorm = 0, finalResNorm = 0, trueResNorm = 0;
    int numIter = 0, outerIter = 0;
};

// la::LinearSystem<scalar, localIdx> (bench_implicitOperators.cpp:109-125):
// values[nnz], rhs[N] on the device; A.x - b is the operators applied to x, times V.
class LinearSystem {
   public:
    explicit LinearSystem(const SparsityPattern& sp) : sp_(sp) {
        fv_linsys* l = nullptr;
        check(fv_linsys_create(sp.get(), &l), "fv_linsys_create");
        ls_ = std::shared_ptr<fv_linsys>(l, [](fv_linsys* p) { fv_linsys_destroy(p); });
    }
    fv_linsys* get() const { return ls_.get(); }
    void zero() { check(fv_linsys_zero(ls_.get()), "fv_linsys_zero"); }
    double* values() { return static_cast<double*>(fv_linsys_values(ls_.get())); }
    double* rhs() { return static_cast<double*>(fv_linsys_rhs(ls_.get())); }

   private:
    SparsityPattern sp_;
    std::shared_ptr<fv_linsys> ls_;
};

// la::computeResidual via applyOperator (expression.hpp:204-220): res = A.x - b
inline void computeResidual(LinearSystem& ls, const Vector<double>& x, Vector<double>& res) {
    check(fv_residual(ls.get(), x.data(), res.data()), "fv_residual");
}

}  // namespace la

// ---- expressions and solve ----------------------------------------------------------

// Dictionary of an fvSolution solver entry or an fvSchemes selection, values as
// written in the case files ("PCG", "1e-10", "1000", "Gauss upwind", ...).
using Dictionary = std::map<std::string, std::string>;

namespace dsl {

struct Term {
    enum Kind { Ddt, Laplacian, Div, Source, ExpDiv, ExpLaplacian } kind;
    double coeff = 1.0;
    std::string scheme = "linear";
    fvcc::VolumeField* psi = nullptr;
    const fvcc::SurfaceField* face = nullptr;  // gamma (laplacian) or phi (div)
    const Vector<double>* cell = nullptr;      // source coefficient field
    fvcc::VolumeVectorField* vpsi = nullptr;   // the unknown of a vector equation
};

// dsl::Expression: a sum of implicit and explicit terms
class Expression {
   public:
    Expression() = default;
    Expression(const Term& t) : terms_{t} {}  // NOLINT: a term is an expression
    Expression& operator+=(const Expression& o) {
        terms_.insert(terms_.end(), o.terms_.begin(), o.terms_.end());
        return *this;
    }
    Expression& operator-=(const Expression& o) {
        for (Term t : o.terms_) {
            t.coeff = -t.coeff;
            terms_.push_back(t);
        }
        return *this;
    }
    const std::vector<Term>& terms() const { return terms_; }
    std::vector<Term>& terms() { return terms_; }
    // Expression::read: divSchemes from fvSchemes ("div" -> "Gauss upwind")
    void read(const Dictionary& fvSchemes) {
        auto it = fvSchemes.find("div");
        if (it == fvSchemes.end()) return;
        const std::string s = it->second.substr(it->second.rfind(' ') + 1);
        for (Term& t : terms_)
            if (t.kind == Term::Div || t.kind == Term::ExpDiv) t.scheme = s;
    }

   private:
    std::vector<Term> terms_;
};

inline Expression operator+(Expression a, const Expression& b) { return a += b; }
inline Expression operator-(Expression a, const Expression& b) { return a -= b; }
inline Expression operator*(double c, Expression e) {
    for (Term& t : e.terms()) t.coeff *= c;
    return e;
}

namespace imp {
inline Term ddt(fvcc::VolumeField& psi) { return Term{Term::Ddt, 1.0, "linear", &psi, nullptr, nullptr}; }
inline Term laplacian(const fvcc::SurfaceField& gamma, fvcc::VolumeField& psi) {
    return Term{Term::Laplacian, 1.0, "linear", &psi, &gamma, nullptr};
}
inline Term div(const fvcc::SurfaceField& phi, fvcc::VolumeField& psi, std::string scheme = "linear") {
    return Term{Term::Div, 1.0, std::move(scheme), &psi, &phi, nullptr};
}
inline Term source(const Vector<double>& coeffField, fvcc::VolumeField& psi) {
    return Term{Term::Source, 1.0, "linear", &psi, nullptr, &coeffField};
}
// the same operators on a vector field (PDESolver<Vec3>, neoIcoFoam.cpp:94-96)
inline Term ddt(fvcc::VolumeVectorField& U) {
    return Term{Term::Ddt, 1.0, "linear", nullptr, nullptr, nullptr, &U};
}
inline Term laplacian(const fvcc::SurfaceField& gamma, fvcc::VolumeVectorField& U) {
    return Term{Term::Laplacian, 1.0, "linear", nullptr, &gamma, nullptr, &U};
}
inline Term div(const fvcc::SurfaceField& phi, fvcc::VolumeVectorField& U,
                std::string scheme = "linear") {
    return Term{Term::Div, 1.0, std::move(scheme), nullptr, &phi, nullptr, &U};
}
}  // namespace imp

namespace exp {
inline Term div(const fvcc::SurfaceField& phi, fvcc::VolumeField& psi, std::string scheme = "linear") {
    return Term{Term::ExpDiv, 1.0, std::move(scheme), &psi, &phi, nullptr};
}
// single-argument dsl::exp::div(phi): surfaceIntegrate (neoIcoFoam.cpp:145)
inline Term div(const fvcc::SurfaceField& phi) {
    return Term{Term::ExpDiv, 1.0, "linear", nullptr, &phi, nullptr};
}
inline Term laplacian(const fvcc::SurfaceField& gamma, fvcc::VolumeField& psi) {
    return Term{Term::ExpLaplacian, 1.0, "linear", &psi, &gamma, nullptr};
}
}  // namespace exp

// Expression::assemble(t, dt, sparsity, ls) (expression.hpp:69-73)
inline void assemble(const Expression& e, fvcc::VolumeField& psi, double /*t*/, double dt,
                     la::LinearSystem& ls) {
    ls.zero();
    const UnstructuredMesh& m = psi.mesh();
    std::unique_ptr<Vector<double>> tmp;
    for (const Term& t : e.terms()) {
        switch (t.kind) {
            case Term::Ddt:
                check(fv_assemble_ddt_euler(ls.get(), t.psi->oldTime().data(), dt, t.coeff),
                      "fv_assemble_ddt_euler");
                break;
            case Term::Laplacian: {
                const Vector<double>* g = t.psi->boundaryGradient();
                check(fv_assemble_laplacian(ls.get(), t.face->internalVector().data(),
                                            t.psi->bcs().get(), t.psi->boundaryValues().data(),
                                            g ? g->data() : nullptr, t.coeff),
                      "fv_assemble_laplacian");
                break;
            }
            case Term::Div: {
                const Vector<double>* g = t.psi->boundaryGradient();
                check(fv_assemble_div(ls.get(), fvcc::schemeOf({t.scheme}),
                                      t.face->internalVector().data(), t.psi->bcs().get(),
                                      t.psi->boundaryValues().data(), g ? g->data() : nullptr,
                                      t.coeff),
                      "fv_assemble_div");
                break;
            }
            case Term::Source:
                check(fv_assemble_source(ls.get(), t.cell->data(), t.coeff), "fv_assemble_source");
                break;
            case Term::ExpDiv:
            case Term::ExpLaplacian: {
                if (!tmp) tmp = std::make_unique<Vector<double>>(m.exec(), (size_t)m.nCells());
                if (t.kind == Term::ExpDiv) {
                    fvcc::GaussGreenDiv d(m.exec(), m, {t.scheme});
                    if (t.psi)
                        d.div(*tmp, *t.face, *t.psi, 1.0);
                    else
                        d.div(*tmp, *t.face, 1.0);
                } else {
                    fvcc::GaussGreenLaplacian(m.exec(), m).laplacian(*tmp, *t.face, *t.psi, 1.0);
                }
                check(fv_assemble_explicit(ls.get(), tmp->data(), t.coeff), "fv_assemble_explicit");
                break;
            }
        }
    }
}

// The equation of a vector field U, component by component on one matrix (the
// components share their boundary-condition kinds, so the matrix is the same
// for all three): each component's right-hand side goes to rhs3 [3N]
// (interleaved); ls keeps component 0's system.
inline void assembleVector(const Expression& e, fvcc::VolumeVectorField& U, double t, double dt,
                           la::LinearSystem& ls, Vector<double>& rhs3) {
    const UnstructuredMesh& m = U.mesh();
    const size_t N = (size_t)m.nCells();
    for (const Term& term : e.terms())
        if (term.kind == Term::Ddt && term.vpsi == &U) U.oldTime();
    for (int k = 2; k >= 0; --k) {
        fvcc::VolumeField c(m, U.name, U.bcsPtr());
        U.component(k, c);
        Expression ek = e;
        for (Term& term : ek.terms())
            if (term.vpsi == &U) {
                term.psi = &c;
                term.vpsi = nullptr;
            }
        assemble(ek, c, t, dt, ls);
        Vector<double> b(m.exec(), N);
        if (N) check(fv_memcpy_d2d(m.exec().get(), b.data(), ls.rhs(), N * sizeof(double)),
                     "fv_memcpy_d2d");
        moveComponent(b, rhs3, k, N, 1);
    }
}

// fvSolution solver entry -> solve options (src/compatibility/fvSolution.cpp:19-159):
// solver PCG -> CG, PBiCGStab / smoothSolver -> BiCGStab; relTol, maxIter (1000),
// tolerance; "precision mixed" selects FP32 CG inside FP64 refinement
inline fv_solve_opts solveOptions(const Dictionary& d, bool* bicgstab) {
    auto num = [&](const char* k, double def) {
        auto it = d.find(k);
        return it == d.end() ? def : std::stod(it->second);
    };
    fv_solve_opts o{};
    o.rel_tol = num("relTol", 1e-6);
    o.max_iter = (int32_t)num("maxIter", 1000);
    o.abs_tol = num("tolerance", 0.0);
    o.check_every = (int32_t)num("checkEvery", 32);
    auto p = d.find("precision");
    o.mode = (p != d.end() && p->second == "mixed") ? FV_SOLVE_F32_IR : FV_SOLVE_F64;
    auto s = d.find("solver");
    const std::string name = s == d.end() ? "PCG" : s->second;
    if (name == "PCG" || name == "CG")
        *bicgstab = false;
    else if (name == "PBiCGStab" || name == "smoothSolver" || name == "BiCGStab")
        *bicgstab = true;
    else
        throw Error(FV_EINVAL, "solver " + name + " is not supported (PCG, PBiCGStab)");
    return o;
}

inline la::SolverStats solveSystem(la::LinearSystem& ls, Vector<double>& x, const Dictionary& solver) {
    bool bicg = false;
    const fv_solve_opts o = solveOptions(solver, &bicg);
    fv_solver_stats st{};
    const int rc = bicg ? fv_solve_bicgstab(ls.get(), x.data(), &o, &st)
                        : fv_solve_cg(ls.get(), x.data(), &o, &st);
    check(rc, bicg ? "fv_solve_bicgstab" : "fv_solve_cg");
    return la::SolverStats{st.init_res, st.final_res, st.true_res, st.iters, st.outer_iters};
}

inline la::LinearSystem& systemOf(fvcc::VolumeField& psi) {
    if (!psi.linearSystemCache)
        psi.linearSystemCache =
            std::make_shared<la::LinearSystem>(la::SparsityPattern(psi.mesh()));
    return *psi.linearSystemCache;
}

// dsl::solve(expr, psi, t, dt, fvSchemes, solverDict) (heatTransfer.cpp:69-76):
// assemble, solve, correct psi's boundary values and advance its old time
inline la::SolverStats solve(Expression expr, fvcc::VolumeField& psi, double t, double dt,
                             const Dictionary& fvSchemes, const Dictionary& solverDict) {
    expr.read(fvSchemes);
    la::LinearSystem& ls = systemOf(psi);
    assemble(expr, psi, t, dt, ls);
    la::SolverStats st = solveSystem(ls, psi.internalVector(), solverDict);
    psi.correctBoundaryConditions();
    for (const Term& term : expr.terms())
        if (term.kind == Term::Ddt) {
            psi.oldTime().copyFrom(psi.internalVector());
            break;
        }
    return st;
}

}  // namespace dsl

// FoamAdapter::PDESolver<T> (include/FoamAdapter/datastructures/expression.hpp:23-179):
// assemble, optional SetReference, solve.  PDESolver<scalar> (also written
// PDESolver, its arguments deciding) and PDESolver<Vec3>.
template <typename T = scalar>
class PDESolver {
   public:
    PDESolver(dsl::Expression expr, fvcc::VolumeField& psi, double t = 0.0, double dt = 1.0,
              Dictionary fvSchemes = {}, Dictionary solverDict = {})
        : expr_(std::move(expr)), psi_(psi), t_(t), dt_(dt), solver_(std::move(solverDict)),
          ls_(la::SparsityPattern(psi.mesh())) {
        expr_.read(fvSchemes);
    }
    la::LinearSystem& assemble() {
        dsl::assemble(expr_, psi_, t_, dt_, ls_);
        return ls_;
    }
    void setReference(int cell, double value) {
        refCell_ = cell;
        refValue_ = value;
    }
    la::LinearSystem& linearSystem() { return ls_; }
    fvcc::VolumeField& psi() { return psi_; }
    // RunTime::t and RunTime::dt the next assembly uses
    void setTime(double t, double dt) {
        t_ = t;
        dt_ = dt;
    }
    la::SolverStats solve() {
        assemble();
        if (refCell_ >= 0) check(fv_set_reference(ls_.get(), refCell_, refValue_), "fv_set_reference");
        la::SolverStats st = dsl::solveSystem(ls_, psi_.internalVector(), solver_);
        psi_.correctBoundaryConditions();
        return st;
    }

   private:
    dsl::Expression expr_;
    fvcc::VolumeField& psi_;
    double t_, dt_;
    Dictionary solver_;
    la::LinearSystem ls_;
    int refCell_ = -1;
    double refValue_ = 0.0;
};

// PDESolver<Vec3>: the matrix once, three right-hand sides, the components
// solved in turn (BiCGStab for the non-symmetric momentum equation when the
// solver dictionary says PBiCGStab / smoothSolver).
template <>
class PDESolver<Vec3> {
   public:
    PDESolver(dsl::Expression expr, fvcc::VolumeVectorField& U, double t = 0.0, double dt = 1.0,
              Dictionary fvSchemes = {}, Dictionary solverDict = {})
        : expr_(std::move(expr)), U_(U), t_(t), dt_(dt), solver_(std::move(solverDict)),
          ls_(la::SparsityPattern(U.mesh())), rhs3_(U.mesh().exec(), 3 * (size_t)U.mesh().nCells()) {
        expr_.read(fvSchemes);
    }
    la::LinearSystem& assemble() {
        dsl::assembleVector(expr_, U_, t_, dt_, ls_, rhs3_);
        return ls_;
    }
    la::LinearSystem& linearSystem() { return ls_; }
    const Vector<double>& rhs3() const { return rhs3_; }  // the three right-hand sides, xyz
    fvcc::VolumeVectorField& psi() { return U_; }
    void setTime(double t, double dt) {
        t_ = t;
        dt_ = dt;
    }
    std::vector<la::SolverStats> solve() {
        assemble();
        const UnstructuredMesh& m = U_.mesh();
        const size_t N = (size_t)m.nCells();
        std::vector<la::SolverStats> st;
        Vector<double> x(m.exec(), N), b(m.exec(), N);
        for (int k = 0; k < 3; ++k) {
            moveComponent(rhs3_, b, k, N, 0);
            if (N) check(fv_memcpy_d2d(m.exec().get(), ls_.rhs(), b.data(), N * sizeof(double)),
                         "fv_memcpy_d2d");
            moveComponent(U_.internalVector(), x, k, N, 0);
            st.push_back(dsl::solveSystem(ls_, x, solver_));
            moveComponent(x, U_.internalVector(), k, N, 1);
        }
        U_.correctBoundaryConditions();
        return st;
    }

   private:
    dsl::Expression expr_;
    fvcc::VolumeVectorField& U_;
    double t_, dt_;
    Dictionary solver_;
    la::LinearSystem ls_;
    Vector<double> rhs3_;
};

// ---- pressure-velocity coupling (src/algorithms/pressureVelocityCoupling.cpp) ----

// computeRAUandHByA(UEqn) (:63-119): rAU = V / diag(A) and HbyA = (b - sum_nb
// A U_nb) rAU / V, boundary values extrapolated; HbyA takes U's conditions
inline std::pair<fvcc::VolumeField, fvcc::VolumeVectorField> computeRAUandHByA(
    PDESolver<Vec3>& UEqn) {
    fvcc::VolumeVectorField& U = UEqn.psi();
    const UnstructuredMesh& m = U.mesh();
    fvcc::VolumeField rAU(m, "rAU", nullptr);
    fvcc::VolumeVectorField HbyA(m, "HbyA", U.bcsPtr());
    la::LinearSystem& ls = UEqn.linearSystem();
    check(fv_compute_rau(ls.get(), rAU.internalVector().data(), rAU.boundaryValues().data()),
          "fv_compute_rau");
    check(fv_compute_hbya(ls.get(), UEqn.rhs3().data(), U.internalVector().data(),
                          rAU.internalVector().data(), HbyA.internalVector().data(),
                          HbyA.boundaryValues().data()),
          "fv_compute_hbya");
    return {std::move(rAU), std::move(HbyA)};
}

// constrainHbyA (:14-36): HbyA_b = U_b on U's fixedValue patches
inline void constrainHbyA(const fvcc::VolumeVectorField& U, const fvcc::VolumeField& /*p*/,
                          fvcc::VolumeVectorField& HbyA) {
    check(fv_constrain_hbya(U.bcs().get(), U.boundaryValues().data(), HbyA.boundaryValues().data()),
          "fv_constrain_hbya");
}

// flux(U) (:213-267): phi = Sf . interpolate(U), boundary Sf_b . U_b
inline fvcc::SurfaceField flux(const fvcc::VolumeVectorField& U) {
    const UnstructuredMesh& m = U.mesh();
    fvcc::SurfaceField phi(m, "phi");
    check(fv_flux(m.exec().get(), m.get(), FV_F64, U.internalVector().data(),
                  U.boundaryValues().data(), phi.internalVector().data()),
          "fv_flux");
    return phi;
}

// updateFaceVelocity(phiHbyA, pEqn, phi) (:122-195): phi = phiHbyA - pEqn.flux()
inline void updateFaceVelocity(const fvcc::SurfaceField& phiHbyA, PDESolver<scalar>& pEqn,
                               fvcc::SurfaceField& phi) {
    check(fv_update_face_velocity(pEqn.linearSystem().get(), phiHbyA.internalVector().data(),
                                  pEqn.psi().internalVector().data(), phi.internalVector().data()),
          "fv_update_face_velocity");
}

// updateVelocity(HbyA, rAU, p, U) (:197-211): U = HbyA - rAU grad(p)
inline void updateVelocity(const fvcc::VolumeVectorField& HbyA, const fvcc::VolumeField& rAU,
                           const fvcc::VolumeField& p, fvcc::VolumeVectorField& U) {
    const UnstructuredMesh& m = U.mesh();
    Vector<double> gradP(m.exec(), 3 * (size_t)m.nCells());
    fvcc::GaussGreenGrad(m.exec(), m).grad(p, dsl::Coeff(), gradP);
    check(fv_update_velocity(m.exec().get(), m.get(), HbyA.internalVector().data(),
                             rAU.internalVector().data(), gradP.data(), U.internalVector().data()),
          "fv_update_velocity");
}

// computeCoNum(phi, dt) (scalarAdvection.cpp:69) and setDeltaT (setup.cpp:13-22)
inline double computeCoNum(const fvcc::SurfaceField& phi, double dt) {
    double co = 0.0;
    check(fv_courant_number(phi.mesh().exec().get(), phi.mesh().get(),
                            phi.internalVector().data(), dt, &co),
          "fv_courant_number");
    return co;
}
inline double setDeltaT(double dt, double coNum, double maxCo, double maxDeltaT) {
    double out = dt;
    check(fv_set_delta_t(dt, coNum, maxCo, maxDeltaT, &out), "fv_set_delta_t");
    return out;
}

}  // namespace fvhip

#endif  // FVHIP_HPP
