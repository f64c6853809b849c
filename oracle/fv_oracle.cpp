// fv_oracle.cpp -- CPU restatement of the FoamAdapter/NeoN finite-volume hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker against which the HIP
// library (foamadapter_amd/csrc) is compared.  Only tests/, the smoke() entry
// and the cpu_baseline leg of bench.py may load it; the product path never
// calls it and has no CPU fallback.
//
// NeoN (the library that defines the reference's fvcc/dsl/la classes) is an
// empty git submodule in the reference (.gitmodules:1-3), so the arithmetic
// below restates the OpenFOAM v2406 schemes that NeoN reproduces and that the
// reference's tests compare against.  Every routine is written serially, in
// OpenFOAM's face order, with no FMA contraction (build with
// -ffp-contract=off), so that its results define the expected bits.
//
// Pinning (what fixes each routine independently of this file):
//   face/cell geometry ...... closed forms on uniform hex meshes (exact on dyadic
//                             spacings), on affine images of them (skewed,
//                             non-orthogonal) and on tapered meshes (trapezoid
//                             faces: shoelace centroids, integrated cell
//                             centroids); on jittered meshes sum(V) = domain
//                             volume, sum over a cell's faces of +-Sf = 0, quad
//                             Sf = half the diagonals' cross product, V = 1/3
//                             sum Sf.Cf.  The centre weighting of non-planar faces
//                             beyond these is "parity unpinned" (it needs
//                             OpenFOAM's primitiveMesh to compare against).
//   interpolate/grad/div .... golden OpenFOAM fields test/setup_operator/0/ofGradT,
//                             ofDivT (and NeoN-serial gradT_Serial, divT_Serial).
//   laplacian/snGrad ........ closed forms: lap(linear)=0, lap(x^2)=2, snGrad of a
//                             linear field = its normal derivative.
//   sparsity ................ invariants of test/test_sparsityPattern.cpp:42-62.
//   assembly ................ A.x-b equals the explicit operator times V (the
//                             identity of test/test_implicitOperators.cpp:92-100,
//                             254-262), symmetry and row sums.
//   Jacobi-PCG .............. scipy.sparse.linalg.spsolve on small meshes and the
//                             explicit residual ||b-Ax||/||b|| <= tol.
//   rAU / HbyA .............. the matrix form rAU = V/diag(A), HbyA = (b - (A -
//                             diag A) U) rAU/V evaluated with scipy.
//   Courant number .......... closed form u*dt/h for uniform flow on a uniform mesh
//   face/cell velocity update the corrected flux of a solved pressure equation
//                             is divergence-free (surfaceIntegrate(phi) ~ 0);
//                             U = HbyA - rAU grad(p) against the grad oracle.
//                             The boundary coefficients (NeoN's
//                             BoundaryCoefficients, NeoN absent) are "parity
//                             unpinned" beyond that conservation identity.
//
// Reference files followed:
//   src/datastructures/meshAdapter.cpp:59-136   mesh arrays, flattened boundary
//   include/FoamAdapter/auxiliary/readers.hpp:43-95,175-217  BC kinds, surface layout
//   src/algorithms/pressureVelocityCoupling.cpp:87-112,242-264  owner/neighbour offsets,
//                                                interpolation form w*(P-N)+N
//   include/FoamAdapter/datastructures/expression.hpp:95-111,181-220
//                                                SetReference, diag, A.x-b
//   src/compatibility/fvSolution.cpp:19-159     PCG + point Jacobi, relTol, maxIter
//   src/algorithms/pressureVelocityCoupling.cpp:14-211  constrainHbyA, computeRAU,
//                                                computeRAUandHByA, updateFaceVelocity,
//                                                updateVelocity
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <algorithm>

namespace {

// boundary condition kinds per patch (readers.hpp:43-95 maps zeroGradient ->
// fixedGradient 0, extrapolatedCalculated -> calculated, noSlip -> fixedValue 0)
enum BcKind : int32_t { BC_FIXED_VALUE = 0, BC_FIXED_GRADIENT = 1, BC_CALCULATED = 2 };

struct V3 { double x, y, z; };
inline V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 vmul(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
// OpenFOAM's vector/scalar divides each component
inline V3 vdiv(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 vcross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double vmag(V3 a) { return std::sqrt(vdot(a, a)); }
inline V3 ld(const double* p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline void st(double* p, int64_t i, V3 v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }

const double VSMALL = 1.0e-300;
const double ROOTVSMALL = 1.0e-150;

// patch index of boundary face b (patch offsets are ascending)
inline int patch_of(int b, int nPatches, const int32_t* poff) {
    int p = 0;
    while (p + 1 < nPatches && b >= poff[p + 1]) ++p;
    return p;
}

template <typename S>
S bnd_sngrad(int kind, S dcb, S Tb, S TP, S g) {
    // fvPatchField::snGrad(): deltaCoeffs*(value - internal); fixedGradient returns
    // its gradient (OpenFOAM fixedGradientFvPatchField::snGrad)
    if (kind == BC_FIXED_GRADIENT) return g;
    return dcb * (Tb - TP);
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// geometry (OpenFOAM primitiveMesh::makeFaceCentresAndAreas / makeCellCentresAndVols,
// surfaceInterpolation::makeWeights / makeDeltaCoeffs / makeNonOrthDeltaCoeffs)
// ---------------------------------------------------------------------------

int orc_face_geometry(int nFaces, const int32_t* fOff, const int32_t* fVerts, const double* pts,
                      double* Cf, double* Sf) {
    for (int f = 0; f < nFaces; ++f) {
        const int n = fOff[f + 1] - fOff[f];
        const int32_t* v = fVerts + fOff[f];
        if (n == 3) {
            V3 p0 = ld(pts, v[0]), p1 = ld(pts, v[1]), p2 = ld(pts, v[2]);
            st(Cf, f, vmul(1.0 / 3.0, vadd(vadd(p0, p1), p2)));
            st(Sf, f, vmul(0.5, vcross(vsub(p1, p0), vsub(p2, p0))));
            continue;
        }
        V3 fc = ld(pts, v[0]);
        for (int i = 1; i < n; ++i) fc = vadd(fc, ld(pts, v[i]));
        fc = vdiv(fc, (double)n);
        V3 sumN{0, 0, 0}, sumAc{0, 0, 0};
        double sumA = 0.0;
        for (int i = 0; i < n; ++i) {
            V3 a = ld(pts, v[i]);
            V3 b = ld(pts, v[i + 1 == n ? 0 : i + 1]);
            V3 c = vadd(vadd(a, b), fc);
            V3 nn = vcross(vsub(b, a), vsub(fc, a));
            double m = vmag(nn);
            sumN = vadd(sumN, nn);
            sumA += m;
            sumAc = vadd(sumAc, vmul(m, c));
        }
        if (sumA < ROOTVSMALL) {
            st(Cf, f, fc);
            st(Sf, f, {0, 0, 0});
        } else {
            st(Cf, f, vdiv(vmul(1.0 / 3.0, sumAc), sumA));
            st(Sf, f, vmul(0.5, sumN));
        }
    }
    return 0;
}

int orc_cell_geometry(int nCells, int nFaces, int nInternal, const int32_t* owner,
                      const int32_t* nbr, const double* Cf, const double* Sf, double* C,
                      double* V) {
    std::vector<V3> cEst(nCells, V3{0, 0, 0});
    std::vector<int> nCellFaces(nCells, 0);
    for (int f = 0; f < nFaces; ++f) {
        cEst[owner[f]] = vadd(cEst[owner[f]], ld(Cf, f));
        nCellFaces[owner[f]]++;
    }
    for (int f = 0; f < nInternal; ++f) {
        cEst[nbr[f]] = vadd(cEst[nbr[f]], ld(Cf, f));
        nCellFaces[nbr[f]]++;
    }
    for (int c = 0; c < nCells; ++c) cEst[c] = vdiv(cEst[c], (double)nCellFaces[c]);
    std::vector<V3> cc(nCells, V3{0, 0, 0});
    std::vector<double> vol(nCells, 0.0);
    for (int f = 0; f < nFaces; ++f) {
        int o = owner[f];
        V3 fc = ld(Cf, f);
        double pyr3Vol = vdot(ld(Sf, f), vsub(fc, cEst[o]));
        V3 pc = vadd(vmul(0.75, fc), vmul(0.25, cEst[o]));
        cc[o] = vadd(cc[o], vmul(pyr3Vol, pc));
        vol[o] += pyr3Vol;
    }
    for (int f = 0; f < nInternal; ++f) {
        int nb = nbr[f];
        V3 fc = ld(Cf, f);
        double pyr3Vol = vdot(ld(Sf, f), vsub(cEst[nb], fc));
        V3 pc = vadd(vmul(0.75, fc), vmul(0.25, cEst[nb]));
        cc[nb] = vadd(cc[nb], vmul(pyr3Vol, pc));
        vol[nb] += pyr3Vol;
    }
    for (int c = 0; c < nCells; ++c) {
        if (std::fabs(vol[c]) > VSMALL)
            st(C, c, vdiv(cc[c], vol[c]));
        else
            st(C, c, cEst[c]);
        V[c] = vol[c] * (1.0 / 3.0);
    }
    return 0;
}

// internal-face interpolation weights and delta coefficients
int orc_internal_coeffs(int nInternal, const int32_t* own, const int32_t* nbr, const double* C,
                        const double* Cf, const double* Sf, double* magSf, double* w,
                        double* dc, double* nodc) {
    for (int f = 0; f < nInternal; ++f) {
        V3 sf = ld(Sf, f), cf = ld(Cf, f), cp = ld(C, own[f]), cn = ld(C, nbr[f]);
        double m = vmag(sf);
        magSf[f] = m;
        double sfdOwn = std::fabs(vdot(sf, vsub(cf, cp)));
        double sfdNei = std::fabs(vdot(sf, vsub(cn, cf)));
        w[f] = (sfdOwn + sfdNei > VSMALL) ? sfdNei / (sfdOwn + sfdNei) : 0.5;
        V3 d = vsub(cn, cp);
        dc[f] = 1.0 / vmag(d);
        V3 unitArea = vdiv(sf, m);
        nodc[f] = 1.0 / std::max(vdot(unitArea, d), 0.05 * vmag(d));
    }
    return 0;
}

// boundary faces: nf, patch-normal delta, weights (1), deltaCoeffs = 1/|delta|,
// nonOrthDeltaCoeffs = 1/max(nf.delta, 0.05|delta|)  (fvPatch::delta, fvPatch::nf)
int orc_boundary_coeffs(int nB, const int32_t* faceCells, const double* C, const double* Cfb,
                        const double* Sfb, double* magSfb, double* nfb, double* deltab,
                        double* wb, double* dcb, double* nodcb) {
    for (int b = 0; b < nB; ++b) {
        V3 sf = ld(Sfb, b);
        double m = vmag(sf);
        magSfb[b] = m;
        V3 nf = vdiv(sf, m);
        st(nfb, b, nf);
        V3 d = vsub(ld(Cfb, b), ld(C, faceCells[b]));
        V3 delta = vmul(vdot(nf, d), nf);
        st(deltab, b, delta);
        wb[b] = 1.0;
        dcb[b] = 1.0 / vmag(delta);
        nodcb[b] = 1.0 / std::max(vdot(nf, delta), 0.05 * vmag(delta));
    }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// explicit face operators, templated on the storage/compute precision
// ---------------------------------------------------------------------------
namespace {

template <typename S>
void interp(int scheme, int nInt, int nB, const int32_t* own, const int32_t* nbr, const S* w,
            const S* phi, const S* T, const S* Tb, S* Tf) {
    // surfaceInterpolationScheme::interpolate: lambda*(vf[P]-vf[N]) + vf[N];
    // upwind lambda = pos0(phi)
    for (int f = 0; f < nInt; ++f) {
        S lam = scheme == 0 ? w[f] : (phi[f] >= S(0) ? S(1) : S(0));
        Tf[f] = lam * (T[own[f]] - T[nbr[f]]) + T[nbr[f]];
    }
    for (int b = 0; b < nB; ++b) Tf[nInt + b] = Tb[b];
}

template <typename S>
void grad(int nCells, int nInt, int nB, const int32_t* own, const int32_t* nbr,
          const int32_t* faceCells, const S* w, const S* Sf, const S* Sfb, const S* V,
          const S* T, const S* Tb, S* out) {
    std::vector<S> acc(3 * (size_t)nCells, S(0));
    for (int f = 0; f < nInt; ++f) {
        S tf = w[f] * (T[own[f]] - T[nbr[f]]) + T[nbr[f]];
        for (int k = 0; k < 3; ++k) {
            S v = Sf[3 * f + k] * tf;
            acc[3 * (size_t)own[f] + k] += v;
            acc[3 * (size_t)nbr[f] + k] -= v;
        }
    }
    for (int b = 0; b < nB; ++b)
        for (int k = 0; k < 3; ++k) acc[3 * (size_t)faceCells[b] + k] += Sfb[3 * b + k] * Tb[b];
    for (int c = 0; c < nCells; ++c)
        for (int k = 0; k < 3; ++k) out[3 * (size_t)c + k] = acc[3 * (size_t)c + k] / V[c];
}

template <typename S>
void div(int scheme, int nCells, int nInt, int nB, const int32_t* own, const int32_t* nbr,
         const int32_t* faceCells, const S* w, const S* V, const S* phi, const S* T,
         const S* Tb, S* out) {
    // gaussConvectionScheme::fvcDiv = surfaceIntegrate(phi*interpolate(T))
    std::vector<S> acc(nCells, S(0));
    for (int f = 0; f < nInt; ++f) {
        S lam = scheme == 0 ? w[f] : (phi[f] >= S(0) ? S(1) : S(0));
        S tf = lam * (T[own[f]] - T[nbr[f]]) + T[nbr[f]];
        S v = phi[f] * tf;
        acc[own[f]] += v;
        acc[nbr[f]] -= v;
    }
    for (int b = 0; b < nB; ++b) acc[faceCells[b]] += phi[nInt + b] * Tb[b];
    for (int c = 0; c < nCells; ++c) out[c] = acc[c] / V[c];
}

template <typename S>
void sngrad(int nInt, int nB, int nPatches, const int32_t* poff, const int32_t* pkind,
            const int32_t* own, const int32_t* nbr, const int32_t* faceCells, const S* nodc,
            const S* dcb, const S* T, const S* Tb, const S* Tg, S* out) {
    // snGradScheme::snGrad: deltaCoeffs*(vf[N]-vf[P]); boundary pvf.snGrad()
    for (int f = 0; f < nInt; ++f) out[f] = nodc[f] * (T[nbr[f]] - T[own[f]]);
    for (int b = 0; b < nB; ++b) {
        int k = pkind[patch_of(b, nPatches, poff)];
        out[nInt + b] = bnd_sngrad<S>(k, dcb[b], Tb[b], T[faceCells[b]], Tg ? Tg[b] : S(0));
    }
}

template <typename S>
void laplacian(int nCells, int nInt, int nB, int nPatches, const int32_t* poff,
               const int32_t* pkind, const int32_t* own, const int32_t* nbr,
               const int32_t* faceCells, const S* magSf, const S* nodc, const S* magSfb,
               const S* dcb, const S* V, const S* gamma, const S* T, const S* Tb, const S* Tg,
               S* out) {
    // gaussLaplacianScheme::fvcLaplacian = surfaceIntegrate((gamma*magSf)*snGrad(T))
    std::vector<S> acc(nCells, S(0));
    for (int f = 0; f < nInt; ++f) {
        S sn = nodc[f] * (T[nbr[f]] - T[own[f]]);
        S v = (gamma[f] * magSf[f]) * sn;
        acc[own[f]] += v;
        acc[nbr[f]] -= v;
    }
    for (int b = 0; b < nB; ++b) {
        int k = pkind[patch_of(b, nPatches, poff)];
        S sn = bnd_sngrad<S>(k, dcb[b], Tb[b], T[faceCells[b]], Tg ? Tg[b] : S(0));
        acc[faceCells[b]] += (gamma[nInt + b] * magSfb[b]) * sn;
    }
    for (int c = 0; c < nCells; ++c) out[c] = acc[c] / V[c];
}

}  // namespace

extern "C" {

#define ORC_OPS(S, SUF)                                                                          \
    int orc_interp_##SUF(int scheme, int nInt, int nB, const int32_t* own, const int32_t* nbr,   \
                         const S* w, const S* phi, const S* T, const S* Tb, S* Tf) {             \
        interp<S>(scheme, nInt, nB, own, nbr, w, phi, T, Tb, Tf);                                \
        return 0;                                                                                \
    }                                                                                            \
    int orc_grad_##SUF(int nCells, int nInt, int nB, const int32_t* own, const int32_t* nbr,     \
                       const int32_t* fc, const S* w, const S* Sf, const S* Sfb, const S* V,     \
                       const S* T, const S* Tb, S* out) {                                        \
        grad<S>(nCells, nInt, nB, own, nbr, fc, w, Sf, Sfb, V, T, Tb, out);                      \
        return 0;                                                                                \
    }                                                                                            \
    int orc_div_##SUF(int scheme, int nCells, int nInt, int nB, const int32_t* own,              \
                      const int32_t* nbr, const int32_t* fc, const S* w, const S* V,             \
                      const S* phi, const S* T, const S* Tb, S* out) {                           \
        div<S>(scheme, nCells, nInt, nB, own, nbr, fc, w, V, phi, T, Tb, out);                   \
        return 0;                                                                                \
    }                                                                                            \
    int orc_sngrad_##SUF(int nInt, int nB, int nPatches, const int32_t* poff,                    \
                         const int32_t* pkind, const int32_t* own, const int32_t* nbr,           \
                         const int32_t* fc, const S* nodc, const S* dcb, const S* T,             \
                         const S* Tb, const S* Tg, S* out) {                                     \
        sngrad<S>(nInt, nB, nPatches, poff, pkind, own, nbr, fc, nodc, dcb, T, Tb, Tg, out);     \
        return 0;                                                                                \
    }                                                                                            \
    int orc_laplacian_##SUF(int nCells, int nInt, int nB, int nPatches, const int32_t* poff,     \
                            const int32_t* pkind, const int32_t* own, const int32_t* nbr,        \
                            const int32_t* fc, const S* magSf, const S* nodc, const S* magSfb,   \
                            const S* dcb, const S* V, const S* gamma, const S* T, const S* Tb,   \
                            const S* Tg, S* out) {                                               \
        laplacian<S>(nCells, nInt, nB, nPatches, poff, pkind, own, nbr, fc, magSf, nodc, magSfb, \
                     dcb, V, gamma, T, Tb, Tg, out);                                             \
        return 0;                                                                                \
    }

ORC_OPS(double, f64)
ORC_OPS(float, f32)

// ---------------------------------------------------------------------------
// sparsity pattern (la::SparsityPattern): CSR of the face graph with ascending
// columns per row, diagOffset[N], ownerOffset[F] (slot of column nbr in row own),
// neighbourOffset[F] (slot of column own in row nbr)
// usage: pressureVelocityCoupling.cpp:87-107, expression.hpp:95-107
// ---------------------------------------------------------------------------
int orc_sparsity(int nCells, int nInt, const int32_t* own, const int32_t* nbr, int32_t* rowOffs,
                 int32_t* colIdxs, int32_t* diagOffset, int32_t* ownerOffset,
                 int32_t* neighbourOffset) {
    std::vector<int32_t> nLower(nCells, 0), nUpper(nCells, 0);
    for (int f = 0; f < nInt; ++f) {
        nUpper[own[f]]++;
        nLower[nbr[f]]++;
    }
    rowOffs[0] = 0;
    for (int c = 0; c < nCells; ++c) rowOffs[c + 1] = rowOffs[c] + nLower[c] + 1 + nUpper[c];
    std::vector<int32_t> lowFill(nCells, 0), upFill(nCells, 0);
    for (int c = 0; c < nCells; ++c) {
        diagOffset[c] = nLower[c];
        colIdxs[rowOffs[c] + nLower[c]] = c;
    }
    // faces are ordered by owner then neighbour, so within a row the lower
    // neighbours (faces where the row is the neighbour) arrive in ascending
    // owner order and the upper neighbours in ascending neighbour order
    for (int f = 0; f < nInt; ++f) {
        int o = own[f], n = nbr[f];
        int so = nLower[o] + 1 + upFill[o]++;
        colIdxs[rowOffs[o] + so] = n;
        ownerOffset[f] = so;
        int sn = lowFill[n]++;
        colIdxs[rowOffs[n] + sn] = o;
        neighbourOffset[f] = sn;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// implicit assembly into values[nnz] / rhs[N] with the convention A.x - b =
// (operator applied to x) * V  (test_implicitOperators.cpp:92-100,254-262)
// ---------------------------------------------------------------------------

// EulerDdtScheme::fvmDdt: diag = rDeltaT*V, source = rDeltaT*T_old*V
int orc_assemble_ddt(int nCells, const int32_t* rowOffs, const int32_t* diagOffset,
                     const double* V, const double* Told, double dt, double coeff,
                     double* values, double* rhs) {
    const double rdt = 1.0 / dt;
    for (int c = 0; c < nCells; ++c) {
        values[rowOffs[c] + diagOffset[c]] += coeff * (rdt * V[c]);
        rhs[c] += coeff * ((rdt * Told[c]) * V[c]);
    }
    return 0;
}

// SourceTerm implicit: diag += coeff*c_P*V_P (test_implicitOperators.cpp:141)
int orc_assemble_source(int nCells, const int32_t* rowOffs, const int32_t* diagOffset,
                        const double* V, const double* cfield, double coeff, double* values) {
    for (int c = 0; c < nCells; ++c) values[rowOffs[c] + diagOffset[c]] += coeff * (cfield[c] * V[c]);
    return 0;
}

// fvmLaplacianUncorrected: upper = nonOrthDeltaCoeffs*(gamma*magSf), negSumDiag,
// boundary folded (fixedValue: diag -= g|Sf|dc, b -= g|Sf|dc*Tb;
// fixedGradient: b -= g|Sf|*grad)
int orc_assemble_laplacian(int nCells, int nInt, int nB, int nPatches, const int32_t* poff,
                           const int32_t* pkind, const int32_t* own, const int32_t* nbr,
                           const int32_t* faceCells, const double* magSf, const double* nodc,
                           const double* magSfb, const double* dcb, const double* gamma,
                           const double* Tb, const double* Tg, const int32_t* rowOffs,
                           const int32_t* diagOffset, const int32_t* ownerOffset,
                           const int32_t* neighbourOffset, double coeff, double* values,
                           double* rhs) {
    std::vector<double> dsum(nCells, 0.0), bsum(nCells, 0.0);
    for (int f = 0; f < nInt; ++f) {
        double up = nodc[f] * (gamma[f] * magSf[f]);
        values[rowOffs[own[f]] + ownerOffset[f]] += coeff * up;
        values[rowOffs[nbr[f]] + neighbourOffset[f]] += coeff * up;
        dsum[own[f]] -= up;
        dsum[nbr[f]] -= up;
    }
    for (int b = 0; b < nB; ++b) {
        int k = pkind[patch_of(b, nPatches, poff)];
        double gm = gamma[nInt + b] * magSfb[b];
        int c = faceCells[b];
        if (k == BC_FIXED_GRADIENT) {
            bsum[c] -= gm * (Tg ? Tg[b] : 0.0);
        } else {
            double ic = gm * dcb[b];
            dsum[c] -= ic;
            bsum[c] -= ic * Tb[b];
        }
    }
    for (int c = 0; c < nCells; ++c) {
        values[rowOffs[c] + diagOffset[c]] += coeff * dsum[c];
        rhs[c] += coeff * bsum[c];
    }
    return 0;
}

// gaussConvectionScheme::fvmDiv: lower = -w*phi, upper = lower + phi, negSumDiag;
// boundary: fixedValue b -= phi_b*Tb; fixedGradient diag += phi_b,
// b -= phi_b*grad/deltaCoeffs
int orc_assemble_div(int scheme, int nCells, int nInt, int nB, int nPatches, const int32_t* poff,
                     const int32_t* pkind, const int32_t* own, const int32_t* nbr,
                     const int32_t* faceCells, const double* w, const double* dcb,
                     const double* phi, const double* Tb, const double* Tg,
                     const int32_t* rowOffs, const int32_t* diagOffset,
                     const int32_t* ownerOffset, const int32_t* neighbourOffset, double coeff,
                     double* values, double* rhs) {
    std::vector<double> dsum(nCells, 0.0), bsum(nCells, 0.0);
    for (int f = 0; f < nInt; ++f) {
        double lam = scheme == 0 ? w[f] : (phi[f] >= 0.0 ? 1.0 : 0.0);
        double lower = -lam * phi[f];
        double upper = lower + phi[f];
        values[rowOffs[own[f]] + ownerOffset[f]] += coeff * upper;
        values[rowOffs[nbr[f]] + neighbourOffset[f]] += coeff * lower;
        dsum[own[f]] -= lower;
        dsum[nbr[f]] -= upper;
    }
    for (int b = 0; b < nB; ++b) {
        int k = pkind[patch_of(b, nPatches, poff)];
        double pf = phi[nInt + b];
        int c = faceCells[b];
        if (k == BC_FIXED_GRADIENT) {
            dsum[c] += pf;
            bsum[c] -= pf * ((Tg ? Tg[b] : 0.0) / dcb[b]);
        } else {
            bsum[c] -= pf * Tb[b];
        }
    }
    for (int c = 0; c < nCells; ++c) {
        values[rowOffs[c] + diagOffset[c]] += coeff * dsum[c];
        rhs[c] += coeff * bsum[c];
    }
    return 0;
}

// PDESolver::SetReference (expression.hpp:102-111)
int orc_set_reference(const int32_t* rowOffs, const int32_t* diagOffset, int cell, double value,
                      double* values, double* rhs) {
    int d = rowOffs[cell] + diagOffset[cell];
    double dv = values[d];
    rhs[cell] += dv * value;
    values[d] += dv;
    return 0;
}

// la::computeResidual: res = A.x - b (expression.hpp:204-220)
int orc_residual(int nCells, const int32_t* rowOffs, const int32_t* colIdxs, const double* values,
                 const double* rhs, const double* x, double* res) {
    for (int r = 0; r < nCells; ++r) {
        double s = 0.0;
        for (int k = rowOffs[r]; k < rowOffs[r + 1]; ++k) s += values[k] * x[colIdxs[k]];
        res[r] = s - rhs[r];
    }
    return 0;
}

// Jacobi-preconditioned CG, stop when ||r||/||b|| <= relTol or iter >= maxIter
// (fvSolution.cpp:22-28,50-55,106-142: PCG -> Cg, DIC -> point Jacobi).
// stats = {initial rel. residual, final rel. residual (recurrence), iterations}
int orc_cg(int nCells, const int32_t* rowOffs, const int32_t* colIdxs, const double* values,
           const double* b, double* x, double relTol, int maxIter, double* stats) {
    const int n = nCells;
    std::vector<double> r(n), z(n), p(n), q(n), dinv(n);
    auto spmv = [&](const double* in, double* out) {
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int k = rowOffs[i]; k < rowOffs[i + 1]; ++k) s += values[k] * in[colIdxs[k]];
            out[i] = s;
        }
    };
    auto dot = [&](const double* a, const double* c) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += a[i] * c[i];
        return s;
    };
    for (int i = 0; i < n; ++i) {
        double d = 0.0;
        for (int k = rowOffs[i]; k < rowOffs[i + 1]; ++k)
            if (colIdxs[k] == i) d = values[k];
        dinv[i] = 1.0 / d;
    }
    spmv(x, q.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] - q[i];
    double bnorm = std::sqrt(dot(b, b));
    if (bnorm == 0.0) bnorm = 1.0;
    double res = std::sqrt(dot(r.data(), r.data())) / bnorm;
    stats[0] = res;
    int it = 0;
    if (res > relTol) {
        for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
        for (int i = 0; i < n; ++i) p[i] = z[i];
        double rz = dot(r.data(), z.data());
        while (it < maxIter) {
            spmv(p.data(), q.data());
            double alpha = rz / dot(p.data(), q.data());
            for (int i = 0; i < n; ++i) x[i] += alpha * p[i];
            for (int i = 0; i < n; ++i) r[i] -= alpha * q[i];
            ++it;
            res = std::sqrt(dot(r.data(), r.data())) / bnorm;
            if (res <= relTol) break;
            for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
            double rzn = dot(r.data(), z.data());
            double beta = rzn / rz;
            rz = rzn;
            for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        }
    }
    stats[1] = res;
    stats[2] = it;
    return res <= relTol ? 0 : 1;
}

// Boundary coefficients of coeff * laplacian(gamma, T) per boundary face, the
// face flux being bRhs - bMat * T_P (NeoN's la::BoundaryCoefficients as read
// by pressureVelocityCoupling.cpp:180-192): fixedValue bMat = coeff*gamma|Sf|dc,
// bRhs = coeff*gamma|Sf|dc*Tb; fixedGradient bMat = 0, bRhs = coeff*gamma|Sf|*g.
int orc_laplacian_bcoeffs(int nInt, int nB, int nPatches, const int32_t* poff,
                          const int32_t* pkind, const double* magSfb, const double* dcb,
                          const double* gamma, const double* Tb, const double* Tg, double coeff,
                          double* bMat, double* bRhs) {
    for (int b = 0; b < nB; ++b) {
        int k = pkind[patch_of(b, nPatches, poff)];
        double gm = gamma[nInt + b] * magSfb[b];
        if (k == BC_FIXED_GRADIENT) {
            bMat[b] += 0.0;
            bRhs[b] += coeff * (gm * (Tg ? Tg[b] : 0.0));
        } else {
            double ic = gm * dcb[b];
            bMat[b] += coeff * ic;
            bRhs[b] += coeff * (ic * Tb[b]);
        }
    }
    return 0;
}

// computeRAU (pressureVelocityCoupling.cpp:38-61): rAU = V / diag; boundary
// values extrapolated (rAU_b = rAU_P)
int orc_rau(int nCells, int nB, const int32_t* rowOffs, const int32_t* diagOffset,
            const double* values, const double* V, const int32_t* faceCells, double* rau,
            double* raub) {
    for (int c = 0; c < nCells; ++c) rau[c] = V[c] / values[rowOffs[c] + diagOffset[c]];
    for (int b = 0; b < nB; ++b) raub[b] = rau[faceCells[b]];
    return 0;
}

// computeRAUandHByA (pressureVelocityCoupling.cpp:63-119), with OpenFOAM's serial
// fvMatrix::H face loop in place of the atomics: H = 0; per face
// H[nbr] -= A(nbr,own) U[own], H[own] -= A(own,nbr) U[nbr]; then
// HbyA = (H + b) * (rAU / V); boundary values extrapolated.
int orc_hbya(int nCells, int nInt, int nB, const int32_t* own, const int32_t* nbr,
             const int32_t* rowOffs, const int32_t* ownerOffset, const int32_t* neighbourOffset,
             const double* values, const double* b3, const double* U, const double* rau,
             const double* V, const int32_t* faceCells, double* hbya, double* hbyab) {
    for (int i = 0; i < 3 * nCells; ++i) hbya[i] = 0.0;
    for (int f = 0; f < nInt; ++f) {
        const int o = own[f], n = nbr[f];
        const double lower = values[rowOffs[n] + neighbourOffset[f]];
        const double upper = values[rowOffs[o] + ownerOffset[f]];
        for (int j = 0; j < 3; ++j) {
            hbya[3 * n + j] -= lower * U[3 * o + j];
            hbya[3 * o + j] -= upper * U[3 * n + j];
        }
    }
    for (int c = 0; c < nCells; ++c) {
        const double s = rau[c] / V[c];
        for (int j = 0; j < 3; ++j) {
            hbya[3 * c + j] += b3[3 * c + j];
            hbya[3 * c + j] *= s;
        }
    }
    for (int b = 0; b < nB; ++b)
        for (int j = 0; j < 3; ++j) hbyab[3 * b + j] = hbya[3 * faceCells[b] + j];
    return 0;
}

// updateFaceVelocity (pressureVelocityCoupling.cpp:122-195):
// phi_f = pred_f - (A(nbr,own) p_N - A(own,nbr) p_P); boundary
// phi_b = pred_b - (bRhs - bMat p_P)
int orc_update_face_velocity(int nInt, int nB, const int32_t* own, const int32_t* nbr,
                             const int32_t* rowOffs, const int32_t* ownerOffset,
                             const int32_t* neighbourOffset, const double* values,
                             const int32_t* faceCells, const double* bMat, const double* bRhs,
                             const double* pred, const double* p, double* phi) {
    for (int f = 0; f < nInt; ++f) {
        const int o = own[f], n = nbr[f];
        const double upper = values[rowOffs[n] + neighbourOffset[f]];
        const double lower = values[rowOffs[o] + ownerOffset[f]];
        phi[f] = pred[f] - (upper * p[n] - lower * p[o]);
    }
    for (int b = 0; b < nB; ++b) {
        const double bflux = bRhs[b] - bMat[b] * p[faceCells[b]];
        phi[nInt + b] = pred[nInt + b] - bflux;
    }
    return 0;
}

// updateVelocity (pressureVelocityCoupling.cpp:197-211): U = HbyA - rAU grad(p)
int orc_update_velocity(int nCells, const double* hbya, const double* rau, const double* gradp,
                        double* U) {
    for (int c = 0; c < nCells; ++c)
        for (int j = 0; j < 3; ++j) U[3 * c + j] = hbya[3 * c + j] - rau[c] * gradp[3 * c + j];
    return 0;
}

// constrainHbyA (pressureVelocityCoupling.cpp:14-36): on patches whose velocity
// condition is not assignable (fixedValue) the boundary HbyA is the boundary U
int orc_constrain_hbya(int nB, int nPatches, const int32_t* poff, const int32_t* ukind,
                       const double* Ub, double* hbyab) {
    for (int b = 0; b < nB; ++b)
        if (ukind[patch_of(b, nPatches, poff)] == BC_FIXED_VALUE)
            for (int j = 0; j < 3; ++j) hbyab[3 * b + j] = Ub[3 * b + j];
    return 0;
}

// Courant number (OpenFOAM CourantNo.H): 0.5 * max_c(sum over the cell's faces
// of |phi| / V_c) * dt, internal and boundary faces, summed in face order
double orc_courant(int nCells, int nInt, int nB, const int32_t* own, const int32_t* nbr,
                   const int32_t* faceCells, const double* phi, const double* V, double dt) {
    std::vector<double> s(nCells, 0.0);
    for (int f = 0; f < nInt; ++f) {
        s[own[f]] += std::fabs(phi[f]);
        s[nbr[f]] += std::fabs(phi[f]);
    }
    for (int b = 0; b < nB; ++b) s[faceCells[b]] += std::fabs(phi[nInt + b]);
    double mx = 0.0;
    for (int c = 0; c < nCells; ++c) mx = std::max(mx, s[c] / V[c]);
    return 0.5 * mx * dt;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Multi-core restatement (the CPU baseline of the benchmark, the role of
// NeoN's OpenMP executor).  Cell-result operators walk each cell's faces in
// ascending face order -- internal faces in face order with the owner (+) /
// neighbour (-) sign, then the cell's boundary faces -- which is the order in
// which the serial face loops above accumulate into that cell, so any thread
// count returns the serial bits (checked by the CPU tests).  Face-result
// operators split the face range.
// ---------------------------------------------------------------------------
namespace {

struct OmpStencil {
    int nCells, nInt;
    std::vector<int32_t> off, idx;  // entry: f (owner side), ~f (neighbour side), nInt + b
};

}  // namespace

extern "C" {

int orc_omp_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void* orc_omp_stencil(int nCells, int nInt, int nB, const int32_t* own, const int32_t* nbr,
                      const int32_t* faceCells) {
    auto* s = new OmpStencil{nCells, nInt, std::vector<int32_t>(nCells + 1, 0), {}};
    for (int f = 0; f < nInt; ++f) {
        ++s->off[own[f] + 1];
        ++s->off[nbr[f] + 1];
    }
    for (int b = 0; b < nB; ++b) ++s->off[faceCells[b] + 1];
    for (int c = 0; c < nCells; ++c) s->off[c + 1] += s->off[c];
    s->idx.resize(s->off[nCells]);
    std::vector<int32_t> pos(s->off.begin(), s->off.end() - 1);
    for (int f = 0; f < nInt; ++f) {
        s->idx[pos[own[f]]++] = f;
        s->idx[pos[nbr[f]]++] = ~f;
    }
    for (int b = 0; b < nB; ++b) s->idx[pos[faceCells[b]]++] = nInt + b;
    return s;
}

void orc_omp_free(void* s) { delete static_cast<OmpStencil*>(s); }

int orc_omp_interp(int nInt, int nB, const int32_t* own, const int32_t* nbr, const double* w,
                   const double* T, const double* Tb, double* Tf) {
#pragma omp parallel for schedule(static)
    for (int f = 0; f < nInt; ++f) Tf[f] = w[f] * (T[own[f]] - T[nbr[f]]) + T[nbr[f]];
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nB; ++b) Tf[nInt + b] = Tb[b];
    return 0;
}

int orc_omp_sngrad(int nInt, int nB, int nPatches, const int32_t* poff, const int32_t* pkind,
                   const int32_t* own, const int32_t* nbr, const int32_t* faceCells,
                   const double* nodc, const double* dcb, const double* T, const double* Tb,
                   double* out) {
#pragma omp parallel for schedule(static)
    for (int f = 0; f < nInt; ++f) out[f] = nodc[f] * (T[nbr[f]] - T[own[f]]);
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nB; ++b) {
        const int k = pkind[patch_of(b, nPatches, poff)];
        out[nInt + b] = bnd_sngrad<double>(k, dcb[b], Tb[b], T[faceCells[b]], 0.0);
    }
    return 0;
}

int orc_omp_grad(const void* sp, const int32_t* own, const int32_t* nbr, const double* w,
                 const double* Sf, const double* Sfb, const double* V, const double* T,
                 const double* Tb, double* out) {
    const auto& s = *static_cast<const OmpStencil*>(sp);
    const int F = s.nInt;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < s.nCells; ++c) {
        double a[3] = {0.0, 0.0, 0.0};
        for (int k = s.off[c]; k < s.off[c + 1]; ++k) {
            const int e = s.idx[k];
            if (e >= F) {
                const int b = e - F;
                for (int j = 0; j < 3; ++j) a[j] += Sfb[3 * b + j] * Tb[b];
            } else {
                const int f = e >= 0 ? e : ~e;
                const double tf = w[f] * (T[own[f]] - T[nbr[f]]) + T[nbr[f]];
                for (int j = 0; j < 3; ++j) {
                    const double v = Sf[3 * (size_t)f + j] * tf;
                    if (e >= 0)
                        a[j] += v;
                    else
                        a[j] -= v;
                }
            }
        }
        for (int j = 0; j < 3; ++j) out[3 * (size_t)c + j] = a[j] / V[c];
    }
    return 0;
}

int orc_omp_div(const void* sp, int scheme, const int32_t* own, const int32_t* nbr,
                const double* w, const double* V, const double* phi, const double* T,
                const double* Tb, double* out) {
    const auto& s = *static_cast<const OmpStencil*>(sp);
    const int F = s.nInt;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < s.nCells; ++c) {
        double a = 0.0;
        for (int k = s.off[c]; k < s.off[c + 1]; ++k) {
            const int e = s.idx[k];
            if (e >= F) {
                a += phi[e] * Tb[e - F];
            } else {
                const int f = e >= 0 ? e : ~e;
                const double lam = scheme == 0 ? w[f] : (phi[f] >= 0.0 ? 1.0 : 0.0);
                const double v = phi[f] * (lam * (T[own[f]] - T[nbr[f]]) + T[nbr[f]]);
                if (e >= 0)
                    a += v;
                else
                    a -= v;
            }
        }
        out[c] = a / V[c];
    }
    return 0;
}

int orc_omp_laplacian(const void* sp, int nPatches, const int32_t* poff, const int32_t* pkind,
                      const int32_t* own, const int32_t* nbr, const double* magSf,
                      const double* nodc, const double* magSfb, const double* dcb,
                      const double* V, const double* gamma, const double* T, const double* Tb,
                      double* out) {
    const auto& s = *static_cast<const OmpStencil*>(sp);
    const int F = s.nInt;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < s.nCells; ++c) {
        double a = 0.0;
        for (int k = s.off[c]; k < s.off[c + 1]; ++k) {
            const int e = s.idx[k];
            if (e >= F) {
                const int b = e - F;
                const int kk = pkind[patch_of(b, nPatches, poff)];
                const double sn = bnd_sngrad<double>(kk, dcb[b], Tb[b], T[c], 0.0);
                a += (gamma[e] * magSfb[b]) * sn;
            } else {
                const int f = e >= 0 ? e : ~e;
                const double v = (gamma[f] * magSf[f]) * (nodc[f] * (T[nbr[f]] - T[own[f]]));
                if (e >= 0)
                    a += v;
                else
                    a -= v;
            }
        }
        out[c] = a / V[c];
    }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Multi-core Jacobi-PCG: the algorithm of orc_cg (same recurrence, same stop
// test) with the row loops split over OpenMP threads.  Dot products sum fixed
// blocks of kDotBlock rows serially and then the block sums in block order, so
// the result is the same bit for bit at any thread count (it differs from the
// serial orc_cg in the summation order of the dot products only).  This is the
// implicit solves' CPU baseline (the role of Ginkgo's OpenMP executor).
// ---------------------------------------------------------------------------
namespace {
constexpr int kDotBlock = 8192;
}

extern "C" {

int orc_omp_cg(int nCells, const int32_t* rowOffs, const int32_t* colIdxs, const double* values,
               const double* b, double* x, double relTol, int maxIter, double* stats) {
    const int n = nCells;
    const int nb = (n + kDotBlock - 1) / kDotBlock;
    std::vector<double> r(n), z(n), p(n), q(n), dinv(n), part(nb);
    auto spmv = [&](const double* in, double* out) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int k = rowOffs[i]; k < rowOffs[i + 1]; ++k) s += values[k] * in[colIdxs[k]];
            out[i] = s;
        }
    };
    auto dot = [&](const double* a, const double* c) {
#pragma omp parallel for schedule(static)
        for (int j = 0; j < nb; ++j) {
            double s = 0.0;
            const int i1 = std::min(n, (j + 1) * kDotBlock);
            for (int i = j * kDotBlock; i < i1; ++i) s += a[i] * c[i];
            part[j] = s;
        }
        double s = 0.0;
        for (int j = 0; j < nb; ++j) s += part[j];
        return s;
    };
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double d = 0.0;
        for (int k = rowOffs[i]; k < rowOffs[i + 1]; ++k)
            if (colIdxs[k] == i) d = values[k];
        dinv[i] = 1.0 / d;
    }
    spmv(x, q.data());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) r[i] = b[i] - q[i];
    double bnorm = std::sqrt(dot(b, b));
    if (bnorm == 0.0) bnorm = 1.0;
    double res = std::sqrt(dot(r.data(), r.data())) / bnorm;
    stats[0] = res;
    int it = 0;
    if (res > relTol) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            z[i] = dinv[i] * r[i];
            p[i] = z[i];
        }
        double rz = dot(r.data(), z.data());
        while (it < maxIter) {
            spmv(p.data(), q.data());
            const double alpha = rz / dot(p.data(), q.data());
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) {
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
            }
            ++it;
            res = std::sqrt(dot(r.data(), r.data())) / bnorm;
            if (res <= relTol) break;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
            const double rzn = dot(r.data(), z.data());
            const double beta = rzn / rz;
            rz = rzn;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        }
    }
    stats[1] = res;
    stats[2] = it;
    return res <= relTol ? 0 : 1;
}

// Sets the OpenMP thread count of the multi-core restatement (0: leave it).
int orc_omp_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

}  // extern "C"
