// mesh.hip -- device mesh construction.
//
// fv_mesh_create uploads NeoN-layout geometry (what readOpenFOAMMesh builds,
// meshAdapter.cpp:59-136).  fv_mesh_create_poly computes that geometry on the
// GPU from polyMesh primitives with OpenFOAM's algorithms:
//   face centre/area: triangle fan about the point average, area-weighted
//     (primitiveMeshFaceCentresAndAreas);
//   cell centre/volume: pyramids on each face about the face-centre average
//     (primitiveMeshCellCentresAndVols);
//   weights w = |Sf.(C_N-Cf)| / (|Sf.(Cf-C_P)| + |Sf.(C_N-Cf)|), deltaCoeffs =
//     1/|C_N-C_P|, nonOrthDeltaCoeffs = 1/max(n.d, 0.05|d|) (surfaceInterpolation);
//   boundary: nf = Sf/|Sf|, delta = nf (nf.(Cf-C_P)), w = 1, deltaCoeffs = 1/|delta|.
// Each thread follows the serial per-cell accumulation order of the CPU
// algorithm (owned faces ascending, then neighbour faces ascending), compiled
// without FMA contraction, so the results are the serial ones bit for bit.
#include "mesh.hpp"

#include <algorithm>
#include <cmath>

using namespace fvhip;

namespace {

struct V3 { double x, y, z; };
__device__ inline V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 vmul(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ inline V3 vdiv(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ inline double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 vcross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline double vmag(V3 a) { return sqrt(vdot(a, a)); }
__device__ inline V3 ld3(const double* p, int64_t i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ inline void st3(double* p, int64_t i, V3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}

__global__ void k_face_geom(int nFaces, const int32_t* __restrict__ fOff,
                            const int32_t* __restrict__ fPts, const double* __restrict__ pts,
                            double* __restrict__ Cf, double* __restrict__ Sf) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nFaces) return;
    const int b = fOff[f], n = fOff[f + 1] - b;
    const int32_t* v = fPts + b;
    if (n == 3) {
        V3 p0 = ld3(pts, v[0]), p1 = ld3(pts, v[1]), p2 = ld3(pts, v[2]);
        st3(Cf, f, vmul(1.0 / 3.0, vadd(vadd(p0, p1), p2)));
        st3(Sf, f, vmul(0.5, vcross(vsub(p1, p0), vsub(p2, p0))));
        return;
    }
    V3 fc = ld3(pts, v[0]);
    for (int i = 1; i < n; ++i) fc = vadd(fc, ld3(pts, v[i]));
    fc = vdiv(fc, (double)n);
    V3 sumN{0, 0, 0}, sumAc{0, 0, 0};
    double sumA = 0.0;
    for (int i = 0; i < n; ++i) {
        V3 a = ld3(pts, v[i]);
        V3 c2 = ld3(pts, v[i + 1 == n ? 0 : i + 1]);
        V3 c = vadd(vadd(a, c2), fc);
        V3 nn = vcross(vsub(c2, a), vsub(fc, a));
        double m = vmag(nn);
        sumN = vadd(sumN, nn);
        sumA += m;
        sumAc = vadd(sumAc, vmul(m, c));
    }
    if (sumA < 1.0e-150) {
        st3(Cf, f, fc);
        st3(Sf, f, V3{0, 0, 0});
    } else {
        st3(Cf, f, vdiv(vmul(1.0 / 3.0, sumAc), sumA));
        st3(Sf, f, vmul(0.5, sumN));
    }
}

// geoIdx: per cell, owned faces ascending (entry f) then neighbour faces ascending (~f)
__global__ void k_cell_geom(int nCells, const int32_t* __restrict__ gOff,
                            const int32_t* __restrict__ gIdx, const double* __restrict__ Cf,
                            const double* __restrict__ Sf, double* __restrict__ C,
                            double* __restrict__ V) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nCells) return;
    const int b = gOff[c], e = gOff[c + 1];
    V3 est{0, 0, 0};
    for (int k = b; k < e; ++k) {
        int g = gIdx[k];
        int f = g >= 0 ? g : ~g;
        est = vadd(est, ld3(Cf, f));
    }
    est = vdiv(est, (double)(e - b));
    V3 cc{0, 0, 0};
    double vol = 0.0;
    for (int k = b; k < e; ++k) {
        int g = gIdx[k];
        int f = g >= 0 ? g : ~g;
        V3 fc = ld3(Cf, f);
        double pyr3 = g >= 0 ? vdot(ld3(Sf, f), vsub(fc, est)) : vdot(ld3(Sf, f), vsub(est, fc));
        V3 pc = vadd(vmul(0.75, fc), vmul(0.25, est));
        cc = vadd(cc, vmul(pyr3, pc));
        vol += pyr3;
    }
    st3(C, c, fabs(vol) > 1.0e-300 ? vdiv(cc, vol) : est);
    V[c] = vol * (1.0 / 3.0);
}

__global__ void k_internal_coeffs(int nInt, const int32_t* __restrict__ own,
                                  const int32_t* __restrict__ nbr, const double* __restrict__ C,
                                  const double* __restrict__ Cf, const double* __restrict__ Sf,
                                  double* __restrict__ magSf, double* __restrict__ w,
                                  double* __restrict__ dc, double* __restrict__ nodc) {
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nInt) return;
    V3 sf = ld3(Sf, f), cf = ld3(Cf, f), cp = ld3(C, own[f]), cn = ld3(C, nbr[f]);
    double m = vmag(sf);
    magSf[f] = m;
    double sOwn = fabs(vdot(sf, vsub(cf, cp)));
    double sNei = fabs(vdot(sf, vsub(cn, cf)));
    w[f] = (sOwn + sNei > 1.0e-300) ? sNei / (sOwn + sNei) : 0.5;
    V3 d = vsub(cn, cp);
    dc[f] = 1.0 / vmag(d);
    V3 ua = vdiv(sf, m);
    nodc[f] = 1.0 / fmax(vdot(ua, d), 0.05 * vmag(d));
}

// bFace[b] = global polyMesh face of fv boundary face b
__global__ void k_boundary_coeffs(int nB, const int32_t* __restrict__ bFace,
                                  const int32_t* __restrict__ owner, const double* __restrict__ C,
                                  const double* __restrict__ CfAll, const double* __restrict__ SfAll,
                                  int32_t* __restrict__ faceCells, double* __restrict__ Sfb,
                                  double* __restrict__ Cfb, double* __restrict__ magSfb,
                                  double* __restrict__ nfb, double* __restrict__ deltab,
                                  double* __restrict__ wb, double* __restrict__ dcb,
                                  double* __restrict__ nodcb) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nB) return;
    int g = bFace[b];
    int c = owner[g];
    faceCells[b] = c;
    V3 sf = ld3(SfAll, g), cf = ld3(CfAll, g);
    st3(Sfb, b, sf);
    st3(Cfb, b, cf);
    double m = vmag(sf);
    magSfb[b] = m;
    V3 nf = vdiv(sf, m);
    st3(nfb, b, nf);
    V3 d = vsub(cf, ld3(C, c));
    V3 delta = vmul(vdot(nf, d), nf);
    st3(deltab, b, delta);
    wb[b] = 1.0;
    dcb[b] = 1.0 / vmag(delta);
    nodcb[b] = 1.0 / fmax(vdot(nf, delta), 0.05 * vmag(delta));
}

template <typename D, typename S>
__global__ void k_cvt(D* __restrict__ d, const S* __restrict__ s, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) d[i] = (D)s[i];
}

template <typename T>
void up(fv_ctx* ctx, DBuf<T>& b, const T* h, size_t n, const char* name) {
    FV_REQUIRE(h || n == 0, std::string("mesh array ") + name + " is NULL");
    b.upload(ctx, h, n);
}

void to_f32(fv_ctx* ctx, DBuf<float>& d, const DBuf<double>& s) {
    d.alloc(ctx, s.n);
    if (!s.n) return;
    k_cvt<float, double><<<grid_for((int64_t)s.n, 256, 8192), 256, 0, ctx->stream>>>(d.p, s.p, (int64_t)s.n);
    FV_LAUNCH_CHECK();
}

}  // namespace

namespace fvhip {

void build_stencil(fv_mesh* m, const int32_t* h_own, const int32_t* h_nbr,
                   const int32_t* h_faceCells) {
    const int64_t N = m->nCells, F = m->nInt, B = m->nB;
    std::vector<int32_t> off(N + 1, 0);
    for (int64_t f = 0; f < F; ++f) {
        FV_REQUIRE(h_own[f] >= 0 && h_own[f] < N && h_nbr[f] >= 0 && h_nbr[f] < N,
                   "owner/neighbour index out of range");
        off[h_own[f] + 1]++;
        off[h_nbr[f] + 1]++;
    }
    for (int64_t b = 0; b < B; ++b) {
        FV_REQUIRE(h_faceCells[b] >= 0 && h_faceCells[b] < N, "faceCells index out of range");
        off[h_faceCells[b] + 1]++;
    }
    for (int64_t c = 0; c < N; ++c) off[c + 1] += off[c];
    std::vector<int32_t> idx(2 * F + B), oth(2 * F + B), fill(off.begin(), off.end() - 1);
    for (int64_t f = 0; f < F; ++f) {
        int32_t o = h_own[f], n = h_nbr[f];
        int32_t ko = fill[o]++, kn = fill[n]++;
        idx[ko] = (int32_t)f;
        oth[ko] = n;
        idx[kn] = ~(int32_t)f;
        oth[kn] = o;
    }
    for (int64_t b = 0; b < B; ++b) {
        int32_t c = h_faceCells[b];
        int32_t k = fill[c]++;
        idx[k] = (int32_t)(F + b);
        oth[k] = c;
    }
    m->cfOff.upload(m->ctx, off.data(), off.size());
    m->cfIdx.upload(m->ctx, idx.data(), idx.size());
    m->cfOther.upload(m->ctx, oth.data(), oth.size());
    // split stencil when the faces are in upper-triangular order
    bool upper = true;
    for (int64_t f = 0; f < F && upper; ++f) {
        if (h_own[f] >= h_nbr[f]) upper = false;
        if (f > 0 && (h_own[f] < h_own[f - 1] ||
                      (h_own[f] == h_own[f - 1] && h_nbr[f] <= h_nbr[f - 1])))
            upper = false;
    }
    m->upperTri = upper;
    if (upper) {
        std::vector<int32_t> ownStart(N + 1, 0), lowOff(N + 1, 0), bOff(N + 1, 0);
        for (int64_t f = 0; f < F; ++f) {
            ownStart[h_own[f] + 1]++;
            lowOff[h_nbr[f] + 1]++;
        }
        for (int64_t b = 0; b < B; ++b) bOff[h_faceCells[b] + 1]++;
        for (int64_t c = 0; c < N; ++c) {
            ownStart[c + 1] += ownStart[c];
            lowOff[c + 1] += lowOff[c];
            bOff[c + 1] += bOff[c];
        }
        std::vector<int32_t> lowFace(F), lowOwn(F), bFace(B);
        std::vector<int32_t> lf(lowOff.begin(), lowOff.end() - 1), bf(bOff.begin(), bOff.end() - 1);
        for (int64_t f = 0; f < F; ++f) {
            int32_t k = lf[h_nbr[f]]++;
            lowFace[k] = (int32_t)f;
            lowOwn[k] = h_own[f];
        }
        for (int64_t b = 0; b < B; ++b) bFace[bf[h_faceCells[b]]++] = (int32_t)b;
        m->ownStart.upload(m->ctx, ownStart.data(), ownStart.size());
        m->lowOff.upload(m->ctx, lowOff.data(), lowOff.size());
        m->lowFace.upload(m->ctx, lowFace.data(), lowFace.size());
        m->lowOwn.upload(m->ctx, lowOwn.data(), lowOwn.size());
        m->bOff.upload(m->ctx, bOff.data(), bOff.size());
        m->bFace.upload(m->ctx, bFace.data(), bFace.size());
    }
    FV_HIP(hipStreamSynchronize(m->ctx->stream));
}

template <typename S>
__global__ void k_pack_ext_geo(int64_t n, const int32_t* __restrict__ face, const S* __restrict__ w,
                               const S* __restrict__ magSf, const S* __restrict__ nodc,
                               const S* __restrict__ Sf, S* __restrict__ ow, S* __restrict__ om,
                               S* __restrict__ on, S* __restrict__ os) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t f = face[k];
    ow[k] = w[f];
    om[k] = magSf[f];
    on[k] = nodc[f];
    for (int j = 0; j < 3; ++j) os[3 * k + j] = Sf[3 * f + j];
}

template <typename S>
void pack_ext_geo(fv_ctx* ctx, const DBuf<int32_t>& face, int64_t n, const S* w, const S* magSf,
                  const S* nodc, const S* Sf, DBuf<S>& ow, DBuf<S>& om, DBuf<S>& on, DBuf<S>& os) {
    ow.alloc(ctx, n);
    om.alloc(ctx, n);
    on.alloc(ctx, n);
    os.alloc(ctx, 3 * n);
    if (n)
        k_pack_ext_geo<S><<<grid_for(n, 256), 256, 0, ctx->stream>>>(
            n, face.p, w, magSf, nodc, Sf, ow.p, om.p, on.p, os.p);
    FV_LAUNCH_CHECK();
    FV_HIP(hipStreamSynchronize(ctx->stream));
}

const GatherLists& gather_lists(fv_mesh* m, int cells, bool f32) {
    const int slot = cells == 256 ? 0 : cells == 512 ? 1 : 2;
    GatherLists& L = m->glists[slot];
    if (L.cells == cells) {
        if (f32 && !L.ready32) {
            ensure_f32(m);
            const MeshF32& g = m->f32;
            pack_ext_geo<float>(m->ctx, L.extFace, L.nExt, g.w.p, g.magSf.p,
                                g.nodc.p, g.Sf.p, L.exW32, L.exMagSf32, L.exNodc32, L.exSf32);
            L.ready32 = true;
        }
        return L;
    }
    FV_REQUIRE(m->upperTri, "staged gather lists need upper-triangular face order");
    const int64_t N = m->nCells, F = m->nInt, B = m->nB;
    std::vector<int32_t> lowOff(N + 1), lowFace(F), lowOwn(F), bOff(N + 1), bFace(B);
    m->lowOff.download(lowOff.data());
    m->lowFace.download(lowFace.data());
    m->lowOwn.download(lowOwn.data());
    m->bOff.download(bOff.data());
    m->bFace.download(bFace.data());
    std::vector<int32_t> ownStart(N + 1);
    m->ownStart.download(ownStart.data());
    const int64_t nBlocks = (N + cells - 1) / cells;
    std::vector<int32_t> eo(nBlocks + 1, 0), ef, ew, bo(nBlocks + 1, 0), bf;
    std::vector<uint16_t> ec, bc;
    int64_t nOverflow = 0;
    for (int64_t blk = 0; blk < nBlocks; ++blk) {
        const int64_t c0 = blk * cells, c1 = std::min<int64_t>(c0 + cells, N);
        int innerMax = 0;
        for (int64_t c = c0; c < c1; ++c) {
            // a cell's lower faces ascend; those owned before the block come first
            int32_t k = lowOff[c];
            for (; k < lowOff[c + 1] && lowOwn[k] < c0; ++k) {
                ef.push_back(lowFace[k]);
                ew.push_back(lowOwn[k]);
                ec.push_back((uint16_t)(c - c0));
            }
            innerMax = std::max(innerMax, lowOff[c + 1] - k);
            for (int32_t k = bOff[c]; k < bOff[c + 1]; ++k) {
                bf.push_back(bFace[k]);
                bc.push_back((uint16_t)(c - c0));
            }
        }
        eo[blk + 1] = (int32_t)ef.size();
        bo[blk + 1] = (int32_t)bf.size();
        const bool fits = ownStart[c1] - ownStart[c0] <= gather_max_up(cells) &&
                          eo[blk + 1] - eo[blk] <= gather_max_ext(cells) &&
                          bo[blk + 1] - bo[blk] <= gather_max_bnd(cells) &&
                          innerMax <= gather_max_low(cells);
        nOverflow += fits ? 0 : 1;
    }
    L.nBlocks = nBlocks;
    L.nOverflow = nOverflow;
    L.nExt = (int64_t)ef.size();
    // the zero records past the last block
    ef.push_back(0);
    ew.push_back(0);
    ec.push_back(0);
    bf.push_back(0);
    bc.push_back(0);
    fv_ctx* ctx = m->ctx;
    L.extOff.upload(ctx, eo.data(), eo.size());
    L.extFace.upload(ctx, ef.data(), ef.size());
    L.extOwn.upload(ctx, ew.data(), ew.size());
    L.extCell.upload(ctx, ec.data(), ec.size());
    L.bndOff.upload(ctx, bo.data(), bo.size());
    L.bndFace.upload(ctx, bf.data(), bf.size());
    L.bndCell.upload(ctx, bc.data(), bc.size());
    FV_HIP(hipStreamSynchronize(ctx->stream));
    pack_ext_geo<double>(ctx, L.extFace, L.nExt, m->w.p, m->magSf.p, m->nodc.p,
                         m->Sf.p, L.exW, L.exMagSf, L.exNodc, L.exSf);
    L.ready32 = false;
    L.cells = cells;
    if (f32) return gather_lists(m, cells, true);
    return L;
}

void ensure_f32(fv_mesh* m) {
    if (m->f32.ready) return;
    fv_ctx* c = m->ctx;
    to_f32(c, m->f32.V, m->V);
    to_f32(c, m->f32.Sf, m->Sf);
    to_f32(c, m->f32.magSf, m->magSf);
    to_f32(c, m->f32.w, m->w);
    to_f32(c, m->f32.nodc, m->nodc);
    to_f32(c, m->f32.Sfb, m->Sfb);
    to_f32(c, m->f32.magSfb, m->magSfb);
    to_f32(c, m->f32.dcb, m->dcb);
    m->f32.ready = true;
}

}  // namespace fvhip

extern "C" {

int fv_mesh_create(fv_ctx* ctx, const fv_mesh_desc* d, fv_mesh** out) {
    fv_mesh* m = nullptr;
    int rc = guarded([&] {
        FV_REQUIRE(ctx && d && out, "NULL argument");
        FV_REQUIRE(d->n_cells > 0 && d->n_internal_faces >= 0 && d->n_boundary_faces >= 0 &&
                       d->n_patches >= 0,
                   "bad mesh sizes");
        m = new fv_mesh();
        m->ctx = ctx;
        const size_t N = d->n_cells, F = d->n_internal_faces, B = d->n_boundary_faces;
        m->nCells = d->n_cells;
        m->nInt = d->n_internal_faces;
        m->nB = d->n_boundary_faces;
        m->nPatches = d->n_patches;
        FV_REQUIRE(d->h_patch_offsets, "patch offsets NULL");
        m->patchOff.assign(d->h_patch_offsets, d->h_patch_offsets + d->n_patches + 1);
        FV_REQUIRE(m->patchOff[0] == 0 && m->patchOff.back() == (int32_t)B,
                   "patch offsets must span the boundary faces");
        up(ctx, m->V, d->h_V, N, "V");
        up(ctx, m->C, d->h_C, 3 * N, "C");
        up(ctx, m->own, d->h_owner, F, "owner");
        up(ctx, m->nbr, d->h_neighbour, F, "neighbour");
        up(ctx, m->Sf, d->h_Sf, 3 * F, "Sf");
        up(ctx, m->Cf, d->h_Cf, 3 * F, "Cf");
        up(ctx, m->magSf, d->h_magSf, F, "magSf");
        up(ctx, m->w, d->h_weights, F, "weights");
        up(ctx, m->dc, d->h_delta_coeffs, F, "deltaCoeffs");
        up(ctx, m->nodc, d->h_nonorth_delta_coeffs, F, "nonOrthDeltaCoeffs");
        up(ctx, m->faceCells, d->h_face_cells, B, "faceCells");
        up(ctx, m->Sfb, d->h_Sf_b, 3 * B, "Sf_b");
        up(ctx, m->Cfb, d->h_Cf_b, 3 * B, "Cf_b");
        up(ctx, m->magSfb, d->h_magSf_b, B, "magSf_b");
        up(ctx, m->nfb, d->h_nf_b, 3 * B, "nf_b");
        up(ctx, m->deltab, d->h_delta_b, 3 * B, "delta_b");
        up(ctx, m->wb, d->h_weights_b, B, "weights_b");
        up(ctx, m->dcb, d->h_delta_coeffs_b, B, "deltaCoeffs_b");
        up(ctx, m->nodcb, d->h_nonorth_delta_coeffs_b, B, "nonOrthDeltaCoeffs_b");
        m->patchOffD.upload(ctx, m->patchOff.data(), m->patchOff.size());
        build_stencil(m, d->h_owner, d->h_neighbour, d->h_face_cells);
        *out = m;
    });
    if (rc != FV_OK) delete m;
    return rc;
}

int fv_mesh_create_poly(fv_ctx* ctx, const fv_polymesh_desc* d, fv_mesh** out) {
    fv_mesh* m = nullptr;
    int rc = guarded([&] {
        FV_REQUIRE(ctx && d && out, "NULL argument");
        FV_REQUIRE(d->n_cells > 0 && d->n_faces >= d->n_internal_faces && d->n_points > 0 &&
                       d->n_patches >= 0,
                   "bad polyMesh sizes");
        FV_REQUIRE(d->h_points && d->h_face_offsets && d->h_face_points && d->h_owner &&
                       (d->h_neighbour || d->n_internal_faces == 0),
                   "NULL polyMesh array");
        const int64_t nP = d->n_points, nF = d->n_faces, F = d->n_internal_faces, N = d->n_cells;
        // validate face lists and owner/neighbour on the host before any launch
        FV_REQUIRE(d->h_face_offsets[0] == 0, "face offsets must start at 0");
        for (int64_t f = 0; f < nF; ++f) {
            int a = d->h_face_offsets[f], b = d->h_face_offsets[f + 1];
            FV_REQUIRE(b - a >= 3, "face with fewer than 3 points");
            FV_REQUIRE(d->h_owner[f] >= 0 && d->h_owner[f] < N, "owner out of range");
        }
        const int64_t nFP = d->h_face_offsets[nF];
        for (int64_t i = 0; i < nFP; ++i)
            FV_REQUIRE(d->h_face_points[i] >= 0 && d->h_face_points[i] < nP, "face point out of range");
        for (int64_t f = 0; f < F; ++f)
            FV_REQUIRE(d->h_neighbour[f] >= 0 && d->h_neighbour[f] < N, "neighbour out of range");
        // fv boundary faces: non-empty patches, in patch order
        std::vector<int32_t> bFace, poff(1, 0);
        for (int p = 0; p < d->n_patches; ++p) {
            int32_t s = d->h_patch_start[p], n = d->h_patch_size[p];
            FV_REQUIRE(s >= F && n >= 0 && s + (int64_t)n <= nF, "patch range out of bounds");
            if (!(d->h_patch_empty && d->h_patch_empty[p]))
                for (int32_t i = 0; i < n; ++i) bFace.push_back(s + i);
            poff.push_back((int32_t)bFace.size());
        }
        // per-cell geometry lists: owned faces ascending, then neighbour faces ascending
        std::vector<int32_t> gOff(N + 1, 0);
        for (int64_t f = 0; f < nF; ++f) gOff[d->h_owner[f] + 1]++;
        for (int64_t f = 0; f < F; ++f) gOff[d->h_neighbour[f] + 1]++;
        for (int64_t c = 0; c < N; ++c) {
            FV_REQUIRE(gOff[c + 1] > 0, "cell without faces");
            gOff[c + 1] += gOff[c];
        }
        std::vector<int32_t> gIdx(gOff[N]), fill(gOff.begin(), gOff.end() - 1);
        for (int64_t f = 0; f < nF; ++f) gIdx[fill[d->h_owner[f]]++] = (int32_t)f;
        for (int64_t f = 0; f < F; ++f) gIdx[fill[d->h_neighbour[f]]++] = ~(int32_t)f;

        m = new fv_mesh();
        m->ctx = ctx;
        m->nCells = (int32_t)N;
        m->nInt = (int32_t)F;
        m->nB = (int32_t)bFace.size();
        m->nPatches = d->n_patches;
        m->patchOff = poff;
        const int64_t B = m->nB;
        hipStream_t s = ctx->stream;
        DBuf<double> pts, CfAll, SfAll;
        DBuf<int32_t> fOff, fPts, ownAll, gOffD, gIdxD, bFaceD;
        pts.upload(ctx, d->h_points, 3 * nP);
        fOff.upload(ctx, d->h_face_offsets, nF + 1);
        fPts.upload(ctx, d->h_face_points, nFP);
        ownAll.upload(ctx, d->h_owner, nF);
        gOffD.upload(ctx, gOff.data(), gOff.size());
        gIdxD.upload(ctx, gIdx.data(), gIdx.size());
        bFaceD.upload(ctx, bFace.data(), bFace.size());
        CfAll.alloc(ctx, 3 * nF);
        SfAll.alloc(ctx, 3 * nF);
        m->V.alloc(ctx, N);
        m->C.alloc(ctx, 3 * N);
        k_face_geom<<<grid_for(nF, 256), 256, 0, s>>>((int)nF, fOff.p, fPts.p, pts.p, CfAll.p, SfAll.p);
        FV_LAUNCH_CHECK();
        k_cell_geom<<<grid_for(N, 256), 256, 0, s>>>((int)N, gOffD.p, gIdxD.p, CfAll.p, SfAll.p,
                                                     m->C.p, m->V.p);
        FV_LAUNCH_CHECK();
        m->own.upload(ctx, d->h_owner, F);
        m->nbr.upload(ctx, d->h_neighbour, F);
        m->Sf.alloc(ctx, 3 * F);
        m->Cf.alloc(ctx, 3 * F);
        if (F) {
            FV_HIP(hipMemcpyAsync(m->Sf.p, SfAll.p, 3 * F * sizeof(double), hipMemcpyDeviceToDevice, s));
            FV_HIP(hipMemcpyAsync(m->Cf.p, CfAll.p, 3 * F * sizeof(double), hipMemcpyDeviceToDevice, s));
        }
        m->magSf.alloc(ctx, F);
        m->w.alloc(ctx, F);
        m->dc.alloc(ctx, F);
        m->nodc.alloc(ctx, F);
        if (F) {
            k_internal_coeffs<<<grid_for(F, 256), 256, 0, s>>>((int)F, m->own.p, m->nbr.p, m->C.p,
                                                               m->Cf.p, m->Sf.p, m->magSf.p, m->w.p,
                                                               m->dc.p, m->nodc.p);
            FV_LAUNCH_CHECK();
        }
        m->faceCells.alloc(ctx, B);
        m->Sfb.alloc(ctx, 3 * B);
        m->Cfb.alloc(ctx, 3 * B);
        m->magSfb.alloc(ctx, B);
        m->nfb.alloc(ctx, 3 * B);
        m->deltab.alloc(ctx, 3 * B);
        m->wb.alloc(ctx, B);
        m->dcb.alloc(ctx, B);
        m->nodcb.alloc(ctx, B);
        if (B) {
            k_boundary_coeffs<<<grid_for(B, 256), 256, 0, s>>>(
                (int)B, bFaceD.p, ownAll.p, m->C.p, CfAll.p, SfAll.p, m->faceCells.p, m->Sfb.p,
                m->Cfb.p, m->magSfb.p, m->nfb.p, m->deltab.p, m->wb.p, m->dcb.p, m->nodcb.p);
            FV_LAUNCH_CHECK();
        }
        m->patchOffD.upload(ctx, m->patchOff.data(), m->patchOff.size());
        std::vector<int32_t> hfc(B);
        for (int64_t b = 0; b < B; ++b) hfc[b] = d->h_owner[bFace[b]];
        build_stencil(m, d->h_owner, d->h_neighbour, hfc.data());
        FV_HIP(hipStreamSynchronize(s));
        *out = m;
    });
    if (rc != FV_OK) delete m;
    return rc;
}

int fv_mesh_destroy(fv_mesh* m) {
    return guarded([&] { delete m; });
}

int fv_mesh_sizes(const fv_mesh* m, int32_t* s) {
    return guarded([&] {
        FV_REQUIRE(m && s, "NULL argument");
        s[0] = m->nCells;
        s[1] = m->nInt;
        s[2] = m->nB;
        s[3] = m->nPatches;
    });
}

int fv_mesh_get(const fv_mesh* m, fv_geom which, void* h) {
    return guarded([&] {
        FV_REQUIRE(m && h, "NULL argument");
        switch (which) {
            case FV_G_V: m->V.download((double*)h); break;
            case FV_G_C: m->C.download((double*)h); break;
            case FV_G_OWNER: m->own.download((int32_t*)h); break;
            case FV_G_NEIGHBOUR: m->nbr.download((int32_t*)h); break;
            case FV_G_SF: m->Sf.download((double*)h); break;
            case FV_G_CF: m->Cf.download((double*)h); break;
            case FV_G_MAGSF: m->magSf.download((double*)h); break;
            case FV_G_WEIGHTS: m->w.download((double*)h); break;
            case FV_G_DELTA_COEFFS: m->dc.download((double*)h); break;
            case FV_G_NONORTH_DELTA_COEFFS: m->nodc.download((double*)h); break;
            case FV_G_FACE_CELLS: m->faceCells.download((int32_t*)h); break;
            case FV_G_SF_B: m->Sfb.download((double*)h); break;
            case FV_G_CF_B: m->Cfb.download((double*)h); break;
            case FV_G_MAGSF_B: m->magSfb.download((double*)h); break;
            case FV_G_NF_B: m->nfb.download((double*)h); break;
            case FV_G_DELTA_B: m->deltab.download((double*)h); break;
            case FV_G_WEIGHTS_B: m->wb.download((double*)h); break;
            case FV_G_DELTA_COEFFS_B: m->dcb.download((double*)h); break;
            case FV_G_NONORTH_DELTA_COEFFS_B: m->nodcb.download((double*)h); break;
            case FV_G_PATCH_OFFSETS:
                std::copy(m->patchOff.begin(), m->patchOff.end(), (int32_t*)h);
                break;
            case FV_G_CELL_FACE_OFFSETS: m->cfOff.download((int32_t*)h); break;
            case FV_G_CELL_FACES: {
                int32_t* o = (int32_t*)h;
                m->cfIdx.download(o);
                for (size_t k = 0; k < m->cfIdx.n; ++k)
                    if (o[k] < 0) o[k] = ~o[k];
                break;
            }
            default: FV_REQUIRE(false, "unknown geometry id");
        }
    });
}

int fv_mesh_set_gather(fv_mesh* m, int mode) {
    return guarded([&] {
        FV_REQUIRE(m && mode >= 0 && mode <= 4, "bad argument");
        m->gatherMode = mode;
    });
}

int fv_mesh_set_gather_run(fv_mesh* m, int run) {
    return guarded([&] {
        FV_REQUIRE(m && run >= 1 && run <= 4096, "run must be in [1, 4096]");
        m->gatherRun = run;
    });
}

int fv_mesh_get_gather(const fv_mesh* m, int* mode) {
    return guarded([&] {
        FV_REQUIRE(m && mode, "NULL argument");
        *mode = gather_cells(m) == 0 ? 1 : gather_cells(m) == 256 ? 2 : gather_cells(m) == 512 ? 3 : 4;
    });
}

int fv_mesh_gather_stats(fv_mesh* m, int cells, int64_t* counts2) {
    return guarded([&] {
        FV_REQUIRE(m && counts2, "NULL argument");
        FV_REQUIRE(cells == 256 || cells == 512 || cells == 1024, "cells must be 256, 512 or 1024");
        FV_REQUIRE(m->upperTri, "the staged gather needs upper-triangular face order");
        const GatherLists& L = gather_lists(m, cells);
        counts2[0] = L.nBlocks;
        counts2[1] = L.nOverflow;
    });
}

int fv_bc_create(fv_mesh* m, const int32_t* kinds, fv_bc** out) {
    fv_bc* bc = nullptr;
    int rc = guarded([&] {
        FV_REQUIRE(m && out && (kinds || m->nPatches == 0), "NULL argument");
        bc = new fv_bc();
        bc->mesh = m;
        bc->kinds.assign(kinds, kinds + m->nPatches);
        std::vector<int8_t> fk(m->nB);
        for (int p = 0; p < m->nPatches; ++p) {
            FV_REQUIRE(kinds[p] >= FV_BC_FIXED_VALUE && kinds[p] <= FV_BC_CALCULATED,
                       "unknown boundary kind");
            for (int b = m->patchOff[p]; b < m->patchOff[p + 1]; ++b) fk[b] = (int8_t)kinds[p];
        }
        bc->faceKind.upload(m->ctx, fk.data(), fk.size());
        FV_HIP(hipStreamSynchronize(m->ctx->stream));
        *out = bc;
    });
    if (rc != FV_OK) delete bc;
    return rc;
}

int fv_bc_destroy(fv_bc* bc) {
    return guarded([&] { delete bc; });
}

}  // extern "C"
