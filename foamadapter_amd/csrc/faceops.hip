// faceops.hip -- explicit finite-volume face operators (NeoN fvcc:
// SurfaceInterpolation, GaussGreenGrad, GaussGreenDiv, GaussGreenLaplacian,
// FaceNormalGradient; FoamAdapter::flux).
//
// Cell-result operators (grad, div, laplacian) are computed as a deterministic
// cell gather: each cell walks its faces in ascending face order (the cell->face
// stencil), recomputes the face value from the face stream and the cell on the
// other side, and accumulates it with the owner (+) / neighbour (-) sign, then
// the boundary faces, then divides by the volume.  That is the serial
// accumulation order of OpenFOAM's surfaceIntegrate / gaussGrad::gradf, so no
// float atomics are used and the result does not depend on scheduling.
// Face-result operators (interpolate, snGrad, flux) are one coalesced pass
// over the faces.
#include <type_traits>

#include "mesh.hpp"

using namespace fvhip;

namespace {

constexpr int kBlock = 256;

template <typename S>
struct Geo {
    const S *V, *Sf, *magSf, *w, *nodc, *Sfb, *magSfb, *dcb;
};

Geo<double> geo64(const fv_mesh* m) {
    return {m->V.p, m->Sf.p, m->magSf.p, m->w.p, m->nodc.p, m->Sfb.p, m->magSfb.p, m->dcb.p};
}
Geo<float> geo32(fv_mesh* m) {
    ensure_f32(m);
    auto& g = m->f32;
    return {g.V.p, g.Sf.p, g.magSf.p, g.w.p, g.nodc.p, g.Sfb.p, g.magSfb.p, g.dcb.p};
}

// Geometry of the staged gather's external lower faces, packed in list order
// (record k of the per-workgroup lists): w, magSf, nonOrthDeltaCoeffs, Sf.
template <typename S>
struct ExGeo {
    const S *w, *magSf, *nodc, *Sf;
};

// ---- cell-result operators -------------------------------------------------

template <typename S>
__global__ __launch_bounds__(kBlock) void k_grad(int N, int F, const int32_t* __restrict__ off,
                                                 const int32_t* __restrict__ idx,
                                                 const int32_t* __restrict__ oth, Geo<S> g,
                                                 const S* __restrict__ T, const S* __restrict__ Tb,
                                                 S coeff, S* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const S Tc = T[c];
    S ax = 0, ay = 0, az = 0;
    const int e1 = off[c + 1];
    for (int k = off[c]; k < e1; ++k) {
        const int e = idx[k];
        if (e >= F) {
            const int b = e - F;
            const S tb = Tb[b];
            ax += g.Sfb[3 * b] * tb;
            ay += g.Sfb[3 * b + 1] * tb;
            az += g.Sfb[3 * b + 2] * tb;
        } else {
            const S To = T[oth[k]];
            if (e >= 0) {
                const S tf = g.w[e] * (Tc - To) + To;
                ax += g.Sf[3 * (int64_t)e] * tf;
                ay += g.Sf[3 * (int64_t)e + 1] * tf;
                az += g.Sf[3 * (int64_t)e + 2] * tf;
            } else {
                const int f = ~e;
                const S tf = g.w[f] * (To - Tc) + Tc;
                ax -= g.Sf[3 * (int64_t)f] * tf;
                ay -= g.Sf[3 * (int64_t)f + 1] * tf;
                az -= g.Sf[3 * (int64_t)f + 2] * tf;
            }
        }
    }
    const S v = g.V[c];
    out[3 * (int64_t)c] = coeff * (ax / v);
    out[3 * (int64_t)c + 1] = coeff * (ay / v);
    out[3 * (int64_t)c + 2] = coeff * (az / v);
}

// scheme: 0 linear, 1 upwind; T == nullptr: surfaceIntegrate(phi)
template <typename S>
__global__ __launch_bounds__(kBlock) void k_div(int N, int F, int scheme,
                                                const int32_t* __restrict__ off,
                                                const int32_t* __restrict__ idx,
                                                const int32_t* __restrict__ oth, Geo<S> g,
                                                const S* __restrict__ phi, const S* __restrict__ T,
                                                const S* __restrict__ Tb, S coeff,
                                                S* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    S acc = 0;
    const int e1 = off[c + 1];
    if (T) {
        const S Tc = T[c];
        for (int k = off[c]; k < e1; ++k) {
            const int e = idx[k];
            if (e >= F) {
                const int b = e - F;
                acc += phi[e] * Tb[b];
            } else {
                const S To = T[oth[k]];
                const int f = e >= 0 ? e : ~e;
                const S ph = phi[f];
                const S lam = scheme == 0 ? g.w[f] : (ph >= S(0) ? S(1) : S(0));
                // owner side: P = this cell; neighbour side: P = other cell
                const S TP = e >= 0 ? Tc : To, TN = e >= 0 ? To : Tc;
                const S v = ph * (lam * (TP - TN) + TN);
                if (e >= 0) acc += v; else acc -= v;
            }
        }
    } else {
        for (int k = off[c]; k < e1; ++k) {
            const int e = idx[k];
            if (e >= F) acc += phi[e];
            else if (e >= 0) acc += phi[e];
            else acc -= phi[~e];
        }
    }
    out[c] = coeff * (acc / g.V[c]);
}

template <typename S>
__device__ inline S bnd_sngrad(int8_t kind, S dcb, S Tb, S TP, const S* Tg, int b) {
    if (kind == FV_BC_FIXED_GRADIENT) return Tg ? Tg[b] : S(0);
    return dcb * (Tb - TP);
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_laplacian(
    int N, int F, const int32_t* __restrict__ off, const int32_t* __restrict__ idx,
    const int32_t* __restrict__ oth, Geo<S> g, const int8_t* __restrict__ kind,
    const S* __restrict__ gamma, const S* __restrict__ T, const S* __restrict__ Tb,
    const S* __restrict__ Tg, S coeff, S* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const S Tc = T[c];
    S acc = 0;
    const int e1 = off[c + 1];
    for (int k = off[c]; k < e1; ++k) {
        const int e = idx[k];
        if (e >= F) {
            const int b = e - F;
            const S sn = bnd_sngrad<S>(kind[b], g.dcb[b], Tb[b], Tc, Tg, b);
            acc += (gamma[e] * g.magSfb[b]) * sn;
        } else {
            const S To = T[oth[k]];
            const int f = e >= 0 ? e : ~e;
            const S TP = e >= 0 ? Tc : To, TN = e >= 0 ? To : Tc;
            const S v = (gamma[f] * g.magSf[f]) * (g.nodc[f] * (TN - TP));
            if (e >= 0) acc += v; else acc -= v;
        }
    }
    out[c] = coeff * (acc / g.V[c]);
}

// ---- LDS-staged cell gathers (meshes in upper-triangular face order) --------
//
// A workgroup owns CELLS consecutive cells [c0, c1) (256, 512 or 1024).  In
// OpenFOAM face order the internal faces those cells own are one contiguous
// range [fs, fe), streamed with coalesced loads; the faces owned by cells
// before c0 whose neighbour is in [c0, c1) (external lower faces) and the
// boundary faces of [c0, c1) come from per-workgroup lists built once per mesh
// (GatherLists, with the external faces' geometry packed in list order).  Every
// contribution is computed once into LDS, then each cell accumulates its faces
// in ascending face order -- the serial face order of the other kernels, so the
// result is the same bit for bit.  On a Z-order renumbered mesh a workgroup is a
// compact brick (8x8x4, 8x8x8 or 16x8x8 cells), and only the faces on its lower
// sides are external.
// The hardware deals workgroups round-robin over the 8 XCDs (block h runs on
// XCD h % 8), each with its own L2.  With `run` > 1 the block index is mapped
// so that each XCD takes runs of `run` consecutive bricks (XCD x's s-th block
// gets brick ((s / run) * 8 + x) * run + s % run): a brick's lower neighbours
// in Z-order (bricks b-1, b-2, b-4, ...), whose faces and cells it reads
// again, then ran shortly before on the same XCD and are still in its L2.
// Measured at 256^3, runs of 8 cut the L2 misses of the gathers by 6-18% at
// the same speed; run = 1 is the natural order.
__device__ __forceinline__ int brick_of(int h, int run) {
    if (run <= 1) return h;
    const int x = h & 7, s = h >> 3;
    return ((s / run) * 8 + x) * run + s % run;
}
inline int gather_grid(int nBlocks, int run) {
    if (run <= 1) return nBlocks;
    const int g = 8 * run;
    return (nBlocks + g - 1) / g * g;
}

// LDS face slots of a CELLS-cell workgroup: 3 owned faces per hex cell plus slack;
// faces beyond it (unusual meshes) are recomputed in phase 2
template <int CELLS>
constexpr int max_up() { return gather_max_up(CELLS); }

struct Split {
    const int32_t *own, *nbr, *ownStart, *lowOff, *lowFace, *lowOwn, *bOff, *bFace;
};

// An operator gives each face a contribution of NC components.  For the
// staged gather it splits a face into its loads (load/eload/bload: global reads
// only, into an In/BIn record; load = fload, the face streams, then tload, the
// cell values by index) and the arithmetic on them (make/bmake), so the
// kernel can issue the loads of every face it holds -- own, external lower and
// boundary -- before it waits for any.  The arithmetic is the serial kernels'
// own expression, so the bits are the same.  face/bface (make of load) serve
// the per-cell fallback.
template <typename S, typename D, int NCOMP>
struct FaceOp {
    static constexpr int NC = NCOMP;
    static constexpr int kDefaultCells = 256;
    // own faces per thread and pass of k_gather (measured at 256^3, 256-cell
    // workgroups: 3 beats 4 and 2 -- loads in flight against registers)
    static constexpr int kFacesPerPass = 3;
    __device__ const D& self() const { return *static_cast<const D*>(this); }
    __device__ void face(int f, int o, int n, S* v) const { self().make(self().load(f, o, n), v); }
    __device__ void bface(int b, int c, S* v) const { self().bmake(self().bload(b, c), v); }
};

template <typename S>
struct GradOp : FaceOp<S, GradOp<S>, 3> {
    const S *T, *w, *Sf, *Sfb, *Tb;
    struct In {
        S w, TP, TN, sx, sy, sz;
    };
    struct BIn {
        S tb, sx, sy, sz;
    };
    __device__ In load(int f, int o, int n) const {
        const int64_t f3 = 3 * (int64_t)f;
        return {w[f], T[o], T[n], Sf[f3], Sf[f3 + 1], Sf[f3 + 2]};
    }
    __device__ In fload(int f) const {
        const int64_t f3 = 3 * (int64_t)f;
        return {w[f], S(0), S(0), Sf[f3], Sf[f3 + 1], Sf[f3 + 2]};
    }
    __device__ void tload(In& a, int o, int n) const {
        a.TP = T[o];
        a.TN = T[n];
    }
    __device__ In eload(int k, int, int o, int n, const ExGeo<S>& x) const {
        const int64_t k3 = 3 * (int64_t)k;
        return {x.w[k], T[o], T[n], x.Sf[k3], x.Sf[k3 + 1], x.Sf[k3 + 2]};
    }
    __device__ void make(const In& a, S* v) const {
        const S s0 = a.w * (a.TP - a.TN) + a.TN;
        v[0] = a.sx * s0;
        v[1] = a.sy * s0;
        v[2] = a.sz * s0;
    }
    __device__ BIn bload(int b, int) const { return {Tb[b], Sfb[3 * b], Sfb[3 * b + 1], Sfb[3 * b + 2]}; }
    __device__ void bmake(const BIn& a, S* v) const {
        v[0] = a.sx * a.tb;
        v[1] = a.sy * a.tb;
        v[2] = a.sz * a.tb;
    }
};

template <typename S, int SCHEME>  // 0 linear, 1 upwind
struct DivOp : FaceOp<S, DivOp<S, SCHEME>, 1> {
    // measured at 256^3: 512-cell workgroups are 3-5% faster than 256 in FP64,
    // 1024-cell ones 2-4% in FP32; the FP32 workgroup's 3072 own faces go out in
    // one pass of six per thread (upwind +10%, linear +4% over two passes of three)
    static constexpr int kDefaultCells = sizeof(S) == 8 ? 512 : 1024;
    static constexpr int kFacesPerPass = sizeof(S) == 8 ? 3 : 6;
    const S *T, *w, *phi, *Tb;
    int F;
    struct In {
        S ph, w, TP, TN;
    };
    struct BIn {
        S ph, tb;
    };
    __device__ In load(int f, int o, int n) const {
        return {phi[f], SCHEME == 0 ? w[f] : S(0), T[o], T[n]};
    }
    __device__ In fload(int f) const { return {phi[f], SCHEME == 0 ? w[f] : S(0), S(0), S(0)}; }
    __device__ void tload(In& a, int o, int n) const {
        a.TP = T[o];
        a.TN = T[n];
    }
    __device__ In eload(int k, int f, int o, int n, const ExGeo<S>& x) const {
        return {phi[f], SCHEME == 0 ? x.w[k] : S(0), T[o], T[n]};
    }
    __device__ void make(const In& a, S* v) const {
        const S lam = SCHEME == 0 ? a.w : (a.ph >= S(0) ? S(1) : S(0));
        v[0] = a.ph * (lam * (a.TP - a.TN) + a.TN);
    }
    __device__ BIn bload(int b, int) const { return {phi[F + b], Tb[b]}; }
    __device__ void bmake(const BIn& a, S* v) const { v[0] = a.ph * a.tb; }
};

template <typename S>
struct IntegrateOp : FaceOp<S, IntegrateOp<S>, 1> {  // surfaceIntegrate(phi): div with a unit field
    const S* phi;
    int F;
    struct In {
        S ph;
    };
    using BIn = In;
    __device__ In load(int f, int, int) const { return {phi[f]}; }
    __device__ In fload(int f) const { return {phi[f]}; }
    __device__ void tload(In&, int, int) const {}
    __device__ In eload(int, int f, int, int, const ExGeo<S>&) const { return {phi[f]}; }
    __device__ void make(const In& a, S* v) const { v[0] = a.ph; }
    __device__ BIn bload(int b, int) const { return {phi[F + b]}; }
    __device__ void bmake(const BIn& a, S* v) const { v[0] = a.ph; }
};

template <typename S>
struct LapOp : FaceOp<S, LapOp<S>, 1> {
    // measured at 256^3: in FP32 1024-cell workgroups are 5% faster than 256,
    // and six own faces per thread in one pass 4% faster than two passes of three
    static constexpr int kDefaultCells = sizeof(S) == 8 ? 256 : 1024;
    static constexpr int kFacesPerPass = sizeof(S) == 8 ? 3 : 6;
    const S *T, *gamma, *magSf, *nodc, *magSfb, *dcb, *Tb, *Tg;
    const int8_t* kind;
    int F;
    struct In {
        S g, m, nd, TN, TP;
    };
    struct BIn {
        S g, m, tg, d, tb, tc;
        int k;
    };
    __device__ In load(int f, int o, int n) const {
        return {gamma[f], magSf[f], nodc[f], T[n], T[o]};
    }
    __device__ In fload(int f) const { return {gamma[f], magSf[f], nodc[f], S(0), S(0)}; }
    __device__ void tload(In& a, int o, int n) const {
        a.TN = T[n];
        a.TP = T[o];
    }
    __device__ In eload(int k, int f, int o, int n, const ExGeo<S>& x) const {
        return {gamma[f], x.magSf[k], x.nodc[k], T[n], T[o]};
    }
    __device__ void make(const In& a, S* v) const { v[0] = (a.g * a.m) * (a.nd * (a.TN - a.TP)); }
    __device__ BIn bload(int b, int c) const {
        return {gamma[F + b], magSfb[b], Tg ? Tg[b] : S(0), dcb[b], Tb[b], T[c], (int)kind[b]};
    }
    __device__ void bmake(const BIn& a, S* v) const {
        const S sn = a.k == FV_BC_FIXED_GRADIENT ? a.tg : a.d * (a.tb - a.tc);
        v[0] = (a.g * a.m) * sn;
    }
};

// One cell of the per-cell path: LDS holds the contributions (NC components,
// `stride` apart) of the workgroup's first nUp own faces; the others are
// recomputed.
template <typename S, typename Op>
__device__ inline void gather_cell(int c, int fs, int nUp, const Split& t, const Op& op,
                                   S vol, S coeff, S* __restrict__ out, const S* st0, int stride) {
    constexpr int NC = Op::NC;
    S acc[NC];
    for (int j = 0; j < NC; ++j) acc[j] = S(0);
    const int l1 = t.lowOff[c + 1];
    for (int k = t.lowOff[c]; k < l1; ++k) {
        const int f = t.lowFace[k];
        S v[NC];
        const int i = f - fs;
        if (i >= 0 && i < nUp) {
            for (int j = 0; j < NC; ++j) v[j] = st0[j * stride + i];
        } else {
            op.face(f, t.lowOwn[k], c, v);
        }
        for (int j = 0; j < NC; ++j) acc[j] -= v[j];
    }
    const int u1 = t.ownStart[c + 1];
    for (int f = t.ownStart[c]; f < u1; ++f) {
        S v[NC];
        const int i = f - fs;
        if (i < nUp) {
            for (int j = 0; j < NC; ++j) v[j] = st0[j * stride + i];
        } else {
            op.face(f, c, t.nbr[f], v);
        }
        for (int j = 0; j < NC; ++j) acc[j] += v[j];
    }
    const int b1 = t.bOff[c + 1];
    for (int k = t.bOff[c]; k < b1; ++k) {
        S v[NC];
        op.bface(t.bFace[k], c, v);
        for (int j = 0; j < NC; ++j) acc[j] += v[j];
    }
    for (int j = 0; j < NC; ++j) out[(int64_t)NC * c + j] = coeff * (acc[j] / vol);
}

struct Lists {
    const int32_t *extOff, *extFace, *extOwn, *bndOff, *bndFace;
    const uint16_t *extCell, *bndCell;
};

// Sorts 8 (or 4) keys ascending in registers: the 19-comparator network for 8
// inputs, layers (0,2)(1,3)(4,6)(5,7) / (0,4)(1,5)(2,6)(3,7) /
// (0,1)(2,3)(4,5)(6,7) / (2,4)(3,5) / (1,4)(3,6) / (1,2)(3,4)(5,6); the
// 5-comparator one for 4, (0,1)(2,3) / (0,2)(1,3) / (1,2).
template <int K>
__device__ inline void sort_keys(uint16_t (&k)[K]) {
    static_assert(K == 8 || K == 4, "per-cell lower-face lists hold 8 or 4 entries");
    auto cx = [&](int a, int b) {
        const uint16_t x = k[a], y = k[b];
        k[a] = x < y ? x : y;
        k[b] = x < y ? y : x;
    };
    if constexpr (K == 8) {
        cx(0, 2); cx(1, 3); cx(4, 6); cx(5, 7);
        cx(0, 4); cx(1, 5); cx(2, 6); cx(3, 7);
        cx(0, 1); cx(2, 3); cx(4, 5); cx(6, 7);
        cx(2, 4); cx(3, 5);
        cx(1, 4); cx(3, 6);
        cx(1, 2); cx(3, 4); cx(5, 6);
    } else {
        cx(0, 1); cx(2, 3);
        cx(0, 2); cx(1, 3);
        cx(1, 2);
    }
}

// A workgroup of CELLS cells.  Phase 1 computes, with coalesced loads, every
// face contribution the workgroup needs into LDS -- its own faces (streamed),
// its external lower faces and its boundary faces (from the per-workgroup
// lists) -- and builds each cell's face lists in LDS: the own-face range from
// the cells' face offsets, the lower faces from inside the workgroup by LDS atomics
// (then sorted, so the order is fixed), the external and boundary ranges from
// the lists (sorted by cell).  Phase 2 accumulates each cell in ascending face
// order (external lower, inner lower, own, boundary) and divides by V.  A
// workgroup whose faces do not fit the LDS arrays takes the per-cell path of
// gather_cell instead; both give the same bits.
template <typename S, typename Op, int CELLS, int THREADS>
__global__ __launch_bounds__(THREADS) void k_gather(int N, int nBlocks, int run, Split t,
                                                    Lists L, ExGeo<S> xg, Op op,
                                                    const S* __restrict__ V, S coeff,
                                                    S* __restrict__ out) {
    constexpr int NC = Op::NC;
    constexpr int kUL = Op::kFacesPerPass;
    constexpr int kMaxUp = max_up<CELLS>();
    constexpr int kMaxExt = gather_max_ext(CELLS);
    constexpr int kMaxBnd = gather_max_bnd(CELLS);
    constexpr int kMaxLow = gather_max_low(CELLS);  // inner lower faces per cell, in LDS
    __shared__ S fluxU[NC][kMaxUp];
    __shared__ S fluxE[NC][kMaxExt];
    __shared__ S fluxB[NC][kMaxBnd];
    // upS[l], upE[l]: the own-face range of cell l (local indices).  A
    // 1024-cell workgroup keeps no upE (its end is upS[l + 1]): the 2 KB let
    // four FP32 workgroups share a CU's LDS.  (Smaller workgroups keep it:
    // reading the neighbouring cell's start made them 2-3 % slower.)
    constexpr bool kUpE = CELLS != 1024;
    __shared__ uint16_t upS[CELLS + 1], upE[kUpE ? CELLS : 1], exS[CELLS], exE[CELLS], bS[CELLS],
        bE[CELLS];
    __shared__ uint16_t lowL[kMaxLow][CELLS];  // slot-major: lanes read consecutive cells
    __shared__ int lowN[CELLS];
    __shared__ int overflow;
    const int lb = brick_of((int)blockIdx.x, run);
    if (lb >= nBlocks) return;
    const int c0 = lb * CELLS, c1 = min(c0 + CELLS, N);
    const int fs = t.ownStart[c0], fe = t.ownStart[c1];
    const int nF = fe - fs;
    const int e0 = L.extOff[lb], e1 = L.extOff[lb + 1];
    const int b0 = L.bndOff[lb], b1 = L.bndOff[lb + 1];
    const bool fits = nF <= kMaxUp && e1 - e0 <= kMaxExt && b1 - b0 <= kMaxBnd;
    // the volumes phase 2 divides by, loaded now so their latency hides behind phase 1
    constexpr int kPer = CELLS / THREADS;
    S vols[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int c = c0 + (int)threadIdx.x + q * THREADS;
        vols[q] = c < c1 ? V[c] : S(1);
    }
    using In = typename Op::In;
    using BIn = typename Op::BIn;
    // A workgroup whose faces do not fit stages nothing: the per-cell path
    // recomputes every face it needs.
    const int nUp = fits ? nF : 0;
    // The own faces' owners are not read from memory: in upper-triangular
    // order the faces of cell c are [ownStart[c], ownStart[c+1]), so each cell
    // writes its own-face range, and its local index into the LDS slot
    // fluxU[0][i] of each of its faces, which phase 1 overwrites with the
    // face's contribution in the one thread that reads it.  Everything that
    // needs no index -- the cells' offsets, the first pass's neighbour indices
    // and face streams, the list records -- is loaded before the barrier, in
    // one round trip; after it only the cell values (by owner and neighbour
    // index) and the list faces' inputs are.
    int oa[kPer], ob[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int c = c0 + (int)threadIdx.x + q * THREADS;
        oa[q] = ob[q] = fe;  // cells past the last one: empty ranges at the end
        if (fits && c < c1) {
            oa[q] = t.ownStart[c];
            ob[q] = t.ownStart[c + 1];
        }
    }
    int nn0[kUL] = {};
    In fin0[kUL] = {};  // the first pass's face streams (the cell values come after the barrier)
    if (nUp > 0) {
#pragma unroll
        for (int u = 0; u < kUL; ++u) nn0[u] = t.nbr[fs + min((int)threadIdx.x + u * THREADS, nUp - 1)];
#pragma unroll
        for (int u = 0; u < kUL; ++u) fin0[u] = op.fload(fs + min((int)threadIdx.x + u * THREADS, nUp - 1));
    }
    // Each thread takes kUL own faces per pass (index i0 + u*THREADS).  The
    // first pass also takes this thread's list records (at most one external
    // lower face and one boundary face: the lists of a fitting workgroup hold
    // no more records than it has threads).
    static_assert(kMaxExt <= THREADS && kMaxBnd <= THREADS, "one list record per thread");
    const int ke = e0 + (int)threadIdx.x, kb = b0 + (int)threadIdx.x;
    const bool hasE = fits && ke < e1, hasB = fits && kb < b1;
    // Every lane reads records (past its workgroup's range: the first record
    // after it, at most the zero record past the last workgroup), so no lane
    // branches around these loads; lanes without a record ignore them.
    const int kec = min(ke, e1), kbc = min(kb, b1);
    const int ef = L.extFace[kec], eo = L.extOwn[kec], ecl = L.extCell[kec];
    const int ecp = L.extCell[max(kec - 1, e0)], ecn = L.extCell[max(min(kec + 1, e1 - 1), e0)];
    const int bf = L.bndFace[kbc], bcl = L.bndCell[kbc];
    const int bcp = L.bndCell[max(kbc - 1, b0)], bcn = L.bndCell[max(min(kbc + 1, b1 - 1), b0)];
    for (int l = threadIdx.x; l < CELLS; l += THREADS) {
        exS[l] = exE[l] = bS[l] = bE[l] = 0;
        lowN[l] = 0;
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int l = (int)threadIdx.x + q * THREADS;
        const int a = oa[q] - fs, b = ob[q] - fs;  // empty unless fits and c < c1
        upS[l] = (uint16_t)a;
        if constexpr (kUpE) upE[l] = (uint16_t)b;
        for (int i = a; i < b; ++i) fluxU[0][i] = S(l);
    }
    if (threadIdx.x == 0) {
        overflow = fits ? 0 : 1;
        upS[CELLS] = (uint16_t)nF;
    }
    __syncthreads();
    // This thread's list faces: loads, contribution into LDS, the cell ranges.
    auto list_faces = [&] {
        In ein{};
        BIn bin{};
        if (hasE) ein = op.eload(ke, ef, eo, c0 + ecl, xg);
        if (hasB) bin = op.bload(bf, c0 + bcl);
        if (hasE) {
            S sv[NC];
            op.make(ein, sv);
            for (int j = 0; j < NC; ++j) fluxE[j][ke - e0] = sv[j];
            if (ke == e0 || ecp != ecl) exS[ecl] = (uint16_t)(ke - e0);
            if (ke == e1 - 1 || ecn != ecl) exE[ecl] = (uint16_t)(ke - e0 + 1);
        }
        if (hasB) {
            S sv[NC];
            op.bmake(bin, sv);
            for (int j = 0; j < NC; ++j) fluxB[j][kb - b0] = sv[j];
            if (kb == b0 || bcp != bcl) bS[bcl] = (uint16_t)(kb - b0);
            if (kb == b1 - 1 || bcn != bcl) bE[bcl] = (uint16_t)(kb - b0 + 1);
        }
    };
    // One pass over this thread's kUL own faces from i0 (nUp > 0).  The loads
    // of every face the pass holds, then in the first pass the list faces',
    // go out before any arithmetic waits, and no lane branches around a load
    // (the compiler waits for every load in flight where such a branch joins):
    // lanes past the range read the map slot and load the inputs of the last
    // face, take owner c0 and store nothing.
    auto own_pass = [&](int i0, const int (&nn)[kUL], In (&fin)[kUL], auto first) {
        constexpr bool kFirst = decltype(first)::value;
#pragma unroll
        for (int u = 0; u < kUL; ++u) {
            const int iu = min(i0 + u * THREADS, nUp - 1);
            const S sl = fluxU[0][iu];  // the owner map
            const int ol = i0 + u * THREADS < nUp ? (int)sl : 0;
            op.tload(fin[u], c0 + ol, nn[u]);
        }
        if constexpr (kFirst) list_faces();
#pragma unroll
        for (int u = 0; u < kUL; ++u) {
            const int i = i0 + u * THREADS;
            if (i >= nUp) break;
            const int n = nn[u];
            S sv[NC];
            op.make(fin[u], sv);
            for (int j = 0; j < NC; ++j) fluxU[j][i] = sv[j];
            if (n < c1) {
                const int nl = n - c0;
                const int slot = atomicAdd(&lowN[nl], 1);
                if (slot < kMaxLow)
                    lowL[slot][nl] = (uint16_t)i;
                else
                    overflow = 1;
            }
        }
    };
    if (nUp > 0) {
        own_pass((int)threadIdx.x, nn0, fin0, std::true_type{});
        for (int i0 = (int)threadIdx.x + THREADS * kUL; i0 < nUp; i0 += THREADS * kUL) {
            int nn[kUL];
            In fin[kUL];
#pragma unroll
            for (int u = 0; u < kUL; ++u) {
                const int f = fs + min(i0 + u * THREADS, nUp - 1);
                nn[u] = t.nbr[f];
                fin[u] = op.fload(f);
            }
            own_pass(i0, nn, fin, std::false_type{});
        }
    } else {
        list_faces();  // a workgroup that owns no face (none when it does not fit)
    }
    __syncthreads();
    if (overflow) {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int c = c0 + (int)threadIdx.x + q * THREADS;
            if (c < c1) gather_cell<S, Op>(c, fs, nUp, t, op, vols[q], coeff, out, fluxU[0], kMaxUp);
        }
        return;
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int c = c0 + (int)threadIdx.x + q * THREADS;
        if (c >= c1) break;
        const int l = c - c0;
        S acc[NC];
        for (int j = 0; j < NC; ++j) acc[j] = S(0);
        S v[NC];
        for (int k = exS[l]; k < exE[l]; ++k) {
            for (int j = 0; j < NC; ++j) v[j] = fluxE[j][k];
            for (int j = 0; j < NC; ++j) acc[j] -= v[j];
        }
        // the inner lower faces arrived in any order: sort this cell's list in
        // registers (unused slots sort last)
        const int n = lowN[l];
        uint16_t key[kMaxLow];
#pragma unroll
        for (int a = 0; a < kMaxLow; ++a) key[a] = a < n ? lowL[a][l] : uint16_t(0xFFFF);
        sort_keys<kMaxLow>(key);
#pragma unroll
        for (int a = 0; a < kMaxLow; ++a)
            if (a < n) {
                const int i = key[a];
                for (int j = 0; j < NC; ++j) v[j] = fluxU[j][i];
                for (int j = 0; j < NC; ++j) acc[j] -= v[j];
            }
        const int ue = kUpE ? upE[l] : upS[l + 1];
        for (int i = upS[l]; i < ue; ++i) {
            for (int j = 0; j < NC; ++j) v[j] = fluxU[j][i];
            for (int j = 0; j < NC; ++j) acc[j] += v[j];
        }
        for (int k = bS[l]; k < bE[l]; ++k) {
            for (int j = 0; j < NC; ++j) v[j] = fluxB[j][k];
            for (int j = 0; j < NC; ++j) acc[j] += v[j];
        }
        const S vol = vols[q];
        for (int j = 0; j < NC; ++j) out[(int64_t)NC * c + j] = coeff * (acc[j] / vol);
    }
}

Split split_of(const fv_mesh* m) {
    return {m->own.p,     m->nbr.p,    m->ownStart.p, m->lowOff.p,
            m->lowFace.p, m->lowOwn.p, m->bOff.p,     m->bFace.p};
}

template <typename S, typename Op, int CELLS, int THREADS>
void launch_gather_t(fv_ctx* ctx, const fv_mesh* m, const Op& op, const S* V, S coeff, S* out) {
    const int nBlocks = (m->nCells + CELLS - 1) / CELLS;
    // the lists are built once per mesh and workgroup size (a cache on the mesh)
    const GatherLists& gl = gather_lists(const_cast<fv_mesh*>(m), CELLS, sizeof(S) == 4);
    const Lists L{gl.extOff.p, gl.extFace.p, gl.extOwn.p, gl.bndOff.p, gl.bndFace.p,
                  gl.extCell.p, gl.bndCell.p};
    ExGeo<S> xg;
    if constexpr (sizeof(S) == 8)
        xg = {gl.exW.p, gl.exMagSf.p, gl.exNodc.p, gl.exSf.p};
    else
        xg = {gl.exW32.p, gl.exMagSf32.p, gl.exNodc32.p, gl.exSf32.p};
    k_gather<S, Op, CELLS, THREADS><<<gather_grid(nBlocks, m->gatherRun), THREADS, 0, ctx->stream>>>(
        m->nCells, nBlocks, m->gatherRun, split_of(m), L, xg, op, V, coeff, out);
    FV_LAUNCH_CHECK();
}

template <typename S, typename Op>
void launch_gather(fv_ctx* ctx, const fv_mesh* m, const Op& op, const S* V, S coeff, S* out) {
    // default workgroup size per operator (measured at 256^3)
    int cells = m->gatherMode == 0 ? Op::kDefaultCells : gather_cells(m);
    if (m->gatherMode == 0) {
        // A mesh whose Z-order runs are ragged (sides that are not powers of
        // two, irregular cells) can overflow the larger workgroups' lists: more
        // than 1 % of the workgroups taking the per-cell path moves the
        // operator to the next smaller size (same bits; the lists are cached).
        while (cells > 256) {
            const GatherLists& gl = gather_lists(const_cast<fv_mesh*>(m), cells, sizeof(S) == 4);
            if (gl.nOverflow * 100 <= gl.nBlocks) break;
            cells /= 2;
        }
    }
    switch (cells) {
        case 1024: launch_gather_t<S, Op, 1024, 512>(ctx, m, op, V, coeff, out); break;
        case 512: launch_gather_t<S, Op, 512, 512>(ctx, m, op, V, coeff, out); break;
        default: launch_gather_t<S, Op, 256, 256>(ctx, m, op, V, coeff, out); break;
    }
}

// ---- face-result operators -------------------------------------------------

template <typename S>
__global__ __launch_bounds__(kBlock) void k_interp(int F, int B, int scheme,
                                                   const int32_t* __restrict__ own,
                                                   const int32_t* __restrict__ nbr,
                                                   const S* __restrict__ w,
                                                   const S* __restrict__ phi,
                                                   const S* __restrict__ T,
                                                   const S* __restrict__ Tb, S* __restrict__ Tf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F) {
        const S lam = scheme == 0 ? w[i] : (phi[i] >= S(0) ? S(1) : S(0));
        const S TN = T[nbr[i]];
        Tf[i] = lam * (T[own[i]] - TN) + TN;
    } else if (i < F + B) {
        Tf[i] = Tb[i - F];
    }
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_sngrad(int F, int B, const int32_t* __restrict__ own,
                                                   const int32_t* __restrict__ nbr,
                                                   const int32_t* __restrict__ fc,
                                                   const S* __restrict__ nodc,
                                                   const S* __restrict__ dcb,
                                                   const int8_t* __restrict__ kind,
                                                   const S* __restrict__ T,
                                                   const S* __restrict__ Tb,
                                                   const S* __restrict__ Tg, S* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F) {
        out[i] = nodc[i] * (T[nbr[i]] - T[own[i]]);
    } else if (i < F + B) {
        const int b = i - F;
        out[i] = bnd_sngrad<S>(kind[b], dcb[b], Tb[b], T[fc[b]], Tg, b);
    }
}

// Four consecutive internal faces per thread with 16-byte loads and stores of
// the face streams (owner, neighbour, weights/coefficients, result), then the
// tail faces and the boundary faces one per thread; the per-face arithmetic is
// that of k_interp / k_sngrad.  Used when the face streams are 16-byte aligned.
template <typename T>
struct alignas(16) Vec4 {
    T v[4];
};
template <typename T>
__device__ inline Vec4<T> ld4(const T* p) {
    return *reinterpret_cast<const Vec4<T>*>(p);
}
template <typename T>
__device__ inline void st4(T* p, const Vec4<T>& x) {
    *reinterpret_cast<Vec4<T>*>(p) = x;
}
inline bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

template <typename S>
__global__ __launch_bounds__(kBlock) void k_interp4(int F, int B, int scheme,
                                                    const int32_t* __restrict__ own,
                                                    const int32_t* __restrict__ nbr,
                                                    const S* __restrict__ w,
                                                    const S* __restrict__ phi,
                                                    const S* __restrict__ T,
                                                    const S* __restrict__ Tb, S* __restrict__ Tf) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nv = F / 4;
    if (t < nv) {
        const int64_t f = 4 * t;
        const Vec4<int32_t> o = ld4(own + f), n = ld4(nbr + f);
        Vec4<S> lam;
        if (scheme == 0) {
            lam = ld4(w + f);
        } else {
            const Vec4<S> ph = ld4(phi + f);
            for (int k = 0; k < 4; ++k) lam.v[k] = ph.v[k] >= S(0) ? S(1) : S(0);
        }
        S TO[4], TN[4];
        for (int k = 0; k < 4; ++k) {
            TN[k] = T[n.v[k]];
            TO[k] = T[o.v[k]];
        }
        Vec4<S> out;
        for (int k = 0; k < 4; ++k) out.v[k] = lam.v[k] * (TO[k] - TN[k]) + TN[k];
        st4(Tf + f, out);
        return;
    }
    const int64_t i = 4 * nv + (t - nv);  // tail faces, then boundary faces
    if (i < F) {
        const S lam = scheme == 0 ? w[i] : (phi[i] >= S(0) ? S(1) : S(0));
        const S TN = T[nbr[i]];
        Tf[i] = lam * (T[own[i]] - TN) + TN;
    } else if (i < (int64_t)F + B) {
        Tf[i] = Tb[i - F];
    }
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_sngrad4(int F, int B, const int32_t* __restrict__ own,
                                                    const int32_t* __restrict__ nbr,
                                                    const int32_t* __restrict__ fc,
                                                    const S* __restrict__ nodc,
                                                    const S* __restrict__ dcb,
                                                    const int8_t* __restrict__ kind,
                                                    const S* __restrict__ T,
                                                    const S* __restrict__ Tb,
                                                    const S* __restrict__ Tg, S* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nv = F / 4;
    if (t < nv) {
        const int64_t f = 4 * t;
        const Vec4<int32_t> o = ld4(own + f), n = ld4(nbr + f);
        const Vec4<S> d = ld4(nodc + f);
        S TO[4], TN[4];
        for (int k = 0; k < 4; ++k) {
            TN[k] = T[n.v[k]];
            TO[k] = T[o.v[k]];
        }
        Vec4<S> r;
        for (int k = 0; k < 4; ++k) r.v[k] = d.v[k] * (TN[k] - TO[k]);
        st4(out + f, r);
        return;
    }
    const int64_t i = 4 * nv + (t - nv);
    if (i < F) {
        out[i] = nodc[i] * (T[nbr[i]] - T[own[i]]);
    } else if (i < (int64_t)F + B) {
        const int b = (int)(i - F);
        out[i] = bnd_sngrad<S>(kind[b], dcb[b], Tb[b], T[fc[b]], Tg, b);
    }
}

// threads of a k_*4 launch over F internal and B boundary faces
inline int64_t face4_threads(int F, int B) { return F / 4 + (F % 4) + (int64_t)B; }

template <typename S>
__global__ __launch_bounds__(kBlock) void k_flux(int F, int B, const int32_t* __restrict__ own,
                                                 const int32_t* __restrict__ nbr,
                                                 const S* __restrict__ w, const S* __restrict__ Sf,
                                                 const S* __restrict__ Sfb,
                                                 const S* __restrict__ U, const S* __restrict__ Ub,
                                                 S* __restrict__ phi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F) {
        const int64_t o = own[i], n = nbr[i];
        const S wf = w[i];
        const S ux = wf * (U[3 * o] - U[3 * n]) + U[3 * n];
        const S uy = wf * (U[3 * o + 1] - U[3 * n + 1]) + U[3 * n + 1];
        const S uz = wf * (U[3 * o + 2] - U[3 * n + 2]) + U[3 * n + 2];
        phi[i] = Sf[3 * (int64_t)i] * ux + Sf[3 * (int64_t)i + 1] * uy + Sf[3 * (int64_t)i + 2] * uz;
    } else if (i < F + B) {
        const int64_t b = i - F;
        phi[i] = Sfb[3 * b] * Ub[3 * b] + Sfb[3 * b + 1] * Ub[3 * b + 1] + Sfb[3 * b + 2] * Ub[3 * b + 2];
    }
}

// fixedGradient patches: Tb = T_P + g/deltaCoeffs (fixedGradientFvPatchField::evaluate);
// fixedValue and calculated faces keep their stored value
template <typename S>
__global__ __launch_bounds__(kBlock) void k_correct_bc(int B, const int32_t* __restrict__ fc,
                                                       const double* __restrict__ dcb,
                                                       const int8_t* __restrict__ kind,
                                                       const S* __restrict__ T,
                                                       const S* __restrict__ Tg,
                                                       S* __restrict__ Tb) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || kind[b] != FV_BC_FIXED_GRADIENT) return;
    const S g = Tg ? Tg[b] : S(0);
    Tb[b] = T[fc[b]] + g / (S)dcb[b];
}

// Operator drivers: the LDS-staged gather on upper-triangular meshes, the
// stencil walk otherwise (same arithmetic, same order).
template <typename S>
void run_grad(fv_ctx* ctx, const fv_mesh* m, Geo<S> g, const S* T, const S* Tb, S coeff, S* out) {
    if (gather_cells(m)) {
        launch_gather<S>(ctx, m, GradOp<S>{{}, T, g.w, g.Sf, g.Sfb, Tb}, g.V, coeff, out);
        return;
    }
    const int N = m->nCells;
    k_grad<S><<<grid_for(N, kBlock), kBlock, 0, ctx->stream>>>(
        N, m->nInt, m->cfOff.p, m->cfIdx.p, m->cfOther.p, g, T, Tb, coeff, out);
    FV_LAUNCH_CHECK();
}

template <typename S>
void run_div(fv_ctx* ctx, const fv_mesh* m, int scheme, Geo<S> g, const S* phi, const S* T,
             const S* Tb, S coeff, S* out) {
    if (gather_cells(m)) {
        if (!T)
            launch_gather<S>(ctx, m, IntegrateOp<S>{{}, phi, m->nInt}, g.V, coeff, out);
        else if (scheme == FV_LINEAR)
            launch_gather<S>(ctx, m, DivOp<S, 0>{{}, T, g.w, phi, Tb, m->nInt}, g.V, coeff, out);
        else
            launch_gather<S>(ctx, m, DivOp<S, 1>{{}, T, g.w, phi, Tb, m->nInt}, g.V, coeff, out);
        return;
    }
    const int N = m->nCells;
    k_div<S><<<grid_for(N, kBlock), kBlock, 0, ctx->stream>>>(
        N, m->nInt, scheme, m->cfOff.p, m->cfIdx.p, m->cfOther.p, g, phi, T, Tb, coeff, out);
    FV_LAUNCH_CHECK();
}

template <typename S>
void run_lap(fv_ctx* ctx, const fv_mesh* m, Geo<S> g, const int8_t* kind, const S* gamma,
             const S* T, const S* Tb, const S* Tg, S coeff, S* out) {
    if (gather_cells(m)) {
        launch_gather<S>(ctx, m,
                         LapOp<S>{{}, T, gamma, g.magSf, g.nodc, g.magSfb, g.dcb, Tb, Tg, kind, m->nInt},
                         g.V, coeff, out);
        return;
    }
    const int N = m->nCells;
    k_laplacian<S><<<grid_for(N, kBlock), kBlock, 0, ctx->stream>>>(
        N, m->nInt, m->cfOff.p, m->cfIdx.p, m->cfOther.p, g, kind, gamma, T, Tb, Tg, coeff, out);
    FV_LAUNCH_CHECK();
}

void check_common(fv_ctx* ctx, const fv_mesh* m, fv_prec prec) {
    FV_REQUIRE(ctx && m, "NULL ctx/mesh");
    FV_REQUIRE(m->ctx == ctx, "mesh belongs to another context");
    FV_REQUIRE(prec == FV_F64 || prec == FV_F32, "unknown precision");
}

void check_bc(const fv_mesh* m, const fv_bc* bc) {
    FV_REQUIRE(bc, "boundary conditions are NULL");
    FV_REQUIRE(bc->mesh == m, "boundary conditions belong to another mesh");
}

}  // namespace

extern "C" {

int fv_interpolate(fv_ctx* ctx, const fv_mesh* mc, fv_scheme scheme, fv_prec prec,
                   const void* phi, const void* T, const void* Tb, void* Tf) {
    return guarded([&] {
        fv_mesh* m = const_cast<fv_mesh*>(mc);
        check_common(ctx, m, prec);
        FV_REQUIRE(scheme == FV_LINEAR || scheme == FV_UPWIND, "unknown scheme");
        FV_REQUIRE(T && Tf && (Tb || m->nB == 0), "NULL field");
        FV_REQUIRE(scheme == FV_LINEAR || phi, "upwind interpolation needs phi");
        const int n = m->nInt + m->nB;
        if (!n) return;
        // the four-face kernel needs 16-byte aligned face streams (the caller's
        // phi and Tf may be offset views)
        const bool vec = aligned16(Tf) && (scheme == FV_LINEAR || aligned16(phi));
        const int64_t n4 = face4_threads(m->nInt, m->nB);
        if (prec == FV_F64) {
            if (vec)
                k_interp4<double><<<grid_for(n4, kBlock), kBlock, 0, ctx->stream>>>(
                    m->nInt, m->nB, scheme, m->own.p, m->nbr.p, m->w.p, (const double*)phi,
                    (const double*)T, (const double*)Tb, (double*)Tf);
            else
                k_interp<double><<<grid_for(n, kBlock), kBlock, 0, ctx->stream>>>(
                    m->nInt, m->nB, scheme, m->own.p, m->nbr.p, m->w.p, (const double*)phi,
                    (const double*)T, (const double*)Tb, (double*)Tf);
        } else {
            auto g = geo32(m);
            if (vec)
                k_interp4<float><<<grid_for(n4, kBlock), kBlock, 0, ctx->stream>>>(
                    m->nInt, m->nB, scheme, m->own.p, m->nbr.p, g.w, (const float*)phi,
                    (const float*)T, (const float*)Tb, (float*)Tf);
            else
                k_interp<float><<<grid_for(n, kBlock), kBlock, 0, ctx->stream>>>(
                    m->nInt, m->nB, scheme, m->own.p, m->nbr.p, g.w, (const float*)phi,
                    (const float*)T, (const float*)Tb, (float*)Tf);
        }
        FV_LAUNCH_CHECK();
    });
}

int fv_grad_gauss(fv_ctx* ctx, const fv_mesh* mc, fv_prec prec, double coeff, const void* T,
                  const void* Tb, void* out) {
    return guarded([&] {
        fv_mesh* m = const_cast<fv_mesh*>(mc);
        check_common(ctx, m, prec);
        FV_REQUIRE(T && out && (Tb || m->nB == 0), "NULL field");
        if (prec == FV_F64)
            run_grad<double>(ctx, m, geo64(m), (const double*)T, (const double*)Tb, coeff,
                             (double*)out);
        else
            run_grad<float>(ctx, m, geo32(m), (const float*)T, (const float*)Tb, (float)coeff,
                            (float*)out);
    });
}

int fv_div_gauss(fv_ctx* ctx, const fv_mesh* mc, fv_scheme scheme, fv_prec prec, double coeff,
                 const void* phi, const void* T, const void* Tb, void* out) {
    return guarded([&] {
        fv_mesh* m = const_cast<fv_mesh*>(mc);
        check_common(ctx, m, prec);
        FV_REQUIRE(scheme == FV_LINEAR || scheme == FV_UPWIND, "unknown scheme");
        FV_REQUIRE(phi && out, "NULL field");
        FV_REQUIRE(!T || Tb || m->nB == 0, "boundary values are NULL");
        if (prec == FV_F64)
            run_div<double>(ctx, m, scheme, geo64(m), (const double*)phi, (const double*)T,
                            (const double*)Tb, coeff, (double*)out);
        else
            run_div<float>(ctx, m, scheme, geo32(m), (const float*)phi, (const float*)T,
                           (const float*)Tb, (float)coeff, (float*)out);
    });
}

int fv_laplacian_uncorrected(fv_ctx* ctx, const fv_mesh* mc, fv_prec prec, double coeff,
                             const void* gamma, const void* T, const fv_bc* bc, const void* Tb,
                             const void* Tg, void* out) {
    return guarded([&] {
        fv_mesh* m = const_cast<fv_mesh*>(mc);
        check_common(ctx, m, prec);
        check_bc(m, bc);
        FV_REQUIRE(gamma && T && out && (Tb || m->nB == 0), "NULL field");
        if (prec == FV_F64)
            run_lap<double>(ctx, m, geo64(m), bc->faceKind.p, (const double*)gamma,
                            (const double*)T, (const double*)Tb, (const double*)Tg, coeff,
                            (double*)out);
        else
            run_lap<float>(ctx, m, geo32(m), bc->faceKind.p, (const float*)gamma,
                           (const float*)T, (const float*)Tb, (const float*)Tg, (float)coeff,
                           (float*)out);
    });
}

int fv_sngrad_uncorrected(fv_ctx* ctx, const fv_mesh* mc, fv_prec prec, const void* T,
                          const fv_bc* bc, const void* Tb, const void* Tg, void* out) {
    return guarded([&] {
        fv_mesh* m = const_cast<fv_mesh*>(mc);
        check_common(ctx, m, prec);
        check_bc(m, bc);
        FV_REQUIRE(T && out && (Tb || m->nB == 0), "NULL field");
        const int n = m->nInt + m->nB;
        if (!n) return;
        const bool vec = aligned16(out);  // the four-face kernel stores 16 bytes at a time
        const int64_t n4 = face4_threads(m->nInt, m->nB);
        if (prec == FV_F64) {
            if (vec)
                k_sngrad4<double><<<grid_for(n4, kBlock), kBlock, 0, ctx->stream>>>(
                    m->nInt, m->nB, m->own.p, m->nbr.p, m->faceCells.p, m->nodc.p, m->dcb.p,
                    bc->faceKind.p, (const double*)T, (const double*)Tb, (const double*)Tg,
                    (double*)out);
            else
                k_sngrad<double><<<grid_for(n, kBlock), kBlock, 0, ctx->stream>>>(
                    m->nInt, m->nB, m->own.p, m->nbr.p, m->faceCells.p, m->nodc.p, m->dcb.p,
                    bc->faceKind.p, (const double*)T, (const double*)Tb, (const double*)Tg,
                    (double*)out);
        } else {
            auto g = geo32(m);
            if (vec)
                k_sngrad4<float><<<grid_for(n4, kBlock), kBlock, 0, ctx->stream>>>(
                    m->nInt, m->nB, m->own.p, m->nbr.p, m->faceCells.p, g.nodc, g.dcb,
                    bc->faceKind.p, (const float*)T, (const float*)Tb, (const float*)Tg,
                    (float*)out);
            else
                k_sngrad<float><<<grid_for(n, kBlock), kBlock, 0, ctx->stream>>>(
                    m->nInt, m->nB, m->own.p, m->nbr.p, m->faceCells.p, g.nodc, g.dcb,
                    bc->faceKind.p, (const float*)T, (const float*)Tb, (const float*)Tg,
                    (float*)out);
        }
        FV_LAUNCH_CHECK();
    });
}

int fv_flux(fv_ctx* ctx, const fv_mesh* mc, fv_prec prec, const void* U, const void* Ub,
            void* phi) {
    return guarded([&] {
        fv_mesh* m = const_cast<fv_mesh*>(mc);
        check_common(ctx, m, prec);
        FV_REQUIRE(U && phi && (Ub || m->nB == 0), "NULL field");
        const int n = m->nInt + m->nB;
        if (!n) return;
        if (prec == FV_F64)
            k_flux<double><<<grid_for(n, kBlock), kBlock, 0, ctx->stream>>>(
                m->nInt, m->nB, m->own.p, m->nbr.p, m->w.p, m->Sf.p, m->Sfb.p, (const double*)U,
                (const double*)Ub, (double*)phi);
        else {
            auto g = geo32(m);
            k_flux<float><<<grid_for(n, kBlock), kBlock, 0, ctx->stream>>>(
                m->nInt, m->nB, m->own.p, m->nbr.p, g.w, g.Sf, g.Sfb, (const float*)U,
                (const float*)Ub, (float*)phi);
        }
        FV_LAUNCH_CHECK();
    });
}

int fv_correct_boundary(fv_ctx* ctx, const fv_mesh* mc, fv_prec prec, const void* T,
                        const fv_bc* bc, const void* Tg, void* Tb) {
    return guarded([&] {
        fv_mesh* m = const_cast<fv_mesh*>(mc);
        check_common(ctx, m, prec);
        check_bc(m, bc);
        FV_REQUIRE(T && (Tb || m->nB == 0), "NULL field");
        if (!m->nB) return;
        if (prec == FV_F64)
            k_correct_bc<double><<<grid_for(m->nB, kBlock), kBlock, 0, ctx->stream>>>(
                m->nB, m->faceCells.p, m->dcb.p, bc->faceKind.p, (const double*)T,
                (const double*)Tg, (double*)Tb);
        else
            k_correct_bc<float><<<grid_for(m->nB, kBlock), kBlock, 0, ctx->stream>>>(
                m->nB, m->faceCells.p, m->dcb.p, bc->faceKind.p, (const float*)T,
                (const float*)Tg, (float*)Tb);
        FV_LAUNCH_CHECK();
    });
}

}  // extern "C"
