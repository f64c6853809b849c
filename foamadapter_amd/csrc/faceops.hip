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
// A workgroup owns CELLS consecutive cells [c0, c1) (256, 512 or 1024).  In OpenFOAM face order the
// internal faces those cells own are one contiguous range [fs, fe), so:
//   phase 1: the workgroup streams that face range with coalesced loads,
//            computes every face contribution once and keeps it in LDS;
//   phase 2: each thread takes cells in turn and accumulates, in ascending face order, its
//            lower faces (owned by lower-numbered cells: read from LDS when the
//            owner is in this workgroup, recomputed otherwise), its own faces
//            (from LDS), then its boundary faces, and divides by V.
// The accumulation order is the serial face order of the other kernels, so the
// result is the same bit for bit.  On a Z-order renumbered mesh a workgroup is a
// compact brick (8x8x4, 8x8x8 or 16x8x8 cells) and only the faces on its lower
// sides are recomputed; larger bricks recompute fewer faces per cell.
// Workgroups run in their natural order, which the hardware deals round-robin
// over the 8 XCDs: measured at 256^3 this is as fast as or faster (FP64
// gradient +7%) than remapping each XCD onto one contiguous range of bricks.

constexpr int kU = 4;  // faces per thread whose loads a staged-gather phase 1 keeps in flight
// LDS face slots of a CELLS-cell workgroup: 3 owned faces per hex cell plus slack;
// faces beyond it (unusual meshes) are recomputed in phase 2
template <int CELLS>
constexpr int max_up() { return CELLS * 3 + CELLS / 4; }

struct Split {
    const int32_t *own, *nbr, *ownStart, *lowOff, *lowFace, *lowOwn, *bOff, *bFace;
};

// An operator gives each face a contribution of NC components.  The staged
// gather keeps a face's NS-component state in LDS (state/bstate) and expands it
// to the contribution when it accumulates (contrib/bcontrib): the gradient keeps
// the face value (8 bytes instead of 24) and multiplies by Sf when it adds.
template <typename S>
struct GradOp {
    static constexpr int NC = 3, NS = 1;
    static constexpr bool kLists = false;  // runs k_gather_stencil
    static constexpr int kDefaultCells = 512;
    const S *T, *w, *Sf, *Sfb, *Tb;
    __device__ void state(int f, int o, int n, S* s) const {
        const S TP = T[o], TN = T[n];
        s[0] = w[f] * (TP - TN) + TN;
    }
    __device__ void contrib(int f, const S* s, S* v) const {
        v[0] = Sf[3 * (int64_t)f] * s[0];
        v[1] = Sf[3 * (int64_t)f + 1] * s[0];
        v[2] = Sf[3 * (int64_t)f + 2] * s[0];
    }
    __device__ void bstate(int b, int, S* s) const { s[0] = Tb[b]; }
    __device__ void bcontrib(int b, const S* s, S* v) const {
        v[0] = Sfb[3 * b] * s[0];
        v[1] = Sfb[3 * b + 1] * s[0];
        v[2] = Sfb[3 * b + 2] * s[0];
    }
    __device__ void face(int f, int o, int n, S* v) const {
        S s[NS];
        state(f, o, n, s);
        contrib(f, s, v);
    }
    __device__ void bface(int b, int c, S* v) const {
        S s[NS];
        bstate(b, c, s);
        bcontrib(b, s, v);
    }
};

// The gradient on the LDS-list kernel: the state is the whole contribution
// (three components), so phase 2 reads no geometry
template <typename S>
struct GradListOp : GradOp<S> {
    static constexpr int NS = 3;
    static constexpr bool kLists = true;
    static constexpr int kDefaultCells = 256;
    __device__ void state(int f, int o, int n, S* s) const { GradOp<S>::face(f, o, n, s); }
    __device__ void contrib(int, const S* s, S* v) const {
        v[0] = s[0];
        v[1] = s[1];
        v[2] = s[2];
    }
    __device__ void bstate(int b, int c, S* s) const { GradOp<S>::bface(b, c, s); }
    __device__ void bcontrib(int, const S* s, S* v) const {
        v[0] = s[0];
        v[1] = s[1];
        v[2] = s[2];
    }
};

// operators whose state is the contribution itself
template <typename S, typename D>
struct PlainOp {
    static constexpr int NC = 1, NS = 1;
    static constexpr bool kLists = true;  // runs k_gather
    static constexpr int kDefaultCells = 256;
    __device__ void state(int f, int o, int n, S* s) const {
        static_cast<const D*>(this)->face(f, o, n, s);
    }
    __device__ void contrib(int, const S* s, S* v) const { v[0] = s[0]; }
    __device__ void bstate(int b, int c, S* s) const { static_cast<const D*>(this)->bface(b, c, s); }
    __device__ void bcontrib(int, const S* s, S* v) const { v[0] = s[0]; }
};

template <typename S, int SCHEME>  // 0 linear, 1 upwind
struct DivOp : PlainOp<S, DivOp<S, SCHEME>> {
    const S *T, *w, *phi, *Tb;
    int F;
    __device__ void face(int f, int o, int n, S* v) const {
        const S ph = phi[f];
        const S lam = SCHEME == 0 ? w[f] : (ph >= S(0) ? S(1) : S(0));
        const S TP = T[o], TN = T[n];
        v[0] = ph * (lam * (TP - TN) + TN);
    }
    __device__ void bface(int b, int, S* v) const { v[0] = phi[F + b] * Tb[b]; }
};

template <typename S>
struct IntegrateOp : PlainOp<S, IntegrateOp<S>> {  // surfaceIntegrate(phi): div with a unit field
    const S* phi;
    int F;
    __device__ void face(int f, int, int, S* v) const { v[0] = phi[f]; }
    __device__ void bface(int b, int, S* v) const { v[0] = phi[F + b]; }
};

template <typename S>
struct LapOp : PlainOp<S, LapOp<S>> {
    const S *T, *gamma, *magSf, *nodc, *magSfb, *dcb, *Tb, *Tg;
    const int8_t* kind;
    int F;
    __device__ void face(int f, int o, int n, S* v) const {
        v[0] = (gamma[f] * magSf[f]) * (nodc[f] * (T[n] - T[o]));
    }
    __device__ void bface(int b, int c, S* v) const {
        const S sn = kind[b] == FV_BC_FIXED_GRADIENT ? (Tg ? Tg[b] : S(0)) : dcb[b] * (Tb[b] - T[c]);
        v[0] = (gamma[F + b] * magSfb[b]) * sn;
    }
};

// One cell of the per-cell path: LDS holds, for the workgroup's first nUp own
// faces, the face state (kState, NS components) or the contribution (NC).
template <typename S, typename Op, bool kState>
__device__ inline void gather_cell(int c, int fs, int nUp, const Split& t, const Op& op,
                                   S vol, S coeff, S* __restrict__ out, const S* st0, int stride) {
    constexpr int NC = Op::NC, NS = kState ? Op::NS : Op::NC;
    S acc[NC];
    for (int j = 0; j < NC; ++j) acc[j] = S(0);
    const int l1 = t.lowOff[c + 1];
    for (int k = t.lowOff[c]; k < l1; ++k) {
        const int f = t.lowFace[k];
        S v[NC];
        const int i = f - fs;
        if (i >= 0 && i < nUp) {
            S sv[NS];
            for (int j = 0; j < NS; ++j) sv[j] = st0[j * stride + i];
            if (kState)
                op.contrib(f, sv, v);
            else
                for (int j = 0; j < NC; ++j) v[j] = sv[j];
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
            S sv[NS];
            for (int j = 0; j < NS; ++j) sv[j] = st0[j * stride + i];
            if (kState)
                op.contrib(f, sv, v);
            else
                for (int j = 0; j < NC; ++j) v[j] = sv[j];
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

// Staged gather with the per-cell stencil lists read from HBM: phase 1 keeps the
// contribution (NC components) of every own face in LDS, phase 2 runs
// gather_cell per cell.  Lighter in LDS and in phase-1 work than k_gather, it
// is the faster variant for the gradient (three components per face), whose
// phase 2 would otherwise re-read Sf per face.
template <typename S, typename Op, int CELLS, int THREADS>
__global__ __launch_bounds__(THREADS) void k_gather_stencil(int N, int nBlocks,
                                                            Split t, Op op,
                                                            const S* __restrict__ V, S coeff,
                                                            S* __restrict__ out) {
    constexpr int NC = Op::NC;
    constexpr int kMaxUp = max_up<CELLS>();
    __shared__ S flux[NC][kMaxUp];
    const int lb = (int)blockIdx.x;
    if (lb >= nBlocks) return;
    const int c0 = lb * CELLS, c1 = min(c0 + CELLS, N);
    const int fs = t.ownStart[c0], fe = t.ownStart[c1];
    const int nUp = min(fe - fs, kMaxUp);
    // the volumes phase 2 divides by, loaded now so their latency hides behind phase 1
    constexpr int kPer = CELLS / THREADS;
    S vols[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int c = c0 + (int)threadIdx.x + q * THREADS;
        vols[q] = c < c1 ? V[c] : S(1);
    }
    // kU faces per thread in flight: their loads are issued together (lanes past
    // the range work on the last face and store nothing)
    for (int i0 = threadIdx.x; i0 < nUp; i0 += THREADS * kU) {
        int oo[kU], nn[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int f = fs + min(i0 + u * THREADS, nUp - 1);
            oo[u] = t.own[f];
            nn[u] = t.nbr[f];
        }
        S v[kU][NC];
#pragma unroll
        for (int u = 0; u < kU; ++u) op.face(fs + min(i0 + u * THREADS, nUp - 1), oo[u], nn[u], v[u]);
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = i0 + u * THREADS;
            if (i < nUp)
                for (int j = 0; j < NC; ++j) flux[j][i] = v[u][j];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int c = c0 + (int)threadIdx.x + q * THREADS;
        if (c < c1) gather_cell<S, Op, false>(c, fs, nUp, t, op, vols[q], coeff, out, flux[0], kMaxUp);
    }
}

struct Lists {
    const int32_t *extOff, *extFace, *extOwn, *bndOff, *bndFace;
    const uint16_t *extCell, *bndCell;
};

constexpr int kMaxLow = 8;  // lower faces per cell from inside the workgroup, in LDS

// A workgroup of CELLS cells.  Phase 1 computes, with coalesced loads, every
// face contribution the workgroup needs into LDS -- its own faces (streamed),
// its external lower faces and its boundary faces (from the per-workgroup
// lists) -- and builds each cell's face lists in LDS: the own-face range from
// the owner sequence, the lower faces from inside the workgroup by LDS atomics
// (then sorted, so the order is fixed), the external and boundary ranges from
// the lists (sorted by cell).  Phase 2 accumulates each cell in ascending face
// order (external lower, inner lower, own, boundary) and divides by V.  A
// workgroup whose faces do not fit the LDS arrays takes the per-cell path of
// gather_cell instead; both give the same bits.
template <typename S, typename Op, int CELLS, int THREADS>
__global__ __launch_bounds__(THREADS) void k_gather(int N, int nBlocks, Split t,
                                                    Lists L, Op op, const S* __restrict__ V,
                                                    S coeff, S* __restrict__ out) {
    constexpr int NC = Op::NC, NS = Op::NS;
    constexpr int kMaxUp = max_up<CELLS>();
    constexpr int kMaxExt = CELLS / 2;
    constexpr int kMaxBnd = CELLS / 4;
    __shared__ S fluxU[NS][kMaxUp];
    __shared__ S fluxE[NS][kMaxExt];
    __shared__ S fluxB[NS][kMaxBnd];
    __shared__ uint16_t upS[CELLS], upE[CELLS], exS[CELLS], exE[CELLS], bS[CELLS], bE[CELLS];
    __shared__ uint16_t lowL[CELLS][kMaxLow];
    __shared__ int lowN[CELLS];
    __shared__ int overflow;
    const int lb = (int)blockIdx.x;
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
    for (int l = threadIdx.x; l < CELLS; l += THREADS) {
        upS[l] = upE[l] = exS[l] = exE[l] = bS[l] = bE[l] = 0;
        lowN[l] = 0;
    }
    if (threadIdx.x == 0) overflow = fits ? 0 : 1;
    __syncthreads();
    const int nUp = min(nF, kMaxUp);
    for (int i0 = threadIdx.x; i0 < nUp; i0 += THREADS * kU) {
      // kU faces per thread in flight (see k_gather_stencil)
      int oo[kU], nn[kU];
      S svs[kU][NS];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
          const int f = fs + min(i0 + u * THREADS, nUp - 1);
          oo[u] = t.own[f];
          nn[u] = t.nbr[f];
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) op.state(fs + min(i0 + u * THREADS, nUp - 1), oo[u], nn[u], svs[u]);
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int i = i0 + u * THREADS;
        if (i >= nUp) break;
        const int f = fs + i;
        const int o = oo[u], n = nn[u];
        for (int j = 0; j < NS; ++j) fluxU[j][i] = svs[u][j];
        if (fits) {
            const int ol = o - c0;
            if (i == 0 || t.own[f - 1] != o) upS[ol] = (uint16_t)i;
            if (i == nF - 1 || t.own[f + 1] != o) upE[ol] = (uint16_t)(i + 1);
            if (n < c1) {
                const int nl = n - c0;
                const int slot = atomicAdd(&lowN[nl], 1);
                if (slot < kMaxLow)
                    lowL[nl][slot] = (uint16_t)i;
                else
                    overflow = 1;
            }
        }
      }
    }
    if (fits) {
        for (int k = e0 + threadIdx.x; k < e1; k += THREADS) {
            const int cl = L.extCell[k];
            S sv[NS];
            op.state(L.extFace[k], L.extOwn[k], c0 + cl, sv);
            for (int j = 0; j < NS; ++j) fluxE[j][k - e0] = sv[j];
            if (k == e0 || L.extCell[k - 1] != cl) exS[cl] = (uint16_t)(k - e0);
            if (k == e1 - 1 || L.extCell[k + 1] != cl) exE[cl] = (uint16_t)(k - e0 + 1);
        }
        for (int k = b0 + threadIdx.x; k < b1; k += THREADS) {
            const int cl = L.bndCell[k];
            S sv[NS];
            op.bstate(L.bndFace[k], c0 + cl, sv);
            for (int j = 0; j < NS; ++j) fluxB[j][k - b0] = sv[j];
            if (k == b0 || L.bndCell[k - 1] != cl) bS[cl] = (uint16_t)(k - b0);
            if (k == b1 - 1 || L.bndCell[k + 1] != cl) bE[cl] = (uint16_t)(k - b0 + 1);
        }
    }
    __syncthreads();
    if (overflow) {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int c = c0 + (int)threadIdx.x + q * THREADS;
            if (c < c1) gather_cell<S, Op, true>(c, fs, nUp, t, op, vols[q], coeff, out, fluxU[0], kMaxUp);
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
        S sv[NS], v[NC];
        for (int k = exS[l]; k < exE[l]; ++k) {
            for (int j = 0; j < NS; ++j) sv[j] = fluxE[j][k];
            op.contrib(L.extFace[e0 + k], sv, v);
            for (int j = 0; j < NC; ++j) acc[j] -= v[j];
        }
        // the inner lower faces arrived in any order: sort this cell's list
        const int n = lowN[l];
        uint16_t* lst = lowL[l];
        for (int a = 1; a < n; ++a) {
            const uint16_t v = lst[a];
            int b = a - 1;
            while (b >= 0 && lst[b] > v) {
                lst[b + 1] = lst[b];
                --b;
            }
            lst[b + 1] = v;
        }
        for (int q = 0; q < n; ++q) {
            const int i = lst[q];
            for (int j = 0; j < NS; ++j) sv[j] = fluxU[j][i];
            op.contrib(fs + i, sv, v);
            for (int j = 0; j < NC; ++j) acc[j] -= v[j];
        }
        for (int i = upS[l]; i < upE[l]; ++i) {
            for (int j = 0; j < NS; ++j) sv[j] = fluxU[j][i];
            op.contrib(fs + i, sv, v);
            for (int j = 0; j < NC; ++j) acc[j] += v[j];
        }
        for (int k = bS[l]; k < bE[l]; ++k) {
            for (int j = 0; j < NS; ++j) sv[j] = fluxB[j][k];
            op.bcontrib(L.bndFace[b0 + k], sv, v);
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
    if (!Op::kLists) {
        k_gather_stencil<S, Op, CELLS, THREADS><<<nBlocks, THREADS, 0, ctx->stream>>>(
            m->nCells, nBlocks, split_of(m), op, V, coeff, out);
        FV_LAUNCH_CHECK();
        return;
    }
    // the lists are built once per mesh and workgroup size (a cache on the mesh)
    const GatherLists& gl = gather_lists(const_cast<fv_mesh*>(m), CELLS);
    const Lists L{gl.extOff.p, gl.extFace.p, gl.extOwn.p, gl.bndOff.p, gl.bndFace.p,
                  gl.extCell.p, gl.bndCell.p};
    k_gather<S, Op, CELLS, THREADS><<<nBlocks, THREADS, 0, ctx->stream>>>(
        m->nCells, nBlocks, split_of(m), L, op, V, coeff, out);
    FV_LAUNCH_CHECK();
}

template <typename S, typename Op>
void launch_gather(fv_ctx* ctx, const fv_mesh* m, const Op& op, const S* V, S coeff, S* out) {
    // default workgroup size per operator (measured at 256^3: the gradient's
    // stencil-list kernel is fastest with 512 cells, the LDS-list kernel with 256)
    const int cells = m->gatherMode == 0 ? Op::kDefaultCells : gather_cells(m);
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
        // measured at 256^3: FP64 is fastest on the stencil-list kernel (512-cell
        // workgroups), FP32 on the LDS-list kernel (256): 3.7 vs 3.6 TB/s and
        // 3.5 vs 4.1 TB/s algorithmic
        if (sizeof(S) == 8)
            launch_gather<S>(ctx, m, GradOp<S>{T, g.w, g.Sf, g.Sfb, Tb}, g.V, coeff, out);
        else
            launch_gather<S>(ctx, m, GradListOp<S>{{T, g.w, g.Sf, g.Sfb, Tb}}, g.V, coeff, out);
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
        if (prec == FV_F64) {
            k_interp<double><<<grid_for(n, kBlock), kBlock, 0, ctx->stream>>>(
                m->nInt, m->nB, scheme, m->own.p, m->nbr.p, m->w.p, (const double*)phi,
                (const double*)T, (const double*)Tb, (double*)Tf);
        } else {
            auto g = geo32(m);
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
        if (prec == FV_F64)
            k_sngrad<double><<<grid_for(n, kBlock), kBlock, 0, ctx->stream>>>(
                m->nInt, m->nB, m->own.p, m->nbr.p, m->faceCells.p, m->nodc.p, m->dcb.p,
                bc->faceKind.p, (const double*)T, (const double*)Tb, (const double*)Tg,
                (double*)out);
        else {
            auto g = geo32(m);
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
