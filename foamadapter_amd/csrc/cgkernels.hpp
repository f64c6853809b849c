// cgkernels.hpp -- the Jacobi-preconditioned CG iteration as row-tile kernels,
// shared by the single-system solver (fv_solve_cg) and the batched engine.
//
// Jacobi preconditioning is applied as a symmetric scaling: with g_i the signed
// square root of the diagonal a_ii (g_i |g_i| = a_ii) and G = diag(g), the
// kernels run plain CG on  As y = bs,  As = G^-1 A |G|^-1,  bs = G^-1 b,
// x = |G|^-1 y.  As is symmetric with a unit diagonal when A's diagonal has one
// sign (every symmetric FV operator: ddt + laplacian, laplacian), and the
// iterates are those of Jacobi-PCG on A (r = G rs, z = D^-1 r = |G|^-1 rs,
// p = |G|^-1 ps, r.z = rs.rs, p.Ap = ps.As ps): the preconditioner costs no
// D^-1 loads and no z stores, and the SpMV reads no diagonal entries (the
// sliced ELL holds the off-diagonal entries; the unit diagonal is added).  The
// stop test uses the unscaled residual, ||r||^2 = sum (g_i rs_i)^2.  A zero,
// non-finite or opposite-sign diagonal entry sets g_i to NaN, so such a
// system's solve fails (non-finite residual) rather than running another method.
//
// A launch covers a list of tiles; a tile is a contiguous range of one
// system's rows (a multiple of 64) and owns one partial-sum slot.  The single-
// system solver uses implicit tiles (logical block b covers rows [b*R, b*R+R));
// the batched engine passes an explicit tile table over many systems.  Logical
// blocks are numbered XCD-contiguously (hardware block i runs on XCD i % 8), so
// neighbouring row tiles -- and the SpMV gathers between them -- share an L2.
//
// A chunk of n iterations is 2n + 1 kernels:
//   spmv       q = As p (sliced ELL + unit diagonal), partials of p.q
//                                                            (first iteration)
//   update     alpha from the p.q partials; d += alpha p, r -= alpha q,
//              partials of ||G r||^2 and r.r
//   dirspmv    beta and the stop decision from those partials; the new
//              direction p' = r + beta p of its own rows (stored) and of the
//              SpMV's columns (recomputed from r and p, the same expression),
//              q = As p', partials of p'.q                   (iterations 2..n)
//   update
//   direction  decision and p = r + beta p after the last update
// The search direction alternates between two buffers inside a chunk (dirspmv
// reads one and writes the other, since its SpMV reads the old direction of
// neighbouring rows); the chunk starts and ends with p in the first buffer.
// The results are bitwise those of a separate direction kernel per iteration,
// which large FP64 systems use instead (spmv, update, direction per iteration;
// see cg_fuse_direction).
// The FP64 residual replacement of the mixed solver is
//   accum      x += |G|^-1 d, d = 0
//   resid      r64 = b - A x (FP64 CSR, unscaled), partials of r.r and b.b
//   replace    relative residual from those partials; r = G^-1 r64,
//              partials of ||G r||^2 and r.r (a direction kernel in mode 1 follows)
// The replacement runs only for systems that stopped (stop == 1: the FP32
// residual fell by the trigger factor, its estimate reached the tolerance, or
// the iteration limit); the others go on iterating across chunks.
// Every dot product is summed by each consuming block from the producer's
// slots in a fixed order: deterministic, no atomics, no fences.
#pragma once

#include "cgstate.hpp"

namespace fvhip {
// internal linkage: every translation unit that launches these owns its copies
namespace {

template <typename S>
struct CgSys {
    int32_t N, tile0, ntiles, pad;  // rows; the system's slots [tile0, tile0 + ntiles)
    const int64_t* sellOff;         // sliced ELL of the off-diagonal entries, 64-row
    const int32_t* sellCol;         // slices, entries column-major
    const S* sval;                  // their scaled values a_ij / (g_i |g_j|)
    const int32_t* sellSrc;         // CSR index of each sliced-ELL entry (-1: padding)
    const int32_t* diag;            // CSR slot of the diagonal within each row
    const int32_t* rowOffs;         // FP64 CSR (residual replacement)
    const int32_t* colIdxs;
    const double* vals;
    const double* rhs;
    double* x;
    double* xold;                   // optional copy target after the solve (may be null)
    S *d, *r, *p, *q, *g;           // d: scaled correction (x += |G|^-1 d); g: signed
                                    // square roots of the diagonal
    S* p2;                          // second search-direction buffer
    double* r64;
    CgState* st;
};

struct CgTile {
    int32_t prob, row0, row1, slot;  // prob < 0: empty entry
};

template <typename S>
struct CgGrid {
    const CgTile* tiles;   // null: implicit tiles of probs[0]
    const CgSys<S>* probs;
    int32_t nTiles;        // tiles (explicit) or blocks with rows (implicit)
    int32_t rowsPerBlock;  // implicit tiles
};

// Per-slot partial sums: p.q (1 per slot), r.r and r.z (2), FP64 r.r and b.b (2)
struct CgPartials {
    double *pq, *rz, *rb;
};

// The launch description lives in device memory, so a captured hipGraph stays
// valid when the batch's tile table changes (only the table and the count are
// rewritten); a grid larger than the tile count leaves the extra blocks idle.
template <typename S>
__device__ inline bool cg_locate(const CgGrid<S>* __restrict__ gp, const CgSys<S>*& P, int& row0,
                                 int& row1, int& slot) {
    const CgGrid<S> g = *gp;
    const int per = gridDim.x >> 3;  // the grid is a multiple of 8 blocks
    const int lb = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (lb >= g.nTiles) return false;
    if (g.tiles) {
        const CgTile t = g.tiles[lb];
        if (t.prob < 0) return false;  // padding entry of the launch table
        P = g.probs + t.prob;
        row0 = t.row0;
        row1 = t.row1;
        slot = t.slot;
    } else {
        P = g.probs;
        row0 = lb * g.rowsPerBlock;
        row1 = min(row0 + g.rowsPerBlock, (P->N + 63) & ~63);
        slot = lb;
    }
    return true;
}

// rows per thread whose vector loads the update and direction kernels issue together
constexpr int kCgU = 4;

template <typename S>
__device__ inline S* cg_pbuf(const CgSys<S>& P, int b) {
    return b ? P.p2 : P.p;
}

// Software pipeline of the sliced-ELL rows of one thread: fetch() issues a
// row's slice offsets and its first K values and columns into the "next"
// registers, promote() makes them the current row.  The caller issues the
// current row's gathers and fetches the row after it before it waits for
// them, so a row costs about one memory round trip instead of three (measured
// at 128^3: FP64 solve 72 -> 67 ms, mixed 49 -> 45 ms; fetching two rows
// ahead in FP32 gains nothing more).
template <typename S, int K>
struct CgRowPipe {
    S v[K] = {}, vn[K] = {};
    int col[K] = {}, cn[K] = {};
    int64_t b = 0, bn = 0;  // entry 0 of the row (slice start + lane)
    int w = 0, wn = 0;      // slice width
    __device__ void fetch(const int64_t* __restrict__ soff, const int32_t* __restrict__ scol,
                          const S* __restrict__ sval, int r) {
        const int sl = r >> 6;
        const int64_t s0 = soff[sl];
        bn = s0 + (r & 63);
        wn = (int)((soff[sl + 1] - s0) >> 6);
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (k < wn) {
                vn[k] = sval[bn + 64 * (int64_t)k];
                cn[k] = scol[bn + 64 * (int64_t)k];
            }
    }
    __device__ void promote() {
        b = bn;
        w = wn;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            v[k] = vn[k];
            col[k] = cn[k];
        }
    }
};

// MODE-0 stop decision from ||r|| (rn) after `it` iterations
__device__ inline int cg_decide(const CgState* st, double rn, int it) {
    const double est = rn / st->bnorm;
    return (est <= st->tol || rn <= st->atol || it >= st->maxit || !(est == est) ||
            (st->delta > 0.0 && rn < st->delta * st->rrep))
               ? 1
               : 0;
}

template <typename S>
__global__ __launch_bounds__(kCgBlock) void kc_spmv(const CgGrid<S>* __restrict__ g, CgPartials red) {
    __shared__ double sh[kCgShared];
    const CgSys<S>* Pp;
    int row0, row1, slot;
    if (!cg_locate(g, Pp, row0, row1, slot)) return;
    const CgSys<S>& P = *Pp;
    if (P.st->stop) return;
    // One row per thread and pass, the first kSpE off-diagonal entries with their loads
    // issued together (values and columns, then the gathers of p), and the
    // rows software-pipelined: the next row's slice offsets, values and
    // columns are in flight while this row's gathers complete, so a row costs
    // about one memory round trip instead of three.  The row sums its entries
    // in column order.
    constexpr int kSpE = 8;
    const int64_t* __restrict__ soff = P.sellOff;
    const int32_t* __restrict__ scol = P.sellCol;
    const S* __restrict__ sval = P.sval;
    const S* __restrict__ pvec = P.p;
    S* __restrict__ qvec = P.q;
    const int N = P.N;
    CgRowPipe<S, kSpE> rp;
    double pq = 0.0;
    const int r0 = row0 + (int)threadIdx.x;
    if (r0 < row1) rp.fetch(soff, scol, sval, r0);
    for (int r = r0; r < row1; r += kCgBlock) {
        rp.promote();
        S pv[kSpE];
        const S own = r < N ? pvec[r] : S(0);
#pragma unroll
        for (int k = 0; k < kSpE; ++k)
            if (k < rp.w) pv[k] = pvec[rp.col[k]];
        if (r + kCgBlock < row1) rp.fetch(soff, scol, sval, r + kCgBlock);
        S sum = own;  // the unit diagonal, then the off-diagonal entries in column order
#pragma unroll
        for (int k = 0; k < kSpE; ++k)
            if (k < rp.w) sum += rp.v[k] * pv[k];
        for (int k = kSpE; k < rp.w; ++k) {  // rows wider than kSpE entries
            const int64_t pos = rp.b + 64 * (int64_t)k;
            sum += sval[pos] * pvec[scol[pos]];
        }
        if (r < N) {
            qvec[r] = sum;
            pq += (double)own * (double)sum;
        }
    }
    double v[1] = {pq};
    write_slot<1>(v, red.pq, slot, sh);
}

// Direction of the previous iteration fused with this iteration's SpMV.  pOld
// and pNew select the search-direction buffers (uniform over the launch).  A
// thread's first row is loaded (values, columns, then z and the old p of the
// columns) before the reduction, which hides the reduction's latency.
template <typename S>
__global__ __launch_bounds__(kCgBlock) void kc_dirspmv(const CgGrid<S>* __restrict__ g,
                                                       CgPartials red, int pOld, int pNew) {
    __shared__ double sh[kCgShared];
    const CgSys<S>* Pp;
    int row0, row1, slot;
    if (!cg_locate(g, Pp, row0, row1, slot)) return;
    const CgSys<S>& P = *Pp;
    CgState* st = P.st;
    if (st->done) return;
    constexpr int kSpE = 8;
    const int64_t* __restrict__ soff = P.sellOff;
    const int32_t* __restrict__ scol = P.sellCol;
    const S* __restrict__ sval = P.sval;
    const S* __restrict__ zv = P.r;  // z = r on the scaled system
    const S* __restrict__ po = cg_pbuf(P, pOld);
    S* __restrict__ pn = cg_pbuf(P, pNew);
    S* __restrict__ qvec = P.q;
    const int N = P.N;
    // rows software-pipelined as in kc_spmv: gather() promotes the fetched row
    // and issues its gathers of z and p; the row after it is fetched before
    // the current one is summed
    CgRowPipe<S, kSpE> rp;
    S zc[kSpE], pc[kSpE], zr = S(0), pr = S(0);
    auto gather = [&](int r) {
        rp.promote();
        if (r < N) {
            zr = zv[r];
            pr = po[r];
        }
#pragma unroll
        for (int k = 0; k < kSpE; ++k)
            if (k < rp.w) {
                zc[k] = zv[rp.col[k]];
                pc[k] = po[rp.col[k]];
            }
    };
    const int r0 = row0 + (int)threadIdx.x;
    if (r0 < row1) {
        rp.fetch(soff, scol, sval, r0);
        gather(r0);
        if (r0 + kCgBlock < row1) rp.fetch(soff, scol, sval, r0 + kCgBlock);
    }
    double tot[2];
    sum_slots<2>(red.rz, P.tile0, P.ntiles, sh, tot);
    const double rn = sqrt(tot[0]);
    const int stop = cg_decide(st, rn, st->it);
    if (row0 == 0 && threadIdx.x == 0) {
        st->rho = tot[1];
        st->res = rn / st->bnorm;
        st->stop = stop;
    }
    if (stop) return;
    const S beta = (S)(tot[1] / st->rhoPrev);
    double pq = 0.0;
    for (int r = r0; r < row1; r += kCgBlock) {
        if (r != r0) {
            gather(r);
            if (r + kCgBlock < row1) rp.fetch(soff, scol, sval, r + kCgBlock);
        }
        const S own = zr + beta * pr;
        S sum = r < N ? own : S(0);
#pragma unroll
        for (int k = 0; k < kSpE; ++k)
            if (k < rp.w) sum += rp.v[k] * (zc[k] + beta * pc[k]);
        for (int k = kSpE; k < rp.w; ++k) {  // rows wider than kSpE entries
            const int c = scol[rp.b + 64 * (int64_t)k];
            sum += sval[rp.b + 64 * (int64_t)k] * (zv[c] + beta * po[c]);
        }
        if (r < N) {
            pn[r] = own;
            qvec[r] = sum;
            pq += (double)own * (double)sum;
        }
    }
    double vv[1] = {pq};
    write_slot<1>(vv, red.pq, slot, sh);
}

template <typename S>
__global__ __launch_bounds__(kCgBlock) void kc_update(const CgGrid<S>* __restrict__ g, CgPartials red,
                                                      int pb) {
    __shared__ double sh[kCgShared];
    const CgSys<S>* Pp;
    int row0, row1, slot;
    if (!cg_locate(g, Pp, row0, row1, slot)) return;
    const CgSys<S>& P = *Pp;
    CgState* st = P.st;
    const int stop = st->stop;
    if (row0 == 0 && threadIdx.x == 0) {
        st->done = stop;
        if (!stop) {
            st->it += 1;
            st->rhoPrev = st->rho;
            st->pbuf = pb;
        }
    }
    if (stop) return;
    const S* __restrict__ pvec = cg_pbuf(P, pb);
    // (the other vectors are addressed through P: reading their addresses
    // once, as the other kernels do, costs this kernel 6 VGPRs and a wave per
    // SIMD in FP64, 2.5 % of a 128^3 solve)
    // rows are taken kCgU per thread at a time, their loads issued together;
    // the first group's loads go out before the reduction of p.q
    const int r1 = min(row1, P.N);
    const int i0 = row0 + (int)threadIdx.x;
    S dv[kCgU], pv[kCgU], rv[kCgU], qv[kCgU], gv[kCgU];
    auto load = [&](int base) {
#pragma unroll
        for (int u = 0; u < kCgU; ++u) {
            const int i = base + u * kCgBlock;
            if (i < r1) {
                dv[u] = P.d[i];
                pv[u] = pvec[i];
                rv[u] = P.r[i];
                qv[u] = P.q[i];
                gv[u] = P.g[i];
            }
        }
    };
    load(i0);
    double tot[1];
    sum_slots<1>(red.pq, P.tile0, P.ntiles, sh, tot);
    const double alpha = st->rho / tot[0];
    if (row0 == 0 && threadIdx.x == 0) st->alpha = alpha;
    const S a = (S)alpha;
    double rr = 0.0, rz = 0.0;
    for (int base = i0; base < r1; base += kCgU * kCgBlock) {
        if (base != i0) load(base);
#pragma unroll
        for (int u = 0; u < kCgU; ++u) {
            const int i = base + u * kCgBlock;
            if (i < r1) {
                P.d[i] = dv[u] + a * pv[u];
                const S ri = rv[u] - a * qv[u];
                P.r[i] = ri;
                const double ur = (double)ri * (double)gv[u];  // unscaled residual
                rr += ur * ur;
                rz += (double)ri * (double)ri;
            }
        }
    }
    double v[2] = {rr, rz};
    write_slot<2>(v, red.rz, slot, sh);
}

// MODE 0: after an update (stop on the recurrence estimate, the iteration limit,
//         NaN or the residual-drop trigger);
// MODE 1: after a residual replacement, for the replaced systems only (stop on
//         the FP64 residual: 2 converged, 3 non-finite, 1 iteration limit);
// MODE 2: after the start (replacement of the initial guess; p = z already).
template <typename S, int MODE>
__global__ __launch_bounds__(kCgBlock) void kc_direction(const CgGrid<S>* __restrict__ g, CgPartials red,
                                                         int pOld) {
    __shared__ double sh[kCgShared];
    const CgSys<S>* Pp;
    int row0, row1, slot;
    if (!cg_locate(g, Pp, row0, row1, slot)) return;
    const CgSys<S>& P = *Pp;
    CgState* st = P.st;
    const int done = st->done;
    if ((MODE == 0 && done) || (MODE == 1 && done != 1)) return;
    // the old direction: the chunk's last buffer (mode 0) or that of the
    // system's last update (mode 1); the new one goes to the first buffer
    const S* po = cg_pbuf(P, MODE == 1 ? st->pbuf : pOld);  // may be the new buffer
    // the vector addresses are read from the system description once: read
    // through P inside the row loop they would be re-read after every store
    // (a store might change the description), each re-read waiting for the
    // loads before it
    const S* __restrict__ rvec = P.r;
    S* pn = P.p;
    // the first kCgU rows' r (= z) and p are loaded before the reduction
    const int r1 = min(row1, P.N);
    const int i0 = row0 + (int)threadIdx.x;
    S zv[kCgU], pv[kCgU];
    auto load = [&](int base) {
#pragma unroll
        for (int u = 0; u < kCgU; ++u) {
            const int i = base + u * kCgBlock;
            if (i < r1) {
                zv[u] = rvec[i];
                pv[u] = po[i];
            }
        }
    };
    if (MODE != 2) load(i0);
    double tot[2];
    sum_slots<2>(red.rz, P.tile0, P.ntiles, sh, tot);
    const int it = st->it;
    const double rn = sqrt(tot[0]);
    int stop;
    if (MODE == 0) {
        stop = cg_decide(st, rn, it);
    } else {
        const double res = st->res;
        const bool conv = res <= st->tol || res * st->bnorm <= st->atol;
        stop = conv ? 2 : !(res <= 1.7976931348623157e308) ? 3 : it >= st->maxit ? 1 : 0;
    }
    if (row0 == 0 && threadIdx.x == 0) {
        st->rho = tot[1];
        if (MODE == 0)
            st->res = rn / st->bnorm;
        else
            st->rrep = rn;
        st->stop = stop;
    }
    if (MODE == 2 || stop != 0 || it == 0) return;
    const S beta = (S)(tot[1] / st->rhoPrev);
    for (int base = i0; base < r1; base += kCgU * kCgBlock) {
        if (base != i0) load(base);
#pragma unroll
        for (int u = 0; u < kCgU; ++u) {
            const int i = base + u * kCgBlock;
            if (i < r1) pn[i] = zv[u] + beta * pv[u];
        }
    }
}

// x += |G|^-1 d, d = 0: for the systems that stopped (ALL = 0), or for every
// system (ALL = 1: the end of a plain FP64 solve)
template <typename S, int ALL>
__global__ __launch_bounds__(kCgBlock) void kc_accum(const CgGrid<S>* __restrict__ g) {
    const CgSys<S>* Pp;
    int row0, row1, slot;
    if (!cg_locate(g, Pp, row0, row1, slot)) return;
    const CgSys<S>& P = *Pp;
    if (!ALL && P.st->stop != 1) return;
    const int r1 = min(row1, P.N);
    double* __restrict__ x = P.x;  // addresses read once (see kc_direction)
    S* __restrict__ d = P.d;
    const S* __restrict__ gv = P.g;
    for (int i = row0 + threadIdx.x; i < r1; i += kCgBlock) {
        x[i] += (double)d[i] / fabs((double)gv[i]);
        d[i] = S(0);
    }
}

// x_old = x for the systems that have a copy target (after a batched solve)
template <typename S>
__global__ __launch_bounds__(kCgBlock) void kc_copy_old(const CgGrid<S>* __restrict__ g) {
    const CgSys<S>* Pp;
    int row0, row1, slot;
    if (!cg_locate(g, Pp, row0, row1, slot)) return;
    const CgSys<S>& P = *Pp;
    double* xo = P.xold;
    if (!xo) return;
    const double* x = P.x;
    const int r1 = min(row1, P.N);
    for (int i = row0 + threadIdx.x; i < r1; i += kCgBlock) xo[i] = x[i];
}

// r64 = b - A x in FP64 (CSR rows in column order), partials of r.r and b.b
template <typename S, int START>
__global__ __launch_bounds__(kCgBlock) void kc_resid(const CgGrid<S>* __restrict__ g, CgPartials red) {
    __shared__ double sh[kCgShared];
    const CgSys<S>* Pp;
    int row0, row1, slot;
    if (!cg_locate(g, Pp, row0, row1, slot)) return;
    const CgSys<S>& P = *Pp;
    if (!START && P.st->stop != 1) return;
    // a row's first kRE entries have their loads issued together (values and
    // columns, then the gathers of x), as in the SpMV; same summation order
    constexpr int kRE = 8;
    const int32_t* __restrict__ ro = P.rowOffs;
    const int32_t* __restrict__ ci = P.colIdxs;
    const double* __restrict__ va = P.vals;
    const double* __restrict__ xv = P.x;
    const double* __restrict__ rhs = P.rhs;
    double* __restrict__ r64 = P.r64;
    double rr = 0.0, bb = 0.0;
    const int r1 = min(row1, P.N);
    for (int i = row0 + threadIdx.x; i < r1; i += kCgBlock) {
        const int k0 = ro[i], k1 = ro[i + 1];
        const double bi = rhs[i];
        double v[kRE], xg[kRE];
        int c[kRE];
#pragma unroll
        for (int k = 0; k < kRE; ++k)
            if (k0 + k < k1) {
                v[k] = va[k0 + k];
                c[k] = ci[k0 + k];
            }
#pragma unroll
        for (int k = 0; k < kRE; ++k)
            if (k0 + k < k1) xg[k] = xv[c[k]];
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < kRE; ++k)
            if (k0 + k < k1) s += v[k] * xg[k];
        for (int k = k0 + kRE; k < k1; ++k) s += va[k] * xv[ci[k]];
        const double ri = bi - s;
        r64[i] = ri;
        rr += ri * ri;
        bb += bi * bi;
    }
    double v[2] = {rr, bb};
    write_slot<2>(v, red.rb, slot, sh);
}

// relative FP64 residual from the resid partials; r = (S) (r64 / g) (z = r),
// START: p = r and ||b||; partials of ||G r||^2 and r.r
template <typename S, int START>
__global__ __launch_bounds__(kCgBlock) void kc_replace(const CgGrid<S>* __restrict__ g, CgPartials red) {
    __shared__ double sh[kCgShared];
    const CgSys<S>* Pp;
    int row0, row1, slot;
    if (!cg_locate(g, Pp, row0, row1, slot)) return;
    const CgSys<S>& P = *Pp;
    CgState* st = P.st;
    const int stop = st->stop;
    if (!START && stop != 1) return;
    double tot[2];
    sum_slots<2>(red.rb, P.tile0, P.ntiles, sh, tot);
    double bn;
    if (START) {
        bn = sqrt(tot[1]);
        if (!(bn > 0.0)) bn = 1.0;
    } else {
        bn = st->bnorm;
    }
    const double res = sqrt(tot[0]) / bn;
    if (row0 == 0 && threadIdx.x == 0) {
        st->res = res;
        st->done = stop;  // tells the mode-1 direction kernel that this system was replaced
        if (START) {
            st->bnorm = bn;
            st->res0 = res;
        } else {
            st->nrep += 1;
        }
    }
    double rr = 0.0, rz = 0.0;
    const int r1 = min(row1, P.N);
    const S* __restrict__ gv = P.g;  // addresses read once (see kc_direction)
    const double* __restrict__ r64 = P.r64;
    S* __restrict__ rv = P.r;
    S* __restrict__ pv = P.p;
    for (int i = row0 + threadIdx.x; i < r1; i += kCgBlock) {
        const double gi = (double)gv[i];
        const S ri = (S)(r64[i] / gi);
        rv[i] = ri;
        if (START) pv[i] = ri;
        const double ur = (double)ri * gi;
        rr += ur * ur;
        rz += (double)ri * (double)ri;
    }
    double v[2] = {rr, rz};
    write_slot<2>(v, red.rz, slot, sh);
}

// One block: ||r64|| / ||b|| of system 0 from the resid partials into st.res
template <typename S>
__global__ __launch_bounds__(kCgBlock) void kc_resid_norm(const CgSys<S>* __restrict__ P,
                                                          CgPartials red, CgState* out) {
    __shared__ double sh[kCgShared];
    double tot[2];
    sum_slots<2>(red.rb, P->tile0, P->ntiles, sh, tot);
    if (threadIdx.x == 0) {
        double bn = sqrt(tot[1]);
        if (!(bn > 0.0)) bn = 1.0;
        out->bnorm = bn;
        out->res = sqrt(tot[0]) / bn;
    }
}

// Start parameters of one system (batched start)
struct CgStartParam {
    double tol, delta, atol;
    int32_t maxit, pad;
};

// Start, one workgroup per tile of the (re)starting systems: g_i (the signed
// square root of the diagonal, NaN for a zero, non-finite or opposite-sign
// entry), d = 0, and (STATE, first tile) the system's CG state from prm.
// Tile rows are multiples of 64, so a tile owns whole slices.
template <typename S, int STATE>
__global__ __launch_bounds__(kCgBlock) void kc_start_tile(const CgGrid<S>* __restrict__ g,
                                                          const CgStartParam* __restrict__ prm) {
    const CgSys<S>* Pp;
    int row0, row1, slot;
    if (!cg_locate(g, Pp, row0, row1, slot)) return;
    const CgSys<S>& P = *Pp;
    const double d0 = P.vals[P.rowOffs[0] + P.diag[0]];
    const int r1 = min(row1, P.N);
    for (int i = row0 + threadIdx.x; i < r1; i += kCgBlock) {
        const double di = P.vals[P.rowOffs[i] + P.diag[i]];
        const bool ok = di != 0.0 && (di > 0.0) == (d0 > 0.0) && fabs(di) <= 1.7976931348623157e308;
        P.g[i] = ok ? (S)copysign(sqrt(fabs(di)), di) : (S)NAN;
        P.d[i] = S(0);
    }
    if (STATE && row0 == 0 && threadIdx.x == 0) {
        const CgStartParam q = prm[Pp - g->probs];
        CgState s{};
        s.tol = q.tol;
        s.atol = q.atol;
        s.maxit = q.maxit;
        s.delta = q.delta;
        *P.st = s;
    }
}

// Scaled sliced-ELL values a_ij / (g_i |g_j|) in S of the tile's rows (after
// kc_start_tile has written g); one thread per row, so the stores of an entry
// index are coalesced over a slice.
template <typename S>
__global__ __launch_bounds__(kCgBlock) void kc_scale_tile(const CgGrid<S>* __restrict__ g) {
    const CgSys<S>* Pp;
    int row0, row1, slot;
    if (!cg_locate(g, Pp, row0, row1, slot)) return;
    const CgSys<S>& P = *Pp;
    S* sval = const_cast<S*>(P.sval);
    for (int i = row0 + threadIdx.x; i < row1; i += kCgBlock) {
        const int sl = i >> 6;
        const int64_t b = P.sellOff[sl] + (i & 63);
        const int w = (int)((P.sellOff[sl + 1] - P.sellOff[sl]) >> 6);
        const double gi = i < P.N ? (double)P.g[i] : 1.0;
        for (int k = 0; k < w; ++k) {
            const int64_t pos = b + 64 * (int64_t)k;
            const int32_t j = P.sellSrc[pos];
            sval[pos] = j >= 0 ? (S)(P.vals[j] / (gi * fabs((double)P.g[P.sellCol[pos]]))) : S(0);
        }
    }
}

template <int = 0>
__global__ void kc_init_state(CgState* st, int n, double tol, int maxit, double delta,
                              double atol = 0.0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        CgState s{};
        s.tol = tol;
        s.atol = atol;
        s.maxit = maxit;
        s.delta = delta;
        st[i] = s;
    }
}

// Whether a chunk fuses each direction into the next SpMV (two kernels per
// iteration instead of three).  The fused kernel gathers r and the old p for
// every column instead of p alone; measured on MI355X with the scaled
// iteration, that second gather costs more than the launch it saves once a
// launch covers 2^20 rows or more, in both precisions (128^3 Poisson: FP32
// 39.2 -> 37.1 ms and FP64 unfused as before; the served stream +7 %, the
// 512-problem batch +2 %), while small systems, whose iterations are launch-
// and reduction-bound, keep the fused form (32^3 heat step unchanged).
template <typename S>
inline bool cg_fuse_direction(int64_t rows) {
    return rows < (int64_t(1) << 20);
}

// Grid for `tiles` logical blocks: a multiple of 8 (one share per XCD)
inline int cg_grid(int tiles) { return ((tiles + 7) / 8) * 8; }

// Rows per tile of an N-row system: at least `minRows`, and at most kCgMaxTiles
// tiles per system (bounds the partial sums every block re-reads)
constexpr int kCgMaxTiles = 1024;
inline int cg_tile_rows(int N, int minRows) {
    const int64_t need = ((int64_t)N + kCgMaxTiles - 1) / kCgMaxTiles;
    const int r = (int)(((need + 63) / 64) * 64);
    return r > minRows ? r : minRows;
}

}  // namespace
}  // namespace fvhip
