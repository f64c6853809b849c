// bicgstab.hip -- Jacobi-preconditioned BiCGStab for non-symmetric systems
// (Ginkgo solver::Bicgstab + preconditioner::Jacobi{max_block_size 1}, the
// mapping of PBiCGStab / smoothSolver in src/compatibility/fvSolution.cpp:22-27
// and the backward-Euler advection solve of test/test_advection.cpp:171-178),
// in FP64.
//
// Right-preconditioned BiCGStab (van der Vorst), M = D = diag(A), run as
// BiCGStab on the column-scaled matrix As = A D^-1 (unit diagonal) with
// u = D x: the iterates and the residual r = b - A x are the same, while the
// preconditioner costs no y = M^-1 p or z = M^-1 s vectors and the SpMVs read
// no diagonal entries (the off-diagonal sliced ELL, a_ij / a_jj; the unit
// diagonal is added):
//   r = b - A x (CSR), rhat = r, p = r at the first iteration, then per iteration i
//   K1  rho_i = rhat.r, stop test on ||r||; beta = (rho_i/rho_{i-1})(alpha/omega)
//       p = r + beta (p - omega v)
//   K2  v = As p                                  partials rhat.v
//   K3  alpha = rho_i / rhat.v, s = r - alpha v
//   K4  t = As s                                  partials t.s, t.t
//   K5  omega = t.s / t.t, x += D^-1 (alpha p + omega s), r = s - omega t
//                                                 partials r.r, rhat.r
// on the row tiles of cgkernels.hpp: every consuming block sums the producer's
// per-tile partials itself in one fixed order (no atomics, no fences), each
// state field has one writer per kernel and is not read by that kernel.
#include "cgkernels.hpp"

#include <cmath>

using namespace fvhip;

namespace {

constexpr int kB = kCgBlock;

struct BiState {
    double rho[2], alpha, omega, bnorm, res, res0, tol, atol;
    int32_t it, stop, maxit, pad;
};

struct BiSys {
    int32_t N, ntiles, rowsPerBlock, pad;
    const int64_t* sellOff;  // off-diagonal sliced ELL
    const int32_t* sellCol;
    const double* sval;      // a_ij / a_jj
    const int32_t* rowOffs;  // the CSR of A (the explicit residuals)
    const int32_t* colIdxs;
    const double* vals;
    const double* rhs;
    const double* dinv;
    double *x, *r, *rhat, *p, *v, *s, *t;
    double *pv, *pt, *pr, *pb;  // partial slots: rhat.v | t.s, t.t | r.r, rhat.r | b.b
    BiState* st;
};

__device__ inline bool bi_locate(const BiSys& P, int& row0, int& row1, int& slot) {
    const int per = gridDim.x >> 3;
    const int lb = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (lb >= P.ntiles) return false;
    row0 = lb * P.rowsPerBlock;
    row1 = min(row0 + P.rowsPerBlock, (P.N + 63) & ~63);
    slot = lb;
    return true;
}

// (As vec)_i for this thread's rows of [row0, row1), handed to emit(i, value)
// for i < N: the unit diagonal, then the off-diagonal entries.  The rows are software-pipelined as in the CG's SpMV: a row's
// first kSpE values and columns are loaded together, its gathers are issued,
// and the next row's slice offsets, values and columns are fetched before the
// gathers are waited for.  Entries are summed in column order.
template <typename F>
__device__ inline void sell_rows(const BiSys& P, int row0, int row1,
                                 const double* __restrict__ vec, F&& emit) {
    constexpr int kSpE = 8;
    const int64_t* __restrict__ soff = P.sellOff;
    const int32_t* __restrict__ scol = P.sellCol;
    const double* __restrict__ sval = P.sval;
    CgRowPipe<double, kSpE> rp;
    const int r0 = row0 + (int)threadIdx.x;
    if (r0 < row1) rp.fetch(soff, scol, sval, r0);
    for (int r = r0; r < row1; r += kB) {
        rp.promote();
        double g[kSpE];
#pragma unroll
        for (int k = 0; k < kSpE; ++k)
            if (k < rp.w) g[k] = vec[rp.col[k]];
        const double own = r < P.N ? vec[r] : 0.0;
        if (r + kB < row1) rp.fetch(soff, scol, sval, r + kB);
        double sum = own;
#pragma unroll
        for (int k = 0; k < kSpE; ++k)
            if (k < rp.w) sum += rp.v[k] * g[k];
        for (int k = kSpE; k < rp.w; ++k) {  // rows wider than kSpE entries
            const int64_t pos = rp.b + 64 * (int64_t)k;
            sum += sval[pos] * vec[scol[pos]];
        }
        if (r < P.N) emit(r, sum);
    }
}

// r = b - A x (CSR rows in column order), rhat = r; partials r.r, rhat.r
// (= r.r) and b.b
__global__ __launch_bounds__(kB) void kbi_start(const BiSys* __restrict__ Pp) {
    __shared__ double sh[kCgShared];
    const BiSys& P = *Pp;
    int row0, row1, slot;
    if (!bi_locate(P, row0, row1, slot)) return;
    double rr = 0.0, bb = 0.0;
    const int r1 = min(row1, P.N);
    for (int i = row0 + threadIdx.x; i < r1; i += kB) {
        const int k0 = P.rowOffs[i], k1 = P.rowOffs[i + 1];
        double ax = 0.0;
        for (int k = k0; k < k1; ++k) ax += P.vals[k] * P.x[P.colIdxs[k]];
        const double bi = P.rhs[i];
        const double ri = bi - ax;
        P.r[i] = ri;
        P.rhat[i] = ri;
        rr += ri * ri;
        bb += bi * bi;
    }
    double v[2] = {rr, rr}, w[1] = {bb};
    write_slot<2>(v, P.pr, slot, sh);
    write_slot<1>(w, P.pb, slot, sh);
}

// one block: ||b|| and the initial residual
__global__ __launch_bounds__(kB) void kbi_start_state(const BiSys* __restrict__ Pp, double tol,
                                                      double atol, int maxit) {
    __shared__ double sh[kCgShared];
    const BiSys& P = *Pp;
    double tb[1], tr[2];
    sum_slots<1>(P.pb, 0, P.ntiles, sh, tb);
    sum_slots<2>(P.pr, 0, P.ntiles, sh, tr);
    if (threadIdx.x == 0) {
        BiState s{};
        double bn = sqrt(tb[0]);
        if (!(bn > 0.0)) bn = 1.0;
        s.bnorm = bn;
        s.res0 = sqrt(tr[0]) / bn;
        s.res = s.res0;
        s.tol = tol;
        s.atol = atol;
        s.maxit = maxit;
        s.rho[1] = 1.0;
        s.alpha = 1.0;
        s.omega = 1.0;
        *P.st = s;
    }
}

__global__ __launch_bounds__(kB) void kbi_dir(const BiSys* __restrict__ Pp) {
    __shared__ double sh[kCgShared];
    const BiSys& P = *Pp;
    int row0, row1, slot;
    if (!bi_locate(P, row0, row1, slot)) return;
    BiState* st = P.st;
    double tr[2];
    sum_slots<2>(P.pr, 0, P.ntiles, sh, tr);
    const int it = st->it;
    const double rn = sqrt(tr[0]), rho = tr[1];
    const double res = rn / st->bnorm;
    // 2 converged, 1 iteration limit, 3 breakdown (rho or omega zero, or NaN)
    int stop = 0;
    if (res <= st->tol || rn <= st->atol)
        stop = 2;
    else if (!(res == res) || rho == 0.0 || (it > 0 && st->omega == 0.0))
        stop = 3;
    else if (it >= st->maxit)
        stop = 1;
    if (row0 == 0 && threadIdx.x == 0) {
        st->rho[it & 1] = rho;
        st->res = res;
        st->stop = stop;
    }
    if (stop) return;
    const int r1 = min(row1, P.N);
    if (it == 0) {
        for (int i = row0 + threadIdx.x; i < r1; i += kB) P.p[i] = P.r[i];
    } else {
        const double beta = (rho / st->rho[(it - 1) & 1]) * (st->alpha / st->omega);
        const double om = st->omega;
        for (int i = row0 + threadIdx.x; i < r1; i += kB)
            P.p[i] = P.r[i] + beta * (P.p[i] - om * P.v[i]);
    }
}

// v = As p; partials rhat.v
__global__ __launch_bounds__(kB) void kbi_spmv_v(const BiSys* __restrict__ Pp) {
    __shared__ double sh[kCgShared];
    const BiSys& P = *Pp;
    int row0, row1, slot;
    if (!bi_locate(P, row0, row1, slot)) return;
    if (P.st->stop) return;
    double rv = 0.0;
    sell_rows(P, row0, row1, P.p, [&](int i, double vi) {
        P.v[i] = vi;
        rv += P.rhat[i] * vi;
    });
    double a[1] = {rv};
    write_slot<1>(a, P.pv, slot, sh);
}

// alpha = rho / rhat.v; s = r - alpha v
__global__ __launch_bounds__(kB) void kbi_s(const BiSys* __restrict__ Pp) {
    __shared__ double sh[kCgShared];
    const BiSys& P = *Pp;
    int row0, row1, slot;
    if (!bi_locate(P, row0, row1, slot)) return;
    BiState* st = P.st;
    if (st->stop) return;
    double tv[1];
    sum_slots<1>(P.pv, 0, P.ntiles, sh, tv);
    const double alpha = st->rho[st->it & 1] / tv[0];
    if (row0 == 0 && threadIdx.x == 0) st->alpha = alpha;
    const int r1 = min(row1, P.N);
    for (int i = row0 + threadIdx.x; i < r1; i += kB) {
        P.s[i] = P.r[i] - alpha * P.v[i];
    }
}

// t = As s; partials t.s, t.t
__global__ __launch_bounds__(kB) void kbi_spmv_t(const BiSys* __restrict__ Pp) {
    __shared__ double sh[kCgShared];
    const BiSys& P = *Pp;
    int row0, row1, slot;
    if (!bi_locate(P, row0, row1, slot)) return;
    if (P.st->stop) return;
    double ts = 0.0, tt = 0.0;
    sell_rows(P, row0, row1, P.s, [&](int i, double ti) {
        P.t[i] = ti;
        ts += ti * P.s[i];
        tt += ti * ti;
    });
    double a[2] = {ts, tt};
    write_slot<2>(a, P.pt, slot, sh);
}

// omega = t.s / t.t; x += D^-1 (alpha p + omega s); r = s - omega t; partials
// r.r, rhat.r
__global__ __launch_bounds__(kB) void kbi_update(const BiSys* __restrict__ Pp) {
    __shared__ double sh[kCgShared];
    const BiSys& P = *Pp;
    int row0, row1, slot;
    if (!bi_locate(P, row0, row1, slot)) return;
    BiState* st = P.st;
    if (st->stop) return;
    double tt[2];
    sum_slots<2>(P.pt, 0, P.ntiles, sh, tt);
    const double omega = tt[1] > 0.0 ? tt[0] / tt[1] : 0.0;
    const double alpha = st->alpha;
    if (row0 == 0 && threadIdx.x == 0) {
        st->omega = omega;
        st->it = st->it + 1;
    }
    double rr = 0.0, hr = 0.0;
    const int r1 = min(row1, P.N);
    for (int i = row0 + threadIdx.x; i < r1; i += kB) {
        const double si = P.s[i];
        P.x[i] += P.dinv[i] * (alpha * P.p[i] + omega * si);
        const double ri = si - omega * P.t[i];
        P.r[i] = ri;
        rr += ri * ri;
        hr += P.rhat[i] * ri;
    }
    double a[2] = {rr, hr};
    write_slot<2>(a, P.pr, slot, sh);
}

__global__ void kbi_dinv(int N, const int32_t* __restrict__ rowOffs, const int32_t* __restrict__ diag,
                         const double* __restrict__ values, double* __restrict__ dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) dinv[i] = 1.0 / values[rowOffs[i] + diag[i]];
}

// Column-scaled off-diagonal sliced-ELL values a_ij / a_jj (padding 0)
__global__ void kbi_scale(int64_t n, const int32_t* __restrict__ src, const int32_t* __restrict__ col,
                          const double* __restrict__ values, const double* __restrict__ dinv,
                          double* __restrict__ out) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t j = src[e];
        out[e] = j >= 0 ? values[j] * dinv[col[e]] : 0.0;
    }
}

// Workspace kept with the linear system between solves
struct BiWork : fv_solver_cache {
    DBuf<double> sval, dinv, r, rhat, p, v, s, t, parts;
    DBuf<BiState> st;
    DBuf<BiSys> sys;
    BiSys h{};
    int G = 0;
    BiWork(fv_ctx* c, const fv_linsys* ls) {
        const fv_sparsity* sp = ls->sp;
        const int N = sp->nRows;
        const int R = cg_tile_rows(N, kB);
        const int tiles = (N + R - 1) / R;
        G = cg_grid(tiles);
        sval.alloc(c, sp->sellTotal);
        for (DBuf<double>* b : {&dinv, &r, &rhat, &p, &v, &s, &t}) b->alloc(c, N);
        parts.alloc(c, 6 * (size_t)tiles);
        st.alloc(c, 1);
        sys.alloc(c, 1);
        h.N = N;
        h.ntiles = tiles;
        h.rowsPerBlock = R;
        h.sellOff = sp->sellOff.p;
        h.sellCol = sp->sellCol.p;
        h.sval = sval.p;
        h.rowOffs = sp->rowOffs.p;
        h.colIdxs = sp->colIdxs.p;
        h.vals = ls->values.p;
        h.rhs = ls->rhs.p;
        h.dinv = dinv.p;
        h.r = r.p;
        h.rhat = rhat.p;
        h.p = p.p;
        h.v = v.p;
        h.s = s.p;
        h.t = t.p;
        h.pv = parts.p;
        h.pt = parts.p + tiles;
        h.pr = parts.p + 3 * (int64_t)tiles;
        h.pb = parts.p + 5 * (int64_t)tiles;
        h.st = st.p;
    }
};

}  // namespace

extern "C" {

int fv_solve_bicgstab(fv_linsys* ls, void* d_x, const fv_solve_opts* o, fv_solver_stats* stats) {
    fv_solver_stats local{};
    int rc = guarded([&] {
        FV_REQUIRE(ls && ls->sp && d_x && o, "NULL argument");
        FV_REQUIRE(o->rel_tol >= 0.0 && o->max_iter >= 0 && o->abs_tol >= 0.0,
                   "bad solver options");
        FV_REQUIRE(o->mode == FV_SOLVE_F64, "BiCGStab runs in FP64 only");
        fv_sparsity* sp = ls->sp;
        fv_ctx* ctx = sp->mesh->ctx;
        hipStream_t s = ctx->stream;
        auto& slot = ls->cg[2];
        if (!slot) slot.reset(new BiWork(ctx, ls));
        BiWork& w = *static_cast<BiWork*>(slot.get());
        w.h.x = (double*)d_x;
        FV_HIP(hipMemcpyAsync(w.sys.p, &w.h, sizeof w.h, hipMemcpyHostToDevice, s));
        kbi_dinv<<<grid_for(sp->nRows, kB), kB, 0, s>>>(sp->nRows, sp->rowOffs.p, sp->diagOffset.p,
                                                        ls->values.p, w.dinv.p);
        kbi_scale<<<grid_for(sp->sellTotal, kB, 8192), kB, 0, s>>>(
            sp->sellTotal, sp->sellSrc.p, sp->sellCol.p, ls->values.p, w.dinv.p, w.sval.p);
        kbi_start<<<w.G, kB, 0, s>>>(w.sys.p);
        kbi_start_state<<<1, kB, 0, s>>>(w.sys.p, o->rel_tol, o->abs_tol, o->max_iter);
        FV_LAUNCH_CHECK();
        const int check = o->check_every > 0 ? o->check_every : 8;
        BiState h{};
        int launched = 0, batch = std::min(2, check);
        while (true) {
            const int n = std::min(batch, std::max(1, o->max_iter - launched + 1));
            batch = std::min(2 * batch, check);
            for (int i = 0; i < n; ++i) {
                kbi_dir<<<w.G, kB, 0, s>>>(w.sys.p);
                kbi_spmv_v<<<w.G, kB, 0, s>>>(w.sys.p);
                kbi_s<<<w.G, kB, 0, s>>>(w.sys.p);
                kbi_spmv_t<<<w.G, kB, 0, s>>>(w.sys.p);
                kbi_update<<<w.G, kB, 0, s>>>(w.sys.p);
            }
            FV_LAUNCH_CHECK();
            launched += n;
            FV_HIP(hipMemcpyAsync(&h, w.st.p, sizeof h, hipMemcpyDeviceToHost, s));
            FV_HIP(hipStreamSynchronize(s));
            if (h.stop) break;
            FV_REQUIRE(launched <= o->max_iter + 1, "BiCGStab iteration bookkeeping");
        }
        local.init_res = h.res0;
        local.final_res = h.res;
        local.iters = h.it;
        local.outer_iters = 1;
        // explicit FP64 residual b - A x of the result
        kbi_start<<<w.G, kB, 0, s>>>(w.sys.p);
        kbi_start_state<<<1, kB, 0, s>>>(w.sys.p, o->rel_tol, o->abs_tol, o->max_iter);
        FV_LAUNCH_CHECK();
        BiState h2{};
        FV_HIP(hipMemcpyAsync(&h2, w.st.p, sizeof h2, hipMemcpyDeviceToHost, s));
        FV_HIP(hipStreamSynchronize(s));
        local.true_res = h2.res0;
        if (h.stop == 3)
            throw Error{FV_ENOCONVERGE, "BiCGStab breakdown at iteration " + std::to_string(h.it)};
        if (h.stop != 2)
            throw Error{FV_ENOCONVERGE, "BiCGStab stopped at relative residual " +
                                            std::to_string(h.res) + " after " +
                                            std::to_string(h.it) + " iterations"};
    });
    if (stats) *stats = local;
    return rc;
}

}  // extern "C"
