// batch.hip -- many independent implicit solves in one launch sequence.
//
// A batch holds assembled linear systems (fv_linsys, any sizes) and their
// solution vectors.  Their rows are cut into tiles of a fixed number of rows
// (a multiple of 64, never straddling two systems); every kernel of the CG
// iteration runs over the tile list of the whole batch, one workgroup per tile,
// so small and large systems share the machine and one launch serves all of
// them.  Dot products are reduced per system: each tile writes its partials to
// its own slot and the last tile of a system (integer ticket in that system's
// state) sums the system's slots in tile order -- so a system's results depend
// only on its own tiling, bit for bit the same alone or in any mix.
//
// Precision: the CG runs in S (float: mixed, double: FP64) on a sliced-ELL copy
// of the matrix; every chunk of iterations ends with an FP64 residual
// replacement (x += d, r = b - A x in FP64, convergence test in FP64, CG
// continues from the replaced residual without restarting its direction), so
// each system stops on its true FP64 relative residual.  A chunk is captured
// once as a hipGraph and replayed until every system has stopped.
#include "cgstate.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace fvhip;

namespace {

constexpr int kB = kCgBlock;

template <typename S>
struct BProb {
    int32_t N, tile0, ntiles, pad;
    const int64_t* sellOff;
    const int32_t* sellCol;
    const S* sval;
    const int32_t* rowOffs;
    const int32_t* colIdxs;
    const double* vals;
    const double* rhs;
    double* x;
    double* xold;  // optional copy target after convergence (Euler T_old), may be null
    S *d, *r, *z, *p, *q, *dinv;
    double* r64;
    CgState* st;
};

struct BTile {
    int32_t prob, row0, row1, slot;
};

template <typename S>
__global__ __launch_bounds__(kB) void kb_spmv(const BTile* __restrict__ tiles, int nTiles,
                                              const BProb<S>* __restrict__ probs,
                                              double* __restrict__ partials) {
    __shared__ double sh[kB / 64];
    if ((int)blockIdx.x >= nTiles) return;
    const BTile t = tiles[blockIdx.x];
    const BProb<S>& P = probs[t.prob];
    CgState* st = P.st;
    if (st->done) return;
    double pq = 0.0;
    for (int r = t.row0 + threadIdx.x; r < t.row1; r += kB) {
        const int sl = r >> 6, l = r & 63;
        const int64_t b = P.sellOff[sl];
        const int w = (int)((P.sellOff[sl + 1] - b) >> 6);
        S s = 0;
        for (int k = 0; k < w; ++k) {
            const int64_t pos = b + 64 * (int64_t)k + l;
            s += P.sval[pos] * P.p[P.sellCol[pos]];
        }
        if (r < P.N) {
            P.q[r] = s;
            pq += (double)P.p[r] * (double)s;
        }
    }
    double v[1] = {pq}, tot[1];
    if (publish_slots<1>(v, partials, st, sh, P.tile0, P.ntiles, t.slot, tot) && threadIdx.x == 0)
        st->alpha = st->rz[st->it & 1] / tot[0];
}

template <typename S>
__global__ __launch_bounds__(kB) void kb_update(const BTile* __restrict__ tiles, int nTiles,
                                                const BProb<S>* __restrict__ probs,
                                                double* __restrict__ partials) {
    __shared__ double sh[kB / 64];
    if ((int)blockIdx.x >= nTiles) return;
    const BTile t = tiles[blockIdx.x];
    const BProb<S>& P = probs[t.prob];
    CgState* st = P.st;
    if (st->done) return;
    const S a = (S)st->alpha;
    double rr = 0.0, rz = 0.0;
    const int r1 = min(t.row1, P.N);
    for (int i = t.row0 + threadIdx.x; i < r1; i += kB) {
        P.d[i] += a * P.p[i];
        const S ri = P.r[i] - a * P.q[i];
        P.r[i] = ri;
        const S zi = P.dinv[i] * ri;
        P.z[i] = zi;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
    }
    double v[2] = {rr, rz}, tot[2];
    if (publish_slots<2>(v, partials, st, sh, P.tile0, P.ntiles, t.slot, tot) && threadIdx.x == 0) {
        const int it = st->it + 1;
        st->it = it;
        st->rz[it & 1] = tot[1];
        const double rn = sqrt(tot[0]);
        const double est = rn / st->bnorm;
        if (est <= st->tol || it >= st->maxit || !(est == est) ||
            (st->delta > 0.0 && rn < st->delta * st->rrep))
            st->done = 1;
    }
}

template <typename S>
__global__ __launch_bounds__(kB) void kb_direction(const BTile* __restrict__ tiles, int nTiles,
                                                   const BProb<S>* __restrict__ probs) {
    if ((int)blockIdx.x >= nTiles) return;
    const BTile t = tiles[blockIdx.x];
    const BProb<S>& P = probs[t.prob];
    const CgState* st = P.st;
    if (st->done) return;
    const int it = st->it;
    const S b = (S)(st->rz[it & 1] / st->rz[(it - 1) & 1]);
    const int r1 = min(t.row1, P.N);
    for (int i = t.row0 + threadIdx.x; i < r1; i += kB) P.p[i] = P.z[i] + b * P.p[i];
}

// x += d, d = 0 (FP64 accumulation of the chunk's correction)
template <typename S>
__global__ __launch_bounds__(kB) void kb_accum(const BTile* __restrict__ tiles, int nTiles,
                                               const BProb<S>* __restrict__ probs) {
    if ((int)blockIdx.x >= nTiles) return;
    const BTile t = tiles[blockIdx.x];
    const BProb<S>& P = probs[t.prob];
    if (P.st->done == 2) return;
    const int r1 = min(t.row1, P.N);
    for (int i = t.row0 + threadIdx.x; i < r1; i += kB) {
        P.x[i] += (double)P.d[i];
        P.d[i] = S(0);
    }
}

// r64 = b - A x in FP64 (CSR rows in column order); start: also ||b||
template <typename S>
__global__ __launch_bounds__(kB) void kb_resid(const BTile* __restrict__ tiles, int nTiles,
                                               const BProb<S>* __restrict__ probs,
                                               double* __restrict__ partials, int start) {
    __shared__ double sh[kB / 64];
    if ((int)blockIdx.x >= nTiles) return;
    const BTile t = tiles[blockIdx.x];
    const BProb<S>& P = probs[t.prob];
    CgState* st = P.st;
    if (!start && st->done == 2) return;
    double rr = 0.0, bb = 0.0;
    const int r1 = min(t.row1, P.N);
    for (int i = t.row0 + threadIdx.x; i < r1; i += kB) {
        double s = 0.0;
        const int e = P.rowOffs[i + 1];
        for (int k = P.rowOffs[i]; k < e; ++k) s += P.vals[k] * P.x[P.colIdxs[k]];
        const double ri = P.rhs[i] - s;
        P.r64[i] = ri;
        rr += ri * ri;
        bb += P.rhs[i] * P.rhs[i];
    }
    double v[2] = {rr, bb}, tot[2];
    if (publish_slots<2>(v, partials, st, sh, P.tile0, P.ntiles, t.slot, tot) && threadIdx.x == 0) {
        if (start) {
            const double bn = sqrt(tot[1]);
            st->bnorm = bn > 0.0 ? bn : 1.0;
        }
        st->res = sqrt(tot[0]) / st->bnorm;
        if (start) st->res0 = st->res;
    }
}

// r = (S) r64; z = D^-1 r; start: p = z; sets rz[it&1], rrep, done (2: converged)
template <typename S>
__global__ __launch_bounds__(kB) void kb_replace(const BTile* __restrict__ tiles, int nTiles,
                                                 const BProb<S>* __restrict__ probs,
                                                 double* __restrict__ partials, int start) {
    __shared__ double sh[kB / 64];
    if ((int)blockIdx.x >= nTiles) return;
    const BTile t = tiles[blockIdx.x];
    const BProb<S>& P = probs[t.prob];
    CgState* st = P.st;
    if (!start && st->done == 2) return;
    double rr = 0.0, rz = 0.0;
    const int r1 = min(t.row1, P.N);
    for (int i = t.row0 + threadIdx.x; i < r1; i += kB) {
        const S ri = (S)P.r64[i];
        const S zi = P.dinv[i] * ri;
        P.r[i] = ri;
        P.z[i] = zi;
        if (start) P.p[i] = zi;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
    }
    double v[2] = {rr, rz}, tot[2];
    if (publish_slots<2>(v, partials, st, sh, P.tile0, P.ntiles, t.slot, tot) && threadIdx.x == 0) {
        st->rrep = sqrt(tot[0]);
        st->rz[st->it & 1] = tot[1];
        const bool conv = st->res <= st->tol;
        st->done = conv ? 2 : 0;
        if (!conv && st->it >= st->maxit) st->done = 1;
    }
}

// after convergence: xold = x (Euler T_old <- T), once
template <typename S>
__global__ __launch_bounds__(kB) void kb_store_old(const BTile* __restrict__ tiles, int nTiles,
                                                   const BProb<S>* __restrict__ probs) {
    if ((int)blockIdx.x >= nTiles) return;
    const BTile t = tiles[blockIdx.x];
    const BProb<S>& P = probs[t.prob];
    if (!P.xold) return;
    const int r1 = min(t.row1, P.N);
    for (int i = t.row0 + threadIdx.x; i < r1; i += kB) P.xold[i] = P.x[i];
}

template <typename S>
__global__ void kb_sell_values(int64_t n, const int32_t* __restrict__ src,
                               const double* __restrict__ values, S* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t j = src[i];
        out[i] = j >= 0 ? (S)values[j] : (S)0;
    }
}

template <typename S>
__global__ void kb_dinv(int N, const int32_t* __restrict__ rowOffs,
                        const int32_t* __restrict__ diag, const double* __restrict__ values,
                        S* __restrict__ dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) dinv[i] = (S)(1.0 / values[rowOffs[i] + diag[i]]);
}

__global__ void kb_init_states(CgState* st, int n, double tol, int maxit, double delta) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        CgState s{};
        s.tol = tol;
        s.maxit = maxit;
        s.delta = delta;
        st[i] = s;
    }
}

__global__ void kb_zero(int64_t n, uint8_t* __restrict__ p) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

}  // namespace

// ---- host side ----------------------------------------------------------------

struct BMember {
    fv_linsys* ls = nullptr;
    double* x = nullptr;
    double* xold = nullptr;
    int id = -1;
    int N = 0;
    DBuf<uint8_t> vec;   // d, r, z, p, q, dinv (S) + r64 (double)
    DBuf<uint8_t> sval;  // sliced-ELL values in S
    fv_solver_stats stats{};
    bool finished = false;
};

struct fv_batch {
    fv_ctx* ctx = nullptr;
    int prec = FV_SOLVE_F32_IR;  // inner precision of the whole batch
    int tileRows = 1024;         // rows per tile (multiple of 64)
    int chunk = 16;              // CG iterations per replacement
    int nextId = 0;
    std::vector<BMember*> members;
    // device tables, rebuilt when membership changes
    DBuf<uint8_t> probs;
    DBuf<BTile> tiles;
    DBuf<CgState> states;
    DBuf<double> partials;
    int nTiles = 0;
    bool dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    int graphTiles = -1, graphChunk = -1;
    ~fv_batch() {
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto* m : members) delete m;
    }
};

namespace {

size_t elem(int prec) { return prec == FV_SOLVE_F32_IR ? sizeof(float) : sizeof(double); }

template <typename S>
void build_tables(fv_batch* b) {
    fv_ctx* ctx = b->ctx;
    const int P = (int)b->members.size();
    std::vector<BProb<S>> hp(P);
    std::vector<BTile> ht;
    b->states.alloc(ctx, std::max(P, 1));
    for (int i = 0; i < P; ++i) {
        BMember* m = b->members[i];
        const fv_sparsity* sp = m->ls->sp;
        const int N = m->N;
        BProb<S>& p = hp[i];
        p.N = N;
        p.tile0 = (int)ht.size();
        for (int r0 = 0; r0 < N; r0 += b->tileRows) {
            BTile t;
            t.prob = i;
            t.row0 = r0;
            t.row1 = std::min(r0 + b->tileRows, ((N + 63) / 64) * 64);
            t.slot = (int)ht.size();
            ht.push_back(t);
        }
        p.ntiles = (int)ht.size() - p.tile0;
        p.sellOff = sp->sellOff.p;
        p.sellCol = sp->sellCol.p;
        p.sval = (const S*)m->sval.p;
        p.rowOffs = sp->rowOffs.p;
        p.colIdxs = sp->colIdxs.p;
        p.vals = m->ls->values.p;
        p.rhs = m->ls->rhs.p;
        p.x = m->x;
        p.xold = m->xold;
        S* v = (S*)m->vec.p;
        p.d = v;
        p.r = v + N;
        p.z = v + 2 * (int64_t)N;
        p.p = v + 3 * (int64_t)N;
        p.q = v + 4 * (int64_t)N;
        p.dinv = v + 5 * (int64_t)N;
        p.r64 = (double*)(m->vec.p + 6 * (int64_t)N * sizeof(S));
        p.st = b->states.p + i;
    }
    b->nTiles = (int)ht.size();
    b->probs.alloc(ctx, std::max<size_t>(1, P * sizeof(BProb<S>)));
    if (P) FV_HIP(hipMemcpyAsync(b->probs.p, hp.data(), P * sizeof(BProb<S>), hipMemcpyHostToDevice, ctx->stream));
    b->tiles.alloc(ctx, std::max<size_t>(1, ht.size()));
    if (!ht.empty()) FV_HIP(hipMemcpyAsync(b->tiles.p, ht.data(), ht.size() * sizeof(BTile), hipMemcpyHostToDevice, ctx->stream));
    b->partials.alloc(ctx, 2 * std::max<size_t>(1, ht.size()));
    FV_HIP(hipStreamSynchronize(ctx->stream));
    b->dirty = false;
    if (b->graphExec) {
        (void)hipGraphExecDestroy(b->graphExec);
        b->graphExec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
}

template <typename S>
void enqueue_chunk(fv_batch* b, int chunk) {
    hipStream_t s = b->ctx->stream;
    const auto* probs = (const BProb<S>*)b->probs.p;
    const int nt = b->nTiles;
    for (int i = 0; i < chunk; ++i) {
        kb_spmv<S><<<nt, kB, 0, s>>>(b->tiles.p, nt, probs, b->partials.p);
        kb_update<S><<<nt, kB, 0, s>>>(b->tiles.p, nt, probs, b->partials.p);
        if (i + 1 < chunk) kb_direction<S><<<nt, kB, 0, s>>>(b->tiles.p, nt, probs);
    }
    kb_accum<S><<<nt, kB, 0, s>>>(b->tiles.p, nt, probs);
    kb_resid<S><<<nt, kB, 0, s>>>(b->tiles.p, nt, probs, b->partials.p, 0);
    kb_replace<S><<<nt, kB, 0, s>>>(b->tiles.p, nt, probs, b->partials.p, 0);
    kb_direction<S><<<nt, kB, 0, s>>>(b->tiles.p, nt, probs);
}

// Capture one chunk as a graph (kernels only) and replay it.
template <typename S>
void run_chunk(fv_batch* b, bool use_graph) {
    hipStream_t s = b->ctx->stream;
    if (!use_graph) {
        enqueue_chunk<S>(b, b->chunk);
        FV_LAUNCH_CHECK();
        return;
    }
    if (!b->graphExec || b->graphTiles != b->nTiles || b->graphChunk != b->chunk) {
        if (b->graphExec) (void)hipGraphExecDestroy(b->graphExec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->graphExec = nullptr;
        b->graph = nullptr;
        FV_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue_chunk<S>(b, b->chunk);
        FV_HIP(hipStreamEndCapture(s, &b->graph));
        FV_HIP(hipGraphInstantiate(&b->graphExec, b->graph, nullptr, nullptr, 0));
        b->graphTiles = b->nTiles;
        b->graphChunk = b->chunk;
    }
    FV_HIP(hipGraphLaunch(b->graphExec, s));
}

template <typename S>
void prepare_members(fv_batch* b) {
    hipStream_t s = b->ctx->stream;
    for (BMember* m : b->members) {
        const fv_sparsity* sp = m->ls->sp;
        kb_sell_values<S><<<grid_for(sp->sellTotal, kB, 8192), kB, 0, s>>>(
            sp->sellTotal, sp->sellSrc.p, m->ls->values.p, (S*)m->sval.p);
        kb_dinv<S><<<grid_for(m->N, kB), kB, 0, s>>>(m->N, sp->rowOffs.p, sp->diagOffset.p,
                                                     m->ls->values.p,
                                                     (S*)m->vec.p + 5 * (int64_t)m->N);
        kb_zero<<<grid_for(m->N * (int64_t)sizeof(S), kB, 4096), kB, 0, s>>>(
            m->N * (int64_t)sizeof(S), m->vec.p);  // d = 0
    }
    FV_LAUNCH_CHECK();
}

template <typename S>
void solve_all(fv_batch* b, double tol, int maxit, double delta, bool use_graph) {
    fv_ctx* ctx = b->ctx;
    hipStream_t s = ctx->stream;
    if (b->dirty) build_tables<S>(b);
    const int P = (int)b->members.size();
    if (!P) return;
    prepare_members<S>(b);
    const auto* probs = (const BProb<S>*)b->probs.p;
    const int nt = b->nTiles;
    kb_init_states<<<grid_for(P, 64), 64, 0, s>>>(b->states.p, P, tol, maxit, delta);
    kb_resid<S><<<nt, kB, 0, s>>>(b->tiles.p, nt, probs, b->partials.p, 1);
    kb_replace<S><<<nt, kB, 0, s>>>(b->tiles.p, nt, probs, b->partials.p, 1);
    FV_LAUNCH_CHECK();
    std::vector<CgState> hs(P);
    while (true) {
        FV_HIP(hipMemcpyAsync(hs.data(), b->states.p, P * sizeof(CgState), hipMemcpyDeviceToHost, s));
        FV_HIP(hipStreamSynchronize(s));
        bool running = false;
        for (const CgState& st : hs)
            if (st.done != 2 && st.it < st.maxit) running = true;
        if (!running) break;
        run_chunk<S>(b, use_graph);
    }
    for (int i = 0; i < P; ++i) {
        BMember* m = b->members[i];
        m->stats.init_res = hs[i].res0;
        m->stats.final_res = hs[i].res;
        m->stats.true_res = hs[i].res;
        m->stats.iters = hs[i].it;
        m->stats.outer_iters = 0;
        m->finished = true;
    }
    kb_store_old<S><<<nt, kB, 0, s>>>(b->tiles.p, nt, probs);
    FV_LAUNCH_CHECK();
}

}  // namespace

extern "C" {

int fv_batch_create(fv_ctx* ctx, int mode, int tile_rows, int chunk, fv_batch** out) {
    fv_batch* b = nullptr;
    int rc = guarded([&] {
        FV_REQUIRE(ctx && out, "NULL argument");
        FV_REQUIRE(mode == FV_SOLVE_F64 || mode == FV_SOLVE_F32_IR, "unknown mode");
        FV_REQUIRE(tile_rows >= 64 && tile_rows % 64 == 0 && tile_rows <= (1 << 20),
                   "tile_rows must be a positive multiple of 64");
        FV_REQUIRE(chunk >= 1 && chunk <= 4096, "chunk must be in [1, 4096]");
        b = new fv_batch();
        b->ctx = ctx;
        b->prec = mode;
        b->tileRows = tile_rows;
        b->chunk = chunk;
        *out = b;
    });
    if (rc != FV_OK) delete b;
    return rc;
}

int fv_batch_destroy(fv_batch* b) {
    return guarded([&] { delete b; });
}

int fv_batch_add(fv_batch* b, fv_linsys* ls, void* d_x, void* d_xold, int* id) {
    return guarded([&] {
        FV_REQUIRE(b && ls && ls->sp && d_x && id, "NULL argument");
        FV_REQUIRE(ls->sp->mesh->ctx == b->ctx, "linear system belongs to another context");
        auto* m = new BMember();
        try {
            m->ls = ls;
            m->x = (double*)d_x;
            m->xold = (double*)d_xold;
            m->N = ls->sp->nRows;
            const size_t e = elem(b->prec);
            m->vec.alloc(b->ctx, 6 * (size_t)m->N * e + (size_t)m->N * sizeof(double));
            m->sval.alloc(b->ctx, (size_t)ls->sp->sellTotal * e);
        } catch (...) {
            delete m;
            throw;
        }
        m->id = b->nextId++;
        b->members.push_back(m);
        b->dirty = true;
        *id = m->id;
    });
}

int fv_batch_clear(fv_batch* b) {
    return guarded([&] {
        FV_REQUIRE(b, "NULL batch");
        for (auto* m : b->members) delete m;
        b->members.clear();
        b->dirty = true;
    });
}

int fv_batch_size(const fv_batch* b, int* n) {
    return guarded([&] {
        FV_REQUIRE(b && n, "NULL argument");
        *n = (int)b->members.size();
    });
}

int fv_batch_set_shape(fv_batch* b, int tile_rows, int chunk) {
    return guarded([&] {
        FV_REQUIRE(b, "NULL batch");
        FV_REQUIRE(tile_rows >= 64 && tile_rows % 64 == 0, "tile_rows must be a multiple of 64");
        FV_REQUIRE(chunk >= 1 && chunk <= 4096, "chunk must be in [1, 4096]");
        if (tile_rows != b->tileRows) b->dirty = true;
        b->tileRows = tile_rows;
        b->chunk = chunk;
    });
}

int fv_batch_solve(fv_batch* b, const fv_solve_opts* o, int use_graph, fv_solver_stats* stats) {
    return guarded([&] {
        FV_REQUIRE(b && o, "NULL argument");
        FV_REQUIRE(o->rel_tol >= 0.0 && o->max_iter >= 0, "bad solver options");
        const double delta = o->inner_rel_tol > 0.0 ? o->inner_rel_tol : 1e-3;
        if (b->prec == FV_SOLVE_F32_IR)
            solve_all<float>(b, o->rel_tol, o->max_iter, delta, use_graph != 0);
        else
            solve_all<double>(b, o->rel_tol, o->max_iter, 0.0, use_graph != 0);
        int bad = 0;
        for (size_t i = 0; i < b->members.size(); ++i) {
            if (stats) stats[i] = b->members[i]->stats;
            if (!(b->members[i]->stats.final_res <= o->rel_tol)) ++bad;
        }
        if (bad)
            throw Error{FV_ENOCONVERGE, std::to_string(bad) + " system(s) of the batch did not converge"};
    });
}

}  // extern "C"
