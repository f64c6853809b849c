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
// once as a hipGraph (per membership) and replayed until every system has
// stopped.  Members can join and leave between chunks (continuous batching):
// joining members are started by their own tiles, the others carry on.
#include "batch.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
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

fv_batch::~fv_batch() {
    if (graphExec) (void)hipGraphExecDestroy(graphExec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto* m : members) delete m;
}

namespace {

size_t elem(int prec) { return prec == FV_SOLVE_F32_IR ? sizeof(float) : sizeof(double); }

void drop_graph(fv_batch* b) {
    if (b->graphExec) (void)hipGraphExecDestroy(b->graphExec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graphExec = nullptr;
    b->graph = nullptr;
}

template <typename S>
void build_tables(fv_batch* b) {
    fv_ctx* ctx = b->ctx;
    const int P = (int)b->members.size();
    std::vector<BProb<S>> hp(P);
    std::vector<BTile> ht;
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
        p.st = b->states.p + m->slot;
    }
    b->nTiles = (int)ht.size();
    b->probs.alloc(ctx, std::max<size_t>(1, P * sizeof(BProb<S>)));
    if (P)
        FV_HIP(hipMemcpyAsync(b->probs.p, hp.data(), P * sizeof(BProb<S>), hipMemcpyHostToDevice,
                              ctx->stream));
    b->tiles.alloc(ctx, std::max<size_t>(1, ht.size() * sizeof(BTile)));
    if (!ht.empty())
        FV_HIP(hipMemcpyAsync(b->tiles.p, ht.data(), ht.size() * sizeof(BTile),
                              hipMemcpyHostToDevice, ctx->stream));
    b->partials.alloc(ctx, 2 * std::max<size_t>(1, ht.size()));
    FV_HIP(hipStreamSynchronize(ctx->stream));
    b->dirty = false;
    drop_graph(b);
}

template <typename S>
void enqueue_chunk(fv_batch* b) {
    hipStream_t s = b->ctx->stream;
    const auto* probs = (const BProb<S>*)b->probs.p;
    const auto* tiles = (const BTile*)b->tiles.p;
    const int nt = b->nTiles;
    for (int i = 0; i < b->chunk; ++i) {
        kb_spmv<S><<<nt, kB, 0, s>>>(tiles, nt, probs, b->partials.p);
        kb_update<S><<<nt, kB, 0, s>>>(tiles, nt, probs, b->partials.p);
        // the chunk ends after an update: its direction update uses the replaced residual
        if (i + 1 < b->chunk) kb_direction<S><<<nt, kB, 0, s>>>(tiles, nt, probs);
    }
    kb_accum<S><<<nt, kB, 0, s>>>(tiles, nt, probs);
    kb_resid<S><<<nt, kB, 0, s>>>(tiles, nt, probs, b->partials.p, 0);
    kb_replace<S><<<nt, kB, 0, s>>>(tiles, nt, probs, b->partials.p, 0);
    kb_direction<S><<<nt, kB, 0, s>>>(tiles, nt, probs);
}

template <typename S>
void run_chunk(fv_batch* b) {
    hipStream_t s = b->ctx->stream;
    if (!b->useGraph) {
        enqueue_chunk<S>(b);
        FV_LAUNCH_CHECK();
        return;
    }
    if (!b->graphExec) {
        // the chunk holds kernels only; it is captured once per membership
        FV_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue_chunk<S>(b);
        hipError_t e = hipStreamEndCapture(s, &b->graph);
        if (e != hipSuccess) {
            b->graph = nullptr;
            throw Error{FV_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e)};
        }
        FV_HIP(hipGraphInstantiate(&b->graphExec, b->graph, nullptr, nullptr, 0));
    }
    FV_HIP(hipGraphLaunch(b->graphExec, s));
}

template <typename S>
void start_members(fv_batch* b) {
    fv_ctx* ctx = b->ctx;
    hipStream_t s = ctx->stream;
    std::vector<BTile> st;
    // tile list of the members that (re)start, with their global slots
    int slot = 0;
    for (int i = 0; i < (int)b->members.size(); ++i) {
        BMember* m = b->members[i];
        const int nt = (m->N + b->tileRows - 1) / b->tileRows;
        if (m->needStart) {
            for (int k = 0; k < nt; ++k) {
                BTile t;
                t.prob = i;
                t.row0 = k * b->tileRows;
                t.row1 = std::min(t.row0 + b->tileRows, ((m->N + 63) / 64) * 64);
                t.slot = slot + k;
                st.push_back(t);
            }
            const fv_sparsity* sp = m->ls->sp;
            kb_sell_values<S><<<grid_for(sp->sellTotal, kB, 8192), kB, 0, s>>>(
                sp->sellTotal, sp->sellSrc.p, m->ls->values.p, (S*)m->sval.p);
            kb_dinv<S><<<grid_for(m->N, kB), kB, 0, s>>>(m->N, sp->rowOffs.p, sp->diagOffset.p,
                                                         m->ls->values.p,
                                                         (S*)m->vec.p + 5 * (int64_t)m->N);
            kb_zero<<<grid_for(m->N * (int64_t)sizeof(S), kB, 4096), kB, 0, s>>>(
                m->N * (int64_t)sizeof(S), m->vec.p);  // d = 0
            kb_init_states<<<1, 64, 0, s>>>(b->states.p + m->slot, 1, b->tol, b->maxit, b->delta);
            m->needStart = false;
        }
        slot += nt;
    }
    if (st.empty()) return;
    b->startTiles.alloc(ctx, st.size() * sizeof(BTile));
    FV_HIP(hipMemcpyAsync(b->startTiles.p, st.data(), st.size() * sizeof(BTile),
                          hipMemcpyHostToDevice, s));
    const auto* probs = (const BProb<S>*)b->probs.p;
    const auto* tiles = (const BTile*)b->startTiles.p;
    const int nt = (int)st.size();
    kb_resid<S><<<nt, kB, 0, s>>>(tiles, nt, probs, b->partials.p, 1);
    kb_replace<S><<<nt, kB, 0, s>>>(tiles, nt, probs, b->partials.p, 1);
    FV_LAUNCH_CHECK();
    // the upload buffer must outlive the launches
    FV_HIP(hipStreamSynchronize(s));
}

template <typename S>
void step_t(fv_batch* b, std::vector<CgState>& out) {
    if (b->dirty) build_tables<S>(b);
    out.clear();
    if (b->members.empty()) return;
    start_members<S>(b);
    run_chunk<S>(b);
    std::vector<CgState> all(b->slotCap);
    hipStream_t s = b->ctx->stream;
    FV_HIP(hipMemcpyAsync(all.data(), b->states.p, b->slotCap * sizeof(CgState),
                          hipMemcpyDeviceToHost, s));
    FV_HIP(hipStreamSynchronize(s));
    for (BMember* m : b->members) out.push_back(all[m->slot]);
}

int grow_slots(fv_batch* b) {
    if (b->freeSlots.empty()) {
        const int cap = std::max(8, 2 * b->slotCap);
        DBuf<CgState> ns;
        ns.alloc(b->ctx, cap);
        if (b->slotCap)
            FV_HIP(hipMemcpyAsync(ns.p, b->states.p, b->slotCap * sizeof(CgState),
                                  hipMemcpyDeviceToDevice, b->ctx->stream));
        FV_HIP(hipStreamSynchronize(b->ctx->stream));
        b->states = std::move(ns);
        for (int i = cap - 1; i >= b->slotCap; --i) b->freeSlots.push_back(i);
        b->slotCap = cap;
        b->dirty = true;
    }
    const int s = b->freeSlots.back();
    b->freeSlots.pop_back();
    return s;
}

}  // namespace

namespace fvhip {

BMember* batch_add(fv_batch* b, fv_linsys* ls, double* x, double* xold, void* user) {
    FV_REQUIRE(ls && ls->sp && x, "NULL system or vector");
    FV_REQUIRE(ls->sp->mesh->ctx == b->ctx, "linear system belongs to another context");
    auto* m = new BMember();
    try {
        m->ls = ls;
        m->x = x;
        m->xold = xold;
        m->N = ls->sp->nRows;
        m->user = user;
        const size_t e = elem(b->prec);
        m->vec.alloc(b->ctx, 6 * (size_t)m->N * e + (size_t)m->N * sizeof(double));
        m->sval.alloc(b->ctx, (size_t)ls->sp->sellTotal * e);
        m->slot = grow_slots(b);
    } catch (...) {
        delete m;
        throw;
    }
    m->id = b->nextId++;
    m->needStart = true;
    b->members.push_back(m);
    b->dirty = true;
    return m;
}

void batch_remove(fv_batch* b, BMember* m) {
    for (size_t i = 0; i < b->members.size(); ++i)
        if (b->members[i] == m) {
            b->members.erase(b->members.begin() + i);
            b->freeSlots.push_back(m->slot);
            delete m;
            b->dirty = true;
            return;
        }
}

void batch_restart(fv_batch*, BMember* m) { m->needStart = true; }

void batch_step(fv_batch* b, std::vector<CgState>& states) {
    if (b->prec == FV_SOLVE_F32_IR)
        step_t<float>(b, states);
    else
        step_t<double>(b, states);
}

void batch_copy_old(fv_batch* b, BMember* m) {
    if (!m->xold) return;
    FV_HIP(hipMemcpyAsync(m->xold, m->x, m->N * sizeof(double), hipMemcpyDeviceToDevice,
                          b->ctx->stream));
}

}  // namespace fvhip

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
        FV_REQUIRE(b && id, "NULL argument");
        BMember* m = batch_add(b, ls, (double*)d_x, (double*)d_xold, nullptr);
        *id = m->id;
    });
}

int fv_batch_clear(fv_batch* b) {
    return guarded([&] {
        FV_REQUIRE(b, "NULL batch");
        while (!b->members.empty()) batch_remove(b, b->members.back());
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
        FV_REQUIRE(tile_rows >= 64 && tile_rows % 64 == 0 && tile_rows <= (1 << 20),
                   "tile_rows must be a positive multiple of 64");
        FV_REQUIRE(chunk >= 1 && chunk <= 4096, "chunk must be in [1, 4096]");
        if (tile_rows != b->tileRows) b->dirty = true;
        if (chunk != b->chunk) drop_graph(b);
        b->tileRows = tile_rows;
        b->chunk = chunk;
    });
}

}  // extern "C"

// ---- autotuning of the batch shape ---------------------------------------------

namespace {

std::string tune_key(const fv_batch* b) {
    int64_t total = 0, mx = 0;
    for (const BMember* m : b->members) {
        total += m->N;
        mx = std::max<int64_t>(mx, m->N);
    }
    auto lg = [](int64_t v) {
        int l = 0;
        while ((int64_t(1) << (l + 1)) <= v) ++l;
        return l;
    };
    return std::string(b->prec == FV_SOLVE_F32_IR ? "mixed" : "fp64") + "_P" +
           std::to_string(lg((int64_t)b->members.size())) + "_N" + std::to_string(lg(total)) +
           "_M" + std::to_string(lg(mx));
}

// minimal JSON cache: {"<key>": {"tile_rows": T, "chunk": C, "ms": M}, ...}
bool cache_lookup(const char* path, const std::string& key, int* tile, int* chunk) {
    if (!path) return false;
    FILE* f = fopen(path, "r");
    if (!f) return false;
    std::string s;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
    fclose(f);
    const size_t k = s.find("\"" + key + "\"");
    if (k == std::string::npos) return false;
    const size_t t = s.find("\"tile_rows\":", k), c = s.find("\"chunk\":", k);
    if (t == std::string::npos || c == std::string::npos) return false;
    *tile = atoi(s.c_str() + t + 12);
    *chunk = atoi(s.c_str() + c + 8);
    return *tile >= 64 && *tile % 64 == 0 && *chunk >= 1;
}

void cache_store(const char* path, const std::string& key, int tile, int chunk, double ms) {
    if (!path) return;
    std::string s;
    if (FILE* f = fopen(path, "r")) {
        char buf[4096];
        size_t n;
        while ((n = fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
        fclose(f);
    }
    const size_t e = s.rfind('}');
    char entry[256];
    snprintf(entry, sizeof entry, "\"%s\": {\"tile_rows\": %d, \"chunk\": %d, \"ms\": %.4f}",
             key.c_str(), tile, chunk, ms);
    if (e == std::string::npos || s.find('{') == std::string::npos) {
        s = std::string("{\n  ") + entry + "\n}\n";
    } else {
        const bool empty = s.find(':') == std::string::npos;
        s = s.substr(0, e);
        while (!s.empty() && (s.back() == '\n' || s.back() == ' ')) s.pop_back();
        s += std::string(empty ? "\n  " : ",\n  ") + entry + "\n}\n";
    }
    if (FILE* f = fopen(path, "w")) {
        fwrite(s.data(), 1, s.size(), f);
        fclose(f);
    }
}

}  // namespace

extern "C" {

int fv_batch_autotune(fv_batch* b, const fv_solve_opts* o, const char* cache_path, int* tile_rows,
                      int* chunk) {
    return guarded([&] {
        FV_REQUIRE(b && o && tile_rows && chunk, "NULL argument");
        FV_REQUIRE(!b->members.empty(), "empty batch");
        const std::string key = tune_key(b);
        int t = 0, c = 0;
        if (cache_lookup(cache_path, key, &t, &c)) {
            FV_HIP(hipStreamSynchronize(b->ctx->stream));
            *tile_rows = t;
            *chunk = c;
            int rc = fv_batch_set_shape(b, t, c);
            if (rc != FV_OK) throw Error{rc, fv_last_error()};
            return;
        }
        fv_ctx* ctx = b->ctx;
        hipStream_t s = ctx->stream;
        // every trial starts from the same x (and x_old)
        std::vector<DBuf<double>> save(2 * b->members.size());
        for (size_t i = 0; i < b->members.size(); ++i) {
            BMember* m = b->members[i];
            save[2 * i].alloc(ctx, m->N);
            FV_HIP(hipMemcpyAsync(save[2 * i].p, m->x, m->N * sizeof(double), hipMemcpyDeviceToDevice, s));
            if (m->xold) {
                save[2 * i + 1].alloc(ctx, m->N);
                FV_HIP(hipMemcpyAsync(save[2 * i + 1].p, m->xold, m->N * sizeof(double),
                                      hipMemcpyDeviceToDevice, s));
            }
        }
        auto restore = [&] {
            for (size_t i = 0; i < b->members.size(); ++i) {
                BMember* m = b->members[i];
                FV_HIP(hipMemcpyAsync(m->x, save[2 * i].p, m->N * sizeof(double), hipMemcpyDeviceToDevice, s));
                if (m->xold)
                    FV_HIP(hipMemcpyAsync(m->xold, save[2 * i + 1].p, m->N * sizeof(double),
                                          hipMemcpyDeviceToDevice, s));
            }
        };
        hipEvent_t e0, e1;
        FV_HIP(hipEventCreate(&e0));
        FV_HIP(hipEventCreate(&e1));
        double best = 1e300;
        int bt = b->tileRows, bc = b->chunk;
        const int tiles[] = {256, 512, 1024, 2048, 4096};
        const int chunks[] = {8, 16, 32};
        for (int tr : tiles)
            for (int ch : chunks) {
                restore();
                if (fv_batch_set_shape(b, tr, ch) != FV_OK) continue;
                // first run builds tables and captures the graph, second is timed
                (void)fv_batch_solve(b, o, b->useGraph ? 1 : 0, nullptr);
                restore();
                FV_HIP(hipEventRecord(e0, s));
                const int rc = fv_batch_solve(b, o, b->useGraph ? 1 : 0, nullptr);
                FV_HIP(hipEventRecord(e1, s));
                FV_HIP(hipEventSynchronize(e1));
                float ms = 0;
                FV_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (rc == FV_OK && ms < best) {
                    best = ms;
                    bt = tr;
                    bc = ch;
                }
            }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        restore();
        FV_HIP(hipStreamSynchronize(s));
        int rc = fv_batch_set_shape(b, bt, bc);
        if (rc != FV_OK) throw Error{rc, fv_last_error()};
        cache_store(cache_path, key, bt, bc, best);
        *tile_rows = bt;
        *chunk = bc;
    });
}

int fv_batch_solve(fv_batch* b, const fv_solve_opts* o, int use_graph, fv_solver_stats* stats) {
    return guarded([&] {
        FV_REQUIRE(b && o, "NULL argument");
        FV_REQUIRE(o->rel_tol >= 0.0 && o->max_iter >= 0, "bad solver options");
        b->tol = o->rel_tol;
        b->maxit = o->max_iter;
        b->delta = b->prec == FV_SOLVE_F32_IR ? (o->inner_rel_tol > 0.0 ? o->inner_rel_tol : 1e-3)
                                              : 0.0;
        if (b->useGraph != (use_graph != 0)) drop_graph(b);
        b->useGraph = use_graph != 0;
        for (BMember* m : b->members) m->needStart = true;
        std::vector<CgState> hs;
        int chunks = 0;
        while (true) {
            batch_step(b, hs);
            ++chunks;
            bool running = false;
            for (const CgState& st : hs)
                if (st.done != 2 && st.it < st.maxit) running = true;
            if (!running) break;
        }
        int bad = 0;
        for (size_t i = 0; i < b->members.size(); ++i) {
            BMember* m = b->members[i];
            m->stats.init_res = hs[i].res0;
            m->stats.final_res = hs[i].res;
            m->stats.true_res = hs[i].res;
            m->stats.iters = hs[i].it;
            m->stats.outer_iters = chunks;
            batch_copy_old(b, m);
            if (stats) stats[i] = m->stats;
            if (!(hs[i].res <= o->rel_tol)) ++bad;
        }
        FV_HIP(hipStreamSynchronize(b->ctx->stream));
        if (bad)
            throw Error{FV_ENOCONVERGE, std::to_string(bad) + " system(s) of the batch did not converge"};
    });
}

}  // extern "C"
