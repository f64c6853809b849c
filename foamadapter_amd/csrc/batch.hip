// batch.hip -- many independent implicit solves in one launch sequence.
//
// A batch holds assembled linear systems (fv_linsys, any sizes) and their
// solution vectors.  Their rows are cut into tiles (a multiple of 64 rows, never
// straddling two systems, at most kCgMaxTiles per system); every kernel of the
// CG iteration (cgkernels.hpp) runs over the tile list of the whole batch, one
// workgroup per tile, so small and large systems share the machine and one
// launch serves all of them.  Dot products are reduced per system: each tile
// writes its partials to its own slot and every consuming tile sums its
// system's slots in one fixed order -- so a system's results depend only on its
// own tiling, bit for bit the same alone or in any mix.
//
// Precision: the CG runs in S (float: mixed, double: FP64) on a sliced-ELL copy
// of the matrix; a system whose CG stopped in a chunk (FP32 residual dropped by
// the trigger factor, estimate at the tolerance, iteration limit) gets an FP64
// residual replacement at the end of that chunk (x += d, r = b - A x in FP64,
// convergence test in FP64, CG continues from the replaced residual without
// restarting its direction), so each system stops on its true FP64 relative
// residual while the others iterate on.  A chunk is captured
// once as a hipGraph (per membership) and replayed until every system has
// stopped.  Members can join and leave between chunks (continuous batching):
// joining members are started by their own tiles, the others carry on.
#include "batch.hpp"
#include "cgkernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

using namespace fvhip;

namespace {

constexpr int kB = kCgBlock;

}  // namespace
// ---- host side ----------------------------------------------------------------

fv_batch::~fv_batch() {
    if (graphExec) (void)hipGraphExecDestroy(graphExec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
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

// rows per tile of a member and its tile count
int member_rows(const fv_batch* b, const BMember* m) { return cg_tile_rows(m->N, b->tileRows); }
int member_tiles(const fv_batch* b, const BMember* m) {
    const int R = member_rows(b, m);
    return (m->N + R - 1) / R;
}

template <typename S>
void member_tile_list(const fv_batch* b, int i, int slot0, std::vector<CgTile>& out) {
    const BMember* m = b->members[i];
    const int R = member_rows(b, m), npad = ((m->N + 63) / 64) * 64;
    for (int r0 = 0, k = 0; r0 < m->N; r0 += R, ++k) {
        CgTile t;
        t.prob = i;
        t.row0 = r0;
        t.row1 = std::min(r0 + R, npad);
        t.slot = slot0 + k;
        out.push_back(t);
    }
}

template <typename S>
CgPartials partials_of(fv_batch* b) {
    const int64_t n = std::max(1, b->tileCap);
    return CgPartials{b->partials.p, b->partials.p + n, b->partials.p + 3 * n};
}

constexpr int kXcdGroups = 8;

int next_pow2(int v) {
    int c = 64;
    while (c < v) c *= 2;
    return c;
}

template <typename S>
void build_tables(fv_batch* b) {
    fv_ctx* ctx = b->ctx;
    const int P = (int)b->members.size();
    std::vector<CgSys<S>> hp(P);
    std::vector<CgTile> ht;
    for (int i = 0; i < P; ++i) {
        BMember* m = b->members[i];
        const fv_sparsity* sp = m->ls->sp;
        const int N = m->N;
        CgSys<S>& p = hp[i];
        p.N = N;
        p.tile0 = (int)ht.size();
        member_tile_list<S>(b, i, p.tile0, ht);
        p.ntiles = (int)ht.size() - p.tile0;
        p.sellOff = sp->sellOff.p;
        p.sellCol = sp->sellCol.p;
        p.sellSrc = sp->sellSrc.p;
        p.diag = sp->diagOffset.p;
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
        p.p = v + 2 * (int64_t)N;
        p.q = v + 3 * (int64_t)N;
        p.g = v + 4 * (int64_t)N;
        p.p2 = v + 5 * (int64_t)N;
        p.r64 = (double*)(m->vec.p + 6 * (int64_t)N * sizeof(S));
        p.st = b->states.p + m->slot;
    }
    // Launch table: logical blocks are numbered XCD-contiguously (block i runs on
    // XCD i % 8), so the tiles are split into 8 contiguous groups of about equal
    // rows -- one per XCD, padded to equal length with empty entries -- so that
    // members with large tiles do not load one XCD more than the others.
    int64_t total = 0;
    for (const CgTile& t : ht) total += t.row1 - t.row0;
    int gs[kXcdGroups + 1];
    {
        int k = 0;
        int64_t acc = 0;
        gs[0] = 0;
        for (int i = 0; i < (int)ht.size() && k < kXcdGroups - 1; ++i) {
            if (acc >= total * (k + 1) / kXcdGroups) gs[++k] = i;
            acc += ht[i].row1 - ht[i].row0;
        }
        while (k < kXcdGroups) gs[++k] = (int)ht.size();
    }
    int per = 1;
    for (int k = 0; k < kXcdGroups; ++k) per = std::max(per, gs[k + 1] - gs[k]);
    // the table holds kXcdGroups groups of perCap entries (the graph's grid);
    // perCap follows per within [per, 1.25 per]: growing past it or shrinking
    // below half of it reallocates and invalidates the captured graph
    ++b->tableBuilds;
    const int perCap = b->tileCap / kXcdGroups;
    const bool regrow = per > perCap || per < perCap / 2 || P > b->probCap || !b->gridDev.p;
    if (regrow) {
        const int pc = (per + per / 4 + 7) / 8 * 8;
        b->tileCap = kXcdGroups * pc;
        b->probCap = std::max(b->probCap, next_pow2(P));
        b->probs.alloc(ctx, b->probCap * sizeof(CgSys<S>));
        b->tiles.alloc(ctx, b->tileCap * sizeof(CgTile));
        b->partials.alloc(ctx, 5 * (size_t)b->tileCap);
        b->gridDev.alloc(ctx, sizeof(CgGrid<S>));
        drop_graph(b);
    }
    const int stride = b->tileCap / kXcdGroups;
    std::vector<CgTile> lt((size_t)b->tileCap, CgTile{-1, 0, 0, 0});
    for (int k = 0; k < kXcdGroups; ++k)
        for (int j = gs[k]; j < gs[k + 1]; ++j) lt[(size_t)k * stride + (j - gs[k])] = ht[j];
    b->nTiles = (int)lt.size();
    if (P)
        FV_HIP(hipMemcpyAsync(b->probs.p, hp.data(), P * sizeof(CgSys<S>), hipMemcpyHostToDevice,
                              ctx->stream));
    if (!lt.empty())
        FV_HIP(hipMemcpyAsync(b->tiles.p, lt.data(), lt.size() * sizeof(CgTile),
                              hipMemcpyHostToDevice, ctx->stream));
    const CgGrid<S> g{(const CgTile*)b->tiles.p, (const CgSys<S>*)b->probs.p, b->nTiles, 0};
    FV_HIP(hipMemcpyAsync(b->gridDev.p, &g, sizeof g, hipMemcpyHostToDevice, ctx->stream));
    FV_HIP(hipStreamSynchronize(ctx->stream));
    b->dirty = false;
}

template <typename S>
void enqueue_chunk(fv_batch* b) {
    hipStream_t s = b->ctx->stream;
    const auto* g = (const CgGrid<S>*)b->gridDev.p;
    const CgPartials red = partials_of<S>(b);
    const int G = cg_grid(b->tileCap);
    // the first iteration with a plain SpMV, the others with the previous
    // direction fused into the SpMV, then the last direction (or, for large
    // FP64 batches, three kernels per iteration)
    const int n = b->chunk;
    int64_t rows = 0;
    for (const BMember* m : b->members) rows += m->N;
    if (!cg_fuse_direction<S>(rows)) {
        for (int j = 0; j < n; ++j) {
            kc_spmv<S><<<G, kB, 0, s>>>(g, red);
            kc_update<S><<<G, kB, 0, s>>>(g, red, 0);
            kc_direction<S, 0><<<G, kB, 0, s>>>(g, red, 0);
        }
    } else {
        kc_spmv<S><<<G, kB, 0, s>>>(g, red);
        kc_update<S><<<G, kB, 0, s>>>(g, red, 0);
        for (int j = 1; j < n; ++j) {
            kc_dirspmv<S><<<G, kB, 0, s>>>(g, red, (j - 1) & 1, j & 1);
            kc_update<S><<<G, kB, 0, s>>>(g, red, j & 1);
        }
        kc_direction<S, 0><<<G, kB, 0, s>>>(g, red, (n - 1) & 1);
    }
    // FP64 residual replacement of the systems that stopped in this chunk
    kc_accum<S, 0><<<G, kB, 0, s>>>(g);
    kc_resid<S, 0><<<G, kB, 0, s>>>(g, red);
    kc_replace<S, 0><<<G, kB, 0, s>>>(g, red);
    kc_direction<S, 1><<<G, kB, 0, s>>>(g, red, 0);
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
        ++b->graphBuilds;
    }
    FV_HIP(hipGraphLaunch(b->graphExec, s));
}

template <typename S>
void start_members(fv_batch* b) {
    fv_ctx* ctx = b->ctx;
    hipStream_t s = ctx->stream;
    std::vector<CgTile> st;
    std::vector<CgStartParam> prm(b->members.size());
    // tile list of the members that (re)start, with their global slots, and
    // their start parameters; one batched start launch covers all of them
    int slot = 0;
    for (int i = 0; i < (int)b->members.size(); ++i) {
        BMember* m = b->members[i];
        if (m->needStart) {
            member_tile_list<S>(b, i, slot, st);
            prm[i] = CgStartParam{m->tol > 0.0 ? m->tol : b->tol, b->delta, b->atol,
                                  m->maxit > 0 ? m->maxit : b->maxit, 0};
            m->needStart = false;
        }
        slot += member_tiles(b, m);
    }
    if (st.empty()) return;
    if (b->startTiles.n < st.size() * sizeof(CgTile))
        b->startTiles.alloc(ctx, next_pow2((int)st.size()) * sizeof(CgTile));
    if (b->startParams.n < prm.size() * sizeof(CgStartParam))
        b->startParams.alloc(ctx, next_pow2((int)prm.size()) * sizeof(CgStartParam));
    if (!b->startGridDev.p) b->startGridDev.alloc(ctx, sizeof(CgGrid<S>));
    FV_HIP(hipMemcpyAsync(b->startTiles.p, st.data(), st.size() * sizeof(CgTile),
                          hipMemcpyHostToDevice, s));
    FV_HIP(hipMemcpyAsync(b->startParams.p, prm.data(), prm.size() * sizeof(CgStartParam),
                          hipMemcpyHostToDevice, s));
    const CgGrid<S> hg{(const CgTile*)b->startTiles.p, (const CgSys<S>*)b->probs.p,
                       (int)st.size(), 0};
    FV_HIP(hipMemcpyAsync(b->startGridDev.p, &hg, sizeof hg, hipMemcpyHostToDevice, s));
    const auto* g = (const CgGrid<S>*)b->startGridDev.p;
    const CgPartials red = partials_of<S>(b);
    const int G = cg_grid((int)st.size());
    FV_HIP(hipEventRecord(b->ev[2], s));
    kc_start_tile<S, 1><<<G, kB, 0, s>>>(g, (const CgStartParam*)b->startParams.p);
    kc_scale_tile<S><<<G, kB, 0, s>>>(g);
    kc_resid<S, 1><<<G, kB, 0, s>>>(g, red);
    kc_replace<S, 1><<<G, kB, 0, s>>>(g, red);
    kc_direction<S, 2><<<G, kB, 0, s>>>(g, red, 0);
    FV_LAUNCH_CHECK();
    FV_HIP(hipEventRecord(b->ev[3], s));
    // the upload buffer must outlive the launches
    FV_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    FV_HIP(hipEventElapsedTime(&ms, b->ev[2], b->ev[3]));
    b->startMs += ms;
    ++b->starts;
}

void make_events(fv_batch* b) {
    for (hipEvent_t& e : b->ev)
        if (!e) FV_HIP(hipEventCreate(&e));
}

template <typename S>
void step_t(fv_batch* b, std::vector<CgState>& out) {
    if (b->dirty) build_tables<S>(b);
    out.clear();
    if (b->members.empty()) return;
    make_events(b);
    start_members<S>(b);
    hipStream_t s = b->ctx->stream;
    FV_HIP(hipEventRecord(b->ev[0], s));
    run_chunk<S>(b);
    FV_HIP(hipEventRecord(b->ev[1], s));
    std::vector<CgState> all(b->slotCap);
    FV_HIP(hipMemcpyAsync(all.data(), b->states.p, b->slotCap * sizeof(CgState),
                          hipMemcpyDeviceToHost, s));
    FV_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    FV_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    b->chunkMs += ms;
    ++b->chunks;
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

template <typename S>
void copy_old_t(fv_batch* b) {
    const int G = cg_grid(b->tileCap);
    kc_copy_old<S><<<G, kB, 0, b->ctx->stream>>>((const CgGrid<S>*)b->gridDev.p);
    FV_LAUNCH_CHECK();
}

// x_old = x of every member that has a copy target, in one launch over the tiles
void batch_copy_old(fv_batch* b) {
    if (b->members.empty()) return;
    if (b->prec == FV_SOLVE_F32_IR)
        copy_old_t<float>(b);
    else
        copy_old_t<double>(b);
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

int fv_batch_timing(const fv_batch* b, double* h_ms2, int64_t* h_counts2) {
    return guarded([&] {
        FV_REQUIRE(b && h_ms2 && h_counts2, "NULL argument");
        h_ms2[0] = b->chunkMs;
        h_ms2[1] = b->startMs;
        h_counts2[0] = b->chunks;
        h_counts2[1] = b->starts;
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
        FV_REQUIRE(o->rel_tol >= 0.0 && o->max_iter >= 0 && o->abs_tol >= 0.0,
                   "bad solver options");
        b->tol = o->rel_tol;
        b->atol = o->abs_tol;
        b->maxit = o->max_iter;
        b->delta = b->prec == FV_SOLVE_F32_IR ? (o->inner_rel_tol > 0.0 ? o->inner_rel_tol : kIrDelta)
                                              : 0.0;
        if (b->useGraph != (use_graph != 0)) drop_graph(b);
        b->useGraph = use_graph != 0;
        for (BMember* m : b->members) m->needStart = true;
        std::vector<CgState> hs;
        while (true) {
            batch_step(b, hs);
            bool running = false;
            for (const CgState& st : hs)
                if (st.stop < 2 && st.it < st.maxit) running = true;
            if (!running) break;
        }
        batch_copy_old(b);
        int bad = 0;
        for (size_t i = 0; i < b->members.size(); ++i) {
            BMember* m = b->members[i];
            m->stats.init_res = hs[i].res0;
            m->stats.final_res = hs[i].res;
            m->stats.true_res = hs[i].res;
            m->stats.iters = hs[i].it;
            m->stats.outer_iters = hs[i].nrep;  // FP64 residual replacements, as fv_solve_cg
            if (stats) stats[i] = m->stats;
            // converged on the relative or (when set) the absolute residual, as
            // fv_solve_cg judges it
            if (!(hs[i].res <= o->rel_tol) &&
                !(o->abs_tol > 0.0 && hs[i].res * hs[i].bnorm <= o->abs_tol))
                ++bad;
        }
        FV_HIP(hipStreamSynchronize(b->ctx->stream));
        if (bad)
            throw Error{FV_ENOCONVERGE, std::to_string(bad) + " system(s) of the batch did not converge"};
    });
}

}  // extern "C"
