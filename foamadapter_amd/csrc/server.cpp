// server.cpp -- resident request loop for implicit solves (continuous batching).
//
// Requests arrive from any host thread (fv_server_submit) and queue.  A worker
// thread owns the server's context: it admits queued requests while the
// resident meshes plus the estimated device memory of the requests in flight
// fit the HBM budget (and the in-flight limit),
// builds each admitted request's system on its resident mesh (fields,
// assembly), and then either
//   * batched: steps the batched CG engine -- one chunk of iterations for every
//     system in flight, followed by the FP64 residual replacement -- and after
//     each step retires converged systems (a heat request re-assembles for its
//     next Euler step and rejoins; a finished request is read back and its
//     memory freed), so requests that arrive while others run join at the next
//     chunk; or
//   * plain: runs one request at a time through fv_solve_cg (the drop-in path).
// A batched server runs the requests on meshes of at least solo_rows cells --
// systems that fill the GPU by themselves -- alone through fv_solve_cg, on a
// second lane: its own context (stream and memory pool, with those meshes
// resident a second time) and its own worker thread, so the device interleaves
// these HBM-bound solves with the batch's latency-bound chunks.
// Latency is measured from submission to the result being on the host.
#include "batch.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

using namespace fvhip;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct MeshEntry {
    fv_mesh* mesh = nullptr;
    fv_sparsity* sp = nullptr;
    fv_bc* bcHeat = nullptr;     // fixedValue on every patch
    fv_bc* bcPoisson = nullptr;  // fixedValue on patch 0, zeroGradient elsewhere
    DBuf<int32_t> cellMap;       // device cell -> caller cell (renumbered meshes)
    bool renumbered = false;
    int32_t N = 0, F = 0, B = 0;
    int64_t nnz = 0, sellTotal = 0;
    // the same mesh on the solo lane's context (meshes of at least solo_rows cells)
    std::unique_ptr<MeshEntry> solo;
};

struct Req {
    int64_t id = 0;
    fv_request d{};
    std::vector<double> wall, T0;  // owned copies of the caller's arrays
    int state = FV_REQ_QUEUED;
    std::atomic<bool> cancel{false};  // set by fv_server_cancel, read by the worker
    bool waited = false;              // a fv_server_wait holds the request
    int status = FV_OK;
    Clock::time_point tSubmit, tStart, tDone;
    size_t estimate = 0;
    // the lane that runs it: the server's context and mesh entry, or the solo
    // lane's (set at admission)
    fv_ctx* ctx = nullptr;
    MeshEntry* me = nullptr;
    // device resources while running
    fv_linsys* ls = nullptr;
    DBuf<double> x, xold, Tb, gamma, scratch;
    BMember* member = nullptr;
    int step = 0, iters = 0;
    double res = 0.0;
    // the solution in the caller's numbering, in pinned host memory (from the
    // server's pool; returned to it when the request is collected or dropped)
    double* result = nullptr;
    size_t resultN = 0;
};

}  // namespace

struct fv_server {
    fv_ctx* ctx = nullptr;
    fv_server_opts opts{};
    std::vector<std::unique_ptr<MeshEntry>> meshes;
    std::mutex qmu;               // queue, request table, statistics
    std::condition_variable qcv;  // work arrived / request finished
    std::mutex gpumu;             // users of the context's stream
    std::deque<Req*> pending;
    std::unordered_map<int64_t, Req*> reqs;
    std::vector<Req*> active;
    // the solo lane (batched servers): its context, queue, admitted requests,
    // the users of its stream, and its worker
    fv_ctx* soloCtx = nullptr;
    std::deque<Req*> soloQ;
    std::vector<Req*> soloActive;
    std::mutex solomu;
    std::thread soloWorker;
    int64_t nextId = 1;
    bool stop = false;
    std::thread worker;
    fv_batch* batch = nullptr;
    // statistics
    int64_t submitted = 0, completed = 0, failed = 0, rejected = 0, cancelled = 0;
    std::vector<double> latencies;
    bool haveFirst = false;
    Clock::time_point tFirst, tLast;
    double admitMs = 0, gpuMs = 0, advanceMs = 0;  // worker time split (under qmu when read)
    int64_t chunks = 0, tableBuilds = 0, graphBuilds = 0;  // copies of the batch's counters
    int64_t soloRuns = 0;                                   // requests that ran alone
    double soloMs = 0;                                      // the solo lane's solve time
    size_t laneBytes[2] = {0, 0};  // device bytes in use per lane (each written by its thread)
    double chunkEventMs = 0, startEventMs = 0;             // and of its event timings
    size_t peakBytes = 0;
    int peakActive = 0;
    size_t residentBytes = 0;  // meshes, their sparsity patterns and boundary conditions
    size_t inFlightBytes = 0;  // estimates of the admitted, unfinished requests
    // Pinned host buffers for results, cached by element count: a device-to-host
    // copy into pinned memory runs at the link's rate (pageable memory goes
    // through the runtime's staging, 3-4x slower), and hipHostMalloc itself is
    // too slow to call per request.
    std::mutex pinmu;
    std::unordered_multimap<size_t, double*> pinFree;
    size_t pinCached = 0;  // bytes in pinFree
};

namespace {

constexpr size_t kPinCacheBytes = size_t(1) << 30;  // pinned result buffers kept for reuse

double* pinned_take(fv_server* s, size_t n) {
    {
        std::lock_guard<std::mutex> lk(s->pinmu);
        auto it = s->pinFree.find(n);
        if (it != s->pinFree.end()) {
            double* p = it->second;
            s->pinFree.erase(it);
            s->pinCached -= n * sizeof(double);
            return p;
        }
    }
    void* p = nullptr;
    FV_HIP(hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(double), hipHostMallocDefault));
    return (double*)p;
}

void pinned_give(fv_server* s, double* p, size_t n) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(s->pinmu);
        if (s->pinCached + n * sizeof(double) <= kPinCacheBytes) {
            s->pinFree.emplace(n, p);
            s->pinCached += n * sizeof(double);
            return;
        }
    }
    (void)hipHostFree(p);
}

// Frees a request and the pinned buffer of its result.
void drop_request(fv_server* s, Req* r) {
    pinned_give(s, r->result, r->resultN);
    delete r;
}

// Device bytes of one request: its linear system, fields and face coefficient,
// and the solver's matrix copy and Krylov vectors in the server's precision.
size_t estimate_bytes(const fv_server* s, const MeshEntry& me, int kind) {
    const size_t N = me.N, F = me.F, B = me.B, nnz = me.nnz;
    const size_t e = s->opts.mode == FV_SOLVE_F32_IR ? 4 : 8;
    size_t b = 8 * nnz + 8 * N              // linear system
               + 8 * N * 2 + 8 * B           // x, x_old, boundary values
               + 8 * (F + B)                 // face coefficient (kappa or 1)
               + e * (size_t)me.sellTotal    // sliced-ELL matrix copy
               + 6 * e * N + 8 * N;          // Krylov vectors, correction, FP64 residual
    if (kind == FV_REQ_POISSON) b += 24 * N + 24 * B + 8 * (F + B) + 8 * N;  // U, phi, div
    return b + (1 << 16);
}

void free_request_device(Req* r) {
    if (r->ls) {
        fv_linsys_destroy(r->ls);
        r->ls = nullptr;
    }
    r->x.release();
    r->xold.release();
    r->Tb.release();
    r->gamma.release();
    r->scratch.release();
}

void check(int rc) {
    if (rc != FV_OK) throw Error{rc, fv_last_error()};
}

// Assemble the request's system for its current step.
void assemble(fv_server* s, Req* r) {
    (void)s;
    MeshEntry& me = *r->me;
    if (r->d.kind == FV_REQ_HEAT) {
        // ddt(T) - laplacian(kappa, T), one fused pass
        check(fv_assemble_ddt_laplacian(r->ls, r->xold.p, r->d.dt, r->gamma.p, me.bcHeat, r->Tb.p,
                                        nullptr, -1.0));
    } else {
        check(fv_linsys_zero(r->ls));
        check(fv_assemble_laplacian(r->ls, r->gamma.p, me.bcPoisson, r->Tb.p, nullptr, 1.0));
        // explicit div(phi) computed into scratch (after U, phi)
        double* E = r->scratch.p + 3 * (size_t)me.N + 3 * (size_t)me.B + (me.F + me.B);
        check(fv_assemble_explicit(r->ls, E, -1.0));
    }
}

// Build a request's fields and system on its mesh (its lane's stream).
void setup(fv_server* s, Req* r) {
    MeshEntry& me = *r->me;
    fv_ctx* ctx = r->ctx;
    const int32_t N = me.N, F = me.F, B = me.B;
    check(fv_linsys_create(me.sp, &r->ls));
    r->x.alloc(ctx, N);
    r->xold.alloc(ctx, N);
    r->Tb.alloc(ctx, B);
    r->gamma.alloc(ctx, F + B);
    if (r->d.kind == FV_REQ_HEAT) {
        // initial field: caller's (mapped into the mesh numbering) or seeded
        if (!r->T0.empty()) {
            DBuf<double> tmp;
            tmp.upload(ctx, r->T0.data(), N);
            if (me.renumbered)
                check(fv_permute(ctx, FV_F64, N, 1, 0, 0, me.cellMap.p, tmp.p, r->x.p));
            else
                check(fv_memcpy_d2d(ctx, r->x.p, tmp.p, N * sizeof(double)));
            check(fv_sync(ctx));
        } else {
            check(fv_fill_uniform(ctx, FV_F64, r->x.p, N, 300.0, 400.0, r->d.seed, 0));
        }
        check(fv_memcpy_d2d(ctx, r->xold.p, r->x.p, N * sizeof(double)));
        // wall values, patch by patch
        std::vector<int32_t> poff(me.mesh->nPatches + 1);
        check(fv_mesh_get(me.mesh, FV_G_PATCH_OFFSETS, poff.data()));
        std::vector<double> tb(B, 350.0);
        for (int p = 0; p < me.mesh->nPatches; ++p)
            for (int b = poff[p]; b < poff[p + 1]; ++b)
                tb[b] = r->wall.empty() ? 350.0 : r->wall[p];
        if (B) check(fv_memcpy_h2d(ctx, r->Tb.p, tb.data(), B * sizeof(double)));
        check(fv_fill(ctx, FV_F64, r->gamma.p, F + B, r->d.kappa));
    } else {
        check(fv_fill(ctx, FV_F64, r->x.p, N, 0.0));
        check(fv_fill(ctx, FV_F64, r->xold.p, N, 0.0));
        check(fv_fill(ctx, FV_F64, r->Tb.p, B, 0.0));
        check(fv_fill(ctx, FV_F64, r->gamma.p, F + B, 1.0));
        // velocity U ~ U[-1,1) in cells and on boundary faces -> phi = Sf.U_f -> div(phi)
        r->scratch.alloc(ctx, 3 * (size_t)N + 3 * (size_t)B + (F + B) + N);
        double* U = r->scratch.p;
        double* Ub = U + 3 * (size_t)N;
        double* phi = Ub + 3 * (size_t)B;
        double* E = phi + (F + B);
        check(fv_fill_uniform(ctx, FV_F64, U, 3 * (int64_t)N, -1.0, 1.0, r->d.seed, 1));
        check(fv_fill_uniform(ctx, FV_F64, Ub, 3 * (int64_t)B, -1.0, 1.0, r->d.seed, 2));
        check(fv_flux(ctx, me.mesh, FV_F64, U, Ub, phi));
        check(fv_div_gauss(ctx, me.mesh, FV_LINEAR, FV_F64, 1.0, phi, nullptr, nullptr, E));
    }
    r->step = 0;
    assemble(s, r);
}

void finish(fv_server* s, Req* r, int status) {
    MeshEntry& me = *r->me;
    fv_ctx* ctx = r->ctx;
    if (status == FV_OK) {
        r->result = pinned_take(s, me.N);
        r->resultN = me.N;
        if (me.renumbered) {
            DBuf<double> tmp;
            tmp.alloc(ctx, me.N);
            check(fv_permute(ctx, FV_F64, me.N, 1, 1, 0, me.cellMap.p, r->x.p, tmp.p));
            tmp.download(r->result);
        } else {
            r->x.download(r->result);
        }
    }
    free_request_device(r);
    std::lock_guard<std::mutex> lk(s->qmu);
    s->inFlightBytes -= r->estimate;
    r->status = status;
    r->state = FV_REQ_DONE;
    r->tDone = Clock::now();
    s->tLast = r->tDone;
    if (status == FV_OK) {
        ++s->completed;
        s->latencies.push_back(ms_between(r->tSubmit, r->tDone));
    } else if (status == FV_ECANCELLED) {
        ++s->cancelled;
    } else {
        ++s->failed;
    }
    s->qcv.notify_all();
}

// Admit queued requests that fit; returns them (already moved to RUNNING).
std::vector<Req*> admit(fv_server* s, bool one_at_a_time) {
    std::vector<Req*> out;
    std::lock_guard<std::mutex> lk(s->qmu);
    size_t inUse = s->residentBytes + s->inFlightBytes;
    while (!s->pending.empty()) {
        Req* r = s->pending.front();
        if (r->cancel) {
            s->pending.pop_front();
            r->state = FV_REQ_DONE;
            r->status = FV_ECANCELLED;
            ++s->cancelled;
            s->qcv.notify_all();
            continue;
        }
        const int nActive = (int)(s->active.size() + s->soloActive.size() + out.size());
        if (one_at_a_time && nActive >= 1) break;
        if (s->opts.max_active > 0 && nActive >= s->opts.max_active) break;
        if (s->opts.hbm_budget && inUse + r->estimate > s->opts.hbm_budget && nActive > 0) break;
        s->pending.pop_front();
        r->state = FV_REQ_RUNNING;
        r->tStart = Clock::now();
        inUse += r->estimate;
        s->inFlightBytes += r->estimate;
        out.push_back(r);
    }
    return out;
}

void run_plain(fv_server* s, Req* r) {
    fv_solve_opts o{};
    o.rel_tol = r->d.rel_tol;
    o.max_iter = r->d.max_iter;
    o.mode = s->opts.mode;
    o.inner_rel_tol = 0.0;  // the library's default trigger
    o.check_every = s->opts.chunk > 0 ? s->opts.chunk : 32;
    const int steps = r->d.kind == FV_REQ_HEAT ? std::max(1, r->d.steps) : 1;
    for (int st = 0; st < steps; ++st) {
        if (st > 0) {
            check(fv_memcpy_d2d(r->ctx, r->xold.p, r->x.p, r->x.n * sizeof(double)));
            assemble(s, r);
        }
        fv_solver_stats ss{};
        int rc = fv_solve_cg(r->ls, r->x.p, &o, &ss);
        r->iters += ss.iters;
        r->res = ss.final_res;
        if (rc != FV_OK) throw Error{rc, fv_last_error()};
        if (r->cancel) throw Error{FV_ECANCELLED, "cancelled"};
    }
}

void worker_loop(fv_server* s) {
    const bool batched = s->opts.batched != 0;
    std::vector<CgState> states;
    while (true) {
        {
            std::unique_lock<std::mutex> lk(s->qmu);
            s->qcv.wait(lk, [&] { return s->stop || !s->pending.empty() || !s->active.empty(); });
            if (s->stop) return;
        }
        std::lock_guard<std::mutex> g(s->gpumu);
        // admission
        const Clock::time_point t0 = Clock::now();
        std::vector<Req*> adm = admit(s, !batched);
        for (Req* r : adm) {
            // a system large enough to fill the GPU runs alone through the
            // single-system solver on the solo lane (its polling and replay beat
            // the batch tables' per-chunk round trips); the others join the batch
            MeshEntry* base = s->meshes[r->d.mesh_id].get();
            const bool alone = batched && base->solo != nullptr;
            r->ctx = alone ? s->soloCtx : s->ctx;
            r->me = alone ? base->solo.get() : base;
            if (alone) {
                std::lock_guard<std::mutex> lk(s->qmu);
                s->soloActive.push_back(r);
                s->soloQ.push_back(r);
                s->qcv.notify_all();
                continue;
            }
            try {
                setup(s, r);
                if (batched) {
                    r->member = batch_add(s->batch, r->ls, r->x.p, r->xold.p, r);
                    r->member->tol = r->d.rel_tol;
                    r->member->maxit = r->d.max_iter;
                }
                std::lock_guard<std::mutex> lk(s->qmu);
                s->active.push_back(r);
            } catch (const Error& e) {
                set_error(e.msg);
                finish(s, r, e.code);
            }
        }
        Clock::time_point t1 = Clock::now();
        {
            std::lock_guard<std::mutex> lk(s->qmu);
            s->peakActive =
                std::max(s->peakActive, (int)(s->active.size() + s->soloActive.size()));
            s->laneBytes[0] = s->ctx->in_use;
            s->peakBytes = std::max(s->peakBytes, s->laneBytes[0] + s->laneBytes[1]);
            s->admitMs += ms_between(t0, t1);
        }
        if (!batched) {
            // plain: the next request, alone
            Req* r = nullptr;
            {
                std::lock_guard<std::mutex> lk(s->qmu);
                if (!s->active.empty()) r = s->active.front();
            }
            if (!r) continue;
            int status = FV_OK;
            try {
                run_plain(s, r);
            } catch (const Error& e) {
                set_error(e.msg);
                status = e.code;
            }
            {
                std::lock_guard<std::mutex> lk(s->qmu);
                s->active.erase(std::find(s->active.begin(), s->active.end(), r));
                s->gpuMs += ms_between(t1, Clock::now());
            }
            try {
                finish(s, r, status);
            } catch (const Error& e) {
                set_error(e.msg);
                finish(s, r, e.code);
            }
            continue;
        }
        if (s->batch->members.empty()) continue;
        try {
            batch_step(s->batch, states);
            std::lock_guard<std::mutex> lk(s->qmu);
            s->gpuMs += ms_between(t1, Clock::now());
            ++s->chunks;
            s->tableBuilds = s->batch->tableBuilds;
            s->graphBuilds = s->batch->graphBuilds;
            s->chunkEventMs = s->batch->chunkMs;
            s->startEventMs = s->batch->startMs;
        } catch (const Error& e) {
            // a device error fails everything in flight
            set_error(e.msg);
            std::vector<Req*> act;
            {
                std::lock_guard<std::mutex> lk(s->qmu);
                act.swap(s->active);
            }
            for (Req* r : act) {
                if (r->member) batch_remove(s->batch, r->member);
                r->member = nullptr;
                finish(s, r, e.code);
            }
            continue;
        }
        // retire / advance members (states are in member order)
        const Clock::time_point t2 = Clock::now();
        std::vector<BMember*> members = s->batch->members;
        for (size_t i = 0; i < members.size(); ++i) {
            BMember* m = members[i];
            Req* r = (Req*)m->user;
            const CgState& st = states[i];
            const bool converged = st.stop == 2;
            const bool failed =
                !converged && (st.stop == 3 || st.it >= st.maxit || !std::isfinite(st.res));
            if (!converged && !failed && !r->cancel) continue;
            r->iters += st.it;
            r->res = st.res;
            int status = FV_OK;
            bool done = true;
            if (r->cancel) {
                status = FV_ECANCELLED;
            } else if (failed) {
                status = FV_ENOCONVERGE;
            } else if (r->d.kind == FV_REQ_HEAT && r->step + 1 < std::max(1, r->d.steps)) {
                // next Euler step: T_old <- T, re-assemble, restart this member
                r->step += 1;
                try {
                    check(fv_memcpy_d2d(s->ctx, r->xold.p, r->x.p, r->x.n * sizeof(double)));
                    assemble(s, r);
                    batch_restart(s->batch, m);
                    done = false;
                } catch (const Error& e) {
                    set_error(e.msg);
                    status = e.code;
                }
            }
            if (done) {
                batch_remove(s->batch, m);
                r->member = nullptr;
                {
                    std::lock_guard<std::mutex> lk(s->qmu);
                    s->active.erase(std::find(s->active.begin(), s->active.end(), r));
                }
                try {
                    finish(s, r, status);
                } catch (const Error& e) {
                    set_error(e.msg);
                    finish(s, r, e.code);
                }
            }
        }
        std::lock_guard<std::mutex> lk(s->qmu);
        s->advanceMs += ms_between(t2, Clock::now());
    }
}

// The solo lane: requests on the largest meshes, one at a time through
// fv_solve_cg on the lane's own context, while the worker steps the batch.
void solo_loop(fv_server* s) {
    while (true) {
        Req* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->qmu);
            s->qcv.wait(lk, [&] { return s->stop || !s->soloQ.empty(); });
            if (s->stop) return;
            r = s->soloQ.front();
            s->soloQ.pop_front();
        }
        std::lock_guard<std::mutex> g(s->solomu);
        const Clock::time_point t0 = Clock::now();
        int status = FV_OK;
        try {
            if (r->cancel) throw Error{FV_ECANCELLED, "cancelled"};
            setup(s, r);
            run_plain(s, r);
        } catch (const Error& e) {
            set_error(e.msg);
            status = e.code;
        }
        {
            std::lock_guard<std::mutex> lk(s->qmu);
            s->soloActive.erase(std::find(s->soloActive.begin(), s->soloActive.end(), r));
            s->soloMs += ms_between(t0, Clock::now());
            ++s->soloRuns;
            s->laneBytes[1] = s->soloCtx->in_use;
            s->peakBytes = std::max(s->peakBytes, s->laneBytes[0] + s->laneBytes[1]);
        }
        try {
            finish(s, r, status);
        } catch (const Error& e) {
            set_error(e.msg);
            finish(s, r, e.code);
        }
    }
}

// A resident mesh on context ctx from a (possibly renumbered) description:
// device mesh, cell map, sparsity, the two boundary-condition sets and the
// staged-gather lists.  Returns the device bytes it took.
size_t make_entry(fv_ctx* ctx, const fv_polymesh_desc* d, const std::vector<int32_t>* cmap,
                  MeshEntry& me) {
    const size_t used0 = ctx->in_use;
    check(fv_mesh_create_poly(ctx, d, &me.mesh));
    if (cmap) {
        me.cellMap.upload(ctx, cmap->data(), cmap->size());
        check(fv_sync(ctx));
        me.renumbered = true;
    }
    check(fv_sparsity_create(me.mesh, &me.sp));
    const int np = me.mesh->nPatches;
    std::vector<int32_t> kh(np, FV_BC_FIXED_VALUE), kp(np, FV_BC_FIXED_GRADIENT);
    if (np) kp[0] = FV_BC_FIXED_VALUE;
    check(fv_bc_create(me.mesh, kh.data(), &me.bcHeat));
    check(fv_bc_create(me.mesh, kp.data(), &me.bcPoisson));
    {
        // A Poisson request's explicit surface integral builds the mesh's
        // staged-gather lists on first use (24 ms at 128^3, 55 ms at 160^3):
        // build them with the mesh, by the same call, not in a request.
        DBuf<double> phi, out;
        phi.alloc(ctx, (size_t)me.mesh->nInt + me.mesh->nB);
        out.alloc(ctx, (size_t)me.mesh->nCells);
        check(fv_fill(ctx, FV_F64, phi.p, (int64_t)phi.n, 0.0));
        check(fv_div_gauss(ctx, me.mesh, FV_LINEAR, FV_F64, 1.0, phi.p, nullptr, nullptr, out.p));
        check(fv_sync(ctx));
    }
    me.N = me.mesh->nCells;
    me.F = me.mesh->nInt;
    me.B = me.mesh->nB;
    me.nnz = me.sp->nnz;
    me.sellTotal = me.sp->sellTotal;
    return ctx->in_use - used0;
}

void destroy_entry(MeshEntry& me) {
    if (me.solo) destroy_entry(*me.solo);
    if (me.bcHeat) fv_bc_destroy(me.bcHeat);
    if (me.bcPoisson) fv_bc_destroy(me.bcPoisson);
    if (me.sp) fv_sparsity_destroy(me.sp);
    me.cellMap.release();
    if (me.mesh) fv_mesh_destroy(me.mesh);
}

MeshEntry& mesh_of(fv_server* s, int id) {
    FV_REQUIRE(id >= 0 && id < (int)s->meshes.size(), "unknown mesh id");
    return *s->meshes[id];
}

}  // namespace

extern "C" {

int fv_server_create(int device, const fv_server_opts* opts, fv_server** out) {
    fv_server* s = nullptr;
    int rc = guarded([&] {
        FV_REQUIRE(opts && out, "NULL argument");
        FV_REQUIRE(opts->mode == FV_SOLVE_F64 || opts->mode == FV_SOLVE_F32_IR, "unknown mode");
        s = new fv_server();
        s->opts = *opts;
        if (s->opts.solo_rows == 0) s->opts.solo_rows = 1 << 20;
        check(fv_ctx_create(device, 0, &s->ctx));
        // measured on the served stream of jittered cubes, n ~ logU[16,160]:
        // 4096-row tiles 114 solves/s, 2048 112, 1024 107, 8192 112
        const int tr = opts->tile_rows > 0 ? opts->tile_rows : 4096;
        const int ch = opts->chunk > 0 ? opts->chunk : 16;
        check(fv_batch_create(s->ctx, opts->mode, tr, ch, &s->batch));
        s->batch->useGraph = opts->use_graph != 0;
        if (s->opts.batched && s->opts.solo_rows > 0) {
            check(fv_ctx_create(device, 0, &s->soloCtx));
            s->soloWorker = std::thread(solo_loop, s);
        }
        s->worker = std::thread(worker_loop, s);
        *out = s;
    });
    if (rc != FV_OK && s) {
        if (s->soloWorker.joinable()) {
            {
                std::lock_guard<std::mutex> lk(s->qmu);
                s->stop = true;
                s->qcv.notify_all();
            }
            s->soloWorker.join();
        }
        if (s->batch) fv_batch_destroy(s->batch);
        if (s->soloCtx) fv_ctx_destroy(s->soloCtx);
        if (s->ctx) fv_ctx_destroy(s->ctx);
        delete s;
    }
    return rc;
}

int fv_server_add_mesh(fv_server* s, const fv_polymesh_desc* d, int renumber, int* mesh_id) {
    return guarded([&] {
        FV_REQUIRE(s && d && mesh_id, "NULL argument");
        auto me = std::make_unique<MeshEntry>();
        fv_polymesh_desc r = *d;
        std::vector<int32_t> foff, fpts, own, nbr, cmap, fmap;
        if (renumber) {
            const int64_t nF = d->n_faces, N = d->n_cells;
            foff.resize(nF + 1);
            fpts.resize(d->h_face_offsets[nF]);
            own.resize(nF);
            nbr.resize(d->n_internal_faces);
            cmap.resize(N);
            fmap.resize(nF);
            check(fv_polymesh_renumber(d, 1, foff.data(), fpts.data(), own.data(), nbr.data(),
                                       cmap.data(), fmap.data()));
            r.h_face_offsets = foff.data();
            r.h_face_points = fpts.data();
            r.h_owner = own.data();
            r.h_neighbour = nbr.data();
        }
        size_t bytes = 0;
        {
            std::lock_guard<std::mutex> g(s->gpumu);
            bytes += make_entry(s->ctx, &r, renumber ? &cmap : nullptr, *me);
        }
        if (s->soloCtx && me->N >= s->opts.solo_rows) {
            // the solo lane runs requests on this mesh: resident there too
            auto solo = std::make_unique<MeshEntry>();
            std::lock_guard<std::mutex> g(s->solomu);
            try {
                bytes += make_entry(s->soloCtx, &r, renumber ? &cmap : nullptr, *solo);
            } catch (...) {
                destroy_entry(*solo);
                std::lock_guard<std::mutex> g2(s->gpumu);
                destroy_entry(*me);
                throw;
            }
            me->solo = std::move(solo);
        }
        std::lock_guard<std::mutex> lk(s->qmu);
        s->residentBytes += bytes;
        *mesh_id = (int)s->meshes.size();
        s->meshes.push_back(std::move(me));
    });
}

int fv_server_mesh_cells(fv_server* s, int mesh_id, int32_t* n) {
    return guarded([&] {
        FV_REQUIRE(s && n, "NULL argument");
        std::lock_guard<std::mutex> lk(s->qmu);
        *n = mesh_of(s, mesh_id).N;
    });
}

int fv_server_submit(fv_server* s, const fv_request* q, int64_t* id) {
    return guarded([&] {
        FV_REQUIRE(s && q && id, "NULL argument");
        FV_REQUIRE(q->kind == FV_REQ_HEAT || q->kind == FV_REQ_POISSON, "unknown request kind");
        FV_REQUIRE(q->rel_tol > 0.0 && q->max_iter > 0, "bad solver settings");
        FV_REQUIRE(q->kind != FV_REQ_HEAT || (q->dt > 0.0 && q->steps >= 1), "bad time step");
        auto* r = new Req();
        r->d = *q;
        std::unique_lock<std::mutex> lk(s->qmu);
        MeshEntry& me = mesh_of(s, q->mesh_id);
        if (q->h_wall_values) r->wall.assign(q->h_wall_values, q->h_wall_values + me.mesh->nPatches);
        if (q->h_T0) r->T0.assign(q->h_T0, q->h_T0 + me.N);
        r->d.h_wall_values = nullptr;
        r->d.h_T0 = nullptr;
        r->estimate = estimate_bytes(s, me, q->kind);
        r->tSubmit = Clock::now();
        r->id = s->nextId++;
        if (!s->haveFirst) {
            s->haveFirst = true;
            s->tFirst = r->tSubmit;
        }
        ++s->submitted;
        if (s->opts.hbm_budget && s->residentBytes + r->estimate > s->opts.hbm_budget) {
            ++s->rejected;
            delete r;
            throw Error{FV_ENOMEM_BUDGET, "request needs more device memory than the server budget"};
        }
        s->reqs[r->id] = r;
        s->pending.push_back(r);
        *id = r->id;
        s->qcv.notify_all();
    });
}

int fv_server_poll(fv_server* s, int64_t id, int* state) {
    return guarded([&] {
        FV_REQUIRE(s && state, "NULL argument");
        std::lock_guard<std::mutex> lk(s->qmu);
        auto it = s->reqs.find(id);
        FV_REQUIRE(it != s->reqs.end(), "unknown request id");
        *state = it->second->state;
    });
}

int fv_server_wait(fv_server* s, int64_t id, double* h_result, fv_request_stats* stats) {
    int status = FV_OK;
    int rc = guarded([&] {
        FV_REQUIRE(s, "NULL server");
        std::unique_lock<std::mutex> lk(s->qmu);
        auto it = s->reqs.find(id);
        FV_REQUIRE(it != s->reqs.end(), "unknown request id");
        Req* r = it->second;
        // one waiter per request: the waiter frees it
        FV_REQUIRE(!r->waited, "request is already being waited on");
        r->waited = true;
        s->qcv.wait(lk, [&] { return r->state == FV_REQ_DONE; });
        if (stats) {
            stats->status = r->status;
            stats->iters = r->iters;
            stats->final_res = r->res;
            stats->queue_ms = r->tStart.time_since_epoch().count() ? ms_between(r->tSubmit, r->tStart) : 0.0;
            stats->latency_ms = ms_between(r->tSubmit, r->tDone);
        }
        if (h_result && r->status == FV_OK) std::copy(r->result, r->result + r->resultN, h_result);
        status = r->status;
        s->reqs.erase(it);
        lk.unlock();
        drop_request(s, r);
    });
    if (rc != FV_OK) return rc;
    if (status != FV_OK) set_error("request finished with status " + std::to_string(status));
    return status;
}

int fv_server_cancel(fv_server* s, int64_t id) {
    return guarded([&] {
        FV_REQUIRE(s, "NULL server");
        std::lock_guard<std::mutex> lk(s->qmu);
        auto it = s->reqs.find(id);
        FV_REQUIRE(it != s->reqs.end(), "unknown request id");
        it->second->cancel = true;
        s->qcv.notify_all();
    });
}

int fv_server_request_bytes(fv_server* s, int mesh_id, int kind, size_t* bytes) {
    return guarded([&] {
        FV_REQUIRE(s && bytes, "NULL argument");
        FV_REQUIRE(kind == FV_REQ_HEAT || kind == FV_REQ_POISSON, "unknown request kind");
        std::lock_guard<std::mutex> lk(s->qmu);
        *bytes = estimate_bytes(s, mesh_of(s, mesh_id), kind);
    });
}

int fv_server_get_stats(fv_server* s, fv_server_stats* o) {
    return guarded([&] {
        FV_REQUIRE(s && o, "NULL argument");
        std::lock_guard<std::mutex> lk(s->qmu);
        *o = fv_server_stats{};
        o->submitted = s->submitted;
        o->completed = s->completed;
        o->failed = s->failed;
        o->rejected = s->rejected;
        o->cancelled = s->cancelled;
        o->peak_bytes = s->peakBytes;
        o->peak_active = s->peakActive;
        o->resident_bytes = s->residentBytes;
        o->admit_ms = s->admitMs;
        o->gpu_ms = s->gpuMs;
        o->advance_ms = s->advanceMs;
        o->chunks = s->chunks;
        o->table_builds = s->tableBuilds;
        o->graph_builds = s->graphBuilds;
        o->chunk_event_ms = s->chunkEventMs;
        o->start_event_ms = s->startEventMs;
        o->solo_runs = s->soloRuns;
        o->solo_ms = s->soloMs;
        if (s->haveFirst && s->completed) o->wall_s = ms_between(s->tFirst, s->tLast) * 1e-3;
        if (o->wall_s > 0) o->solves_per_s = s->completed / o->wall_s;
        if (!s->latencies.empty()) {
            std::vector<double> l = s->latencies;
            std::sort(l.begin(), l.end());
            auto q = [&](double p) {
                const double pos = p * (l.size() - 1);
                const size_t i = (size_t)pos;
                const double f = pos - i;
                return i + 1 < l.size() ? l[i] * (1 - f) + l[i + 1] * f : l[i];
            };
            o->p50_ms = q(0.50);
            o->p99_ms = q(0.99);
            double sum = 0;
            for (double v : l) sum += v;
            o->mean_ms = sum / l.size();
            o->max_ms = l.back();
        }
    });
}

int fv_server_destroy(fv_server* s) {
    return guarded([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->qmu);
            s->stop = true;
            s->qcv.notify_all();
        }
        if (s->worker.joinable()) s->worker.join();
        if (s->soloWorker.joinable()) s->soloWorker.join();
        for (Req* r : s->active) free_request_device(r);
        for (Req* r : s->soloActive) free_request_device(r);
        for (auto& kv : s->reqs) drop_request(s, kv.second);
        for (auto& kv : s->pinFree) (void)hipHostFree(kv.second);
        s->pinFree.clear();
        if (s->batch) fv_batch_destroy(s->batch);
        for (auto& me : s->meshes) destroy_entry(*me);
        if (s->soloCtx) fv_ctx_destroy(s->soloCtx);
        fv_ctx_destroy(s->ctx);
        delete s;
    });
}

}  // extern "C"
