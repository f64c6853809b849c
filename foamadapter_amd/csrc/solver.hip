// solver.hip -- Jacobi-preconditioned conjugate gradients on the assembled CSR
// system (Ginkgo solver::Cg + preconditioner::Jacobi{max_block_size 1} with the
// criteria {iteration, relative_residual_norm}, src/compatibility/fvSolution.cpp:22-28,
// 50-55,106-142), in FP64, and in FP32 inside FP64 iterative refinement.
//
// The iteration runs as the row-tile kernels of cgkernels.hpp (CG on the
// symmetrically Jacobi-scaled system, whose iterates are Jacobi-PCG's) over
// implicit tiles of the one system (at most kCgMaxTiles tiles, rows in multiples of 64).
// All scalars live on the device; kernels turn into no-ops once the stop flag is
// set, and the host polls the flag every `check` iterations, so a solve needs
// no host round trip per iteration.  The workspace stays with the linear
// system, so repeated solves (time steps) allocate nothing, and a repeated
// solve launches the previous solve's schedule at once -- as a captured
// hipGraph once the same schedule has come five times in a row.
#include "cgkernels.hpp"

#include <cmath>
#include <deque>
#include <vector>

using namespace fvhip;

namespace {

constexpr int kB = kCgBlock;

// Workspace of one CG in value type S on one linear system.
template <typename S>
struct CgWork : fv_solver_cache {
    DBuf<S> dbuf, r, p, p2, q, g, sval;
    DBuf<double> r64, pq, rz, rb;
    DBuf<CgState> st, st2;
    DBuf<CgSys<S>> sys;
    CgSys<S> h{};
    CgGrid<S> grid{};
    DBuf<CgGrid<S>> gdev;  // device copy of grid (the kernels' launch description)
    CgPartials red{};
    int G = 0;  // launch grid (multiple of 8)
    // the captured replay of a repeated solve and the schedules it was made for
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<double> gkey, seenKey;
    int seenCount = 0;  // consecutive solves with seenKey
    int lastIters = 0;  // iterations of the previous solve (time steps repeat them)
    std::vector<int> lastRounds;  // FP32 iterations per round of the previous mixed solve

    CgWork(fv_ctx* c, const fv_linsys* ls) {
        const fv_sparsity* sp = ls->sp;
        const int N = sp->nRows;
        const int R = cg_tile_rows(N, kB);
        const int tiles = (N + R - 1) / R;
        G = cg_grid(tiles);
        r.alloc(c, N);
        p.alloc(c, N);
        p2.alloc(c, N);
        q.alloc(c, N);
        g.alloc(c, N);
        dbuf.alloc(c, N);
        sval.alloc(c, sp->sellTotal);
        r64.alloc(c, N);
        pq.alloc(c, tiles);
        rz.alloc(c, 2 * (size_t)tiles);
        rb.alloc(c, 2 * (size_t)tiles);
        st.alloc(c, 1);
        st2.alloc(c, 1);
        sys.alloc(c, 1);
        h.N = N;
        h.tile0 = 0;
        h.ntiles = tiles;
        h.sellOff = sp->sellOff.p;
        h.sellCol = sp->sellCol.p;
        h.sellSrc = sp->sellSrc.p;
        h.diag = sp->diagOffset.p;
        h.sval = sval.p;
        h.rowOffs = sp->rowOffs.p;
        h.colIdxs = sp->colIdxs.p;
        h.vals = ls->values.p;
        h.rhs = ls->rhs.p;
        h.r = r.p;
        h.p = p.p;
        h.p2 = p2.p;
        h.q = q.p;
        h.g = g.p;
        h.d = dbuf.p;
        h.r64 = r64.p;
        h.st = st.p;
        grid.tiles = nullptr;
        grid.probs = sys.p;
        grid.nTiles = tiles;
        grid.rowsPerBlock = R;
        gdev.upload(c, &grid, 1);
        FV_HIP(hipStreamSynchronize(c->stream));
        red = CgPartials{pq.p, rz.p, rb.p};
    }

    ~CgWork() override { drop_graph(); }
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }

    // Binds the solution vector (the device copy of the system description
    // the kernels read).
    void bind(fv_ctx* c, double* x) {
        h.x = x;
        h.xold = nullptr;
        FV_HIP(hipMemcpyAsync(sys.p, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
    }
    // Loads the matrix (values may have changed since the last solve): the
    // diagonal's signed square roots g, d = 0, and the scaled off-diagonal
    // values in S.
    void prep(hipStream_t s) {
        kc_start_tile<S, 0><<<G, kB, 0, s>>>(gdev.p, nullptr);
        kc_scale_tile<S><<<G, kB, 0, s>>>(gdev.p);
    }
    // Runs enqueue() -- kernel launches only, whose by-value arguments `key`
    // determines -- eagerly until the same key has come kCaptureAfter times in
    // a row, then as a captured hipGraph, replayed while the key keeps coming.
    // Measured on MI355X: a capture costs about 0.15 ms (3.3 ms for the first
    // in a process), a replay saves about 0.03 ms on a 32^3 heat step, so only
    // long runs of one schedule gain: short marches (5-step requests) and
    // schedules that change every solve (the pressure solves inside PISO)
    // never capture.
    static constexpr int kCaptureAfter = 5;
    template <typename F>
    void replay(hipStream_t s, const std::vector<double>& key, F&& enqueue) {
        if (gexec && key == gkey) {
            FV_HIP(hipGraphLaunch(gexec, s));
            return;
        }
        if (key != seenKey) {
            seenKey = key;
            seenCount = 0;
        }
        if (++seenCount < kCaptureAfter) {
            enqueue();
            return;
        }
        // the graph being replaced has no launch in flight: every solve ends
        // with a synchronisation on the stream
        drop_graph();
        FV_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        enqueue();
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (e != hipSuccess) {
            graph = nullptr;
            throw Error{FV_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e)};
        }
        FV_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        gkey = key;
        FV_HIP(hipGraphLaunch(gexec, s));
    }

    // FP64 residual of the current x, r = (S) (r64 / g), p = r, r.r
    void start(hipStream_t s, double tol, int maxit, double delta, double atol) {
        kc_init_state<<<1, 64, 0, s>>>(st.p, 1, tol, maxit, delta, atol);
        kc_resid<S, 1><<<G, kB, 0, s>>>(gdev.p, red);
        kc_replace<S, 1><<<G, kB, 0, s>>>(gdev.p, red);
        kc_direction<S, 2><<<G, kB, 0, s>>>(gdev.p, red, 0);
    }
    // n iterations: the first with a plain SpMV, the others with the previous
    // direction fused into the SpMV, then the last direction (or, for large
    // FP64 systems, three kernels per iteration)
    void chunk(hipStream_t s, int n) {
        if (!cg_fuse_direction<S>(h.N)) {
            for (int j = 0; j < n; ++j) {
                kc_spmv<S><<<G, kB, 0, s>>>(gdev.p, red);
                kc_update<S><<<G, kB, 0, s>>>(gdev.p, red, 0);
                kc_direction<S, 0><<<G, kB, 0, s>>>(gdev.p, red, 0);
            }
            return;
        }
        kc_spmv<S><<<G, kB, 0, s>>>(gdev.p, red);
        kc_update<S><<<G, kB, 0, s>>>(gdev.p, red, 0);
        for (int j = 1; j < n; ++j) {
            kc_dirspmv<S><<<G, kB, 0, s>>>(gdev.p, red, (j - 1) & 1, j & 1);
            kc_update<S><<<G, kB, 0, s>>>(gdev.p, red, j & 1);
        }
        kc_direction<S, 0><<<G, kB, 0, s>>>(gdev.p, red, (n - 1) & 1);
    }
    void replacement(hipStream_t s) {
        kc_accum<S, 0><<<G, kB, 0, s>>>(gdev.p);
        kc_resid<S, 0><<<G, kB, 0, s>>>(gdev.p, red);
        kc_replace<S, 0><<<G, kB, 0, s>>>(gdev.p, red);
        kc_direction<S, 1><<<G, kB, 0, s>>>(gdev.p, red, 0);
    }
    CgState state(hipStream_t s) {
        CgState o{};
        FV_HIP(hipMemcpyAsync(&o, st.p, sizeof o, hipMemcpyDeviceToHost, s));
        FV_HIP(hipStreamSynchronize(s));
        return o;
    }
    // x += |G|^-1 d, d = 0 (the plain FP64 solve's correction into x)
    void finish(hipStream_t s) { kc_accum<S, 1><<<G, kB, 0, s>>>(gdev.p); }
    // explicit FP64 relative residual ||b - A x|| / ||b|| of the bound x into st2
    void enqueue_true_residual(hipStream_t s) {
        kc_resid<S, 1><<<G, kB, 0, s>>>(gdev.p, red);
        kc_resid_norm<S><<<1, kB, 0, s>>>(sys.p, red, st2.p);
    }
    double true_residual(hipStream_t s) {
        enqueue_true_residual(s);
        FV_LAUNCH_CHECK();
        CgState o{};
        FV_HIP(hipMemcpyAsync(&o, st2.p, sizeof o, hipMemcpyDeviceToHost, s));
        FV_HIP(hipStreamSynchronize(s));
        return o.res;
    }
    // the state and (when enqueued) the explicit residual, one synchronisation
    CgState state2(hipStream_t s, CgState* o2) {
        CgState o{};
        FV_HIP(hipMemcpyAsync(&o, st.p, sizeof o, hipMemcpyDeviceToHost, s));
        FV_HIP(hipMemcpyAsync(o2, st2.p, sizeof *o2, hipMemcpyDeviceToHost, s));
        FV_HIP(hipStreamSynchronize(s));
        return o;
    }
};

template <typename S>
CgWork<S>& work_of(fv_linsys* ls, bool mixed) {
    auto& slot = ls->cg[mixed ? 1 : 0];
    if (!slot) slot.reset(new CgWork<S>(ls->sp->mesh->ctx, ls));
    return *static_cast<CgWork<S>*>(slot.get());
}

// Iterations a CG run still needs for its relative residual to fall from r to
// a target, at the rate (log-residual decrease per iteration) observed between
// the two latest polls; -1 while no rate is known.
struct ConvRate {
    int it = -1;
    double r = 0.0, rate = 0.0;
    void observe(int it1, double r1) {
        if (it >= 0 && it1 > it && r1 > 0.0 && r > r1) rate = std::log(r / r1) / (it1 - it);
        it = it1;
        r = r1;
    }
    int remaining(double r1, double target) const {
        if (!(rate > 0.0) || !(r1 > 0.0)) return -1;
        if (r1 <= target) return 0;
        const double n = std::ceil(std::log(r1 / target) / rate);
        return n > 1e6 ? 1000000 : (int)n;
    }
};

// Extra iterations launched beyond a predicted stop (CG residuals are not
// monotone, so the prediction errs both ways).
constexpr int kPredictMargin = 2;

// Chunks of iterations are queued with their state copied to pinned memory
// behind an event; at most two such polls are in flight, so the device works on
// one chunk while the host reads the state of the one before.  Once two polls
// have given a convergence rate, each chunk is sized to end just after the
// predicted stop (the iterations still missing minus those already queued), so
// a solve overshoots its stop by a few no-op iterations, not by a whole chunk;
// the first chunks ramp up 4, 8, 16, ...  Kernels of a stopped system are
// no-ops (about 2 us each).
struct PollPipe {
    CgState* mirror;
    fv_ctx* ctx;
    int pend[2] = {-1, -1}, npend = 0, slot = 0;
    void push(hipStream_t s, const CgState* dev) {
        FV_HIP(hipMemcpyAsync(&mirror[slot], dev, sizeof(CgState), hipMemcpyDeviceToHost, s));
        FV_HIP(hipEventRecord(ctx->pollEv[slot], s));
        pend[npend++] = slot;
        slot ^= 1;
    }
    CgState pop() {
        const int q = pend[0];
        FV_HIP(hipEventSynchronize(ctx->pollEv[q]));
        pend[0] = pend[1];
        --npend;
        return mirror[q];
    }
};

// Plain FP64 CG from x; stops on the recurrence residual estimate.  The
// iteration accumulates its scaled correction d; finish() adds it to x before
// any explicit residual (adding it early, as the repeated-solve read does, only
// moves where d restarts from zero).  *true_res receives the explicit FP64
// residual of the result.  A solve of a system whose
// previous solve (a time step before) took k iterations first launches k at
// once and reads the state with the explicit residual: one synchronisation when
// the iteration count repeats (the whole sequence, from the matrix scaling to
// the explicit residual, replayed as a hipGraph unless `eager`).
CgState run_cg(fv_ctx* ctx, CgWork<double>& w, double tol, int maxit, int check, double atol,
               bool eager, double* true_res) {
    hipStream_t s = ctx->stream;
    static_assert(2 * sizeof(CgState) <= fv_ctx::kPollBytes, "poll area too small");
    PollPipe pp{static_cast<CgState*>(ctx_poll_host(ctx)), ctx};
    CgState h{}, t{};
    int launched = 0;
    bool haveTrue = false;
    check = std::max(check, 1);
    if (w.lastIters > 0) {
        const int n = std::min(w.lastIters, maxit);
        auto seq = [&] {
            w.prep(s);
            w.start(s, tol, maxit, 0.0, atol);
            w.chunk(s, n);
            w.finish(s);
            w.enqueue_true_residual(s);
        };
        if (eager)
            seq();
        else
            w.replay(s, {(double)n, tol, (double)maxit, atol}, seq);
        FV_LAUNCH_CHECK();
        launched = n;
        h = w.state2(s, &t);
        haveTrue = true;
    } else {
        w.prep(s);
        w.start(s, tol, maxit, 0.0, atol);
        FV_LAUNCH_CHECK();
    }
    ConvRate cr;
    cr.observe(h.it, h.res);
    int ramp = 4;
    while (!h.stop) {
        bool queued = false;
        if (pp.npend < 2 && launched < maxit) {
            const double target = std::max(tol, atol > 0.0 ? atol / h.bnorm : 0.0);
            const int R = h.it > 0 ? cr.remaining(h.res, target) : -1;
            int need;
            if (R < 0) {
                need = ramp;
                ramp = std::min(2 * ramp, check);
            } else {
                need = R + kPredictMargin - (launched - h.it);
            }
            if (pp.npend == 0) need = std::max(need, 1);
            need = std::min(std::min(need, check), maxit - launched);
            if (need > 0) {
                w.chunk(s, need);
                FV_LAUNCH_CHECK();
                launched += need;
                pp.push(s, w.st.p);
                queued = true;
                haveTrue = false;
            }
        }
        if (pp.npend == 0) break;               // the iteration limit is launched and read
        if (queued && pp.npend < 2) continue;   // keep a second chunk queued before waiting
        h = pp.pop();
        cr.observe(h.it, h.res);
    }
    while (pp.npend) h = pp.pop();  // chunks queued behind the stop ran as no-ops
    // a system that needed no iteration polls (blocking) after the first
    w.lastIters = h.stop ? std::max(h.it, 1) : h.it;
    if (!haveTrue) w.finish(s);
    *true_res = haveTrue ? t.res : w.true_residual(s);
    return h;
}

// ---- mixed precision: FP32 CG with FP64 residual replacement ---------------
//
// The FP32 CG accumulates its correction d (FP32); it stops when the FP32
// residual has dropped by `delta` since the last replacement, its estimate
// reaches the tolerance, or the iteration limit.  A round is a chunk of FP32
// iterations followed by the replacement kernels, which act only on a stopped
// CG: in FP64 x += d, r = b - A x, the convergence test on ||r||/||b||; the
// FP32 residual is replaced by the true one and CG continues with its search
// direction (beta from the replaced residual), so the Krylov space is not
// restarted.
//
// Rounds are queued and read like run_cg's chunks (two polls in flight), and
// once a rate is known each round is sized to end just after the predicted
// stop of the FP32 CG (drop trigger or tolerance), where its replacement
// follows.  The prediction runs through the rounds already queued: a segment
// (from one replacement to the next trigger) whose stop they cover is followed
// by the next segment, about as long (the residual falls by the same factor),
// so the next round is sized for that one and the queue does not drain at
// every replacement.  The rounds of a solve are recorded: the next solve of the same
// system (a time step) queues the same rounds at once and reads the state
// once, continuing adaptively only if that did not converge.
struct IrOut {
    double init_res, res, bnorm;
    int iters, outer;  // outer: FP64 residual replacements
};

IrOut run_mixed(fv_ctx* ctx, CgWork<float>& w, double tol, int maxit, int chunk, double delta,
                int max_outer, double atol, bool eager) {
    hipStream_t s = ctx->stream;
    PollPipe pp{static_cast<CgState*>(ctx_poll_host(ctx)), ctx};
    CgState h{};
    int launched = 0, rounds = 0;
    // iterations that ran in each round (no-op launches behind a stop do not
    // count): what the next solve replays
    std::vector<int> sched;
    int done = 0;  // iterations accounted in sched
    std::deque<int> inflight;  // sizes of the rounds queued after the polled state
    ConvRate cr;
    auto round = [&](int n) {
        w.chunk(s, n);
        w.replacement(s);  // only once the FP32 CG stopped (trigger, estimate, limit)
        FV_LAUNCH_CHECK();
        launched += n;
        ++rounds;
    };
    // Iterations the next round should launch, from the polled state h and the
    // rounds queued after it; -1 while no convergence rate is known, 0 when the
    // queued rounds are predicted to reach the tolerance.
    auto plan = [&](const CgState& h, double fin) -> int {
        double tgt = fin;
        if (h.bnorm > 0.0 && h.rrep > 0.0) tgt = std::max(tgt, delta * h.rrep / h.bnorm);
        int left = cr.remaining(h.res, tgt);  // to the stop of the segment in progress
        if (left < 0) return -1;
        for (int n : inflight) {
            if (left > n) {
                left -= n;
                continue;
            }
            // the segment stops in this round; the replacement at its end
            // starts the next one (or the solve ends at the tolerance)
            if (tgt <= fin) return 0;
            const double next = std::max(fin, tgt * delta);
            left = cr.remaining(tgt, next);
            tgt = next;
        }
        return left + kPredictMargin;
    };
    chunk = std::max(chunk, 1);
    if (!w.lastRounds.empty()) {
        // the rounds of the previous solve, at once (the whole sequence from the
        // matrix scaling on replayed as a hipGraph unless `eager`)
        std::vector<int> rr;
        for (int n : w.lastRounds) {
            if (launched >= maxit || (int)rr.size() >= max_outer) break;
            rr.push_back(std::min(n, maxit - launched));
            launched += rr.back();
        }
        rounds = (int)rr.size();
        auto seq = [&] {
            w.prep(s);
            w.start(s, tol, maxit, delta, atol);
            for (int n : rr) {
                w.chunk(s, n);
                w.replacement(s);
            }
        };
        if (eager) {
            seq();
        } else {
            std::vector<double> key{tol, (double)maxit, delta, atol};
            for (int n : rr) key.push_back(n);
            w.replay(s, key, seq);
        }
        FV_LAUNCH_CHECK();
        h = w.state(s);
        // the replayed rounds keep their sizes; the last one ends where the state is
        for (int i = 0; i + 1 < rounds; ++i) {
            sched.push_back(rr[i]);
            done += rr[i];
        }
        if (h.it > done) sched.push_back(h.it - done);
        done = h.it;
    } else {
        w.prep(s);
        w.start(s, tol, maxit, delta, atol);
        FV_LAUNCH_CHECK();
    }
    cr.observe(h.it, h.res);
    int ramp = 8;
    while (h.stop < 2 && h.it < maxit) {
        bool queued = false;
        if (pp.npend < 2 && launched < maxit && rounds < max_outer) {
            // the final target; the FP32 CG also stops at each drop trigger
            const double fin = std::max(tol, atol > 0.0 && h.bnorm > 0.0 ? atol / h.bnorm : 0.0);
            int need = h.it > 0 ? plan(h, fin) : -1;
            if (need < 0) {
                need = ramp;
                ramp = std::min(2 * ramp, chunk);
            }
            if (pp.npend == 0) need = std::max(need, 1);
            need = std::min(std::min(need, chunk), maxit - launched);
            if (need > 0) {
                round(need);
                pp.push(s, w.st.p);
                inflight.push_back(need);
                queued = true;
            }
        }
        if (pp.npend == 0) {
            if (launched >= maxit || rounds >= max_outer) break;
            continue;
        }
        if (queued && pp.npend < 2) continue;
        h = pp.pop();
        inflight.pop_front();
        cr.observe(h.it, h.res);
        if (h.it > done) sched.push_back(h.it - done);  // one poll per round, in order
        done = std::max(done, h.it);
    }
    while (pp.npend) h = pp.pop();
    if (rounds == 0) h = w.state(s);
    // the schedule of a converged solve is what the next solve replays
    if (h.stop == 2) w.lastRounds = sched;
    return IrOut{h.res0, h.res, h.bnorm, h.it, h.nrep};
}

}  // namespace

extern "C" {

int fv_solve_cg(fv_linsys* ls, void* d_x, const fv_solve_opts* o, fv_solver_stats* stats) {
    fv_solver_stats local{};
    int rc = guarded([&] {
        FV_REQUIRE(ls && ls->sp && d_x && o, "NULL argument");
        FV_REQUIRE(o->rel_tol >= 0.0 && o->max_iter >= 0 && o->abs_tol >= 0.0,
                   "bad solver options");
        fv_sparsity* sp = ls->sp;
        fv_ctx* ctx = sp->mesh->ctx;
        const int check = o->check_every > 0 ? o->check_every : 8;
        double* x = (double*)d_x;
        double bnorm = 0.0;
        if (o->mode == FV_SOLVE_F64) {
            CgWork<double>& w = work_of<double>(ls, false);
            w.bind(ctx, x);
            double tr = 0.0;
            CgState h =
                run_cg(ctx, w, o->rel_tol, o->max_iter, check, o->abs_tol, o->eager != 0, &tr);
            local.init_res = h.res0;
            local.final_res = h.res;
            bnorm = h.bnorm;
            local.iters = h.it;
            local.outer_iters = 1;
            local.true_res = tr;
        } else if (o->mode == FV_SOLVE_F32_IR) {
            CgWork<float>& w = work_of<float>(ls, true);
            w.bind(ctx, x);
            IrOut r = run_mixed(ctx, w, o->rel_tol, o->max_iter,
                                o->check_every > 0 ? o->check_every : 64,
                                o->inner_rel_tol > 0.0 ? o->inner_rel_tol : kIrDelta,
                                o->max_outer > 0 ? o->max_outer : (1 << 30), o->abs_tol,
                                o->eager != 0);
            local.init_res = r.init_res;
            local.final_res = r.res;
            bnorm = r.bnorm;
            local.true_res = r.res;
            local.iters = r.iters;
            local.outer_iters = r.outer;
        } else {
            FV_REQUIRE(false, "unknown solver mode");
        }
        // convergence is judged on the solver's residual: the CG recurrence for
        // FP64 (as Ginkgo's residual-norm criterion), the explicit FP64
        // residual b - A x for iterative refinement
        if (!(local.final_res <= o->rel_tol) &&
            !(o->abs_tol > 0.0 && local.final_res * bnorm <= o->abs_tol))
            throw Error{FV_ENOCONVERGE,
                        "CG stopped at relative residual " + std::to_string(local.final_res) +
                            " after " + std::to_string(local.iters) + " iterations" +
                            (local.final_res == local.final_res
                                 ? ""
                                 : " (not finite: a zero, non-finite or mixed-sign matrix "
                                   "diagonal, which the Jacobi scaling cannot take, or a "
                                   "non-finite matrix or right-hand side)")};
    });
    if (stats) *stats = local;
    return rc;
}

}  // extern "C"
