// batch.hpp -- the batched CG engine's host interface, shared by the C ABI
// (fv_batch_*) and the request server.
#pragma once

#include "cgstate.hpp"

#include <vector>

struct BMember {
    fv_linsys* ls = nullptr;
    double* x = nullptr;
    double* xold = nullptr;
    int id = -1;
    int N = 0;
    int slot = -1;           // index of the member's CgState (stable while it is in the batch)
    bool needStart = true;   // matrix/right-hand side (re)loaded, CG must (re)start
    double tol = 0.0;        // FP64 relative residual target (0: the batch's)
    int maxit = 0;           // CG iteration limit per solve (0: the batch's)
    fvhip::DBuf<uint8_t> vec;   // d, r, p, q, g, p2 (S) + r64 (double)
    fvhip::DBuf<uint8_t> sval;  // scaled off-diagonal sliced-ELL values in S
    fv_solver_stats stats{};
    void* user = nullptr;    // owner's tag (the server's request)
};

struct fv_batch {
    fv_ctx* ctx = nullptr;
    int prec = FV_SOLVE_F32_IR;  // inner precision of the whole batch
    int tileRows = 1024;         // rows per tile (multiple of 64)
    int chunk = 16;              // CG iterations per FP64 residual replacement
    double tol = 1e-8, delta = fvhip::kIrDelta;
    double atol = 0.0;           // absolute residual target of fv_batch_solve (0: off)
    int maxit = 1000;
    int nextId = 0;
    std::vector<BMember*> members;
    std::vector<int> freeSlots;
    int slotCap = 0;
    // device tables, rewritten when membership changes; their capacities only
    // grow, and the captured chunk graph stays valid until they do
    fvhip::DBuf<uint8_t> probs;
    fvhip::DBuf<uint8_t> tiles;
    fvhip::DBuf<uint8_t> startTiles, startParams;
    fvhip::DBuf<uint8_t> gridDev, startGridDev;  // launch descriptions (CgGrid)
    fvhip::DBuf<fvhip::CgState> states;
    fvhip::DBuf<double> partials;
    int nTiles = 0, nStartTiles = 0;
    int tileCap = 0, probCap = 0;
    bool dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    int64_t graphBuilds = 0, tableBuilds = 0;  // counters (server statistics)
    bool useGraph = true;
    // GPU time of the chunks (graph launch or kernel sequence) and of the
    // batched starts, from HIP events on the batch's stream around each
    int64_t chunks = 0, starts = 0;
    double chunkMs = 0.0, startMs = 0.0;
    hipEvent_t ev[4] = {};
    ~fv_batch();
};

namespace fvhip {

BMember* batch_add(fv_batch* b, fv_linsys* ls, double* x, double* xold, void* user);
void batch_remove(fv_batch* b, BMember* m);
// Mark a member whose system was re-assembled (and x possibly changed): its
// CG restarts from x at the next step.
void batch_restart(fv_batch* b, BMember* m);
// One step: (re)start members that need it, then one chunk of CG iterations
// for every member followed by the FP64 residual replacement.  Returns the
// host copy of every member's state (in member order).
void batch_step(fv_batch* b, std::vector<CgState>& states);
// Launch-only part of a step (no host synchronisation), for callers that
// overlap the next step's host work.
void batch_copy_old(fv_batch* b);  // x_old = x of every member with a copy target

}  // namespace fvhip
