// cgstate.hpp -- device-resident CG scalars and the deterministic reductions
// shared by the single-system solver and the batched engine.
#pragma once

#include "implicit.hpp"

namespace fvhip {

constexpr int kCgBlock = 256;

// Device-resident CG scalars.  rz[it & 1] holds r_it.z_it, so the direction
// update reads beta = rz[it&1] / rz[(it-1)&1] without a second writer; a
// residual replacement overwrites rz[it&1] with the replaced residual's value.
// done: 0 running, 1 stopped (converged estimate, iteration limit, NaN, or the
// residual-drop trigger of the mixed solver), 2 converged in FP64.
struct CgState {
    double rz[2], alpha, bnorm, res, res0, rrep, delta;
    int32_t it, done, maxit, pad;
    double tol;
    uint32_t ticket;
    uint32_t pad2;
};

__device__ inline double block_sum(double v, double* sh) {
    // fixed-shape tree: wavefront shuffle then across the block's wavefronts
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
    }
    __syncthreads();
    return r;  // valid in thread 0
}

__device__ inline double load_coherent(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Each of `count` participating blocks writes its NV partials to slot `slot`
// (slots [base, base+count)); the block that finishes last (integer ticket in
// st) sums the slots in slot order and returns true with the totals in thread 0.
template <int NV>
__device__ bool publish_slots(const double (&v)[NV], double* partials, CgState* st, double* sh,
                              int base, int count, int slot, double (&tot)[NV]) {
    __shared__ bool last;
    double s[NV];
    for (int j = 0; j < NV; ++j) s[j] = block_sum(v[j], sh);
    if (threadIdx.x == 0) {
        for (int j = 0; j < NV; ++j) partials[(int64_t)slot * NV + j] = s[j];
        __threadfence();
        const uint32_t t = atomicAdd(&st->ticket, 1u);
        last = (t == (uint32_t)count - 1);
    }
    __syncthreads();
    if (!last) return false;
    __threadfence();
    for (int j = 0; j < NV; ++j) {
        double a = 0.0;
        for (int b = threadIdx.x; b < count; b += blockDim.x)
            a += load_coherent(&partials[(int64_t)(base + b) * NV + j]);
        a = block_sum(a, sh);
        if (threadIdx.x == 0) tot[j] = a;
    }
    if (threadIdx.x == 0) st->ticket = 0;
    return true;
}

// The single-system grid: every block of the grid participates.
template <int NV>
__device__ bool publish(const double (&v)[NV], double* partials, CgState* st, double* sh,
                        double (&tot)[NV]) {
    return publish_slots<NV>(v, partials, st, sh, 0, (int)gridDim.x, (int)blockIdx.x, tot);
}

}  // namespace fvhip
