// cgstate.hpp -- device-resident CG scalars and the deterministic, fence-free
// block reductions shared by the single-system solver and the batched engine.
#pragma once

#include "implicit.hpp"

namespace fvhip {

constexpr int kCgBlock = 256;

// Default residual-drop trigger of the mixed solver (fv_solve_opts::inner_rel_tol
// = 0): the FP32 CG runs until its residual fell by this factor since the last
// FP64 residual replacement.  Measured against 1e-3: 128^3 Poisson 838 vs 865
// iterations (43.5 vs 44.5 ms), 160^3 1111 vs 1140, the 32^3 heat step 0.34 vs
// 0.50 ms (two replacements instead of three), batched heat +6 %; 1e-4 and
// 1e-6 took more iterations at 128^3 (986, 1091).
constexpr double kIrDelta = 1e-5;

// Device-resident CG scalars of one system.  Each field has one writer per
// kernel (the thread 0 of the system's first tile) and is never read by the
// kernel that writes it, so no fence or atomic is needed: the kernel boundary
// orders a write before every read of it.
//   stop     decision of the latest direction kernel: 0 running, 1 stopped
//            (recurrence estimate at the tolerance, iteration limit, NaN, or the
//            residual-drop trigger of the mixed solver), 2 converged in FP64,
//            3 failed (non-finite FP64 residual); read by SpMV, update and the
//            replacement kernels
//   done     copy of stop taken by the update kernel (and by the replacement
//            kernel, for the direction kernel after it); read by the direction
//            kernels
//   it       completed iterations (the update kernel advances it)
//   rho      r.z of the latest residual, written by the direction kernels; the
//            update divides it by p.q for alpha
//   rhoPrev  rho as of the latest update (written by the update kernel); the
//            direction kernels divide the new r.z by it for beta
//   pbuf     which of the two search-direction buffers holds the p of the
//            latest update (written by the update kernel)
//   nrep     FP64 residual replacements so far (advanced by the replacement
//            kernel; read by the host)
struct CgState {
    double rho, rhoPrev, alpha, bnorm, res, res0, rrep, delta;
    int32_t it, done, maxit, stop;
    double tol, atol;  // relative and absolute (0: off) residual targets
    int32_t pbuf, nrep;
};

// Fixed-shape block reduction of NV values at once: wavefront shuffles, then
// thread 0 adds the wavefronts' sums in order.  sh holds NV * (blockDim/64)
// doubles.  The totals are valid in thread 0.
template <int NV>
__device__ inline void block_sum_n(double (&v)[NV], double* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
    for (int j = 0; j < NV; ++j)
        for (int o = 32; o > 0; o >>= 1) v[j] += __shfl_down(v[j], o, 64);
    if (lane == 0)
        for (int j = 0; j < NV; ++j) sh[j * nw + wid] = v[j];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int j = 0; j < NV; ++j) {
            double r = 0.0;
            for (int w = 0; w < nw; ++w) r += sh[j * nw + w];
            v[j] = r;
        }
    __syncthreads();
}

// LDS the reductions of a kCgBlock-thread kernel need (up to 2 values each)
constexpr int kCgShared = 2 * (kCgBlock / 64);

__device__ inline double block_sum(double v, double* sh) {
    double a[1] = {v};
    block_sum_n<1>(a, sh);
    return a[0];  // valid in thread 0
}

// The block's NV partial sums go to slot `slot` of `partials` (NV doubles per
// slot).  Plain stores: the consumer is a later kernel.
template <int NV>
__device__ inline void write_slot(const double (&v)[NV], double* partials, int slot, double* sh) {
    double s[NV];
    for (int j = 0; j < NV; ++j) s[j] = v[j];
    block_sum_n<NV>(s, sh);
    if (threadIdx.x == 0)
        for (int j = 0; j < NV; ++j) partials[(int64_t)slot * NV + j] = s[j];
}

// Every block of the consuming kernel sums the producer's slots [base,
// base+count) itself, in one fixed order (strided per thread, then the
// fixed-shape block tree), so all blocks obtain bitwise the same totals and the
// result depends only on the system's own slot count.  The NV values of a slot
// are loaded together.
template <int NV>
__device__ inline void sum_slots(const double* partials, int base, int count, double* sh,
                                 double (&tot)[NV]) {
    __shared__ double bc[NV];
    double a[NV];
    for (int j = 0; j < NV; ++j) a[j] = 0.0;
    for (int b = threadIdx.x; b < count; b += blockDim.x)
        for (int j = 0; j < NV; ++j) a[j] += partials[(int64_t)(base + b) * NV + j];
    block_sum_n<NV>(a, sh);
    if (threadIdx.x == 0)
        for (int j = 0; j < NV; ++j) bc[j] = a[j];
    __syncthreads();
    for (int j = 0; j < NV; ++j) tot[j] = bc[j];
    __syncthreads();
}

}  // namespace fvhip
