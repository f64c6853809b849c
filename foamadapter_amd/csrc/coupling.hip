// coupling.hip -- pressure-velocity coupling helpers of the PISO/SIMPLE loop
// (src/algorithms/pressureVelocityCoupling.cpp:14-211): computeRAU,
// computeRAUandHByA, updateFaceVelocity, updateVelocity, constrainHbyA; and the
// time-step control of the transient solvers: the Courant number and setDeltaT
// (src/auxiliary/setup.cpp:13-22).
//
// The reference scatters the off-diagonal products into HbyA with float
// atomics (pressureVelocityCoupling.cpp:109-110), so its result depends on the
// thread schedule.  Here each cell gathers its own products over the
// cell->face stencil in ascending face order -- the order of OpenFOAM's serial
// fvMatrix::H face loop -- and the result repeats bit for bit.  The momentum
// matrix is the scalar system (the reference reads component 0 of its Vec3
// coefficients, "all the diagonal coefficients are the same"), the right-hand
// side and the fields are Vec3 (xyz interleaved).
#include "implicit.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

using namespace fvhip;

// Block maxima of the Courant kernel for fv_courant_number_noalloc: one double
// per 256-cell block, so meshes of up to kCoScratchBlocks * 256 = 2^25 cells
// (1 MiB; every GPU the library runs on holds its own copy until the process
// ends).  Host code never names this array in a launch or a copy: it uses the
// device address hipGetSymbolAddress returns for the current device.
constexpr int kCoScratchBlocks = 1 << 17;
static __device__ __align__(256) double scratch[kCoScratchBlocks];

namespace {

constexpr int kBlock = 256;

// rAU = V / diag(A); boundary values extrapolated (rAU_b = rAU_P)
__global__ __launch_bounds__(kBlock) void k_rau(int N, int B, const int32_t* __restrict__ rowOffs,
                                                const int32_t* __restrict__ diag,
                                                const double* __restrict__ values,
                                                const double* __restrict__ V,
                                                const int32_t* __restrict__ faceCells,
                                                double* __restrict__ rau, double* __restrict__ raub) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) rau[i] = V[i] / values[rowOffs[i] + diag[i]];
    if (raub && i < B) {
        const int c = faceCells[i];
        raub[i] = V[c] / values[rowOffs[c] + diag[c]];
    }
}

// HbyA_c = (b_c - sum_f A(c, other) U_other) * rAU_c / V_c, the sum in ascending
// face order starting from zero (H is accumulated, then b added)
__global__ __launch_bounds__(kBlock) void k_hbya(int N, int F, const int32_t* __restrict__ off,
                                                 const int32_t* __restrict__ idx,
                                                 const int32_t* __restrict__ oth,
                                                 const int32_t* __restrict__ slot,
                                                 const int32_t* __restrict__ rowOffs,
                                                 const double* __restrict__ values,
                                                 const double* __restrict__ b3,
                                                 const double* __restrict__ U,
                                                 const double* __restrict__ rau,
                                                 const double* __restrict__ V,
                                                 double* __restrict__ hbya) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const int row = rowOffs[c];
    double h0 = 0.0, h1 = 0.0, h2 = 0.0;
    const int k0 = off[c], e1 = off[c + 1];
    // the first kE stencil entries: indices, then their coefficients and the
    // neighbours' velocities, loaded in two rounds before the sums, which run
    // in stencil order (boundary faces carry no off-diagonal)
    constexpr int kE = 8;
    int ie[kE], oe[kE], se[kE];
#pragma unroll
    for (int j = 0; j < kE; ++j)
        if (k0 + j < e1) {
            ie[j] = idx[k0 + j];
            oe[j] = oth[k0 + j];
            se[j] = slot[k0 + j];
        }
    double av[kE], u0[kE], u1[kE], u2[kE];
#pragma unroll
    for (int j = 0; j < kE; ++j)
        if (k0 + j < e1 && ie[j] < F) {
            const int64_t o = 3 * (int64_t)oe[j];
            av[j] = values[row + se[j]];
            u0[j] = U[o];
            u1[j] = U[o + 1];
            u2[j] = U[o + 2];
        }
#pragma unroll
    for (int j = 0; j < kE; ++j)
        if (k0 + j < e1 && ie[j] < F) {
            h0 -= av[j] * u0[j];
            h1 -= av[j] * u1[j];
            h2 -= av[j] * u2[j];
        }
    for (int k = k0 + kE; k < e1; ++k) {  // cells with more than kE faces
        if (idx[k] >= F) continue;
        const double a = values[row + slot[k]];
        const int64_t o = 3 * (int64_t)oth[k];
        h0 -= a * U[o];
        h1 -= a * U[o + 1];
        h2 -= a * U[o + 2];
    }
    const double s = rau[c] / V[c];
    const int64_t i = 3 * (int64_t)c;
    hbya[i] = (h0 + b3[i]) * s;
    hbya[i + 1] = (h1 + b3[i + 1]) * s;
    hbya[i + 2] = (h2 + b3[i + 2]) * s;
}

// boundary values of a cell field, extrapolated: out_b = in_P (ncomp components)
__global__ __launch_bounds__(kBlock) void k_extrapolate(int B, int ncomp,
                                                        const int32_t* __restrict__ faceCells,
                                                        const double* __restrict__ in,
                                                        double* __restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int64_t c = faceCells[b];
    for (int j = 0; j < ncomp; ++j) out[(int64_t)ncomp * b + j] = in[ncomp * c + j];
}

// phi_f = pred_f - (A(nbr,own) p_N - A(own,nbr) p_P); phi_b = pred_b - (bRhs - bMat p_P)
__global__ __launch_bounds__(kBlock) void k_face_velocity(
    int F, int B, const int32_t* __restrict__ own, const int32_t* __restrict__ nbr,
    const int32_t* __restrict__ rowOffs, const int32_t* __restrict__ ownerOffset,
    const int32_t* __restrict__ neighbourOffset, const double* __restrict__ values,
    const int32_t* __restrict__ faceCells, const double* __restrict__ bMat,
    const double* __restrict__ bRhs, const double* __restrict__ pred,
    const double* __restrict__ p, double* __restrict__ phi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < F) {
        const int o = own[i], n = nbr[i];
        const double upper = values[rowOffs[n] + neighbourOffset[i]];
        const double lower = values[rowOffs[o] + ownerOffset[i]];
        phi[i] = pred[i] - (upper * p[n] - lower * p[o]);
    } else if (i < F + B) {
        const int b = i - F;
        const double bflux = bRhs[b] - bMat[b] * p[faceCells[b]];
        phi[i] = pred[i] - bflux;
    }
}

// U = HbyA - rAU * grad(p)
__global__ __launch_bounds__(kBlock) void k_velocity(int N, const double* __restrict__ hbya,
                                                     const double* __restrict__ rau,
                                                     const double* __restrict__ gradp,
                                                     double* __restrict__ U) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const double r = rau[c];
    for (int j = 0; j < 3; ++j) U[3 * (int64_t)c + j] = hbya[3 * (int64_t)c + j] - r * gradp[3 * (int64_t)c + j];
}

// HbyA_b = U_b on faces whose velocity condition is fixedValue (not assignable)
__global__ __launch_bounds__(kBlock) void k_constrain(int B, const int8_t* __restrict__ kind,
                                                      const double* __restrict__ Ub,
                                                      double* __restrict__ hbyab) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || kind[b] != FV_BC_FIXED_VALUE) return;
    for (int j = 0; j < 3; ++j) hbyab[3 * (int64_t)b + j] = Ub[3 * (int64_t)b + j];
}

// Courant number (OpenFOAM CourantNo.H, NeoN computeCoNum used at
// examples/scalarAdvection/scalarAdvection.cpp:69): per cell the sum of |phi|
// over all its faces (internal and boundary, fvc::surfaceSum(mag(phi))) over V;
// each block writes the maximum of its cells (max is exact, so any order gives
// the same result)
__global__ __launch_bounds__(kBlock) void k_conum(int N, int F, const int32_t* __restrict__ off,
                                                  const int32_t* __restrict__ idx,
                                                  const double* __restrict__ phi,
                                                  const double* __restrict__ V,
                                                  double* __restrict__ blockMax) {
    __shared__ double sh[kBlock];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (c < N) {
        double s = 0.0;
        const int e1 = off[c + 1];
        for (int k = off[c]; k < e1; ++k) {
            const int e = idx[k];
            s += fabs(phi[e >= 0 ? e : ~e]);  // e >= F: boundary face e (phi is [F+B])
        }
        m = s / V[c];
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) blockMax[blockIdx.x] = sh[0];
}

// one block: the maximum of the n block maxima, times dt/2, into *co (device)
__global__ __launch_bounds__(kBlock) void k_conum_final(int n, const double* __restrict__ blockMax,
                                                        double dt, double* __restrict__ co) {
    __shared__ double sh[kBlock];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) m = fmax(m, blockMax[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *co = 0.5 * sh[0] * dt;
}

// Device address of `scratch` on each device, looked up once per device with
// that device current (a fixed table: the lookup allocates nothing).
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchDevices] = {};

double* scratch_on(int device) {
    FV_REQUIRE(device >= 0 && device < kScratchDevices, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        FV_HIP(hipGetDevice(&cur));
        if (cur != device) FV_HIP(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) FV_HIP(hipSetDevice(cur));
        if (e != hipSuccess || !p)
            throw Error{FV_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                        hipGetErrorString(e)};
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    return g_scratch_ptr[device];
}

void check_ls(const fv_linsys* ls) { FV_REQUIRE(ls && ls->sp && ls->sp->mesh, "NULL linear system"); }

}  // namespace

extern "C" {

int fv_compute_rau(fv_linsys* ls, void* d_rau, void* d_rau_b) {
    return guarded([&] {
        check_ls(ls);
        FV_REQUIRE(d_rau, "NULL output");
        const fv_sparsity* sp = ls->sp;
        const fv_mesh* m = sp->mesh;
        const int n = std::max(m->nCells, m->nB);
        k_rau<<<grid_for(n, kBlock), kBlock, 0, m->ctx->stream>>>(
            m->nCells, m->nB, sp->rowOffs.p, sp->diagOffset.p, ls->values.p, m->V.p,
            m->faceCells.p, (double*)d_rau, (double*)d_rau_b);
        FV_LAUNCH_CHECK();
    });
}

int fv_compute_hbya(fv_linsys* ls, const void* d_b3, const void* d_U, const void* d_rau,
                    void* d_hbya, void* d_hbya_b) {
    return guarded([&] {
        check_ls(ls);
        FV_REQUIRE(d_b3 && d_U && d_rau && d_hbya, "NULL argument");
        const fv_sparsity* sp = ls->sp;
        const fv_mesh* m = sp->mesh;
        hipStream_t s = m->ctx->stream;
        k_hbya<<<grid_for(m->nCells, kBlock), kBlock, 0, s>>>(
            m->nCells, m->nInt, m->cfOff.p, m->cfIdx.p, m->cfOther.p, sp->entrySlot.p,
            sp->rowOffs.p, ls->values.p, (const double*)d_b3, (const double*)d_U,
            (const double*)d_rau, m->V.p, (double*)d_hbya);
        if (d_hbya_b && m->nB)
            k_extrapolate<<<grid_for(m->nB, kBlock), kBlock, 0, s>>>(
                m->nB, 3, m->faceCells.p, (const double*)d_hbya, (double*)d_hbya_b);
        FV_LAUNCH_CHECK();
    });
}

int fv_update_face_velocity(fv_linsys* ls, const void* d_phi_pred, const void* d_p, void* d_phi) {
    return guarded([&] {
        check_ls(ls);
        FV_REQUIRE(d_phi_pred && d_p && d_phi, "NULL argument");
        const fv_sparsity* sp = ls->sp;
        const fv_mesh* m = sp->mesh;
        k_face_velocity<<<grid_for((int64_t)m->nInt + m->nB, kBlock), kBlock, 0, m->ctx->stream>>>(
            m->nInt, m->nB, m->own.p, m->nbr.p, sp->rowOffs.p, sp->ownerOffset.p,
            sp->neighbourOffset.p, ls->values.p, m->faceCells.p, ls->bMat.p, ls->bRhs.p,
            (const double*)d_phi_pred, (const double*)d_p, (double*)d_phi);
        FV_LAUNCH_CHECK();
    });
}

int fv_update_velocity(fv_ctx* ctx, const fv_mesh* m, const void* d_hbya, const void* d_rau,
                       const void* d_gradp, void* d_U) {
    return guarded([&] {
        FV_REQUIRE(ctx && m && m->ctx == ctx, "bad ctx/mesh");
        FV_REQUIRE(d_hbya && d_rau && d_gradp && d_U, "NULL argument");
        k_velocity<<<grid_for(m->nCells, kBlock), kBlock, 0, ctx->stream>>>(
            m->nCells, (const double*)d_hbya, (const double*)d_rau, (const double*)d_gradp,
            (double*)d_U);
        FV_LAUNCH_CHECK();
    });
}

int fv_constrain_hbya(const fv_bc* ubc, const void* d_Ub, void* d_hbya_b) {
    return guarded([&] {
        FV_REQUIRE(ubc && ubc->mesh && d_Ub && d_hbya_b, "NULL argument");
        const fv_mesh* m = ubc->mesh;
        if (!m->nB) return;
        k_constrain<<<grid_for(m->nB, kBlock), kBlock, 0, m->ctx->stream>>>(
            m->nB, ubc->faceKind.p, (const double*)d_Ub, (double*)d_hbya_b);
        FV_LAUNCH_CHECK();
    });
}

int fv_courant_number(fv_ctx* ctx, const fv_mesh* m, const void* d_phi, double dt, double* co) {
    return guarded([&] {
        FV_REQUIRE(ctx && m && m->ctx == ctx && d_phi && co, "bad arguments");
        const int G = grid_for(m->nCells, kBlock);
        DBuf<double> bm;
        bm.alloc(ctx, G);
        k_conum<<<G, kBlock, 0, ctx->stream>>>(m->nCells, m->nInt, m->cfOff.p, m->cfIdx.p,
                                                (const double*)d_phi, m->V.p, bm.p);
        FV_LAUNCH_CHECK();
        std::vector<double> h(G);
        bm.download(h.data());
        double mx = 0.0;
        for (double v : h) mx = std::max(mx, v);
        *co = 0.5 * mx * dt;
    });
}

int fv_courant_number_noalloc(fv_ctx* ctx, const fv_mesh* m, const void* d_phi, double dt,
                              double* d_co) {
    return guarded([&] {
        FV_REQUIRE(ctx && m && m->ctx == ctx && d_phi && d_co, "bad arguments");
        const int G = grid_for(m->nCells, kBlock);
        FV_REQUIRE(G <= kCoScratchBlocks,
                   "mesh too large for fv_courant_number_noalloc: at most 33554432 cells "
                   "(use fv_courant_number)");
        double* bm = scratch_on(ctx->device);
        // every block writes its slot, so nothing in the scratch needs clearing
        k_conum<<<G, kBlock, 0, ctx->stream>>>(m->nCells, m->nInt, m->cfOff.p, m->cfIdx.p,
                                                (const double*)d_phi, m->V.p, bm);
        k_conum_final<<<1, kBlock, 0, ctx->stream>>>(G, bm, dt, (double*)d_co);
        FV_LAUNCH_CHECK();
    });
}

int fv_set_delta_t(double dt, double co, double max_co, double max_dt, double* new_dt) {
    return guarded([&] {
        FV_REQUIRE(new_dt && dt > 0.0 && max_co > 0.0 && max_dt > 0.0, "bad arguments");
        const double small = 1e-15;  // Foam::SMALL
        const double maxFact = max_co / (co + small);
        const double fact = std::min(std::min(maxFact, 1.0 + 0.1 * maxFact), 1.2);
        *new_dt = std::min(fact * dt, max_dt);
    });
}

int fv_linsys_boundary_coeffs(fv_linsys* ls, void** d_bmat, void** d_brhs) {
    return guarded([&] {
        check_ls(ls);
        FV_REQUIRE(d_bmat && d_brhs, "NULL argument");
        *d_bmat = ls->bMat.p;
        *d_brhs = ls->bRhs.p;
    });
}

}  // extern "C"
