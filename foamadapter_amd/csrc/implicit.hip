// implicit.hip -- la::SparsityPattern, la::LinearSystem and the implicit
// operator assembly (ddt Euler, laplacian Gauss linear uncorrected, div Gauss
// linear/upwind, source, explicit source, SetReference, A.x - b).
//
// Every assembly kernel is a cell gather over the mesh's cell->face stencil in
// ascending face order: a cell writes the off-diagonal coefficients of its own
// row (each CSR slot has exactly one writer) and accumulates its diagonal and
// right-hand side in the serial face order, so no atomics are used and the
// results repeat bit for bit.
#include "implicit.hpp"

#include <algorithm>

using namespace fvhip;

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_ddt(int N, const int32_t* __restrict__ rowOffs,
                                                const int32_t* __restrict__ diag,
                                                const double* __restrict__ V,
                                                const double* __restrict__ Told, double rdt,
                                                double coeff, double* __restrict__ values,
                                                double* __restrict__ rhs) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    // EulerDdtScheme::fvmDdt: diag = rDeltaT*V, source = rDeltaT*T_old*V
    const double v = V[c];
    values[rowOffs[c] + diag[c]] += coeff * (rdt * v);
    rhs[c] += coeff * ((rdt * Told[c]) * v);
}

__global__ __launch_bounds__(kBlock) void k_source(int N, const int32_t* __restrict__ rowOffs,
                                                   const int32_t* __restrict__ diag,
                                                   const double* __restrict__ V,
                                                   const double* __restrict__ cf, double coeff,
                                                   double* __restrict__ values) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    values[rowOffs[c] + diag[c]] += coeff * (cf[c] * V[c]);
}

__global__ __launch_bounds__(kBlock) void k_explicit(int N, const double* __restrict__ V,
                                                     const double* __restrict__ E, double coeff,
                                                     double* __restrict__ rhs) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    rhs[c] -= coeff * (E[c] * V[c]);
}

struct Stencil {
    const int32_t *off, *idx, *slot;
};

// The first kAsmE stencil entries of a cell have their loads issued in two
// rounds (entries, then the face data of all of them) before any arithmetic;
// the arithmetic and its order are those of the loop over the stencil, which
// the entries past kAsmE (cells with more faces) take.
constexpr int kAsmE = 8;

__global__ __launch_bounds__(kBlock) void k_laplacian(
    int N, int F, Stencil st, const int32_t* __restrict__ rowOffs,
    const int32_t* __restrict__ diag, const double* __restrict__ magSf,
    const double* __restrict__ nodc, const double* __restrict__ magSfb,
    const double* __restrict__ dcb, const int8_t* __restrict__ kind,
    const double* __restrict__ gamma, const double* __restrict__ Tb,
    const double* __restrict__ Tg, double coeff, double* __restrict__ values,
    double* __restrict__ rhs, double* __restrict__ bMat, double* __restrict__ bRhs) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const int row = rowOffs[c], dg = diag[c];
    const int k0 = st.off[c], k1 = st.off[c + 1];
    int ev[kAsmE], sl[kAsmE];
#pragma unroll
    for (int j = 0; j < kAsmE; ++j)
        if (k0 + j < k1) {
            ev[j] = st.idx[k0 + j];
            sl[j] = st.slot[k0 + j];
        }
    double ga[kAsmE], ma[kAsmE], da[kAsmE], tb[kAsmE], tg[kAsmE];
    int kd[kAsmE];
#pragma unroll
    for (int j = 0; j < kAsmE; ++j)
        if (k0 + j < k1) {
            const int e = ev[j];
            if (e >= F) {
                const int b = e - F;
                ga[j] = gamma[e];
                ma[j] = magSfb[b];
                kd[j] = kind[b];
                da[j] = dcb[b];
                tb[j] = Tb[b];
                tg[j] = Tg ? Tg[b] : 0.0;
            } else {
                const int f = e >= 0 ? e : ~e;
                ga[j] = gamma[f];
                ma[j] = magSf[f];
                da[j] = nodc[f];
            }
        }
    double dsum = 0.0, bsum = 0.0;
    auto entry = [&](int e, int slot, double g0, double m0, double d0, int k, double tb0,
                     double tg0) {
        if (e >= F) {
            // boundary coefficients: the cell is the only writer of its faces
            const int b = e - F;
            const double gm = g0 * m0;
            if (k == FV_BC_FIXED_GRADIENT) {
                const double g = gm * tg0;
                bsum -= g;
                bMat[b] += 0.0;
                bRhs[b] += coeff * g;
            } else {
                const double ic = gm * d0;
                const double it = ic * tb0;
                dsum -= ic;
                bsum -= it;
                bMat[b] += coeff * ic;
                bRhs[b] += coeff * it;
            }
        } else {
            const double up = d0 * (g0 * m0);
            values[row + slot] += coeff * up;
            dsum -= up;
        }
    };
#pragma unroll
    for (int j = 0; j < kAsmE; ++j)
        if (k0 + j < k1) {
            const bool bnd = ev[j] >= F;
            entry(ev[j], sl[j], ga[j], ma[j], da[j], bnd ? kd[j] : 0, bnd ? tb[j] : 0.0,
                  bnd ? tg[j] : 0.0);
        }
    for (int k = k0 + kAsmE; k < k1; ++k) {
        const int e = st.idx[k];
        if (e >= F) {
            const int b = e - F;
            entry(e, 0, gamma[e], magSfb[b], dcb[b], kind[b], Tb[b], Tg ? Tg[b] : 0.0);
        } else {
            const int f = e >= 0 ? e : ~e;
            entry(e, st.slot[k], gamma[f], magSf[f], nodc[f], 0, 0.0, 0.0);
        }
    }
    values[row + dg] += coeff * dsum;
    rhs[c] += coeff * bsum;
}

// Inputs of the fused heat assembly of one system (the kernel arguments of
// fv_assemble_ddt_laplacian, a table entry of its batched form).
struct HeatDesc {
    int N, F;
    Stencil st;
    const int32_t *rowOffs, *diag;
    const double *V, *Told;
    double rdt;
    const double *magSf, *nodc, *magSfb, *dcb;
    const int8_t* kind;
    const double *gamma, *Tb, *Tg;
    double coeff;
    double *values, *rhs, *bMat, *bRhs;
};

// Fused Euler ddt + coeff * laplacian of cell c, overwriting its row and rhs:
// the same arithmetic as zero -> ddt -> laplacian in that order (0 + x is
// exact), in one pass over the cell's stencil.
__device__ inline void ddt_laplacian_cell(const HeatDesc& D, int c) {
    // The first kSE stencil entries have their loads issued in two rounds
    // (entries, then the face data of all of them) before any arithmetic; the
    // arithmetic and its order are those of the loop over the stencil.
    constexpr int kSE = 8;
    const int row = D.rowOffs[c], dg = D.diag[c];
    const double v = D.V[c], told = D.Told[c];
    const int k0 = D.st.off[c], k1 = D.st.off[c + 1];
    int ev[kSE], sl[kSE];
#pragma unroll
    for (int j = 0; j < kSE; ++j)
        if (k0 + j < k1) {
            ev[j] = D.st.idx[k0 + j];
            sl[j] = D.st.slot[k0 + j];
        }
    double ga[kSE], ma[kSE], da[kSE], tb[kSE], tg[kSE];
    int kd[kSE];
#pragma unroll
    for (int j = 0; j < kSE; ++j)
        if (k0 + j < k1) {
            const int e = ev[j];
            if (e >= D.F) {
                const int b = e - D.F;
                ga[j] = D.gamma[e];
                ma[j] = D.magSfb[b];
                kd[j] = D.kind[b];
                da[j] = D.dcb[b];
                tb[j] = D.Tb[b];
                tg[j] = D.Tg ? D.Tg[b] : 0.0;
            } else {
                const int f = e >= 0 ? e : ~e;
                ga[j] = D.gamma[f];
                ma[j] = D.magSf[f];
                da[j] = D.nodc[f];
            }
        }
    double dsum = 0.0, bsum = 0.0;
    auto entry = [&](int e, int slot, double g0, double m0, double d0, int k, double tb0,
                     double tg0) {
        if (e >= D.F) {
            const int b = e - D.F;
            const double gm = g0 * m0;
            if (k == FV_BC_FIXED_GRADIENT) {
                const double g = gm * tg0;
                bsum -= g;
                D.bMat[b] = 0.0;
                D.bRhs[b] = D.coeff * g;
            } else {
                const double ic = gm * d0;
                const double it = ic * tb0;
                dsum -= ic;
                bsum -= it;
                D.bMat[b] = D.coeff * ic;
                D.bRhs[b] = D.coeff * it;
            }
        } else {
            const double up = d0 * (g0 * m0);
            D.values[row + slot] = D.coeff * up;
            dsum -= up;
        }
    };
#pragma unroll
    for (int j = 0; j < kSE; ++j)
        if (k0 + j < k1) {
            const bool bnd = ev[j] >= D.F;
            entry(ev[j], sl[j], ga[j], ma[j], da[j], bnd ? kd[j] : 0, bnd ? tb[j] : 0.0,
                  bnd ? tg[j] : 0.0);
        }
    for (int k = k0 + kSE; k < k1; ++k) {  // cells with more than kSE faces
        const int e = D.st.idx[k];
        if (e >= D.F) {
            const int b = e - D.F;
            entry(e, 0, D.gamma[e], D.magSfb[b], D.dcb[b], D.kind[b], D.Tb[b], D.Tg ? D.Tg[b] : 0.0);
        } else {
            const int f = e >= 0 ? e : ~e;
            entry(e, D.st.slot[k], D.gamma[f], D.magSf[f], D.nodc[f], 0, 0.0, 0.0);
        }
    }
    D.values[row + dg] = (D.rdt * v) + D.coeff * dsum;
    D.rhs[c] = ((D.rdt * told) * v) + D.coeff * bsum;
}

__global__ __launch_bounds__(kBlock) void k_ddt_laplacian(HeatDesc D) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < D.N) ddt_laplacian_cell(D, c);
}

// Batched form: block b assembles cells [c0, c0 + kBlock) of system sys.
struct HeatBlock {
    int32_t sys, c0;
};

__global__ __launch_bounds__(kBlock) void k_ddt_laplacian_many(const HeatDesc* __restrict__ desc,
                                                               const HeatBlock* __restrict__ blocks) {
    const HeatBlock hb = blocks[blockIdx.x];
    const HeatDesc D = desc[hb.sys];
    const int c = hb.c0 + (int)threadIdx.x;
    if (c < D.N) ddt_laplacian_cell(D, c);
}

__global__ __launch_bounds__(kBlock) void k_div(
    int N, int F, int scheme, Stencil st, const int32_t* __restrict__ rowOffs,
    const int32_t* __restrict__ diag, const double* __restrict__ w,
    const double* __restrict__ dcb, const int8_t* __restrict__ kind,
    const double* __restrict__ phi, const double* __restrict__ Tb,
    const double* __restrict__ Tg, double coeff, double* __restrict__ values,
    double* __restrict__ rhs) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const int row = rowOffs[c], dg = diag[c];
    const int k0 = st.off[c], k1 = st.off[c + 1];
    // loads in two rounds for the first kAsmE entries, as in k_laplacian
    int ev[kAsmE], sl[kAsmE];
#pragma unroll
    for (int j = 0; j < kAsmE; ++j)
        if (k0 + j < k1) {
            ev[j] = st.idx[k0 + j];
            sl[j] = st.slot[k0 + j];
        }
    double pa[kAsmE], wa[kAsmE], tb[kAsmE], tg[kAsmE], da[kAsmE];
    int kd[kAsmE];
#pragma unroll
    for (int j = 0; j < kAsmE; ++j)
        if (k0 + j < k1) {
            const int e = ev[j];
            if (e >= F) {
                const int b = e - F;
                pa[j] = phi[e];
                kd[j] = kind[b];
                tb[j] = Tb[b];
                tg[j] = Tg ? Tg[b] : 0.0;
                da[j] = dcb[b];
            } else {
                const int f = e >= 0 ? e : ~e;
                pa[j] = phi[f];
                wa[j] = scheme == 0 ? w[f] : 0.0;
            }
        }
    double dsum = 0.0, bsum = 0.0;
    auto entry = [&](int e, int slot, double pf, double wf, int k, double tb0, double tg0,
                     double d0) {
        if (e >= F) {
            if (k == FV_BC_FIXED_GRADIENT) {
                dsum += pf;
                bsum -= pf * (tg0 / d0);
            } else {
                bsum -= pf * tb0;
            }
        } else {
            const double lam = scheme == 0 ? wf : (pf >= 0.0 ? 1.0 : 0.0);
            const double lower = -lam * pf;
            const double upper = lower + pf;
            // gaussConvectionScheme::fvmDiv: owner row gets upper in column N,
            // neighbour row gets lower in column P; negSumDiag
            if (e >= 0) {
                values[row + slot] += coeff * upper;
                dsum -= lower;
            } else {
                values[row + slot] += coeff * lower;
                dsum -= upper;
            }
        }
    };
#pragma unroll
    for (int j = 0; j < kAsmE; ++j)
        if (k0 + j < k1) {
            const bool bnd = ev[j] >= F;
            entry(ev[j], sl[j], pa[j], bnd ? 0.0 : wa[j], bnd ? kd[j] : 0, bnd ? tb[j] : 0.0,
                  bnd ? tg[j] : 0.0, bnd ? da[j] : 1.0);
        }
    for (int k = k0 + kAsmE; k < k1; ++k) {
        const int e = st.idx[k];
        if (e >= F) {
            const int b = e - F;
            entry(e, 0, phi[e], 0.0, kind[b], Tb[b], Tg ? Tg[b] : 0.0, dcb[b]);
        } else {
            const int f = e >= 0 ? e : ~e;
            entry(e, st.slot[k], phi[f], scheme == 0 ? w[f] : 0.0, 0, 0.0, 0.0, 1.0);
        }
    }
    values[row + dg] += coeff * dsum;
    rhs[c] += coeff * bsum;
}

__global__ void k_set_reference(const int32_t* __restrict__ rowOffs,
                                const int32_t* __restrict__ diag, int cell, double value,
                                double* __restrict__ values, double* __restrict__ rhs) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const int d = rowOffs[cell] + diag[cell];
        const double dv = values[d];
        rhs[cell] += dv * value;
        values[d] += dv;
    }
}

// res = A.x - b, one thread per row, row entries summed in column order
__global__ __launch_bounds__(kBlock) void k_residual(int N, const int32_t* __restrict__ rowOffs,
                                                     const int32_t* __restrict__ col,
                                                     const double* __restrict__ values,
                                                     const double* __restrict__ b,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ res) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    double s = 0.0;
    const int e = rowOffs[r + 1];
    for (int k = rowOffs[r]; k < e; ++k) s += values[k] * x[col[k]];
    res[r] = s - b[r];
}

fv_mesh* mesh_of(const fv_linsys* ls) { return ls->sp->mesh; }

void check_ls(const fv_linsys* ls) { FV_REQUIRE(ls && ls->sp && ls->sp->mesh, "NULL linear system"); }

void check_bc(const fv_linsys* ls, const fv_bc* bc) {
    FV_REQUIRE(bc && bc->mesh == mesh_of(ls), "boundary conditions belong to another mesh");
}

HeatDesc heat_desc(fv_linsys* ls, const double* Told, double dt, const double* gamma,
                   const fv_bc* bc, const double* Tb, const double* Tg, double coeff) {
    check_ls(ls);
    check_bc(ls, bc);
    fv_mesh* m = mesh_of(ls);
    FV_REQUIRE(Told && gamma && (Tb || m->nB == 0) && dt > 0.0, "bad arguments");
    fv_sparsity* sp = ls->sp;
    return HeatDesc{m->nCells, m->nInt, Stencil{m->cfOff.p, m->cfIdx.p, sp->entrySlot.p},
                    sp->rowOffs.p, sp->diagOffset.p, m->V.p, Told, 1.0 / dt, m->magSf.p,
                    m->nodc.p, m->magSfb.p, m->dcb.p, bc->faceKind.p, gamma, Tb, Tg, coeff,
                    ls->values.p, ls->rhs.p, ls->bMat.p, ls->bRhs.p};
}

}  // namespace

extern "C" {

int fv_sparsity_create(fv_mesh* m, fv_sparsity** out) {
    fv_sparsity* sp = nullptr;
    int rc = guarded([&] {
        FV_REQUIRE(m && out, "NULL argument");
        const int64_t N = m->nCells, F = m->nInt;
        std::vector<int32_t> off(N + 1), idx(m->cfIdx.n), oth(m->cfOther.n);
        m->cfOff.download(off.data());
        m->cfIdx.download(idx.data());
        m->cfOther.download(oth.data());
        std::vector<int32_t> rowOffs(N + 1, 0), diag(N), oo(F), no(F), slot(idx.size(), -1);
        for (int64_t c = 0; c < N; ++c) {
            int32_t nint = 0;
            for (int32_t k = off[c]; k < off[c + 1]; ++k)
                if (idx[k] < F) ++nint;
            rowOffs[c + 1] = rowOffs[c] + nint + 1;
        }
        std::vector<int32_t> col(rowOffs[N]);
        for (int64_t c = 0; c < N; ++c) {
            int32_t j = 0, row = rowOffs[c];
            int phase = 0;  // 0: lower entries, 1: upper entries, 2: boundary entries
            bool diagDone = false;
            for (int32_t k = off[c]; k < off[c + 1]; ++k) {
                const int32_t e = idx[k];
                const int ph = e < 0 ? 0 : (e < F ? 1 : 2);
                FV_REQUIRE(ph >= phase,
                           "mesh faces are not in upper-triangular order (owner < neighbour, "
                           "sorted by owner then neighbour)");
                phase = ph;
                if (ph >= 1 && !diagDone) {
                    diag[c] = j;
                    col[row + j++] = (int32_t)c;
                    diagDone = true;
                }
                if (ph == 0) {
                    no[~e] = j;
                    slot[k] = j;
                    col[row + j++] = oth[k];
                } else if (ph == 1) {
                    oo[e] = j;
                    slot[k] = j;
                    col[row + j++] = oth[k];
                }
            }
            if (!diagDone) {
                diag[c] = j;
                col[row + j++] = (int32_t)c;
            }
            for (int32_t t = 1; t < j; ++t)
                FV_REQUIRE(col[row + t] > col[row + t - 1], "duplicate or unsorted columns in a row");
        }
        sp = new fv_sparsity();
        sp->mesh = m;
        sp->nRows = (int32_t)N;
        sp->nnz = rowOffs[N];
        fv_ctx* ctx = m->ctx;
        sp->rowOffs.upload(ctx, rowOffs.data(), rowOffs.size());
        sp->colIdxs.upload(ctx, col.data(), col.size());
        sp->diagOffset.upload(ctx, diag.data(), diag.size());
        sp->ownerOffset.upload(ctx, oo.data(), oo.size());
        sp->neighbourOffset.upload(ctx, no.data(), no.size());
        sp->entrySlot.upload(ctx, slot.data(), slot.size());
        // sliced ELL (64-row slices, column-major inside a slice) of the rows
        // without their diagonal entry: the solvers' scaled matrices have a unit
        // diagonal, which their SpMVs add themselves
        const int64_t C = fv_sparsity::kSlice, nS = (N + C - 1) / C;
        auto sliced = [&](fvhip::DBuf<int64_t>& offOut, fvhip::DBuf<int32_t>& colOut,
                          fvhip::DBuf<int32_t>& srcOut) -> int64_t {
            std::vector<int64_t> soff(nS + 1, 0);
            for (int64_t s = 0; s < nS; ++s) {
                int32_t wmax = 0;
                for (int64_t r = s * C; r < std::min<int64_t>(N, (s + 1) * C); ++r)
                    wmax = std::max(wmax, rowOffs[r + 1] - rowOffs[r] - 1);
                soff[s + 1] = soff[s] + C * wmax;
            }
            std::vector<int32_t> scol(soff[nS]), ssrc(soff[nS]);
            for (int64_t s = 0; s < nS; ++s) {
                const int64_t wdt = (soff[s + 1] - soff[s]) / C;
                for (int64_t l = 0; l < C; ++l) {
                    const int64_t r = s * C + l;
                    int64_t k = 0;
                    if (r < N)
                        for (int32_t j = 0; j < rowOffs[r + 1] - rowOffs[r]; ++j) {
                            if (j == diag[r]) continue;
                            const int64_t pos = soff[s] + k++ * C + l;
                            scol[pos] = col[rowOffs[r] + j];
                            ssrc[pos] = rowOffs[r] + j;
                        }
                    for (; k < wdt; ++k) {
                        const int64_t pos = soff[s] + k * C + l;
                        scol[pos] = r < N ? (int32_t)r : 0;
                        ssrc[pos] = -1;
                    }
                }
            }
            offOut.upload(ctx, soff.data(), soff.size());
            colOut.upload(ctx, scol.data(), scol.size());
            srcOut.upload(ctx, ssrc.data(), ssrc.size());
            FV_HIP(hipStreamSynchronize(ctx->stream));  // the host vectors end here
            return soff[nS];
        };
        sp->sellTotal = sliced(sp->sellOff, sp->sellCol, sp->sellSrc);
        FV_HIP(hipStreamSynchronize(ctx->stream));
        *out = sp;
    });
    if (rc != FV_OK) delete sp;
    return rc;
}

int fv_sparsity_destroy(fv_sparsity* sp) {
    return guarded([&] { delete sp; });
}

int fv_sparsity_sizes(const fv_sparsity* sp, int64_t* s) {
    return guarded([&] {
        FV_REQUIRE(sp && s, "NULL argument");
        s[0] = sp->nRows;
        s[1] = sp->nnz;
    });
}

int fv_sparsity_get(const fv_sparsity* sp, int which, int32_t* h) {
    return guarded([&] {
        FV_REQUIRE(sp && h, "NULL argument");
        switch (which) {
            case 0: sp->rowOffs.download(h); break;
            case 1: sp->colIdxs.download(h); break;
            case 2: sp->diagOffset.download(h); break;
            case 3: sp->ownerOffset.download(h); break;
            case 4: sp->neighbourOffset.download(h); break;
            default: FV_REQUIRE(false, "unknown sparsity array");
        }
    });
}

int fv_linsys_create(fv_sparsity* sp, fv_linsys** out) {
    fv_linsys* ls = nullptr;
    int rc = guarded([&] {
        FV_REQUIRE(sp && out, "NULL argument");
        fv_ctx* ctx = sp->mesh->ctx;
        ls = new fv_linsys();
        ls->sp = sp;
        ls->values.alloc(ctx, sp->nnz);
        ls->rhs.alloc(ctx, sp->nRows);
        ls->bMat.alloc(ctx, std::max<int32_t>(1, sp->mesh->nB));
        ls->bRhs.alloc(ctx, std::max<int32_t>(1, sp->mesh->nB));
        FV_HIP(hipMemsetAsync(ls->values.p, 0, sp->nnz * sizeof(double), ctx->stream));
        FV_HIP(hipMemsetAsync(ls->rhs.p, 0, sp->nRows * sizeof(double), ctx->stream));
        FV_HIP(hipMemsetAsync(ls->bMat.p, 0, ls->bMat.n * sizeof(double), ctx->stream));
        FV_HIP(hipMemsetAsync(ls->bRhs.p, 0, ls->bRhs.n * sizeof(double), ctx->stream));
        *out = ls;
    });
    if (rc != FV_OK) delete ls;
    return rc;
}

int fv_linsys_destroy(fv_linsys* ls) {
    return guarded([&] { delete ls; });
}

void* fv_linsys_values(fv_linsys* ls) { return ls ? (void*)ls->values.p : nullptr; }
void* fv_linsys_rhs(fv_linsys* ls) { return ls ? (void*)ls->rhs.p : nullptr; }

int fv_linsys_zero(fv_linsys* ls) {
    return guarded([&] {
        check_ls(ls);
        hipStream_t s = mesh_of(ls)->ctx->stream;
        FV_HIP(hipMemsetAsync(ls->values.p, 0, ls->values.n * sizeof(double), s));
        FV_HIP(hipMemsetAsync(ls->rhs.p, 0, ls->rhs.n * sizeof(double), s));
        FV_HIP(hipMemsetAsync(ls->bMat.p, 0, ls->bMat.n * sizeof(double), s));
        FV_HIP(hipMemsetAsync(ls->bRhs.p, 0, ls->bRhs.n * sizeof(double), s));
    });
}

int fv_assemble_ddt_euler(fv_linsys* ls, const void* Told, double dt, double coeff) {
    return guarded([&] {
        check_ls(ls);
        FV_REQUIRE(Told && dt > 0.0, "bad ddt arguments");
        fv_mesh* m = mesh_of(ls);
        fv_sparsity* sp = ls->sp;
        k_ddt<<<grid_for(m->nCells, kBlock), kBlock, 0, m->ctx->stream>>>(
            m->nCells, sp->rowOffs.p, sp->diagOffset.p, m->V.p, (const double*)Told, 1.0 / dt,
            coeff, ls->values.p, ls->rhs.p);
        FV_LAUNCH_CHECK();
    });
}

int fv_assemble_source(fv_linsys* ls, const void* cf, double coeff) {
    return guarded([&] {
        check_ls(ls);
        FV_REQUIRE(cf, "NULL coefficient field");
        fv_mesh* m = mesh_of(ls);
        fv_sparsity* sp = ls->sp;
        k_source<<<grid_for(m->nCells, kBlock), kBlock, 0, m->ctx->stream>>>(
            m->nCells, sp->rowOffs.p, sp->diagOffset.p, m->V.p, (const double*)cf, coeff,
            ls->values.p);
        FV_LAUNCH_CHECK();
    });
}

int fv_assemble_explicit(fv_linsys* ls, const void* E, double coeff) {
    return guarded([&] {
        check_ls(ls);
        FV_REQUIRE(E, "NULL explicit field");
        fv_mesh* m = mesh_of(ls);
        k_explicit<<<grid_for(m->nCells, kBlock), kBlock, 0, m->ctx->stream>>>(
            m->nCells, m->V.p, (const double*)E, coeff, ls->rhs.p);
        FV_LAUNCH_CHECK();
    });
}

int fv_assemble_laplacian(fv_linsys* ls, const void* gamma, const fv_bc* bc, const void* Tb,
                          const void* Tg, double coeff) {
    return guarded([&] {
        check_ls(ls);
        check_bc(ls, bc);
        fv_mesh* m = mesh_of(ls);
        FV_REQUIRE(gamma && (Tb || m->nB == 0), "NULL field");
        fv_sparsity* sp = ls->sp;
        Stencil st{m->cfOff.p, m->cfIdx.p, sp->entrySlot.p};
        k_laplacian<<<grid_for(m->nCells, kBlock), kBlock, 0, m->ctx->stream>>>(
            m->nCells, m->nInt, st, sp->rowOffs.p, sp->diagOffset.p, m->magSf.p, m->nodc.p,
            m->magSfb.p, m->dcb.p, bc->faceKind.p, (const double*)gamma, (const double*)Tb,
            (const double*)Tg, coeff, ls->values.p, ls->rhs.p, ls->bMat.p, ls->bRhs.p);
        FV_LAUNCH_CHECK();
    });
}

int fv_assemble_ddt_laplacian(fv_linsys* ls, const void* Told, double dt, const void* gamma,
                              const fv_bc* bc, const void* Tb, const void* Tg, double coeff) {
    return guarded([&] {
        const HeatDesc d = heat_desc(ls, (const double*)Told, dt, (const double*)gamma, bc,
                                     (const double*)Tb, (const double*)Tg, coeff);
        k_ddt_laplacian<<<grid_for(d.N, kBlock), kBlock, 0, mesh_of(ls)->ctx->stream>>>(d);
        FV_LAUNCH_CHECK();
    });
}

int fv_assemble_ddt_laplacian_many(int n, fv_linsys* const* ls, const void* const* Told,
                                   const double* dt, const void* const* gamma,
                                   const fv_bc* const* bc, const void* const* Tb,
                                   const void* const* Tg, double coeff) {
    return guarded([&] {
        FV_REQUIRE(n >= 0 && (n == 0 || (ls && Told && dt && gamma && bc && Tb)), "NULL argument");
        if (n == 0) return;
        std::vector<HeatDesc> desc(n);
        std::vector<HeatBlock> blocks;
        fv_ctx* ctx = nullptr;
        for (int i = 0; i < n; ++i) {
            desc[i] = heat_desc(ls[i], (const double*)Told[i], dt[i], (const double*)gamma[i],
                                bc[i], (const double*)Tb[i], Tg ? (const double*)Tg[i] : nullptr,
                                coeff);
            fv_ctx* c = mesh_of(ls[i])->ctx;
            FV_REQUIRE(!ctx || c == ctx, "systems of one call must share a context");
            ctx = c;
            for (int c0 = 0; c0 < desc[i].N; c0 += kBlock) blocks.push_back(HeatBlock{i, c0});
        }
        if (blocks.empty()) return;
        DBuf<HeatDesc> dd;
        DBuf<HeatBlock> db;
        dd.upload(ctx, desc.data(), desc.size());
        db.upload(ctx, blocks.data(), blocks.size());
        k_ddt_laplacian_many<<<(unsigned)blocks.size(), kBlock, 0, ctx->stream>>>(dd.p, db.p);
        FV_LAUNCH_CHECK();
        // the tables' blocks return to the context's pool: later work on the
        // same stream is ordered after this launch
    });
}

int fv_assemble_div(fv_linsys* ls, fv_scheme scheme, const void* phi, const fv_bc* bc,
                    const void* Tb, const void* Tg, double coeff) {
    return guarded([&] {
        check_ls(ls);
        check_bc(ls, bc);
        FV_REQUIRE(scheme == FV_LINEAR || scheme == FV_UPWIND, "unknown scheme");
        fv_mesh* m = mesh_of(ls);
        FV_REQUIRE(phi && (Tb || m->nB == 0), "NULL field");
        fv_sparsity* sp = ls->sp;
        Stencil st{m->cfOff.p, m->cfIdx.p, sp->entrySlot.p};
        k_div<<<grid_for(m->nCells, kBlock), kBlock, 0, m->ctx->stream>>>(
            m->nCells, m->nInt, scheme, st, sp->rowOffs.p, sp->diagOffset.p, m->w.p, m->dcb.p,
            bc->faceKind.p, (const double*)phi, (const double*)Tb, (const double*)Tg, coeff,
            ls->values.p, ls->rhs.p);
        FV_LAUNCH_CHECK();
    });
}

int fv_set_reference(fv_linsys* ls, int cell, double value) {
    return guarded([&] {
        check_ls(ls);
        fv_mesh* m = mesh_of(ls);
        FV_REQUIRE(cell >= 0 && cell < m->nCells, "reference cell out of range");
        k_set_reference<<<1, 64, 0, m->ctx->stream>>>(ls->sp->rowOffs.p, ls->sp->diagOffset.p,
                                                      cell, value, ls->values.p, ls->rhs.p);
        FV_LAUNCH_CHECK();
    });
}

int fv_residual(fv_linsys* ls, const void* x, void* res) {
    return guarded([&] {
        check_ls(ls);
        FV_REQUIRE(x && res && x != res, "NULL or aliased vector");
        fv_mesh* m = mesh_of(ls);
        k_residual<<<grid_for(m->nCells, kBlock), kBlock, 0, m->ctx->stream>>>(
            m->nCells, ls->sp->rowOffs.p, ls->sp->colIdxs.p, ls->values.p, ls->rhs.p,
            (const double*)x, (double*)res);
        FV_LAUNCH_CHECK();
    });
}

}  // extern "C"
