// implicit.hpp -- sparsity pattern and linear system handles.
#pragma once

#include "mesh.hpp"

#include <memory>

// la::SparsityPattern (expression.hpp:35; test_sparsityPattern.cpp:36-62):
// CSR of the face graph, columns ascending per row; row c is
// [lower neighbours | c | upper neighbours] in the cell's stencil order.
struct fv_sparsity {
    fv_mesh* mesh = nullptr;
    int32_t nRows = 0;
    int64_t nnz = 0;
    fvhip::DBuf<int32_t> rowOffs, colIdxs, diagOffset, ownerOffset, neighbourOffset;
    // per stencil entry k (cell->face stencil of the mesh): CSR slot of the entry's
    // coefficient in the cell's row (-1 for boundary entries)
    fvhip::DBuf<int32_t> entrySlot;
    // sliced-ELL copy of the off-diagonal structure for the solvers' SpMVs,
    // whose scaled matrices have a unit diagonal (slices of 64 rows, one
    // wavefront each; entry k of row 64s+l at sellOff[s] + 64k + l): column and
    // source CSR index (-1 for padding)
    static constexpr int kSlice = 64;
    int64_t sellTotal = 0;
    fvhip::DBuf<int64_t> sellOff;
    fvhip::DBuf<int32_t> sellCol, sellSrc;
};

// Solver workspace kept with a linear system between solves (one per precision).
struct fv_solver_cache {
    virtual ~fv_solver_cache() = default;
};

// la::LinearSystem<scalar, localIdx>: values[nnz], rhs[N] (FP64), with the
// convention A.x - b = (operator applied to x) * V.
struct fv_linsys {
    fv_sparsity* sp = nullptr;
    fvhip::DBuf<double> values, rhs;
    // boundary coefficients of the assembled laplacian terms, per boundary face:
    // the face flux is bRhs - bMat * x_P (NeoN's la::BoundaryCoefficients,
    // read by updateFaceVelocity, pressureVelocityCoupling.cpp:180-192)
    fvhip::DBuf<double> bMat, bRhs;
    std::unique_ptr<fv_solver_cache> cg[3];  // FP64 CG, FP32 CG in FP64 refinement, BiCGStab
};
