// mesh.hpp -- device-resident finite-volume mesh (NeoN UnstructuredMesh +
// BoundaryMesh arrays, meshAdapter.cpp:59-136) plus the gather structures the
// deterministic kernels use.
#pragma once

#include "common.hpp"

// Layout in HBM (all device arrays, int32 indices):
//   cells:   V[N], C[3N]
//   internal faces (OpenFOAM order): own[F], nbr[F], Sf[3F], Cf[3F], magSf[F],
//            w[F], dc[F], nodc[F]
//   boundary faces (flattened patches): faceCells[B], Sfb[3B], Cfb[3B], magSfb[B],
//            nfb[3B], deltab[3B], wb[B], dcb[B], nodcb[B]
//   cell->face stencil in ascending face order (CellToFaceStencil):
//            cfOff[N+1], cfIdx[2F+B]: entry e >= 0 && e < F: internal face e with
//            the cell as owner; e < 0: internal face ~e with the cell as
//            neighbour; e >= F: boundary face e-F.  cfOther[2F+B] holds the cell
//            on the other side of an internal face (the cell itself for a
//            boundary face) so the kernels read it as a stream.
//   FP32 mirror of every floating array, built on first FP32 use.
struct MeshF32 {
    fvhip::DBuf<float> V, Sf, magSf, w, nodc, Sfb, magSfb, dcb;
    bool ready = false;
};

// Per-workgroup face lists of the LDS-staged gather for one workgroup size
// (`cells` consecutive cells per workgroup, block b = cells [b*cells, ...)):
//   extOff[nBlocks+1]; extFace, extOwn, extCell (cell - c0): the block's lower
//     faces whose owner lies in an earlier block, sorted by (cell, face)
//   bndOff[nBlocks+1]; bndFace (boundary face b), bndCell: the block's boundary
//     faces sorted by (cell, b)
// Each record array holds one zero record past its last block, so a kernel can
// read a record at any index up to extOff[nBlocks] / bndOff[nBlocks] without a
// branch.
struct GatherLists {
    int cells = 0;  // 0: not built
    int64_t nExt = 0;  // external lower-face records (without the zero record)
    fvhip::DBuf<int32_t> extOff, extFace, extOwn, bndOff, bndFace;
    fvhip::DBuf<uint16_t> extCell, bndCell;
    // geometry of the external lower faces packed in record order (w, magSf,
    // nonOrthDeltaCoeffs, Sf xyz), so phase 1 reads it coalesced instead of
    // gathering it by face index; FP32 copies built on first FP32 use
    fvhip::DBuf<double> exW, exMagSf, exNodc, exSf;
    fvhip::DBuf<float> exW32, exMagSf32, exNodc32, exSf32;
    bool ready32 = false;
    // workgroups whose faces do not fit the kernel's LDS lists (more owned,
    // external lower or boundary faces than it holds, or a cell with more
    // than gather_max_low lower faces from inside the workgroup): they take the
    // per-cell path, same bits, slower
    int64_t nBlocks = 0, nOverflow = 0;
};

// LDS capacities of the staged gather for a CELLS-cell workgroup (faces owned,
// external lower faces, boundary faces, and inner lower faces per cell).  A
// Z-order brick of 8x8x4 / 8x8x8 / 16x8x8 cells has 128 / 192 / 320 faces on
// its three lower sides, and as many boundary faces at a corner of the domain.
// The external list holds one record per thread of the workgroup (256 / 512 /
// 512), room for the ragged bricks of a domain whose sides are not powers of
// two (a 17 x 23 x 55 box: a median of 140 such faces per 256-cell brick); the
// boundary list holds CELLS/2, so corner and edge bricks stay on the fast path.
constexpr int gather_max_up(int cells) { return cells * 3 + cells / 4; }
constexpr int gather_max_ext(int cells) { return cells < 512 ? cells : 512; }
constexpr int gather_max_bnd(int cells) { return cells / 2; }
// Inner lower faces per cell: 8 (a hex cell has 3; polyhedra up to 8), 4 for
// 1024-cell workgroups, whose LDS then fits four workgroups per CU (a mesh with
// more such cells moves to 512-cell workgroups, see launch_gather).
constexpr int gather_max_low(int cells) { return cells == 1024 ? 4 : 8; }

struct fv_mesh {
    fv_ctx* ctx = nullptr;
    int32_t nCells = 0, nInt = 0, nB = 0, nPatches = 0;
    std::vector<int32_t> patchOff;  // host copy [nPatches+1]
    fvhip::DBuf<double> V, C;
    fvhip::DBuf<int32_t> own, nbr;
    fvhip::DBuf<double> Sf, Cf, magSf, w, dc, nodc;
    fvhip::DBuf<int32_t> faceCells;
    fvhip::DBuf<double> Sfb, Cfb, magSfb, nfb, deltab, wb, dcb, nodcb;
    fvhip::DBuf<int32_t> patchOffD;
    fvhip::DBuf<int32_t> cfOff, cfIdx, cfOther;
    // Split stencil for meshes in upper-triangular face order (owner < neighbour,
    // sorted by owner then neighbour), used by the LDS-staged face kernels:
    //   ownStart[N+1]: cell c owns internal faces [ownStart[c], ownStart[c+1])
    //   lowOff[N+1], lowFace[F], lowOwn[F]: faces where c is the neighbour,
    //     ascending, with their owner
    //   bOff[N+1], bFace[B]: boundary faces of c, ascending
    bool upperTri = false;
    int gatherMode = 0;  // cell-result kernels: 0 default, 1 stencil walk, 2/3/4 staged 256/512/1024
    int gatherRun = 8;   // staged gather: consecutive bricks per XCD (1: natural order)
    fvhip::DBuf<int32_t> ownStart, lowOff, lowFace, lowOwn, bOff, bFace;
    GatherLists glists[3];  // workgroups of 256, 512, 1024 cells (built on first use)
    MeshF32 f32;
};

struct fv_bc {
    fv_mesh* mesh = nullptr;
    std::vector<int32_t> kinds;        // per patch
    fvhip::DBuf<int8_t> faceKind;      // per boundary face
};

namespace fvhip {
// Build cfOff/cfIdx/cfOther on the host from owner/neighbour/faceCells and upload.
void build_stencil(fv_mesh* m, const int32_t* h_own, const int32_t* h_nbr,
                   const int32_t* h_faceCells);
// FP32 copies of the geometry (lazy).
void ensure_f32(fv_mesh* m);
// The staged gather's per-workgroup lists for `cells` cells per workgroup
// (with the FP32 packed geometry too when f32).
const GatherLists& gather_lists(fv_mesh* m, int cells, bool f32 = false);
// Cells per workgroup of the LDS-staged gather for the cell-result operators,
// 0 when the per-cell stencil walk runs instead (faces not in upper-triangular
// order, or forced by fv_mesh_set_gather); kDefaultGatherCells stands for the
// per-operator default of the staged gather (its Op::kDefaultCells).
constexpr int kDefaultGatherCells = 512;
inline int gather_cells(const fv_mesh* m) {
    if (!m->upperTri) return 0;
    switch (m->gatherMode) {
        case 1: return 0;
        case 2: return 256;
        case 3: return 512;
        case 4: return 1024;
        default: return kDefaultGatherCells;
    }
}
}  // namespace fvhip
