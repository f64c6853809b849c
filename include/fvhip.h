/*
 * fvhip.h -- C ABI of the MI355X-native finite-volume hot path (libfvhip.so).
 *
 * This is the drop-in boundary that replaces the NeoN calls FoamAdapter makes
 * on its finite-volume path (NeoN's fvcc operators, la::SparsityPattern /
 * LinearSystem, dsl::solve with Ginkgo CG + Jacobi).  Each entry point names
 * the reference interface it replaces as file:line of the reference
 * repository.  Conventions:
 *   - every call returns an int status (FV_OK = 0); fv_last_error() gives a
 *     thread-local message for the last failure on the calling thread;
 *   - arrays named d_* are DEVICE pointers (allocated with fv_device_alloc or
 *     by any HIP allocator on the context's device); arrays named h_* are host
 *     pointers; the caller owns both, library handles own their own buffers;
 *   - calls are ordered on the context's stream and asynchronous unless they
 *     return host data (fv_sync waits);
 *   - indices are int32 (Foam::label with label=32, readers.hpp:24); floating
 *     arrays are double (FV_F64) or float (FV_F32) as the prec argument says;
 *   - vectors (Sf, C, grad) are stored as interleaved xyz triples, as NeoN's
 *     Vec3 (src/auxiliary/convert.cpp:9).
 *   - a context is used by one host thread at a time.
 */
#ifndef FVHIP_H
#define FVHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    FV_OK = 0,
    FV_EINVAL = 1,         /* bad argument or size                          */
    FV_ENOMEM_BUDGET = 2,  /* allocation would exceed the context HBM budget */
    FV_ENOCONVERGE = 3,    /* solver hit max iterations above tolerance      */
    FV_EDEVICE = 4,        /* HIP runtime error                              */
    FV_EBUSY = 5,          /* server queue full                              */
    FV_ECANCELLED = 6,     /* request cancelled before it ran                */
    FV_ENOTREADY = 7       /* request not finished yet (poll)                */
};

typedef enum { FV_F64 = 0, FV_F32 = 1 } fv_prec;
/* interpolation schemes: "linear", "upwind" (fvSchemes divSchemes Gauss <scheme>) */
typedef enum { FV_LINEAR = 0, FV_UPWIND = 1 } fv_scheme;
/* boundary kinds after FoamAdapter's mapping (readers.hpp:43-95):
 * zeroGradient -> fixedGradient 0, extrapolatedCalculated -> calculated,
 * noSlip -> fixedValue 0; empty patches carry no faces. */
typedef enum { FV_BC_FIXED_VALUE = 0, FV_BC_FIXED_GRADIENT = 1, FV_BC_CALCULATED = 2 } fv_bc_kind;

typedef struct fv_ctx fv_ctx;
typedef struct fv_mesh fv_mesh;
typedef struct fv_bc fv_bc;

/* ---- context, memory ---------------------------------------------------- */

/* Last error message of the calling thread ("" if none). */
const char* fv_last_error(void);
/* Library version string. */
const char* fv_version(void);

/* Executor selection: replaces createExecutor("GPU") (src/auxiliary/setup.cpp:51-80).
 * hbm_budget_bytes = 0 means unlimited; allocations through the context are
 * accounted against it and fail with FV_ENOMEM_BUDGET beyond it. */
int fv_ctx_create(int device, size_t hbm_budget_bytes, fv_ctx** out);
int fv_ctx_destroy(fv_ctx* ctx);
/* Wait for all work on the context's stream (Kokkos::fence, bench_explicitOperators.cpp:116-119). */
int fv_sync(fv_ctx* ctx);
/* The context's HIP stream (hipStream_t), for callers that order their own work. */
void* fv_ctx_stream(fv_ctx* ctx);
/* Bytes currently accounted against the budget. */
size_t fv_ctx_bytes_in_use(fv_ctx* ctx);
/* Timing on the context's stream (the hipEvent counterpart of the Catch2
 * BENCHMARK + Kokkos::fence pattern, bench_explicitOperators.cpp:104-121):
 * record event slot (0..63), then read the elapsed milliseconds between two
 * recorded slots (waits for the later one). */
int fv_event_record(fv_ctx* ctx, int slot);
int fv_event_elapsed(fv_ctx* ctx, int slot_begin, int slot_end, float* h_ms);

/* Device memory owned by the caller (NeoN::Vector storage on the GPU executor). */
int fv_device_alloc(fv_ctx* ctx, size_t bytes, void** d_out);
int fv_device_free(fv_ctx* ctx, void* d_ptr);
/* Copies; h2d/d2h are synchronous with respect to the host buffer
 * (fromFoamField readers.hpp:17-29 copies host->device, copyToHost writers.hpp:28). */
int fv_memcpy_h2d(fv_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int fv_memcpy_d2h(fv_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);
int fv_memcpy_d2d(fv_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
/* Fill n elements of d_ptr with a constant (NeoN::fill). */
int fv_fill(fv_ctx* ctx, fv_prec prec, void* d_ptr, int64_t n, double value);
/* Convert between precisions on the device (n elements). */
int fv_convert(fv_ctx* ctx, fv_prec to, void* d_dst, fv_prec from, const void* d_src, int64_t n);
/* One component of an interleaved field of n items with ncomp components:
 * mode 0 copies it out (dst[i] = src[i*ncomp + comp]), mode 1 in (dst[i*ncomp +
 * comp] = src[i]).  A Vec3 equation (PDESolver<Vec3>, the momentum predictor of
 * examples/neoIcoFoam/neoIcoFoam.cpp:94-96) is assembled and solved component
 * by component on the shared matrix through it. */
int fv_component(fv_ctx* ctx, fv_prec prec, int64_t n, int ncomp, int comp, int mode,
                 const void* d_src, void* d_dst);
/* Counter-based uniform random numbers in [lo,hi): element i gets
 * lo + (hi-lo)*u(seed, stream, i) with u = splitmix64-hash(...) >> 11 * 2^-53,
 * identical on host and device, so CPU checkers can regenerate any input. */
int fv_fill_uniform(fv_ctx* ctx, fv_prec prec, void* d_ptr, int64_t n, double lo, double hi,
                    uint64_t seed, uint64_t stream);

/* Streaming reference kernels (n doubles, 16-byte aligned arrays, n even):
 * mode 0 copy d_a = d_b, mode 1 triad d_a = d_b + s * d_c, mode 2 read (d_a
 * receives one partial sum of d_b per 2048 doubles).  No counterpart in
 * the reference: the measured HBM ceiling of a kernel that only streams, which
 * the benchmark reports beside the face operators. */
int fv_stream(fv_ctx* ctx, int mode, int64_t n, const double* d_b, const double* d_c, double* d_a,
              double s);

/* ---- mesh ---------------------------------------------------------------- */

/* OpenFOAM polyMesh primitives (host arrays): replaces the fvMesh that
 * MeshAdapter wraps (include/FoamAdapter/datastructures/meshAdapter.hpp:30-89).
 * Faces: internal faces first (owner < neighbour, ordered by owner then
 * neighbour), then boundary faces patch by patch. */
typedef struct {
    int32_t n_points, n_faces, n_internal_faces, n_cells, n_patches;
    const double* h_points;        /* [3*n_points] */
    const int32_t* h_face_offsets; /* [n_faces+1] into h_face_points */
    const int32_t* h_face_points;  /* [face_offsets[n_faces]] */
    const int32_t* h_owner;        /* [n_faces] */
    const int32_t* h_neighbour;    /* [n_internal_faces] */
    const int32_t* h_patch_start;  /* [n_patches] first face of each patch */
    const int32_t* h_patch_size;   /* [n_patches] */
    const int32_t* h_patch_empty;  /* [n_patches] 1 for 'empty' patches (no fv faces) */
} fv_polymesh_desc;

/* Geometry arrays as NeoN::UnstructuredMesh/BoundaryMesh hold them
 * (src/datastructures/meshAdapter.cpp:59-136); host arrays.  Internal-face
 * arrays are [n_internal_faces], boundary arrays [n_boundary_faces] flattened
 * patch by patch over non-empty patches (flatBCField, meshAdapter.cpp:12-29). */
typedef struct {
    int32_t n_cells, n_internal_faces, n_boundary_faces, n_patches;
    const double* h_V;            /* [N] cell volumes */
    const double* h_C;            /* [3N] cell centres */
    const int32_t* h_owner;       /* [F] */
    const int32_t* h_neighbour;   /* [F] */
    const double* h_Sf;           /* [3F] */
    const double* h_Cf;           /* [3F] */
    const double* h_magSf;        /* [F] */
    const double* h_weights;      /* [F] linear interpolation weights */
    const double* h_delta_coeffs; /* [F] 1/|C_N-C_P| */
    const double* h_nonorth_delta_coeffs; /* [F] 1/max(n.d, 0.05|d|) */
    const int32_t* h_face_cells;  /* [B] */
    const double* h_Sf_b;         /* [3B] */
    const double* h_Cf_b;         /* [3B] */
    const double* h_magSf_b;      /* [B] */
    const double* h_nf_b;         /* [3B] */
    const double* h_delta_b;      /* [3B] */
    const double* h_weights_b;    /* [B] */
    const double* h_delta_coeffs_b; /* [B] */
    const double* h_nonorth_delta_coeffs_b; /* [B] */
    const int32_t* h_patch_offsets; /* [n_patches+1] offsets into the boundary arrays */
} fv_mesh_desc;

/* Structured hex block meshes in polyMesh form (the 3DCube blockMesh cases,
 * benchmarks/benchmarkSuite/templates/3DCube/system/blockMeshDict;
 * tutorials/heatTranfer3D/system/blockMeshDict).  sizes6 = {n_points, n_cells,
 * n_internal_faces, n_boundary_faces, n_patches(=6), n_face_points}.
 * Patches in order x0, x1, y0, y1, z0, z1.  Interior points are moved by
 * (2u-1)*jitter*h per axis (jitter in [0,0.5)); perm_seed != 0 renumbers the
 * cells with a seeded Fisher-Yates shuffle; faces come out in OpenFOAM
 * upper-triangular order either way.  Host arrays sized from sizes6. */
int fv_hexmesh_sizes(int32_t nx, int32_t ny, int32_t nz, int64_t* h_sizes6);
int fv_hexmesh_generate(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz,
                        double jitter, uint64_t jitter_seed, uint64_t perm_seed,
                        double* h_points, int32_t* h_face_offsets, int32_t* h_face_points,
                        int32_t* h_owner, int32_t* h_neighbour, int32_t* h_patch_start,
                        int32_t* h_patch_size);

/* Locality renumbering of a polyMesh (host), the counterpart of OpenFOAM's
 * renumberMesh that a case runs before the solver: method 0 reverse
 * Cuthill-McKee (a breadth-first sweep of the face graph), method 1 Z-order
 * (Morton code of the cell centres, each axis quantised in units of the
 * mesh's typical cell spacing along it: aligned runs of cells are compact
 * bricks, whole bricks of cells on a structured mesh of any size, which the
 * face kernels exploit); internal
 * faces rewritten in upper-triangular order of the new numbering (a face whose
 * owner/neighbour swap has its point list reversed), boundary faces unchanged.
 * Outputs are sized like the input (points and patches are unchanged):
 * cell_new_to_old[n_cells]; face_new_to_old[n_faces] = old face, or ~old when
 * the face was flipped (oriented face fields change sign). */
int fv_polymesh_renumber(const fv_polymesh_desc* in, int method, int32_t* h_face_offsets_out,
                         int32_t* h_face_points_out, int32_t* h_owner_out,
                         int32_t* h_neighbour_out, int32_t* h_cell_new_to_old,
                         int32_t* h_face_new_to_old);
/* Device permutation of n items of ncomp values (field transfer between a
 * caller's numbering and a renumbered mesh):
 *   mode 0 (gather):  dst[i] = s * src[j]
 *   mode 1 (scatter): dst[j] = s * src[i]
 * with e = d_map[i], j = e >= 0 ? e : ~e, s = e >= 0 ? +1 : -1 when `oriented`
 * is non-zero, else +1. */
int fv_permute(fv_ctx* ctx, fv_prec prec, int64_t n, int ncomp, int mode, int oriented,
               const int32_t* d_map, const void* d_src, void* d_dst);

/* Upload NeoN-layout geometry (readOpenFOAMMesh, meshAdapter.cpp:59-136). */
int fv_mesh_create(fv_ctx* ctx, const fv_mesh_desc* desc, fv_mesh** out);
/* Build a mesh from polyMesh primitives, computing face centres/areas, cell
 * centres/volumes, weights and delta coefficients on the GPU (the fvMesh
 * geometry that MeshAdapter reads, meshAdapter.cpp:59-136 and GeometryScheme,
 * test/test_unstructuredMesh.cpp:164-208). */
int fv_mesh_create_poly(fv_ctx* ctx, const fv_polymesh_desc* desc, fv_mesh** out);
int fv_mesh_destroy(fv_mesh* mesh);
/* sizes[4] = {n_cells, n_internal_faces, n_boundary_faces, n_patches} */
int fv_mesh_sizes(const fv_mesh* mesh, int32_t* h_sizes4);

/* Geometry array identifiers for fv_mesh_get (download, host arrays sized as in fv_mesh_desc). */
typedef enum {
    FV_G_V = 0, FV_G_C, FV_G_OWNER, FV_G_NEIGHBOUR, FV_G_SF, FV_G_CF, FV_G_MAGSF, FV_G_WEIGHTS,
    FV_G_DELTA_COEFFS, FV_G_NONORTH_DELTA_COEFFS, FV_G_FACE_CELLS, FV_G_SF_B, FV_G_CF_B,
    FV_G_MAGSF_B, FV_G_NF_B, FV_G_DELTA_B, FV_G_WEIGHTS_B, FV_G_DELTA_COEFFS_B,
    FV_G_NONORTH_DELTA_COEFFS_B, FV_G_PATCH_OFFSETS,
    /* cell->face stencil, ascending face order (fvcc::CellToFaceStencil,
     * test/test_stencils.cpp:34-54): offsets [N+1], faces [2F+B] (internal faces
     * then boundary faces numbered F+b) */
    FV_G_CELL_FACE_OFFSETS, FV_G_CELL_FACES
} fv_geom;
int fv_mesh_get(const fv_mesh* mesh, fv_geom which, void* h_out);

/* Kernel choice for the cell-result operators (grad, div, laplacian): mode 0
 * picks the LDS-staged face-block gather (default workgroup size) when the
 * faces are in upper-triangular order, mode 1 forces the per-cell stencil walk,
 * modes 2/3/4 the staged gather with 256/512/1024 cells per workgroup (same
 * arithmetic and order: bit-identical results).  fv_mesh_get_gather reports the
 * kernel in use: 1 stencil walk, 2/3/4 staged with 256/512/1024 cells.  No
 * counterpart in the reference: a tuning knob. */
int fv_mesh_set_gather(fv_mesh* mesh, int mode);
int fv_mesh_get_gather(const fv_mesh* mesh, int* mode);
/* Order in which the staged gather's workgroups take the mesh's bricks: each
 * of the 8 XCDs takes runs of `run` consecutive bricks (default 8; 1 is the
 * natural order, brick b on XCD b % 8).  Results are the same bits for any
 * run.  No counterpart in the reference: a tuning knob. */
int fv_mesh_set_gather_run(fv_mesh* mesh, int run);
/* How the staged gather with `cells` cells per workgroup (256, 512 or 1024)
 * runs on this mesh (its lists are built if needed): h_counts2 = {workgroups,
 * workgroups whose faces exceed its LDS lists and take the per-cell path --
 * same results, slower}.  No counterpart in the reference: a diagnostic. */
int fv_mesh_gather_stats(fv_mesh* mesh, int cells, int64_t* h_counts2);

/* Boundary conditions of a field: one kind per patch
 * (fvcc::VolumeBoundary list built by readVolBoundaryConditions, readers.hpp:31-108). */
int fv_bc_create(fv_mesh* mesh, const int32_t* h_patch_kinds, fv_bc** out);
int fv_bc_destroy(fv_bc* bc);

/* ---- explicit face operators (config: 256^3 operator sweep) --------------
 * T: [N] cell values, Tb: [B] boundary values (VolumeField::boundaryData().value()),
 * phi/gamma: surface fields of length F+B (internal then boundary faces,
 * readers.hpp:186-213).  coeff multiplies the result (dsl::Coeff). */

/* SurfaceInterpolation<scalar>::interpolate (test_operators.cpp:52-59;
 * bench_explicitOperators.cpp:394): Tf[F+B]; d_phi only read for FV_UPWIND. */
int fv_interpolate(fv_ctx* ctx, const fv_mesh* mesh, fv_scheme scheme, fv_prec prec,
                   const void* d_phi, const void* d_T, const void* d_Tb, void* d_Tf);
/* GaussGreenGrad::grad (test_operators.cpp:72; bench_explicitOperators.cpp:283): out[3N]. */
int fv_grad_gauss(fv_ctx* ctx, const fv_mesh* mesh, fv_prec prec, double coeff, const void* d_T,
                  const void* d_Tb, void* d_gradT);
/* GaussGreenDiv<scalar>::div(out, phi, T, coeff) (test_operators.cpp:109-110;
 * bench_explicitOperators.cpp:114): out[N]. d_T == NULL gives the single-argument
 * dsl::exp::div(phi) = surfaceIntegrate(phi) (neoIcoFoam.cpp:145). */
int fv_div_gauss(fv_ctx* ctx, const fv_mesh* mesh, fv_scheme scheme, fv_prec prec, double coeff,
                 const void* d_phi, const void* d_T, const void* d_Tb, void* d_out);
/* GaussGreenLaplacian<scalar>::laplacian(out, gamma, T, coeff) with
 * "linear uncorrected" (bench_explicitOperators.cpp:212): out[N].
 * d_Tgrad: [B] fixedGradient values or NULL (zeroGradient). */
int fv_laplacian_uncorrected(fv_ctx* ctx, const fv_mesh* mesh, fv_prec prec, double coeff,
                             const void* d_gamma, const void* d_T, const fv_bc* bc,
                             const void* d_Tb, const void* d_Tgrad, void* d_out);
/* FaceNormalGradient<scalar> "uncorrected" faceNormalGrad (bench_explicitOperators.cpp:478):
 * out[F+B]. */
int fv_sngrad_uncorrected(fv_ctx* ctx, const fv_mesh* mesh, fv_prec prec, const void* d_T,
                          const fv_bc* bc, const void* d_Tb, const void* d_Tgrad, void* d_out);
/* VolumeField::correctBoundaryConditions for the scalar BC kinds
 * (readers.hpp:123): fixedGradient faces get Tb = T_P + g/deltaCoeffs_b
 * (g = d_Tgrad[b], or 0 when d_Tgrad is NULL: zeroGradient); fixedValue and
 * calculated faces keep their stored values.  Updates d_Tb[B] in place. */
int fv_correct_boundary(fv_ctx* ctx, const fv_mesh* mesh, fv_prec prec, const void* d_T,
                        const fv_bc* bc, const void* d_Tgrad, void* d_Tb);
/* Face flux phi = Sf . interpolate_linear(U) (FoamAdapter::flux,
 * src/algorithms/pressureVelocityCoupling.cpp:242-264): U[3N], Ub[3B] -> phi[F+B]. */
int fv_flux(fv_ctx* ctx, const fv_mesh* mesh, fv_prec prec, const void* d_U, const void* d_Ub,
            void* d_phi);

/* ---- implicit path: sparsity pattern, linear system, assembly, solve ----- */

typedef struct fv_sparsity fv_sparsity;
typedef struct fv_linsys fv_linsys;

/* la::SparsityPattern::readOrCreate(mesh) / la::createSparsity (expression.hpp:35;
 * test_sparsityPattern.cpp:36-62): CSR of the face graph with ascending columns
 * per row; requires OpenFOAM upper-triangular face order. */
int fv_sparsity_create(fv_mesh* mesh, fv_sparsity** out);
int fv_sparsity_destroy(fv_sparsity* sp);
/* sizes2 = {n_rows, nnz} (nnz = N + 2F) */
int fv_sparsity_sizes(const fv_sparsity* sp, int64_t* h_sizes2);
/* which: 0 rowOffs[N+1], 1 colIdxs[nnz], 2 diagOffset[N], 3 ownerOffset[F]
 * (slot of column neighbour in row owner), 4 neighbourOffset[F] (slot of
 * column owner in row neighbour) (pressureVelocityCoupling.cpp:87-107). */
int fv_sparsity_get(const fv_sparsity* sp, int which, int32_t* h_out);

/* la::createEmptyLinearSystem<scalar, localIdx>(mesh, sparsity)
 * (bench_implicitOperators.cpp:109-125): FP64 values[nnz] and rhs[N], zeroed.
 * Convention: A.x - b equals the assembled operators applied to x, times V
 * (test_implicitOperators.cpp:92-100,254-262). */
int fv_linsys_create(fv_sparsity* sp, fv_linsys** out);
int fv_linsys_destroy(fv_linsys* ls);
/* Device pointers of matrix().values() and rhs() (LinearSystem::view). */
void* fv_linsys_values(fv_linsys* ls);
void* fv_linsys_rhs(fv_linsys* ls);
/* NeoN::fill(values, 0); fill(rhs, 0) (bench_implicitOperators.cpp:116-117). */
int fv_linsys_zero(fv_linsys* ls);

/* DdtOperator(Implicit).implicitOperation(ls, t, dt), Euler
 * (test_implicitOperators.cpp:67-73): diag += coeff*V/dt, rhs += coeff*V/dt*T_old. */
int fv_assemble_ddt_euler(fv_linsys* ls, const void* d_Told, double dt, double coeff);
/* SourceTerm(Implicit, coeffField).implicitOperation(ls)
 * (test_implicitOperators.cpp:115-153): diag += coeff*c_P*V_P. */
int fv_assemble_source(fv_linsys* ls, const void* d_coeff_field, double coeff);
/* An explicit operator's per-volume result E[N] moved to the right-hand side of
 * the equation (dsl::exp terms of an Expression, neoIcoFoam.cpp:145):
 * rhs -= coeff*E*V. */
int fv_assemble_explicit(fv_linsys* ls, const void* d_E, double coeff);
/* LaplacianOperator(Implicit, gamma, T, "Gauss linear uncorrected").implicitOperation(ls)
 * (test_implicitOperators.cpp:242-249): off-diagonals coeff*gamma|Sf|nonOrthDeltaCoeffs,
 * diagonal minus their sum, boundary faces folded into diag/rhs. */
int fv_assemble_laplacian(fv_linsys* ls, const void* d_gamma, const fv_bc* bc, const void* d_Tb,
                          const void* d_Tgrad, double coeff);
/* The heat-transfer system dsl::Expression(imp::ddt(T) + coeff*imp::laplacian(gamma, T))
 * (heatTransfer.cpp:67, coeff = -1) assembled in one pass, OVERWRITING values and
 * rhs: bit-identical to fv_linsys_zero + fv_assemble_ddt_euler(.., 1) +
 * fv_assemble_laplacian(.., coeff). */
int fv_assemble_ddt_laplacian(fv_linsys* ls, const void* d_Told, double dt, const void* d_gamma,
                              const fv_bc* bc, const void* d_Tb, const void* d_Tgrad,
                              double coeff);
/* fv_assemble_ddt_laplacian of n systems (any meshes of one context) in one
 * launch, each bit-identical to its own call: entry i of every array belongs to
 * system i (Tgrad may be NULL, or hold NULL entries).  No counterpart in the
 * reference, which assembles one system per call; for the batched and served
 * heat steps, whose per-system launches would otherwise dominate small solves. */
int fv_assemble_ddt_laplacian_many(int n, fv_linsys* const* ls, const void* const* d_Told,
                                   const double* dt, const void* const* d_gamma,
                                   const fv_bc* const* bc, const void* const* d_Tb,
                                   const void* const* d_Tgrad, double coeff);
/* DivOperator(Implicit, phi, T, "Gauss <scheme>").implicitOperation(ls)
 * (test_implicitOperators.cpp:188-195): lower = -w phi, upper = lower + phi. */
int fv_assemble_div(fv_linsys* ls, fv_scheme scheme, const void* d_phi, const fv_bc* bc,
                    const void* d_Tb, const void* d_Tgrad, double coeff);
/* PDESolver::SetReference (expression.hpp:102-111): rhs[c] += A_cc*value; A_cc *= 2. */
int fv_set_reference(fv_linsys* ls, int cell, double value);
/* la::computeResidual via FoamAdapter::applyOperator (expression.hpp:204-220):
 * res = A.x - b. */
int fv_residual(fv_linsys* ls, const void* d_x, void* d_res);

typedef enum { FV_SOLVE_F64 = 0, FV_SOLVE_F32_IR = 1 } fv_solve_mode;
typedef struct {
    double rel_tol;       /* stop when ||r||/||b|| <= rel_tol (criteria relative_residual_norm) */
    int32_t max_iter;     /* criteria iteration (default 1000, fvSolution.cpp:114-118) */
    int32_t mode;         /* fv_solve_mode */
    double inner_rel_tol; /* F32_IR: FP32 residual drop since the last FP64 residual
                             replacement that triggers the next one; 0: 1e-5 */
    int32_t max_outer;    /* F32_IR: at most this many rounds (a chunk of FP32
                             iterations and a replacement); 0: no limit */
    int32_t check_every;  /* host polls the stop flag every this many iterations */
    double abs_tol;       /* also stop when ||r|| <= abs_tol (criteria absolute_residual_norm,
                             fvSolution.cpp:139-144); 0: off */
    int32_t eager;        /* fv_solve_cg: 1 launches a repeated solve's kernels one by one;
                             0 (default) replays them as a hipGraph once the same
                             schedule repeats (same bits either way) */
} fv_solve_opts;
typedef struct {
    double init_res;      /* SolverStats::initResNorm, relative */
    double final_res;     /* SolverStats::finalResNorm, relative */
    int32_t iters;        /* SolverStats::numIter (inner iterations for F32_IR) */
    int32_t outer_iters;  /* FP64 residual replacements (F32_IR; 1 for FP64) */
    double true_res;      /* explicit FP64 ||b - A x|| / ||b|| at exit */
} fv_solver_stats;
/* Jacobi-preconditioned CG on the assembled system, x as initial guess and
 * result (dsl::solve -> Ginkgo solver::Cg + preconditioner::Jacobi,
 * src/compatibility/fvSolution.cpp:22-28,50-55,106-142).  FV_SOLVE_F32_IR runs
 * the CG in FP32 for the correction and keeps r = b - A x, the update of x and
 * the convergence test in FP64.  Returns FV_ENOCONVERGE (stats filled) when
 * the tolerance is not reached. */
int fv_solve_cg(fv_linsys* ls, void* d_x, const fv_solve_opts* opts, fv_solver_stats* stats);
/* Jacobi-preconditioned BiCGStab on the assembled (non-symmetric) system, FP64
 * (solver::Bicgstab + preconditioner::Jacobi, the mapping of PBiCGStab and
 * smoothSolver in src/compatibility/fvSolution.cpp:22-27; used for implicit
 * advection, test/test_advection.cpp:171-178).  Stops on the recurrence residual
 * (relative or absolute); stats.true_res is the explicit FP64 ||b - A x|| / ||b||.
 * FV_ENOCONVERGE on the iteration limit or a breakdown (rho or omega zero). */
int fv_solve_bicgstab(fv_linsys* ls, void* d_x, const fv_solve_opts* opts,
                      fv_solver_stats* stats);

/* ---- pressure-velocity coupling (src/algorithms/pressureVelocityCoupling.cpp) ---
 * FP64.  The momentum matrix is the scalar system ls (the reference reads
 * component 0 of its Vec3 coefficients, :54-57); Vec3 fields are xyz
 * interleaved [3N] / [3B]. */
/* computeRAU (:38-61): rAU = V / diag(A); boundary values extrapolated into
 * d_rau_b [B] (may be NULL). */
int fv_compute_rau(fv_linsys* ls, void* d_rau, void* d_rau_b);
/* computeRAUandHByA (:63-119): HbyA = (b3 - sum_nb A(c,nb) U_nb) * rAU / V with
 * the sum in ascending face order (a deterministic gather in place of the
 * reference's atomics); d_b3 is the momentum right-hand side [3N]; boundary
 * values extrapolated into d_hbya_b [3B] (may be NULL). */
int fv_compute_hbya(fv_linsys* ls, const void* d_b3, const void* d_U, const void* d_rau,
                    void* d_hbya, void* d_hbya_b);
/* updateFaceVelocity (:122-195) on the assembled pressure system ls:
 * phi_f = pred_f - (A(N,P) p_N - A(P,N) p_P) on internal faces and
 * phi_b = pred_b - (bRhs_b - bMat_b p_P) on boundary faces, where bMat/bRhs are
 * the boundary coefficients the laplacian assembly of ls recorded (the face
 * flux of the laplacian term is bRhs - bMat p_P).  Surface arrays [F+B]. */
int fv_update_face_velocity(fv_linsys* ls, const void* d_phi_pred, const void* d_p, void* d_phi);
/* updateVelocity (:197-211): U = HbyA - rAU * gradP (gradP from fv_grad_gauss). */
int fv_update_velocity(fv_ctx* ctx, const fv_mesh* mesh, const void* d_hbya, const void* d_rau,
                       const void* d_gradp, void* d_U);
/* constrainHbyA (:14-36): HbyA_b = U_b on the faces of patches whose velocity
 * condition is not assignable (fixedValue); other faces keep their value. */
int fv_constrain_hbya(const fv_bc* u_bc, const void* d_Ub, void* d_hbya_b);
/* Courant number (NeoN computeCoNum, examples/scalarAdvection/scalarAdvection.cpp:69;
 * OpenFOAM CourantNo.H): co = 0.5 * max_c(sum_faces |phi| / V_c) * dt over the
 * internal and boundary faces of each cell, d_phi [F+B] FP64. */
int fv_courant_number(fv_ctx* ctx, const fv_mesh* mesh, const void* d_phi, double dt, double* co);
/* fv_courant_number for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  No
 * counterpart in the reference.  Allocates nothing: the per-block maxima live in
 * a fixed 1 MiB device array of the library (one per GPU, held until the
 * process ends), so meshes of at most 2^25 = 33,554,432 cells (FV_EINVAL
 * beyond); the result co = 0.5 * max_c(sum |phi| / V) * dt is written to the
 * DEVICE double d_co, stream-ordered and asynchronous (no host wait), with the
 * bits of fv_courant_number.  Calls that overlap on one GPU (from several
 * streams or threads) share that array and race. */
int fv_courant_number_noalloc(fv_ctx* ctx, const fv_mesh* mesh, const void* d_phi, double dt,
                              double* d_co);
/* setDeltaT (src/auxiliary/setup.cpp:13-22): new_dt = min(f * dt, max_dt) with
 * f = min(min(maxCo/(co + SMALL), 1 + 0.1 maxCo/(co + SMALL)), 1.2).  Host only. */
int fv_set_delta_t(double dt, double co, double max_co, double max_dt, double* new_dt);
/* The boundary coefficients [B] the laplacian assembly of ls recorded
 * (la::BoundaryCoefficients{matrixValues, rhsValues}, read at :180-192);
 * zeroed by fv_linsys_zero, accumulated by fv_assemble_laplacian, overwritten
 * by fv_assemble_ddt_laplacian. */
int fv_linsys_boundary_coeffs(fv_linsys* ls, void** d_bmat, void** d_brhs);

/* ---- batched solves (many independent systems in one launch sequence) ----
 * No counterpart in the reference (each NeoN solve is a separate Ginkgo call);
 * this is the batching layer the served and batched workloads run on.  A batch
 * holds assembled systems of any sizes; fv_batch_solve runs Jacobi-CG on all of
 * them together (row tiles of tile_rows rows over every system, per-system
 * deterministic reductions, FP64 residual replacement every `chunk`
 * iterations, a chunk captured once as a hipGraph when use_graph != 0).  A
 * system's result does not depend on the other members of the batch.
 * mode: FV_SOLVE_F64 or FV_SOLVE_F32_IR (FP32 CG, FP64 residual and test). */
typedef struct fv_batch fv_batch;
int fv_batch_create(fv_ctx* ctx, int mode, int tile_rows, int chunk, fv_batch** out);
int fv_batch_destroy(fv_batch* b);
/* d_x: FP64 initial guess and result [N]; d_xold: optional [N] that receives x
 * after the solve (Euler T_old <- T), or NULL.  The system must stay alive
 * while it is in the batch. */
int fv_batch_add(fv_batch* b, fv_linsys* ls, void* d_x, void* d_xold, int* id);
int fv_batch_clear(fv_batch* b);
int fv_batch_size(const fv_batch* b, int* n);
int fv_batch_set_shape(fv_batch* b, int tile_rows, int chunk);
/* stats: array of fv_batch_size entries (in insertion order) or NULL;
 * stats.outer_iters counts the member's FP64 residual replacements.
 * opts->inner_rel_tol: FP32 residual drop that ends a chunk early (mixed;
 * 0: 1e-5).  A member has converged when its FP64 relative residual is at most
 * opts->rel_tol or, when opts->abs_tol > 0, its absolute residual at most
 * opts->abs_tol (as fv_solve_cg). */
int fv_batch_solve(fv_batch* b, const fv_solve_opts* opts, int use_graph,
                   fv_solver_stats* stats);
/* GPU time the batch has spent so far, from HIP events on its stream:
 * h_ms2 = {chunks (the replayed chunk graphs or kernel sequences, each with
 * its FP64 residual replacement), batched starts}, h_counts2 = {chunks run,
 * starts run}.  No counterpart in the reference: kernel-level timing of the
 * batched and served solves without a profiler. */
int fv_batch_timing(const fv_batch* b, double* h_ms2, int64_t* h_counts2);
/* Autotuner: times full batch solves for tile_rows in {256..4096} x chunk in
 * {8,16,32} from the batch's current x (restored after every trial), applies
 * the fastest shape and records it in the JSON file cache_path (may be NULL)
 * under a key built from the precision, the member count and the row counts;
 * a later call with a matching key applies the cached shape without trials. */
int fv_batch_autotune(fv_batch* b, const fv_solve_opts* opts, const char* cache_path,
                      int* tile_rows, int* chunk);

/* ---- serving: a resident request loop for implicit solves ----------------
 * No counterpart in the reference (its solves run inside an application's time
 * loop).  The server owns a context (one HIP stream) on `device` and a worker
 * thread.  Meshes are added once and stay resident; requests (heat transfer or
 * pressure Poisson on a resident mesh) are admitted while their estimated
 * device memory fits the HBM budget, and run either by continuous batching on
 * the batched CG engine (systems join and leave between chunks) or one at a
 * time through fv_solve_cg (the plain drop-in path).  Results are returned in
 * the caller's cell numbering. */
typedef struct fv_server fv_server;
enum { FV_REQ_HEAT = 0, FV_REQ_POISSON = 1 };
enum { FV_REQ_QUEUED = 0, FV_REQ_RUNNING = 1, FV_REQ_DONE = 2 };
typedef struct {
    size_t hbm_budget;   /* admission budget in bytes: resident meshes + requests in
                            flight (0: unlimited) */
    int32_t mode;        /* FV_SOLVE_F32_IR or FV_SOLVE_F64 */
    int32_t batched;     /* 1: continuous batching, 0: one request at a time via fv_solve_cg */
    int32_t max_active;  /* systems in flight (batched), 0: no limit */
    int32_t tile_rows;   /* batch tile rows (multiple of 64), 0: 4096 */
    int32_t chunk;       /* CG iterations per FP64 residual replacement, 0: 16 */
    int32_t use_graph;   /* replay chunks as hipGraphs */
    int32_t solo_rows;   /* batched: a request on a mesh of at least this many cells runs
                            alone through fv_solve_cg on a second lane -- its own context
                            and worker thread, the mesh resident there too -- concurrently
                            with the batch (a system that size fills the GPU by itself);
                            0: 1 << 20, negative: never */
} fv_server_opts;
typedef struct {
    int32_t mesh_id;
    int32_t kind;        /* FV_REQ_HEAT: Euler steps of ddt(T) - laplacian(kappa,T), all walls
                            fixedValue; FV_REQ_POISSON: laplacian(1,p) = div(Sf.U_f), p = 0 on
                            patch 0, zeroGradient elsewhere */
    int32_t steps;       /* heat: Euler steps */
    int32_t max_iter;    /* CG iterations per solve */
    double dt, kappa;    /* heat */
    double rel_tol;      /* FP64 relative residual per solve */
    const double* h_wall_values; /* heat: one value per patch, or NULL (350) */
    const double* h_T0;  /* heat: initial field [N] in the caller's numbering, or NULL */
    uint64_t seed;       /* synthetic inputs: T0 ~ U[300,400) (heat), U ~ U[-1,1) (Poisson) */
} fv_request;
typedef struct {
    int32_t status;      /* FV_OK, FV_ENOCONVERGE, FV_ECANCELLED, FV_ENOMEM_BUDGET, FV_EDEVICE */
    int32_t iters;       /* CG iterations summed over the request's solves */
    double final_res;    /* FP64 relative residual of the last solve */
    double queue_ms;     /* submit -> admission */
    double latency_ms;   /* submit -> result */
} fv_request_stats;
typedef struct {
    int64_t submitted, completed, failed, rejected, cancelled;
    double wall_s;       /* first submission -> last completion */
    double solves_per_s; /* completed requests per second of wall_s */
    double p50_ms, p99_ms, mean_ms, max_ms;
    size_t peak_bytes;   /* peak device bytes in use */
    int32_t peak_active; /* peak requests in flight */
    size_t resident_bytes; /* device bytes of the resident meshes */
    /* where the worker's time went (ms): admitting requests (allocation, inputs,
     * assembly), GPU chunks (batched: one chunk of every member, with its
     * synchronisation; plain: whole solves; the batched server's solo lane
     * reports its solves in solo_ms), and advancing requests between
     * chunks (next Euler step, results, release); chunks run, launch tables
     * built and chunk graphs captured (batched) */
    double admit_ms, gpu_ms, advance_ms;
    int64_t chunks, table_builds, graph_builds;
    /* batched: GPU time of the chunks and of the batched starts from HIP events
     * around each (the part of gpu_ms the device was working) */
    double chunk_event_ms, start_event_ms;
    int64_t solo_runs;   /* batched: requests that ran alone (solo_rows) */
    double solo_ms;      /* batched: time of those solves on the solo lane (its own
                            context and worker thread, concurrent with the batch) */
} fv_server_stats;
int fv_server_create(int device, const fv_server_opts* opts, fv_server** out);
/* Adds a resident mesh (renumber != 0: Z-order renumbered internally). */
int fv_server_add_mesh(fv_server* s, const fv_polymesh_desc* mesh, int renumber, int* mesh_id);
int fv_server_mesh_cells(fv_server* s, int mesh_id, int32_t* n_cells);
/* Device bytes one request of `kind` on the mesh is admitted with.  The budget
 * holds the resident meshes plus these estimates of the requests in flight (at
 * least one request runs whenever it fits beside the meshes). */
int fv_server_request_bytes(fv_server* s, int mesh_id, int kind, size_t* bytes);
/* Queues a request; FV_ENOMEM_BUDGET if it can never fit the budget. */
int fv_server_submit(fv_server* s, const fv_request* req, int64_t* req_id);
int fv_server_poll(fv_server* s, int64_t req_id, int* state);
/* Waits for a request, copies its result field [N] (caller numbering; may be
 * NULL) and statistics, and releases it.  Returns the request's status.  One
 * waiter per request: a second wait on a request that is already waited on
 * returns FV_EINVAL (and a wait on a released request "unknown request id"). */
int fv_server_wait(fv_server* s, int64_t req_id, double* h_result, fv_request_stats* stats);
/* Cancels a queued or running request (its wait returns FV_ECANCELLED). */
int fv_server_cancel(fv_server* s, int64_t req_id);
int fv_server_get_stats(fv_server* s, fv_server_stats* out);
/* Stops the worker and frees every request; no other call on the server may
 * run concurrently with (or after) it. */
int fv_server_destroy(fv_server* s);

#ifdef __cplusplus
}
#endif
#endif /* FVHIP_H */
