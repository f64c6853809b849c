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
/* Counter-based uniform random numbers in [lo,hi): element i gets
 * lo + (hi-lo)*u(seed, stream, i) with u = splitmix64-hash(...) >> 11 * 2^-53,
 * identical on host and device, so CPU checkers can regenerate any input. */
int fv_fill_uniform(fv_ctx* ctx, fv_prec prec, void* d_ptr, int64_t n, double lo, double hi,
                    uint64_t seed, uint64_t stream);

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

#ifdef __cplusplus
}
#endif
#endif /* FVHIP_H */
