// hexmesh.cpp -- structured hex block meshes in OpenFOAM polyMesh form (the
// blockMesh cases the reference runs: tutorials/heatTranfer3D/system/blockMeshDict,
// benchmarks/benchmarkSuite/templates/3DCube/system/blockMeshDict), with seeded
// interior-vertex jitter and a seeded random cell numbering.  Host code: it
// produces the polyMesh primitives that fv_mesh_create_poly turns into device
// geometry.
//
// Layout produced (OpenFOAM upper-triangular order):
//   * points p(i,j,k) = (i/nx, j/ny, k/nz) * L, id = i + (nx+1)*(j + (ny+1)*k);
//     points with all coordinates strictly inside are moved by
//     (2u-1)*jitter*h per coordinate, u = rng_u01(seed, axis, point id);
//   * cells numbered c = i + nx*(j + ny*k), then renumbered by a seeded
//     Fisher-Yates permutation;
//   * internal faces with owner < neighbour, sorted by owner then neighbour,
//     points ordered so the normal points from owner to neighbour;
//   * boundary patches x0, x1, y0, y1, z0, z1 (outward normals).
#include <algorithm>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.hpp"

using namespace fvhip;

namespace {

struct Dims {
    int64_t nx, ny, nz;
    int64_t pid(int64_t i, int64_t j, int64_t k) const { return i + (nx + 1) * (j + (ny + 1) * k); }
    int64_t cid(int64_t i, int64_t j, int64_t k) const { return i + nx * (j + ny * k); }
};

// quad points with normal +axis at plane index (on that axis) q and cell-corner (a,b)
inline void quad(const Dims& d, int axis, int64_t i, int64_t j, int64_t k, int32_t* out) {
    // i,j,k is the lowest corner point of the face
    if (axis == 0) {
        out[0] = (int32_t)d.pid(i, j, k);
        out[1] = (int32_t)d.pid(i, j + 1, k);
        out[2] = (int32_t)d.pid(i, j + 1, k + 1);
        out[3] = (int32_t)d.pid(i, j, k + 1);
    } else if (axis == 1) {
        out[0] = (int32_t)d.pid(i, j, k);
        out[1] = (int32_t)d.pid(i, j, k + 1);
        out[2] = (int32_t)d.pid(i + 1, j, k + 1);
        out[3] = (int32_t)d.pid(i + 1, j, k);
    } else {
        out[0] = (int32_t)d.pid(i, j, k);
        out[1] = (int32_t)d.pid(i + 1, j, k);
        out[2] = (int32_t)d.pid(i + 1, j + 1, k);
        out[3] = (int32_t)d.pid(i, j + 1, k);
    }
}

// OpenFOAM face::reverseFace keeps the first point: (v0, v3, v2, v1)
inline void reverse4(int32_t* q) { std::swap(q[1], q[3]); }

}  // namespace

extern "C" {

int fv_hexmesh_sizes(int32_t nx, int32_t ny, int32_t nz, int64_t* h_sizes6) {
    return guarded([&] {
        FV_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && h_sizes6, "bad dimensions");
        int64_t X = nx, Y = ny, Z = nz;
        int64_t nP = (X + 1) * (Y + 1) * (Z + 1);
        int64_t nC = X * Y * Z;
        int64_t nI = (X - 1) * Y * Z + X * (Y - 1) * Z + X * Y * (Z - 1);
        int64_t nB = 2 * (Y * Z + X * Z + X * Y);
        FV_REQUIRE(nP < INT32_MAX && nI + nB < INT32_MAX / 4, "mesh too large for int32 labels");
        h_sizes6[0] = nP;
        h_sizes6[1] = nC;
        h_sizes6[2] = nI;
        h_sizes6[3] = nB;
        h_sizes6[4] = 6;                  // patches
        h_sizes6[5] = 4 * (nI + nB);      // face point entries
    });
}

int fv_hexmesh_generate(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz,
                        double jitter, uint64_t jitter_seed, uint64_t perm_seed,
                        double* h_points, int32_t* h_face_offsets, int32_t* h_face_points,
                        int32_t* h_owner, int32_t* h_neighbour, int32_t* h_patch_start,
                        int32_t* h_patch_size) {
    return guarded([&] {
        FV_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "bad dimensions");
        FV_REQUIRE(jitter >= 0.0 && jitter < 0.5, "jitter must be in [0, 0.5)");
        FV_REQUIRE(h_points && h_face_offsets && h_face_points && h_owner && h_neighbour &&
                       h_patch_start && h_patch_size,
                   "NULL output");
        const Dims d{nx, ny, nz};
        const int64_t X = nx, Y = ny, Z = nz;
        const int64_t nC = X * Y * Z;
        const double hx = lx / X, hy = ly / Y, hz = lz / Z;
        // points
        for (int64_t k = 0; k <= Z; ++k)
            for (int64_t j = 0; j <= Y; ++j)
                for (int64_t i = 0; i <= X; ++i) {
                    int64_t p = d.pid(i, j, k);
                    double x = lx * ((double)i / (double)X);
                    double y = ly * ((double)j / (double)Y);
                    double z = lz * ((double)k / (double)Z);
                    bool interior = i > 0 && i < X && j > 0 && j < Y && k > 0 && k < Z;
                    if (interior && jitter > 0.0) {
                        x += (2.0 * rng_u01(jitter_seed, 0, (uint64_t)p) - 1.0) * jitter * hx;
                        y += (2.0 * rng_u01(jitter_seed, 1, (uint64_t)p) - 1.0) * jitter * hy;
                        z += (2.0 * rng_u01(jitter_seed, 2, (uint64_t)p) - 1.0) * jitter * hz;
                    }
                    h_points[3 * p] = x;
                    h_points[3 * p + 1] = y;
                    h_points[3 * p + 2] = z;
                }
        // cell renumbering newOfOld
        std::vector<int32_t> newOf(nC);
        std::iota(newOf.begin(), newOf.end(), 0);
        if (perm_seed) {
            for (int64_t i = nC - 1; i > 0; --i) {
                int64_t j = (int64_t)(rng_u64(perm_seed, 7, (uint64_t)i) % (uint64_t)(i + 1));
                std::swap(newOf[i], newOf[j]);
            }
        }
        // internal faces, bucketed by owner (counting sort), then by neighbour
        const int64_t nI = (X - 1) * Y * Z + X * (Y - 1) * Z + X * Y * (Z - 1);
        std::vector<int32_t> cnt(nC + 1, 0);
        auto for_each_internal = [&](auto&& fn) {
            for (int64_t k = 0; k < Z; ++k)
                for (int64_t j = 0; j < Y; ++j)
                    for (int64_t i = 0; i < X; ++i) {
                        int64_t c = d.cid(i, j, k);
                        if (i + 1 < X) fn(c, d.cid(i + 1, j, k), 0, i + 1, j, k);
                        if (j + 1 < Y) fn(c, d.cid(i, j + 1, k), 1, i, j + 1, k);
                        if (k + 1 < Z) fn(c, d.cid(i, j, k + 1), 2, i, j, k + 1);
                    }
        };
        for_each_internal([&](int64_t a, int64_t b, int, int64_t, int64_t, int64_t) {
            int32_t na = newOf[a], nb = newOf[b];
            cnt[std::min(na, nb) + 1]++;
        });
        for (int64_t c = 0; c < nC; ++c) cnt[c + 1] += cnt[c];
        std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
        for_each_internal([&](int64_t a, int64_t b, int axis, int64_t i, int64_t j, int64_t k) {
            int32_t na = newOf[a], nb = newOf[b];
            int32_t o = std::min(na, nb), n = std::max(na, nb);
            int64_t f = fill[o]++;
            h_owner[f] = o;
            h_neighbour[f] = n;
            int32_t* q = h_face_points + 4 * f;
            quad(d, axis, i, j, k, q);
            if (na > nb) reverse4(q);
        });
        // sort each owner bucket by neighbour (buckets hold at most 6 faces)
        for (int64_t c = 0; c < nC; ++c) {
            for (int64_t f = cnt[c] + 1; f < cnt[c + 1]; ++f) {
                int64_t g = f;
                while (g > cnt[c] && h_neighbour[g - 1] > h_neighbour[g]) {
                    std::swap(h_neighbour[g - 1], h_neighbour[g]);
                    for (int t = 0; t < 4; ++t) std::swap(h_face_points[4 * (g - 1) + t], h_face_points[4 * g + t]);
                    --g;
                }
            }
        }
        // boundary patches
        int64_t f = nI;
        auto patch = [&](int pi, auto&& body) {
            h_patch_start[pi] = (int32_t)f;
            int64_t f0 = f;
            body();
            h_patch_size[pi] = (int32_t)(f - f0);
        };
        patch(0, [&] {  // x0
            for (int64_t k = 0; k < Z; ++k)
                for (int64_t j = 0; j < Y; ++j) {
                    int32_t* q = h_face_points + 4 * f;
                    quad(d, 0, 0, j, k, q);
                    reverse4(q);
                    h_owner[f++] = newOf[d.cid(0, j, k)];
                }
        });
        patch(1, [&] {  // x1
            for (int64_t k = 0; k < Z; ++k)
                for (int64_t j = 0; j < Y; ++j) {
                    quad(d, 0, X, j, k, h_face_points + 4 * f);
                    h_owner[f++] = newOf[d.cid(X - 1, j, k)];
                }
        });
        patch(2, [&] {  // y0
            for (int64_t k = 0; k < Z; ++k)
                for (int64_t i = 0; i < X; ++i) {
                    int32_t* q = h_face_points + 4 * f;
                    quad(d, 1, i, 0, k, q);
                    reverse4(q);
                    h_owner[f++] = newOf[d.cid(i, 0, k)];
                }
        });
        patch(3, [&] {  // y1
            for (int64_t k = 0; k < Z; ++k)
                for (int64_t i = 0; i < X; ++i) {
                    quad(d, 1, i, Y, k, h_face_points + 4 * f);
                    h_owner[f++] = newOf[d.cid(i, Y - 1, k)];
                }
        });
        patch(4, [&] {  // z0
            for (int64_t j = 0; j < Y; ++j)
                for (int64_t i = 0; i < X; ++i) {
                    int32_t* q = h_face_points + 4 * f;
                    quad(d, 2, i, j, 0, q);
                    reverse4(q);
                    h_owner[f++] = newOf[d.cid(i, j, 0)];
                }
        });
        patch(5, [&] {  // z1
            for (int64_t j = 0; j < Y; ++j)
                for (int64_t i = 0; i < X; ++i) {
                    quad(d, 2, i, j, Z, h_face_points + 4 * f);
                    h_owner[f++] = newOf[d.cid(i, j, Z - 1)];
                }
        });
        for (int64_t g = 0; g <= f; ++g) h_face_offsets[g] = (int32_t)(4 * g);
    });
}

}  // extern "C"
