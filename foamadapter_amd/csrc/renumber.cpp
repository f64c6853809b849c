// renumber.cpp -- locality renumbering of a polyMesh (host), the counterpart of
// OpenFOAM's renumberMesh utility.
//
// Method 0 (renumberMesh's default Cuthill-McKee): cells ordered by a
// breadth-first sweep of the face graph started from a pseudo-peripheral cell,
// visiting neighbours by increasing degree, the sweep order reversed.
// Method 1 (Z-order): cells sorted by the Morton code of their centres, so that
// any aligned run of 2^k cells is a compact brick of space -- the property the
// LDS-staged face kernels use (a 256-cell workgroup owns most of its faces).
// Internal faces are then rewritten
// in upper-triangular order of the new numbering (owner < neighbour, sorted by
// owner then neighbour); a face whose owner and neighbour change sides has its
// point list reversed (face::reverseFace keeps the first point), so its area
// vector flips and oriented face fields (fluxes) change sign.  Boundary faces
// keep their positions and patches.
//
// Outputs map the new numbering to the input one:
//   cell_new_to_old[N]
//   face_new_to_old[nFaces]: old face index, or ~old if the face was flipped.
#include <algorithm>
#include <numeric>
#include <vector>

#include "common.hpp"

using namespace fvhip;

namespace {

struct Graph {
    std::vector<int32_t> off, adj;
};

Graph build_graph(int64_t N, int64_t F, const int32_t* own, const int32_t* nbr) {
    Graph g;
    g.off.assign(N + 1, 0);
    for (int64_t f = 0; f < F; ++f) {
        g.off[own[f] + 1]++;
        g.off[nbr[f] + 1]++;
    }
    for (int64_t c = 0; c < N; ++c) g.off[c + 1] += g.off[c];
    g.adj.resize(2 * F);
    std::vector<int32_t> fill(g.off.begin(), g.off.end() - 1);
    for (int64_t f = 0; f < F; ++f) {
        g.adj[fill[own[f]]++] = nbr[f];
        g.adj[fill[nbr[f]]++] = own[f];
    }
    return g;
}

// BFS levels from `start`; returns the last visited cell and fills order
int32_t bfs(const Graph& g, int32_t start, std::vector<int32_t>& order, std::vector<int32_t>& mark,
            int32_t stamp, bool sort_by_degree) {
    size_t head = order.size();
    order.push_back(start);
    mark[start] = stamp;
    std::vector<int32_t> nb;
    while (head < order.size()) {
        int32_t c = order[head++];
        nb.clear();
        for (int32_t k = g.off[c]; k < g.off[c + 1]; ++k)
            if (mark[g.adj[k]] != stamp) {
                mark[g.adj[k]] = stamp;
                nb.push_back(g.adj[k]);
            }
        if (sort_by_degree)
            std::stable_sort(nb.begin(), nb.end(), [&](int32_t a, int32_t b) {
                return g.off[a + 1] - g.off[a] < g.off[b + 1] - g.off[b];
            });
        order.insert(order.end(), nb.begin(), nb.end());
    }
    return order.back();
}

// spread the low 21 bits of v to every third bit
inline uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

// Z-order (Morton) of approximate cell centres (average of the cell's face-point
// averages).  Each axis is quantised in units of the mesh's typical cell
// spacing along it -- the median distance between face-neighbour centres whose
// offset points mainly along that axis -- from the points' bounding-box corner,
// with the same number of sub-cell levels on every axis.  On a structured mesh
// (a box of any size, its vertices jittered or not) the top bits of a code are
// then a cell's integer coordinates, so a run of 2^k consecutive cells is a
// brick of cells whatever the box's sides; on an unstructured mesh the spacing
// is an average and the runs stay compact.
void morton_order(const fv_polymesh_desc* d, std::vector<int32_t>& order) {
    const int64_t N = d->n_cells, nF = d->n_faces, F = d->n_internal_faces, nP = d->n_points;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t p = 0; p < nP; ++p)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], d->h_points[3 * p + k]);
            hi[k] = std::max(hi[k], d->h_points[3 * p + k]);
        }
    std::vector<double> cs(3 * N, 0.0);
    std::vector<int32_t> cnt(N, 0);
    for (int64_t f = 0; f < nF; ++f) {
        const int32_t a = d->h_face_offsets[f], n = d->h_face_offsets[f + 1] - a;
        double fc[3] = {0, 0, 0};
        for (int32_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) fc[k] += d->h_points[3 * (int64_t)d->h_face_points[a + i] + k];
        for (int k = 0; k < 3; ++k) fc[k] /= n;
        const int32_t o = d->h_owner[f];
        for (int k = 0; k < 3; ++k) cs[3 * o + k] += fc[k];
        cnt[o]++;
        if (f < F) {
            const int32_t nb = d->h_neighbour[f];
            for (int k = 0; k < 3; ++k) cs[3 * nb + k] += fc[k];
            cnt[nb]++;
        }
    }
    for (int64_t c = 0; c < N; ++c)
        for (int k = 0; k < 3; ++k) cs[3 * c + k] /= cnt[c];
    // typical spacing per axis from (at most about 2^20 sampled) internal faces
    double h[3];
    {
        std::vector<double> dk[3];
        const int64_t stride = std::max<int64_t>(1, F >> 20);
        for (int64_t f = 0; f < F; f += stride) {
            const int32_t o = d->h_owner[f], nb = d->h_neighbour[f];
            double v[3];
            int ax = 0;
            for (int k = 0; k < 3; ++k) {
                v[k] = std::abs(cs[3 * nb + k] - cs[3 * o + k]);
                if (v[k] > v[ax]) ax = k;
            }
            if (v[ax] > 0.0) dk[ax].push_back(v[ax]);
        }
        double hmax = 0.0;
        for (int k = 0; k < 3; ++k) {
            h[k] = 0.0;
            if (!dk[k].empty()) {
                auto mid = dk[k].begin() + dk[k].size() / 2;
                std::nth_element(dk[k].begin(), mid, dk[k].end());
                h[k] = *mid;
            }
            hmax = std::max(hmax, h[k]);
        }
        // an axis with no neighbours along it (one cell thick) or no faces at
        // all: quantise it like the widest spacing, or the box
        double ext = 0.0;
        for (int k = 0; k < 3; ++k) ext = std::max(ext, hi[k] - lo[k]);
        if (!(ext > 0.0)) ext = 1.0;
        for (int k = 0; k < 3; ++k)
            if (!(h[k] > 0.0)) h[k] = hmax > 0.0 ? hmax : ext;
    }
    // sub-cell levels: the most cells along an axis times 2^b stays below 2^21
    double cells = 1.0;
    for (int k = 0; k < 3; ++k) cells = std::max(cells, (hi[k] - lo[k]) / h[k] + 1.0);
    int b = 0;
    while (b < 20 && cells * double(1 << (b + 1)) < 2097152.0) ++b;
    const double top = 2097151.0;
    std::vector<std::pair<uint64_t, int32_t>> key(N);
    for (int64_t c = 0; c < N; ++c) {
        uint64_t code = 0;
        for (int k = 0; k < 3; ++k) {
            const double u = (cs[3 * c + k] - lo[k]) / h[k] * double(1 << b);
            const uint64_t q = u <= 0.0 ? 0 : (u >= top ? (uint64_t)top : (uint64_t)u);
            code |= spread3(q) << (2 - k);
        }
        key[c] = {code, (int32_t)c};
    }
    std::sort(key.begin(), key.end());
    order.resize(N);
    for (int64_t i = 0; i < N; ++i) order[i] = key[i].second;
}

}  // namespace

extern "C" {

// method 0: reverse Cuthill-McKee; method 1: Z-order (Morton) of cell centres
int fv_polymesh_renumber(const fv_polymesh_desc* d, int method, int32_t* h_face_offsets_out,
                         int32_t* h_face_points_out, int32_t* h_owner_out,
                         int32_t* h_neighbour_out, int32_t* h_cell_new_to_old,
                         int32_t* h_face_new_to_old) {
    return guarded([&] {
        FV_REQUIRE(d && h_face_offsets_out && h_face_points_out && h_owner_out &&
                       h_neighbour_out && h_cell_new_to_old && h_face_new_to_old,
                   "NULL argument");
        FV_REQUIRE(method == 0 || method == 1, "unknown renumbering method");
        const int64_t N = d->n_cells, nF = d->n_faces, F = d->n_internal_faces;
        for (int64_t f = 0; f < nF; ++f)
            FV_REQUIRE(d->h_owner[f] >= 0 && d->h_owner[f] < N, "owner out of range");
        for (int64_t f = 0; f < F; ++f)
            FV_REQUIRE(d->h_neighbour[f] >= 0 && d->h_neighbour[f] < N, "neighbour out of range");
        std::vector<int32_t> order;
        if (method == 1) {
            FV_REQUIRE(d->h_points && d->h_face_offsets && d->h_face_points, "NULL points/faces");
            morton_order(d, order);
        } else {
            Graph g = build_graph(N, F, d->h_owner, d->h_neighbour);
            // order all connected components
            order.reserve(N);
            std::vector<int32_t> mark(N, -1);
            int32_t stamp = 0;
            for (int64_t seed = 0; seed < N; ++seed) {
                if (mark[seed] >= 0) continue;
                // pseudo-peripheral start: two BFS sweeps from the component seed
                std::vector<int32_t> tmp;
                int32_t far = bfs(g, (int32_t)seed, tmp, mark, stamp++, false);
                for (int32_t c : tmp) mark[c] = -1;
                tmp.clear();
                far = bfs(g, far, tmp, mark, stamp++, false);
                for (int32_t c : tmp) mark[c] = -1;
                bfs(g, far, order, mark, stamp++, true);
            }
            FV_REQUIRE((int64_t)order.size() == N, "renumbering lost cells");
            std::reverse(order.begin(), order.end());
        }
        std::vector<int32_t> newOf(N);
        for (int64_t i = 0; i < N; ++i) {
            h_cell_new_to_old[i] = order[i];
            newOf[order[i]] = (int32_t)i;
        }
        // internal faces: new owner/neighbour, bucket by owner, sort by neighbour
        std::vector<int32_t> cnt(N + 1, 0);
        for (int64_t f = 0; f < F; ++f)
            cnt[std::min(newOf[d->h_owner[f]], newOf[d->h_neighbour[f]]) + 1]++;
        for (int64_t c = 0; c < N; ++c) cnt[c + 1] += cnt[c];
        std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1), slotFace(F), slotNbr(F);
        for (int64_t f = 0; f < F; ++f) {
            int32_t a = newOf[d->h_owner[f]], b = newOf[d->h_neighbour[f]];
            int32_t o = std::min(a, b);
            int32_t s = fill[o]++;
            slotFace[s] = a < b ? (int32_t)f : ~(int32_t)f;
            slotNbr[s] = std::max(a, b);
        }
        for (int64_t c = 0; c < N; ++c) {
            // buckets are small: insertion sort by neighbour
            for (int32_t s = cnt[c] + 1; s < cnt[c + 1]; ++s)
                for (int32_t t = s; t > cnt[c] && slotNbr[t - 1] > slotNbr[t]; --t) {
                    std::swap(slotNbr[t - 1], slotNbr[t]);
                    std::swap(slotFace[t - 1], slotFace[t]);
                }
            for (int32_t s = cnt[c]; s < cnt[c + 1]; ++s) {
                h_owner_out[s] = (int32_t)c;
                h_neighbour_out[s] = slotNbr[s];
                h_face_new_to_old[s] = slotFace[s];
            }
        }
        for (int64_t f = F; f < nF; ++f) {
            h_owner_out[f] = newOf[d->h_owner[f]];
            h_face_new_to_old[f] = (int32_t)f;
        }
        // face point lists in the new face order, reversed for flipped faces
        h_face_offsets_out[0] = 0;
        for (int64_t s = 0; s < nF; ++s) {
            int32_t e = h_face_new_to_old[s];
            int32_t f = e >= 0 ? e : ~e;
            int32_t a = d->h_face_offsets[f], n = d->h_face_offsets[f + 1] - a;
            int32_t o = h_face_offsets_out[s];
            h_face_offsets_out[s + 1] = o + n;
            if (e >= 0) {
                for (int32_t i = 0; i < n; ++i) h_face_points_out[o + i] = d->h_face_points[a + i];
            } else {
                h_face_points_out[o] = d->h_face_points[a];
                for (int32_t i = 1; i < n; ++i) h_face_points_out[o + i] = d->h_face_points[a + n - i];
            }
        }
    });
}

}  // extern "C"
