// runtime.hip -- context, device memory accounting, copies, fills and the
// counter-based uniform generator.
#include "common.hpp"

namespace fvhip {

static thread_local std::string g_last_error;

// live contexts and their generations
static std::mutex g_ctx_mu;
static std::unordered_map<fv_ctx*, uint64_t> g_live;
static uint64_t g_next_gen = 1;

uint64_t ctx_generation(fv_ctx* ctx) {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    auto it = g_live.find(ctx);
    return it == g_live.end() ? 0 : it->second;
}

void* ctx_poll_host(fv_ctx* ctx) {
    if (!ctx->pollHost) {
        FV_HIP(hipSetDevice(ctx->device));
        FV_HIP(hipHostMalloc(&ctx->pollHost, fv_ctx::kPollBytes, hipHostMallocDefault));
        for (auto& e : ctx->pollEv) FV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    return ctx->pollHost;
}

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

// Device memory is pooled per context: freed blocks are kept, keyed by their
// size class (multiples of 1/8 of the power of two below the request, at least
// 256 bytes), and handed out again for a request of the same class.  Reuse
// needs no synchronisation: a context's work is ordered on its one stream.  The
// cache is trimmed (largest blocks first) beyond kCacheLimit and emptied before
// a failed hipMalloc is retried.
constexpr size_t kCacheLimit = size_t(8) << 30;

size_t size_class(size_t bytes) {
    if (bytes <= 256) return 256;
    size_t p = 1;
    while (p <= bytes / 2) p <<= 1;  // largest power of two <= bytes
    const size_t step = std::max<size_t>(256, p / 8);
    return (bytes + step - 1) / step * step;
}

void trim_cache(fv_ctx* ctx, size_t keep) {
    if (ctx->cached <= keep) return;
    (void)hipStreamSynchronize(ctx->stream);
    while (ctx->cached > keep && !ctx->cache.empty()) {
        auto it = std::prev(ctx->cache.end());
        ctx->cached -= it->first;
        (void)hipFree(it->second);
        ctx->cache.erase(it);
    }
}

void* ctx_alloc(fv_ctx* ctx, size_t bytes) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->budget && ctx->in_use + bytes > ctx->budget)
        throw Error{FV_ENOMEM_BUDGET, "HBM budget exceeded: in use " + std::to_string(ctx->in_use) +
                                          " + request " + std::to_string(bytes) + " > budget " +
                                          std::to_string(ctx->budget)};
    const size_t cls = size_class(bytes);
    void* p = nullptr;
    auto hit = ctx->cache.find(cls);
    if (hit != ctx->cache.end()) {
        p = hit->second;
        ctx->cached -= cls;
        ctx->cache.erase(hit);
    } else {
        FV_HIP(hipSetDevice(ctx->device));
        hipError_t e = hipMalloc(&p, cls);
        if (e != hipSuccess && !ctx->cache.empty()) {
            (void)hipGetLastError();
            trim_cache(ctx, 0);
            e = hipMalloc(&p, cls);
        }
        if (e != hipSuccess)
            throw Error{FV_EDEVICE, std::string("hipMalloc(") + std::to_string(cls) + "): " +
                                        hipGetErrorString(e)};
    }
    ctx->sizes[p] = {bytes, cls};
    ctx->in_use += bytes;
    return p;
}

void ctx_free(fv_ctx* ctx, void* p) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->sizes.find(p);
    if (it == ctx->sizes.end()) return;
    ctx->in_use -= it->second.first;
    const size_t cls = it->second.second;
    ctx->sizes.erase(it);
    ctx->cache.emplace(cls, p);
    ctx->cached += cls;
    trim_cache(ctx, kCacheLimit);
}

template <typename S>
__global__ void k_fill(S* __restrict__ p, int64_t n, S v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t s = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) p[i] = v;
}

template <typename D, typename S>
__global__ void k_convert(D* __restrict__ d, const S* __restrict__ s, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) d[i] = (D)s[i];
}

template <typename S>
__global__ void k_uniform(S* __restrict__ p, int64_t n, double lo, double hi, uint64_t seed,
                          uint64_t stream) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t s = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) p[i] = (S)(lo + (hi - lo) * rng_u01(seed, stream, (uint64_t)i));
}

template <typename S>
__global__ void k_permute(int64_t n, int nc, int mode, int oriented,
                          const int32_t* __restrict__ map, const S* __restrict__ src,
                          S* __restrict__ dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) {
        const int32_t e = map[i];
        const int64_t j = e >= 0 ? e : ~e;
        const bool neg = oriented && e < 0;
        for (int k = 0; k < nc; ++k) {
            if (mode == 0) {
                const S v = src[j * nc + k];
                dst[i * nc + k] = neg ? -v : v;
            } else {
                const S v = src[i * nc + k];
                dst[j * nc + k] = neg ? -v : v;
            }
        }
    }
}

// k_permute for NC = 1 or 3 components: each thread takes kPermU elements
// (i = base + u * blockDim), their map entries and values loaded together
// before any store, so a thread has that many random accesses in flight.
constexpr int kPermU = 4;
template <typename S, int NC, int MODE>
__global__ __launch_bounds__(256) void k_permute_nc(int64_t n, int oriented,
                                                    const int32_t* __restrict__ map,
                                                    const S* __restrict__ src, S* __restrict__ dst) {
    const int64_t base = (int64_t)blockIdx.x * blockDim.x * kPermU + threadIdx.x;
    int32_t e[kPermU];
#pragma unroll
    for (int u = 0; u < kPermU; ++u) {
        const int64_t i = base + (int64_t)u * blockDim.x;
        e[u] = i < n ? map[i] : 0;
    }
    S v[kPermU][NC];
#pragma unroll
    for (int u = 0; u < kPermU; ++u) {
        const int64_t i = base + (int64_t)u * blockDim.x;
        const int64_t j = e[u] >= 0 ? e[u] : ~e[u];
        const int64_t from = MODE == 0 ? j : i;
#pragma unroll
        for (int k = 0; k < NC; ++k) v[u][k] = i < n ? src[from * NC + k] : S(0);
    }
#pragma unroll
    for (int u = 0; u < kPermU; ++u) {
        const int64_t i = base + (int64_t)u * blockDim.x;
        if (i >= n) break;
        const int64_t j = e[u] >= 0 ? e[u] : ~e[u];
        const int64_t to = MODE == 0 ? i : j;
        const bool neg = oriented && e[u] < 0;
#pragma unroll
        for (int k = 0; k < NC; ++k) dst[to * NC + k] = neg ? -v[u][k] : v[u][k];
    }
}

template <typename S>
void launch_permute(hipStream_t s, int64_t n, int nc, int mode, int oriented, const int32_t* map,
                    const S* src, S* dst) {
    const int64_t per = 256 * (int64_t)kPermU;
    const unsigned g = (unsigned)((n + per - 1) / per);
    if (nc == 1 && mode == 0)
        k_permute_nc<S, 1, 0><<<g, 256, 0, s>>>(n, oriented, map, src, dst);
    else if (nc == 1)
        k_permute_nc<S, 1, 1><<<g, 256, 0, s>>>(n, oriented, map, src, dst);
    else if (nc == 3 && mode == 0)
        k_permute_nc<S, 3, 0><<<g, 256, 0, s>>>(n, oriented, map, src, dst);
    else if (nc == 3)
        k_permute_nc<S, 3, 1><<<g, 256, 0, s>>>(n, oriented, map, src, dst);
    else
        k_permute<S><<<grid_for(n, 256, 8192), 256, 0, s>>>(n, nc, mode, oriented, map, src, dst);
}

// one component of an interleaved field: mode 0 dst[i] = src[i*nc + k] (out),
// mode 1 dst[i*nc + k] = src[i] (in)
template <typename S>
__global__ void k_component(int64_t n, int nc, int k, int mode, const S* __restrict__ src,
                            S* __restrict__ dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += st) {
        if (mode == 0)
            dst[i] = src[i * nc + k];
        else
            dst[i * nc + k] = src[i];
    }
}

// Streaming reference kernels: what a kernel that only streams HBM reaches on
// this part (the ceiling the face operators are judged against).  Each thread
// moves kStreamU 16-byte elements, lanes consecutive (coalesced), one pass:
//   mode 0 copy   a = b          (16 bytes read, 16 written per element)
//   mode 1 triad  a = b + s * c  (32 read, 16 written)
//   mode 2 read   a[block] = sum of b over the block (16 read; one double
//                 written per 1024 elements)
constexpr int kStreamU = 4;

template <int MODE>
__global__ __launch_bounds__(256) void k_stream(int64_t n2, const double2* __restrict__ b,
                                                const double2* __restrict__ c,
                                                double2* __restrict__ a, double s) {
    const int64_t base = (int64_t)blockIdx.x * (256 * kStreamU) + threadIdx.x;
    double2 vb[kStreamU], vc[kStreamU];
#pragma unroll
    for (int u = 0; u < kStreamU; ++u) {
        const int64_t i = base + u * 256;
        vb[u] = make_double2(0.0, 0.0);
        if (i < n2) {
            vb[u] = b[i];
            if (MODE == 1) vc[u] = c[i];
        }
    }
    if (MODE == 2) {
        __shared__ double sh[256];
        double t = 0.0;
#pragma unroll
        for (int u = 0; u < kStreamU; ++u) t += vb[u].x + vb[u].y;
        sh[threadIdx.x] = t;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) reinterpret_cast<double*>(a)[blockIdx.x] = sh[0];
        return;
    }
#pragma unroll
    for (int u = 0; u < kStreamU; ++u) {
        const int64_t i = base + u * 256;
        if (i < n2) {
            if (MODE == 1)
                a[i] = make_double2(vb[u].x + s * vc[u].x, vb[u].y + s * vc[u].y);
            else
                a[i] = vb[u];
        }
    }
}

}  // namespace fvhip

using namespace fvhip;

extern "C" {

const char* fv_last_error(void) { return g_last_error.c_str(); }
const char* fv_version(void) { return "fvhip 0.1 (gfx950)"; }

int fv_ctx_create(int device, size_t hbm_budget_bytes, fv_ctx** out) {
    return guarded([&] {
        FV_REQUIRE(out, "out is NULL");
        int n = 0;
        FV_HIP(hipGetDeviceCount(&n));
        FV_REQUIRE(device >= 0 && device < n, "device index out of range");
        FV_HIP(hipSetDevice(device));
        auto* c = new fv_ctx();
        c->device = device;
        c->budget = hbm_budget_bytes;
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            throw Error{FV_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e)};
        }
        {
            std::lock_guard<std::mutex> lk(g_ctx_mu);
            c->gen = g_next_gen++;
            g_live[c] = c->gen;
        }
        *out = c;
    });
}

int fv_ctx_destroy(fv_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        {
            std::lock_guard<std::mutex> lk(g_ctx_mu);
            g_live.erase(ctx);
        }
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto& kv : ctx->sizes) (void)hipFree(kv.first);
        ctx->sizes.clear();
        for (auto& kv : ctx->cache) (void)hipFree(kv.second);
        ctx->cache.clear();
        ctx->cached = 0;
        for (auto& e : ctx->events)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : ctx->pollEv)
            if (e) (void)hipEventDestroy(e);
        if (ctx->pollHost) (void)hipHostFree(ctx->pollHost);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int fv_sync(fv_ctx* ctx) {
    return guarded([&] {
        FV_REQUIRE(ctx, "ctx is NULL");
        FV_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int fv_event_record(fv_ctx* ctx, int slot) {
    return guarded([&] {
        FV_REQUIRE(ctx && slot >= 0 && slot < fv_ctx::kEvents, "bad event slot");
        if (!ctx->events[slot]) FV_HIP(hipEventCreate(&ctx->events[slot]));
        FV_HIP(hipEventRecord(ctx->events[slot], ctx->stream));
    });
}

int fv_event_elapsed(fv_ctx* ctx, int a, int b, float* ms) {
    return guarded([&] {
        FV_REQUIRE(ctx && ms && a >= 0 && b >= 0 && a < fv_ctx::kEvents && b < fv_ctx::kEvents,
                   "bad event slot");
        FV_REQUIRE(ctx->events[a] && ctx->events[b], "event slot not recorded");
        FV_HIP(hipEventSynchronize(ctx->events[b]));
        FV_HIP(hipEventElapsedTime(ms, ctx->events[a], ctx->events[b]));
    });
}

void* fv_ctx_stream(fv_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }
size_t fv_ctx_bytes_in_use(fv_ctx* ctx) { return ctx ? ctx->in_use : 0; }

int fv_device_alloc(fv_ctx* ctx, size_t bytes, void** d_out) {
    return guarded([&] {
        FV_REQUIRE(ctx && d_out, "NULL argument");
        *d_out = ctx_alloc(ctx, bytes);
    });
}

int fv_device_free(fv_ctx* ctx, void* d_ptr) {
    return guarded([&] {
        FV_REQUIRE(ctx, "ctx is NULL");
        if (d_ptr) {
            FV_REQUIRE(ctx->sizes.count(d_ptr), "pointer was not allocated by this context");
            ctx_free(ctx, d_ptr);
        }
    });
}

int fv_memcpy_h2d(fv_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
    return guarded([&] {
        FV_REQUIRE(ctx, "ctx is NULL");
        if (!bytes) return;
        FV_REQUIRE(d_dst && h_src, "NULL pointer");
        FV_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
        FV_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int fv_memcpy_d2h(fv_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
    return guarded([&] {
        FV_REQUIRE(ctx, "ctx is NULL");
        if (!bytes) return;
        FV_REQUIRE(h_dst && d_src, "NULL pointer");
        FV_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        FV_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int fv_memcpy_d2d(fv_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
    return guarded([&] {
        FV_REQUIRE(ctx, "ctx is NULL");
        if (!bytes) return;
        FV_REQUIRE(d_dst && d_src, "NULL pointer");
        FV_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    });
}

int fv_fill(fv_ctx* ctx, fv_prec prec, void* d_ptr, int64_t n, double value) {
    return guarded([&] {
        FV_REQUIRE(ctx && n >= 0, "bad argument");
        if (!n) return;
        FV_REQUIRE(d_ptr, "NULL pointer");
        int g = grid_for(n, 256, 4096);
        if (prec == FV_F64)
            k_fill<double><<<g, 256, 0, ctx->stream>>>((double*)d_ptr, n, value);
        else
            k_fill<float><<<g, 256, 0, ctx->stream>>>((float*)d_ptr, n, (float)value);
        FV_LAUNCH_CHECK();
    });
}

int fv_convert(fv_ctx* ctx, fv_prec to, void* d_dst, fv_prec from, const void* d_src, int64_t n) {
    return guarded([&] {
        FV_REQUIRE(ctx && n >= 0, "bad argument");
        if (!n) return;
        FV_REQUIRE(d_dst && d_src, "NULL pointer");
        int g = grid_for(n, 256, 4096);
        if (to == FV_F32 && from == FV_F64)
            k_convert<float, double><<<g, 256, 0, ctx->stream>>>((float*)d_dst, (const double*)d_src, n);
        else if (to == FV_F64 && from == FV_F32)
            k_convert<double, float><<<g, 256, 0, ctx->stream>>>((double*)d_dst, (const float*)d_src, n);
        else if (to == FV_F64)
            k_convert<double, double><<<g, 256, 0, ctx->stream>>>((double*)d_dst, (const double*)d_src, n);
        else
            k_convert<float, float><<<g, 256, 0, ctx->stream>>>((float*)d_dst, (const float*)d_src, n);
        FV_LAUNCH_CHECK();
    });
}

int fv_permute(fv_ctx* ctx, fv_prec prec, int64_t n, int ncomp, int mode, int oriented,
               const int32_t* map, const void* src, void* dst) {
    return guarded([&] {
        FV_REQUIRE(ctx && n >= 0 && ncomp >= 1 && (mode == 0 || mode == 1), "bad argument");
        if (!n) return;
        FV_REQUIRE(map && src && dst && src != dst, "NULL or aliased pointer");
        FV_REQUIRE(n <= (int64_t)0xFFFFFFFF * 256 * kPermU, "too many elements");
        if (prec == FV_F64)
            launch_permute<double>(ctx->stream, n, ncomp, mode, oriented, map,
                                   (const double*)src, (double*)dst);
        else
            launch_permute<float>(ctx->stream, n, ncomp, mode, oriented, map, (const float*)src,
                                  (float*)dst);
        FV_LAUNCH_CHECK();
    });
}

int fv_component(fv_ctx* ctx, fv_prec prec, int64_t n, int ncomp, int comp, int mode,
                 const void* d_src, void* d_dst) {
    return guarded([&] {
        FV_REQUIRE(ctx && n >= 0 && ncomp >= 1 && comp >= 0 && comp < ncomp &&
                       (mode == 0 || mode == 1),
                   "bad argument");
        if (!n) return;
        FV_REQUIRE(d_src && d_dst && d_src != d_dst, "NULL or aliased pointer");
        const int g = grid_for(n, 256, 8192);
        if (prec == FV_F64)
            k_component<double><<<g, 256, 0, ctx->stream>>>(n, ncomp, comp, mode,
                                                            (const double*)d_src, (double*)d_dst);
        else
            k_component<float><<<g, 256, 0, ctx->stream>>>(n, ncomp, comp, mode,
                                                           (const float*)d_src, (float*)d_dst);
        FV_LAUNCH_CHECK();
    });
}

int fv_stream(fv_ctx* ctx, int mode, int64_t n, const double* d_b, const double* d_c, double* d_a,
              double s) {
    return guarded([&] {
        FV_REQUIRE(ctx && n >= 0 && mode >= 0 && mode <= 2, "bad argument");
        FV_REQUIRE(n % 2 == 0, "n must be even (16-byte elements)");
        if (!n) return;
        FV_REQUIRE(d_a && d_b && (mode == 0 || d_c), "NULL pointer");
        FV_REQUIRE(((uintptr_t)d_a | (uintptr_t)d_b | (uintptr_t)(mode ? d_c : d_b)) % 16 == 0,
                   "arrays must be 16-byte aligned");
        const int64_t n2 = n / 2;
        const int64_t g = (n2 + 256 * kStreamU - 1) / (256 * kStreamU);
        FV_REQUIRE(g < (int64_t(1) << 31), "array too long");
        if (mode == 0)
            k_stream<0><<<(unsigned)g, 256, 0, ctx->stream>>>(n2, (const double2*)d_b, nullptr,
                                                           (double2*)d_a, s);
        else if (mode == 2)
            k_stream<2><<<(unsigned)g, 256, 0, ctx->stream>>>(n2, (const double2*)d_b, nullptr,
                                                           (double2*)d_a, s);
        else
            k_stream<1><<<(unsigned)g, 256, 0, ctx->stream>>>(n2, (const double2*)d_b,
                                                           (const double2*)d_c, (double2*)d_a, s);
        FV_LAUNCH_CHECK();
    });
}

int fv_fill_uniform(fv_ctx* ctx, fv_prec prec, void* d_ptr, int64_t n, double lo, double hi,
                    uint64_t seed, uint64_t stream) {
    return guarded([&] {
        FV_REQUIRE(ctx && n >= 0, "bad argument");
        if (!n) return;
        FV_REQUIRE(d_ptr, "NULL pointer");
        int g = grid_for(n, 256, 8192);
        if (prec == FV_F64)
            k_uniform<double><<<g, 256, 0, ctx->stream>>>((double*)d_ptr, n, lo, hi, seed, stream);
        else
            k_uniform<float><<<g, 256, 0, ctx->stream>>>((float*)d_ptr, n, lo, hi, seed, stream);
        FV_LAUNCH_CHECK();
    });
}

}  // extern "C"
