// common.hpp -- shared internals of libfvhip: context, error plumbing, device
// buffers and the counter-based RNG used for synthetic inputs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
#include <mutex>
#include <map>
#include <unordered_map>

#include "../../include/fvhip.h"

namespace fvhip {

// ---- error plumbing --------------------------------------------------------
void set_error(const std::string& msg);
void clear_error();

struct Error {
    int code;
    std::string msg;
};

#define FV_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::fvhip::Error{FV_EDEVICE, std::string(#call) + ": " +                 \
                                                 hipGetErrorString(e_)};                \
    } while (0)

#define FV_REQUIRE(cond, msg)                                                           \
    do {                                                                                \
        if (!(cond)) throw ::fvhip::Error{FV_EINVAL, std::string(msg)};                 \
    } while (0)

// Run f() and translate exceptions into a status code + thread-local message.
template <typename F>
int guarded(F&& f) {
    try {
        clear_error();
        f();
        return FV_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return FV_EINVAL;
    } catch (const std::exception& e) {
        set_error(e.what());
        return FV_EINVAL;
    }
}

// ---- counter-based RNG (splitmix64 finaliser over a keyed counter) --------
__host__ __device__ inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__host__ __device__ inline uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(mix64(seed ^ mix64(stream)) + i);
}
__host__ __device__ inline double rng_u01(uint64_t seed, uint64_t stream, uint64_t i) {
    return (double)(rng_u64(seed, stream, i) >> 11) * (1.0 / 9007199254740992.0);
}

// ---- context -----------------------------------------------------------------
}  // namespace fvhip

struct fv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t budget = 0;  // 0 = unlimited
    size_t in_use = 0;                                       // live bytes (as requested)
    std::unordered_map<void*, std::pair<size_t, size_t>> sizes;  // live: requested, block size
    std::multimap<size_t, void*> cache;                      // freed blocks by block size
    size_t cached = 0;                                       // bytes held in the cache
    std::mutex mu;
    static constexpr int kEvents = 64;
    hipEvent_t events[kEvents] = {};
    uint64_t gen = 0;  // unique per context of the process
    // pinned host area and events for polls that do not stall the stream
    // (solver state copied back while the next iterations are queued)
    static constexpr size_t kPollBytes = 4096;
    void* pollHost = nullptr;
    hipEvent_t pollEv[2] = {};
};

namespace fvhip {

// Allocate device memory accounted against the context's HBM budget.
void* ctx_alloc(fv_ctx* ctx, size_t bytes);
void ctx_free(fv_ctx* ctx, void* p);
// Generation of a live context (0 if it was destroyed); a buffer released after
// its context is gone was freed with the context and is dropped.
uint64_t ctx_generation(fv_ctx* ctx);
// The context's pinned poll area (kPollBytes) and its two events, made on first use.
void* ctx_poll_host(fv_ctx* ctx);

// RAII device buffer owned by a library handle.
template <typename T>
struct DBuf {
    fv_ctx* ctx = nullptr;
    uint64_t gen = 0;
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : ctx(o.ctx), gen(o.gen), p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            release();
            ctx = o.ctx;
            gen = o.gen;
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(fv_ctx* c, size_t count) {
        release();
        ctx = c;
        gen = ctx_generation(c);
        n = count;
        p = count ? static_cast<T*>(ctx_alloc(c, count * sizeof(T))) : nullptr;
    }
    void upload(fv_ctx* c, const T* h, size_t count) {
        alloc(c, count);
        if (count) FV_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
    }
    void download(T* h) const {
        if (n) {
            FV_HIP(hipMemcpyAsync(h, p, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
            FV_HIP(hipStreamSynchronize(ctx->stream));
        }
    }
    void release() {
        if (p && ctx && gen && ctx_generation(ctx) == gen) ctx_free(ctx, p);
        p = nullptr;
        n = 0;
    }
};

inline int grid_for(int64_t n, int block, int max_blocks = 1 << 30) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > max_blocks) g = max_blocks;
    return (int)g;
}

#define FV_LAUNCH_CHECK() FV_HIP(hipGetLastError())

}  // namespace fvhip
