// Context, errors and device buffers of libepsx.
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "epsx_common.h"

namespace epsx {

int set_error(epsx_ctx *ctx, int code, const char *fmt, ...)
{
    if (!ctx) return code;
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->err = buf;
    return code;
}

int hip_fail(epsx_ctx *ctx, hipError_t e, const char *what, const char *file, int line)
{
    int code = (e == hipErrorOutOfMemory) ? EPSX_ERR_OOM : EPSX_ERR_DEVICE;
    return set_error(ctx, code, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e),
                     what, file, line);
}

// A context keeps at most this much released temporary memory for reuse.
constexpr size_t POOL_MAX = size_t(32) << 30;

static void raw_free(epsx_ctx *ctx, void *p, size_t bytes)
{
    (void)hipFree(p);
    if (ctx) ctx->dev_bytes -= bytes;
}

void dev_free(DevBuf &b)
{
    if (b.p && b.tmp && b.owner && b.owner->pool_bytes + b.bytes <= POOL_MAX) {
        b.owner->pool.emplace(b.bytes, b.p);
        b.owner->pool_bytes += b.bytes;
    } else if (b.p) {
        raw_free(b.owner, b.p, b.bytes);
    }
    b.p = nullptr;
    b.bytes = 0;
    b.owner = nullptr;
    b.tmp = false;
}

void pool_clear(epsx_ctx *ctx)
{
    for (auto &kv : ctx->pool) raw_free(ctx, kv.second, kv.first);
    ctx->pool.clear();
    ctx->pool_bytes = 0;
}

static hipError_t try_malloc(epsx_ctx *ctx, void **p, size_t bytes)
{
    if (ctx->alloc_cap && ctx->dev_bytes + bytes > ctx->alloc_cap) return hipErrorOutOfMemory;
    return hipMalloc(p, bytes);
}

int ctx_malloc(epsx_ctx *ctx, void **p, size_t bytes)
{
    *p = nullptr;
    hipError_t e = try_malloc(ctx, p, bytes);
    if (e != hipSuccess && !ctx->pool.empty()) {
        // the pool's released buffers may be what is missing: a released buffer may still be read
        // by work queued on either stream, so wait for both before freeing
        (void)hipGetLastError();
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->side) (void)hipStreamSynchronize(ctx->side);
        pool_clear(ctx);
        e = try_malloc(ctx, p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return set_error(ctx, EPSX_ERR_OOM, "device allocation of %zu bytes failed: %s", bytes,
                         hipGetErrorString(e));
    }
    ctx->dev_bytes += bytes;
    return EPSX_OK;
}

int dev_alloc(epsx_ctx *ctx, DevBuf &b, size_t bytes)
{
    if (b.p && b.bytes >= bytes) return EPSX_OK;
    dev_free(b);
    if (bytes == 0) bytes = 8;
    EPSX_TRY(ctx_malloc(ctx, &b.p, bytes));
    b.bytes = bytes;
    b.owner = ctx;
    b.tmp = false;
    bool seen = false;
    for (auto *q : ctx->owned) seen |= (q == &b);
    if (!seen) ctx->owned.push_back(&b);
    return EPSX_OK;
}

int tmp_alloc(epsx_ctx *ctx, DevBuf &b, size_t bytes)
{
    dev_free(b);
    if (bytes == 0) bytes = 8;
    const auto hit = ctx->pool.find(bytes);
    if (hit != ctx->pool.end()) {
        b.p = hit->second;
        ctx->pool.erase(hit);
        ctx->pool_bytes -= bytes;
    } else {
        EPSX_TRY(ctx_malloc(ctx, &b.p, bytes));
    }
    b.bytes = bytes;
    b.owner = ctx;
    b.tmp = true;
    return EPSX_OK;
}

int tmp_upload(epsx_ctx *ctx, Tmp &t, const void *host, size_t bytes)
{
    EPSX_TRY(tmp_alloc(ctx, t.b, bytes));
    if (bytes) EPSX_HIP(ctx, hipMemcpyAsync(t.b.p, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    return EPSX_OK;
}

void parallel_copy(void *dst, const void *src, size_t n)
{
    unsigned hw = std::thread::hardware_concurrency();
    hw = hw == 0 ? 1 : (hw > 8 ? 8 : hw);
    const size_t want = n >> 20;
    const unsigned nt = want < 1 ? 1u : (want < hw ? (unsigned)want : hw);
    const size_t per = (n + nt - 1) / nt;
    char *d = static_cast<char *>(dst);
    const char *h = static_cast<const char *>(src);
    std::vector<std::thread> th;
    for (unsigned k = 1; k < nt; k++) {
        const size_t a = per * k, b = a + per < n ? a + per : n;
        if (a < b) th.emplace_back([=] { std::memcpy(d + a, h + a, b - a); });
    }
    std::memcpy(d, h, per < n ? per : n);
    for (auto &t : th) t.join();
}

int read_back(epsx_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    constexpr size_t STAGE = size_t(64) << 20;
    const size_t stage = bytes < STAGE ? bytes : STAGE;
    if (bytes == 0) return EPSX_OK;
    if (ctx->pin_bytes < stage) {
        if (ctx->pin) (void)hipHostFree(ctx->pin);
        ctx->pin = nullptr;
        ctx->pin_bytes = 0;
        EPSX_HIP(ctx, hipHostMalloc(&ctx->pin, stage, hipHostMallocDefault));
        ctx->pin_bytes = stage;
    }
    for (size_t off = 0; off < bytes; off += stage) {
        const size_t n = bytes - off < stage ? bytes - off : stage;
        EPSX_HIP(ctx, hipMemcpyAsync(ctx->pin, static_cast<const char *>(src) + off, n, hipMemcpyDeviceToHost,
                                     ctx->stream));
        EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        parallel_copy(static_cast<char *>(dst) + off, ctx->pin, n);
    }
    return EPSX_OK;
}

int write_from_host(epsx_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    constexpr size_t STAGE = size_t(64) << 20;
    const size_t stage = bytes < STAGE ? bytes : STAGE;
    if (bytes == 0) return EPSX_OK;
    // the staging buffer may still feed an earlier copy on the stream
    EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->pin_bytes < stage) {
        if (ctx->pin) (void)hipHostFree(ctx->pin);
        ctx->pin = nullptr;
        ctx->pin_bytes = 0;
        EPSX_HIP(ctx, hipHostMalloc(&ctx->pin, stage, hipHostMallocDefault));
        ctx->pin_bytes = stage;
    }
    for (size_t off = 0; off < bytes; off += stage) {
        const size_t n = bytes - off < stage ? bytes - off : stage;
        parallel_copy(ctx->pin, static_cast<const char *>(src) + off, n);
        EPSX_HIP(ctx, hipMemcpyAsync(static_cast<char *>(dst) + off, ctx->pin, n, hipMemcpyHostToDevice, ctx->stream));
        EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return EPSX_OK;
}

int upload(epsx_ctx *ctx, DevBuf &b, const void *host, size_t bytes)
{
    EPSX_TRY(dev_alloc(ctx, b, bytes));
    if (bytes) EPSX_HIP(ctx, hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    return EPSX_OK;
}

}  // namespace epsx

extern "C" {

int epsx_abi_version(void) { return EPSX_ABI_VERSION; }

int epsx_open(int device, epsx_ctx **out)
{
    if (!out) return EPSX_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return EPSX_ERR_DEVICE;
    if (device < 0 || device >= n) return EPSX_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return EPSX_ERR_DEVICE;
    auto *ctx = new epsx_ctx();
    ctx->device = device;
    if (const char *cap = getenv("EPSX_ALLOC_CAP_MB")) ctx->alloc_cap = size_t(atoll(cap)) << 20;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return EPSX_ERR_DEVICE;
    }
    if (epsx::kk_scratch_init(ctx) != EPSX_OK) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return EPSX_ERR_DEVICE;
    }
    *out = ctx;
    return EPSX_OK;
}

void epsx_close(epsx_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    epsx::blas_destroy(ctx);
    epsx::blaslt_destroy(ctx);
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    if (ctx->side) {
        (void)hipStreamSynchronize(ctx->side);
        (void)hipStreamDestroy(ctx->side);
    }
    if (ctx->pin_side) (void)hipHostFree(ctx->pin_side);
    if (ctx->pin_slots) (void)hipHostFree(ctx->pin_slots);
    if (ctx->pin_prec) (void)hipHostFree(ctx->pin_prec);
    if (ctx->pin_book) (void)hipHostFree(ctx->pin_book);
    for (auto &e : ctx->run_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto *b : ctx->owned) epsx::dev_free(*b);
    epsx::pool_clear(ctx);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *epsx_last_error(const epsx_ctx *ctx) { return ctx ? ctx->err.c_str() : ""; }

int epsx_device_memory(const epsx_ctx *ctx, int64_t *held, int64_t *pooled)
{
    if (!ctx) return EPSX_ERR_ARG;
    if (held) *held = (int64_t)ctx->dev_bytes;
    if (pooled) *pooled = (int64_t)ctx->pool_bytes;
    return EPSX_OK;
}

int epsx_last_timings(const epsx_ctx *ctx, double *seconds8, int64_t *launches8)
{
    if (!ctx) return EPSX_ERR_ARG;
    for (int k = 0; k < 8; k++) {
        if (seconds8) seconds8[k] = ctx->t_class[k];
        if (launches8) launches8[k] = ctx->n_class[k];
    }
    return EPSX_OK;
}

}  // extern "C"
