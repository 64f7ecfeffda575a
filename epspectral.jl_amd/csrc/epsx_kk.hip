// Kramers-Kronig on the device: the spline KK matrix (src/frequency_grids.jl:5-44) built by
// one thread per entry, applied by the FP64 MFMA GEMM (src/scgd0.jl:289-292).
#include <cmath>
#include <vector>

#include "epsx_common.h"
#include "epsx_device.h"

namespace epsx {
namespace {

__device__ inline double xlogx(double x) { return x == 0.0 ? 0.0 : x * log(fabs(x)); }

// K[i, j] = linear_spline_kramers_kronig(xs[i+1], xs[j], xs[j+1], xs[j+2]) / π on the grid
// padded by one end cell per side (src/frequency_grids.jl:30-44).
__global__ void kk_build_kernel(int64_t n, const double *__restrict__ ws, double *__restrict__ K)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= n) return;
    auto xs = [&](int64_t q) -> double {  // padded grid, q in 0..n+1
        if (q == 0) return ws[0] - (ws[1] - ws[0]);
        if (q == n + 1) return ws[n - 1] + (ws[n - 1] - ws[n - 2]);
        return ws[q - 1];
    };
    const double x = xs(i + 1), x1 = xs(j), x2 = xs(j + 1), x3 = xs(j + 2);
    const double v = xlogx(x1 - x) / (x2 - x1) - xlogx(x2 - x) * (x3 - x1) / (x2 - x1) / (x3 - x2) +
                     xlogx(x3 - x) / (x3 - x2);
    K[i + n * j] = v / 3.141592653589793;
}

// Ko = K with a zero diagonal, Kd = diag(K): the principal-value self-term kept apart so a
// low-precision product can leave it in FP64.
__global__ void kk_split_kernel(int64_t n, const double *__restrict__ K, double *__restrict__ Ko,
                                double *__restrict__ Kd)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n * n) return;
    const int64_t i = q % n, j = q / n;
    Ko[q] = i == j ? 0.0 : K[q];
    if (i == j) Kd[i] = K[q];
}

// The off-diagonal KK in low precision, transposed for the k-contiguous product:
// out[k + np m] = Ko[m + n k] (zero for n <= k < np), E = BF16 bits or float.
template <typename E>
__global__ void kk_lp_kernel(int64_t n, int64_t np, const double *__restrict__ Ko, E *__restrict__ out)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= np * n) return;
    const int64_t k = q % np, m = q / np;
    const double v = k < n ? Ko[m + n * k] : 0.0;
    if constexpr (sizeof(E) == 2) out[q] = to_bf16(v);
    else out[q] = (float)v;
}

// Columns of the slots at one precision level in low precision, rows padded to np:
// dst[r + np c + dstride t] = src[r + ld c + sstride t] (zero for nrow <= r < np), t = tidx[z].
template <typename E>
__global__ void lp_cols_kernel(int64_t nrow, int64_t np, int64_t ncol, const double *__restrict__ src, int64_t ld,
                               int64_t sstride, E *__restrict__ dst, int64_t dstride, const int32_t *__restrict__ tidx,
                               const int32_t *__restrict__ gate)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (q >= np * ncol) return;
    const int64_t t = tidx[z];
    if (gate && !gate[t]) return;
    const int64_t r = q % np, c = q / np;
    const double v = r < nrow ? src[r + ld * c + sstride * t] : 0.0;
    if constexpr (sizeof(E) == 2) dst[q + dstride * t] = to_bf16(v);
    else dst[q + dstride * t] = (float)v;
}

}  // namespace

int kk_lp_build(epsx_ctx *ctx, int precision, int64_t n, int64_t np, const double *d_Ko, void *out)
{
    const dim3 grid((unsigned)((np * n + 255) / 256)), blk(256);
    if (precision == EPSX_PREC_BF16)
        hipLaunchKernelGGL(kk_lp_kernel<uint16_t>, grid, blk, 0, ctx->stream, n, np, d_Ko, static_cast<uint16_t *>(out));
    else
        hipLaunchKernelGGL(kk_lp_kernel<float>, grid, blk, 0, ctx->stream, n, np, d_Ko, static_cast<float *>(out));
    EPSX_HIP(ctx, hipGetLastError());
    return EPSX_OK;
}

int lp_cols(epsx_ctx *ctx, int precision, int64_t nrow, int64_t np, int64_t ncol, const double *src, int64_t ld,
            int64_t sstride, void *dst, int64_t dstride, int nslots, const int32_t *tidx, const int32_t *gate)
{
    const dim3 grid((unsigned)((np * ncol + 255) / 256), (unsigned)nslots), blk(256);
    if (precision == EPSX_PREC_BF16)
        hipLaunchKernelGGL(lp_cols_kernel<uint16_t>, grid, blk, 0, ctx->stream, nrow, np, ncol, src, ld, sstride,
                           static_cast<uint16_t *>(dst), dstride, tidx, gate);
    else
        hipLaunchKernelGGL(lp_cols_kernel<float>, grid, blk, 0, ctx->stream, nrow, np, ncol, src, ld, sstride,
                           static_cast<float *>(dst), dstride, tidx, gate);
    EPSX_HIP(ctx, hipGetLastError());
    return EPSX_OK;
}

int kk_split(epsx_ctx *ctx, int64_t n, const double *d_K, double *d_Ko, double *d_Kd)
{
    hipLaunchKernelGGL(kk_split_kernel, dim3((unsigned)((n * n + 255) / 256)), dim3(256), 0, ctx->stream, n, d_K,
                       d_Ko, d_Kd);
    EPSX_HIP(ctx, hipGetLastError());
    return EPSX_OK;
}

// The fixed device workspace of epsx_kramers_kronig_noalloc: the KK matrix (n² doubles) and, behind
// it, the product's split-k partials.  40 MiB per device, held until the process ends; every call on
// one device shares it.  Host code never uses this name as an address: epsx_open looks up this
// device's copy with hipGetSymbolAddress and keeps it in the context (kk_scratch).
constexpr int64_t KK_SCRATCH_DOUBLES = int64_t(5) << 20;
constexpr int64_t KK_NOALLOC_MAX_N = 2048;   // n² <= 4 Mi doubles, the rest for partials
static_assert(KK_NOALLOC_MAX_N * KK_NOALLOC_MAX_N < KK_SCRATCH_DOUBLES, "workspace too small");
static __device__ __align__(256) double scratch[KK_SCRATCH_DOUBLES];

int kk_scratch_init(epsx_ctx *ctx)
{
    void *p = nullptr;
    EPSX_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ctx->kk_scratch = static_cast<double *>(p);
    return EPSX_OK;
}

int kk_build(epsx_ctx *ctx, int64_t n, const double *d_ws, double *d_K)
{
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)n);
    hipLaunchKernelGGL(kk_build_kernel, grid, dim3(256), 0, ctx->stream, n, d_ws, d_K);
    EPSX_HIP(ctx, hipGetLastError());
    return EPSX_OK;
}

}  // namespace epsx

extern "C" {

int epsx_kramers_kronig_matrix(epsx_ctx *ctx, int64_t n, const double *ws, double *K)
{
    if (!ctx) return EPSX_ERR_ARG;
    if (n < 2 || !ws || !K)
        return epsx::set_error(ctx, EPSX_ERR_ARG, "kramers_kronig_matrix needs at least two nodes");
    EPSX_HIP(ctx, hipSetDevice(ctx->device));
    epsx::Tmp dws, dK;
    EPSX_TRY(epsx::tmp_upload(ctx, dws, ws, sizeof(double) * n));
    EPSX_TRY(epsx::tmp_alloc(ctx, dK.b, sizeof(double) * n * n));
    EPSX_TRY(epsx::kk_build(ctx, n, dws.ptr<double>(), dK.ptr<double>()));
    EPSX_HIP(ctx, hipMemcpyAsync(K, dK.b.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EPSX_OK;
}

int epsx_kramers_kronig(epsx_ctx *ctx, int64_t n, const double *ws, int64_t ncol, const double *zs,
                        double *ys)
{
    if (!ctx) return EPSX_ERR_ARG;
    if (n < 2 || ncol < 0 || !ws || (ncol && (!zs || !ys)))
        return epsx::set_error(ctx, EPSX_ERR_ARG, "kramers_kronig: bad arguments");
    if (ncol == 0) return EPSX_OK;
    EPSX_HIP(ctx, hipSetDevice(ctx->device));
    epsx::Tmp dws, dK, dz, dy;
    EPSX_TRY(epsx::tmp_upload(ctx, dws, ws, sizeof(double) * n));
    EPSX_TRY(epsx::tmp_alloc(ctx, dK.b, sizeof(double) * n * n));
    EPSX_TRY(epsx::tmp_upload(ctx, dz, zs, sizeof(double) * n * ncol));
    EPSX_TRY(epsx::tmp_alloc(ctx, dy.b, sizeof(double) * n * ncol));
    EPSX_TRY(epsx::kk_build(ctx, n, dws.ptr<double>(), dK.ptr<double>()));
    EPSX_TRY(epsx::dgemm_nn(ctx, n, ncol, n, dK.ptr<double>(), n, dz.ptr<double>(), n,
                            dy.ptr<double>(), n, false));
    EPSX_HIP(ctx, hipMemcpyAsync(ys, dy.b.p, sizeof(double) * n * ncol, hipMemcpyDeviceToHost, ctx->stream));
    EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EPSX_OK;
}

int epsx_kramers_kronig_noalloc(epsx_ctx *ctx, int64_t n, const double *d_ws, int64_t ncol,
                                const double *d_zs, double *d_ys, void *stream)
{
    if (!ctx) return EPSX_ERR_ARG;
    if (n < 2 || ncol < 0 || !d_ws || (ncol && (!d_zs || !d_ys)))
        return epsx::set_error(ctx, EPSX_ERR_ARG, "kramers_kronig_noalloc: bad arguments");
    if (n > epsx::KK_NOALLOC_MAX_N)
        return epsx::set_error(ctx, EPSX_ERR_ARG,
                               "kramers_kronig_noalloc: %lld nodes exceed the fixed workspace's limit of %lld",
                               (long long)n, (long long)epsx::KK_NOALLOC_MAX_N);
    if (!ctx->kk_scratch) return epsx::set_error(ctx, EPSX_ERR_STATE, "kramers_kronig_noalloc: no workspace");
    if (ncol == 0) return EPSX_OK;
    EPSX_HIP(ctx, hipSetDevice(ctx->device));
    // the launches go to the caller's stream for this call (every launcher uses ctx->stream)
    struct StreamSwap {
        epsx_ctx *c; hipStream_t saved;
        ~StreamSwap() { c->stream = saved; }
    } swap{ctx, ctx->stream};
    if (stream) ctx->stream = static_cast<hipStream_t>(stream);
    double *K = ctx->kk_scratch;
    EPSX_TRY(epsx::kk_build(ctx, n, d_ws, K));
    epsx::GemmBatch bt;
    bt.tune = false;   // no timing runs: the tuned tile shape if one is known, else the default
    bt.part = K + n * n;
    bt.part_cap = epsx::KK_SCRATCH_DOUBLES - n * n;
    return epsx::dgemm(ctx, n, ncol, n, K, n, d_zs, n, d_ys, n, false, bt);
}

}  // extern "C"
