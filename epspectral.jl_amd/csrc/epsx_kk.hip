// Kramers-Kronig on the device: the spline KK matrix (src/frequency_grids.jl:5-44) built by
// one thread per entry, and the FP64 MFMA GEMM that applies it (src/scgd0.jl:289-292).
#include <cmath>
#include <vector>

#include "epsx_common.h"

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

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16;

// C[m, ncol] (+)= A[m, k] B[k, ncol], column-major FP64.  256 threads = 4 waves in a 2x2
// arrangement; each wave owns a 32x32 sub-tile as 2x2 v_mfma_f64_16x16x4 accumulators.
// Operand tiles are staged through LDS (k-major, padded against bank conflicts); edges are
// zero-filled, so any shape is valid.
__global__ __launch_bounds__(256) void dgemm_nn_kernel(int64_t m, int64_t ncol, int64_t k,
                                                       const double *__restrict__ A, int64_t lda,
                                                       const double *__restrict__ B, int64_t ldb,
                                                       double *__restrict__ C, int64_t ldc,
                                                       int accumulate, const int32_t *__restrict__ gate)
{
    if (gate && *gate == 0) return;  // a masked (converged) instance
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t row0 = (int64_t)blockIdx.x * BM, col0 = (int64_t)blockIdx.y * BN;
    double4_t acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = (double4_t){0, 0, 0, 0};

    for (int64_t k0 = 0; k0 < k; k0 += BK) {
        // A tile: 64 x 16, thread loads rows tid % 64, k = tid / 64 + 4 r
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int ii = tid & 63, kk = (tid >> 6) + 4 * r;
            const int64_t gi = row0 + ii, gk = k0 + kk;
            As[kk][ii] = (gi < m && gk < k) ? A[gi + lda * gk] : 0.0;
        }
        // B tile: 16 x 64, thread loads k = tid % 16, column tid / 16 + 16 r
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int kk = tid & 15, jj = (tid >> 4) + 16 * r;
            const int64_t gk = k0 + kk, gj = col0 + jj;
            Bs[kk][jj] = (gk < k && gj < ncol) ? B[gk + ldb * gj] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 4) {
            const int kk = ks + (lane >> 4);
            double a0 = As[kk][wm * 32 + (lane & 15)];
            double a1 = As[kk][wm * 32 + 16 + (lane & 15)];
            double b0 = Bs[kk][wn * 32 + (lane & 15)];
            double b1 = Bs[kk][wn * 32 + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    // D[i][j] of a 16x16 f64 tile: element v of lane l holds i = 4 v + l / 16, j = l % 16
    // (one row per register group, four input blocks)
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t gi = row0 + wm * 32 + a * 16 + 4 * v + (lane >> 4);
                const int64_t gj = col0 + wn * 32 + b * 16 + (lane & 15);
                if (gi < m && gj < ncol) {
                    double *c = C + gi + ldc * gj;
                    *c = accumulate ? *c + acc[a][b][v] : acc[a][b][v];
                }
            }
}

}  // namespace

int kk_build(epsx_ctx *ctx, int64_t n, const double *d_ws, double *d_K)
{
    dim3 grid((unsigned)((n + 255) / 256), (unsigned)n);
    hipLaunchKernelGGL(kk_build_kernel, grid, dim3(256), 0, ctx->stream, n, d_ws, d_K);
    EPSX_HIP(ctx, hipGetLastError());
    return EPSX_OK;
}

int dgemm_nn(epsx_ctx *ctx, int64_t m, int64_t ncol, int64_t k, const double *A, int64_t lda,
             const double *B, int64_t ldb, double *C, int64_t ldc, bool accumulate, const int32_t *gate)
{
    if (m <= 0 || ncol <= 0) return EPSX_OK;
    dim3 grid((unsigned)((m + BM - 1) / BM), (unsigned)((ncol + BN - 1) / BN));
    hipLaunchKernelGGL(dgemm_nn_kernel, grid, dim3(256), 0, ctx->stream, m, ncol, k, A, lda, B,
                       ldb, C, ldc, accumulate ? 1 : 0, gate);
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

}  // extern "C"
