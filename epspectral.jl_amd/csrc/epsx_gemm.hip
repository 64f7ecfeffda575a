// FP64 GEMM on the MFMA units (v_mfma_f64_16x16x4f64) for the dense contractions of the path:
// Kramers-Kronig (ReΣ = KK·ImΣ, src/scgd0.jl:289-292), the star Fan-Migdal product h·G̃, and
// the ladder's shift-table products (src/ladder.jl:283-290, src/ladder_kernel.jl:296-322).
//
// Layout: column-major operands, C[m, n] (+)= A[m, k] B[k, n], for a batch of instances that may
// share an operand (stride 0) and may be picked through an index array (the active temperatures).
// One workgroup of 4 waves owns a 64x64 tile of one instance and one split of the k range:
//   * operand tiles of 64x32 are staged through LDS; the next tile's global loads are issued
//     into registers before the current tile's MFMAs, so HBM/L2 latency hides behind compute;
//   * each wave owns a 32x32 sub-tile as 2x2 accumulators of 16x16;
//   * split-k partials go to a workspace and are summed in ascending split order by a second
//     kernel, so results are bitwise reproducible;
//   * workgroups are enumerated XCD-major (dispatch sends workgroup b to XCD b mod 8), so the
//     tiles one XCD runs are neighbours that share B panels in that XCD's L2.
#include "epsx_common.h"

namespace epsx {
namespace {

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int GM = 64, GN = 64, GK = 32, GT = 256, NXCD = 8;

struct GemmArgs {
    int64_t m, n, k, lda, ldb, ldc, sA, sB, sC;
    const double *A, *B;
    double *C;
    double2 *Z;           // complex epilogue: Z = complex(result, Zim), C unused
    const double *Zim;
    double *part;         // split-k partials [m, n] per (instance, split)
    const int32_t *idx;   // instance -> operand index for the operands flagged in imap
    const int32_t *gate;  // per operand index: 0 = skip
    int imap;             // bit 0: A, bit 1: B, bit 2: C/Z/Zim
    int S, mt, nt, ktiles, accumulate;
    int64_t units;
};

__device__ inline int64_t opi(const GemmArgs &p, int64_t z, int bit) { return (p.idx && (p.imap & bit)) ? p.idx[z] : z; }

__device__ inline void store_out(const GemmArgs &p, int64_t zc, int64_t gi, int64_t gj, double v)
{
    const int64_t o = gi + p.ldc * gj + p.sC * zc;
    if (p.Z) {
        p.Z[o] = make_double2(v, p.Zim[o]);
    } else {
        double *c = p.C + o;
        *c = p.accumulate ? *c + v : v;
    }
}

__global__ __launch_bounds__(GT, 2) void dgemm_kernel(GemmArgs p)
{
    const int64_t per_xcd = (int64_t)gridDim.x / NXCD;
    const int64_t v = (int64_t)(blockIdx.x % NXCD) * per_xcd + blockIdx.x / NXCD;
    if (v >= p.units) return;
    int64_t r = v;
    const int tm = (int)(r % p.mt); r /= p.mt;
    const int tn = (int)(r % p.nt); r /= p.nt;
    const int s = (int)(r % p.S); r /= p.S;
    const int64_t z = r;
    if (p.gate && p.gate[p.idx ? p.idx[z] : z] == 0) return;  // a masked (converged) instance
    const double *A = p.A + p.sA * opi(p, z, 1);
    const double *B = p.B + p.sB * opi(p, z, 2);
    const int kt0 = (int)((int64_t)p.ktiles * s / p.S), kt1 = (int)((int64_t)p.ktiles * (s + 1) / p.S);

    __shared__ double As[GK][GM + 1];
    __shared__ double Bs[GK][GN + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t row0 = (int64_t)tm * GM, col0 = (int64_t)tn * GN;
    const int64_t m = p.m, n = p.n, k = p.k, lda = p.lda, ldb = p.ldb;
    double ra[8], rb[8];
    // A tile 64 x 32: row tid % 64, k = tid / 64 + 4 q; B tile 32 x 64: k = tid % 32, column tid / 32 + 8 q
    auto load = [&](int kt) {
        const int64_t k0 = (int64_t)kt * GK;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int64_t gi = row0 + (tid & 63), gk = k0 + (tid >> 6) + 4 * q;
            ra[q] = (gi < m && gk < k) ? A[gi + lda * gk] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int64_t gk = k0 + (tid & 31), gj = col0 + (tid >> 5) + 8 * q;
            rb[q] = (gk < k && gj < n) ? B[gk + ldb * gj] : 0.0;
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int q = 0; q < 8; q++) As[(tid >> 6) + 4 * q][tid & 63] = ra[q];
#pragma unroll
        for (int q = 0; q < 8; q++) Bs[tid & 31][(tid >> 5) + 8 * q] = rb[q];
    };
    double4_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = (double4_t){0, 0, 0, 0};

    if (kt0 < kt1) {
        load(kt0);
        stage();
    }
    __syncthreads();
    for (int kt = kt0; kt < kt1; kt++) {
        const bool more = kt + 1 < kt1;
        if (more) load(kt + 1);
        // A[i][k] of a 16x4 block sits in lane i + 16 k, B[k][j] in lane j + 16 k
#pragma unroll
        for (int ks = 0; ks < GK; ks += 4) {
            const int kk = ks + (lane >> 4);
            const double a0 = As[kk][wm * 32 + (lane & 15)];
            const double a1 = As[kk][wm * 32 + 16 + (lane & 15)];
            const double b0 = Bs[kk][wn * 32 + (lane & 15)];
            const double b1 = Bs[kk][wn * 32 + 16 + (lane & 15)];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
        if (more) stage();
        __syncthreads();
    }
    // D[i][j] of a 16x16 f64 tile: element e of lane l holds i = 4 e + l / 16, j = l % 16
    const int64_t zc = opi(p, z, 4);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int64_t gi = row0 + wm * 32 + a * 16 + 4 * e + (lane >> 4);
                const int64_t gj = col0 + wn * 32 + b * 16 + (lane & 15);
                if (gi < m && gj < n) {
                    if (p.S == 1) store_out(p, zc, gi, gj, acc[a][b][e]);
                    else p.part[gi + m * gj + m * n * (s + (int64_t)p.S * z)] = acc[a][b][e];
                }
            }
}

// C = Σ_s part_s in ascending s (then + C if accumulating), or the complex epilogue.
__global__ __launch_bounds__(256) void dgemm_splitk_reduce_kernel(GemmArgs p)
{
    const int64_t mn = p.m * p.n;
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t z = blockIdx.y;
    if (q >= mn) return;
    if (p.gate && p.gate[p.idx ? p.idx[z] : z] == 0) return;
    const double *src = p.part + mn * (int64_t)p.S * z + q;
    double acc = src[0];
    for (int s = 1; s < p.S; s++) acc += src[mn * s];
    store_out(p, opi(p, z, 4), q % p.m, q / p.m, acc);
}

}  // namespace

int dgemm(epsx_ctx *ctx, int64_t m, int64_t n, int64_t k, const double *A, int64_t lda, const double *B,
          int64_t ldb, double *C, int64_t ldc, bool accumulate, const GemmBatch &bt)
{
    if (m <= 0 || n <= 0 || bt.n_inst <= 0) return EPSX_OK;
    GemmArgs p{};
    p.m = m; p.n = n; p.k = k; p.lda = lda; p.ldb = ldb; p.ldc = ldc;
    p.sA = bt.sA; p.sB = bt.sB; p.sC = bt.sC;
    p.A = A; p.B = B; p.C = C; p.Z = bt.Z; p.Zim = bt.Zim;
    p.idx = bt.idx; p.gate = bt.gate; p.imap = bt.imap;
    p.accumulate = accumulate ? 1 : 0;
    p.mt = (int)((m + GM - 1) / GM);
    p.nt = (int)((n + GN - 1) / GN);
    p.ktiles = (int)((k + GK - 1) / GK);
    // Split k to fill the machine: 256 CUs hold 4 workgroups each (the LDS and register use above),
    // so a launch runs in rounds of 1024 units; pick the split with the least rounds x k-tiles per
    // unit (+2 for the prologue and epilogue), keeping >= 4 k-tiles per split.  The choice depends
    // on the shape and the batch size only, so results are bitwise reproducible.
    const int64_t tiles = (int64_t)p.mt * p.nt;
    constexpr int64_t SLOTS = 256 * 4;
    int S = 1;
    double best = 0;
    for (int c = 1; c <= 16 && (c == 1 || p.ktiles / c >= 4); c++) {
        const int64_t rounds = (tiles * c * bt.n_inst + SLOTS - 1) / SLOTS;
        const double cost = (double)rounds * ((double)p.ktiles / c + 2.0);
        if (c == 1 || cost < best) { best = cost; S = c; }
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    EPSX_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cap));
    if (S > 1) {
        const size_t need = 8 * (size_t)m * n * S * bt.n_inst;
        if (ctx->gemm_ws.bytes < need) {
            if (cap != hipStreamCaptureStatusNone)
                return set_error(ctx, EPSX_ERR_STATE, "GEMM workspace must be sized before graph capture");
            EPSX_TRY(dev_alloc(ctx, ctx->gemm_ws, need));
        }
        p.part = static_cast<double *>(ctx->gemm_ws.p);
    }
    p.S = S;
    p.units = tiles * S * bt.n_inst;
    const int64_t grid = (p.units + NXCD - 1) / NXCD * NXCD;
    hipLaunchKernelGGL(dgemm_kernel, dim3((unsigned)grid), dim3(GT), 0, ctx->stream, p);
    EPSX_HIP(ctx, hipGetLastError());
    if (S > 1) {
        hipLaunchKernelGGL(dgemm_splitk_reduce_kernel, dim3((unsigned)((m * n + 255) / 256), (unsigned)bt.n_inst),
                           dim3(256), 0, ctx->stream, p);
        EPSX_HIP(ctx, hipGetLastError());
    }
    return EPSX_OK;
}

int dgemm_nn(epsx_ctx *ctx, int64_t m, int64_t ncol, int64_t k, const double *A, int64_t lda, const double *B,
             int64_t ldb, double *C, int64_t ldc, bool accumulate, const int32_t *gate)
{
    GemmBatch bt;
    bt.gate = gate;
    return dgemm(ctx, m, ncol, k, A, lda, B, ldb, C, ldc, accumulate, bt);
}

}  // namespace epsx
