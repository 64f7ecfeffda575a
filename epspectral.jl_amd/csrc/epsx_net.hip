// The learned Σ_init warm start: a 1D convolutional network over the frequency grid, with its
// forward, backward and Adam kernels.  The reference has no learned component; its hook is the
// Σ_init argument of the loop (src/scgd0.jl:66-67, :204-212).
//
// Network: n_layers 'same'-padded convolutions of width `kernel` over the grid points of one
// outer state: c_in -> c_hidden -> ... -> c_hidden -> c_out, SiLU after every layer but the last.
// Parameters are FP32 master weights, W_l [cout][cin][kernel] then b_l [cout] per layer.
//
// Layout on the device: an activation is [channels][nb * L] (sample-major, then grid point), FP32.
// Every convolution is an implicit GEMM on the MFMA units, in BF16 (v_mfma_f32_16x16x32_bf16,
// operands rounded to nearest even when staged to LDS, FP32 accumulation) or in FP32
// (v_mfma_f32_16x16x4f32):
//   forward   Y[co][n]          = Σ_{ci,t} W[co][ci,t] X[ci][n + t - h]
//   input     dX[ci][n]         = Σ_{co,t} W[co][ci,t] dZ[co][n - t + h]
//   weights   dW[co][ci,t]      = Σ_n dZ[co][n] X[ci][n + t - h]        (split over n, fixed order)
// with h = kernel / 2 and taps outside the sample's L points read as zero.  All reductions run in
// a fixed order, so training is bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "epsx_common.h"

namespace epsx {
namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int NG_M = 64, NG_N = 64, NG_K = 32, NG_T = 256;
enum { B_IM2COL = 0, B_IM2COL_T = 1 };

struct NetGemm {
    int64_t M, N, K;
    const float *A;
    int64_t sam, sak;          // A[m][k] = A[m * sam + k * sak]
    const float *X;            // the gathered operand
    int64_t ldx;               // channel stride of X
    int L, KW, dir;            // grid points per sample, kernel width, tap direction (+1 / -1)
    float *C;                  // C[m][n] = C[m * N + n]
    float *part;               // split partials [S][M][N]
    int S, mt, nt, ktiles;
    int64_t units;
};

__device__ inline unsigned short f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (unsigned short)((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// B operand element (k, n):
//   B_IM2COL:   k = c * KW + t, value X[c][n + dir (t - h)] (same sample)
//   B_IM2COL_T: k = n', n = c * KW + t, value X[c][n' + t - h]
// x is the grid point of the position (n, or n') within its sample, found by the caller.
template <int BM>
__device__ inline float b_elem(const NetGemm &g, int64_t k, int64_t n, int x)
{
    if (k >= g.K || n >= g.N) return 0.f;
    const int h = g.KW / 2;
    const int kn = BM == B_IM2COL ? (int)k : (int)n;   // c * KW + t (c * KW + t < 2^31)
    const int c = kn / g.KW, t = kn - c * g.KW;
    const int64_t pos = BM == B_IM2COL ? n : k;
    const int d = (BM == B_IM2COL ? g.dir : 1) * (t - h);
    if (x + d < 0 || x + d >= g.L) return 0.f;
    return g.X[c * g.ldx + pos + d];
}

// The grid point (position mod L) of position base + i, 0 <= i < 64, from xb = base mod L: one
// 64-bit remainder per tile instead of one per element.
__device__ inline int grid_point(int xb, int i, int L)
{
    int x = xb + i;
    if (L >= 64) return x >= L ? x - L : x;
    return x % L;
}

// One 64x64 tile of C per workgroup and one split of k; 4 waves, each 32x32 as 2x2 MFMA tiles.
// Operands are staged k-contiguous in LDS: As[m][k], Bs[n][k].
template <bool BF16, int BM>
__global__ __launch_bounds__(NG_T) void net_gemm_kernel(NetGemm g)
{
    const int64_t v = blockIdx.x;
    if (v >= g.units) return;
    const int tm = (int)(v % g.mt), tn = (int)((v / g.mt) % g.nt), s = (int)(v / ((int64_t)g.mt * g.nt));
    const int kt0 = (int)((int64_t)g.ktiles * s / g.S), kt1 = (int)((int64_t)g.ktiles * (s + 1) / g.S);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t row0 = (int64_t)tm * NG_M, col0 = (int64_t)tn * NG_N;
    constexpr int PAD = BF16 ? 8 : 1;
    typedef typename std::conditional<BF16, unsigned short, float>::type st_t;
    __shared__ __align__(16) st_t As[NG_M][NG_K + PAD];
    __shared__ __align__(16) st_t Bs[NG_N][NG_K + PAD];
    f32x4_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
    const int xcol = (int)(col0 % g.L);   // grid point of the tile's first column (B_IM2COL)
    auto put = [](st_t &d, float x) {
        if constexpr (BF16) d = f2bf(x);
        else d = x;
    };
    for (int kt = kt0; kt < kt1; kt++) {
        const int64_t k0 = (int64_t)kt * NG_K;
        // A tile: thread -> (m = e / 32, k = e % 32), consecutive threads along k
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int e = tid + NG_T * q, mm = e >> 5, kk = e & 31;
            const int64_t gm = row0 + mm, gk = k0 + kk;
            put(As[mm][kk], (gm < g.M && gk < g.K) ? g.A[gm * g.sam + gk * g.sak] : 0.f);
        }
        if (BM == B_IM2COL) {
            // consecutive threads along n (contiguous in X)
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int e = tid + NG_T * q, nn = e & 63, kk = e >> 6;
                put(Bs[nn][kk], b_elem<BM>(g, k0 + kk, col0 + nn, grid_point(xcol, nn, g.L)));
            }
        } else {
            // consecutive threads along k (contiguous in X)
            const int xk = (int)(k0 % g.L);
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int e = tid + NG_T * q, kk = e & 31, nn = e >> 5;
                put(Bs[nn][kk], b_elem<BM>(g, k0 + kk, col0 + nn, grid_point(xk, kk, g.L)));
            }
        }
        __syncthreads();
        if constexpr (BF16) {
            // lane l holds A[l % 16][8 (l / 16) .. + 7] and B[8 (l / 16) .. + 7][l % 16]
            const int kb = 8 * (lane >> 4);
            bf16x8_t af[2], bfr[2];
#pragma unroll
            for (int a = 0; a < 2; a++) {
                const int r = wm * 32 + a * 16 + (lane & 15);
                __builtin_memcpy(&af[a], &As[r][kb], 16);
            }
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const int c = wn * 32 + b * 16 + (lane & 15);
                __builtin_memcpy(&bfr[b], &Bs[c][kb], 16);
            }
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        } else {
#pragma unroll
            for (int ks = 0; ks < NG_K; ks += 4) {
                const int kk = ks + (lane >> 4);
                float af[2], bfr[2];
#pragma unroll
                for (int a = 0; a < 2; a++) af[a] = As[wm * 32 + a * 16 + (lane & 15)][kk];
#pragma unroll
                for (int b = 0; b < 2; b++) bfr[b] = Bs[wn * 32 + b * 16 + (lane & 15)][kk];
#pragma unroll
                for (int a = 0; a < 2; a++)
#pragma unroll
                    for (int b = 0; b < 2; b++)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bfr[b], acc[a][b], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // C[i][j] of a 16x16 tile: element e of lane l holds i = 4 (l / 16) + e, j = l % 16
    float *out = g.S == 1 ? g.C : g.part + (int64_t)s * g.M * g.N;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int64_t gi = row0 + wm * 32 + a * 16 + 4 * (lane >> 4) + e;
                const int64_t gj = col0 + wn * 32 + b * 16 + (lane & 15);
                if (gi < g.M && gj < g.N) out[gi * g.N + gj] = acc[a][b][e];
            }
}

__global__ void net_splitk_reduce_kernel(int64_t mn, int S, const float *__restrict__ part, float *__restrict__ C)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= mn) return;
    float acc = part[q];
    for (int s = 1; s < S; s++) acc += part[q + mn * s];
    C[q] = acc;
}

__device__ inline float sigmoidf_(float z) { return 1.f / (1.f + expf(-z)); }

// Z = Y + b[c] (kept for the backward pass); X_next = SiLU(Z), or Z itself for the last layer
__global__ void net_bias_act_kernel(int64_t nb_len, int64_t nch, const float *__restrict__ Y, const float *__restrict__ b,
                                    float *__restrict__ Z, float *__restrict__ Xn, int act)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nb_len * nch) return;
    const float z = Y[q] + b[q / nb_len];
    Z[q] = z;
    Xn[q] = act ? z * sigmoidf_(z) : z;
}

// dZ = dA SiLU'(Z), SiLU'(z) = s (1 + z (1 - s))
__global__ void net_act_bwd_kernel(int64_t n, const float *__restrict__ Z, float *__restrict__ dA)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float z = Z[q], s = sigmoidf_(z);
    dA[q] = dA[q] * (s * (1.f + z * (1.f - s)));
}

// db[c] = Σ_n dZ[c][n], one block per channel, fixed-order tree
__global__ __launch_bounds__(256) void net_bias_grad_kernel(int64_t nb_len, const float *__restrict__ dZ,
                                                            float *__restrict__ db)
{
    __shared__ float red[256];
    const int64_t c = blockIdx.x;
    float acc = 0.f;
    for (int64_t q = threadIdx.x; q < nb_len; q += 256) acc += dZ[c * nb_len + q];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}

// per-block sums of (pred - target)^2 and dPred = 2 (pred - target) / n_total
__global__ __launch_bounds__(256) void net_mse_kernel(int64_t n, const float *__restrict__ P, const float *__restrict__ T,
                                                      float *__restrict__ dP, double *__restrict__ bsum)
{
    __shared__ double red[256];
    double acc = 0;
    for (int64_t q = blockIdx.x * 256 + threadIdx.x; q < n; q += (int64_t)gridDim.x * 256) {
        const float d = P[q] - T[q];
        dP[q] = 2.f * d / (float)n;
        acc += (double)d * d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

__global__ void net_sum_kernel(int nb, const double *__restrict__ bsum, int64_t n, double *__restrict__ loss)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0;
    for (int b = 0; b < nb; b++) s += bsum[b];
    *loss = s / (double)n;
}

// Wt[ci][co * KW + t] = W[co][ci][t]
__global__ void net_wt_kernel(int cout, int cin, int KW, const float *__restrict__ W, float *__restrict__ Wt)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= (int64_t)cout * cin * KW) return;
    const int t = (int)(q % KW), ci = (int)((q / KW) % cin), co = (int)(q / ((int64_t)KW * cin));
    Wt[(int64_t)ci * cout * KW + (int64_t)co * KW + t] = W[q];
}

// X0[c][b * L + x] = data[idx[b]][c][x]  (data: [n][c][L])
__global__ void net_gather_kernel(int64_t nb, int64_t nch, int64_t L, const int64_t *__restrict__ idx,
                                  const float *__restrict__ data, float *__restrict__ X)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nb * nch * L) return;
    const int64_t x = q % L, b = (q / L) % nb, c = q / (L * nb);
    X[q] = data[(idx[b] * nch + c) * L + x];
}

// [c][b * L + x] -> out[b][c][x]
__global__ void net_scatter_kernel(int64_t nb, int64_t nch, int64_t L, const float *__restrict__ Y, float *__restrict__ out)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nb * nch * L) return;
    const int64_t x = q % L, b = (q / L) % nb, c = q / (L * nb);
    out[(b * nch + c) * L + x] = Y[q];
}

// Adam with decoupled weight decay on the FP32 master weights
__global__ void net_adam_kernel(int64_t n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                                float *__restrict__ v, float lr, float b1, float b2, float eps, float wd, float c1,
                                float c2)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.f - b1) * gq;
    const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * ((mq / c1) / (sqrtf(vq / c2) + eps) + wd * p[q]);
}

inline unsigned nblk(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

// Per-layer shapes and offsets of the parameter vector
static void net_layout(const NetDev &n, std::vector<int> &cin, std::vector<int> &cout, std::vector<int64_t> &woff,
                       std::vector<int64_t> &boff)
{
    cin.clear(); cout.clear(); woff.clear(); boff.clear();
    int64_t o = 0;
    for (int l = 0; l < n.layers; l++) {
        const int ci = l == 0 ? n.c_in : n.c_hidden;
        const int co = l == n.layers - 1 ? n.c_out : n.c_hidden;
        cin.push_back(ci);
        cout.push_back(co);
        woff.push_back(o);
        o += (int64_t)co * ci * n.kernel;
        boff.push_back(o);
        o += co;
    }
}

static int net_gemm(epsx_ctx *ctx, int bmode, bool bf16, int64_t M, int64_t N, int64_t K, const float *A, int64_t sam,
                    int64_t sak, const float *X, int64_t ldx, int L, int KW, int dir, float *C)
{
    NetDev &nd = ctx->net;
    NetGemm g{};
    g.M = M; g.N = N; g.K = K; g.A = A; g.sam = sam; g.sak = sak; g.X = X; g.ldx = ldx;
    g.L = L; g.KW = KW; g.dir = dir; g.C = C;
    g.mt = (int)((M + NG_M - 1) / NG_M);
    g.nt = (int)((N + NG_N - 1) / NG_N);
    g.ktiles = (int)((K + NG_K - 1) / NG_K);
    const int64_t tiles = (int64_t)g.mt * g.nt;
    int S = 1;
    while (tiles * S < 512 && g.ktiles / (2 * S) >= 4) S *= 2;
    g.S = S;
    if (S > 1) {
        EPSX_TRY(dev_alloc(ctx, nd.part, 4 * (size_t)S * M * N));
        g.part = static_cast<float *>(nd.part.p);
    }
    g.units = tiles * S;
    const dim3 grid((unsigned)g.units), blk(NG_T);
    if (bmode == B_IM2COL) {
        if (bf16) hipLaunchKernelGGL((net_gemm_kernel<true, B_IM2COL>), grid, blk, 0, ctx->stream, g);
        else hipLaunchKernelGGL((net_gemm_kernel<false, B_IM2COL>), grid, blk, 0, ctx->stream, g);
    } else {
        if (bf16) hipLaunchKernelGGL((net_gemm_kernel<true, B_IM2COL_T>), grid, blk, 0, ctx->stream, g);
        else hipLaunchKernelGGL((net_gemm_kernel<false, B_IM2COL_T>), grid, blk, 0, ctx->stream, g);
    }
    EPSX_HIP(ctx, hipGetLastError());
    if (S > 1) {
        hipLaunchKernelGGL(net_splitk_reduce_kernel, dim3(nblk(M * N)), dim3(256), 0, ctx->stream, M * N, S,
                           static_cast<const float *>(nd.part.p), C);
        EPSX_HIP(ctx, hipGetLastError());
    }
    return EPSX_OK;
}

// Activation buffers for nb samples
static int net_buffers(epsx_ctx *ctx, int64_t nb)
{
    NetDev &n = ctx->net;
    if (n.cap_nb >= nb) return EPSX_OK;
    const int64_t NB = nb * n.length;
    const int cmax = std::max(std::max(n.c_in, n.c_hidden), n.c_out);
    for (int l = 0; l <= n.layers; l++) EPSX_TRY(dev_alloc(ctx, n.X[l], 4 * (size_t)cmax * NB));
    for (int l = 0; l < n.layers; l++) EPSX_TRY(dev_alloc(ctx, n.Z[l], 4 * (size_t)cmax * NB));
    EPSX_TRY(dev_alloc(ctx, n.Y, 4 * (size_t)cmax * NB));
    EPSX_TRY(dev_alloc(ctx, n.dA, 4 * (size_t)cmax * NB));
    EPSX_TRY(dev_alloc(ctx, n.dB, 4 * (size_t)cmax * NB));
    EPSX_TRY(dev_alloc(ctx, n.T, 4 * (size_t)n.c_out * NB));
    EPSX_TRY(dev_alloc(ctx, n.idx, 8 * (size_t)nb));
    n.cap_nb = nb;
    return EPSX_OK;
}

// ---- inference: one fused convolution per layer ---------------------------------------------------
// X_next[co][pos] = act(b[co] + Σ_{ci,t} W[co][ci][t] X[ci][pos + t - h]) for c_in, c_out <= 64, as a
// BF16 MFMA product (v_mfma_f32_16x16x32_bf16, FP32 sums) with the reduction index tap-major,
// k = t CI + ci (CI = c_in rounded up to 8): eight consecutive k are eight channels at one tap, so a
// lane's B operand is one 16-byte LDS read of the staged window.  One workgroup per IC_NT positions
// of one sample (the last tile of a sample may be partial): the window X[:, n0 - h .. n0 + IC_NT + h) is staged in LDS
// as BF16 [position][channel] (zeros past the sample's ends and past c_in); wave w owns the output
// channels 16 w .. 16 w + 15 with its A operand (the packed weights) held in registers for the
// whole k range.  Bias and SiLU in the epilogue.
constexpr int IC_NT = 128, IC_MAXK = 18 * 32;   // k up to 9 taps x 64 channels

// Wp[m][k] (IC_KP = Kp columns, 64 rows): BF16 of W[m][ci][t] at k = t CI + ci, zero elsewhere
__global__ void net_pack_weights_kernel(int cout, int cin, int KW, int CI, int Kp, const float *__restrict__ W,
                                        uint16_t *__restrict__ Wp)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 64 * Kp) return;
    const int m = q / Kp, k = q - m * Kp, t = k / CI, ci = k - t * CI;
    const float w = (m < cout && ci < cin && t < KW) ? W[((int64_t)m * cin + ci) * KW + t] : 0.f;
    Wp[q] = f2bf(w);
}

template <int KS>   // k steps of 32 (Kp = 32 KS): the A operand stays in registers
__global__ __launch_bounds__(256) void net_conv_fwd_kernel(const uint16_t *__restrict__ Wp, const float *__restrict__ bias,
                                                           int cout, int cin, int CI, int KW, int L,
                                                           const float *__restrict__ X, int64_t ldx,
                                                           float *__restrict__ Y, int act)
{
    constexpr int Kp = 32 * KS;
    __shared__ __align__(16) uint16_t xw[(IC_NT + 16) * (64 + 8)];
    const int PW = CI + 8;                       // row pitch (elements): 16-byte reads on distinct banks
    const int h = KW / 2, NW = IC_NT + KW - 1;   // window positions
    const int tiles = (L + IC_NT - 1) / IC_NT;   // tiles per sample
    const int xs = (int)(blockIdx.x % tiles) * IC_NT;   // grid point of the tile's first position
    const int64_t n0 = (int64_t)(blockIdx.x / tiles) * L + xs;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // the window, eight channels of one position per item (one 16-byte LDS store); q / NW as a
    // multiply-high by ceil(2^32 / NW), exact while q NW < 2^32 (q < NW CI / 8 <= 144 x 8)
    const unsigned rnw = 0xFFFFFFFFu / (unsigned)NW + 1u;
    for (int q = tid; q < NW * (CI / 8); q += 256) {
        const int cg = (int)__umulhi((unsigned)q, rnw), pw = q - cg * NW;
        const int x = xs - h + pw;
        const bool in = x >= 0 && x < L;
        bf16x8_t v;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = 8 * cg + j;
            v[j] = (__bf16)((in && c < cin) ? X[c * ldx + n0 - h + pw] : 0.f);
        }
        __builtin_memcpy(xw + pw * PW + 8 * cg, &v, 16);
    }
    const unsigned rci = 0xFFFFFFFFu / (unsigned)CI + 1u;   // k / CI likewise (k < 18 x 32)
    const int row = 16 * wave + (lane & 15), kb = 8 * (lane >> 4);
    bf16x8_t af[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) __builtin_memcpy(&af[ks], Wp + (int64_t)row * Kp + ks * 32 + kb, 16);
    // the fragments stay in registers for the whole k range (the empty asm keeps the compiler from
    // reloading each one from memory next to its MFMA)
#pragma unroll
    for (int ks = 0; ks < KS; ks++) asm volatile("" : "+v"(af[ks]));
    __syncthreads();
    if (16 * wave >= cout) return;
    f32x4_t acc[IC_NT / 16];
#pragma unroll
    for (int j = 0; j < IC_NT / 16; j++) acc[j] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
        const int k = ks * 32 + kb;              // eight channels ci .. ci + 7 at tap t
        const int t0 = (int)__umulhi((unsigned)k, rci), ci = k - t0 * CI;
        const int t = t0 < KW ? t0 : KW - 1;     // taps past KW: zero weights, any finite window value
#pragma unroll
        for (int j = 0; j < IC_NT / 16; j++) {
            bf16x8_t bfr;
            __builtin_memcpy(&bfr, xw + (j * 16 + (lane & 15) + t) * PW + ci, 16);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[ks], bfr, acc[j], 0, 0, 0);
        }
    }
    // C[i][j] of a 16x16 tile: element e of lane l holds i = 4 (l / 16) + e, j = l % 16
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int co = 16 * wave + 4 * (lane >> 4) + e;
        if (co >= cout) continue;
        const float b = bias[co];
#pragma unroll
        for (int j = 0; j < IC_NT / 16; j++) {
            const int x = j * 16 + (lane & 15);
            if (xs + x >= L) continue;   // past the sample's last position
            const float z = acc[j][e] + b;
            Y[co * ldx + n0 + x] = act ? z * sigmoidf_(z) : z;
        }
    }
}

// Whether the fused inference convolution serves this network
static bool net_infer_fits(const NetDev &n)
{
    if (n.kernel > 17) return false;
    const int cm = std::max(std::max(n.c_in, n.c_hidden), n.c_out);
    const int CI = (cm + 7) / 8 * 8;
    return cm <= 64 && (int64_t)n.kernel * CI <= IC_MAXK;
}

// BF16 inference of nb samples in X[0] into X[layers] (no Z kept): the fused convolution when it
// fits, else the training forward pass
static int net_infer_dev(epsx_ctx *ctx, int64_t nb);

// Forward pass of nb samples already gathered into X[0]; output in X[layers]
static int net_forward_dev(epsx_ctx *ctx, int64_t nb, bool bf16)
{
    NetDev &n = ctx->net;
    std::vector<int> cin, cout;
    std::vector<int64_t> woff, boff;
    net_layout(n, cin, cout, woff, boff);
    const int64_t NB = nb * n.length;
    const float *P = static_cast<const float *>(n.params.p);
    for (int l = 0; l < n.layers; l++) {
        EPSX_TRY(net_gemm(ctx, B_IM2COL, bf16, cout[l], NB, (int64_t)cin[l] * n.kernel, P + woff[l],
                          (int64_t)cin[l] * n.kernel, 1, dptr<float>(n.X[l]), NB, (int)n.length, n.kernel, 1,
                          dptr<float>(n.Y)));
        hipLaunchKernelGGL(net_bias_act_kernel, dim3(nblk(NB * cout[l])), dim3(256), 0, ctx->stream, NB,
                           (int64_t)cout[l], dptr<float>(n.Y), P + boff[l], dptr<float>(n.Z[l]),
                           dptr<float>(n.X[l + 1]), l < n.layers - 1 ? 1 : 0);
        EPSX_HIP(ctx, hipGetLastError());
    }
    return EPSX_OK;
}

static int net_infer_dev(epsx_ctx *ctx, int64_t nb)
{
    NetDev &n = ctx->net;
    if (!net_infer_fits(n)) return net_forward_dev(ctx, nb, true);
    std::vector<int> cin, cout;
    std::vector<int64_t> woff, boff;
    net_layout(n, cin, cout, woff, boff);
    const int64_t NB = nb * n.length;
    const float *P = static_cast<const float *>(n.params.p);
    const int Kp = (n.kernel * 64 + 31) / 32 * 32;   // the widest layer's k, rounded to 32
    if (!n.wp_ok) {   // the packed weights of every layer, once per set of parameters
        EPSX_TRY(dev_alloc(ctx, n.Wp, 2 * (size_t)n.layers * 64 * Kp));
        for (int l = 0; l < n.layers; l++) {
            const int CI = (cin[l] + 7) / 8 * 8, Kl = (n.kernel * CI + 31) / 32 * 32;
            hipLaunchKernelGGL(net_pack_weights_kernel, dim3(nblk(64 * Kl)), dim3(256), 0, ctx->stream, cout[l], cin[l],
                               n.kernel, CI, Kl, P + woff[l], static_cast<uint16_t *>(n.Wp.p) + (size_t)l * 64 * Kp);
            EPSX_HIP(ctx, hipGetLastError());
        }
        n.wp_ok = true;
    }
    for (int l = 0; l < n.layers; l++) {
        const int CI = (cin[l] + 7) / 8 * 8, Kl = (n.kernel * CI + 31) / 32 * 32;
        const uint16_t *W = static_cast<const uint16_t *>(n.Wp.p) + (size_t)l * 64 * Kp;
        const dim3 grid((unsigned)(nb * ((n.length + IC_NT - 1) / IC_NT)));
        const float *b = P + boff[l], *X = dptr<float>(n.X[l]);
        float *Y = dptr<float>(n.X[l + 1]);
        const int act = l < n.layers - 1 ? 1 : 0;
        switch (Kl / 32) {
#define EPSX_CONV_CASE(q)                                                                                         \
    case q:                                                                                                        \
        hipLaunchKernelGGL(net_conv_fwd_kernel<q>, grid, dim3(256), 0, ctx->stream, W, b, cout[l], cin[l], CI,    \
                           n.kernel, (int)n.length, X, NB, Y, act);                                                \
        break;
            EPSX_CONV_CASE(1) EPSX_CONV_CASE(2) EPSX_CONV_CASE(3) EPSX_CONV_CASE(4) EPSX_CONV_CASE(5)
            EPSX_CONV_CASE(6) EPSX_CONV_CASE(7) EPSX_CONV_CASE(8) EPSX_CONV_CASE(9) EPSX_CONV_CASE(10)
            EPSX_CONV_CASE(11) EPSX_CONV_CASE(12) EPSX_CONV_CASE(13) EPSX_CONV_CASE(14) EPSX_CONV_CASE(15)
            EPSX_CONV_CASE(16) EPSX_CONV_CASE(17) EPSX_CONV_CASE(18)
#undef EPSX_CONV_CASE
        default: return set_error(ctx, EPSX_ERR_STATE, "inference convolution: %d k steps", Kl / 32);
        }
        EPSX_HIP(ctx, hipGetLastError());
    }
    return EPSX_OK;
}

// Loss and gradient of nb samples (inputs in X[0], targets in T) into grad (device, P entries)
static int net_backward_dev(epsx_ctx *ctx, int64_t nb, bool bf16, float *grad, double *d_loss)
{
    NetDev &n = ctx->net;
    std::vector<int> cin, cout;
    std::vector<int64_t> woff, boff;
    net_layout(n, cin, cout, woff, boff);
    const int64_t NB = nb * n.length;
    const float *P = static_cast<const float *>(n.params.p);
    EPSX_TRY(net_forward_dev(ctx, nb, bf16));
    const int64_t nout = NB * n.c_out;
    const int nbl = 256;
    EPSX_TRY(dev_alloc(ctx, n.bsum, 8 * (size_t)nbl));
    hipLaunchKernelGGL(net_mse_kernel, dim3(nbl), dim3(256), 0, ctx->stream, nout, dptr<float>(n.X[n.layers]),
                       dptr<float>(n.T), dptr<float>(n.dA), dptr<double>(n.bsum));
    hipLaunchKernelGGL(net_sum_kernel, dim3(1), dim3(64), 0, ctx->stream, nbl, dptr<double>(n.bsum), nout, d_loss);
    EPSX_HIP(ctx, hipGetLastError());
    // dA holds dL/dZ of the last layer (no activation there)
    for (int l = n.layers - 1; l >= 0; l--) {
        if (l < n.layers - 1) {
            hipLaunchKernelGGL(net_act_bwd_kernel, dim3(nblk(NB * cout[l])), dim3(256), 0, ctx->stream,
                               NB * cout[l], dptr<float>(n.Z[l]), dptr<float>(n.dA));
            EPSX_HIP(ctx, hipGetLastError());
        }
        // dW[co][ci, t] = Σ_n dZ[co][n] X[ci][n + t - h]
        EPSX_TRY(net_gemm(ctx, B_IM2COL_T, bf16, cout[l], (int64_t)cin[l] * n.kernel, NB, dptr<float>(n.dA), NB, 1,
                          dptr<float>(n.X[l]), NB, (int)n.length, n.kernel, 1, grad + woff[l]));
        hipLaunchKernelGGL(net_bias_grad_kernel, dim3(cout[l]), dim3(256), 0, ctx->stream, NB, dptr<float>(n.dA),
                           grad + boff[l]);
        EPSX_HIP(ctx, hipGetLastError());
        if (l > 0) {
            // dX[ci][n] = Σ_{co,t} Wt[ci][co, t] dZ[co][n - t + h]
            hipLaunchKernelGGL(net_wt_kernel, dim3(nblk((int64_t)cout[l] * cin[l] * n.kernel)), dim3(256), 0,
                               ctx->stream, cout[l], cin[l], n.kernel, P + woff[l], dptr<float>(n.Wt));
            EPSX_HIP(ctx, hipGetLastError());
            EPSX_TRY(net_gemm(ctx, B_IM2COL, bf16, cin[l], NB, (int64_t)cout[l] * n.kernel, dptr<float>(n.Wt),
                              (int64_t)cout[l] * n.kernel, 1, dptr<float>(n.dA), NB, (int)n.length, n.kernel, -1,
                              dptr<float>(n.dB)));
            std::swap(n.dA, n.dB);
        }
    }
    return EPSX_OK;
}

// ---- the warm start's prediction on the device ---------------------------------------------------
// Inputs per (outer state i, slot z), over the L = nω grid points (the features of
// epspectral.jl_amd/epsx/learn.py, which the network was trained on):
//   0, 1: asinh(Re Σ / SIGMA_SCALE), asinh(Im Σ / SIGMA_SCALE) of the G0D0 self-energy;
//   2: (ω - μ) / ENERGY_SCALE;  3: (ε_i - μ) / ENERGY_SCALE;  4: T / T_SCALE.
// The output (Re, Im) times SIGMA_SCALE is the predicted Σ.
constexpr double WS_SIGMA_SCALE = 0.05, WS_ENERGY_SCALE = 1.0, WS_T_SCALE = 0.04;

// X0[c][b L + x] for the samples b = i + ni z' (z' = z - z0 < nbz) of the slots tidx[z]
__global__ void ws_features_kernel(int64_t L, int64_t ni, int64_t b0, int64_t nb, const double2 *__restrict__ S,
                                   const double *__restrict__ ws, const double *__restrict__ e_i,
                                   const double *__restrict__ mu, const double *__restrict__ T,
                                   const int32_t *__restrict__ tsrc, float *__restrict__ X)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nb * L) return;
    const int64_t x = q % L, bl = q / L, b = b0 + bl;
    const int64_t i = b % ni, t = tsrc[b / ni];
    const double2 v = S[x + L * (i + ni * t)];
    const double m = mu[t];
    const int64_t NB = nb * L;
    X[q] = (float)asinh(v.x / WS_SIGMA_SCALE);
    X[q + NB] = (float)asinh(v.y / WS_SIGMA_SCALE);
    X[q + 2 * NB] = (float)((ws[x] - m) / WS_ENERGY_SCALE);
    X[q + 3 * NB] = (float)((e_i[i] - m) / WS_ENERGY_SCALE);
    X[q + 4 * NB] = (float)(T[t] / WS_T_SCALE);
}

// Σ[x, i, tdst[b / ni]] = complex(Y0, Y1) * SIGMA_SCALE (the scale applied in FP32, as the host
// prediction applies it)
__global__ void ws_output_kernel(int64_t L, int64_t ni, int64_t b0, int64_t nb, const float *__restrict__ Y,
                                 const int32_t *__restrict__ tdst, double2 *__restrict__ S)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nb * L) return;
    const int64_t x = q % L, bl = q / L, b = b0 + bl;
    const int64_t i = b % ni, t = tdst[b / ni];
    const int64_t NB = nb * L;
    const float sc = (float)WS_SIGMA_SCALE;
    S[x + L * (i + ni * t)] = make_double2((double)(Y[q] * sc), (double)(Y[q + NB] * sc));
}

static int net_check(epsx_ctx *ctx, int32_t precision)
{
    if (!ctx) return EPSX_ERR_ARG;
    if (!ctx->net.ready) return set_error(ctx, EPSX_ERR_STATE, "no network: call epsx_net_create first");
    if (precision != EPSX_PREC_FP32 && precision != EPSX_PREC_BF16)
        return set_error(ctx, EPSX_ERR_ARG, "network precision must be FP32 or BF16, got %d", precision);
    return EPSX_OK;
}

static int upload_gather(epsx_ctx *ctx, const float *x, int64_t first, int64_t nb, int64_t nch, float *dst)
{
    NetDev &n = ctx->net;
    Tmp raw;
    EPSX_TRY(tmp_upload(ctx, raw, x + first * nch * n.length, 4 * (size_t)nb * nch * n.length));
    std::vector<int64_t> id(nb);
    for (int64_t b = 0; b < nb; b++) id[b] = b;
    EPSX_HIP(ctx, hipMemcpyAsync(n.idx.p, id.data(), 8 * nb, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(net_gather_kernel, dim3(nblk(nb * nch * n.length)), dim3(256), 0, ctx->stream, nb, nch,
                       n.length, dptr<int64_t>(n.idx), raw.ptr<float>(), dst);
    EPSX_HIP(ctx, hipGetLastError());
    EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EPSX_OK;
}

// The network's prediction Σ[:, :, tdst[z]] from the G0D0 Σ[:, :, tsrc[z]] for the nz slot pairs
// (device tables tsrc / tdst), all on the device; in chunks of at most 512 samples.
int net_predict_sigma(epsx_ctx *ctx, int64_t nw, int64_t ni, int nz, const int32_t *tsrc, const int32_t *tdst,
                      double2 *S, const double *ws, const double *e_i, const double *mu, const double *T, int precision)
{
    EPSX_TRY(net_check(ctx, precision));
    NetDev &n = ctx->net;
    if (n.c_in != 5 || n.c_out != 2 || n.length != nw)
        return set_error(ctx, EPSX_ERR_ARG, "the warm-start network takes 5 input and 2 output channels over the %lld "
                         "nodes of ωs, got %d, %d and %lld", (long long)nw, n.c_in, n.c_out, (long long)n.length);
    const int64_t total = ni * nz, chunk = 512;
    EPSX_TRY(net_buffers(ctx, std::min(chunk, total)));
    for (int64_t b0 = 0; b0 < total; b0 += chunk) {
        const int64_t nb = std::min(chunk, total - b0);
        hipLaunchKernelGGL(ws_features_kernel, dim3(nblk(nb * nw)), dim3(256), 0, ctx->stream, nw, ni, b0, nb,
                           static_cast<const double2 *>(S), ws, e_i, mu, T, tsrc, dptr<float>(n.X[0]));
        EPSX_HIP(ctx, hipGetLastError());
        EPSX_TRY(precision == EPSX_PREC_BF16 ? net_infer_dev(ctx, nb) : net_forward_dev(ctx, nb, false));
        hipLaunchKernelGGL(ws_output_kernel, dim3(nblk(nb * nw)), dim3(256), 0, ctx->stream, nw, ni, b0, nb,
                           dptr<float>(n.X[n.layers]), tdst, S);
        EPSX_HIP(ctx, hipGetLastError());
    }
    return EPSX_OK;
}

}  // namespace epsx

using namespace epsx;

extern "C" {

int epsx_net_create(epsx_ctx *ctx, const epsx_net_config *cfg, uint64_t seed)
{
    if (!ctx) return EPSX_ERR_ARG;
    if (!cfg || cfg->struct_size != (int64_t)sizeof(epsx_net_config))
        return set_error(ctx, EPSX_ERR_ARG, "epsx_net_create: struct size mismatch (ABI %d)", EPSX_ABI_VERSION);
    if (cfg->n_layers < 2 || cfg->n_layers > NET_MAX_LAYERS || cfg->kernel < 1 || cfg->kernel % 2 == 0 || cfg->c_in < 1 || cfg->c_hidden < 1 ||
        cfg->c_out < 1 || cfg->length < 1)
        return set_error(ctx, EPSX_ERR_ARG, "network: need 2..%d layers, an odd kernel width and positive sizes",
                         NET_MAX_LAYERS);
    EPSX_HIP(ctx, hipSetDevice(ctx->device));
    NetDev &n = ctx->net;
    n.layers = cfg->n_layers; n.kernel = cfg->kernel; n.c_in = cfg->c_in; n.c_hidden = cfg->c_hidden;
    n.c_out = cfg->c_out; n.length = cfg->length; n.cap_nb = 0;
    std::vector<int> cin, cout;
    std::vector<int64_t> woff, boff;
    net_layout(n, cin, cout, woff, boff);
    n.nparam = boff.back() + cout.back();
    // Kaiming-uniform weights for SiLU layers (bound sqrt(6 / fan_in)), the last layer scaled by 0.1
    // so the untrained net predicts a small correction; zero biases.
    std::vector<float> p(n.nparam, 0.f);
    uint64_t st = seed;
    for (int l = 0; l < n.layers; l++) {
        const double fan_in = (double)cin[l] * n.kernel;
        const double bound = std::sqrt(6.0 / fan_in) * (l == n.layers - 1 ? 0.1 : 1.0);
        const int64_t nw = (int64_t)cout[l] * cin[l] * n.kernel;
        for (int64_t q = 0; q < nw; q++) {
            const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            p[woff[l] + q] = (float)((2 * u - 1) * bound);
        }
    }
    EPSX_TRY(upload(ctx, n.params, p.data(), 4 * n.nparam));
    n.wp_ok = false;
    EPSX_TRY(dev_alloc(ctx, n.grad, 4 * (size_t)n.nparam));
    EPSX_TRY(dev_alloc(ctx, n.m, 4 * (size_t)n.nparam));
    EPSX_TRY(dev_alloc(ctx, n.v, 4 * (size_t)n.nparam));
    // Adam starts from zero moments at step 0
    EPSX_HIP(ctx, hipMemsetAsync(n.m.p, 0, 4 * n.nparam, ctx->stream));
    EPSX_HIP(ctx, hipMemsetAsync(n.v.p, 0, 4 * n.nparam, ctx->stream));
    n.step = 0;
    int64_t wmax = 0;
    for (int l = 0; l < n.layers; l++) wmax = std::max(wmax, (int64_t)cout[l] * cin[l] * n.kernel);
    EPSX_TRY(dev_alloc(ctx, n.Wt, 4 * (size_t)wmax));
    EPSX_TRY(dev_alloc(ctx, n.loss, 8));
    EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    n.ready = true;
    return EPSX_OK;
}

int epsx_net_param_count(epsx_ctx *ctx, int64_t *count)
{
    if (!ctx || !count) return EPSX_ERR_ARG;
    if (!ctx->net.ready) return set_error(ctx, EPSX_ERR_STATE, "no network: call epsx_net_create first");
    *count = ctx->net.nparam;
    return EPSX_OK;
}

int epsx_net_get_params(epsx_ctx *ctx, float *params)
{
    if (!ctx || !params) return EPSX_ERR_ARG;
    if (!ctx->net.ready) return set_error(ctx, EPSX_ERR_STATE, "no network: call epsx_net_create first");
    EPSX_HIP(ctx, hipMemcpyAsync(params, ctx->net.params.p, 4 * ctx->net.nparam, hipMemcpyDeviceToHost, ctx->stream));
    EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EPSX_OK;
}

int epsx_net_set_params(epsx_ctx *ctx, const float *params)
{
    if (!ctx || !params) return EPSX_ERR_ARG;
    if (!ctx->net.ready) return set_error(ctx, EPSX_ERR_STATE, "no network: call epsx_net_create first");
    EPSX_HIP(ctx, hipMemcpyAsync(ctx->net.params.p, params, 4 * ctx->net.nparam, hipMemcpyHostToDevice, ctx->stream));
    ctx->net.wp_ok = false;
    EPSX_HIP(ctx, hipMemsetAsync(ctx->net.m.p, 0, 4 * ctx->net.nparam, ctx->stream));
    EPSX_HIP(ctx, hipMemsetAsync(ctx->net.v.p, 0, 4 * ctx->net.nparam, ctx->stream));
    EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->net.step = 0;
    return EPSX_OK;
}

int epsx_net_forward(epsx_ctx *ctx, const float *x, int64_t nsamples, int32_t precision, float *y)
{
    EPSX_TRY(net_check(ctx, precision));
    if (!x || !y || nsamples < 0) return set_error(ctx, EPSX_ERR_ARG, "epsx_net_forward: null array");
    NetDev &n = ctx->net;
    const int64_t chunk = 256;
    EPSX_TRY(net_buffers(ctx, std::min(chunk, std::max<int64_t>(nsamples, 1))));
    Tmp out;
    EPSX_TRY(tmp_alloc(ctx, out.b, 4 * (size_t)chunk * n.c_out * n.length));
    for (int64_t b0 = 0; b0 < nsamples; b0 += chunk) {
        const int64_t nb = std::min(chunk, nsamples - b0);
        EPSX_TRY(upload_gather(ctx, x, b0, nb, n.c_in, dptr<float>(n.X[0])));
        EPSX_TRY(precision == EPSX_PREC_BF16 ? net_infer_dev(ctx, nb) : net_forward_dev(ctx, nb, false));
        hipLaunchKernelGGL(net_scatter_kernel, dim3(nblk(nb * n.c_out * n.length)), dim3(256), 0, ctx->stream, nb,
                           (int64_t)n.c_out, n.length, dptr<float>(n.X[n.layers]), out.ptr<float>());
        EPSX_HIP(ctx, hipGetLastError());
        EPSX_HIP(ctx, hipMemcpyAsync(y + b0 * n.c_out * n.length, out.b.p, 4 * (size_t)nb * n.c_out * n.length,
                                     hipMemcpyDeviceToHost, ctx->stream));
        EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return EPSX_OK;
}

int epsx_net_gradient(epsx_ctx *ctx, const float *x, const float *target, int64_t nsamples, int32_t precision,
                      double *loss, float *grad)
{
    EPSX_TRY(net_check(ctx, precision));
    if (!x || !target || !loss || !grad || nsamples < 1)
        return set_error(ctx, EPSX_ERR_ARG, "epsx_net_gradient: null array or no samples");
    NetDev &n = ctx->net;
    EPSX_TRY(net_buffers(ctx, nsamples));
    EPSX_TRY(upload_gather(ctx, x, 0, nsamples, n.c_in, dptr<float>(n.X[0])));
    EPSX_TRY(upload_gather(ctx, target, 0, nsamples, n.c_out, dptr<float>(n.T)));
    EPSX_TRY(net_backward_dev(ctx, nsamples, precision == EPSX_PREC_BF16, dptr<float>(n.grad), dptr<double>(n.loss)));
    EPSX_HIP(ctx, hipMemcpyAsync(grad, n.grad.p, 4 * n.nparam, hipMemcpyDeviceToHost, ctx->stream));
    EPSX_HIP(ctx, hipMemcpyAsync(loss, n.loss.p, 8, hipMemcpyDeviceToHost, ctx->stream));
    EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EPSX_OK;
}

int epsx_net_train(epsx_ctx *ctx, const float *x, const float *target, int64_t nsamples,
                   const epsx_net_train_params *tp, double *epoch_loss)
{
    if (!ctx) return EPSX_ERR_ARG;
    if (!tp || tp->struct_size != (int64_t)sizeof(epsx_net_train_params))
        return set_error(ctx, EPSX_ERR_ARG, "epsx_net_train: struct size mismatch (ABI %d)", EPSX_ABI_VERSION);
    EPSX_TRY(net_check(ctx, tp->precision));
    if (!x || !target || !epoch_loss || nsamples < 1 || tp->batch < 1 || tp->epochs < 0)
        return set_error(ctx, EPSX_ERR_ARG, "epsx_net_train: null array, no samples or a bad batch size");
    NetDev &n = ctx->net;
    const int64_t batch = std::min(tp->batch, nsamples);
    EPSX_TRY(net_buffers(ctx, batch));
    // the whole training set resident on the device
    Tmp dx, dt, dloss;
    EPSX_TRY(tmp_upload(ctx, dx, x, 4 * (size_t)nsamples * n.c_in * n.length));
    EPSX_TRY(tmp_upload(ctx, dt, target, 4 * (size_t)nsamples * n.c_out * n.length));
    const int64_t nbatches = nsamples / batch;   // the remainder is dropped each epoch
    EPSX_TRY(tmp_alloc(ctx, dloss.b, 8 * (size_t)std::max<int64_t>(nbatches, 1)));
    std::vector<int64_t> perm(nsamples);
    uint64_t st = tp->seed;
    std::vector<double> bl(nbatches);
    for (int64_t ep = 0; ep < tp->epochs; ep++) {
        for (int64_t q = 0; q < nsamples; q++) perm[q] = q;
        for (int64_t q = nsamples - 1; q > 0; q--) std::swap(perm[q], perm[splitmix(st) % (uint64_t)(q + 1)]);
        Tmp didx;
        EPSX_TRY(tmp_upload(ctx, didx, perm.data(), 8 * (size_t)nsamples));
        for (int64_t bt = 0; bt < nbatches; bt++) {
            const int64_t *ib = didx.ptr<int64_t>() + bt * batch;
            hipLaunchKernelGGL(net_gather_kernel, dim3(nblk(batch * n.c_in * n.length)), dim3(256), 0, ctx->stream,
                               batch, (int64_t)n.c_in, n.length, ib, dx.ptr<float>(), dptr<float>(n.X[0]));
            hipLaunchKernelGGL(net_gather_kernel, dim3(nblk(batch * n.c_out * n.length)), dim3(256), 0, ctx->stream,
                               batch, (int64_t)n.c_out, n.length, ib, dt.ptr<float>(), dptr<float>(n.T));
            EPSX_HIP(ctx, hipGetLastError());
            EPSX_TRY(net_backward_dev(ctx, batch, tp->precision == EPSX_PREC_BF16, dptr<float>(n.grad),
                                      dloss.ptr<double>() + bt));
            n.step++;
            const float c1 = (float)(1.0 - std::pow(tp->beta1, (double)n.step));
            const float c2 = (float)(1.0 - std::pow(tp->beta2, (double)n.step));
            n.wp_ok = false;
            hipLaunchKernelGGL(net_adam_kernel, dim3(nblk(n.nparam)), dim3(256), 0, ctx->stream, n.nparam,
                               dptr<float>(n.params), dptr<float>(n.grad), dptr<float>(n.m), dptr<float>(n.v),
                               (float)tp->lr, (float)tp->beta1, (float)tp->beta2, (float)tp->eps,
                               (float)tp->weight_decay, c1, c2);
            EPSX_HIP(ctx, hipGetLastError());
        }
        if (nbatches > 0)
            EPSX_HIP(ctx, hipMemcpyAsync(bl.data(), dloss.b.p, 8 * nbatches, hipMemcpyDeviceToHost, ctx->stream));
        EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
        double s = 0;
        for (int64_t b = 0; b < nbatches; b++) s += bl[b];
        epoch_loss[ep] = nbatches > 0 ? s / nbatches : NAN;
        if (tp->verbosity > 0) fprintf(stderr, "net epoch %lld: loss %.6e\n", (long long)ep + 1, epoch_loss[ep]);
    }
    return EPSX_OK;
}

}  // extern "C"
