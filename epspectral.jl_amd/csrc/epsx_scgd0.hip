// The scGD0 iteration on the device (src/scgd0.jl:189-346): set-up of the resident problem and
// the loop of Fan-Migdal sum, Kramers-Kronig, multiplet average, spectral DOS, μ solve and the
// convergence test.  Everything stays on the device; per iteration the host reads back only
// err, μ, min ∫A and a status word per temperature.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "epsx_common.h"
#include "epsx_device.h"

namespace epsx {
namespace {

constexpr double PI_D = 3.141592653589793;
constexpr double SQRT_EPS = 1.4901161193847656e-08;

__device__ inline double occ_fermion_d(double e, double T)
{
    if (T > SQRT_EPS) return 1.0 / (exp(e / T) + 1.0);
    const double s = e > 0 ? 1.0 : (e < 0 ? -1.0 : 0.0);
    return (1.0 - s) / 2.0;
}

// g2p[ν, p, i] = g2[ν, i, fsrc[p]]: the inner states in star order, contiguous per outer state.
__global__ void permute_g2_kernel(int64_t nm, int64_t ni, int64_t nf, const double *__restrict__ g2,
                                  const int32_t *__restrict__ fsrc, double *__restrict__ g2p)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nm * nf * ni) return;
    const int64_t nu = q % nm, p = (q / nm) % nf, i = q / (nm * nf);
    g2p[q] = g2[nu + nm * (i + ni * (int64_t)fsrc[p])];
}

// Σ_new = complex(ReΣ, ImΣ) for the slots of a pass (ReΣ = KK ImΣ from a library product).
__global__ void complex_pack_kernel(int64_t n, const int32_t *__restrict__ tidx, const int32_t *__restrict__ gate,
                                    const double *__restrict__ re, const double *__restrict__ im,
                                    double2 *__restrict__ Z)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int tcol = tidx[blockIdx.y];
    if (q >= n || (gate && !gate[tcol])) return;
    const int64_t o = q + n * (int64_t)tcol;
    Z[o] = make_double2(re[o], im[o]);
}

// ImΣ (FP64, column pitch ldd) from the FP32 result of a low-precision library product (column
// pitch nw), for the slots of a pass: ncol columns per slot.
__global__ void widen_kernel(int64_t nw, int64_t ncol, int64_t ldd, const int32_t *__restrict__ tidx,
                             const int32_t *__restrict__ gate, const float *__restrict__ src, double *__restrict__ dst)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int tcol = tidx[blockIdx.y];
    if (q >= nw * ncol || (gate && !gate[tcol])) return;
    const int64_t w = q % nw, c = q / nw;
    dst[w + ldd * c + ldd * ncol * (int64_t)tcol] = (double)src[q + nw * ncol * (int64_t)tcol];
}

// G̃c[:, c] = Σ_{members m of class c, in order} G̃[:, m] / |c|: the multiplet average of
// src/scgd0.jl:294-300 applied to the columns of the linear map ImΣ = h G̃.
// Gtc has the column pitch Kp >= K (rows K .. Kp - 1 zero), so an even pitch can be chosen for the
// product's 16-byte operand loads.
__global__ void gt_classes_kernel(int64_t K, int64_t Kp, int64_t nc, const int64_t *__restrict__ cptr,
                                  const int32_t *__restrict__ cidx, const double *__restrict__ Gt,
                                  double *__restrict__ Gtc)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= Kp * nc) return;
    const int64_t k = q % Kp, c = q / Kp;
    if (k >= K) { Gtc[q] = 0.0; return; }
    const int64_t a = cptr[c], b = cptr[c + 1];
    double sum = 0;
    for (int64_t j = a; j < b; j++) sum += Gt[k + K * (int64_t)cidx[j]];
    Gtc[q] = b - a > 1 ? sum / (double)(b - a) : sum;
}

// ImΣ[:, i] = ImΣc[:, class(i)] for the active slots (the members of a multiplet bitwise equal);
// ImΣc's columns at pitch ldc.
__global__ void cls_scatter_kernel(int64_t nw, int64_t ldc, int64_t ni, int64_t nc, const int32_t *__restrict__ cls,
                                   const int32_t *__restrict__ tidx, const double *__restrict__ ImSc,
                                   double *__restrict__ ImS, const int32_t *__restrict__ act)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nw * ni) return;
    const int tcol = tidx[blockIdx.y];
    if (act && !act[tcol]) return;
    const int64_t w = q % nw, i = q / nw;
    ImS[q + nw * ni * (int64_t)tcol] = ImSc[w + ldc * (int64_t)cls[i] + ldc * nc * (int64_t)tcol];
}

// Pseudo-random values in [-1, 1) (a hash of the index) for timing the library products'
// candidates: zero operands draw less power than a solve's, and so run at higher clocks.
template <typename E>
__global__ void fill_rand_kernel(E *p, int64_t n)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    uint64_t x = (uint64_t)q * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
    x ^= x >> 31;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 29;
    const double v = (double)(x >> 11) * 0x1.0p-52 - 1.0;
    if constexpr (sizeof(E) == 2) p[q] = (E)(__float_as_uint((float)v) >> 16);
    else p[q] = (E)v;
}

__global__ void fill_sigma_kernel(int64_t n, double2 *S, double eta_init)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < n) S[q] = make_double2(0.0, -eta_init);
}

// Multiplet average (src/scgd0.jl:294-300): members summed in order, divided by the count,
// written to every member (so members are bitwise equal).
__global__ void multiplet_kernel(int64_t nw, int64_t ni, int64_t n_mult, const int64_t *__restrict__ ptr,
                                 const int32_t *__restrict__ idx, const int32_t *__restrict__ tidx,
                                 double2 *__restrict__ Snew, const int32_t *__restrict__ act)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nw * n_mult) return;
    if (act && !act[tidx[blockIdx.y]]) return;
    const int64_t w = q % nw, m = q / nw;
    double2 *St = Snew + nw * ni * (int64_t)tidx[blockIdx.y];
    const int64_t a = ptr[m], b = ptr[m + 1];
    double sr = 0, si = 0;
    for (int64_t k = a; k < b; k++) {
        const double2 v = St[w + nw * (int64_t)(idx[k] - 1)];
        sr += v.x;
        si += v.y;
    }
    const double cnt = (double)(b - a);
    sr /= cnt;
    si /= cnt;
    for (int64_t k = a; k < b; k++) St[w + nw * (int64_t)(idx[k] - 1)] = make_double2(sr, si);
}

template <int NT>
__device__ inline double block_sum(double v, double *red)
{
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
#pragma unroll
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// The spectral DOS of Σ_new (already copied into Σ) in one pass over (a, i) (src/scgd0.jl:304-308):
// A[a, i] = -Im(1/z)/π = Im z / (π |z|²) with z = ωa - ε_i - Σ(ωa), formed in registers, never stored.
// One block per (tile of DOS_TA = 256 a, chunk of DOS_IC = 32 outer states, slot): lane l holds the
// DOS_RA = 4 consecutive a = 256 tile + 4 l + r, the 4 waves take the states i ≡ wave mod 4 of the
// chunk (DOS_KI = 8 each).  It writes the chunk's partial Dp[a, chunk] = Σ_i w_i A[a, i] (each
// wave's sum in i order, the wave sums in order) and, per i, the tile's partial integral
// ip[i, tile] = Σ_a A[a, i] Wa[a]: every lane's four terms in order, then the 64 lane values
// through LDS (lane 8 k + p of the wave sums the values 8 q + p of its k-th state in q order, then
// a butterfly over p).  dos_finish_kernel sums the partials in order.
constexpr int DOS_IC = 32, DOS_RA = 4, DOS_TA = 64 * DOS_RA, DOS_KI = DOS_IC / 4;
static_assert(DOS_KI * 8 == 64, "one ipart group of 8 lanes per state of a wave");
__global__ __launch_bounds__(256) void dos_kernel(int64_t nw, int64_t ni, int64_t nwa, const double *__restrict__ wsa,
                                                  const double *__restrict__ e_i, const int32_t *__restrict__ jsp,
                                                  const double *__restrict__ tsp, const int32_t *__restrict__ tidx,
                                                  const double2 *__restrict__ Sg, const double *__restrict__ w_i,
                                                  const double *__restrict__ Wa, double *__restrict__ Dp,
                                                  double *__restrict__ ipart, const int32_t *__restrict__ act)
{
    __shared__ double vred[4][DOS_KI][64];
    __shared__ double part[4][DOS_TA];
    const int z = blockIdx.z;
    if (act && !act[tidx[z]]) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t ntile = gridDim.x, nchunk = gridDim.y;
    const int64_t abase = (int64_t)blockIdx.x * DOS_TA, a0 = abase + DOS_RA * lane;
    const int64_t i0 = (int64_t)blockIdx.y * DOS_IC;
    const double2 *S = Sg + nw * ni * (int64_t)tidx[z];
    int j[DOS_RA];
    double t[DOS_RA], x[DOS_RA], wa[DOS_RA], acc[DOS_RA];
    bool on[DOS_RA];
#pragma unroll
    for (int r = 0; r < DOS_RA; r++) {
        const int64_t a = a0 + r;
        on[r] = a < nwa;
        j[r] = on[r] ? jsp[a] - 1 : 0;
        t[r] = on[r] ? tsp[a] : 0.0;
        x[r] = on[r] ? wsa[a] : 0.0;
        wa[r] = on[r] ? Wa[a] : 0.0;
        acc[r] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < DOS_KI; k++) {
        const int64_t i = i0 + wave + 4 * k;
        double v = 0;
        if (i < ni) {
            const double2 *St = S + nw * i;
            const double ei = e_i[i], wi = w_i[i];
#pragma unroll
            for (int r = 0; r < DOS_RA; r++) {
                if (!on[r]) continue;
                const double2 lo = St[j[r]], hi = St[j[r] + 1];
                const double sr = (1 - t[r]) * lo.x + t[r] * hi.x, si = (1 - t[r]) * lo.y + t[r] * hi.y;
                const double c = x[r] - ei - sr, d = -si;
                // -Im(1/z) / π = d / (π |z|²): one division
                const double A = d / (PI_D * fma(c, c, d * d));
                acc[r] = fma(A, wi, acc[r]);
                v = fma(A, wa[r], v);
            }
        }
        vred[wave][k][lane] = v;
    }
#pragma unroll
    for (int r = 0; r < DOS_RA; r++) part[wave][DOS_RA * lane + r] = acc[r];
    __syncthreads();
    {
        const int k = lane >> 3, p = lane & 7;
        double s = 0;
#pragma unroll
        for (int q = 0; q < 8; q++) s += vred[wave][k][8 * q + p];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        s += __shfl_xor(s, 4);
        const int64_t i = i0 + wave + 4 * k;
        if (p == 0 && i < ni) ipart[i + ni * (blockIdx.x + ntile * (int64_t)z)] = s;
    }
    const int64_t a = abase + threadIdx.x;
    if (a < nwa)
        Dp[a + nwa * (blockIdx.y + nchunk * (int64_t)z)] =
            ((part[0][threadIdx.x] + part[1][threadIdx.x]) + part[2][threadIdx.x]) + part[3][threadIdx.x];
}

// The sums after dos_kernel for the active slots, in one launch: D[a] = Wa[a] Σ_chunk Dp[a, chunk]
// (in chunk order; the first nblk(nwa) blocks), and ∫A dω of each outer state from the tile
// partials (the other blocks: one wave per (i, slot), lanes over the tiles, lane-strided then a
// butterfly: a fixed order).  it (graph mode): the device iteration counter, bumped once.
__global__ __launch_bounds__(256) void dos_finish_kernel(int64_t nwa, int64_t nchunk, int64_t ni, int64_t ntile,
                                                         const double *__restrict__ Dp, const double *__restrict__ Wa,
                                                         const double *__restrict__ ipart,
                                                         const int32_t *__restrict__ tidx, double *__restrict__ D,
                                                         double *__restrict__ integ, const int32_t *__restrict__ act,
                                                         int32_t *__restrict__ it)
{
    const int z = blockIdx.y;
    if (it && blockIdx.x == 0 && z == 0 && threadIdx.x == 0) *it += 1;
    if (act && !act[tidx[z]]) return;
    const int64_t nbd = (nwa + 255) / 256;
    if (blockIdx.x < nbd) {
        const int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
        if (a >= nwa) return;
        // in chunk order, eight loads in flight at a time
        const double *p = Dp + a + nwa * nchunk * (int64_t)z;
        double sum = 0;
        for (int64_t c0 = 0; c0 < nchunk; c0 += 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = c0 + u < nchunk ? p[nwa * (c0 + u)] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (c0 + u < nchunk) sum += v[u];
        }
        D[a + nwa * (int64_t)z] = sum * Wa[a];
        return;
    }
    const int lane = threadIdx.x & 63;
    const int64_t i = (blockIdx.x - nbd) * 4 + (threadIdx.x >> 6);
    if (i >= ni) return;
    double v = 0;
    for (int64_t k = lane; k < ntile; k += 64) v += ipart[i + ni * (k + ntile * (int64_t)z)];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) integ[i + ni * (int64_t)z] = v;
}

// Device-side bookkeeping of the graph-replayed loop: the history row of this iteration, the
// freeze of a converged temperature (src/scgd0.jl:326-333).
struct GraphBook {
    const int32_t *prec = nullptr;   // [nT] precision level (only FP64 may freeze); null: all FP64
    int32_t *act;            // [nT] convergence mask
    int64_t *frozen;         // [nT] iteration at which T froze
    const int32_t *it;       // iterations begun (dos_finish_kernel bumps it)
    int32_t *status = nullptr;   // [nT] μ status words (after act in one allocation), or null
    int32_t *host = nullptr;     // pinned host slots [2][act nT | status nT] (device address), or null
    double *host_err = nullptr;  // pinned host slots [2][nT] of err (device address), or null
    double *hist_err, *hist_mu, *hist_minA;  // [nT, maxiter]
    const double *mu, *minA;
    int64_t nT, maxiter;
    double tol;
};

// Ordered-bit midpoint of a bracket (Float64 bisection as in Roots.jl), 0 across a sign change.
__device__ inline double bits_middle(double a, double b)
{
    if ((a < 0) != (b < 0) && a != 0 && b != 0) return 0.0;
    if (a == 0 || b == 0) {
        const double o = a == 0 ? b : a;
        const unsigned long long uo = (unsigned long long)__double_as_longlong(fabs(o));
        const double r = __longlong_as_double((long long)(uo / 2));
        return o < 0 ? -r : r;
    }
    const unsigned long long ua = (unsigned long long)__double_as_longlong(fabs(a));
    const unsigned long long ub = (unsigned long long)__double_as_longlong(fabs(b));
    const unsigned long long m = ua / 2 + ub / 2 + (ua & ub & 1ULL);
    const double r = __longlong_as_double((long long)m);
    return a < 0 ? -r : r;
}

struct MuArgs {
    int64_t nwa;
    const double *wsa, *D;   // D [nwa, nz]
    const int32_t *tidx;
    const double *T, *nlist; // per temperature index
    double nelec, spin, W, nsb;
    int is_semi;
    double *mu;              // per temperature index
    int32_t *status;         // per temperature index
    const int32_t *act;      // convergence mask or null
    int64_t ni;
    const double *integ;     // ∫A per outer state [ni, nz]
    double *minA;            // per temperature index: min_i ∫A
};

// The smaller of two ∫A values, NaN winning: the min over a set is then the same in any order.
__device__ inline double nan_min(double m, double v) { return (v < m || v != v) ? v : m; }

// min of slot z's ∫A over the states i ≡ r (mod stride), r = the thread's rank among `nthr`
// (stride = nthr), then over the wave (a butterfly), in every lane; four loads in flight.
__device__ inline double integ_wave_min(int64_t ni, const double *__restrict__ integ, int z, int rank, int nthr)
{
    const double *p = integ + ni * (int64_t)z;
    double m = INFINITY;
    for (int64_t i0 = rank; i0 < ni; i0 += 4 * (int64_t)nthr) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = i0 + u * nthr < ni ? p[i0 + u * nthr] : INFINITY;
#pragma unroll
        for (int u = 0; u < 4; u++) m = nan_min(m, v[u]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = nan_min(m, __shfl_xor(m, o));
    return m;
}

// Sum over a wave by an xor butterfly: every lane adds the same pairs in the same order, so all
// lanes hold the bitwise same total.
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// The μ root of n(μ) = n_target on the bracket extrema(ωs_spectral) ± 0.1 (src/scgd0.jl:310-324,
// :361-366), one block of 16 waves per temperature.  Each n(μ) is a fixed-order sum (per thread,
// per wave, then the 16 wave totals in order), which every thread ends with, so every thread
// takes the same bisection step without a broadcast.  The bracket is bisected on the ordered bit
// pattern until two adjacent floats hold the sign change.  For T > sqrt(eps) each thread keeps
// its NREG values D_a and e^{±ωa/T} (as mantissa and binary exponent) in registers (nωa ≤ 1024
// NREG), so a step costs one exponential per thread: e^{(ωa-μ)/T} = e^{ωa/T} e^{-μ/T}.
template <int NREG>
__device__ __forceinline__ void mu_root(const MuArgs &m, int z, int tcol, double (*part)[16])
{
    const int tid = threadIdx.x, wave = tid >> 6;
    const double T = m.T[tcol];
    const double n_target = m.nlist[tcol] + m.nelec;
    const int holes = m.is_semi && m.nlist[tcol] < 0;
    const double *D = m.D + m.nwa * (int64_t)z;
    const bool regs = T > SQRT_EPS && m.nwa <= 1024 * NREG;
    double Dv[NREG], Mv[NREG];
    int kv[NREG];
    if (regs) {
#pragma unroll
        for (int r = 0; r < NREG; r++) {
            const int64_t a = tid + 1024 * r;
            Dv[r] = 0.0; Mv[r] = 1.0; kv[r] = 0;
            if (a < m.nwa) {
                Dv[r] = D[a];
                split_exp((holes ? -m.wsa[a] : m.wsa[a]) / T, Mv[r], kv[r]);
            }
        }
    }
    int buf = 0;
    auto excess = [&](double mu) -> double {
        double acc = 0;
        if (regs) {
            double Mm;
            int km;
            split_exp((holes ? mu : -mu) / T, Mm, km);
#pragma unroll
            for (int r = 0; r < NREG; r++) {
                int k = kv[r] + km;
                k = k > 900 ? 900 : (k < -900 ? -900 : k);
                // 1/y by the hardware reciprocal and two Newton steps (y in [1, 2^902]: the
                // reciprocal and the residuals stay normal)
                const double y = ldexp(Mv[r] * Mm, k) + 1.0;
                double q = __builtin_amdgcn_rcp(y);
                double e = fma(-y, q, 1.0);
                q = fma(q, e, q);
                e = fma(-y, q, 1.0);
                q = fma(q, e, q);
                acc = fma(Dv[r], q, acc);
            }
        } else {
            for (int64_t a = tid; a < m.nwa; a += 1024) {
                const double x = holes ? mu - m.wsa[a] : m.wsa[a] - mu;
                acc = fma(D[a], occ_fermion_d(x, T), acc);
            }
        }
        acc = wave_sum(acc);
        if ((tid & 63) == 0) part[buf][wave] = acc;
        __syncthreads();
        double s = 0;
#pragma unroll
        for (int w = 0; w < 16; w++) s += part[buf][w];
        buf ^= 1;
        if (holes) s = m.W - s;
        return (s + m.nsb) * m.spin - n_target;
    };
    // The previous iterate's μ is close to this root: two evaluations at μ_prev ∓ δ come first.
    // When they differ in sign, [μ_prev - δ, μ_prev + δ] is the bracket (n is monotone, so the
    // ends of the full bracket would differ in sign too); else the full bracket's ends are
    // evaluated (no root when they agree in sign) and narrowed to the side of μ_prev ∓ δ that
    // holds the sign change.  Every end kept still brackets the one sign change, so the root is
    // unchanged.
    double a = m.wsa[0] - 0.1, b = m.wsa[m.nwa - 1] + 0.1, fa, fb;
    const double mp = m.mu[tcol], dl = 1e-4 * (b - a);
    const double lo = mp - dl, hi = mp + dl;
    const bool near = lo > a && hi < b;
    double flo = 0, fhi = 0;
    if (near) {
        flo = excess(lo);
        fhi = excess(hi);
    }
    if (near && ((flo < 0 && fhi > 0) || (flo > 0 && fhi < 0))) {
        if (tid == 0) m.status[tcol] = 0;
        a = lo; fa = flo;
        b = hi; fb = fhi;
    } else {
        fa = excess(a);
        fb = excess(b);
        if (fa * fb > 0) {
            if (tid == 0) m.status[tcol] = 2;
            return;
        }
        if (tid == 0) m.status[tcol] = 0;
        if (fa == 0 || fb == 0) {
            if (tid == 0) m.mu[tcol] = fa == 0 ? a : b;
            return;
        }
        if (near) {
            if (flo == 0) {
                if (tid == 0) m.mu[tcol] = lo;
                return;
            }
            if ((flo > 0) == (fa > 0)) {
                a = lo; fa = flo;
                if (fhi == 0) {
                    if (tid == 0) m.mu[tcol] = hi;
                    return;
                }
                if ((fhi > 0) == (fa > 0)) { a = hi; fa = fhi; }
                else { b = hi; fb = fhi; }
            } else {
                b = lo; fb = flo;
            }
        }
    }
    // Illinois steps (regula falsi with the retained end's value halved) shrink the bracket until
    // it holds two adjacent floats, with ordered-bit bisection whenever the secant point is not
    // strictly inside (and after 40 steps).  n(μ) is a monotone fixed-order sum, so the adjacent
    // pair holding the sign change, and thus the root returned, is the one plain bisection reaches.
    double ga = fa, gb = fb;  // Illinois-scaled end values
    int side = 0;
    for (int it = 0; it < 300; it++) {
        const double bis = bits_middle(a, b);
        if (bis == a || bis == b) break;
        double x = bis;
        if (it < 40) {
            const double r = (a * gb - b * ga) / (gb - ga);
            if (r > a && r < b) x = r;
        }
        const double fm = excess(x);
        if (fm == 0) { a = b = x; fa = fb = 0; break; }
        if ((fm > 0) == (fa > 0)) {
            a = x; fa = ga = fm;
            if (side == -1) gb *= 0.5;
            side = -1;
        } else {
            b = x; fb = gb = fm;
            if (side == 1) ga *= 0.5;
            side = 1;
        }
    }
    if (tid == 0) m.mu[tcol] = fabs(fa) <= fabs(fb) ? a : b;
}

// The per-block err maxima of slot z at most ERR_NB per slot (64 lanes x 8).
constexpr int ERR_NB = 512;

// This lane's max of slot z's per-block err maxima (NaN wins): all its loads issued together.
__device__ inline double err_lane_max(int nb, const double *__restrict__ bmax, int z)
{
    const int lane = threadIdx.x & 63;
    double o[ERR_NB / 64];
#pragma unroll
    for (int r = 0; r < ERR_NB / 64; r++) {
        const int b = lane + 64 * r;
        o[r] = b < nb ? bmax[b + nb * z] : 0.0;
    }
    double m = 0;
#pragma unroll
    for (int r = 0; r < ERR_NB / 64; r++) m = (o[r] > m || o[r] != o[r]) ? o[r] : m;
    return m;
}

// The max over a wave (NaN wins), in every lane.
__device__ inline double err_wave_max(double m)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double v = __shfl_xor(m, o);
        m = (v > m || v != v) ? v : m;
    }
    return m;
}

// Temperature t's mask word and μ status into the host slot of this iteration (slot = 1-based
// iteration mod 2: the host reads one slot while the next iteration writes the other).  One thread.
__device__ inline void host_report(const GraphBook &g, int t, int32_t act)
{
    if (!g.host) return;
    int32_t *slot = g.host + 2 * g.nT * (*g.it & 1);
    slot[t] = act;
    slot[g.nT + t] = g.status ? g.status[t] : 0;
}

// err of temperature t = m, then in graph mode the history row and the freeze of a converged
// temperature (src/scgd0.jl:326-333), reported to the host.  One thread.
__device__ inline void err_book(double m, int t, double *__restrict__ err, const GraphBook &g)
{
    err[t] = m;
    if (g.host_err) g.host_err[g.nT * (*g.it & 1) + t] = m;
    if (g.act) {
        const int it = *g.it - 1;   // bumped by this iteration's dos_finish_kernel
        if (it < g.maxiter) {
            g.hist_err[t + g.nT * it] = m;
            g.hist_mu[t + g.nT * it] = g.mu[t];
            g.hist_minA[t + g.nT * it] = g.minA[t];
        }
        if (m < g.tol && (!g.prec || g.prec[t] == EPSX_PREC_FP64)) {
            g.act[t] = 0;
            g.frozen[t] = it + 1;
        }
        host_report(g, t, g.act[t]);
    }
}

// μ of every active slot (one block of 16 waves each), min_i ∫A, and the err bookkeeping of the
// iteration after μ (its history row records the new μ).
template <int NREG>
__global__ __launch_bounds__(1024) void mu_kernel(MuArgs m, int nb, const double *__restrict__ bmax,
                                                  double *__restrict__ err, GraphBook g)
{
    __shared__ double part[2][16];
    __shared__ double emax, amin[15];
    const int z = blockIdx.x;
    const int tcol = m.tidx[z];
    if (m.act && !m.act[tcol]) {
        if (threadIdx.x == 0) host_report(g, tcol, 0);
        return;
    }
    // wave 0 reduces the err maxima and waves 1-15 min ∫A first (their loads in flight with those
    // of D); the values wait in LDS, not in registers, during the root search
    if (threadIdx.x < 64) {
        const double e = err_wave_max(err_lane_max(nb, bmax, z));
        if (threadIdx.x == 0) emax = e;
    } else {
        const double a = integ_wave_min(m.ni, m.integ, z, threadIdx.x - 64, 960);
        if ((threadIdx.x & 63) == 0) amin[(threadIdx.x >> 6) - 1] = a;
    }
    mu_root<NREG>(m, z, tcol, part);
    // the root search passed a barrier after emax and amin were written; thread 0 wrote μ, and
    // the bookkeeping reads them on the same thread
    if (threadIdx.x == 0) {
        double a = amin[0];
#pragma unroll
        for (int w = 1; w < 15; w++) a = nan_min(a, amin[w]);
        m.minA[tcol] = a;
        err_book(emax, tcol, err, g);
    }
}

// err = max |Σ_new - Σ| over the slice, then Σ ← Σ_new (src/scgd0.jl:326-328): per-block maxima,
// reduced by err_book (in mu_kernel, or err_final_kernel with a fixed μ).
__global__ __launch_bounds__(256) void err_copy_kernel(int64_t n, const int32_t *__restrict__ tidx,
                                                       const double2 *__restrict__ Snew, double2 *__restrict__ S,
                                                       double *__restrict__ bmax, const int32_t *__restrict__ act)
{
    __shared__ double red[256];
    const int z = blockIdx.y;
    if (act && !act[tidx[z]]) return;
    const int64_t base = n * (int64_t)tidx[z];
    double m = 0;
    for (int64_t q = blockIdx.x * 256 + threadIdx.x; q < n; q += (int64_t)gridDim.x * 256) {
        const double2 a = Snew[base + q], b = S[base + q];
        const double e = hypot(a.x - b.x, a.y - b.y);
        m = (e > m || e != e) ? e : m;
        S[base + q] = a;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double o = red[threadIdx.x + s];
            if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) bmax[blockIdx.x + gridDim.x * z] = red[0];
}

// Σ_new = complex(ReΣ, ImΣ) from the columns of the library products (column cls[i] of outer
// state i: its multiplet class, or i itself with cls null), err = max |Σ_new - Σ| over the slice
// (per-block maxima as err_copy_kernel) and Σ ← Σ_new, for the slots of a pass: the pack, the
// multiplet copy and src/scgd0.jl:326-328 in one pass over Σ.  F: the products' type (FP64, or
// FP32 for the FP32 level, whose Kramers-Kronig product is the off-diagonal only: diag, the FP64
// principal-value diagonal, is added here as ReΣ = re + diag[ω] ImΣ).  re's columns at pitch nw, im's
// at pitch ldi.
template <typename F>
__global__ __launch_bounds__(256) void pack_err_kernel(int32_t nw, int32_t ni, int64_t nc, int64_t ldi,
                                                       const int32_t *__restrict__ cls,
                                                       const int32_t *__restrict__ tidx, const F *__restrict__ re,
                                                       const F *__restrict__ im, const double *__restrict__ diag,
                                                       double2 *__restrict__ S, double *__restrict__ bmax,
                                                       const int32_t *__restrict__ gate)
{
    __shared__ double red[256];
    const int z = blockIdx.y;
    const int tcol = tidx[z];
    if (gate && !gate[tcol]) return;
    const int32_t n = nw * ni;
    double2 *St = S + (int64_t)n * tcol;
    const F *rt = re + nw * nc * (int64_t)tcol, *it = im + ldi * nc * (int64_t)tcol;
    double m = 0;
    for (int32_t q = blockIdx.x * 256 + threadIdx.x; q < n; q += (int32_t)gridDim.x * 256) {
        const int32_t i = q / nw, w = q - i * nw;
        const int64_t c = cls ? cls[i] : i, o = w + (int64_t)nw * c;
        const double ai = (double)it[w + ldi * c];
        const double2 a = make_double2(diag ? (double)rt[o] + diag[w] * ai : (double)rt[o], ai), b = St[q];
        const double e = hypot(a.x - b.x, a.y - b.y);
        m = (e > m || e != e) ? e : m;
        St[q] = a;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double o = red[threadIdx.x + s];
            if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) bmax[blockIdx.x + gridDim.x * z] = red[0];
}

// The precision levels of a mixed-precision run, lowest first (the last is FP64).
struct Levels {
    int n;
    int prec[3];
};

// The product gates of a mixed-precision run: gates[l][t] = active and at level l; row L.n: active
// and needing the FP64 h of the star method (FP64, or every level if the low-precision products
// stage FP64 operands).  act null: every temperature active.
__global__ void phase_gates_kernel(int64_t nT, const int32_t *__restrict__ act, const int32_t *__restrict__ prec,
                                   Levels L, int lp_native, int32_t *__restrict__ gates)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const int a = act ? act[t] != 0 : 1;
    for (int l = 0; l < L.n; l++) gates[t + nT * l] = a && prec[t] == L.prec[l];
    gates[t + nT * L.n] = a && !(lp_native && prec[t] != EPSX_PREC_FP64);
}


// err and min_i ∫A of every active slot and the err bookkeeping, for a run with μ fixed.
__global__ __launch_bounds__(64) void err_final_kernel(int nb, const double *__restrict__ bmax,
                                                      const int32_t *__restrict__ tidx, double *__restrict__ err,
                                                      int64_t ni, const double *__restrict__ integ,
                                                      double *__restrict__ minA, GraphBook g)
{
    const int z = blockIdx.x;
    const int t = tidx[z];
    if (g.act && !g.act[t]) {
        if (threadIdx.x == 0) host_report(g, t, 0);
        return;
    }
    const double am = integ_wave_min(ni, integ, z, threadIdx.x, 64);
    const double e = err_wave_max(err_lane_max(nb, bmax, z));
    if (threadIdx.x == 0) {
        minA[t] = am;
        err_book(e, t, err, g);
    }
}

inline unsigned nblk(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace
}  // namespace epsx

using namespace epsx;

extern "C" int epsx_scgd0_set_problem(epsx_ctx *ctx, const epsx_scgd0_problem *P)
{
    if (!ctx) return EPSX_ERR_ARG;
    if (!P || P->struct_size != (int64_t)sizeof(epsx_scgd0_problem))
        return set_error(ctx, EPSX_ERR_ARG, "epsx_scgd0_problem: struct size mismatch (ABI %d)",
                         EPSX_ABI_VERSION);
    Scgd0Dev &s = ctx->s;
    s.ready = false;
    s.gt_ready = false;
    s.gb_ready = s.gf_ready = false;
    s.kkb_ready = s.kkf_ready = false;
    const int64_t nm = P->nm, ni = P->ni, nf = P->nf, nw = P->nw, nwa = P->nwa;
    if (nm < 1 || ni < 1 || nf < 1 || nw < 2 || nwa < 2 || P->nq < 1 || P->nki < 1 || P->nkf < 1)
        return set_error(ctx, EPSX_ERR_ARG, "scgd0 problem: empty dimension (nm=%lld ni=%lld nf=%lld nw=%lld nwa=%lld)",
                         (long long)nm, (long long)ni, (long long)nf, (long long)nw, (long long)nwa);
    if (!P->g2 || !P->wph || !P->iq_kk || !P->iks_i || !P->iks_f || !P->e_i || !P->w_i || !P->e_f ||
        !P->w_f || !P->f_to_i || !P->w_offsets || !P->wa_offsets)
        return set_error(ctx, EPSX_ERR_ARG, "scgd0 problem: null array");
    if (!(std::isfinite(P->w_origin) && std::isfinite(P->dw_unit) && P->dw_unit > 0) ||
        !(std::isfinite(P->wa_origin) && std::isfinite(P->dwa_unit) && P->dwa_unit > 0))
        return set_error(ctx, EPSX_ERR_ARG, "ω_origin must be finite and dω_unit finite and positive");
    int64_t gmin = INT64_MAX;
    for (int64_t k = 0; k + 1 < nw; k++) {
        const int64_t gp = P->w_offsets[k + 1] - P->w_offsets[k];
        if (gp <= 0) return set_error(ctx, EPSX_ERR_ARG, "offsets must be strictly increasing, got offsets[%lld] >= offsets[%lld]",
                                      (long long)k + 1, (long long)k + 2);
        gmin = gp < gmin ? gp : gmin;
    }
    for (int64_t k = 0; k + 1 < nwa; k++)
        if (P->wa_offsets[k + 1] <= P->wa_offsets[k])
            return set_error(ctx, EPSX_ERR_ARG, "ωs_spectral offsets must be strictly increasing");
    for (int64_t f = 0; f < nf; f++) {
        if (P->f_to_i[f] < 1 || P->f_to_i[f] > ni)
            return set_error(ctx, EPSX_ERR_ARG, "f_to_i[%lld] = %d is outside 1:%lld", (long long)f + 1, P->f_to_i[f], (long long)ni);
        if (P->iks_f[f] < 1 || P->iks_f[f] > P->nkf)
            return set_error(ctx, EPSX_ERR_ARG, "iks_f[%lld] = %d is outside 1:%lld", (long long)f + 1, P->iks_f[f], (long long)P->nkf);
    }
    for (int64_t i = 0; i < ni; i++)
        if (P->iks_i[i] < 1 || P->iks_i[i] > P->nki)
            return set_error(ctx, EPSX_ERR_ARG, "iks_i[%lld] = %d is outside 1:%lld", (long long)i + 1, P->iks_i[i], (long long)P->nki);
    for (int64_t q = 0; q < P->nki * P->nkf; q++)
        if (P->iq_kk[q] < 1 || P->iq_kk[q] > P->nq)
            return set_error(ctx, EPSX_ERR_ARG, "iq_kk holds %d, outside 1:%lld", P->iq_kk[q], (long long)P->nq);
    if (P->n_mult < 0 || (P->n_mult > 0 && (!P->mult_ptr || !P->mult_idx)))
        return set_error(ctx, EPSX_ERR_ARG, "multiplets: bad CSR");
    for (int64_t m = 0; m < P->n_mult; m++) {
        if (P->mult_ptr[m + 1] < P->mult_ptr[m]) return set_error(ctx, EPSX_ERR_ARG, "multiplets: bad CSR");
        for (int64_t k = P->mult_ptr[m]; k < P->mult_ptr[m + 1]; k++)
            if (P->mult_idx[k] < 1 || P->mult_idx[k] > ni)
                return set_error(ctx, EPSX_ERR_ARG, "multiplet member %d is outside 1:%lld", P->mult_idx[k], (long long)ni);
    }
    EPSX_HIP(ctx, hipSetDevice(ctx->device));

    s.nm = nm; s.ni = ni; s.nf = nf; s.nq = P->nq; s.nki = P->nki; s.nkf = P->nkf;
    s.nw = nw; s.nwa = nwa; s.n_mult = P->n_mult;
    s.W = P->W; s.nstates_base = P->nstates_base;
    s.w_origin = P->w_origin; s.dw_unit = P->dw_unit;
    s.dw_min = P->dw_unit * (double)gmin;
    s.ws_host.resize(nw);
    s.wsa_host.resize(nwa);
    grid_nodes(nw, P->w_origin, P->dw_unit, P->w_offsets, s.ws_host.data());
    grid_nodes(nwa, P->wa_origin, P->dwa_unit, P->wa_offsets, s.wsa_host.data());

    // Fan-Migdal grid data
    std::vector<int32_t> off, ltab;
    std::vector<double> rgap;
    if (fm_grid_host(nw, P->w_offsets, off, rgap, ltab, s.span))
        return set_error(ctx, EPSX_ERR_ARG, "ωs spans more than 2^30 steps of dω_unit");
    s.ltab_len = (int64_t)ltab.size();
    EPSX_TRY(upload(ctx, s.ws, s.ws_host.data(), sizeof(double) * nw));
    EPSX_TRY(upload(ctx, s.off, off.data(), sizeof(int32_t) * nw));
    EPSX_TRY(upload(ctx, s.rgap, rgap.data(), sizeof(double) * nw));
    EPSX_TRY(upload(ctx, s.ltab, ltab.data(), sizeof(int32_t) * ltab.size()));
    // spectral grid, its table from ωs and its trapezoid weights (src/scgd0.jl:163, :181-182)
    std::vector<int32_t> jsp(nwa);
    std::vector<double> tsp(nwa), Wa(nwa);
    linear_interpolation_table(nw, s.ws_host.data(), nwa, s.wsa_host.data(), jsp.data(), tsp.data());
    trapezoid_weights(nwa, s.wsa_host.data(), Wa.data());
    EPSX_TRY(upload(ctx, s.wsa, s.wsa_host.data(), sizeof(double) * nwa));
    EPSX_TRY(upload(ctx, s.Wa, Wa.data(), sizeof(double) * nwa));
    EPSX_TRY(upload(ctx, s.jsp, jsp.data(), sizeof(int32_t) * nwa));
    EPSX_TRY(upload(ctx, s.tsp, tsp.data(), sizeof(double) * nwa));
    // KK matrix (src/scgd0.jl:183)
    EPSX_TRY(dev_alloc(ctx, s.KK, sizeof(double) * nw * nw));
    EPSX_TRY(kk_build(ctx, nw, dptr<double>(s.ws), dptr<double>(s.KK)));
    EPSX_TRY(dev_alloc(ctx, s.KKo, sizeof(double) * nw * nw));
    EPSX_TRY(dev_alloc(ctx, s.KKd, sizeof(double) * nw));
    EPSX_TRY(kk_split(ctx, nw, dptr<double>(s.KK), dptr<double>(s.KKo), dptr<double>(s.KKd)));
    {
        // KK with an even column pitch, the A operand of the FP64 product's 16-byte loads
        const int64_t ldk = nw + (nw & 1);
        EPSX_TRY(dev_alloc(ctx, s.KKp, sizeof(double) * ldk * nw));
        EPSX_HIP(ctx, hipMemsetAsync(s.KKp.p, 0, sizeof(double) * ldk * nw, ctx->stream));
        EPSX_HIP(ctx, hipMemcpy2DAsync(s.KKp.p, sizeof(double) * ldk, s.KK.p, sizeof(double) * nw, sizeof(double) * nw,
                                       nw, hipMemcpyDeviceToDevice, ctx->stream));
    }

    // star order of the inner states
    std::vector<int32_t> fsrc;
    std::vector<int64_t> sptr;
    star_order(nf, ni, P->f_to_i, fsrc, sptr);
    std::vector<double> efs(nf), wfs(nf);
    std::vector<int32_t> fis(nf);
    s.star_ok = true;
    for (int64_t p = 0; p < nf; p++) {
        const int64_t f = fsrc[p];
        efs[p] = P->e_f[f];
        wfs[p] = P->w_f[f];
        fis[p] = P->f_to_i[f] - 1;
        if (!(P->e_f[f] == P->e_i[fis[p]])) s.star_ok = false;
    }
    s.nstar = 0;
    for (int64_t i = 0; i < ni; i++) s.nstar += sptr[i + 1] > sptr[i];
    EPSX_TRY(upload(ctx, s.ef_s, efs.data(), sizeof(double) * nf));
    EPSX_TRY(upload(ctx, s.wf_s, wfs.data(), sizeof(double) * nf));
    EPSX_TRY(upload(ctx, s.fi_s, fis.data(), sizeof(int32_t) * nf));
    EPSX_TRY(upload(ctx, s.fsrc_s, fsrc.data(), sizeof(int32_t) * nf));
    EPSX_TRY(upload(ctx, s.star_ptr, sptr.data(), sizeof(int64_t) * (ni + 1)));
    // Einstein modes: ωph[ν, :] constant
    s.einstein = true;
    s.wnu_host.assign(nm, 0.0);
    for (int64_t nu = 0; nu < nm; nu++) {
        s.wnu_host[nu] = P->wph[nu];
        for (int64_t q = 1; q < P->nq; q++)
            if (!(P->wph[nu + nm * q] == P->wph[nu])) s.einstein = false;
    }
    EPSX_TRY(upload(ctx, s.wph, P->wph, sizeof(double) * nm * P->nq));
    EPSX_TRY(upload(ctx, s.wnu, s.wnu_host.data(), sizeof(double) * nm));
    if (s.einstein) {
        std::vector<int32_t> hk;
        std::vector<double> ht;
        star_lookup_host(nw, nm, s.wnu_host.data(), P->dw_unit, off, ltab, rgap, s.span, hk, ht);
        EPSX_TRY(upload(ctx, s.hk, hk.data(), sizeof(int32_t) * hk.size()));
        EPSX_TRY(upload(ctx, s.ht, ht.data(), sizeof(double) * ht.size()));
    }
    {
        std::vector<int32_t> ch;
        s.sr_cap = star_stage_cap(nm, sptr);
        star_chunks(nm, s.sr_cap, sptr, ch);
        s.nchunks = (int64_t)ch.size() - 1;
        EPSX_TRY(upload(ctx, s.chunks, ch.data(), sizeof(int32_t) * ch.size()));
    }
    EPSX_TRY(upload(ctx, s.iq_kk, P->iq_kk, sizeof(int32_t) * P->nki * P->nkf));
    EPSX_TRY(upload(ctx, s.iks_i, P->iks_i, sizeof(int32_t) * ni));
    EPSX_TRY(upload(ctx, s.iks_f, P->iks_f, sizeof(int32_t) * nf));
    EPSX_TRY(upload(ctx, s.e_i, P->e_i, sizeof(double) * ni));
    EPSX_TRY(upload(ctx, s.w_i, P->w_i, sizeof(double) * ni));
    s.w_i_host.assign(P->w_i, P->w_i + ni);
    ctx->l.ready = false;  // a ladder problem belongs to the scGD0 problem it was set on
    ctx->l.bt_T = -1;
    s.nc = 0;
    if (P->n_mult > 0) {
        EPSX_TRY(upload(ctx, s.mult_ptr, P->mult_ptr, sizeof(int64_t) * (P->n_mult + 1)));
        EPSX_TRY(upload(ctx, s.mult_idx, P->mult_idx, sizeof(int32_t) * P->mult_ptr[P->n_mult]));
        // column classes of the star product in order of each class's first state (disjoint
        // multiplets only; otherwise the product keeps every column)
        std::vector<int64_t> mult_of(ni, -1);
        bool disjoint = true;
        for (int64_t m = 0; m < P->n_mult && disjoint; m++)
            for (int64_t k = P->mult_ptr[m]; k < P->mult_ptr[m + 1]; k++) {
                const int64_t i = P->mult_idx[k] - 1;
                if (mult_of[i] >= 0) { disjoint = false; break; }
                mult_of[i] = m;
            }
        if (disjoint) {
            std::vector<int32_t> cls(ni, -1), cidx;
            std::vector<int64_t> cptr{0};
            for (int64_t i = 0; i < ni; i++) {
                if (cls[i] >= 0) continue;
                const int32_t c = (int32_t)(cptr.size() - 1);
                if (mult_of[i] < 0) {
                    cls[i] = c;
                    cidx.push_back((int32_t)i);
                } else {
                    const int64_t m = mult_of[i];
                    for (int64_t k = P->mult_ptr[m]; k < P->mult_ptr[m + 1]; k++) {
                        cls[P->mult_idx[k] - 1] = c;
                        cidx.push_back(P->mult_idx[k] - 1);
                    }
                }
                cptr.push_back((int64_t)cidx.size());
            }
            s.nc = (int64_t)cptr.size() - 1;
            EPSX_TRY(upload(ctx, s.cls, cls.data(), sizeof(int32_t) * ni));
            EPSX_TRY(upload(ctx, s.cls_ptr, cptr.data(), sizeof(int64_t) * cptr.size()));
            EPSX_TRY(upload(ctx, s.cls_idx, cidx.data(), sizeof(int32_t) * cidx.size()));
        }
    }
    // g2: raw upload, then the star-ordered layout [ν, p, i] on the device
    {
        Tmp raw;
        const size_t n = (size_t)nm * ni * nf;
        EPSX_TRY(tmp_upload(ctx, raw, P->g2, sizeof(double) * n));
        EPSX_TRY(dev_alloc(ctx, s.g2p, sizeof(double) * (n + 2)));   // padding read by the reduction
        hipLaunchKernelGGL(permute_g2_kernel, dim3(nblk((int64_t)n)), dim3(256), 0, ctx->stream, nm, ni,
                           nf, raw.ptr<double>(), dptr<int32_t>(s.fsrc_s), dptr<double>(s.g2p));
        EPSX_HIP(ctx, hipGetLastError());
        EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    s.ready = true;
    return EPSX_OK;
}

namespace {

struct RunBufs {
    Tmp S, Snew, ImS, D, Dp, ipart, integ, part, bmax;
    Tmp T, mu, eta, nlist, err, minA, status, tidx;
};

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// The learned warm start inside one batched run: slots [0, nTu) carry the plain iteration from
// -iη_init, slots [nTu, 2 nTu) the warm starts (inactive until the prediction is written after the
// first iteration, from the G0D0 self-energy of the plain slot).  After the second iteration, the
// warm start's first, each temperature keeps the slot with the smaller residual: under the common
// contraction rate of the iteration that is the one fewer iterations away from tol (the plain
// start whenever the prediction does not cut the remaining iterations, or has no μ root); the
// other slot is dropped.
struct WarmHooks {
    int64_t nTu = 0;
    int precision = EPSX_PREC_BF16;
    bool judged = false;
    std::vector<char> dropped;        // [2 nTu]
    std::vector<int32_t> kept;        // [nTu] slot kept per temperature
    std::vector<double> err_pred, err_plain;
    double *Sigma_user = nullptr;     // the caller's Σ [nω, ni, nTu]: the kept slots are read into it
};

static int scgd0_run_impl(epsx_ctx *ctx, const epsx_scgd0_params *par, epsx_scgd0_outputs *out, WarmHooks *wh)
{
    if (!ctx) return EPSX_ERR_ARG;
    const double t_entry = now_s();
    Scgd0Dev &s = ctx->s;
    if (!s.ready) return set_error(ctx, EPSX_ERR_STATE, "epsx_scgd0_run before epsx_scgd0_set_problem");
    if (!par || par->struct_size != (int64_t)sizeof(epsx_scgd0_params) || !out ||
        out->struct_size != (int64_t)sizeof(epsx_scgd0_outputs))
        return set_error(ctx, EPSX_ERR_ARG, "epsx_scgd0_run: struct size mismatch (ABI %d)", EPSX_ABI_VERSION);
    const int64_t nT = par->nT, maxiter = par->maxiter;
    if (nT < 1 || !par->Tlist || !par->nlist || !par->mu_start || !par->eta_min)
        return set_error(ctx, EPSX_ERR_ARG, "scgd0 run: nT must be positive with Tlist, nlist, μ_start and η_min given");
    if (maxiter < 1) return set_error(ctx, EPSX_ERR_ARG, "maxiter must be at least 1, got %lld", (long long)maxiter);
    for (int64_t t = 0; t < nT; t++)
        if (!(par->Tlist[t] >= 0)) return set_error(ctx, EPSX_ERR_ARG, "Temperature must be positive");
    if (!out->Sigma || !out->mu || !out->converged || !out->hist_err || !out->hist_mu || !out->hist_minA ||
        !out->wall || !out->iter_frozen)
        return set_error(ctx, EPSX_ERR_ARG, "scgd0 run: null output array");
    if (par->fm_method < EPSX_FM_DIRECT || par->fm_method > EPSX_FM_AUTO)
        return set_error(ctx, EPSX_ERR_ARG, "unknown Fan-Migdal method %d", par->fm_method);
    if (par->precision != EPSX_PREC_FP64 && par->precision != EPSX_PREC_FP32 && par->precision != EPSX_PREC_BF16)
        return set_error(ctx, EPSX_ERR_ARG, "unknown precision %d", par->precision);
    // The Fan-Migdal method: the factorised forms only behind their preconditions.
    int method = par->fm_method;
    if (method == EPSX_FM_AUTO)
        method = s.einstein ? (s.star_ok ? EPSX_FM_STAR : EPSX_FM_EINSTEIN) : EPSX_FM_DIRECT;
    if (method == EPSX_FM_STAR && !(s.einstein && s.star_ok))
        return set_error(ctx, EPSX_ERR_ARG, "the star Fan-Migdal method needs dispersionless modes (ωph[ν, :] "
                         "constant) and ε_f == ε_i of the representative for every inner state");
    if (method == EPSX_FM_EINSTEIN && !s.einstein)
        return set_error(ctx, EPSX_ERR_ARG, "the Einstein Fan-Migdal method needs dispersionless modes (ωph[ν, :] "
                         "constant)");
    if (method == EPSX_FM_EINSTEIN) {
        // a mode below ω_acoustic contributes a zero column of H, which must meet finite couplings
        // (the term-by-term sum skips such a mode before reading g2, which may hold Inf there)
        bool any_ac = false;
        std::vector<uint8_t> skip(s.nm);
        for (int64_t nu = 0; nu < s.nm; nu++) any_ac |= (skip[nu] = s.wnu_host[nu] < par->w_acoustic) != 0;
        if (any_ac) {
            EPSX_TRY(upload(ctx, s.skip, skip.data(), s.nm));
            bool ok = true;
            EPSX_TRY(acoustic_finite_check(ctx, s.nm, s.nm * s.ni * s.nf, dptr<double>(s.g2p), dptr<uint8_t>(s.skip), &ok));
            if (!ok) {
                if (par->fm_method == EPSX_FM_EINSTEIN)
                    return set_error(ctx, EPSX_ERR_ARG, "the Einstein Fan-Migdal method needs finite couplings in the "
                                     "modes below ω_acoustic");
                method = EPSX_FM_DIRECT;
            }
        }
    }
    if (method == EPSX_FM_DIRECT && par->precision == EPSX_PREC_BF16)
        return set_error(ctx, EPSX_ERR_ARG, "the direct Fan-Migdal kernel runs in FP64 or FP32 (BF16 is for the "
                         "factorised methods' products)");
    EPSX_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t nw = s.nw, ni = s.ni, nwa = s.nwa, nf = s.nf;
    const size_t nS = (size_t)nw * ni * nT;
    const hipStream_t st = ctx->stream;

    RunBufs B;
    EPSX_TRY(tmp_alloc(ctx, B.S.b, 16 * nS));
    EPSX_TRY(tmp_alloc(ctx, B.Snew.b, 16 * nS));
    EPSX_TRY(tmp_alloc(ctx, B.ImS.b, 8 * nS));
    EPSX_TRY(tmp_alloc(ctx, B.D.b, 8 * (size_t)nwa * nT));
    EPSX_TRY(tmp_alloc(ctx, B.Dp.b, 8 * (size_t)nwa * ((ni + DOS_IC - 1) / DOS_IC) * nT));   // per-chunk D partials
    EPSX_TRY(tmp_alloc(ctx, B.ipart.b, 8 * (size_t)ni * ((nwa + DOS_TA - 1) / DOS_TA) * nT));            // per-tile integrals
    EPSX_TRY(tmp_alloc(ctx, B.integ.b, 8 * (size_t)ni * nT));
    size_t npart = 0;  // the largest partial buffer any launch of this run needs
    for (int64_t a = 1; a <= (par->batched ? nT : 1); a++) {
        const size_t e = fm_part_elems(nf, ni, nw, a, par->f_tile);
        npart = e > npart ? e : npart;
    }
    EPSX_TRY(tmp_alloc(ctx, B.part.b, 8 * npart));
    // the GEMM split-k workspace for every product of the run (the k split is at most 16), sized
    // before any graph capture: a change of precision level captures new products
    EPSX_TRY(dev_alloc(ctx, ctx->gemm_ws, 8 * (size_t)nw * ni * 16 * nT));
    for (int k = 0; k < 8; k++) { ctx->t_class[k] = 0; ctx->n_class[k] = 0; }
    Tmp hbuf, hocc, hb, hf, imb, imf, ImSc;
    // the star product's columns: the classes when multiplets are present (G̃c), else every state
    const bool use_cls = method == EPSX_FM_STAR && s.nc > 0;
    const int64_t ncol = use_cls ? s.nc : ni;
    // the star product's class columns at an even pitch (16-byte loads of Kramers-Kronig's B operand)
    const int64_t ldi = use_cls ? nw + (nw & 1) : nw, sI = ldi * ncol;
    // Low-precision levels of the star method: h and G̃ stored in BF16 / FP32, k-contiguous with the
    // row pitch kp (nm ni rounded up to a multiple of 8), multiplied by the low-precision MFMA
    // kernel.  EPSX_LP_STAGED=1 rounds the FP64 operands as they are staged instead (for tests).
    const char *staged_env = getenv("EPSX_LP_STAGED");
    const bool lp_native = method == EPSX_FM_STAR && par->precision != EPSX_PREC_FP64 &&
                           !(staged_env && staged_env[0] == '1');
    const bool use_bf = lp_native && par->precision == EPSX_PREC_BF16;   // BF16 level present
    const bool use_f32 = lp_native;                                       // FP32 level present
    const int64_t kp = (s.nm * ni + 7) / 8 * 8;
    const int64_t kpw = (nw + 7) / 8 * 8;
    // column pitches: h's a multiple of 64 (its builds store whole aligned 512-byte runs per wave,
    // and the FP64 product's 16-byte loads need an even pitch), G̃c's even (Kp); G̃ without classes
    // keeps the star reduction's pitch nm ni
    const int64_t ldh = (nw + 63) / 64 * 64, K = s.nm * ni, Kp = K + (K & 1), ldG = use_cls ? Kp : K;
    // Einstein method: H' for a chunk of ein_fc sorted inner states per slot (gpu_tile if given, else
    // chunks that keep H' of all temperatures within 2 GiB)
    int64_t ein_fc = 0;
    if (method == EPSX_FM_EINSTEIN) {
        const int64_t per_f = 8 * ldh * s.nm * nT;
        ein_fc = par->f_tile > 0 ? par->f_tile : (int64_t(2) << 30) / per_f;
        ein_fc = ein_fc < 1 ? 1 : (ein_fc > nf ? nf : ein_fc);
        EPSX_TRY(tmp_alloc(ctx, hbuf.b, (size_t)per_f * ein_fc));
        EPSX_TRY(tmp_alloc(ctx, hocc.b, 16 * (size_t)nw * s.nm * nT));
    }
    if (method == EPSX_FM_STAR) {
        EPSX_TRY(tmp_alloc(ctx, hbuf.b, 8 * (size_t)ldh * s.nm * ni * nT));
        EPSX_TRY(tmp_alloc(ctx, hocc.b, 16 * (size_t)nw * s.nm * nT));
        if (!(par->reuse_star && s.gt_ready && s.gt_wac == par->w_acoustic)) {
            // G̃[(ν, i'), i] = Σ_{f in star(i')} w_f g2[ν, i, f], acoustic modes zeroed
            std::vector<uint8_t> skip(s.nm);
            for (int64_t nu = 0; nu < s.nm; nu++) skip[nu] = s.wnu_host[nu] < par->w_acoustic;
            EPSX_TRY(upload(ctx, s.skip, skip.data(), s.nm));
            EPSX_TRY(dev_alloc(ctx, s.Gt, 8 * (size_t)s.nm * ni * ni));
            hipEvent_t e0, e1;
            EPSX_HIP(ctx, hipEventCreate(&e0));
            EPSX_HIP(ctx, hipEventCreate(&e1));
            EPSX_HIP(ctx, hipEventRecord(e0, st));
            EPSX_TRY(star_reduce_launch(ctx, s.sr_cap, s.nm, ni, nf, dptr<double>(s.g2p), dptr<double>(s.wf_s),
                                        dptr<int64_t>(s.star_ptr), dptr<int32_t>(s.chunks), s.nchunks,
                                        dptr<uint8_t>(s.skip), dptr<double>(s.Gt)));
            EPSX_HIP(ctx, hipEventRecord(e1, st));
            EPSX_HIP(ctx, hipEventSynchronize(e1));
            float ms = 0;
            EPSX_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            ctx->t_class[4] = ms * 1e-3;
            ctx->n_class[4] = 1;
            if (s.nc > 0) {
                EPSX_TRY(dev_alloc(ctx, s.Gtc, 8 * (size_t)Kp * s.nc));
                hipLaunchKernelGGL(gt_classes_kernel, dim3(nblk(Kp * s.nc)), dim3(256), 0, st, K, Kp, s.nc,
                                   dptr<int64_t>(s.cls_ptr), dptr<int32_t>(s.cls_idx), dptr<double>(s.Gt),
                                   dptr<double>(s.Gtc));
                EPSX_HIP(ctx, hipGetLastError());
            }
            s.gt_ready = true;
            s.gb_ready = s.gf_ready = false;
            s.gt_wac = par->w_acoustic;
        }
        const double *Gprod = use_cls ? dptr<double>(s.Gtc) : dptr<double>(s.Gt);
        if (use_bf && !s.gb_ready) {
            EPSX_TRY(dev_alloc(ctx, s.Gb, 2 * (size_t)kp * ncol));
            EPSX_TRY(gt_lp_launch(ctx, EPSX_PREC_BF16, K, ldG, kp, ncol, Gprod, s.Gb.p));
            s.gb_ready = true;
        }
        if (use_f32 && !s.gf_ready) {
            EPSX_TRY(dev_alloc(ctx, s.Gf, 4 * (size_t)kp * ncol));
            EPSX_TRY(gt_lp_launch(ctx, EPSX_PREC_FP32, K, ldG, kp, ncol, Gprod, s.Gf.p));
            s.gf_ready = true;
        }
        if (use_cls) EPSX_TRY(tmp_alloc(ctx, ImSc.b, 8 * (size_t)ldi * ncol * nT));
        if (use_bf) EPSX_TRY(tmp_alloc(ctx, hb.b, 2 * (size_t)kp * nw * nT));
        if (use_f32) EPSX_TRY(tmp_alloc(ctx, hf.b, 4 * (size_t)kp * nw * nT));
        // Kramers-Kronig's off-diagonal at the low-precision levels: KKo transposed once, ImΣ
        // converted per iteration (rows padded to kpw)
        if (use_bf && !s.kkb_ready) {
            EPSX_TRY(dev_alloc(ctx, s.KKb, 2 * (size_t)kpw * nw));
            EPSX_TRY(kk_lp_build(ctx, EPSX_PREC_BF16, nw, kpw, dptr<double>(s.KKo), s.KKb.p));
            s.kkb_ready = true;
        }
        if (use_f32 && !s.kkf_ready) {
            EPSX_TRY(dev_alloc(ctx, s.KKf, 4 * (size_t)kpw * nw));
            EPSX_TRY(kk_lp_build(ctx, EPSX_PREC_FP32, nw, kpw, dptr<double>(s.KKo), s.KKf.p));
            s.kkf_ready = true;
        }
        if (use_bf) EPSX_TRY(tmp_alloc(ctx, imb.b, 2 * (size_t)kpw * ni * nT));
        if (use_f32) EPSX_TRY(tmp_alloc(ctx, imf.b, 4 * (size_t)kpw * ni * nT));
    }
    const int nbmax = ERR_NB;
    EPSX_TRY(tmp_alloc(ctx, B.bmax.b, 8 * (size_t)nbmax * nT));
    EPSX_TRY(tmp_upload(ctx, B.T, par->Tlist, 8 * nT));
    EPSX_TRY(tmp_upload(ctx, B.mu, par->mu_start, 8 * nT));
    EPSX_TRY(tmp_upload(ctx, B.eta, par->eta_min, 8 * nT));
    EPSX_TRY(tmp_upload(ctx, B.nlist, par->nlist, 8 * nT));
    EPSX_TRY(tmp_alloc(ctx, B.err.b, 8 * nT));
    EPSX_TRY(tmp_alloc(ctx, B.minA.b, 8 * nT));
    EPSX_TRY(tmp_alloc(ctx, B.status.b, 4 * nT));
    EPSX_TRY(tmp_alloc(ctx, B.tidx.b, 4 * nT));
    EPSX_HIP(ctx, hipMemsetAsync(B.status.b.p, 0, 4 * nT, st));

    if (par->Sigma_init) {
        EPSX_TRY(write_from_host(ctx, B.S.b.p, par->Sigma_init, 16 * nS));
    } else {
        hipLaunchKernelGGL(fill_sigma_kernel, dim3(nblk((int64_t)nS)), dim3(256), 0, st, (int64_t)nS,
                           B.S.ptr<double2>(), par->eta_init);
        EPSX_HIP(ctx, hipGetLastError());
    }
    EPSX_HIP(ctx, hipMemcpyAsync(B.Snew.b.p, B.S.b.p, 16 * nS, hipMemcpyDeviceToDevice, st));

    FmGrid g;
    g.nw = nw; g.span = s.span; g.ltab_len = s.ltab_len; g.dw_unit = s.dw_unit;
    g.ws = dptr<double>(s.ws); g.off = dptr<int32_t>(s.off); g.rgap = dptr<double>(s.rgap);
    g.ltab = dptr<int32_t>(s.ltab);
    FmDirectArgs fa{};
    fa.nm = s.nm; fa.ni = ni; fa.nfs = nf;
    fa.g2p = dptr<double>(s.g2p);
    fa.g2p32 = nullptr;
    fa.ef = dptr<double>(s.ef_s); fa.wf = dptr<double>(s.wf_s);
    fa.fi = dptr<int32_t>(s.fi_s); fa.fsrc = dptr<int32_t>(s.fsrc_s);
    fa.wq_pair = nullptr;
    fa.wph = dptr<double>(s.wph); fa.iq_kk = dptr<int32_t>(s.iq_kk);
    fa.iks_i = dptr<int32_t>(s.iks_i); fa.iks_f = dptr<int32_t>(s.iks_f); fa.nki = s.nki;
    fa.S = B.S.ptr<double>(); fa.ncol = ni; fa.w_acoustic = par->w_acoustic;

    for (int64_t q = 0; q < nT * maxiter; q++) out->hist_err[q] = out->hist_mu[q] = out->hist_minA[q] = NAN;
    for (int64_t t = 0; t < nT; t++) out->iter_frozen[t] = 0;
    // (created once per context: each creation costs microseconds of host time per run)
    for (auto &e : ctx->run_ev)
        if (!e) EPSX_HIP(ctx, hipEventCreate(&e));
    hipEvent_t *const ev = ctx->run_ev;
    out->niter = 0;
    out->fm_method_used = method;
    out->n_low_precision = 0;
    int nb_err = (int)((nw * ni + 255) / 256);
    nb_err = nb_err > nbmax ? nbmax : nb_err;

    // One iteration of src/scgd0.jl:252-333 for the nA temperature slots in d_tidx.  With a
    // device mask (graph mode) every kernel skips converged temperatures, and the convergence
    // test writes the history and freezes temperatures on the device.
    // With a low-precision mode every temperature carries its own precision level (lvl): BF16, then
    // FP32, then FP64 for precision BF16; FP32, then FP64 for precision FP32.  The products run once
    // per level present among the active temperatures, gated on the device by level.
    const bool split = par->precision != EPSX_PREC_FP64;
    Levels L{};
    if (par->precision == EPSX_PREC_BF16) L = {3, {EPSX_PREC_BF16, EPSX_PREC_FP32, EPSX_PREC_FP64}};
    else if (par->precision == EPSX_PREC_FP32) L = {2, {EPSX_PREC_FP32, EPSX_PREC_FP64, 0}};
    else L = {1, {EPSX_PREC_FP64, 0, 0}};
    std::vector<int32_t> lvl(nT, 0), prec_t(nT, L.prec[0]);
    Tmp d_prec, d_gates;
    if (split) {
        EPSX_TRY(tmp_upload(ctx, d_prec, prec_t.data(), 4 * nT));
        EPSX_TRY(tmp_alloc(ctx, d_gates.b, 4 * (size_t)nT * (L.n + 1)));
    }
    std::vector<int32_t> active(nT, 1), act_idx;
    if (wh)
        for (int64_t t = wh->nTu; t < nT; t++) active[t] = 0;
    // The FP64 products run in the library's own FP64 MFMA GEMM (autotuned tile shape; on the t2g
    // 24³ shapes the 16-byte-load kernel), gated per temperature on the device: the star method's
    // h·G̃, and Kramers-Kronig's KK·ImΣ over the product's columns into ReS, packed into Σ with err
    // by one kernel.  EPSX_STAR_BLAS=1 hands both to hipBLASLt as plain strided-batched GEMMs (no
    // device mask: the host knows which slots take the FP64 pass, since a graph is captured again
    // whenever the active set changes).  The low-precision levels' products run in hipBLASLt unless
    // EPSX_STAR_BLAS=0 or EPSX_GEMM_SHAPE keeps every product in the library's kernels (and
    // Kramers-Kronig in the GEMM's complex epilogue).
    bool star_lt = false, kk_lt = false, lp_lt[3] = {false, false, false}, kklp_lt = false;
    Tmp ReS;    // KK ImΣ of the library product, per column of the star product (then packed into Σ)
    Tmp ImSf;   // FP32 result of the low-precision library products
    Tmp ReSf;   // the FP32 level's Kramers-Kronig off-diagonal by the library, per star column
    // (a run of one or two iterations does not pay for timing the library's candidates: it keeps
    // the library's own kernels, as it keeps their default tile shape)
    const char *blas_env = getenv("EPSX_STAR_BLAS"), *shape_env = getenv("EPSX_GEMM_SHAPE");
    const bool use_lt = maxiter >= 3 && !(blas_env && blas_env[0] == '0') && !(shape_env && *shape_env);
    const bool star_blas = use_lt && blas_env && blas_env[0] == '1';
    // Kramers-Kronig's FP64 product into ReS by the library's kernel, then packed with err
    const bool kk_own = use_lt && !star_blas && (use_cls || s.n_mult == 0) && nw * ni <= INT32_MAX - 65536 * 256;
    if (kk_own) EPSX_TRY(tmp_alloc(ctx, ReS.b, 8 * nS));
    {
        // Kramers-Kronig's FP64 product KK·ImΣ likewise (KK shared, over the star product's columns,
        // the complex result packed after)
        if (star_blas) {
            EPSX_TRY(tmp_alloc(ctx, ReS.b, 8 * nS));
            // candidates timed on pseudo-random operands (ImΣ is overwritten before its first use)
            hipLaunchKernelGGL(fill_rand_kernel<double>, dim3(nblk(nS)), dim3(256), 0, st, B.ImS.ptr<double>(), nS);
            if (use_cls)
                hipLaunchKernelGGL(fill_rand_kernel<double>, dim3(nblk(ldi * ncol * nT)), dim3(256), 0, st,
                                   ImSc.ptr<double>(), ldi * ncol * nT);
            EPSX_HIP(ctx, hipGetLastError());
            EPSX_TRY(blaslt_prepare(ctx, EPSX_PREC_FP64, nw, ncol, nw, dptr<double>(s.KK), nw, 0,
                                    use_cls ? ImSc.ptr<double>() : B.ImS.ptr<double>(), ldi, sI,
                                    ReS.ptr<double>(), nw, nw * ncol, nT, &kk_lt));
        }
    }
    if (method == EPSX_FM_STAR) {
        if (use_lt) {
            // candidates timed on pseudo-random h (overwritten before its first use); the FP64
            // product only when EPSX_STAR_BLAS=1 asks for it
            if (star_blas) {
                hipLaunchKernelGGL(fill_rand_kernel<double>, dim3(nblk(ldh * K * nT)), dim3(256), 0, st,
                                   hbuf.ptr<double>(), ldh * K * nT);
                EPSX_HIP(ctx, hipGetLastError());
            }
            if (star_blas)
                EPSX_TRY(blaslt_prepare(ctx, EPSX_PREC_FP64, nw, ncol, K, hbuf.ptr<double>(), ldh, ldh * K,
                                        use_cls ? dptr<double>(s.Gtc) : dptr<double>(s.Gt), ldG, 0,
                                        use_cls ? ImSc.ptr<double>() : B.ImS.ptr<double>(), ldi, sI, nT, &star_lt));
            // the low-precision levels' products from the stored BF16 / FP32 h and G̃ (FP32 result,
            // widened into ImΣ)
            if (lp_native) {
                EPSX_TRY(tmp_alloc(ctx, ImSf.b, 4 * nS));
                if (use_bf) {
                    hipLaunchKernelGGL(fill_rand_kernel<uint16_t>, dim3(nblk(kp * nw * nT)), dim3(256), 0, st,
                                       hb.ptr<uint16_t>(), kp * nw * nT);
                    EPSX_HIP(ctx, hipGetLastError());
                    EPSX_TRY(blaslt_prepare(ctx, EPSX_PREC_BF16, nw, ncol, kp, hb.b.p, kp, kp * nw, s.Gb.p, kp, 0,
                                            ImSf.ptr<float>(), nw, nw * ncol, nT, &lp_lt[EPSX_PREC_BF16]));
                }
                if (use_f32) {
                    hipLaunchKernelGGL(fill_rand_kernel<float>, dim3(nblk(kp * nw * nT)), dim3(256), 0, st,
                                       hf.ptr<float>(), kp * nw * nT);
                    EPSX_HIP(ctx, hipGetLastError());
                    EPSX_TRY(blaslt_prepare(ctx, EPSX_PREC_FP32, nw, ncol, kp, hf.b.p, kp, kp * nw, s.Gf.p, kp, 0,
                                            ImSf.ptr<float>(), nw, nw * ncol, nT, &lp_lt[EPSX_PREC_FP32]));
                    // the FP32 level's Kramers-Kronig off-diagonal over the same FP32 columns (the
                    // stored FP32 KK, k-contiguous rows of kpw), packed into Σ with the diagonal
                    if (lp_lt[EPSX_PREC_FP32] && (use_cls || s.n_mult == 0)) {
                        EPSX_TRY(tmp_alloc(ctx, ReSf.b, 4 * (size_t)nw * ncol * nT));
                        EPSX_TRY(blaslt_prepare(ctx, EPSX_PREC_FP32, nw, ncol, nw, s.KKf.p, kpw, 0, ImSf.ptr<float>(),
                                                nw, nw * ncol, ReSf.ptr<float>(), nw, nw * ncol, nT, &kklp_lt));
                    }
                }
            }
        }
    }
    // The FP64 pass packs Σ_new from the library products' columns and writes Σ ← Σ_new with err
    // in one kernel, unless multiplets without classes still need their average taken over Σ_new.
    // (pack_err_kernel indexes a slice in 32 bits)
    const bool kk_fuse = (kk_lt || kk_own) && (use_cls || s.n_mult == 0) && nw * ni <= INT32_MAX - 65536 * 256;
    // the FP32 level likewise: its star and Kramers-Kronig products by the library in FP32, one
    // pack of both into Σ with err
    if (!(nw * ni <= INT32_MAX - 65536 * 256)) kklp_lt = false;
    auto lp_fused = [&](int prec) { return kklp_lt && prec == EPSX_PREC_FP32; };
    // err and Σ ← Σ_new per pass when some pass packs its own
    const bool per_pass = kk_fuse || kklp_lt;
    // the levels present among the active temperatures, as a bit mask
    auto present_levels = [&]() {
        int mask = 0;
        for (int64_t t = 0; t < nT; t++)
            if (active[t]) mask |= 1 << lvl[t];
        return mask;
    };
    auto enqueue = [&](int nA, const int32_t *d_tidx, const std::vector<int32_t> &h_tidx, const int32_t *d_act,
                       const GraphBook &book, bool events) -> int {
        // The library products have no device mask: one strided-batched call per run of consecutive
        // slots that take a pass (active, at that level) with consecutive temperatures.
        auto prec_runs = [&](int prec, auto &&call) -> int {
            auto take = [&](int z) {
                const int32_t t = h_tidx[z];
                return active[t] && (prec == EPSX_PREC_FP64 ? (!split || prec_t[t] == EPSX_PREC_FP64) : prec_t[t] == prec);
            };
            for (int z = 0; z < nA;) {
                if (!take(z)) { z++; continue; }
                int z1 = z + 1;
                while (z1 < nA && take(z1) && h_tidx[z1] == h_tidx[z1 - 1] + 1) z1++;
                EPSX_TRY(call(z, z1 - z));
                z = z1;
            }
            return EPSX_OK;
        };
        auto fp64_runs = [&](auto &&call) { return prec_runs(EPSX_PREC_FP64, call); };
        if (events) EPSX_HIP(ctx, hipEventRecord(ev[0], st));
        // the (gate, precision) passes of the products: one FP64 pass, or one per level present
        struct Pass { const int32_t *gate; int prec; bool tune; };
        // (a run of one or two iterations does not pay for timing the GEMM tile shapes)
        Pass passes[3] = {{d_act, EPSX_PREC_FP64, maxiter >= 3}, {nullptr, 0, false}, {nullptr, 0, false}};
        int npass = 1;
        // FP64 h (FP64 level, or staged low precision) and BF16 / FP32 h (native levels) of the star
        // method
        bool need_h64 = true;
        const int32_t *gate_h64 = d_act, *gate_hbf = nullptr, *gate_hf = nullptr;
        if (split) {
            hipLaunchKernelGGL(phase_gates_kernel, dim3(nblk(nT)), dim3(256), 0, st, nT, d_act,
                               d_prec.ptr<int32_t>(), L, lp_native ? 1 : 0, d_gates.ptr<int32_t>());
            EPSX_HIP(ctx, hipGetLastError());
            const int mask = present_levels();
            npass = 0;
            need_h64 = false;
            gate_h64 = d_gates.ptr<int32_t>() + nT * L.n;
            for (int l = 0; l < L.n; l++) {
                if (!(mask & (1 << l))) continue;
                const int32_t *g = d_gates.ptr<int32_t>() + nT * l;
                passes[npass++] = {g, L.prec[l], false};
                if (lp_native && L.prec[l] == EPSX_PREC_BF16) gate_hbf = g;
                else if (lp_native && L.prec[l] == EPSX_PREC_FP32) gate_hf = g;
                else need_h64 = true;
            }
        }
        FmTemps tp;
        tp.T = B.T.ptr<double>(); tp.mu = B.mu.ptr<double>(); tp.eta_min = B.eta.ptr<double>();
        tp.act = d_act;
        // 1-3. clamp + Fan-Migdal (the clamp is applied where Σin is read)
        if (method == EPSX_FM_STAR) {
            // FP64 h for the FP64 pass; with native BF16, BF16 h for the low-precision pass
            EPSX_TRY(h_build_launch(ctx, nw, s.nm, ni, dptr<double>(s.ws), dptr<int32_t>(s.hk), dptr<double>(s.ht),
                                    dptr<double>(s.wnu), dptr<double>(s.e_i), B.S.ptr<double>(), B.T.ptr<double>(),
                                    B.mu.ptr<double>(), B.eta.ptr<double>(), d_tidx, nA, par->w_acoustic,
                                    hocc.ptr<double>(), need_h64 ? hbuf.ptr<double>() : nullptr, d_act, gate_h64,
                                    gate_hbf ? hb.ptr<uint16_t>() : nullptr, gate_hbf,
                                    gate_hf ? hf.ptr<float>() : nullptr, gate_hf, kp, ldh));
            if (events) EPSX_HIP(ctx, hipEventRecord(ev[5], st));
            // ImΣ[:, :, t] = h_z G̃ for every slot z in one batched launch (G̃ shared); with multiplets
            // the class columns (G̃c) into ImΣc, copied to every member afterwards
            double *Cprod = use_cls ? ImSc.ptr<double>() : B.ImS.ptr<double>();
            const double *Gprod = use_cls ? dptr<double>(s.Gtc) : dptr<double>(s.Gt);
            for (int q = 0; q < npass; q++) {
                GemmBatch bt;
                bt.n_inst = nA;
                bt.sA = ldh * K;
                bt.sC = sI;
                bt.idx = d_tidx;
                bt.imap = 4;
                bt.gate = passes[q].gate;
                bt.precision = passes[q].prec;
                bt.tune = passes[q].tune;
                if (events) bt.mark = ev[6];
                if (lp_native && passes[q].prec != EPSX_PREC_FP64 && lp_lt[passes[q].prec]) {
                    const bool bf = passes[q].prec == EPSX_PREC_BF16;
                    const int64_t esz = bf ? 2 : 4, sC = nw * ncol;
                    const char *hA = static_cast<const char *>(bf ? hb.b.p : hf.b.p);
                    EPSX_TRY(prec_runs(passes[q].prec, [&](int z, int n) {
                        return blaslt_gemm(ctx, passes[q].prec, nw, ncol, kp, hA + esz * kp * nw * z, kp, kp * nw,
                                           bf ? s.Gb.p : s.Gf.p, kp, 0, ImSf.ptr<float>() + sC * h_tidx[z], nw, sC, n);
                    }));
                    if (events) EPSX_HIP(ctx, hipEventRecord(ev[6], st));
                    if (events) ctx->n_class[7] = 1;
                    if (!lp_fused(passes[q].prec)) {
                        hipLaunchKernelGGL(widen_kernel, dim3(nblk(nw * ncol), nA), dim3(256), 0, st, nw, ncol, ldi,
                                           d_tidx, passes[q].gate, ImSf.ptr<float>(), Cprod);
                        EPSX_HIP(ctx, hipGetLastError());
                    }
                } else if (lp_native && passes[q].prec != EPSX_PREC_FP64) {
                    const bool bf = passes[q].prec == EPSX_PREC_BF16;
                    bt.sA = kp * nw;
                    EPSX_TRY(lpgemm_tn(ctx, passes[q].prec, nw, ncol, kp, bf ? hb.b.p : hf.b.p, kp,
                                       bf ? s.Gb.p : s.Gf.p, kp, Cprod, ldi, false, bt));
                    if (events) ctx->n_class[7] = 30;
                } else if (star_lt && passes[q].prec == EPSX_PREC_FP64) {
                    const int64_t sA = ldh * K;
                    EPSX_TRY(fp64_runs([&](int z, int n) {
                        return blaslt_gemm(ctx, EPSX_PREC_FP64, nw, ncol, K, hbuf.ptr<double>() + sA * z, ldh, sA, Gprod,
                                           ldG, 0, Cprod + sI * h_tidx[z], ldi, sI, n);
                    }));
                    if (events) EPSX_HIP(ctx, hipEventRecord(ev[6], st));
                    if (events) ctx->n_class[7] = 1;
                } else {
                    EPSX_TRY(dgemm(ctx, nw, ncol, K, hbuf.ptr<double>(), ldh, Gprod, ldG, Cprod, ldi, false, bt));
                    // (the staged low-precision products of EPSX_LP_STAGED=1 run lpgemm_kernel)
                    if (events) ctx->n_class[7] = passes[q].prec == EPSX_PREC_FP64 ? 2 + ctx->last_gemm_shape : 31;
                }
            }
            // ImΣ per outer state for the products that take it (the fused FP64 pack reads the class
            // columns)
            if (use_cls) {
                for (int q = 0; q < (per_pass ? npass : 1); q++) {
                    if ((kk_fuse && passes[q].prec == EPSX_PREC_FP64) || lp_fused(passes[q].prec)) continue;
                    hipLaunchKernelGGL(cls_scatter_kernel, dim3(nblk(nw * ni), nA), dim3(256), 0, st, nw, ldi, ni, s.nc,
                                       dptr<int32_t>(s.cls), d_tidx, ImSc.ptr<double>(), B.ImS.ptr<double>(),
                                       per_pass ? passes[q].gate : d_act);
                    EPSX_HIP(ctx, hipGetLastError());
                }
            }
        } else if (method == EPSX_FM_EINSTEIN) {
            // the occupations of every (ω, ν) per slot, then per chunk of sorted inner states H' and
            // ImΣ (+)= H' g2p[chunk rows] for every slot in one batched product (g2p shared), once
            // per precision pass (the staged BF16 / FP32 products round H' and g2 as they stage them)
            EPSX_TRY(h_build_launch(ctx, nw, s.nm, ni, dptr<double>(s.ws), dptr<int32_t>(s.hk), dptr<double>(s.ht),
                                    dptr<double>(s.wnu), dptr<double>(s.e_i), B.S.ptr<double>(), B.T.ptr<double>(),
                                    B.mu.ptr<double>(), B.eta.ptr<double>(), d_tidx, nA, par->w_acoustic,
                                    hocc.ptr<double>(), nullptr, d_act, nullptr, nullptr, nullptr, nullptr, nullptr, 0,
                                    ldh));
            for (int64_t p0 = 0; p0 < nf; p0 += ein_fc) {
                const int64_t np = ein_fc < nf - p0 ? ein_fc : nf - p0;
                EPSX_TRY(einstein_h_launch(ctx, nw, s.nm, ni, dptr<double>(s.ws), dptr<int32_t>(s.hk),
                                           dptr<double>(s.ht), dptr<double>(s.wnu), B.S.ptr<double>(), B.eta.ptr<double>(),
                                           d_tidx, nA, par->w_acoustic, hocc.ptr<double>(), p0, np, dptr<double>(s.ef_s),
                                           dptr<double>(s.wf_s), dptr<int32_t>(s.fi_s), hbuf.ptr<double>(), ldh,
                                           d_act));
                for (int q = 0; q < npass; q++) {
                    GemmBatch bt;
                    bt.n_inst = nA;
                    bt.sA = ldh * s.nm * np;
                    bt.sC = nw * ni;
                    bt.idx = d_tidx;
                    bt.imap = 4;
                    bt.gate = passes[q].gate;
                    bt.precision = passes[q].prec;
                    bt.tune = passes[q].tune;
                    EPSX_TRY(dgemm(ctx, nw, ni, s.nm * np, hbuf.ptr<double>(), ldh, dptr<double>(s.g2p) + s.nm * p0,
                                   s.nm * nf, B.ImS.ptr<double>(), nw, p0 > 0, bt));
                }
            }
        } else {
            for (int q = 0; q < npass; q++) {
                tp.act = passes[q].gate;
                if (par->batched) {
                    tp.nT = nA;
                    tp.tidx = d_tidx;
                    EPSX_TRY(fm_direct_launch_sp(ctx, g, fa, tp, dptr<int64_t>(s.star_ptr), par->f_tile,
                                                 passes[q].prec, B.part.ptr<double>(), (int64_t)npart,
                                                 B.ImS.ptr<double>(), false));
                } else {
                    for (int z = 0; z < nA; z++) {
                        tp.nT = 1;
                        tp.tidx = d_tidx + z;
                        EPSX_TRY(fm_direct_launch_sp(ctx, g, fa, tp, dptr<int64_t>(s.star_ptr), par->f_tile,
                                                     passes[q].prec, B.part.ptr<double>(), (int64_t)npart,
                                                     B.ImS.ptr<double>(), false));
                    }
                }
            }
            tp.act = d_act;
        }
        if (events) EPSX_HIP(ctx, hipEventRecord(ev[1], st));
        // 4. Re Σ by Kramers-Kronig, Σ_new = complex(KK ImΣ, ImΣ) in the GEMM's epilogue; the FP64
        // library product runs over the columns of the star product (one per multiplet class) and
        // its pack writes Σ ← Σ_new with err directly
        for (int q = 0; q < npass; q++) {
            GemmBatch bt;
            bt.n_inst = nA;
            bt.sB = bt.sC = nw * ni;
            bt.idx = d_tidx;
            bt.imap = 2 | 4;
            bt.gate = passes[q].gate;
            bt.tune = passes[q].tune;
            bt.Z = B.Snew.ptr<double2>();
            bt.Zim = B.ImS.ptr<double>();
            if (passes[q].prec == EPSX_PREC_FP64 && (kk_lt || kk_own)) {
                const int64_t sC = nw * ncol;
                const double *Cim = use_cls ? ImSc.ptr<double>() : B.ImS.ptr<double>();
                if (kk_lt) {
                    EPSX_TRY(fp64_runs([&](int z, int n) {
                        const int64_t t0 = h_tidx[z];
                        return blaslt_gemm(ctx, EPSX_PREC_FP64, nw, ncol, nw, dptr<double>(s.KK), nw, 0, Cim + sI * t0,
                                           ldi, sI, ReS.ptr<double>() + sC * t0, nw, sC, n);
                    }));
                } else {
                    GemmBatch bk;
                    bk.n_inst = nA;
                    bk.sB = sI;
                    bk.sC = sC;
                    bk.idx = d_tidx;
                    bk.imap = 2 | 4;
                    bk.gate = passes[q].gate;
                    bk.tune = passes[q].tune;
                    EPSX_TRY(dgemm(ctx, nw, ncol, nw, dptr<double>(s.KKp), nw + (nw & 1), Cim, ldi, ReS.ptr<double>(), nw,
                                   false, bk));
                }
                if (kk_fuse)
                    hipLaunchKernelGGL(pack_err_kernel<double>, dim3(nb_err, nA), dim3(256), 0, st, (int32_t)nw,
                                       (int32_t)ni, ncol, ldi, use_cls ? dptr<int32_t>(s.cls) : nullptr, d_tidx,
                                       ReS.ptr<double>(), Cim, (const double *)nullptr, B.S.ptr<double2>(),
                                       B.bmax.ptr<double>(), passes[q].gate);
                else
                    hipLaunchKernelGGL(complex_pack_kernel, dim3(nblk(nw * ni), nA), dim3(256), 0, st, nw * ni, d_tidx,
                                       passes[q].gate, ReS.ptr<double>(), B.ImS.ptr<double>(), B.Snew.ptr<double2>());
                EPSX_HIP(ctx, hipGetLastError());
            } else if (passes[q].prec == EPSX_PREC_FP64) {
                EPSX_TRY(dgemm(ctx, nw, ni, nw, dptr<double>(s.KK), nw, B.ImS.ptr<double>(), nw, nullptr, nw, false, bt));
            } else if (lp_fused(passes[q].prec)) {
                // the off-diagonal in FP32 by the library over the star product's FP32 columns; the
                // pack adds the FP64 principal-value diagonal and writes Σ ← Σ_new with err
                const int64_t sC = nw * ncol;
                EPSX_TRY(prec_runs(EPSX_PREC_FP32, [&](int z, int n) {
                    const int64_t t0 = h_tidx[z];
                    return blaslt_gemm(ctx, EPSX_PREC_FP32, nw, ncol, nw, s.KKf.p, kpw, 0, ImSf.ptr<float>() + sC * t0,
                                       nw, sC, ReSf.ptr<float>() + sC * t0, nw, sC, n);
                }));
                hipLaunchKernelGGL(pack_err_kernel<float>, dim3(nb_err, nA), dim3(256), 0, st, (int32_t)nw, (int32_t)ni,
                                   ncol, nw, use_cls ? dptr<int32_t>(s.cls) : nullptr, d_tidx, ReSf.ptr<float>(),
                                   ImSf.ptr<float>(), dptr<double>(s.KKd), B.S.ptr<double2>(), B.bmax.ptr<double>(),
                                   passes[q].gate);
                EPSX_HIP(ctx, hipGetLastError());
            } else if (lp_native) {
                // the off-diagonal part from stored BF16 / FP32 operands, the principal-value diagonal
                // in FP64 (the epilogue)
                const bool bf = passes[q].prec == EPSX_PREC_BF16;
                void *im = bf ? imb.b.p : imf.b.p;
                EPSX_TRY(lp_cols(ctx, passes[q].prec, nw, kpw, ni, B.ImS.ptr<double>(), nw, nw * ni, im, kpw * ni, nA,
                                 d_tidx, passes[q].gate));
                bt.sB = kpw * ni;
                bt.diag = dptr<double>(s.KKd);
                EPSX_TRY(lpgemm_tn(ctx, passes[q].prec, nw, ni, kpw, bf ? s.KKb.p : s.KKf.p, kpw, im, kpw, nullptr, nw,
                                   false, bt));
            } else {
                // the off-diagonal part in low precision, the principal-value diagonal in FP64
                bt.precision = passes[q].prec;
                bt.diag = dptr<double>(s.KKd);
                EPSX_TRY(dgemm(ctx, nw, ni, nw, dptr<double>(s.KKo), nw, B.ImS.ptr<double>(), nw, nullptr, nw, false, bt));
            }
        }
        if (events) EPSX_HIP(ctx, hipEventRecord(ev[2], st));
        // 5. multiplets (the star product with classes already gave every member the average)
        if (s.n_mult > 0 && !use_cls) {
            hipLaunchKernelGGL(multiplet_kernel, dim3(nblk(nw * s.n_mult), nA), dim3(256), 0, st, nw, ni, s.n_mult,
                               dptr<int64_t>(s.mult_ptr), dptr<int32_t>(s.mult_idx), d_tidx, B.Snew.ptr<double2>(),
                               d_act);
            EPSX_HIP(ctx, hipGetLastError());
        }
        // 8-9. err and Σ ← Σ_new for the slots whose pack did not (Σ_new is read from Σ from here)
        for (int q = 0; q < (per_pass ? npass : 1); q++) {
            if ((kk_fuse && passes[q].prec == EPSX_PREC_FP64) || lp_fused(passes[q].prec)) continue;
            hipLaunchKernelGGL(err_copy_kernel, dim3(nb_err, nA), dim3(256), 0, st, nw * ni, d_tidx,
                               B.Snew.ptr<double2>(), B.S.ptr<double2>(), B.bmax.ptr<double>(),
                               per_pass ? passes[q].gate : d_act);
            EPSX_HIP(ctx, hipGetLastError());
        }
        // 6. spectral DOS and min ∫A
        const unsigned ntile = nblk(nwa, DOS_TA), nchunk = nblk(ni, DOS_IC);
        hipLaunchKernelGGL(dos_kernel, dim3(ntile, nchunk, nA), dim3(256), 0, st, nw, ni, nwa, dptr<double>(s.wsa),
                           dptr<double>(s.e_i), dptr<int32_t>(s.jsp), dptr<double>(s.tsp), d_tidx,
                           B.S.ptr<double2>(), dptr<double>(s.w_i), dptr<double>(s.Wa), B.Dp.ptr<double>(),
                           B.ipart.ptr<double>(), d_act);
        EPSX_HIP(ctx, hipGetLastError());
        hipLaunchKernelGGL(dos_finish_kernel, dim3(nblk(nwa) + nblk(ni, 4), nA), dim3(256), 0, st, nwa,
                           (int64_t)nchunk, ni, (int64_t)ntile, B.Dp.ptr<double>(), dptr<double>(s.Wa),
                           B.ipart.ptr<double>(), d_tidx, B.D.ptr<double>(), B.integ.ptr<double>(), d_act,
                           d_act ? const_cast<int32_t *>(book.it) : nullptr);
        EPSX_HIP(ctx, hipGetLastError());
        // 7. μ (and min ∫A)
        if (!par->fix_mu) {
            MuArgs m;
            m.nwa = nwa; m.wsa = dptr<double>(s.wsa); m.D = B.D.ptr<double>(); m.tidx = d_tidx;
            m.T = B.T.ptr<double>(); m.nlist = B.nlist.ptr<double>();
            m.nelec = par->nelec; m.spin = par->spin_degeneracy; m.W = s.W; m.nsb = s.nstates_base;
            m.is_semi = par->is_semiconductor; m.mu = B.mu.ptr<double>();
            m.status = book.status ? book.status : B.status.ptr<int32_t>();
            m.act = d_act;
            m.ni = ni; m.integ = B.integ.ptr<double>(); m.minA = B.minA.ptr<double>();
            // the spectral points each thread keeps in registers: the fewest that hold nωa
            const auto mu_launch = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(nA), dim3(1024), 0, st, m, nb_err, B.bmax.ptr<double>(),
                                   B.err.ptr<double>(), book);
            };
            if (nwa <= 1024 * 2) mu_launch(mu_kernel<2>);
            else if (nwa <= 1024 * 4) mu_launch(mu_kernel<4>);
            else if (nwa <= 1024 * 6) mu_launch(mu_kernel<6>);
            else mu_launch(mu_kernel<8>);
            EPSX_HIP(ctx, hipGetLastError());
        }
        if (events) EPSX_HIP(ctx, hipEventRecord(ev[3], st));
        if (par->fix_mu) {
            hipLaunchKernelGGL(err_final_kernel, dim3(nA), dim3(64), 0, st, nb_err, B.bmax.ptr<double>(), d_tidx,
                               B.err.ptr<double>(), ni, B.integ.ptr<double>(), B.minA.ptr<double>(), book);
            EPSX_HIP(ctx, hipGetLastError());
        }
        if (events) EPSX_HIP(ctx, hipEventRecord(ev[4], st));
        return EPSX_OK;
    };
    auto add_event_times = [&](int nA) -> int {
        float ms[4];
        for (int k = 0; k < 4; k++) {
            EPSX_HIP(ctx, hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]));
            ctx->t_class[k] += ms[k] * 1e-3;
        }
        if (method == EPSX_FM_STAR) {
            // the main GEMM kernel(s) of the star product, without the split-k sum
            float mg = 0;
            EPSX_HIP(ctx, hipEventElapsedTime(&mg, ev[5], ev[6]));
            ctx->t_class[5] += mg * 1e-3;
            ctx->n_class[5] += 1;
            ctx->n_class[6] += nA;
        }
        if (method == EPSX_FM_EINSTEIN) ctx->n_class[6] += nA;
        ctx->n_class[0] += (par->batched || method != EPSX_FM_DIRECT) ? 1 : nA;
        ctx->n_class[1] += nA;
        ctx->n_class[2] += 1;
        ctx->n_class[3] += 1;
        return EPSX_OK;
    };
    auto no_root = [&](int64_t t) {
        const double n_target = par->nlist[t] + par->nelec;
        return set_error(ctx, EPSX_ERR_NO_MU_ROOT,
                         "no chemical potential in (%.17g, %.17g) gives %.17g electrons per cell at "
                         "T = %.17g: the window holds too few states on one side of the gap, or the "
                         "spectral weight it loses outside ωs_spectral moves the count past the target. "
                         "Pass occ.μlist and fix_μ = true, or widen the window or ωs.",
                         s.wsa_host[0] - 0.1, s.wsa_host[nwa - 1] + 0.1, n_target, par->Tlist[t]);
    };

    std::vector<double> h_err(nT), h_mu(nT), h_minA(nT), prev_err(nT, INFINITY);
    std::vector<int32_t> h_status(nT);
    std::vector<double> mu_host(par->mu_start, par->mu_start + nT);
    int rc = EPSX_OK;
    // Precision schedule (with FP32 / BF16): a temperature runs its Fan-Migdal products (and the
    // off-diagonal of KK) at its level, and may not freeze, until it is within 2 tol (then FP64) or
    // has reached the level's floor (then the next level).  A floor is two consecutive iterations
    // that do not contract (err > 0.95 err_prev) once err is two decades below its first value: the
    // early iterations of a strongly coupled solve may rise once without being near any floor.  At
    // most half of maxiter runs below FP64.
    std::vector<int64_t> lp_iters(nT, 0), lp_flat(nT, 0);
    std::vector<double> first_err(nT, NAN);
    std::vector<char> joined(nT, 0);        // slots that join the loop after its first iteration
    std::vector<int64_t> partner(nT, -1);   // and the slot whose first err they take
    // e: the err of every temperature in the iteration just read; stage: null when the stream is
    // idle at the next call (the levels are copied from prec_t), else a pinned ring of two slots
    // of nT levels, one per parity of the iteration read (the copy from one slot is ordered before
    // the event the host waits for two iterations later, when the slot is written again)
    auto update_phases = [&](const double *e, int32_t *stage, int64_t iter) -> int {
        bool changed = false, any_lp = false;
        for (int64_t t = 0; t < nT; t++) {
            if (lvl[t] == L.n - 1) continue;
            any_lp = true;
            if (joined[t]) { joined[t] = 0; continue; }   // no err of its own yet
            // a slot that joined after the first iteration measures its floor against its
            // partner's first err (the G0D0 residual), not against its own small first one
            if (lp_iters[t]++ == 0) first_err[t] = partner[t] >= 0 ? first_err[partner[t]] : e[t];
            const bool near = e[t] <= 2 * par->tol;
            lp_flat[t] = (e[t] > 0.95 * prev_err[t] && e[t] < 1e-2 * first_err[t]) ? lp_flat[t] + 1 : 0;
            const bool budget = 2 * lp_iters[t] >= par->maxiter;
            int next = lvl[t];
            if (near || budget || !(e[t] == e[t])) next = L.n - 1;
            else if (lp_flat[t] >= 2) next = lvl[t] + 1;
            if (next != lvl[t]) {
                lvl[t] = next;
                prec_t[t] = L.prec[next];
                lp_flat[t] = 0;
                changed = true;
            }
            prev_err[t] = e[t];
        }
        if (any_lp) out->n_low_precision++;
        if (changed) {
            const int32_t *src = prec_t.data();
            if (stage) {
                int32_t *slot = stage + nT * (iter & 1);
                std::copy(prec_t.begin(), prec_t.end(), slot);
                src = slot;
            }
            EPSX_HIP(ctx, hipMemcpyAsync(d_prec.b.p, src, 4 * nT, hipMemcpyHostToDevice, st));
        }
        return EPSX_OK;
    };
    const double t_loop = now_s();
    // Σ slices of temperatures frozen in the batched loop, copied to pinned host memory on the side
    // stream while the others iterate (their Σ no longer changes)
    const size_t slice = 16 * (size_t)nw * ni;
    std::vector<char> sent(nT, 0);
    bool any_sent = false;
    // Every exit waits for the side stream's copies out of Σ once any was issued: Σ is a temporary,
    // and the pool may hand it to the next call's work on the main stream.
    struct SideGuard {
        epsx_ctx *c; const bool *sent;
        ~SideGuard() { if (*sent) (void)hipStreamSynchronize(c->side); }
    } side_guard{ctx, &any_sent};
    if (!par->batched) {
        // The reference's loop: the active temperatures one after another, decisions on the host.
        GraphBook none{};
        for (int64_t iter = 1; iter <= maxiter; iter++) {
            const double t0 = now_s();
            act_idx.clear();
            for (int64_t t = 0; t < nT; t++) if (active[t]) act_idx.push_back((int32_t)t);
            const int nA = (int)act_idx.size();
            EPSX_HIP(ctx, hipMemcpyAsync(B.tidx.b.p, act_idx.data(), 4 * nA, hipMemcpyHostToDevice, st));
            EPSX_TRY(enqueue(nA, B.tidx.ptr<int32_t>(), act_idx, nullptr, none, true));
            EPSX_HIP(ctx, hipMemcpyAsync(h_err.data(), B.err.b.p, 8 * nT, hipMemcpyDeviceToHost, st));
            EPSX_HIP(ctx, hipMemcpyAsync(h_mu.data(), B.mu.b.p, 8 * nT, hipMemcpyDeviceToHost, st));
            EPSX_HIP(ctx, hipMemcpyAsync(h_minA.data(), B.minA.b.p, 8 * nT, hipMemcpyDeviceToHost, st));
            EPSX_HIP(ctx, hipMemcpyAsync(h_status.data(), B.status.b.p, 4 * nT, hipMemcpyDeviceToHost, st));
            EPSX_HIP(ctx, hipStreamSynchronize(st));
            EPSX_TRY(add_event_times(nA));
            for (int z = 0; z < nA; z++) {
                const int64_t t = act_idx[z];
                if (!par->fix_mu && h_status[t] == 2) { rc = no_root(t); break; }
                mu_host[t] = h_mu[t];
                out->hist_minA[t + nT * (iter - 1)] = h_minA[t];
                out->hist_err[t + nT * (iter - 1)] = h_err[t];
                out->hist_mu[t + nT * (iter - 1)] = h_mu[t];
                if (lvl[t] == L.n - 1 && h_err[t] < par->tol) {
                    active[t] = 0;
                    out->iter_frozen[t] = iter;
                }
            }
            if (rc != EPSX_OK) break;
            out->wall[iter - 1] = now_s() - t0;
            out->niter = iter;
            if (split) EPSX_TRY(update_phases(h_err.data(), nullptr, iter));
            if (par->verbosity > 0) {
                fprintf(stderr, "scGD0 iteration %lld: wall %.3f s\n", (long long)iter, out->wall[iter - 1]);
                for (int z = 0; z < nA; z++)
                    fprintf(stderr, "  T=%g err=%.6g mu=%.12g\n", par->Tlist[act_idx[z]], h_err[act_idx[z]], h_mu[act_idx[z]]);
            }
            bool any = false;
            for (int64_t t = 0; t < nT; t++) any |= active[t] != 0;
            if (!any) break;
        }
    } else {
        // Batched: every temperature in each launch, converged ones masked on the device.  The
        // first two iterations run eagerly (the GEMM autotuner sees the shapes in the first, HIP
        // events time the kernel classes in the second); the iteration is then captured in a
        // hipGraph and replayed (captured again when the set of precision levels present changes),
        // the host reading back only the mask, the status words and err.
        std::vector<int32_t> all(nT);
        for (int64_t t = 0; t < nT; t++) all[t] = (int32_t)t;
        Tmp d_act, d_book;
        // the mask and the μ status words in one allocation; the convergence kernel reports both
        // into pinned host slots every iteration
        std::vector<int32_t> act_status(2 * nT, 0);
        std::copy(active.begin(), active.end(), act_status.begin());
        EPSX_TRY(tmp_upload(ctx, d_act, act_status.data(), 8 * nT));
        // (two slots: an iteration may be read back while the next one runs)
        // (fine-grained: the convergence kernel writes them, the host reads them after the
        // iteration's event)
        // (kept in the context for the next run: pinned allocations cost tens of µs each; the
        // previous run's work on the stream has ended, so no kernel still writes them)
        if (ctx->pin_slots_bytes < 32 * (size_t)nT) {
            if (ctx->pin_slots) (void)hipHostFree(ctx->pin_slots);
            ctx->pin_slots = nullptr;
            ctx->pin_slots_bytes = 0;
            EPSX_HIP(ctx, hipHostMalloc(&ctx->pin_slots, 32 * nT, hipHostMallocMapped | hipHostMallocCoherent));
            ctx->pin_slots_bytes = 32 * nT;
        }
        if (ctx->pin_prec_bytes < 8 * (size_t)nT) {
            if (ctx->pin_prec) (void)hipHostFree(ctx->pin_prec);
            ctx->pin_prec = nullptr;
            ctx->pin_prec_bytes = 0;
            EPSX_HIP(ctx, hipHostMalloc(&ctx->pin_prec, 8 * nT, hipHostMallocDefault));
            ctx->pin_prec_bytes = 8 * nT;
        }
        int32_t *h_as = static_cast<int32_t *>(ctx->pin_slots), *d_as = nullptr;
        double *h_es = reinterpret_cast<double *>(static_cast<char *>(ctx->pin_slots) + 16 * nT), *d_es = nullptr;   // and err, for the decision to run ahead
        int32_t *h_pr = static_cast<int32_t *>(ctx->pin_prec);   // the staged levels of a mixed-precision run
        EPSX_HIP(ctx, hipHostGetDevicePointer((void **)&d_es, h_es, 0));
        std::fill(h_es, h_es + 2 * nT, INFINITY);
        EPSX_HIP(ctx, hipHostGetDevicePointer((void **)&d_as, h_as, 0));
        hipEvent_t ev_read[2] = {nullptr, nullptr};
        struct ReadGuard { hipEvent_t *e; ~ReadGuard() { for (int k = 0; k < 2; k++) if (e[k]) (void)hipEventDestroy(e[k]); } } rg{ev_read};
        for (auto &e : ev_read) EPSX_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        // the bookkeeping in one device buffer, uploaded (and at the end read back with μ) through
        // one pinned staging buffer kept in the context: iter_frozen [nT] (zero), the iteration
        // counter (zero, in 8 bytes), then the err, μ and min ∫A histories [nT, maxiter]
        const size_t nh = (size_t)nT * maxiter, book_bytes = 8 * ((size_t)nT + 1 + 3 * nh);
        if (ctx->pin_book_bytes < book_bytes + 8 * (size_t)nT) {
            if (ctx->pin_book) (void)hipHostFree(ctx->pin_book);
            ctx->pin_book = nullptr;
            ctx->pin_book_bytes = 0;
            // at least 256 KB, so a short first run (a prime call) leaves room for the long ones
            const size_t want = std::max(book_bytes + 8 * (size_t)nT, (size_t)256 << 10);
            EPSX_HIP(ctx, hipHostMalloc(&ctx->pin_book, want, hipHostMallocDefault));
            ctx->pin_book_bytes = want;
        }
        char *const hbook = static_cast<char *>(ctx->pin_book);
        std::memset(hbook, 0, 8 * ((size_t)nT + 1));
        std::memcpy(hbook + 8 * (nT + 1), out->hist_err, 8 * nh);
        std::memcpy(hbook + 8 * (nT + 1 + nh), out->hist_mu, 8 * nh);
        std::memcpy(hbook + 8 * (nT + 1 + 2 * nh), out->hist_minA, 8 * nh);
        EPSX_TRY(tmp_alloc(ctx, d_book.b, book_bytes));
        EPSX_HIP(ctx, hipMemcpyAsync(d_book.b.p, hbook, book_bytes, hipMemcpyHostToDevice, st));
        char *const dbook = d_book.ptr<char>();
        EPSX_HIP(ctx, hipMemcpyAsync(B.tidx.b.p, all.data(), 4 * nT, hipMemcpyHostToDevice, st));
        GraphBook book;
        book.act = d_act.ptr<int32_t>(); book.status = d_act.ptr<int32_t>() + nT; book.host = d_as; book.host_err = d_es; book.frozen = reinterpret_cast<int64_t *>(dbook); book.it = reinterpret_cast<int32_t *>(dbook + 8 * nT);
        book.hist_err = reinterpret_cast<double *>(dbook + 8 * (nT + 1)); book.hist_mu = book.hist_err + nh; book.hist_minA = book.hist_mu + nh;
        book.mu = B.mu.ptr<double>(); book.minA = B.minA.ptr<double>();
        book.nT = nT; book.maxiter = maxiter; book.tol = par->tol;
        if (split) book.prec = d_prec.ptr<int32_t>();
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        int graph_levels = 0;
        std::vector<int32_t> graph_active, graph_prec;   // the active set and levels captured for
        // the HIP events time the second iteration (warm caches, the GEMM tile shapes chosen in the
        // first), or the first of a one-iteration run
        auto timed = [&](int64_t iter) { return iter == 2 || (iter == 1 && maxiter == 1); };
        struct GraphGuard {
            hipGraph_t *g; hipGraphExec_t *e; hipStream_t st;
            ~GraphGuard()
            {
                // an iteration launched ahead may still run the graph
                (void)hipStreamSynchronize(st);
                if (*e) (void)hipGraphExecDestroy(*e);
                if (*g) (void)hipGraphDestroy(*g);
            }
        } gg{&graph, &exec, st};
        // Replayed iterations go to the GPU one ahead of the host: iteration k + 1 is
        // launched before the host reads back iteration k's mask, so the GPU does not wait for the
        // host between replays.  An iteration launched ahead of the host's view runs the device-masked
        // kernels on the current mask; the library products (no device mask) keep the active set of
        // the graph, and their output for a temperature frozen meanwhile is never read.  Once some
        // active temperature's err is within 2 tol (it may freeze in the iteration on the GPU), the
        // host waits instead, so that no replay runs products for a temperature just frozen or after
        // the last iteration.  Mixed-precision runs set each temperature's level on the host from
        // the err it has read: with iteration k + 1 already launched, a level chosen from iteration
        // k's err takes effect from iteration k + 2 (from k + 1 while the host waits near tol).
        // an active temperature within 2 tol after the last iteration the host has read
        auto near_tol = [&](int64_t iter_read) {
            const double *e = h_es + nT * (iter_read & 1);
            for (int64_t t = 0; t < nT; t++)
                if (active[t] && !(e[t] >= 2 * par->tol)) return true;
            return false;
        };
        int64_t launched = 0;
        double t_prev = now_s();
        // the replayed iteration captured (and instantiated) into exec: nothing runs while the stream
        // captures, and work enqueued before the capture runs on
        auto capture = [&]() -> int {
            graph_levels = present_levels();
            graph_active = active;
            graph_prec = prec_t;
            EPSX_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int r = enqueue((int)nT, B.tidx.ptr<int32_t>(), all, book.act, book, false);
            hipError_t ec = hipStreamEndCapture(st, &graph);
            if (r != EPSX_OK) return r;
            if (ec != hipSuccess) return hip_fail(ctx, ec, "hipStreamEndCapture", __FILE__, __LINE__);
            EPSX_HIP(ctx, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            return EPSX_OK;
        };
        auto launch = [&](int64_t iter) -> int {
            if (iter <= 2) {
                EPSX_TRY(enqueue((int)nT, B.tidx.ptr<int32_t>(), all, book.act, book, timed(iter)));
            } else {
                // a new set of precision levels present, or (hipBLASLt products, which have no
                // device mask) a new active set or level of some temperature: the launches change,
                // capture again
                if (exec && (present_levels() != graph_levels ||
                             ((star_lt || kk_lt || lp_lt[1] || lp_lt[2]) &&
                              (active != graph_active || prec_t != graph_prec)))) {
                    EPSX_HIP(ctx, hipStreamSynchronize(st));
                    EPSX_HIP(ctx, hipGraphExecDestroy(exec));
                    EPSX_HIP(ctx, hipGraphDestroy(graph));
                    exec = nullptr;
                    graph = nullptr;
                }
                if (!exec) EPSX_TRY(capture());
                EPSX_HIP(ctx, hipGraphLaunch(exec, st));
            }
            EPSX_HIP(ctx, hipEventRecord(ev_read[iter & 1], st));
            launched = iter;
            return EPSX_OK;
        };
        bool side_ok = 16 * nS <= (size_t(4) << 30);
        if (side_ok && !ctx->side) EPSX_HIP(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
        if (side_ok && ctx->pin_side_bytes < 16 * nS) {
            EPSX_HIP(ctx, hipStreamSynchronize(ctx->side));
            if (ctx->pin_side) (void)hipHostFree(ctx->pin_side);
            ctx->pin_side = nullptr;
            ctx->pin_side_bytes = 0;
            EPSX_HIP(ctx, hipHostMalloc(&ctx->pin_side, 16 * nS, hipHostMallocDefault));
            ctx->pin_side_bytes = 16 * nS;
        }
        for (int64_t iter = 1; iter <= maxiter; iter++) {
            if (launched < iter) EPSX_TRY(launch(iter));
            if (iter >= 3 && iter < maxiter && !near_tol(iter - 1)) EPSX_TRY(launch(iter + 1));
            // the replayed iteration captured while the second, eager one runs on the GPU (the host
            // would otherwise capture with the GPU idle before the third); captured again before
            // its first replay if the levels or, with library products, the active set change
            if (iter == 2 && maxiter >= 3 && !exec) EPSX_TRY(capture());
            EPSX_HIP(ctx, hipEventSynchronize(ev_read[iter & 1]));
            const int32_t *slot = h_as + 2 * nT * (iter & 1);
            std::copy(slot, slot + nT, active.begin());
            std::copy(slot + nT, slot + 2 * nT, h_status.begin());
            if (timed(iter)) EPSX_TRY(add_event_times((int)nT));
            if (wh && iter == 2) {
                // the judge: per temperature, the slot whose residual is smaller goes on (a warm slot
                // without a μ root is dropped whatever its residual)
                const int64_t nu = wh->nTu;
                const double *e = h_es + nT * (iter & 1);
                std::vector<int32_t> zeros(nT, 0);
                for (int64_t t = 0; t < nu; t++) {
                    const int64_t w = nu + t;
                    wh->err_plain[t] = e[t];
                    wh->err_pred[t] = h_status[w] == 2 ? INFINITY : e[w];
                    const bool warm = wh->err_pred[t] < wh->err_plain[t];
                    const int64_t drop = warm ? t : w;
                    wh->kept[t] = (int32_t)(warm ? w : t);
                    wh->dropped[drop] = 1;
                    active[drop] = 0;
                    h_status[drop] = 0;
                    EPSX_HIP(ctx, hipMemcpyAsync(book.act + drop, zeros.data(), 4, hipMemcpyHostToDevice, st));
                    EPSX_HIP(ctx, hipMemcpyAsync(book.status + drop, zeros.data(), 4, hipMemcpyHostToDevice, st));
                }
                EPSX_HIP(ctx, hipStreamSynchronize(st));
                wh->judged = true;
            }
            for (int64_t t = 0; t < nT; t++)
                if (!par->fix_mu && h_status[t] == 2) { rc = no_root(t); break; }
            if (rc != EPSX_OK) break;
            if (wh && iter == 1) {
                // the prediction from the G0D0 self-energy of every plain slot into its warm slot, μ
                // with it; the warm slots join the second iteration
                const int64_t nu = wh->nTu;
                std::vector<int32_t> tab(2 * nu), ones(nu, 1);
                for (int64_t t = 0; t < nu; t++) { tab[t] = (int32_t)t; tab[nu + t] = (int32_t)(nu + t); }
                Tmp d_tab;
                EPSX_TRY(tmp_upload(ctx, d_tab, tab.data(), 8 * nu));
                EPSX_TRY(net_predict_sigma(ctx, nw, ni, (int)nu, d_tab.ptr<int32_t>(), d_tab.ptr<int32_t>() + nu,
                                           B.S.ptr<double2>(), dptr<double>(s.ws), dptr<double>(s.e_i), B.mu.ptr<double>(),
                                           B.T.ptr<double>(), wh->precision));
                EPSX_HIP(ctx, hipMemcpyAsync(B.mu.ptr<double>() + nu, B.mu.ptr<double>(), 8 * nu, hipMemcpyDeviceToDevice, st));
                EPSX_HIP(ctx, hipMemcpyAsync(book.act + nu, ones.data(), 4 * nu, hipMemcpyHostToDevice, st));
                EPSX_HIP(ctx, hipStreamSynchronize(st));
                for (int64_t t = nu; t < nT; t++) {
                    active[t] = joined[t] = 1;
                    partner[t] = t - nu;
                }
            }
            for (int64_t t = 0; side_ok && t < nT; t++) {
                if (active[t] || sent[t]) continue;
                if (wh && (!wh->judged || wh->dropped[t])) continue;
                EPSX_HIP(ctx, hipStreamWaitEvent(ctx->side, ev_read[iter & 1], 0));
                EPSX_HIP(ctx, hipMemcpyAsync(static_cast<char *>(ctx->pin_side) + slice * t, B.S.ptr<char>() + slice * t,
                                             slice, hipMemcpyDeviceToHost, ctx->side));
                sent[t] = 1;
                any_sent = true;
            }
            const double t1 = now_s();
            out->wall[iter - 1] = t1 - t_prev;
            t_prev = t1;
            out->niter = iter;
            if (split) {
                // frozen temperatures are not in a phase any more
                for (int64_t t = 0; t < nT; t++)
                    if (!active[t]) { lvl[t] = L.n - 1; prec_t[t] = EPSX_PREC_FP64; }
                EPSX_TRY(update_phases(h_es + nT * (iter & 1), h_pr, iter));
            }
            bool any = false;
            for (int64_t t = 0; t < nT; t++) any |= active[t] != 0;
            if (par->verbosity > 0) fprintf(stderr, "scGD0 iteration %lld: wall %.4f s\n", (long long)iter, out->wall[iter - 1]);
            if (!any) break;
        }
        if (rc == EPSX_OK) {
            EPSX_HIP(ctx, hipMemcpyAsync(hbook, dbook, book_bytes, hipMemcpyDeviceToHost, st));
            EPSX_HIP(ctx, hipMemcpyAsync(hbook + book_bytes, B.mu.b.p, 8 * nT, hipMemcpyDeviceToHost, st));
            EPSX_HIP(ctx, hipStreamSynchronize(st));
            std::memcpy(out->iter_frozen, hbook, 8 * nT);
            std::memcpy(out->hist_err, hbook + 8 * (nT + 1), 8 * nh);
            std::memcpy(out->hist_mu, hbook + 8 * (nT + 1 + nh), 8 * nh);
            std::memcpy(out->hist_minA, hbook + 8 * (nT + 1 + 2 * nh), 8 * nh);
            std::memcpy(mu_host.data(), hbook + book_bytes, 8 * nT);
            // history beyond the last iteration is NaN, as in the reference
            for (int64_t q = nT * out->niter; q < nT * maxiter; q++)
                out->hist_err[q] = out->hist_mu[q] = out->hist_minA[q] = NAN;
        }
    }
    if (rc != EPSX_OK) return rc;
    const double t_read = now_s();
    if (wh) {
        // the warm start: only the slot kept for each temperature comes back, into the caller's Σ
        if (!wh->judged)   // the run ended before the judge (maxiter or every slot frozen): plain
            for (int64_t t = 0; t < wh->nTu; t++) wh->kept[t] = (int32_t)t;
        if (any_sent) EPSX_HIP(ctx, hipStreamSynchronize(ctx->side));
        for (int64_t t = 0; t < wh->nTu; t++) {
            const int64_t k = wh->kept[t];
            char *dst = reinterpret_cast<char *>(wh->Sigma_user) + slice * t;
            if (sent[k]) parallel_copy(dst, static_cast<char *>(ctx->pin_side) + slice * k, slice);
            else EPSX_TRY(read_back(ctx, dst, B.S.ptr<char>() + slice * k, slice));
        }
    } else if (any_sent) {
        EPSX_HIP(ctx, hipStreamSynchronize(ctx->side));
        for (int64_t t = 0; t < nT; t++) {
            char *dst = reinterpret_cast<char *>(out->Sigma) + slice * t;
            if (sent[t]) parallel_copy(dst, static_cast<char *>(ctx->pin_side) + slice * t, slice);
            else EPSX_TRY(read_back(ctx, dst, B.S.ptr<char>() + slice * t, slice));
        }
    } else {
        EPSX_TRY(read_back(ctx, out->Sigma, B.S.b.p, 16 * nS));
    }
    if (par->verbosity > 0)
        fprintf(stderr, "scGD0 run: set-up %.4f s, iterations %.4f s, Σ read back %.4f s\n", t_loop - t_entry,
                t_read - t_loop, now_s() - t_read);
    for (int64_t t = 0; t < nT; t++) {
        out->mu[t] = mu_host[t];
        out->converged[t] = !active[t];
    }
    return EPSX_OK;
}

extern "C" int epsx_scgd0_run(epsx_ctx *ctx, const epsx_scgd0_params *par, epsx_scgd0_outputs *out)
{
    if (!ctx) return EPSX_ERR_ARG;
    if (!par || par->struct_size != (int64_t)sizeof(epsx_scgd0_params) || !out ||
        out->struct_size != (int64_t)sizeof(epsx_scgd0_outputs))
        return set_error(ctx, EPSX_ERR_ARG, "epsx_scgd0_run: struct size mismatch (ABI %d)", EPSX_ABI_VERSION);
    if (!par->warm_start) {
        if (out->warm_kept)
            for (int64_t t = 0; t < par->nT; t++) out->warm_kept[t] = 0;
        return scgd0_run_impl(ctx, par, out, nullptr);
    }
    // the warm start: the run over 2 nT slots (plain, then warm), its results composed per
    // temperature from the slot kept
    const int64_t nT = par->nT, maxiter = par->maxiter;
    if (nT < 1 || !par->Tlist || !par->nlist || !par->mu_start || !par->eta_min)
        return set_error(ctx, EPSX_ERR_ARG, "scgd0 run: nT must be positive with Tlist, nlist, μ_start and η_min given");
    if (par->Sigma_init) return set_error(ctx, EPSX_ERR_ARG, "the warm start replaces Σ_init: pass one or the other");
    if (!par->batched) return set_error(ctx, EPSX_ERR_ARG, "the warm start runs in the batched loop (batched = 1)");
    if (maxiter < 3) return set_error(ctx, EPSX_ERR_ARG, "the warm start needs maxiter >= 3, got %lld", (long long)maxiter);
    if (par->warm_start != EPSX_PREC_BF16 && par->warm_start != EPSX_PREC_FP32)
        return set_error(ctx, EPSX_ERR_ARG, "warm_start must be 0, EPSX_PREC_FP32 or EPSX_PREC_BF16, got %d",
                         par->warm_start);
    if (!ctx->net.ready) return set_error(ctx, EPSX_ERR_STATE, "the warm start needs a network: call epsx_net_create");
    if (!ctx->s.ready) return set_error(ctx, EPSX_ERR_STATE, "epsx_scgd0_run before epsx_scgd0_set_problem");
    if (ctx->net.c_in != 5 || ctx->net.c_out != 2 || ctx->net.length != ctx->s.nw)
        return set_error(ctx, EPSX_ERR_ARG, "the warm-start network takes 5 input and 2 output channels over the %lld "
                         "nodes of ωs, got %d, %d and %lld", (long long)ctx->s.nw, ctx->net.c_in, ctx->net.c_out,
                         (long long)ctx->net.length);
    if (!out->Sigma || !out->mu || !out->converged || !out->hist_err || !out->hist_mu || !out->hist_minA ||
        !out->wall || !out->iter_frozen)
        return set_error(ctx, EPSX_ERR_ARG, "scgd0 run: null output array");
    const int64_t n2 = 2 * nT;
    auto dup = [&](const double *x) {
        std::vector<double> v(n2);
        for (int64_t t = 0; t < n2; t++) v[t] = x[t % nT];
        return v;
    };
    std::vector<double> T2 = dup(par->Tlist), nl2 = dup(par->nlist), mu2 = dup(par->mu_start), eta2 = dup(par->eta_min);
    epsx_scgd0_params p2 = *par;
    p2.nT = n2;
    p2.Tlist = T2.data(); p2.nlist = nl2.data(); p2.mu_start = mu2.data(); p2.eta_min = eta2.data();
    p2.warm_start = 0;
    std::vector<double> m2(n2), he(n2 * maxiter), hm(n2 * maxiter), ha(n2 * maxiter);
    std::vector<int32_t> c2(n2);
    std::vector<int64_t> f2(n2);
    epsx_scgd0_outputs o2 = *out;   // Σ: the run reads the kept slots into out->Sigma itself
    o2.mu = m2.data(); o2.converged = c2.data(); o2.hist_err = he.data(); o2.hist_mu = hm.data();
    o2.hist_minA = ha.data(); o2.iter_frozen = f2.data(); o2.warm_kept = nullptr;
    WarmHooks wh;
    wh.nTu = nT;
    wh.precision = par->warm_start;
    wh.dropped.assign(n2, 0);
    wh.kept.assign(nT, 0);
    wh.err_pred.assign(nT, NAN);
    wh.err_plain.assign(nT, NAN);
    wh.Sigma_user = out->Sigma;
    const int rc = scgd0_run_impl(ctx, &p2, &o2, &wh);
    if (rc != EPSX_OK) return rc;
    out->niter = o2.niter;
    out->fm_method_used = o2.fm_method_used;
    out->n_low_precision = o2.n_low_precision;
    for (int64_t t = 0; t < nT; t++) {
        const int64_t k = wh.kept[t];
        out->mu[t] = m2[k];
        out->converged[t] = c2[k];
        out->iter_frozen[t] = f2[k];
        for (int64_t it = 0; it < maxiter; it++) {
            const int64_t src = it == 0 ? t : k;   // iteration 1: the G0D0 of the plain slot
            out->hist_err[t + nT * it] = he[src + n2 * it];
            out->hist_mu[t + nT * it] = hm[src + n2 * it];
            out->hist_minA[t + nT * it] = ha[src + n2 * it];
        }
        if (out->warm_kept) out->warm_kept[t] = k >= nT;
    }
    return EPSX_OK;
}
