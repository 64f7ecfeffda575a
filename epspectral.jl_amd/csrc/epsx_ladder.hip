// The ladder (vertex-corrected Kubo conductivity) on the device: src/ladder_kernel.jl and the
// application / solve loop of src/ladder.jl:245-379.
//
// The shift sums of apply_ladder_kernel! are reorganised exactly.  The reference forms, for each
// sign s, C_s = X_s B̃ with X_s[(j, q), (m, b, ip)] = (S_{s,m} J_ip^b)[j, q] the shifted vertex
// (q = (c12, c0, re/im)).  The shift S_{s,m} acts on the frequency index only (two taps,
// wlo/whi at klo), so with W_m = J B̃_m, where J[(j', q), (b, ip)] is the vertex itself,
//
//   C_s[(j, q), col] = Σ_m wlo[j,m,s] W_m[(klo[j,m,s], q), col] + whi[j,m,s] W_m[(klo+1, q), col].
//
// W_m does not depend on the sign, so both signs share one GEMM of
// (16 nω) x (3 ni) x (6 ni nshift) -- half the reference's flops, no gather in the operand --
// followed by a two-tap shift-add.  B̃ depends only on T and is built once per temperature.
#include <cmath>
#include <complex>
#include <cstdlib>
#include <vector>

#include <rocblas/rocblas.h>

#include "epsx_common.h"
#include "epsx_device.h"

namespace epsx {
namespace {

constexpr double SQRT_EPS_L = 1.4901161193847656e-08;

__device__ inline double occ_fermion_l(double e, double T)
{
    if (T > SQRT_EPS_L) return 1.0 / (exp(e / T) + 1.0);
    const double s = e > 0 ? 1.0 : (e < 0 ? -1.0 : 0.0);
    return (1.0 - s) / 2.0;
}

__device__ inline double occ_boson_l(double e, double T)
{
    if (T > SQRT_EPS_L) return e == 0.0 ? -0.5 : 1.0 / expm1(e / T);
    return e > 0 ? 0.0 : (e == 0 ? -0.5 : -1.0);
}

// Julia's inv(::ComplexF64) (Smith's algorithm) for finite inputs
__device__ inline double2 jinv(double c, double d)
{
    if (fabs(d) <= fabs(c)) {
        const double r = d / c, z = fma(d, r, c);
        return make_double2(1.0 / z, -r / z);
    }
    const double r = c / d, z = fma(-c, r, -d);
    return make_double2(-r / z, 1.0 / z);
}

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// Gm[ω, c12, i] = G(ωs[ω] - Ω/2), Gp = G(ωs[ω] + Ω/2) with Σ read at ωs ∓ Ω/2 through the host
// tables (src/ladder.jl:322-334; src/ladder_kernel.jl:27-57).
__global__ void greens_kernel(int64_t nw, int64_t ni, const double *__restrict__ ws, const double *__restrict__ e_i,
                              const double2 *__restrict__ S, const int32_t *__restrict__ jm, const double *__restrict__ tm,
                              const int32_t *__restrict__ jp, const double *__restrict__ tp, double Omega, double mu,
                              double T, double2 *__restrict__ Gm, double2 *__restrict__ Gp)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nw * ni) return;
    const int64_t w = q % nw, i = q / nw;
    const double2 *col = S + nw * i;
    for (int side = 0; side < 2; side++) {
        const int64_t j = (side == 0 ? jm[w] : jp[w]) - 1;
        const double t = side == 0 ? tm[w] : tp[w];
        const double sr = (1 - t) * col[j].x + t * col[j + 1].x, si = (1 - t) * col[j].y + t * col[j + 1].y;
        const double x = side == 0 ? ws[w] - Omega / 2 : ws[w] + Omega / 2;
        const double2 GR = jinv(x - e_i[i] - sr, -si);
        const double2 GA = make_double2(GR.x, -GR.y);
        const double k = 1 - 2 * occ_fermion_l(x - mu, T);
        const double2 GK = make_double2((GR.x - GA.x) * k, (GR.y - GA.y) * k);
        double2 *G = side == 0 ? Gm : Gp;
        G[w + nw * (0 + 4 * i)] = make_double2((GR.x + GA.x + GK.x) / 2, (GR.y + GA.y + GK.y) / 2);
        G[w + nw * (1 + 4 * i)] = make_double2((GR.x - GA.x + GK.x) / 2, (GR.y - GA.y + GK.y) / 2);
        G[w + nw * (2 + 4 * i)] = make_double2((-GR.x + GA.x + GK.x) / 2, (-GR.y + GA.y + GK.y) / 2);
        G[w + nw * (3 + 4 * i)] = make_double2((-GR.x - GA.x + GK.x) / 2, (-GR.y - GA.y + GK.y) / 2);
    }
}

#define JIX(w, c12, cart, c0, i) ((w) + nw * ((c12) + 4 * ((cart) + 3 * ((c0) + 2 * (int64_t)(i)))))
#define GIX(w, c, i) ((w) + nw * ((c) + 4 * (int64_t)(i)))

// J0^{c1c2,c0} = Gm^{c1 c0} v Gp^{c0 c2}  (src/ladder_kernel.jl:70-73)
__global__ void bare_vertex_kernel(int64_t nw, int64_t ni, const double2 *__restrict__ Gm, const double *__restrict__ v,
                                   const double2 *__restrict__ Gp, double2 *__restrict__ J0)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nw * 24 * ni) return;
    const int64_t w = q % nw;
    int64_t r = q / nw;
    const int c12 = r % 4; r /= 4;
    const int cart = r % 3; r /= 3;
    const int c0 = r % 2;
    const int64_t i = r / 2;
    const int c1 = c12 % 2, c2 = c12 / 2;
    const double2 a = Gm[GIX(w, c1 + 2 * c0, i)], b = Gp[GIX(w, c0 + 2 * c2, i)];
    const double vv = v[cart + 3 * i];
    const double2 av = make_double2(a.x * vv, a.y * vv);
    J0[q] = cmul(av, b);
}

// Y[b, i] = Σ_ω Wω[ω] (J^{22,c0=1} + J^{11,c0=2})(ω)  (src/ladder_kernel.jl:119), one block per (b, i)
__global__ __launch_bounds__(256) void cond_Y_kernel(int64_t nw, int64_t ni, const double2 *__restrict__ J,
                                                     const double *__restrict__ Ww, double2 *__restrict__ Y)
{
    __shared__ double rr[256], ri[256];
    const int b = blockIdx.x % 3;
    const int64_t i = blockIdx.x / 3;
    double sr = 0, si = 0;
    for (int64_t w = threadIdx.x; w < nw; w += 256) {
        const double2 x = J[JIX(w, 3, b, 0, i)], y = J[JIX(w, 0, b, 1, i)];
        sr = fma(Ww[w], x.x + y.x, sr);
        si = fma(Ww[w], x.y + y.y, si);
    }
    rr[threadIdx.x] = sr;
    ri[threadIdx.x] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            rr[threadIdx.x] += rr[threadIdx.x + s];
            ri[threadIdx.x] += ri[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) Y[b + 3 * i] = make_double2(rr[0], ri[0]);
}

struct BtArgs {
    int64_t nm, ni, nfs, ns;
    const double *g2p;        // [nm, nfs, ni]
    const double *wf;         // [nfs]
    const double *Rs;         // [3, 3, nfs] rotations in star order
    const int64_t *star_ptr;  // [ni + 1]
    const int32_t *fsrc;      // [nfs]
    const double *wph;
    const int32_t *iq_kk, *iks_i, *iks_f;
    int64_t nki;
    double h, T, w_acoustic;
    double *Bt;               // [(b, ip), (a, i, t), m]: (b + 3 ip) + 3 ni ((a + 3 (i + ni t)) + 6 ni m)
};

// B̃[m, b, ip, a, i, t] = Σ_{f in star(ip)} R_f[a, b] w_f Σ_ν g2[ν, i, f] tent(ωq/h - m) (n_B(ωq) for
// t = 2), modes below ω_acoustic skipped (src/ladder_kernel.jl:195-264).  One thread per
// (m, ip, i) with its 18 (a, b, t) sums in registers.
__global__ __launch_bounds__(256) void bt_build_kernel(BtArgs a)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t ni = a.ni, ns = a.ns;
    if (q >= ns * ni * ni) return;
    const int64_t m = q % ns, ip = (q / ns) % ni, i = q / (ns * ni);
    double acc[2][9];
    for (int t = 0; t < 2; t++)
        for (int k = 0; k < 9; k++) acc[t][k] = 0.0;
    for (int64_t p = a.star_ptr[ip]; p < a.star_ptr[ip + 1]; p++) {
        const int64_t fs = a.fsrc[p];
        const int64_t qi = a.iq_kk[(a.iks_i[i] - 1) + a.nki * (int64_t)(a.iks_f[fs] - 1)] - 1;
        double s1 = 0, s2 = 0;
        for (int64_t nu = 0; nu < a.nm; nu++) {
            const double om = a.wph[nu + a.nm * qi];
            if (om < a.w_acoustic) continue;
            const double x = om / a.h;
            const double tent = 1.0 - fabs(x - (double)m);
            if (!(tent > 0)) continue;
            const double c = a.g2p[nu + a.nm * (p + a.nfs * i)] * tent;
            s1 += c;
            s2 += c * occ_boson_l(om, a.T);
        }
        if (s1 == 0 && s2 == 0) continue;
        const double w = a.wf[p];
        const double *R = a.Rs + 9 * p;
        for (int k = 0; k < 9; k++) {
            acc[0][k] = fma(R[k] * w, s1, acc[0][k]);
            acc[1][k] = fma(R[k] * w, s2, acc[1][k]);
        }
    }
    // R[k] with k = a_ + 3 b_
    for (int t = 0; t < 2; t++)
        for (int b = 0; b < 3; b++)
            for (int aa = 0; aa < 3; aa++) {
                const int64_t col = aa + 3 * (i + ni * t);
                a.Bt[(b + 3 * ip) + 3 * ni * (col + 6 * ni * m)] = acc[t][aa + 3 * b];
            }
}

// B̃ in BF16 / FP32 with the row pitch of the low-precision product: out[k + kp c] = Bt[k + K c]
// (zero for K <= k < kp).
// B̃ with its columns regrouped by blocks of cb columns, every shift of a block contiguous: the
// column (c0 + cw, m) of B̃ (Bt[:, c + ncol m]) goes to c0 ns + cw + cbb m, where c0 = cb ⌊c / cb⌋
// and cbb = min(cb, ncol - c0), so that one GEMM over a block's cbb ns columns gives W in the
// layout the shift-add reads; in E = double, float or BF16 bits, rows of pitch kp (zero beyond K).
template <typename E>
__global__ void bt_perm_kernel(int64_t K, int64_t kp, int64_t ncol, int64_t ns, int64_t cb, const double *__restrict__ Bt,
                               E *__restrict__ out)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= kp * ncol * ns) return;
    const int64_t k = q % kp, src = q / kp;
    const int64_t c = src % ncol, m = src / ncol;
    const int64_t c0 = c / cb * cb, cbb = cb < ncol - c0 ? cb : ncol - c0;
    const double v = k < K ? Bt[k + K * src] : 0.0;
    E *o = out + kp * (c0 * ns + (c - c0) + cbb * m) + k;
    if constexpr (sizeof(E) == 2) *o = to_bf16(v);
    else *o = (E)v;
}

// The channels K^{ab} with the contour signs (src/ladder_kernel.jl:296-307), then
// LJ = -Gm K Gp as 2x2 contour matrices (src/ladder.jl:294-295).  One thread per (ω, a, c0, i).
__global__ void exch_attach_kernel(int64_t nw, int64_t ni, const double *__restrict__ Cp, const double *__restrict__ Cm,
                                   const double *__restrict__ KY, const double2 *__restrict__ Gm,
                                   const double2 *__restrict__ Gp, double2 *__restrict__ LJ)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= nw * 3 * 2 * ni) return;
    const int64_t w = q % nw;
    int64_t r = q / nw;
    const int a = r % 3; r /= 3;
    const int c0 = r % 2;
    const int64_t i = r / 2;
    const int64_t nrow = 16 * nw;
    // C[(j, c12, c0, ri), (a, i, t)]
    auto C = [&](const double *X, int c12, int ri, int t) -> double {
        return X[(w + nw * (c12 + 4 * (c0 + 2 * ri))) + nrow * (a + 3 * (i + ni * t))];
    };
    auto Ky = [&](int c12, int ri) -> double { return KY[(w + nw * (c12 + 4 * (c0 + 2 * ri))) + nrow * (a + 3 * i)]; };
    double2 K[4];
    for (int c12 = 0; c12 < 4; c12++) {
        const double2 Up = make_double2(C(Cp, c12, 0, 0), C(Cp, c12, 1, 0));
        const double2 Um = make_double2(C(Cm, c12, 0, 0), C(Cm, c12, 1, 0));
        const double2 Vp = make_double2(C(Cp, c12, 0, 1), C(Cp, c12, 1, 1));
        const double2 Vm = make_double2(C(Cm, c12, 0, 1), C(Cm, c12, 1, 1));
        if (c12 == 0 || c12 == 3) {
            const double2 ky = make_double2(Ky(c12, 0), Ky(c12, 1));
            // (c12 == 1 ? -i : +i) * ky
            const double2 iky = c12 == 0 ? make_double2(ky.y, -ky.x) : make_double2(-ky.y, ky.x);
            K[c12] = make_double2(iky.x - (Up.x + Um.x) / 2 - (Vp.x + Vm.x), iky.y - (Up.y + Um.y) / 2 - (Vp.y + Vm.y));
        } else if (c12 == 1) {
            K[c12] = make_double2(Vp.x + Um.x + Vm.x, Vp.y + Um.y + Vm.y);
        } else {
            K[c12] = make_double2(Up.x + Vp.x + Vm.x, Up.y + Vp.y + Vm.y);
        }
    }
    // J^{c1c2} = Σ_ab Gm^{c1 a} K^{ab} Gp^{b c2}, K^{ab} at c12 = a + 2b
    for (int c12 = 0; c12 < 4; c12++) {
        const int c1 = c12 % 2, c2 = c12 / 2;
        const double2 g1 = Gm[GIX(w, c1, i)], g2 = Gm[GIX(w, c1 + 2, i)];
        const double2 p1 = Gp[GIX(w, 0 + 2 * c2, i)], p2 = Gp[GIX(w, 1 + 2 * c2, i)];
        double2 s = cmul(cmul(g1, K[0]), p1);
        s = cadd(s, cmul(cmul(g2, K[1]), p1));
        s = cadd(s, cmul(cmul(g1, K[2]), p2));
        s = cadd(s, cmul(cmul(g2, K[3]), p2));
        LJ[JIX(w, c12, a, c0, i)] = make_double2(-s.x, -s.y);
    }
}

inline unsigned nb(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

}  // namespace epsx

// ------------------------------------------------------------------ GMRES vector kernels
namespace epsx {
namespace {

constexpr int RED_BLOCKS = 512;

// part[b] = Σ over this block's strided elements of conj(x) y (fixed order)
__global__ __launch_bounds__(256) void cdot_partial_kernel(int64_t n, const double2 *__restrict__ x,
                                                           const double2 *__restrict__ y, double2 *__restrict__ part)
{
    __shared__ double rr[256], ri[256];
    double sr = 0, si = 0;
    for (int64_t q = blockIdx.x * 256 + threadIdx.x; q < n; q += (int64_t)gridDim.x * 256) {
        const double2 a = x[q], b = y[q];
        sr = fma(a.x, b.x, fma(a.y, b.y, sr));
        si = fma(a.x, b.y, fma(-a.y, b.x, si));
    }
    rr[threadIdx.x] = sr;
    ri[threadIdx.x] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            rr[threadIdx.x] += rr[threadIdx.x + s];
            ri[threadIdx.x] += ri[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = make_double2(rr[0], ri[0]);
}

__global__ __launch_bounds__(256) void cdot_final_kernel(int nb, const double2 *__restrict__ part, double2 *__restrict__ out)
{
    __shared__ double rr[256], ri[256];
    double sr = 0, si = 0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        sr += part[b].x;
        si += part[b].y;
    }
    rr[threadIdx.x] = sr;
    ri[threadIdx.x] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            rr[threadIdx.x] += rr[threadIdx.x + s];
            ri[threadIdx.x] += ri[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = make_double2(rr[0], ri[0]);
}

// y -= c x, c read from device memory (the Arnoldi coefficient just computed)
__global__ void caxpy_dev_kernel(int64_t n, const double2 *__restrict__ c, const double2 *__restrict__ x,
                                 double2 *__restrict__ y)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double2 a = *c, b = x[q];
    y[q] = make_double2(y[q].x - (a.x * b.x - a.y * b.y), y[q].y - (a.x * b.y + a.y * b.x));
}

__global__ void caxpy_kernel(int64_t n, double ar, double ai, const double2 *__restrict__ x, double2 *__restrict__ y)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double2 b = x[q];
    y[q] = make_double2(y[q].x + (ar * b.x - ai * b.y), y[q].y + (ar * b.y + ai * b.x));
}

__global__ void cscale_kernel(int64_t n, double s, const double2 *__restrict__ x, double2 *__restrict__ y)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < n) y[q] = make_double2(s * x[q].x, s * x[q].y);
}

// y = a + sa * b + sc * c  (elementwise complex)
__global__ void combine3_kernel(int64_t n, const double2 *__restrict__ a, double sb, const double2 *__restrict__ b,
                                double sc, const double2 *__restrict__ c, double2 *__restrict__ y)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    y[q] = make_double2(a[q].x + sb * b[q].x + sc * c[q].x, a[q].y + sb * b[q].y + sc * c[q].y);
}

// ---- instances of the batched solve: one (Ω, T) each; per-slot vectors of n complex at stride
// `vs`; a launch runs the active instances k < nA of `list` (list[k] = slot), with the stacked
// operands of the shift-table product indexed by k ----

// Jm[(r + nrow k), kk] = Re / Im of the vertex of instance k: the stacked A operand (leading dim M)
__global__ void jperm_b_kernel(int64_t nw, int64_t ni, int64_t M, const int32_t *__restrict__ list,
                               const double2 *__restrict__ Jb, int64_t vs, double *__restrict__ Jm)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nrow = 16 * nw, k = blockIdx.y;
    if (q >= nrow * 3 * ni) return;
    const double2 *J = Jb + vs * list[k];
    const int64_t r = q % nrow, kk = q / nrow;
    const int64_t j = r % nw, c12 = (r / nw) % 4, c0 = (r / (4 * nw)) % 2, ri = r / (8 * nw);
    const double2 z = J[JIX(j, c12, kk % 3, c0, kk / 3)];
    Jm[(r + nrow * k) + M * kk] = ri == 0 ? z.x : z.y;
}

// the stacked vertex as the transposed low-precision operand: JmL[kk + kp (r + nrow k)]
template <typename E>
__global__ __launch_bounds__(256) void jperm_lp_b_kernel(int64_t nw, int64_t ni, int64_t kp, const int32_t *__restrict__ list,
                                                         const double2 *__restrict__ Jb, int64_t vs, E *__restrict__ JmL)
{
    __shared__ double tile[64][65];
    const int64_t nrow = 16 * nw, K = 3 * ni, k = blockIdx.z;
    const double2 *J = Jb + vs * list[k];
    const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
    const int lx = threadIdx.x & 63;
    for (int sk = threadIdx.x >> 6; sk < 64; sk += 4) {
        const int64_t r = r0 + lx, kk = k0 + sk;
        double v = 0.0;
        if (r < nrow && kk < K) {
            const int64_t j = r % nw, c12 = (r / nw) % 4, c0 = (r / (4 * nw)) % 2, ri = r / (8 * nw);
            const double2 z = J[JIX(j, c12, kk % 3, c0, kk / 3)];
            v = ri == 0 ? z.x : z.y;
        }
        tile[sk][lx] = v;
    }
    __syncthreads();
    for (int sr = threadIdx.x >> 6; sr < 64; sr += 4) {
        const int64_t r = r0 + sr, kk = k0 + lx;
        if (r < nrow && kk < kp) {
            const double v = tile[lx][sr];
            E *out = JmL + kp * (r + nrow * k);
            if constexpr (sizeof(E) == 2) out[kk] = to_bf16(v);
            else out[kk] = (float)v;
        }
    }
}

// The shift-add over a block of columns [c0, c0 + cb) for every shift at once: W holds W_m[:, c] of
// all stacked instances, W[row + M (c - c0 + cb m)] with row = (j', q) + nrow k, and
//   C±[(j, q), c] = Σ_m wlo[j, m, ±] W_m[(klo - 1, q), c] + whi[j, m, ±] W_m[(klo, q), c]
// is complete after one pass (ascending m, the order of the reference's sum over the shift grid).
// One wave per (64 frequencies j, q, instance k, group of SA_C columns): the taps of (j, m) are read
// once and serve the group's columns; lanes along j, so every W read of a wave is one contiguous run.
constexpr int SA_C = 8;
template <typename TW>
__global__ __launch_bounds__(64) void shift_add_cols_kernel(int64_t nw, int64_t ns, int64_t ncol, int64_t c0, int64_t cb,
                                                            int64_t M, int ngrp, const TW *__restrict__ W,
                                                            const int32_t *__restrict__ klo, const double *__restrict__ wlo,
                                                            const double *__restrict__ whi, double *__restrict__ Cp,
                                                            double *__restrict__ Cm)
{
    const int64_t j = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (j >= nw) return;
    const int64_t q = blockIdx.y;
    const int64_t k = blockIdx.z / ngrp, cg = (int64_t)(blockIdx.z % ngrp) * SA_C;
    const int64_t nrow = 16 * nw;
    const int nc = (int)(cb - cg < SA_C ? cb - cg : SA_C);
    const TW *Wb = W + nrow * k + nw * q + M * cg;
    double sp[SA_C], sm[SA_C];
#pragma unroll
    for (int c = 0; c < SA_C; c++) sp[c] = sm[c] = 0.0;
    for (int64_t m = 0; m < ns; m++) {
        const int64_t t0 = j + nw * m, t1 = j + nw * (m + ns);
        const int64_t k0 = klo[t0] - 1, k1 = klo[t1] - 1;
        const double lo0 = wlo[t0], hi0 = whi[t0], lo1 = wlo[t1], hi1 = whi[t1];
        const TW *Wm = Wb + M * cb * m;
#pragma unroll
        for (int c = 0; c < SA_C; c++) {
            if (c < nc) {
                const TW *col = Wm + M * c;
                sp[c] += lo0 * (double)col[k0] + hi0 * (double)col[k0 + 1];
                sm[c] += lo1 * (double)col[k1] + hi1 * (double)col[k1 + 1];
            }
        }
    }
    const int64_t r = j + nw * q;
#pragma unroll
    for (int c = 0; c < SA_C; c++) {
        if (c < nc) {
            const int64_t o = r + nrow * (c0 + cg + c) + nrow * ncol * k;
            Cp[o] = sp[c];
            Cm[o] = sm[c];
        }
    }
}

// Y of instance k = blockIdx.y: (U+ - U-)/2 from the type-1 columns of its C±
__global__ void exch_Y_b_kernel(int64_t ny, int64_t ncs, const double *__restrict__ Cp, const double *__restrict__ Cm,
                                double *__restrict__ Y)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, k = blockIdx.y;
    if (q < ny) Y[q + ny * k] = (Cp[q + ncs * k] - Cm[q + ncs * k]) / 2;
}

// per slot z = list[k]: Y_z = A_z + sb B_z (n complex per slot, each array at its own slot stride)
__global__ void combine2_b_kernel(int64_t n, const int32_t *__restrict__ list, const double2 *__restrict__ A, int64_t as,
                                  double sb, const double2 *__restrict__ B, int64_t bs, double2 *__restrict__ Y, int64_t ys)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t z = list[blockIdx.y];
    const double2 a = A[as * z + q], b = B[bs * z + q];
    Y[ys * z + q] = make_double2(a.x + sb * b.x, a.y + sb * b.y);
}

// conj(x_z) y_z partial sums per (block, k), then per k into out[z * ostride + oidx]
__global__ __launch_bounds__(256) void cdot_b_partial_kernel(int64_t n, const int32_t *__restrict__ list, const double2 *__restrict__ X,
                                                             int64_t xs, const double2 *__restrict__ Yv, int64_t ys,
                                                             double2 *__restrict__ part)
{
    __shared__ double rr[256], ri[256];
    const int64_t z = list[blockIdx.y];
    const double2 *x = X + xs * z, *y = Yv + ys * z;
    double sr = 0, si = 0;
    for (int64_t q = blockIdx.x * 256 + threadIdx.x; q < n; q += (int64_t)gridDim.x * 256) {
        const double2 a = x[q], b = y[q];
        sr = fma(a.x, b.x, fma(a.y, b.y, sr));
        si = fma(a.x, b.y, fma(-a.y, b.x, si));
    }
    rr[threadIdx.x] = sr;
    ri[threadIdx.x] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            rr[threadIdx.x] += rr[threadIdx.x + s];
            ri[threadIdx.x] += ri[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x + (int64_t)gridDim.x * blockIdx.y] = make_double2(rr[0], ri[0]);
}

__global__ __launch_bounds__(256) void cdot_b_final_kernel(int nb, const int32_t *__restrict__ list,
                                                           const double2 *__restrict__ part, double2 *__restrict__ out,
                                                           int64_t ostride, int64_t oidx)
{
    __shared__ double rr[256], ri[256];
    const int k = blockIdx.x;
    double sr = 0, si = 0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        sr += part[b + (int64_t)nb * k].x;
        si += part[b + (int64_t)nb * k].y;
    }
    rr[threadIdx.x] = sr;
    ri[threadIdx.x] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            rr[threadIdx.x] += rr[threadIdx.x + s];
            ri[threadIdx.x] += ri[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[ostride * list[k] + oidx] = make_double2(rr[0], ri[0]);
}

// y_z -= c_z x_z with c_z = coef[z * cs + ci] (an Arnoldi coefficient just computed)
__global__ void caxpy_b_kernel(int64_t n, const int32_t *__restrict__ list, const double2 *__restrict__ coef, int64_t cs,
                               int64_t ci, const double2 *__restrict__ X, int64_t xs, double2 *__restrict__ Y, int64_t ys)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t z = list[blockIdx.y];
    const double2 a = coef[cs * z + ci], b = X[xs * z + q];
    double2 &y = Y[ys * z + q];
    y = make_double2(y.x - (a.x * b.x - a.y * b.y), y.y - (a.x * b.y + a.y * b.x));
}

// y_z = s_z x_z with s_z = scale[z]
__global__ void cscale_b_kernel(int64_t n, const int32_t *__restrict__ list, const double *__restrict__ scale,
                                const double2 *__restrict__ X, int64_t xs, double2 *__restrict__ Y, int64_t ys)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t z = list[blockIdx.y];
    const double a = scale[z];
    const double2 b = X[xs * z + q];
    Y[ys * z + q] = make_double2(a * b.x, a * b.y);
}

__device__ inline double2 cdiv(double2 a, double2 b)
{
    // a / b with Smith's scaling
    if (fabs(b.y) <= fabs(b.x)) {
        const double r = b.y / b.x, d = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
    }
    const double r = b.x / b.y, d = b.y + b.x * r;
    return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
__device__ inline double cabs2(double2 a) { return hypot(a.x, a.y); }
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline double2 cscl(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// The Givens step of GMRES for instance k (Krylov.jl's MGS Arnoldi with sym_givens, no restart):
// the new Hessenberg column h_z[0..j] and ||q||² in h_z[j+1].x rotate into R's column j, the
// residual norm |z_{j+1}| is appended to norms, and the status says whether the instance is solved
// (norm <= atol) or broke down (h_{j+1,j} <= btol).  inv_h = 1 / h_{j+1,j} scales the next basis
// vector.  State per slot: R [(cap) x (cap)], cs, sn, zv (cap + 1), at stride gst.
struct GivensState {
    double2 *R, *sn, *zv;
    double *cs, *atol, *inv_h, *norms;   // norms [(maxiter + 1) per slot]
    const double2 *h;                    // [cap + 1] per slot
    int32_t *status;                     // per slot: 0 running, 1 solved, 2 breakdown
    int64_t cap, maxiter;
    double btol;
};
__global__ void givens_b_kernel(int nA, const int32_t *__restrict__ list, GivensState g, int64_t j)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nA) return;
    const int64_t z = list[k], cap = g.cap;
    double2 *R = g.R + cap * cap * z, *sn = g.sn + cap * z, *zv = g.zv + (cap + 1) * z;
    double *cs = g.cs + cap * z;
    const double2 *h = g.h + (cap + 1) * z;
    for (int64_t i = 0; i <= j; i++) R[i + cap * j] = h[i];
    const double Hbis = sqrt(h[j + 1].x);
    for (int64_t i = 0; i < j; i++) {
        const double2 a = R[i + cap * j], b = R[i + 1 + cap * j];
        R[i + cap * j] = cadd(cscl(cs[i], a), cmul(sn[i], b));
        R[i + 1 + cap * j] = csub(cmul(cconj(sn[i]), a), cscl(cs[i], b));
    }
    // sym_givens(R[j, j], Hbis)
    const double2 a = R[j + cap * j];
    const double aa = cabs2(a);
    double c;
    double2 sv, rho;
    if (aa == 0) { c = 0; sv = make_double2(1, 0); rho = make_double2(Hbis, 0); }
    else if (Hbis == 0) { c = 1; sv = make_double2(0, 0); rho = a; }
    else {
        const double nrm = hypot(aa, Hbis);
        c = aa / nrm;
        const double2 ph = make_double2(a.x / aa, a.y / aa);
        sv = cscl(Hbis / nrm, ph);
        rho = cscl(nrm, ph);
    }
    cs[j] = c;
    sn[j] = sv;
    R[j + cap * j] = rho;
    zv[j + 1] = cmul(cconj(sv), zv[j]);
    zv[j] = cscl(c, zv[j]);
    const double rn = cabs2(zv[j + 1]);
    g.norms[(g.maxiter + 1) * z + j + 1] = rn;
    const bool solved = rn <= g.atol[z], breakdown = Hbis <= g.btol;
    g.status[z] = solved ? 1 : (breakdown ? 2 : 0);
    g.inv_h[z] = 1.0 / Hbis;
}

// y = R \ z over the instance's niter columns, then x_z += Σ_i y_i V_z,i (one thread per element)
__global__ void gmres_solve_b_kernel(int nA, const int32_t *__restrict__ list, GivensState g,
                                     const int64_t *__restrict__ niter, double2 *__restrict__ y)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nA) return;
    const int64_t z = list[k], cap = g.cap, m = niter[z];
    const double2 *R = g.R + cap * cap * z, *zv = g.zv + (cap + 1) * z;
    double2 *yz = y + cap * z;
    for (int64_t i = m - 1; i >= 0; i--) {
        double2 acc = zv[i];
        for (int64_t q = i + 1; q < m; q++) acc = csub(acc, cmul(R[i + cap * q], yz[q]));
        yz[i] = cdiv(acc, R[i + cap * i]);
    }
}
__global__ void gmres_update_b_kernel(int64_t n, const int32_t *__restrict__ list, const int64_t *__restrict__ niter,
                                      const double2 *__restrict__ y, int64_t cap, const double2 *__restrict__ V, int64_t vstride,
                                      double2 *__restrict__ X, int64_t xs)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t z = list[blockIdx.y], m = niter[z];
    double2 x = X[xs * z + q];
    for (int64_t i = 0; i < m; i++) {
        const double2 a = y[cap * z + i], b = V[vstride * z + n * i + q];
        x = make_double2(x.x + (a.x * b.x - a.y * b.y), x.y + (a.x * b.y + a.y * b.x));
    }
    X[xs * z + q] = x;
}

}  // namespace

static int ldot(epsx_ctx *ctx, int64_t n, const double2 *x, const double2 *y, double2 *d_out)
{
    LadderDev &l = ctx->l;
    hipLaunchKernelGGL(cdot_partial_kernel, dim3(RED_BLOCKS), dim3(256), 0, ctx->stream, n, x, y, dptr<double2>(l.red));
    hipLaunchKernelGGL(cdot_final_kernel, dim3(1), dim3(256), 0, ctx->stream, RED_BLOCKS, dptr<double2>(l.red), d_out);
    EPSX_HIP(ctx, hipGetLastError());
    return EPSX_OK;
}

// The shift-table product W = J B̃ is a plain GEMM (no epilogue, no batching): it runs in rocBLAS
// (deterministic: atomics off; `rocblas_dgemm`, and `rocblas_gemm_ex` from the stored BF16 / FP32
// operands with FP32 accumulation) on the context's stream; EPSX_LADDER_BLAS=0 runs the library's
// own MFMA kernels instead (for comparisons), EPSX_LADDER_LT=1 the BF16 product in hipBLASLt.
static bool use_blas()
{
    const char *e = getenv("EPSX_LADDER_BLAS");
    return !(e && e[0] == '0');
}

static int blas_handle(epsx_ctx *ctx, rocblas_handle &h)
{
    if (!ctx->blas) {
        rocblas_handle nh = nullptr;
        if (rocblas_create_handle(&nh) != rocblas_status_success)
            return set_error(ctx, EPSX_ERR_DEVICE, "rocblas_create_handle failed");
        if (rocblas_set_atomics_mode(nh, rocblas_atomics_not_allowed) != rocblas_status_success) {
            (void)rocblas_destroy_handle(nh);
            return set_error(ctx, EPSX_ERR_DEVICE, "rocblas_set_atomics_mode failed");
        }
        ctx->blas = nh;
    }
    h = static_cast<rocblas_handle>(ctx->blas);
    if (rocblas_set_stream(h, ctx->stream) != rocblas_status_success)
        return set_error(ctx, EPSX_ERR_DEVICE, "rocblas_set_stream failed");
    return EPSX_OK;
}

static int blas_check(epsx_ctx *ctx, rocblas_status st, const char *what)
{
    if (st != rocblas_status_success)
        return set_error(ctx, EPSX_ERR_DEVICE, "%s failed: %s", what, rocblas_status_to_string(st));
    return EPSX_OK;
}

// B̃ for temperature T (once per T and ω_acoustic)
static int ensure_Bt(epsx_ctx *ctx, double T, double w_ac)
{
    Scgd0Dev &s = ctx->s;
    LadderDev &l = ctx->l;
    if (l.bt_T == T && l.bt_wac == w_ac) return EPSX_OK;
    const int64_t ni = s.ni;
    EPSX_TRY(dev_alloc(ctx, l.Bt, 8 * (size_t)(3 * ni) * (6 * ni) * l.ns));
    ClassTimer tm(ctx, 2);
    BtArgs a;
    a.nm = s.nm; a.ni = ni; a.nfs = s.nf; a.ns = l.ns;
    a.g2p = dptr<double>(s.g2p); a.wf = dptr<double>(s.wf_s); a.Rs = dptr<double>(l.Rs);
    a.star_ptr = dptr<int64_t>(s.star_ptr); a.fsrc = dptr<int32_t>(s.fsrc_s);
    a.wph = dptr<double>(s.wph); a.iq_kk = dptr<int32_t>(s.iq_kk); a.iks_i = dptr<int32_t>(s.iks_i);
    a.iks_f = dptr<int32_t>(s.iks_f); a.nki = s.nki;
    a.h = l.h; a.T = T; a.w_acoustic = w_ac; a.Bt = dptr<double>(l.Bt);
    hipLaunchKernelGGL(bt_build_kernel, dim3(nb(l.ns * ni * ni)), dim3(256), 0, ctx->stream, a);
    EPSX_HIP(ctx, hipGetLastError());
    l.bt_T = T;
    l.bt_wac = w_ac;
    l.btl_prec = -1;
    return tm.stop(1);
}

// Work buffers of the batched application, sized for nb instances.
struct BatchWork {
    Tmp Jm, JmL, W, Cp, Cm, Y, KY;
    int64_t nb = 0, w_bytes = 0, cb = 0;   // instances, W bytes, columns per block of the product
};

static int batch_work(epsx_ctx *ctx, BatchWork &bw, int64_t nb, int precision)
{
    Scgd0Dev &s = ctx->s;
    const int64_t nw = s.nw, ni = s.ni, nrow = 16 * nw, K = 3 * ni, kp = (K + 7) / 8 * 8;
    bw.nb = nb;
    EPSX_TRY(tmp_alloc(ctx, bw.Cp.b, 8 * (size_t)nrow * 6 * ni * nb));
    EPSX_TRY(tmp_alloc(ctx, bw.Cm.b, 8 * (size_t)nrow * 6 * ni * nb));
    EPSX_TRY(tmp_alloc(ctx, bw.Y.b, 8 * (size_t)nrow * 3 * ni * nb));
    EPSX_TRY(tmp_alloc(ctx, bw.KY.b, 8 * (size_t)nrow * 3 * ni * nb));
    if (precision == EPSX_PREC_FP64) EPSX_TRY(tmp_alloc(ctx, bw.Jm.b, 8 * (size_t)nrow * K * nb));
    else EPSX_TRY(tmp_alloc(ctx, bw.JmL.b, (precision == EPSX_PREC_BF16 ? 2 : 4) * (size_t)kp * nrow * nb));
    // W for a block of columns of all instances and every shift: at most 16 GiB (at least one column)
    const size_t per_col = (precision == EPSX_PREC_FP64 ? 8 : 4) * (size_t)nrow * nb * ctx->l.ns;
    const size_t ncol = 6 * (size_t)ni, cap = size_t(16) << 30;
    const size_t cols = per_col >= cap ? 1 : (cap / per_col < ncol ? cap / per_col : ncol);
    bw.cb = (int64_t)(cols >= ncol ? ncol : (cols >= (size_t)SA_C ? cols / SA_C * SA_C : cols));
    bw.w_bytes = (int64_t)(per_col * (size_t)bw.cb);
    EPSX_TRY(tmp_alloc(ctx, bw.W.b, (size_t)bw.w_bytes));
    return EPSX_OK;
}

// LJ_z = L(V_z) for the active instances k < nA (slot list[k]) at the temperature whose B̃ is loaded:
// the vertices stacked into one shift-table product W = [J_1; ...; J_nA] B̃ (B̃ read once for all
// instances), then per instance the shift-add, the exchange and the propagators Gm_z, Gp_z.
static int ladder_apply_batch(epsx_ctx *ctx, BatchWork &bw, int nA, const std::vector<int32_t> &list,
                              const int32_t *d_list, const double2 *V, int64_t vs, double2 *LJ, int64_t ljs,
                              const double2 *Gm, const double2 *Gp, int precision)
{
    Scgd0Dev &s = ctx->s;
    LadderDev &l = ctx->l;
    const int64_t nw = s.nw, ni = s.ni, nrow = 16 * nw, ncol = 6 * ni, K = 3 * ni, M = nrow * nA;
    const hipStream_t st = ctx->stream;
    ClassTimer t_apply(ctx, 1);
    const bool bf = precision == EPSX_PREC_BF16, native = precision != EPSX_PREC_FP64;
    // the regrouped B̃'s column pitch: kp (low precision, a multiple of 8) or K rounded up to even
    // (FP64: the 16-byte operand loads of the library's GEMM)
    const int64_t kp = (K + 7) / 8 * 8, esz = bf ? 2 : 4, Kd = K + (K & 1);
    // the block width of this call's buffers (fixed per BatchWork, so one regrouped B̃ serves it)
    const int64_t cb = bw.cb;
    if (l.btl_prec != precision || l.btl_cb != cb) {
        const int64_t ncols = ncol * l.ns, pitch = native ? kp : Kd, eb = native ? esz : 8;
        EPSX_TRY(dev_alloc(ctx, l.Btl, eb * (size_t)pitch * ncols));
        const dim3 grid((unsigned)nb(pitch * ncols));
        if (!native)
            hipLaunchKernelGGL(bt_perm_kernel<double>, grid, dim3(256), 0, st, K, Kd, ncol, l.ns, cb, dptr<double>(l.Bt),
                               dptr<double>(l.Btl));
        else if (bf)
            hipLaunchKernelGGL(bt_perm_kernel<uint16_t>, grid, dim3(256), 0, st, K, kp, ncol, l.ns, cb, dptr<double>(l.Bt),
                               dptr<uint16_t>(l.Btl));
        else
            hipLaunchKernelGGL(bt_perm_kernel<float>, grid, dim3(256), 0, st, K, kp, ncol, l.ns, cb, dptr<double>(l.Bt),
                               dptr<float>(l.Btl));
        EPSX_HIP(ctx, hipGetLastError());
        l.btl_prec = precision;
        l.btl_cb = cb;
    }
    if (native) {
        const dim3 grid((unsigned)((nrow + 63) / 64), (unsigned)((kp + 63) / 64), (unsigned)nA);
        if (bf)
            hipLaunchKernelGGL(jperm_lp_b_kernel<uint16_t>, grid, dim3(256), 0, st, nw, ni, kp, d_list, V, vs,
                               bw.JmL.ptr<uint16_t>());
        else
            hipLaunchKernelGGL(jperm_lp_b_kernel<float>, grid, dim3(256), 0, st, nw, ni, kp, d_list, V, vs,
                               bw.JmL.ptr<float>());
    } else {
        hipLaunchKernelGGL(jperm_b_kernel, dim3(nb(nrow * K), (unsigned)nA), dim3(256), 0, st, nw, ni, M, d_list, V, vs,
                           bw.Jm.ptr<double>());
    }
    EPSX_HIP(ctx, hipGetLastError());
    // W for a block of columns and every shift (one product over the block's cbb ns regrouped columns,
    // J read once per column tile), then that block's shift-add, which completes its columns of C±
    const bool blas = use_blas();
    for (int64_t c0 = 0; c0 < ncol; c0 += cb) {
        const int64_t cbb = c0 + cb <= ncol ? cb : ncol - c0, nn = cbb * l.ns;
        ClassTimer t_gemm(ctx, 0);
        if (native) {
            const char *Bl = static_cast<const char *>(l.Btl.p) + esz * kp * c0 * l.ns;
            bool lt = false;
            // hipBLASLt only when EPSX_LADDER_LT=1 asks for it: the batched solve's product changes
            // its row count whenever instances converge, and each new size would time hipBLASLt's
            // candidates again inside the solve (on the t2g 16³ forward-coupling C4 run 12.5 s
            // against 8.2 s with rocBLAS's gemm_ex, which needs no timing)
            const char *lte = getenv("EPSX_LADDER_LT");
            if (blas && bf && lte && lte[0] == '1')
                EPSX_TRY(blaslt_prepare(ctx, EPSX_PREC_BF16, M, nn, kp, bw.JmL.b.p, kp, 0, Bl, kp, 0, bw.W.b.p, M, 0, 1,
                                        &lt));
            if (lt) {
                EPSX_TRY(blaslt_gemm(ctx, EPSX_PREC_BF16, M, nn, kp, bw.JmL.b.p, kp, 0, Bl, kp, 0, bw.W.b.p, M, 0, 1));
            } else if (blas) {
                rocblas_handle h;
                EPSX_TRY(blas_handle(ctx, h));
                const float one = 1.0f, zero = 0.0f;
                const rocblas_datatype dt = bf ? rocblas_datatype_bf16_r : rocblas_datatype_f32_r;
                EPSX_TRY(blas_check(ctx, rocblas_gemm_ex(h, rocblas_operation_transpose, rocblas_operation_none, (int)M,
                                                         (int)nn, (int)kp, &one, bw.JmL.b.p, dt, (int)kp, Bl, dt, (int)kp,
                                                         &zero, bw.W.b.p, rocblas_datatype_f32_r, (int)M, bw.W.b.p,
                                                         rocblas_datatype_f32_r, (int)M, rocblas_datatype_f32_r,
                                                         rocblas_gemm_algo_standard, 0, 0),
                                    "rocblas_gemm_ex (ladder product)"));
            } else {
                GemmBatch bt;
                bt.Cf = bw.W.ptr<float>();
                EPSX_TRY(lpgemm_tn(ctx, precision, M, nn, kp, bw.JmL.b.p, kp, Bl, kp, nullptr, M, false, bt));
            }
        } else if (blas) {
            rocblas_handle h;
            EPSX_TRY(blas_handle(ctx, h));
            const double one = 1.0, zero = 0.0;
            EPSX_TRY(blas_check(ctx, rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (int)M, (int)nn, (int)K,
                                                   &one, bw.Jm.ptr<double>(), (int)M, dptr<double>(l.Btl) + Kd * c0 * l.ns,
                                                   (int)Kd, &zero, bw.W.ptr<double>(), (int)M),
                                "rocblas_dgemm (ladder product)"));
        } else {
            GemmBatch bt;
            EPSX_TRY(dgemm(ctx, M, nn, K, bw.Jm.ptr<double>(), M, dptr<double>(l.Btl) + Kd * c0 * l.ns, Kd,
                           bw.W.ptr<double>(), M, false, bt));
        }
        EPSX_TRY(t_gemm.stop(1));
        const int ngrp = (int)((cbb + SA_C - 1) / SA_C);
        const dim3 grid((unsigned)((nw + 63) / 64), 16u, (unsigned)(nA * ngrp));
        if (native)
            hipLaunchKernelGGL(shift_add_cols_kernel<float>, grid, dim3(64), 0, st, nw, l.ns, ncol, c0, cbb, M, ngrp,
                               bw.W.ptr<float>(), dptr<int32_t>(l.klo), dptr<double>(l.wlo), dptr<double>(l.whi),
                               bw.Cp.ptr<double>(), bw.Cm.ptr<double>());
        else
            hipLaunchKernelGGL(shift_add_cols_kernel<double>, grid, dim3(64), 0, st, nw, l.ns, ncol, c0, cbb, M, ngrp,
                               bw.W.ptr<double>(), dptr<int32_t>(l.klo), dptr<double>(l.wlo), dptr<double>(l.whi),
                               bw.Cp.ptr<double>(), bw.Cm.ptr<double>());
        EPSX_HIP(ctx, hipGetLastError());
    }
    // KY = KK (U+ - U-)/2 for all instances in one product, then per instance the channels and
    // the attached propagators
    const int64_t ny = nrow * 3 * ni, ncs = nrow * ncol;
    hipLaunchKernelGGL(exch_Y_b_kernel, dim3(nb(ny), (unsigned)nA), dim3(256), 0, st, ny, ncs, bw.Cp.ptr<double>(),
                       bw.Cm.ptr<double>(), bw.Y.ptr<double>());
    EPSX_HIP(ctx, hipGetLastError());
    EPSX_TRY(dgemm_nn(ctx, nw, ny / nw * nA, nw, dptr<double>(s.KK), nw, bw.Y.ptr<double>(), nw, bw.KY.ptr<double>(), nw,
                      false));
    const int64_t ng = nw * 4 * ni;
    for (int k = 0; k < nA; k++) {
        const int64_t z = list[k];
        hipLaunchKernelGGL(exch_attach_kernel, dim3(nb(nw * 6 * ni)), dim3(256), 0, st, nw, ni, bw.Cp.ptr<double>() + ncs * k,
                           bw.Cm.ptr<double>() + ncs * k, bw.KY.ptr<double>() + ny * k, Gm + ng * z, Gp + ng * z,
                           LJ + ljs * z);
    }
    EPSX_HIP(ctx, hipGetLastError());
    ctx->n_class[3] += nA;
    return t_apply.stop(1);
}

// LJ = L(J) for one vertex at the temperature whose B̃ is loaded and the propagators in l.Gm/l.Gp
// (apply_ladder_kernel!, src/ladder.jl:245-296): the batched application with one instance.
static int ladder_apply_dev(epsx_ctx *ctx, BatchWork &bw, const double2 *J, double2 *LJ, int precision)
{
    static const std::vector<int32_t> one{0};
    return ladder_apply_batch(ctx, bw, 1, one, dptr<int32_t>(ctx->l.zero_idx), J, 0, LJ, 0, dptr<double2>(ctx->l.Gm),
                              dptr<double2>(ctx->l.Gp), precision);
}

static int ladder_work(epsx_ctx *ctx)
{
    Scgd0Dev &s = ctx->s;
    LadderDev &l = ctx->l;
    const int64_t nw = s.nw, ni = s.ni;
    EPSX_TRY(dev_alloc(ctx, l.Gm, 16 * (size_t)nw * 4 * ni));
    EPSX_TRY(dev_alloc(ctx, l.Gp, 16 * (size_t)nw * 4 * ni));
    const int32_t zero = 0;
    EPSX_TRY(upload(ctx, l.zero_idx, &zero, 4));
    EPSX_TRY(dev_alloc(ctx, l.red, 16 * (size_t)RED_BLOCKS));
    EPSX_TRY(dev_alloc(ctx, l.Yc, 16 * (size_t)3 * ni));
    return EPSX_OK;
}

// σ_ab = Re[F/(2πi) Σ_i v_a w_i Y[b, i]] (src/ladder_kernel.jl:116-122)
static void sigma_from_Y(const std::vector<double2> &Y, const std::vector<double> &v, const std::vector<double> &w,
                         int64_t ni, double Omega, double T, double *sigma)
{
    using cd = std::complex<double>;
    cd F;
    if (Omega == 0) F = cd(0, 1) / (4 * T);
    else {
        const double nB = T > 1.4901161193847656e-08 ? 1.0 / std::expm1(Omega / T) : (Omega > 0 ? 0.0 : -1.0);
        F = -1.0 / (cd(0, 2) * Omega * (1 + 2 * nB));
    }
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            cd S = 0;
            for (int64_t i = 0; i < ni; i++) S += v[a + 3 * i] * w[i] * cd(Y[b + 3 * i].x, Y[b + 3 * i].y);
            sigma[a + 3 * b] = (F / (2 * 3.141592653589793 * cd(0, 1)) * S).real();
        }
}

static int sigma_of(epsx_ctx *ctx, const double2 *J, double Omega, double T, double *sigma)
{
    Scgd0Dev &s = ctx->s;
    LadderDev &l = ctx->l;
    hipLaunchKernelGGL(cond_Y_kernel, dim3((unsigned)(3 * s.ni)), dim3(256), 0, ctx->stream, s.nw, s.ni, J,
                       dptr<double>(l.Ww), dptr<double2>(l.Yc));
    EPSX_HIP(ctx, hipGetLastError());
    std::vector<double2> Y(3 * s.ni);
    EPSX_HIP(ctx, hipMemcpyAsync(Y.data(), l.Yc.p, 16 * 3 * s.ni, hipMemcpyDeviceToHost, ctx->stream));
    EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    sigma_from_Y(Y, l.v_host, s.w_i_host, s.ni, Omega, T, sigma);
    return EPSX_OK;
}

}  // namespace epsx

using namespace epsx;

extern "C" int epsx_ladder_set_problem(epsx_ctx *ctx, const epsx_ladder_problem *P)
{
    if (!ctx) return EPSX_ERR_ARG;
    if (!P || P->struct_size != (int64_t)sizeof(epsx_ladder_problem))
        return set_error(ctx, EPSX_ERR_ARG, "epsx_ladder_problem: struct size mismatch (ABI %d)", EPSX_ABI_VERSION);
    Scgd0Dev &s = ctx->s;
    LadderDev &l = ctx->l;
    if (!s.ready) return set_error(ctx, EPSX_ERR_STATE, "epsx_ladder_set_problem before epsx_scgd0_set_problem");
    if (!P->R || !P->v) return set_error(ctx, EPSX_ERR_ARG, "ladder problem: null R or v");
    EPSX_HIP(ctx, hipSetDevice(ctx->device));
    l.ready = false;
    const int64_t nf = s.nf, ni = s.ni, nw = s.nw;
    // rotations in star order
    std::vector<int32_t> fsrc(nf);
    EPSX_HIP(ctx, hipMemcpy(fsrc.data(), s.fsrc_s.p, 4 * nf, hipMemcpyDeviceToHost));
    std::vector<double> Rs(9 * (size_t)nf);
    for (int64_t p = 0; p < nf; p++)
        for (int k = 0; k < 9; k++) Rs[9 * p + k] = P->R[9 * (int64_t)fsrc[p] + k];
    EPSX_TRY(upload(ctx, l.Rs, Rs.data(), 8 * Rs.size()));
    l.v_host.assign(P->v, P->v + 3 * ni);
    EPSX_TRY(upload(ctx, l.v, P->v, 8 * 3 * ni));
    std::vector<double> Ww(nw);
    trapezoid_weights(nw, s.ws_host.data(), Ww.data());
    EPSX_TRY(upload(ctx, l.Ww, Ww.data(), 8 * nw));
    // shift tables on the grid's step (src/ladder.jl:151-152)
    double wmax = 0;
    std::vector<double> wph(s.nm * s.nq);
    EPSX_HIP(ctx, hipMemcpy(wph.data(), s.wph.p, 8 * wph.size(), hipMemcpyDeviceToHost));
    for (double x : wph) wmax = x > wmax ? x : wmax;
    l.h = s.dw_unit;
    l.ns = (int64_t)std::ceil(wmax / l.h) + 2;
    std::vector<int64_t> offs(nw);
    for (int64_t j = 0; j < nw; j++) offs[j] = (int64_t)std::llround((s.ws_host[j] - s.ws_host[0]) / s.dw_unit);
    {
        // the grid's integer offsets, recovered from the relative offsets kept on the device
        std::vector<int32_t> off(nw);
        EPSX_HIP(ctx, hipMemcpy(off.data(), s.off.p, 4 * nw, hipMemcpyDeviceToHost));
        for (int64_t j = 0; j < nw; j++) offs[j] = off[j];
    }
    std::vector<int32_t> klo(nw * l.ns * 2);
    std::vector<double> wlo(nw * l.ns * 2), whi(nw * l.ns * 2);
    EPSX_TRY(epsx_vertex_shift_tables(ctx, nw, offs.data(), l.ns, klo.data(), wlo.data(), whi.data()));
    EPSX_TRY(upload(ctx, l.klo, klo.data(), 4 * klo.size()));
    EPSX_TRY(upload(ctx, l.wlo, wlo.data(), 8 * wlo.size()));
    EPSX_TRY(upload(ctx, l.whi, whi.data(), 8 * whi.size()));
    EPSX_TRY(ladder_work(ctx));
    l.bt_T = -1;
    l.ready = true;
    EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EPSX_OK;
}

extern "C" int epsx_apply_ladder_kernel(epsx_ctx *ctx, const double *J, const double *Gm, const double *Gp, double T,
                                        double w_acoustic, int32_t precision, double *LJ)
{
    if (!ctx) return EPSX_ERR_ARG;
    LadderDev &l = ctx->l;
    Scgd0Dev &s = ctx->s;
    if (!l.ready) return set_error(ctx, EPSX_ERR_STATE, "epsx_apply_ladder_kernel before epsx_ladder_set_problem");
    if (!J || !Gm || !Gp || !LJ) return set_error(ctx, EPSX_ERR_ARG, "apply_ladder_kernel!: null array");
    if (!(w_acoustic > 0)) return set_error(ctx, EPSX_ERR_ARG, "ω_acoustic must be positive, got %g", w_acoustic);
    if (precision != EPSX_PREC_FP64 && precision != EPSX_PREC_FP32 && precision != EPSX_PREC_BF16)
        return set_error(ctx, EPSX_ERR_ARG, "unknown ladder precision %d", precision);
    EPSX_HIP(ctx, hipSetDevice(ctx->device));
    for (int k = 0; k < 8; k++) { ctx->t_class[k] = 0; ctx->n_class[k] = 0; }
    const int64_t n = s.nw * 24 * s.ni, ng = s.nw * 4 * s.ni;
    Tmp dJ, dLJ;
    EPSX_TRY(tmp_upload(ctx, dJ, J, 16 * n));
    EPSX_TRY(tmp_alloc(ctx, dLJ.b, 16 * n));
    EPSX_HIP(ctx, hipMemcpyAsync(l.Gm.p, Gm, 16 * ng, hipMemcpyHostToDevice, ctx->stream));
    EPSX_HIP(ctx, hipMemcpyAsync(l.Gp.p, Gp, 16 * ng, hipMemcpyHostToDevice, ctx->stream));
    EPSX_TRY(ensure_Bt(ctx, T, w_acoustic));
    BatchWork bw;
    EPSX_TRY(batch_work(ctx, bw, 1, precision));
    EPSX_TRY(ladder_apply_dev(ctx, bw, dJ.ptr<double2>(), dLJ.ptr<double2>(), precision));
    EPSX_HIP(ctx, hipMemcpyAsync(LJ, dLJ.b.p, 16 * n, hipMemcpyDeviceToHost, ctx->stream));
    EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EPSX_OK;
}

namespace {

// Krylov.jl's sym_givens for complex a and real b >= 0 (c real, s complex)
void sym_givens(std::complex<double> a, double b, double &c, std::complex<double> &s, std::complex<double> &rho)
{
    const double aa = std::abs(a);
    if (aa == 0) { c = 0; s = 1; rho = b; return; }
    if (b == 0) { c = 1; s = 0; rho = a; return; }
    const double nrm = std::hypot(aa, b);
    c = aa / nrm;
    s = (a / aa) * (b / nrm);
    rho = (a / aa) * nrm;
}

}  // namespace

// The GMRES solves of the probe frequencies of each temperature as instances of one batched solve:
// every instance starts from J = 0 (no warm start from the previous Ω, which would chain the
// instances), the Arnoldi steps of all active instances run together (one stacked shift-table
// product per step, batched dots and updates, the Givens rotations on the device), and an instance
// drops out of the launches once it is solved or breaks down.  The host reads one status word per
// instance per step.  Instances per batch: as many as fit a 96 GiB working set.
static int ladder_run_batched(epsx_ctx *ctx, const epsx_ladder_params *par, epsx_ladder_outputs *out)
{
    LadderDev &l = ctx->l;
    Scgd0Dev &s = ctx->s;
    const int64_t nw = s.nw, ni = s.ni, n = nw * 24 * ni, nT = par->nT, nO = par->nOmega, it_max = par->maxiter;
    const int64_t ng = nw * 4 * ni, cap = it_max + 1, nrow = 16 * nw;
    const hipStream_t st = ctx->stream;
    // memory per instance: J0, x, q, LJ, the basis, the propagators, C±, Y, KY and the stacked
    // operand of the product
    const size_t per_inst = 16 * (size_t)n * (4 + cap) + 32 * (size_t)ng + 8 * (size_t)nrow * ni * (12 + 3 + 3 + 3);
    int64_t nbat = (int64_t)((size_t(96) << 30) / per_inst);
    nbat = nbat < 1 ? 1 : (nbat > nO ? nO : nbat);
    Tmp J0, X, Q, LJ, V, Gm, Gp, d_list, h, R, sn, zv, cs, atol, invh, norms, status, niter_d, y, part;
    EPSX_TRY(tmp_alloc(ctx, J0.b, 16 * (size_t)n * nbat));
    EPSX_TRY(tmp_alloc(ctx, X.b, 16 * (size_t)n * nbat));
    EPSX_TRY(tmp_alloc(ctx, Q.b, 16 * (size_t)n * nbat));
    EPSX_TRY(tmp_alloc(ctx, LJ.b, 16 * (size_t)n * nbat));
    EPSX_TRY(tmp_alloc(ctx, V.b, 16 * (size_t)n * cap * nbat));
    EPSX_TRY(tmp_alloc(ctx, Gm.b, 16 * (size_t)ng * nbat));
    EPSX_TRY(tmp_alloc(ctx, Gp.b, 16 * (size_t)ng * nbat));
    EPSX_TRY(tmp_alloc(ctx, d_list.b, 4 * (size_t)nbat));
    EPSX_TRY(tmp_alloc(ctx, h.b, 16 * (size_t)(cap + 1) * nbat));
    EPSX_TRY(tmp_alloc(ctx, R.b, 16 * (size_t)cap * cap * nbat));
    EPSX_TRY(tmp_alloc(ctx, sn.b, 16 * (size_t)cap * nbat));
    EPSX_TRY(tmp_alloc(ctx, zv.b, 16 * (size_t)(cap + 1) * nbat));
    EPSX_TRY(tmp_alloc(ctx, cs.b, 8 * (size_t)cap * nbat));
    EPSX_TRY(tmp_alloc(ctx, atol.b, 8 * (size_t)nbat));
    EPSX_TRY(tmp_alloc(ctx, invh.b, 8 * (size_t)nbat));
    EPSX_TRY(tmp_alloc(ctx, norms.b, 8 * (size_t)cap * nbat));
    EPSX_TRY(tmp_alloc(ctx, status.b, 4 * (size_t)nbat));
    EPSX_TRY(tmp_alloc(ctx, niter_d.b, 8 * (size_t)nbat));
    EPSX_TRY(tmp_alloc(ctx, y.b, 16 * (size_t)cap * nbat));
    EPSX_TRY(tmp_alloc(ctx, part.b, 16 * (size_t)RED_BLOCKS * nbat));
    BatchWork bw;
    EPSX_TRY(batch_work(ctx, bw, nbat, par->precision));
    double2 *dJ0 = J0.ptr<double2>(), *dX = X.ptr<double2>(), *dQ = Q.ptr<double2>(), *dLJ = LJ.ptr<double2>();
    double2 *dV = V.ptr<double2>(), *dh = h.ptr<double2>();
    const int32_t *dl = d_list.ptr<int32_t>();
    const int64_t vst = n * cap;   // basis stride per slot
    GivensState gs;
    gs.R = R.ptr<double2>(); gs.sn = sn.ptr<double2>(); gs.zv = zv.ptr<double2>(); gs.cs = cs.ptr<double>();
    gs.atol = atol.ptr<double>(); gs.inv_h = invh.ptr<double>(); gs.norms = norms.ptr<double>(); gs.h = dh;
    gs.status = status.ptr<int32_t>(); gs.cap = cap; gs.maxiter = it_max;
    gs.btol = std::pow(2.220446049250313e-16, 0.75);
    std::vector<int32_t> jv(nw);
    std::vector<double> tv(nw), xs(nw);
    auto dots = [&](int nA, const double2 *Xv, int64_t xstr, const double2 *Yv, int64_t ystr, int64_t oidx) -> int {
        hipLaunchKernelGGL(cdot_b_partial_kernel, dim3(RED_BLOCKS, (unsigned)nA), dim3(256), 0, st, n, dl, Xv, xstr, Yv,
                           ystr, part.ptr<double2>());
        hipLaunchKernelGGL(cdot_b_final_kernel, dim3((unsigned)nA), dim3(256), 0, st, RED_BLOCKS, dl, part.ptr<double2>(),
                           dh, cap + 1, oidx);
        EPSX_HIP(ctx, hipGetLastError());
        return EPSX_OK;
    };
    for (int64_t iT = 0; iT < nT; iT++) {
        const double T = par->Tlist[iT], mu = par->mu[iT];
        if (!(T >= 0)) return set_error(ctx, EPSX_ERR_ARG, "Temperature must be positive");
        EPSX_TRY(ensure_Bt(ctx, T, par->w_acoustic));
        EPSX_HIP(ctx, hipMemcpyAsync(l.S.p, par->Sigma + 2 * nw * ni * iT, 16 * nw * ni, hipMemcpyHostToDevice, st));
        for (int64_t o0 = 0; o0 < nO; o0 += nbat) {
            const int nb_ = (int)(o0 + nbat <= nO ? nbat : nO - o0);
            // per instance: the propagators, the bare vertex and σ_bubble
            for (int z = 0; z < nb_; z++) {
                const double Om = par->Omegas[o0 + z];
                for (int64_t w = 0; w < nw; w++) xs[w] = s.ws_host[w] - Om / 2;
                linear_interpolation_table(nw, s.ws_host.data(), nw, xs.data(), jv.data(), tv.data());
                EPSX_HIP(ctx, hipMemcpyAsync(l.jm.p, jv.data(), 4 * nw, hipMemcpyHostToDevice, st));
                EPSX_HIP(ctx, hipMemcpyAsync(l.tm.p, tv.data(), 8 * nw, hipMemcpyHostToDevice, st));
                for (int64_t w = 0; w < nw; w++) xs[w] = s.ws_host[w] + Om / 2;
                linear_interpolation_table(nw, s.ws_host.data(), nw, xs.data(), jv.data(), tv.data());
                EPSX_HIP(ctx, hipMemcpyAsync(l.jp.p, jv.data(), 4 * nw, hipMemcpyHostToDevice, st));
                EPSX_HIP(ctx, hipMemcpyAsync(l.tp.p, tv.data(), 8 * nw, hipMemcpyHostToDevice, st));
                hipLaunchKernelGGL(greens_kernel, dim3(nb(nw * ni)), dim3(256), 0, st, nw, ni, dptr<double>(s.ws),
                                   dptr<double>(s.e_i), dptr<double2>(l.S), dptr<int32_t>(l.jm), dptr<double>(l.tm),
                                   dptr<int32_t>(l.jp), dptr<double>(l.tp), Om, mu, T, Gm.ptr<double2>() + ng * z,
                                   Gp.ptr<double2>() + ng * z);
                hipLaunchKernelGGL(bare_vertex_kernel, dim3(nb(n)), dim3(256), 0, st, nw, ni, Gm.ptr<double2>() + ng * z,
                                   dptr<double>(l.v), Gp.ptr<double2>() + ng * z, dJ0 + n * z);
                EPSX_HIP(ctx, hipGetLastError());
                EPSX_TRY(sigma_of(ctx, dJ0 + n * z, Om, T, out->sigma_bubble + 9 * (o0 + z + nO * iT)));
                // the host tables are read by the next instance's copies only after this one's
                // kernels (one stream)
            }
            // GMRES from x = 0: r = J0, β = ||J0||, atol = rtol β
            std::vector<int32_t> list(nb_), st_h(nb_);
            for (int z = 0; z < nb_; z++) list[z] = z;
            EPSX_HIP(ctx, hipMemcpyAsync(d_list.b.p, list.data(), 4 * nb_, hipMemcpyHostToDevice, st));
            EPSX_TRY(dots(nb_, dJ0, n, dJ0, n, 0));
            std::vector<double2> hb(nb_ * (cap + 1));
            EPSX_HIP(ctx, hipMemcpyAsync(hb.data(), dh, 16 * (size_t)(cap + 1) * nb_, hipMemcpyDeviceToHost, st));
            EPSX_HIP(ctx, hipStreamSynchronize(st));
            std::vector<double> beta(nb_), at(nb_), normsh(cap * nb_, NAN), inv(nb_);
            std::vector<double2> z0((cap + 1) * nb_, make_double2(0, 0));
            std::vector<int64_t> niter(nb_, 0);
            std::vector<int32_t> conv(nb_, 0);
            std::vector<int32_t> active;
            for (int z = 0; z < nb_; z++) {
                beta[z] = std::sqrt(hb[(cap + 1) * z].x);
                at[z] = par->rtol * beta[z];
                normsh[cap * z] = beta[z];
                z0[(cap + 1) * z] = make_double2(beta[z], 0);
                inv[z] = beta[z] > 0 ? 1.0 / beta[z] : 0.0;
                conv[z] = beta[z] <= at[z];
                if (beta[z] > 0 && !conv[z]) active.push_back(z);
            }
            EPSX_HIP(ctx, hipMemsetAsync(X.b.p, 0, 16 * (size_t)n * nb_, st));
            EPSX_HIP(ctx, hipMemcpyAsync(atol.b.p, at.data(), 8 * nb_, hipMemcpyHostToDevice, st));
            EPSX_HIP(ctx, hipMemcpyAsync(invh.b.p, inv.data(), 8 * nb_, hipMemcpyHostToDevice, st));
            EPSX_HIP(ctx, hipMemcpyAsync(norms.b.p, normsh.data(), 8 * (size_t)cap * nb_, hipMemcpyHostToDevice, st));
            EPSX_HIP(ctx, hipMemcpyAsync(zv.b.p, z0.data(), 16 * (size_t)(cap + 1) * nb_, hipMemcpyHostToDevice, st));
            // V_0 = r / β for every instance (the list still holds all of them)
            hipLaunchKernelGGL(cscale_b_kernel, dim3(nb(n), (unsigned)nb_), dim3(256), 0, st, n, dl, invh.ptr<double>(),
                               dJ0, n, dV, vst);
            EPSX_HIP(ctx, hipGetLastError());
            for (int64_t j = 0; j < it_max && !active.empty(); j++) {
                const int nA = (int)active.size();
                EPSX_HIP(ctx, hipMemcpyAsync(d_list.b.p, active.data(), 4 * nA, hipMemcpyHostToDevice, st));
                // q = V_j - L(V_j)
                EPSX_TRY(ladder_apply_batch(ctx, bw, nA, active, dl, dV + n * j, vst, dLJ, n, Gm.ptr<double2>(),
                                            Gp.ptr<double2>(), par->precision));
                hipLaunchKernelGGL(combine2_b_kernel, dim3(nb(n), (unsigned)nA), dim3(256), 0, st, n, dl, dV + n * j, vst,
                                   -1.0, dLJ, n, dQ, n);
                EPSX_HIP(ctx, hipGetLastError());
                for (int64_t i = 0; i <= j; i++) {
                    EPSX_TRY(dots(nA, dV + n * i, vst, dQ, n, i));
                    hipLaunchKernelGGL(caxpy_b_kernel, dim3(nb(n), (unsigned)nA), dim3(256), 0, st, n, dl, dh, cap + 1, i,
                                       dV + n * i, vst, dQ, n);
                }
                EPSX_TRY(dots(nA, dQ, n, dQ, n, j + 1));
                hipLaunchKernelGGL(givens_b_kernel, dim3((unsigned)((nA + 63) / 64)), dim3(64), 0, st, nA, dl, gs, j);
                EPSX_HIP(ctx, hipGetLastError());
                EPSX_HIP(ctx, hipMemcpyAsync(st_h.data(), status.b.p, 4 * nb_, hipMemcpyDeviceToHost, st));
                EPSX_HIP(ctx, hipStreamSynchronize(st));
                std::vector<int32_t> next;
                for (int32_t z : active) {
                    niter[z] = j + 1;
                    if (st_h[z] == 1) conv[z] = 1;
                    else if (st_h[z] == 0 && j + 1 < it_max) next.push_back(z);
                }
                // V_{j+1} = q / h_{j+1,j} for the instances that go on
                if (!next.empty()) {
                    EPSX_HIP(ctx, hipMemcpyAsync(d_list.b.p, next.data(), 4 * next.size(), hipMemcpyHostToDevice, st));
                    hipLaunchKernelGGL(cscale_b_kernel, dim3(nb(n), (unsigned)next.size()), dim3(256), 0, st, n, dl,
                                       invh.ptr<double>(), dQ, n, dV + n * (j + 1), vst);
                    EPSX_HIP(ctx, hipGetLastError());
                }
                active.swap(next);
            }
            // x = V (R \ z) for every instance that ran
            std::vector<int32_t> ran;
            for (int z = 0; z < nb_; z++)
                if (niter[z] > 0) ran.push_back(z);
            if (!ran.empty()) {
                EPSX_HIP(ctx, hipMemcpyAsync(niter_d.b.p, niter.data(), 8 * nb_, hipMemcpyHostToDevice, st));
                EPSX_HIP(ctx, hipMemcpyAsync(d_list.b.p, ran.data(), 4 * ran.size(), hipMemcpyHostToDevice, st));
                const int nr = (int)ran.size();
                hipLaunchKernelGGL(gmres_solve_b_kernel, dim3((unsigned)((nr + 63) / 64)), dim3(64), 0, st, nr, dl, gs,
                                   niter_d.ptr<int64_t>(), y.ptr<double2>());
                hipLaunchKernelGGL(gmres_update_b_kernel, dim3(nb(n), (unsigned)nr), dim3(256), 0, st, n, dl,
                                   niter_d.ptr<int64_t>(), y.ptr<double2>(), cap, dV, vst, dX, n);
                EPSX_HIP(ctx, hipGetLastError());
            }
            EPSX_HIP(ctx, hipMemcpyAsync(normsh.data(), norms.b.p, 8 * (size_t)cap * nb_, hipMemcpyDeviceToHost, st));
            EPSX_HIP(ctx, hipStreamSynchronize(st));
            for (int z = 0; z < nb_; z++) {
                const int64_t slot = o0 + z + nO * iT;
                const double Om = par->Omegas[o0 + z];
                out->converged[slot] = conv[z];
                out->niter[slot] = niter[z];
                double *nm_out = out->norms + (it_max + 1) * slot;
                for (int64_t q = 0; q <= niter[z]; q++) nm_out[q] = normsh[cap * z + q];
                EPSX_TRY(sigma_of(ctx, dX + n * z, Om, T, out->sigma_ladder + 9 * slot));
                if (par->store_vertex && out->J)
                    EPSX_TRY(read_back(ctx, out->J + 2 * n * slot, dX + n * z, 16 * (size_t)n));
                if (par->verbosity > 0)
                    fprintf(stderr, "ladder (batched) T = %g, Ω = %g: niter %lld converged %d\n", T, Om,
                            (long long)niter[z], conv[z]);
            }
        }
    }
    EPSX_HIP(ctx, hipStreamSynchronize(st));
    return EPSX_OK;
}

extern "C" int epsx_ladder_run(epsx_ctx *ctx, const epsx_ladder_params *par, epsx_ladder_outputs *out)
{
    using cd = std::complex<double>;
    if (!ctx) return EPSX_ERR_ARG;
    LadderDev &l = ctx->l;
    Scgd0Dev &s = ctx->s;
    if (!l.ready) return set_error(ctx, EPSX_ERR_STATE, "epsx_ladder_run before epsx_ladder_set_problem");
    if (!par || par->struct_size != (int64_t)sizeof(epsx_ladder_params) || !out ||
        out->struct_size != (int64_t)sizeof(epsx_ladder_outputs))
        return set_error(ctx, EPSX_ERR_ARG, "epsx_ladder_run: struct size mismatch (ABI %d)", EPSX_ABI_VERSION);
    if (par->solver != EPSX_SOLVER_GMRES && par->solver != EPSX_SOLVER_FIXED_POINT)
        return set_error(ctx, EPSX_ERR_ARG, "solver must be :gmres or :fixed_point, got %d", par->solver);
    if (par->maxiter < 1) return set_error(ctx, EPSX_ERR_ARG, "maxiter must be at least 1, got %lld", (long long)par->maxiter);
    if (par->nOmega < 1) return set_error(ctx, EPSX_ERR_ARG, "Ωs must hold at least one probe frequency");
    if (!(par->w_acoustic > 0))
        return set_error(ctx, EPSX_ERR_ARG, "ω_acoustic must be positive, got %g: the shift-grid coupling build has "
                         "no row for a negative phonon frequency", par->w_acoustic);
    if (par->nT < 1 || !par->Tlist || !par->mu || !par->Sigma || !par->Omegas)
        return set_error(ctx, EPSX_ERR_ARG, "ladder run: null input");
    if (par->precision != EPSX_PREC_FP64 && par->precision != EPSX_PREC_FP32 && par->precision != EPSX_PREC_BF16)
        return set_error(ctx, EPSX_ERR_ARG, "unknown ladder precision %d", par->precision);
    if (!out->sigma_bubble || !out->sigma_ladder || !out->converged || !out->niter || !out->norms || !out->sigma_iter)
        return set_error(ctx, EPSX_ERR_ARG, "ladder run: null output");
    EPSX_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t nw = s.nw, ni = s.ni, n = nw * 24 * ni, nT = par->nT, nO = par->nOmega, it_max = par->maxiter;
    const hipStream_t st = ctx->stream;
    EPSX_TRY(dev_alloc(ctx, l.S, 16 * (size_t)nw * ni));
    EPSX_TRY(dev_alloc(ctx, l.jm, 4 * nw));
    EPSX_TRY(dev_alloc(ctx, l.tm, 8 * nw));
    EPSX_TRY(dev_alloc(ctx, l.jp, 4 * nw));
    EPSX_TRY(dev_alloc(ctx, l.tp, 8 * nw));
    const int gm = par->solver == EPSX_SOLVER_GMRES;
    for (int k = 0; k < 8; k++) { ctx->t_class[k] = 0; ctx->n_class[k] = 0; }
    for (int64_t q = 0; q < (it_max + 1) * nO * nT; q++) out->norms[q] = NAN;
    for (int64_t q = 0; q < 9 * it_max * nO * nT; q++) out->sigma_iter[q] = NAN;
    if (gm && par->batched) return ladder_run_batched(ctx, par, out);
    Tmp J0, J, LJ, R, V, hcol;
    EPSX_TRY(tmp_alloc(ctx, J0.b, 16 * n));
    EPSX_TRY(tmp_alloc(ctx, J.b, 16 * n));
    EPSX_TRY(tmp_alloc(ctx, LJ.b, 16 * n));
    EPSX_TRY(tmp_alloc(ctx, R.b, 16 * n));
    if (gm) EPSX_TRY(tmp_alloc(ctx, V.b, 16 * (size_t)n * (it_max + 1)));
    EPSX_TRY(tmp_alloc(ctx, hcol.b, 16 * (size_t)(it_max + 2)));
    const double btol = std::pow(2.220446049250313e-16, 0.75);
    BatchWork bw;
    EPSX_TRY(batch_work(ctx, bw, 1, par->precision));
    std::vector<int32_t> jv(nw);
    std::vector<double> tv(nw), xs(nw);
    double2 *dJ0 = J0.ptr<double2>(), *dJ = J.ptr<double2>(), *dLJ = LJ.ptr<double2>(), *dR = R.ptr<double2>();
    double2 *dh = hcol.ptr<double2>();
    auto launch_n = [&](int64_t m) { return dim3(nb(m)); };
    for (int64_t iT = 0; iT < nT; iT++) {
        const double T = par->Tlist[iT], mu = par->mu[iT];
        if (!(T >= 0)) return set_error(ctx, EPSX_ERR_ARG, "Temperature must be positive");
        EPSX_TRY(ensure_Bt(ctx, T, par->w_acoustic));
        EPSX_HIP(ctx, hipMemcpyAsync(l.S.p, par->Sigma + 2 * nw * ni * iT, 16 * nw * ni, hipMemcpyHostToDevice, st));
        for (int64_t iO = 0; iO < nO; iO++) {
            const double Om = par->Omegas[iO];
            const int64_t slot = iO + nO * iT;
            // Σ at ωs ∓ Ω/2 through host tables, then the propagators and the bare vertex
            for (int64_t w = 0; w < nw; w++) xs[w] = s.ws_host[w] - Om / 2;
            linear_interpolation_table(nw, s.ws_host.data(), nw, xs.data(), jv.data(), tv.data());
            EPSX_HIP(ctx, hipMemcpyAsync(l.jm.p, jv.data(), 4 * nw, hipMemcpyHostToDevice, st));
            EPSX_HIP(ctx, hipMemcpyAsync(l.tm.p, tv.data(), 8 * nw, hipMemcpyHostToDevice, st));
            for (int64_t w = 0; w < nw; w++) xs[w] = s.ws_host[w] + Om / 2;
            linear_interpolation_table(nw, s.ws_host.data(), nw, xs.data(), jv.data(), tv.data());
            EPSX_HIP(ctx, hipMemcpyAsync(l.jp.p, jv.data(), 4 * nw, hipMemcpyHostToDevice, st));
            EPSX_HIP(ctx, hipMemcpyAsync(l.tp.p, tv.data(), 8 * nw, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(greens_kernel, launch_n(nw * ni), dim3(256), 0, st, nw, ni, dptr<double>(s.ws),
                               dptr<double>(s.e_i), dptr<double2>(l.S), dptr<int32_t>(l.jm), dptr<double>(l.tm),
                               dptr<int32_t>(l.jp), dptr<double>(l.tp), Om, mu, T, dptr<double2>(l.Gm),
                               dptr<double2>(l.Gp));
            hipLaunchKernelGGL(bare_vertex_kernel, launch_n(n), dim3(256), 0, st, nw, ni, dptr<double2>(l.Gm),
                               dptr<double>(l.v), dptr<double2>(l.Gp), dJ0);
            EPSX_HIP(ctx, hipGetLastError());
            EPSX_TRY(sigma_of(ctx, dJ0, Om, T, out->sigma_bubble + 9 * slot));
            int64_t niter = 0;
            int conv = 0;
            if (gm) {
                // (I - L) J = J0 by GMRES (Krylov.jl: MGS Arnoldi, Givens QR, no restart), warm-started
                // from the previous Ω at this T, stopping at rtol ||J0||.
                double2 nrm;
                EPSX_TRY(ldot(ctx, n, dJ0, dJ0, dh));
                EPSX_HIP(ctx, hipMemcpyAsync(&nrm, dh, 16, hipMemcpyDeviceToHost, st));
                EPSX_HIP(ctx, hipStreamSynchronize(st));
                const double atol = par->rtol * std::sqrt(nrm.x);
                const bool warm = iO > 0;
                if (warm) {
                    EPSX_TRY(ladder_apply_dev(ctx, bw, dJ, dLJ, par->precision));
                    // r = b - (x - Lx)
                    hipLaunchKernelGGL(combine3_kernel, launch_n(n), dim3(256), 0, st, n, dJ0, -1.0, dJ, 1.0, dLJ, dR);
                } else {
                    EPSX_HIP(ctx, hipMemcpyAsync(dR, dJ0, 16 * n, hipMemcpyDeviceToDevice, st));
                    EPSX_HIP(ctx, hipMemsetAsync(dJ, 0, 16 * n, st));
                }
                EPSX_TRY(ldot(ctx, n, dR, dR, dh));
                EPSX_HIP(ctx, hipMemcpyAsync(&nrm, dh, 16, hipMemcpyDeviceToHost, st));
                EPSX_HIP(ctx, hipStreamSynchronize(st));
                const double beta = std::sqrt(nrm.x);
                double *norms = out->norms + (it_max + 1) * slot;
                int64_t hn = 0;
                norms[hn++] = beta;
                bool solved = beta <= atol;
                if (beta > 0 && !solved) {
                    std::vector<cd> Rm((it_max + 1) * (it_max + 1)), sn(it_max + 1), z(it_max + 2), hc(it_max + 2);
                    std::vector<double> cs(it_max + 1);
                    double2 *Vb = V.ptr<double2>();
                    hipLaunchKernelGGL(cscale_kernel, launch_n(n), dim3(256), 0, st, n, 1.0 / beta, dR, Vb);
                    z[0] = beta;
                    bool breakdown = false;
                    const int64_t cap = it_max + 1;
                    while (!(solved || niter >= it_max || breakdown)) {
                        const int64_t j = niter++;
                        double2 *Vj = Vb + n * j, *q = dR;
                        EPSX_TRY(ladder_apply_dev(ctx, bw, Vj, dLJ, par->precision));
                        hipLaunchKernelGGL(combine3_kernel, launch_n(n), dim3(256), 0, st, n, Vj, -1.0, dLJ, 0.0, dLJ, q);
                        for (int64_t i = 0; i <= j; i++) {
                            EPSX_TRY(ldot(ctx, n, Vb + n * i, q, dh + i));
                            hipLaunchKernelGGL(caxpy_dev_kernel, launch_n(n), dim3(256), 0, st, n, dh + i, Vb + n * i, q);
                        }
                        EPSX_TRY(ldot(ctx, n, q, q, dh + j + 1));
                        std::vector<double2> hv(j + 2);
                        EPSX_HIP(ctx, hipMemcpyAsync(hv.data(), dh, 16 * (j + 2), hipMemcpyDeviceToHost, st));
                        EPSX_HIP(ctx, hipStreamSynchronize(st));
                        for (int64_t i = 0; i <= j; i++) Rm[i + cap * j] = cd(hv[i].x, hv[i].y);
                        const double Hbis = std::sqrt(hv[j + 1].x);
                        for (int64_t i = 0; i < j; i++) {
                            const cd a = Rm[i + cap * j], b = Rm[i + 1 + cap * j];
                            Rm[i + cap * j] = cs[i] * a + sn[i] * b;
                            Rm[i + 1 + cap * j] = std::conj(sn[i]) * a - cs[i] * b;
                        }
                        cd rho;
                        sym_givens(Rm[j + cap * j], Hbis, cs[j], sn[j], rho);
                        Rm[j + cap * j] = rho;
                        z[j + 1] = std::conj(sn[j]) * z[j];
                        z[j] = cs[j] * z[j];
                        const double rn = std::abs(z[j + 1]);
                        norms[hn++] = rn;
                        breakdown = Hbis <= btol;
                        solved = rn <= atol;
                        if (!(solved || niter >= it_max || breakdown))
                            hipLaunchKernelGGL(cscale_kernel, launch_n(n), dim3(256), 0, st, n, 1.0 / Hbis, q, Vb + n * (j + 1));
                    }
                    // y = R \ z; x += V y
                    std::vector<cd> y(niter);
                    for (int64_t i = niter - 1; i >= 0; i--) {
                        cd sacc = z[i];
                        for (int64_t k = i + 1; k < niter; k++) sacc -= Rm[i + cap * k] * y[k];
                        y[i] = sacc / Rm[i + cap * i];
                    }
                    for (int64_t i = 0; i < niter; i++)
                        hipLaunchKernelGGL(caxpy_kernel, launch_n(n), dim3(256), 0, st, n, y[i].real(), y[i].imag(), Vb + n * i, dJ);
                    EPSX_HIP(ctx, hipGetLastError());
                }
                conv = solved;
            } else {
                // J ← J0 + L(J) from J0 until σ changes by less than rtol (src/ladder.jl:354-369)
                EPSX_HIP(ctx, hipMemcpyAsync(dJ, dJ0, 16 * n, hipMemcpyDeviceToDevice, st));
                double *sit = out->sigma_iter + 9 * it_max * slot;
                for (int64_t it = 1; it <= it_max; it++) {
                    EPSX_TRY(ladder_apply_dev(ctx, bw, dJ, dLJ, par->precision));
                    hipLaunchKernelGGL(combine3_kernel, launch_n(n), dim3(256), 0, st, n, dJ0, 1.0, dLJ, 0.0, dLJ, dJ);
                    EPSX_TRY(sigma_of(ctx, dJ, Om, T, sit + 9 * (it - 1)));
                    niter = it;
                    if (it > 1) {
                        double d = 0, nn = 0;
                        for (int k = 0; k < 9; k++) {
                            const double a = sit[9 * (it - 1) + k], b = sit[9 * (it - 2) + k];
                            d += (a - b) * (a - b);
                            nn += a * a;
                        }
                        if (std::sqrt(d) < par->rtol * std::sqrt(nn)) { conv = 1; break; }
                    }
                }
            }
            out->converged[slot] = conv;
            out->niter[slot] = niter;
            EPSX_TRY(sigma_of(ctx, dJ, Om, T, out->sigma_ladder + 9 * slot));
            if (par->store_vertex && out->J) {
                EPSX_HIP(ctx, hipMemcpyAsync(out->J + 2 * n * slot, dJ, 16 * n, hipMemcpyDeviceToHost, st));
                EPSX_HIP(ctx, hipStreamSynchronize(st));
            }
            if (par->verbosity > 0)
                fprintf(stderr, "ladder T = %g, Ω = %g: niter %lld converged %d\n", T, Om, (long long)niter, conv);
        }
    }
    EPSX_HIP(ctx, hipStreamSynchronize(st));
    return EPSX_OK;
}

namespace epsx {

void blas_destroy(epsx_ctx *ctx)
{
    if (ctx->blas) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->blas));
    ctx->blas = nullptr;
}

}  // namespace epsx
