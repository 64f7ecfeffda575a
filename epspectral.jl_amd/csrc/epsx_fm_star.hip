// The factorised Fan-Migdal sum for dispersionless (Einstein) modes and star-invariant
// energies: an exact rewrite of src/scgd0_kernel.jl:27-86 behind two checked preconditions.
//
// With ωq = ω_ν for every pair and ε_f == ε_{i_f} (so that the inner state f enters the term
// only through w_f, g2 and its representative i_f = i'), the sum factorises:
//
//   ImΣ[ω, i] = Σ_{ν, i'} h[ω, ν, i'] G̃[(ν, i'), i]
//   h[ω, ν, i'] = Im G_{i'}(ω+ω_ν) (n_ν + f(ω+ω_ν-μ)) + Im G_{i'}(ω-ω_ν) (n_ν + 1 - f(ω-ω_ν-μ))
//   G̃[(ν, i'), i] = Σ_{f in star(i')} w_f g2[ν, i, f]
//
// G̃ depends on neither T nor the iteration: it is one streaming pass over g2 (HBM-bound, the
// kernel that reads the whole coupling), after which every iteration is an h build
// (nω nm ni terms per T) and an FP64 MFMA GEMM of nω x (nm ni) x ni.  Modes below ω_acoustic
// are zeroed in G̃ without reading g2 (which may hold Inf there).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "epsx_common.h"
#include "epsx_device.h"

namespace epsx {
namespace {

constexpr int SR_THREADS = 256;

// One block per (chunk of whole stars, outer state i).  The chunk's slab of g2p[:, p, i]
// (contiguous, n ≤ cap - 2 doubles, cap ≤ CAP) is read with 16-byte non-temporal loads from the 16-byte
// boundary at or below its start (sh = 0 or 1 leading doubles belong to the previous chunk), all of
// a thread's loads unconditional and in flight at once (indices clamped to the last pair), and
// staged in LDS together with w_f of its rows; then one thread per (ν, i') sums its star in
// ascending p.  A star too large for the stage is a chunk of its own, read from global memory.
// The last pair may read one double past the slab: the g2p allocation carries two doubles of
// padding.
template <int CAP>
__global__ __launch_bounds__(SR_THREADS) void star_reduce_kernel(
    int cap, int64_t nm, int64_t ni, int64_t nfs, const double *__restrict__ g2p, const double *__restrict__ wf,
    const int64_t *__restrict__ star_ptr, const int32_t *__restrict__ chunk_star, const uint8_t *__restrict__ skip_mode,
    double *__restrict__ Gt)
{
    constexpr int PER = CAP / (2 * SR_THREADS);   // 16-byte loads per thread
    __shared__ double2 buf2[CAP / 2];
    __shared__ double wl[CAP / 4];
    const double *buf = reinterpret_cast<const double *>(buf2);
    const int tid = threadIdx.x;
    const int64_t c = blockIdx.x, i = blockIdx.y;
    const int64_t s0 = chunk_star[c], s1 = chunk_star[c + 1];
    const int64_t p0 = star_ptr[s0], p1 = star_ptr[s1];
    const int64_t rows = p1 - p0;
    const int nm32 = (int)nm;
    const int64_t g0 = nm * (p0 + nfs * i);
    const int sh = (int)(g0 & 1);
    const bool staged = rows * nm <= cap - 2 && rows <= cap / 4;   // cap ≤ CAP (0: nothing staged)
    if (staged) {
        const int n = (int)(rows * nm);
        const int np = (n + sh + 1) >> 1;
        const double2 *src = reinterpret_cast<const double2 *>(g2p + (g0 - sh));
        double2 v[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int j = tid + SR_THREADS * q;
            const double2 *ptr = src + (j < np ? j : np - 1);
            v[q].x = __builtin_nontemporal_load(&ptr->x);
            v[q].y = __builtin_nontemporal_load(&ptr->y);
        }
        for (int r = tid; r < (int)rows; r += SR_THREADS) wl[r] = wf[p0 + r];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int j = tid + SR_THREADS * q;
            if (j < np) buf2[j] = v[q];
        }
        __syncthreads();
    }
    const double *src = g2p + g0;
    const int nout = (int)((s1 - s0) * nm);
    for (int o = tid; o < nout; o += SR_THREADS) {
        const int nu = o % nm32;
        const int64_t ip = s0 + o / nm32;
        double acc = 0.0;
        if (!skip_mode[nu]) {
            const int a = (int)(star_ptr[ip] - p0), b = (int)(star_ptr[ip + 1] - p0);
            if (staged) {
#pragma unroll 4
                for (int pl = a; pl < b; pl++) acc += wl[pl] * buf[pl * nm32 + nu + sh];
            } else {
                for (int pl = a; pl < b; pl++) acc += wf[p0 + pl] * src[(int64_t)pl * nm + nu];
            }
        }
        Gt[(nu + nm * ip) + nm * ni * i] = acc;
    }
}

struct HArgs {
    int64_t nw, nm, ni;
    const double *ws;
    const int32_t *hk;        // [2, nw, nm] interpolation cell of ω ± ω_ν
    const double *ht;         // [2, nw, nm] its weight
    const double *wnu;        // [nm]
    const double *e_i;        // [ni]
    const double *S;          // complex [nw, ni, nT_all]
    const double *T, *mu, *eta;  // per temperature index
    const int32_t *tidx;      // slot -> temperature index
    double w_acoustic;
    double *occ;              // [2, nw, nm] per slot: n_ν + f(ω+ω_ν-μ), n_ν + 1 - f(ω-ω_ν-μ)
    double *hout;             // [nw, nm, ni] per slot, columns of pitch ldh
    int64_t ldh;              // column pitch of hout (>= nw)
    const int32_t *act;       // convergence mask per temperature index, or null
};

// The occupation factors of the two terms for every (ω, ν) of a slot (src/utils.jl:14-38):
// n = 1/expm1(ω_ν/T), f(x) = 1/(e^{x/T} + 1) and 1 - f(x) = 1/(e^{-x/T} + 1); step functions for
// T ≤ sqrt(eps).
__global__ __launch_bounds__(256) void h_occ_kernel(HArgs a)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    if (q >= a.nw * a.nm) return;
    const int tcol = a.tidx[z];
    if (a.act && !a.act[tcol]) return;
    const int64_t w = q % a.nw, nu = q / a.nw;
    const double wnu = a.wnu[nu], x0 = a.ws[w];
    const double T = a.T[tcol], mu = a.mu[tcol];
    double fa = 0, fe = 0;
    if (!(wnu < a.w_acoustic)) {
        const double ya = x0 + wnu - mu, ye = x0 - wnu - mu;
        if (T > 1.4901161193847656e-08) {
            const double nq = 1.0 / expm1(wnu / T);
            fa = nq + 1.0 / (1.0 + exp(ya / T));
            fe = nq + 1.0 / (1.0 + exp(-ye / T));
        } else {
            const double nq = wnu > 0 ? 0.0 : (wnu == 0 ? -0.5 : -1.0);
            fa = nq + (ya > 0 ? 0.0 : (ya < 0 ? 1.0 : 0.5));
            fe = nq + (ye > 0 ? 1.0 : (ye < 0 ? 0.0 : 0.5));
        }
    }
    double *o = a.occ + 2 * (q + a.nw * a.nm * (int64_t)z);
    o[0] = fa;
    o[1] = fe;
}

// What h[ω_w, ν, ·] of slot z needs beyond the outer state: the cells and weights of ω ± ω_ν, the
// occupations, ω_w, ω_ν, the clamp η and the slot's Σ.
struct HConst {
    int ka, ke;
    double ta, te, x0, wnu, eta;
    double2 o;
    const double2 *S;
};

__device__ inline HConst h_const(const HArgs &a, int64_t w, int64_t nu, int tcol, int z)
{
    const int64_t nw = a.nw, wn = w + nw * nu;
    HConst c;
    c.ka = a.hk[2 * wn];
    c.ke = a.hk[2 * wn + 1];
    c.ta = a.ht[2 * wn];
    c.te = a.ht[2 * wn + 1];
    c.x0 = a.ws[w];
    c.wnu = a.wnu[nu];
    c.eta = a.eta ? a.eta[tcol] : -INFINITY;
    c.o = reinterpret_cast<const double2 *>(a.occ)[wn + nw * a.nm * (int64_t)z];
    c.S = reinterpret_cast<const double2 *>(a.S) + nw * a.ni * (int64_t)tcol;
    return c;
}

// h[ω_w, ν, i'] for a mode at or above ω_acoustic.
__device__ inline double h_from(const HArgs &a, const HConst &c, int64_t ip)
{
    const double2 *col = c.S + a.nw * ip;
    const double2 la = col[c.ka], ha = col[c.ka + 1], le = col[c.ke], he = col[c.ke + 1];
    const double eta = c.eta;
    const double sar = (1 - c.ta) * la.x + c.ta * ha.x;
    const double sai = (1 - c.ta) * fmin(la.y, -eta) + c.ta * fmin(ha.y, -eta);
    const double ser = (1 - c.te) * le.x + c.te * he.x;
    const double sei = (1 - c.te) * fmin(le.y, -eta) + c.te * fmin(he.y, -eta);
    const double ei = a.e_i[ip];
    const double zar = c.x0 + c.wnu - ei - sar, zai = -sai;
    const double zer = c.x0 - c.wnu - ei - ser, zei = -sei;
    // Im 1/z_a o.x + Im 1/z_e o.y with Im 1/z = -Im z / |z|², over one division (the product of the
    // two squared moduli stays in FP64 range for |z| between 1e-75 and 1e75 eV)
    const double da = zar * zar + zai * zai, de = zer * zer + zei * zei;
    return -(zai * c.o.x * de + zei * c.o.y * da) / (da * de);
}

__device__ inline double h_elem(const HArgs &a, int64_t w, int64_t nu, int64_t ip, int tcol, int z)
{
    return h_from(a, h_const(a, w, nu, tcol, z), ip);
}

// Dispatch sends workgroup b of a launch to XCD b mod 8, and each XCD has its own L2.  With the
// x grid a multiple of 8, block b works as the logical block (b mod 8) gridDim.x / 8 + b / 8, so
// each XCD runs one contiguous eighth of the logical blocks: the blocks that read the same Σ
// columns (all (ω, ν) of a group of states) share them in one L2 instead of fetching them 8 times.
__device__ inline int64_t xcd_major_block()
{
    return (int64_t)(blockIdx.x % 8) * (gridDim.x / 8) + blockIdx.x / 8;
}

static inline unsigned xcd_grid(int64_t threads)
{
    return (unsigned)(((threads + 255) / 256 + 7) / 8 * 8);
}

// h[w + ldh (ν + nm i')] per slot, FP64.  One thread per (ω, ν) and HB_R consecutive outer states:
// the (ω, ν) constants are read once and the HB_R independent Σ reads are in flight together
// (lanes along ω, so every read and write is contiguous across a wave).  The threads sweep the
// padded columns (ldh rows, the rows w >= nω written as zeros), so with ldh a multiple of 64 each
// wave stores whole aligned 512-byte runs.
constexpr int HB_R = 4;
__global__ __launch_bounds__(256) void h_build_kernel(HArgs a)
{
    const int64_t q = xcd_major_block() * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    const int64_t nw = a.nw, ni = a.ni, ldh = a.ldh, ldm = ldh * a.nm, ngrp = (ni + HB_R - 1) / HB_R;
    if (q >= ldm * ngrp) return;
    const int tcol = a.tidx[z];
    if (a.act && !a.act[tcol]) return;
    const int64_t wp = q % ldm, ip0 = (q / ldm) * HB_R;
    const int64_t w = wp % ldh, nu = wp / ldh;
    double *out = a.hout + ldm * ni * (int64_t)z + wp;
    double v[HB_R];
    if (w >= nw || a.wnu[nu] < a.w_acoustic) {
#pragma unroll
        for (int r = 0; r < HB_R; r++) v[r] = 0.0;
    } else {
        const HConst c = h_const(a, w, nu, tcol, z);
#pragma unroll
        for (int r = 0; r < HB_R; r++) v[r] = h_from(a, c, ip0 + r < ni ? ip0 + r : ni - 1);
    }
#pragma unroll
    for (int r = 0; r < HB_R; r++)
        if (ip0 + r < ni) out[ldm * (ip0 + r)] = v[r];
}

// The same h with the Σ columns of HS consecutive outer states staged in LDS (η-clamped as they
// are staged), one workgroup per (state group, slot) sweeping every (ω, ν) of its states (the
// padded columns, as h_build_kernel): Σ is read from memory once per launch and the four
// interpolation reads of every term come from LDS; the (ω, ν) constants are read once per HS
// states.  Same operations in the same order as h_from, so the same bits.
template <int HS>
__global__ __launch_bounds__(512) void h_build_lds_kernel(HArgs a)
{
    extern __shared__ __align__(16) double2 h_lds[];   // [HS][nw]
    const int z = blockIdx.y;
    const int tcol = a.tidx[z];
    if (a.act && !a.act[tcol]) return;
    const int64_t nw = a.nw, ni = a.ni, nwm = nw * a.nm, i0 = (int64_t)blockIdx.x * HS;
    const int64_t ldh = a.ldh, ldm = ldh * a.nm;
    const int nr = (int)(ni - i0 < HS ? ni - i0 : HS);
    const double eta = a.eta ? a.eta[tcol] : -INFINITY;
    const double2 *S = reinterpret_cast<const double2 *>(a.S) + nw * (ni * (int64_t)tcol + i0);
    for (int64_t q = threadIdx.x; q < nw * nr; q += 512) {
        const double2 v = S[q];
        h_lds[q] = make_double2(v.x, fmin(v.y, -eta));
    }
    __syncthreads();
    double *out = a.hout + ldm * (ni * (int64_t)z + i0);
    const double2 *occ = reinterpret_cast<const double2 *>(a.occ) + nwm * (int64_t)z;
    // HU (ω, ν) per thread and pass, their constants loaded together
    constexpr int HU = 4;
    for (int64_t wb = threadIdx.x; wb < ldm; wb += 512 * HU) {
        int ka[HU], ke[HU];
        double ta[HU], te[HU], x0[HU], wnu[HU];
        double2 o[HU];
        bool live[HU], zero[HU];
#pragma unroll
        for (int u = 0; u < HU; u++) {
            const int64_t wp = wb + 512 * u;   // w + ldh ν: the row of the padded output columns
            live[u] = wp < ldm;
            const int64_t wpc = live[u] ? wp : ldm - 1;
            const int64_t wq = wpc % ldh, nu = wpc / ldh;
            const int64_t w = wq < nw ? wq : nw - 1, wc = w + nw * nu;
            wnu[u] = a.wnu[nu];
            zero[u] = wq >= nw || wnu[u] < a.w_acoustic;
            ka[u] = a.hk[2 * wc];
            ke[u] = a.hk[2 * wc + 1];
            ta[u] = a.ht[2 * wc];
            te[u] = a.ht[2 * wc + 1];
            x0[u] = a.ws[w];
            o[u] = occ[wc];
        }
#pragma unroll
        for (int u = 0; u < HU; u++) {
            if (!live[u]) continue;
            const int64_t wp = wb + 512 * u;
            double v[HS];
#pragma unroll
            for (int r = 0; r < HS; r++) {
                const int rr = r < nr ? r : nr - 1;
                const double2 *col = h_lds + nw * rr;
                const double2 la = col[ka[u]], ha = col[ka[u] + 1], le = col[ke[u]], he = col[ke[u] + 1];
                const double sar = (1 - ta[u]) * la.x + ta[u] * ha.x;
                const double sai = (1 - ta[u]) * la.y + ta[u] * ha.y;
                const double ser = (1 - te[u]) * le.x + te[u] * he.x;
                const double sei = (1 - te[u]) * le.y + te[u] * he.y;
                const double ei = a.e_i[i0 + rr];
                const double zar = x0[u] + wnu[u] - ei - sar, zai = -sai;
                const double zer = x0[u] - wnu[u] - ei - ser, zei = -sei;
                const double da = zar * zar + zai * zai, de = zer * zer + zei * zei;
                v[r] = zero[u] ? 0.0 : -(zai * o[u].x * de + zei * o[u].y * da) / (da * de);
            }
#pragma unroll
            for (int r = 0; r < HS; r++)
                if (r < nr) out[wp + ldm * r] = v[r];
        }
    }
}

// The Einstein factorisation (dispersionless modes, any inner-state energies): with ωq = ω_ν
// for every pair, the term of (ω, ν, i, f) is w_f g2[ν, i, f] H_ν[ω, f] with
//   H_ν[ω, f] = Im G_f(ω+ω_ν) (n_ν + f(ω+ω_ν-μ)) + Im G_f(ω-ω_ν) (n_ν + 1 - f(ω-ω_ν-μ)),
// G_f(x) = 1 / (x - ε_f - Σin_{i_f}(x)), so ImΣ = H' g2p with H'[ω, ν + nm p] = w_f H_ν[ω, f] over
// the star-sorted inner states p and g2p[ν + nm p, i] the resident coupling: one GEMM of
// nω x (nm nf) x ni per temperature, in chunks of inner states.  This kernel builds H' for the
// sorted inner states [p0, p0 + np) of every slot: one thread per (ω, ν) and EH_R consecutive
// inner states (the (ω, ν) constants read once), lanes along ω.  Modes below ω_acoustic are zero.
constexpr int EH_R = 4;
__global__ __launch_bounds__(256) void einstein_h_kernel(HArgs a, int64_t p0, int64_t np, const double *__restrict__ ef,
                                                         const double *__restrict__ wf, const int32_t *__restrict__ fi)
{
    const int64_t q = xcd_major_block() * blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    const int64_t nw = a.nw, ldh = a.ldh, ldm = ldh * a.nm, ngrp = (np + EH_R - 1) / EH_R;
    if (q >= ldm * ngrp) return;
    const int tcol = a.tidx[z];
    if (a.act && !a.act[tcol]) return;
    const int64_t wp = q % ldm, pl0 = (q / ldm) * EH_R;   // the padded columns, as h_build_kernel
    const int64_t w = wp % ldh, nu = wp / ldh;
    double *out = a.hout + ldm * np * (int64_t)z + wp;
    double v[EH_R];
    if (w >= nw || a.wnu[nu] < a.w_acoustic) {
#pragma unroll
        for (int r = 0; r < EH_R; r++) v[r] = 0.0;
    } else {
        const HConst c = h_const(a, w, nu, tcol, z);
        const double eta = c.eta;
#pragma unroll
        for (int r = 0; r < EH_R; r++) {
            const int64_t p = p0 + (pl0 + r < np ? pl0 + r : np - 1);
            const double2 *col = c.S + a.nw * fi[p];
            const double2 la = col[c.ka], ha = col[c.ka + 1], le = col[c.ke], he = col[c.ke + 1];
            const double sar = (1 - c.ta) * la.x + c.ta * ha.x;
            const double sai = (1 - c.ta) * fmin(la.y, -eta) + c.ta * fmin(ha.y, -eta);
            const double ser = (1 - c.te) * le.x + c.te * he.x;
            const double sei = (1 - c.te) * fmin(le.y, -eta) + c.te * fmin(he.y, -eta);
            const double e = ef[p];
            const double zar = c.x0 + c.wnu - e - sar, zai = -sai;
            const double zer = c.x0 - c.wnu - e - ser, zei = -sei;
            const double da = zar * zar + zai * zai, de = zer * zer + zei * zei;
            v[r] = wf[p] * (-(zai * c.o.x * de + zei * c.o.y * da) / (da * de));
        }
    }
#pragma unroll
    for (int r = 0; r < EH_R; r++)
        if (pl0 + r < np) out[ldm * (pl0 + r)] = v[r];
}

// Any non-finite coupling in a mode below ω_acoustic (skip[ν] = 1): the Einstein product would
// multiply it by a zero H column (NaN), where the term-by-term sum skips the mode before reading g2.
__global__ __launch_bounds__(256) void acoustic_finite_kernel(int64_t nm, int64_t n, const double *__restrict__ g2p,
                                                              const uint8_t *__restrict__ skip, int32_t *bad)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    if (skip[q % nm] && !isfinite(g2p[q])) atomicOr(bad, 1);
}

// h per slot in low precision (E = unsigned short: BF16 bits, or float), transposed for the
// low-precision product: hl[k + kp w] with k = ν + nm i' (zero for nm ni <= k < kp, kp a multiple
// of 8).  One block per (64 frequencies, HL_IP outer states, slot): one thread per (ω, ν) computes
// its HL_IP states with the (ω, ν) constants read once (HL_IP Σ reads in flight) into an LDS row
// of the block's nm HL_IP consecutive k, then whole 16-byte runs of each row are written out.
constexpr int HL_W = 64, HL_IP = 8, HL_NM_MAX = 32;
template <typename E>
__global__ __launch_bounds__(256) void h_build_lp_kernel(HArgs a, E *hl, int64_t kp)
{
    extern __shared__ __align__(16) unsigned char lds[];
    E *tile = reinterpret_cast<E *>(lds);   // [HL_W][nm HL_IP]
    constexpr int RN = 16 / (int)sizeof(E);
    const int nm = (int)a.nm, nk = nm * HL_IP;
    const int z = blockIdx.z;
    const int tcol = a.tidx[z];
    if (a.act && !a.act[tcol]) return;
    const int64_t nw = a.nw, ni = a.ni, w0 = (int64_t)blockIdx.x * HL_W, ip0 = (int64_t)blockIdx.y * HL_IP;
    for (int p = threadIdx.x; p < HL_W * nm; p += 256) {
        const int wl = p % HL_W, nu = p / HL_W;
        const int64_t w = w0 + wl;
        if (w >= nw) continue;
        E *row = tile + wl * nk;
        double v[HL_IP];
        if (a.wnu[nu] < a.w_acoustic) {
#pragma unroll
            for (int r = 0; r < HL_IP; r++) v[r] = 0.0;
        } else {
            const HConst c = h_const(a, w, nu, tcol, z);
#pragma unroll
            for (int r = 0; r < HL_IP; r++) {
                const int64_t ip = ip0 + r;
                const double h = h_from(a, c, ip < ni ? ip : ni - 1);
                v[r] = ip < ni ? h : 0.0;   // k >= nm ni: zero
            }
        }
#pragma unroll
        for (int r = 0; r < HL_IP; r++) {
            if constexpr (sizeof(E) == 2) row[nu + nm * r] = to_bf16(v[r]);
            else row[nu + nm * r] = (float)v[r];
        }
    }
    __syncthreads();
    const int nrun = nk / RN;   // HL_IP = 8: whole runs
    for (int q = threadIdx.x; q < HL_W * nrun; q += 256) {
        const int run = q % nrun, wl = q / nrun;
        const int64_t w = w0 + wl, k0 = (int64_t)nm * ip0 + (int64_t)RN * run;
        if (w >= nw || k0 >= kp) continue;
        uint4 pk;
        __builtin_memcpy(&pk, tile + wl * nk + RN * run, 16);
        *reinterpret_cast<uint4 *>(hl + kp * nw * (int64_t)z + k0 + kp * w) = pk;
    }
}

// The same for nm > HL_NM_MAX (the LDS row would not fit): one thread per (ω, 16-byte run of k),
// lanes along ω, the 8 waves of a block along k.
template <typename E>
__global__ __launch_bounds__(512) void h_build_lp_run_kernel(HArgs a, E *hl, int64_t kp)
{
    constexpr int R = 16 / (int)sizeof(E);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t w = (int64_t)blockIdx.x * 64 + lane;
    const int64_t k0 = ((int64_t)blockIdx.y * 8 + wave) * R;
    const int z = blockIdx.z;
    if (w >= a.nw || k0 >= kp) return;
    const int tcol = a.tidx[z];
    if (a.act && !a.act[tcol]) return;
    const int64_t K = a.nm * a.ni;
    E v[R];
#pragma unroll
    for (int e = 0; e < R; e++) {
        const int64_t k = k0 + e;
        double h = 0.0;
        if (k < K) {
            const int64_t nu = k % a.nm, ip = k / a.nm;
            if (a.wnu[nu] >= a.w_acoustic) h = h_elem(a, w, nu, ip, tcol, z);
        }
        if constexpr (sizeof(E) == 2) v[e] = to_bf16(h);
        else v[e] = (float)h;
    }
    uint4 pk;
    __builtin_memcpy(&pk, v, 16);
    *reinterpret_cast<uint4 *>(hl + kp * a.nw * (int64_t)z + k0 + kp * w) = pk;
}

// G̃ in low precision with the row pitch of the low-precision product: gl[k + kp i] = G̃[k + K i]
// (zero for K <= k < kp).
template <typename E>
__global__ __launch_bounds__(256) void gt_lp_kernel(int64_t K, int64_t ldg, int64_t kp, int64_t ni,
                                                    const double *__restrict__ Gt, E *__restrict__ gl)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= kp * ni) return;
    const int64_t k = q % kp, i = q / kp;
    const double g = k < K ? Gt[k + ldg * i] : 0.0;
    if constexpr (sizeof(E) == 2) gl[q] = to_bf16(g);
    else gl[q] = (float)g;
}

}  // namespace

// The stage size of the reduction: the smallest of 1024, 2048, 4096 doubles that holds the largest
// star (rows * nm ≤ cap - 2, rows ≤ cap / 4), else 4096 (larger stars are read from global memory).
// EPSX_STAR_STAGE=global stages nothing (every star read from global memory; for tests).
int star_stage_cap(int64_t nm, const std::vector<int64_t> &star_ptr)
{
    const char *force = getenv("EPSX_STAR_STAGE");
    if (force && std::string(force) == "global") return 0;
    int64_t rmax = 0;
    for (size_t s = 0; s + 1 < star_ptr.size(); s++) rmax = std::max(rmax, star_ptr[s + 1] - star_ptr[s]);
    for (int cap : {1024, 2048, 4096})
        if (rmax * nm <= cap - 2 && rmax <= cap / 4) return cap;
    return 4096;
}

// Chunks of whole stars that fit a stage of cap doubles (a larger single star is a chunk of its
// own, read from global memory).
void star_chunks(int64_t nm, int cap, const std::vector<int64_t> &star_ptr, std::vector<int32_t> &chunk)
{
    const int64_t ns = (int64_t)star_ptr.size() - 1;
    chunk.clear();
    chunk.push_back(0);
    int64_t s = 0;
    while (s < ns) {
        int64_t e = s + 1;
        while (e < ns && (star_ptr[e + 1] - star_ptr[s]) * nm <= cap - 2 && star_ptr[e + 1] - star_ptr[s] <= cap / 4)
            e++;
        chunk.push_back((int32_t)e);
        s = e;
    }
}

int star_reduce_launch(epsx_ctx *ctx, int cap, int64_t nm, int64_t ni, int64_t nfs, const double *g2p,
                       const double *wf, const int64_t *d_star_ptr, const int32_t *d_chunks, int64_t nchunks,
                       const uint8_t *d_skip, double *Gt)
{
    const dim3 grid((unsigned)nchunks, (unsigned)ni), block(SR_THREADS);
    switch (cap) {
    case 1024:
        hipLaunchKernelGGL(star_reduce_kernel<1024>, grid, block, 0, ctx->stream, cap, nm, ni, nfs, g2p, wf, d_star_ptr,
                           d_chunks, d_skip, Gt);
        break;
    case 2048:
        hipLaunchKernelGGL(star_reduce_kernel<2048>, grid, block, 0, ctx->stream, cap, nm, ni, nfs, g2p, wf, d_star_ptr,
                           d_chunks, d_skip, Gt);
        break;
    default:   // 4096, and 0 (nothing staged: every chunk exceeds the stage)
        hipLaunchKernelGGL(star_reduce_kernel<4096>, grid, block, 0, ctx->stream, cap, nm, ni, nfs, g2p, wf, d_star_ptr,
                           d_chunks, d_skip, Gt);
        break;
    }
    EPSX_HIP(ctx, hipGetLastError());
    return EPSX_OK;
}

// Cell and weight of ω_w ± ω_ν on the integer grid (off: offsets from the first node, ltab: cell of
// every integer position, rgap: 1 / cell width): flat extrapolation outside, as a cell with t = 0
// (below) or t = 1 (above).
void star_lookup_host(int64_t nw, int64_t nm, const double *wnu, double h, const std::vector<int32_t> &off,
                      const std::vector<int32_t> &ltab, const std::vector<double> &rgap, int64_t span,
                      std::vector<int32_t> &hk, std::vector<double> &ht)
{
    hk.assign(2 * nw * nm, 0);
    ht.assign(2 * nw * nm, 0.0);
    for (int64_t nu = 0; nu < nm; nu++)
        for (int64_t w = 0; w < nw; w++)
            for (int sgn = 0; sgn < 2; sgn++) {
                const double x = (sgn == 0 ? wnu[nu] : -wnu[nu]) / h;
                const double fl = std::floor(x);
                const int64_t pos = off[w] + (int64_t)fl;
                const double frac = x - fl;
                int32_t k;
                double t;
                if (pos < 0) { k = 0; t = 0.0; }
                else if (pos >= span) { k = (int32_t)(nw - 2); t = 1.0; }
                else {
                    k = ltab[pos];
                    t = ((double)(pos - off[k]) + frac) * rgap[k];
                }
                hk[2 * (w + nw * nu) + sgn] = k;
                ht[2 * (w + nw * nu) + sgn] = t;
            }
}

int h_build_launch(epsx_ctx *ctx, int64_t nw, int64_t nm, int64_t ni, const double *ws, const int32_t *hk,
                   const double *ht, const double *wnu, const double *e_i, const double *S, const double *T,
                   const double *mu, const double *eta, const int32_t *tidx, int nslots, double w_acoustic,
                   double *occ, double *hout, const int32_t *act, const int32_t *act_h, uint16_t *hb,
                   const int32_t *act_b, float *hf, const int32_t *act_f, int64_t kp, int64_t ldh)
{
    HArgs a;
    a.nw = nw; a.nm = nm; a.ni = ni; a.ws = ws; a.hk = hk; a.ht = ht; a.ldh = ldh;
    a.wnu = wnu; a.e_i = e_i; a.S = S; a.T = T; a.mu = mu; a.eta = eta; a.tidx = tidx;
    a.w_acoustic = w_acoustic; a.occ = occ; a.hout = hout; a.act = act;
    hipLaunchKernelGGL(h_occ_kernel, dim3((unsigned)((nw * nm + 255) / 256), (unsigned)nslots), dim3(256), 0,
                       ctx->stream, a);
    EPSX_HIP(ctx, hipGetLastError());
    if (hout) {
        HArgs ah = a;
        ah.act = act_h;
        // the LDS-staged build while HS columns fit 80 KB (two workgroups per CU), else the
        // gather from memory; a kernel's LDS limit is raised past the default 64 KB once per
        // process (the first build runs eagerly, before any graph capture)
        const size_t col = 16 * (size_t)nw, cap = 80 * 1024;
        if (4 * col <= cap) {
            static bool raised4 = false;
            if (4 * col > 64 * 1024 && !raised4) {
                EPSX_HIP(ctx, hipFuncSetAttribute((const void *)h_build_lds_kernel<4>,
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)cap));
                raised4 = true;
            }
            hipLaunchKernelGGL(h_build_lds_kernel<4>, dim3((unsigned)((ni + 3) / 4), (unsigned)nslots), dim3(512),
                               4 * col, ctx->stream, ah);
        } else if (2 * col <= cap) {
            static bool raised2 = false;
            if (2 * col > 64 * 1024 && !raised2) {
                EPSX_HIP(ctx, hipFuncSetAttribute((const void *)h_build_lds_kernel<2>,
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)cap));
                raised2 = true;
            }
            hipLaunchKernelGGL(h_build_lds_kernel<2>, dim3((unsigned)((ni + 1) / 2), (unsigned)nslots), dim3(512),
                               2 * col, ctx->stream, ah);
        } else {
            const int64_t threads = ldh * nm * ((ni + HB_R - 1) / HB_R);
            hipLaunchKernelGGL(h_build_kernel, dim3(xcd_grid(threads), (unsigned)nslots), dim3(256), 0, ctx->stream, ah);
        }
        EPSX_HIP(ctx, hipGetLastError());
    }
    // the transposed low-precision layouts: blocks of HL_IP states until nm ip covers kp
    const bool rows = nm <= HL_NM_MAX;
    const int64_t ipk = (kp + nm - 1) / nm;
    const dim3 grid_rows((unsigned)((nw + HL_W - 1) / HL_W), (unsigned)(((ipk > ni ? ipk : ni) + HL_IP - 1) / HL_IP),
                         (unsigned)nslots);
    if (hb) {
        HArgs ab = a;
        ab.act = act_b;
        if (rows) {
            hipLaunchKernelGGL(h_build_lp_kernel<uint16_t>, grid_rows, dim3(256), 2 * HL_W * nm * HL_IP, ctx->stream,
                               ab, hb, kp);
        } else {
            const dim3 grid((unsigned)((nw + 63) / 64), (unsigned)((kp + 63) / 64), (unsigned)nslots);
            hipLaunchKernelGGL(h_build_lp_run_kernel<uint16_t>, grid, dim3(512), 0, ctx->stream, ab, hb, kp);
        }
        EPSX_HIP(ctx, hipGetLastError());
    }
    if (hf) {
        HArgs af = a;
        af.act = act_f;
        if (rows) {
            hipLaunchKernelGGL(h_build_lp_kernel<float>, grid_rows, dim3(256), 4 * HL_W * nm * HL_IP, ctx->stream, af,
                               hf, kp);
        } else {
            const dim3 grid((unsigned)((nw + 63) / 64), (unsigned)((kp + 31) / 32), (unsigned)nslots);
            hipLaunchKernelGGL(h_build_lp_run_kernel<float>, grid, dim3(512), 0, ctx->stream, af, hf, kp);
        }
        EPSX_HIP(ctx, hipGetLastError());
    }
    return EPSX_OK;
}

int einstein_h_launch(epsx_ctx *ctx, int64_t nw, int64_t nm, int64_t ni, const double *ws, const int32_t *hk, const double *ht,
                      const double *wnu, const double *S, const double *eta, const int32_t *tidx, int nslots,
                      double w_acoustic, const double *occ, int64_t p0, int64_t np, const double *ef, const double *wf,
                      const int32_t *fi, double *hout, int64_t ldh, const int32_t *act)
{
    HArgs a{};
    a.nw = nw; a.nm = nm; a.ni = ni; a.ws = ws; a.hk = hk; a.ht = ht; a.wnu = wnu; a.S = S; a.eta = eta; a.ldh = ldh;
    a.tidx = tidx; a.w_acoustic = w_acoustic; a.occ = const_cast<double *>(occ); a.hout = hout; a.act = act;
    const int64_t threads = ldh * nm * ((np + EH_R - 1) / EH_R);
    hipLaunchKernelGGL(einstein_h_kernel, dim3(xcd_grid(threads), (unsigned)nslots), dim3(256), 0, ctx->stream, a, p0,
                       np, ef, wf, fi);
    EPSX_HIP(ctx, hipGetLastError());
    return EPSX_OK;
}

int acoustic_finite_check(epsx_ctx *ctx, int64_t nm, int64_t n, const double *g2p, const uint8_t *skip, bool *ok)
{
    Tmp flag;
    const int32_t zero = 0;
    EPSX_TRY(tmp_upload(ctx, flag, &zero, 4));
    hipLaunchKernelGGL(acoustic_finite_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, nm, n, g2p,
                       skip, flag.ptr<int32_t>());
    EPSX_HIP(ctx, hipGetLastError());
    int32_t bad = 0;
    EPSX_HIP(ctx, hipMemcpyAsync(&bad, flag.b.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *ok = bad == 0;
    return EPSX_OK;
}

int gt_lp_launch(epsx_ctx *ctx, int precision, int64_t K, int64_t ldg, int64_t kp, int64_t ni, const double *Gt, void *gl)
{
    const dim3 grid((unsigned)((kp * ni + 255) / 256)), blk(256);
    if (precision == EPSX_PREC_BF16)
        hipLaunchKernelGGL(gt_lp_kernel<uint16_t>, grid, blk, 0, ctx->stream, K, ldg, kp, ni, Gt, static_cast<uint16_t *>(gl));
    else
        hipLaunchKernelGGL(gt_lp_kernel<float>, grid, blk, 0, ctx->stream, K, ldg, kp, ni, Gt, static_cast<float *>(gl));
    EPSX_HIP(ctx, hipGetLastError());
    return EPSX_OK;
}

}  // namespace epsx
