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
#include <cmath>
#include <vector>

#include "epsx_common.h"

namespace epsx {
namespace {

constexpr int SR_THREADS = 256;
constexpr int SR_CAP = 4096;   // doubles of g2 staged per block (whole stars)

// One block per (chunk of whole stars, outer state i): the chunk's slab of g2p[:, p, i] is read
// coalesced into LDS (times w_f), then one thread per (ν, i') sums its star in ascending p.
__global__ __launch_bounds__(SR_THREADS) void star_reduce_kernel(
    int64_t nm, int64_t ni, int64_t nfs, const double *__restrict__ g2p, const double *__restrict__ wf,
    const int64_t *__restrict__ star_ptr, const int32_t *__restrict__ chunk_star, const uint8_t *__restrict__ skip_mode,
    double *__restrict__ Gt)
{
    __shared__ double buf[SR_CAP];
    const int64_t c = blockIdx.x, i = blockIdx.y;
    const int64_t s0 = chunk_star[c], s1 = chunk_star[c + 1];
    const int64_t p0 = star_ptr[s0], p1 = star_ptr[s1];
    const int64_t n = (p1 - p0) * nm;
    const double *src = g2p + nm * (p0 + nfs * i);
    const bool staged = n <= SR_CAP;
    if (staged) {
        for (int64_t e = threadIdx.x; e < n; e += SR_THREADS) {
            const int64_t nu = e % nm;
            buf[e] = skip_mode[nu] ? 0.0 : wf[p0 + e / nm] * src[e];
        }
        __syncthreads();
    }
    const int64_t nout = (s1 - s0) * nm;
    for (int64_t o = threadIdx.x; o < nout; o += SR_THREADS) {
        const int64_t nu = o % nm, ip = s0 + o / nm;
        double acc = 0.0;
        if (!skip_mode[nu]) {
            const int64_t a = star_ptr[ip], b = star_ptr[ip + 1];
            for (int64_t p = a; p < b; p++) {
                const int64_t e = (p - p0) * nm + nu;
                acc += staged ? buf[e] : wf[p] * src[e];
            }
        }
        Gt[(nu + nm * ip) + nm * ni * i] = acc;
    }
}

// e^y = M 2^k (Cody-Waite), as in the direct kernel
__device__ inline void split_exp(double y, double &M, int &k)
{
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    const double log2e = 1.44269504088896338700e+00;
    if (!(y == y)) { M = y; k = 0; return; }
    double yk = rint(y * log2e);
    yk = yk > 1.0e6 ? 1.0e6 : (yk < -1.0e6 ? -1.0e6 : yk);
    const double r = (y - yk * ln2_hi) - yk * ln2_lo;
    M = exp(r);
    k = (int)yk;
}

struct HArgs {
    int64_t nw, nm, ni, span;
    const double *ws;
    const int32_t *off;
    const double *rgap;
    const int32_t *ltab;
    double h;                 // dω_unit
    const double *wnu;        // [nm]
    const double *e_i;        // [ni]
    const double *S;          // complex [nw, ni, nT_all]
    const double *T, *mu, *eta;  // per temperature index
    const int32_t *tidx;      // slot -> temperature index
    double w_acoustic;
    double *hout;             // [nw, nm, ni] per slot
    const int32_t *act;       // convergence mask per temperature index, or null
};

__device__ inline double2 lookup(const double2 *col, int64_t nw, int64_t span, const int32_t *off,
                                 const double *rgap, const int32_t *ltab, int pos, double frac, double eta)
{
    double2 v;
    if (pos < 0) v = col[0];
    else if (pos >= span) v = col[nw - 1];
    else {
        const int k = ltab[pos];
        const double t = ((double)(pos - off[k]) + frac) * rgap[k];
        const double2 lo = col[k], hi = col[k + 1];
        const double lim = lo.y < -eta ? lo.y : -eta, him = hi.y < -eta ? hi.y : -eta;
        return make_double2((1 - t) * lo.x + t * hi.x, (1 - t) * lim + t * him);
    }
    return make_double2(v.x, v.y < -eta ? v.y : -eta);
}

// h[ω, ν, i', slot]: one thread per (ω, ν, i'); Σin is the clamped Σ, read at ω ± ω_ν through
// the integer offsets of the grid.
__global__ __launch_bounds__(256) void h_build_kernel(HArgs a)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int z = blockIdx.y;
    const int64_t nw = a.nw, total = nw * a.nm * a.ni;
    if (q >= total) return;
    const int64_t w = q % nw, nu = (q / nw) % a.nm, ip = q / (nw * a.nm);
    const int tcol = a.tidx[z];
    if (a.act && !a.act[tcol]) return;
    double *out = a.hout + total * (int64_t)z;
    const double wnu = a.wnu[nu];
    if (wnu < a.w_acoustic) { out[q] = 0.0; return; }
    const double T = a.T[tcol], mu = a.mu[tcol], eta = a.eta ? a.eta[tcol] : -INFINITY;
    const bool step = !(T > 1.4901161193847656e-08);
    const double2 *col = reinterpret_cast<const double2 *>(a.S) + nw * a.ni * (int64_t)tcol + nw * ip;
    const double x0 = a.ws[w];
    const int o = a.off[w];
    const double xa = wnu / a.h, xe = -wnu / a.h;
    const double fla = floor(xa), fle = floor(xe);
    const double2 Sa = lookup(col, nw, a.span, a.off, a.rgap, a.ltab, o + (int)fla, xa - fla, eta);
    const double2 Se = lookup(col, nw, a.span, a.off, a.rgap, a.ltab, o + (int)fle, xe - fle, eta);
    const double ei = a.e_i[ip];
    const double zar = x0 + wnu - ei - Sa.x, zai = -Sa.y;
    const double zer = x0 - wnu - ei - Se.x, zei = -Se.y;
    double nq, Ea, Ee;
    if (!step) {
        nq = 1.0 / expm1(wnu / T);
        double Mw, Mq;
        int kw, kq;
        split_exp((x0 - mu) / T, Mw, kw);
        split_exp(wnu / T, Mq, kq);
        int ka = kw + kq, ke = kw - kq;
        ka = ka > 900 ? 900 : (ka < -900 ? -900 : ka);
        ke = ke > 900 ? 900 : (ke < -900 ? -900 : ke);
        Ea = ldexp(Mw * Mq, ka);
        Ee = ldexp(Mw / Mq, ke);
    } else {
        nq = wnu > 0 ? 0.0 : (wnu == 0 ? -0.5 : -1.0);
        const double ya = x0 + wnu - mu, ye = x0 - wnu - mu;
        Ea = ya > 0 ? 1e270 : (ya < 0 ? 0.0 : 1.0);
        Ee = ye > 0 ? 1e270 : (ye < 0 ? 0.0 : 1.0);
    }
    const double ta = -zai * (nq * (1 + Ea) + 1) / ((zar * zar + zai * zai) * (1 + Ea));
    const double te = -zei * (nq * (1 + Ee) + Ee) / ((zer * zer + zei * zei) * (1 + Ee));
    out[q] = ta + te;
}

}  // namespace

// Chunks of whole stars with at most SR_CAP staged doubles each (a larger single star is a chunk
// of its own, read from global memory).
void star_chunks(int64_t nm, const std::vector<int64_t> &star_ptr, std::vector<int32_t> &chunk)
{
    const int64_t ns = (int64_t)star_ptr.size() - 1;
    chunk.clear();
    chunk.push_back(0);
    int64_t s = 0;
    while (s < ns) {
        int64_t e = s + 1;
        while (e < ns && (star_ptr[e + 1] - star_ptr[s]) * nm <= SR_CAP) e++;
        chunk.push_back((int32_t)e);
        s = e;
    }
}

int star_reduce_launch(epsx_ctx *ctx, int64_t nm, int64_t ni, int64_t nfs, const double *g2p, const double *wf,
                       const int64_t *d_star_ptr, const int32_t *d_chunks, int64_t nchunks, const uint8_t *d_skip,
                       double *Gt)
{
    dim3 grid((unsigned)nchunks, (unsigned)ni);
    hipLaunchKernelGGL(star_reduce_kernel, grid, dim3(SR_THREADS), 0, ctx->stream, nm, ni, nfs, g2p, wf,
                       d_star_ptr, d_chunks, d_skip, Gt);
    EPSX_HIP(ctx, hipGetLastError());
    return EPSX_OK;
}

int h_build_launch(epsx_ctx *ctx, const FmGrid &g, int64_t nm, int64_t ni, const double *wnu, const double *e_i,
                   const double *S, const double *T, const double *mu, const double *eta, const int32_t *tidx,
                   int nslots, double w_acoustic, double *hout, const int32_t *act)
{
    HArgs a;
    a.act = act;
    a.nw = g.nw; a.nm = nm; a.ni = ni; a.span = g.span;
    a.ws = g.ws; a.off = g.off; a.rgap = g.rgap; a.ltab = g.ltab; a.h = g.dw_unit;
    a.wnu = wnu; a.e_i = e_i; a.S = S; a.T = T; a.mu = mu; a.eta = eta; a.tidx = tidx;
    a.w_acoustic = w_acoustic; a.hout = hout;
    const int64_t total = g.nw * nm * ni;
    dim3 grid((unsigned)((total + 255) / 256), (unsigned)nslots);
    hipLaunchKernelGGL(h_build_kernel, grid, dim3(256), 0, ctx->stream, a);
    EPSX_HIP(ctx, hipGetLastError());
    return EPSX_OK;
}

}  // namespace epsx
