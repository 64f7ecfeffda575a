// The direct Fan-Migdal Im Σ kernel (src/scgd0_kernel.jl:27-86) for gfx950.
//
// ImΣ[ω, i] += Σ_{f, ν} w_f g2[ν,i,f] { Im G(ω+ωq) (n(ωq) + f(ω+ωq-μ))
//                                      + Im G(ω-ωq) (n(ωq) + 1 - f(ω-ωq-μ)) },
// G(x) = 1/(x - ε_f - Σin_{i_f}(x)), Σin read by flat linear interpolation on ωs.
//
// Work decomposition.  A block owns one outer state i, one temperature and one segment of the
// inner states, which are sorted by their IBZ representative i_f ("stars"), so that the
// clamped Σin column of i_f is staged in LDS once per star and read from LDS for every term.
// Lanes run over ω (each thread R nodes, ω = tid + 256 r), so the sum over (f, ν) is a
// per-lane serial accumulation in a fixed order; the segments of one (i, T) are summed by a
// second, fixed-order kernel, so results are bitwise reproducible run to run.
//
// Per (ν, f) pair, everything that does not depend on ω is computed once per block into an
// LDS record: c = w_f g2 (pairs with ωq < ω_acoustic are flagged and skipped before g2 is
// used, as the reference tests it first), n_B(ωq), ωq/h split into an integer shift and a
// fraction, and e^{±ωq/T} as mantissa and binary exponent.  Per ω the thread keeps
// e^{(ω-μ)/T} the same way, so each Fermi factor costs a multiply and an ldexp instead of an
// exp: e^{(ω±ωq-μ)/T} = e^{(ω-μ)/T} e^{±ωq/T} (exponents clamped at ±900, far past the point
// where f is 0 or 1 to double precision).  The cell of ω ± ωq is found from the grid's
// integer offsets: ω_j ± ωq sits at offset off_j + m + r (m integer, 0 <= r < 1), so the cell
// is a table read at the integer m + off_j and the interpolation weight is
// (off_j + m - off_k + r) / (off_{k+1} - off_k).
#include <cfloat>
#include <cmath>
#include <vector>

#include "epsx_common.h"
#include "epsx_device.h"

namespace epsx {
namespace {

constexpr int FM_THREADS = 256;
constexpr int FM_CP = 64;            // (ν, f) pair records per LDS batch
constexpr int LTAB_LDS_MAX = 8192;   // cell table entries kept in LDS

struct PairRec {
    double c;      // w_f * g2 (0 when skipped)
    double wq;     // phonon frequency
    double nq;     // n_B(ωq, T)
    double ef;     // ε_f
    double fa;     // fraction of +ωq / h
    double fe;     // fraction of -ωq / h
    double Ma;     // mantissa of e^{+ωq/T}
    double Me;     // mantissa of e^{-ωq/T}
    int ia, ie;    // integer part of ±ωq / h
    int ka, ke;    // binary exponents of e^{±ωq/T}
    int skip;
    int pad;
};

__device__ inline double occ_boson_d(double e, double T, bool step)
{
    if (!step) return e == 0.0 ? -0.5 : 1.0 / expm1(e / T);
    if (e > 0) return 0.0;
    if (e == 0) return -0.5;
    return -1.0;
}

// e^{(ω±ωq-μ)/T} from the two split factors, the exponent clamped.
template <typename FT>
__device__ inline FT fermi_E(double Mw, int kw, double Mq, int kq)
{
    // 2^±400 (FP64) and 2^±50 (FP32) are far past where f is 0 or 1 to the working precision, and
    // keep the product of the two denominators below the overflow threshold
    constexpr int KMAX = sizeof(FT) == 4 ? 50 : 400;
    int k = kw + kq;
    k = k > KMAX ? KMAX : (k < -KMAX ? -KMAX : k);
    return (FT)ldexp(Mw * Mq, k);
}

template <int R, typename Acc>
struct Accum;

template <int R>
struct Accum<R, double> {
    double s[R];
    __device__ void zero() {
#pragma unroll
        for (int r = 0; r < R; r++) s[r] = 0.0;
    }
    __device__ void add(int r, double v) { s[r] += v; }
    __device__ double get(int r) const { return s[r]; }
};

// FP32 terms with a compensated (Kahan-Babuska / Neumaier) sum over f
template <int R>
struct Accum<R, float> {
    float s[R], c[R];
    __device__ void zero() {
#pragma unroll
        for (int r = 0; r < R; r++) { s[r] = 0.f; c[r] = 0.f; }
    }
    __device__ void add(int r, float v) {
        float t = s[r] + v;
        if (fabsf(s[r]) >= fabsf(v)) c[r] += (s[r] - t) + v;
        else c[r] += (v - t) + s[r];
        s[r] = t;
    }
    __device__ double get(int r) const { return (double)s[r] + (double)c[r]; }
};

// Flat-extrapolated linear interpolation of the staged column at integer grid position pos plus
// frac, without branches: the cell of the clamped position, and t forced to 0 below the grid and
// 1 above it (so the end value comes back exactly).
__device__ inline double2 lookup_ff(const double2 *Sc, const int32_t *ltab, const int32_t *off, const double *rgap,
                                    int span, int pos, double frac)
{
    const int pc = pos < 0 ? 0 : (pos >= span ? span - 1 : pos);
    const int k = ltab[pc];
    double t = ((double)(pos - off[k]) + frac) * rgap[k];
    t = pos < 0 ? 0.0 : (pos >= span ? 1.0 : t);
    const double2 lo = Sc[k], hi = Sc[k + 1];
    return make_double2((1 - t) * lo.x + t * hi.x, (1 - t) * lo.y + t * hi.y);
}

template <int R, typename FT, bool STEP>
__device__ __forceinline__ void fm_direct_body(FmGrid g, FmDirectArgs a,
                                                               FmTemps tp, int64_t S,
                                                               int64_t seg_len, int64_t nwt,
                                                               const int64_t *__restrict__ star_ptr,
                                                               double *__restrict__ part)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int tid = threadIdx.x;
    const int64_t nw = g.nw;
    // block -> (i, segment, ω tile), temperature slot in blockIdx.y
    int64_t bx = blockIdx.x;
    const int64_t wt = bx % nwt;
    bx /= nwt;
    const int64_t s = bx % S;
    const int64_t i = bx / S;
    const int tz = blockIdx.y;
    const int64_t tcol = tp.tidx[tz];
    if (tp.act && !tp.act[tcol]) return;
    const double T = tp.T[tcol], mu = tp.mu[tcol];
    const double eta = tp.eta_min ? tp.eta_min[tcol] : -INFINITY;
    const bool step = STEP;   // T <= sqrt(eps): the step occupations
    const double h = g.dw_unit;

    // LDS carve-up
    double2 *Sc = reinterpret_cast<double2 *>(lds);
    double *rgap = reinterpret_cast<double *>(Sc + nw);
    int32_t *off = reinterpret_cast<int32_t *>(rgap + nw);
    PairRec *rec = reinterpret_cast<PairRec *>(lds + (((size_t)nw * 16 + (size_t)nw * 8 +
                                                        (size_t)nw * 4 + 15) & ~(size_t)15));
    int32_t *ltab = reinterpret_cast<int32_t *>(rec + FM_CP);
    const bool ltab_lds = g.span <= LTAB_LDS_MAX;
    for (int64_t k = tid; k < nw; k += FM_THREADS) {
        off[k] = g.off[k];
        rgap[k] = g.rgap[k];
    }
    if (ltab_lds)
        for (int64_t k = tid; k < g.span; k += FM_THREADS) ltab[k] = g.ltab[k];
    const int32_t *ltab_g = ltab_lds ? ltab : g.ltab;

    // this thread's frequencies
    double wv[R], Mw[R];
    int ow[R], kw[R];
    bool valid[R];
    const int64_t w0 = wt * (int64_t)(FM_THREADS * R);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t jw = w0 + tid + (int64_t)FM_THREADS * r;
        valid[r] = jw < nw;
        const int64_t jc = valid[r] ? jw : nw - 1;
        wv[r] = g.ws[jc];
        ow[r] = g.off[jc];
        if (!step) split_exp((wv[r] - mu) / T, Mw[r], kw[r]);
        else { Mw[r] = 1.0; kw[r] = 0; }
    }
    Accum<R, FT> acc;
    acc.zero();

    const int64_t fb = s * seg_len, fe_ = fb + seg_len < a.nfs ? fb + seg_len : a.nfs;
    const double *Scol_base = a.S + 2 * nw * a.ncol * tcol;
    const int64_t span = g.span;
    const int64_t nm = a.nm;

    int64_t p = fb;
    while (p < fe_) {
        const int32_t col = a.fi[p];
        const int64_t run_end = star_ptr[col + 1] < fe_ ? star_ptr[col + 1] : fe_;
        __syncthreads();  // previous star's records and column are consumed
        const double *Scol = Scol_base + 2 * nw * col;
        for (int64_t k = tid; k < nw; k += FM_THREADS) {
            const double re = Scol[2 * k], im = Scol[2 * k + 1];
            Sc[k] = make_double2(re, fmin(im, -eta));
        }
        const int64_t npairs_run = (run_end - p) * nm;
        for (int64_t q0 = 0; q0 < npairs_run; q0 += FM_CP) {
            const int npairs = (int)(npairs_run - q0 < FM_CP ? npairs_run - q0 : FM_CP);
            if (q0 != 0) __syncthreads();
            for (int q = tid; q < npairs; q += FM_THREADS) {
                const int64_t qq = q0 + q;
                const int64_t nu = qq % nm, fp = p + qq / nm;
                const int64_t fsrc = a.fsrc[fp];
                double wq;
                if (a.wq_pair) {
                    wq = a.wq_pair[nu + nm * (i + a.ni * fsrc)];
                } else {
                    const int64_t qi = a.iq_kk[(a.iks_i[i] - 1) + a.nki * (int64_t)(a.iks_f[fsrc] - 1)] - 1;
                    wq = a.wph[nu + nm * qi];
                }
                PairRec R_;
                R_.skip = wq < a.w_acoustic;
                R_.wq = wq;
                R_.ef = a.ef[fp];
                if (!R_.skip) {
                    const double g2v = a.g2p32 ? (double)a.g2p32[nu + nm * (fp + a.nfs * i)]
                                               : a.g2p[nu + nm * (fp + a.nfs * i)];
                    R_.c = a.wf[fp] * g2v;
                    R_.nq = occ_boson_d(wq, T, step);
                    const double xa = wq / h, xe = -wq / h;
                    const double fla = floor(xa), fle = floor(xe);
                    R_.ia = (int)fla;
                    R_.fa = xa - fla;
                    R_.ie = (int)fle;
                    R_.fe = xe - fle;
                    if (!step) {
                        split_exp(wq / T, R_.Ma, R_.ka);
                        split_exp(-wq / T, R_.Me, R_.ke);
                    } else {
                        R_.Ma = R_.Me = 1.0;
                        R_.ka = R_.ke = 0;
                    }
                } else {
                    R_.c = 0; R_.nq = 0; R_.ia = R_.ie = 0; R_.fa = R_.fe = 0;
                    R_.Ma = R_.Me = 1.0; R_.ka = R_.ke = 0;
                }
                R_.pad = 0;
                rec[q] = R_;
            }
            __syncthreads();
            for (int q = 0; q < npairs; q++) {
                const PairRec P = rec[q];
                if (P.skip) continue;
                const FT c = (FT)P.c, nq = (FT)P.nq;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    // absorption at ω + ωq, emission at ω - ωq
                    const double2 Sa = lookup_ff(Sc, ltab_g, off, rgap, (int)span, ow[r] + P.ia, P.fa);
                    const double2 Se = lookup_ff(Sc, ltab_g, off, rgap, (int)span, ow[r] + P.ie, P.fe);
                    const FT zar = (FT)(wv[r] + P.wq - P.ef - Sa.x), zai = (FT)(-Sa.y);
                    const FT zer = (FT)(wv[r] - P.wq - P.ef - Se.x), zei = (FT)(-Se.y);
                    FT Ea, Ee;
                    if (!STEP) {
                        Ea = fermi_E<FT>(Mw[r], kw[r], P.Ma, P.ka);
                        Ee = fermi_E<FT>(Mw[r], kw[r], P.Me, P.ke);
                    } else {
                        // T below sqrt(eps): the step occupation of occ_fermion
                        const FT BIG = sizeof(FT) == 4 ? (FT)1e15f : (FT)1e100;
                        const double ya = wv[r] + P.wq - mu, ye = wv[r] - P.wq - mu;
                        Ea = ya > 0 ? BIG : (ya < 0 ? (FT)0 : (FT)1);
                        Ee = ye > 0 ? BIG : (ye < 0 ? (FT)0 : (FT)1);
                    }
                    const FT one = (FT)1;
                    // Im G(ω+ωq)(n + f) + Im G(ω-ωq)(n + 1 - f) with Im(1/z) = -Im z / |z|^2 and
                    // f = 1 / (1 + E): one division for both terms
                    const FT Da = (zar * zar + zai * zai) * (one + Ea);
                    const FT De = (zer * zer + zei * zei) * (one + Ee);
                    const FT Na = -zai * (nq * (one + Ea) + one);
                    const FT Ne = -zei * (nq * (one + Ee) + Ee);
                    // lanes past the grid end compute on the last node and are never written out
                    acc.add(r, c * ((Na * De + Ne * Da) / (Da * De)));
                }
            }
        }
        p = run_end;
    }
    // partial sums of this segment
    double *out = part + (((int64_t)tz * S + s) * a.ni + i) * nw;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int64_t jw = w0 + tid + (int64_t)FM_THREADS * r;
        if (jw < nw) out[jw] = acc.get(r);
    }
}

// The block's temperature picks the occupation form once (T <= sqrt(eps): step functions).
template <int R, typename FT>
__global__ __launch_bounds__(FM_THREADS) void fm_direct_kernel(FmGrid g, FmDirectArgs a, FmTemps tp, int64_t S,
                                                               int64_t seg_len, int64_t nwt,
                                                               const int64_t *__restrict__ star_ptr,
                                                               double *__restrict__ part)
{
    const int64_t tcol = tp.tidx[blockIdx.y];
    if (tp.act && !tp.act[tcol]) return;
    if (tp.T[tcol] > 1.4901161193847656e-08)
        fm_direct_body<R, FT, false>(g, a, tp, S, seg_len, nwt, star_ptr, part);
    else
        fm_direct_body<R, FT, true>(g, a, tp, S, seg_len, nwt, star_ptr, part);
}

// ImS[ω, i, tidx[tz]] (+)= Σ_s part[tz, s, i, ω], s ascending
__global__ void fm_reduce_kernel(int64_t nw, int64_t ni, int64_t S, const int32_t *tidx,
                                 const double *__restrict__ part, double *__restrict__ ImS,
                                 int accumulate, const int32_t *act)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int tz = blockIdx.y;
    if (q >= nw * ni) return;
    if (act && !act[tidx[tz]]) return;
    double acc = 0.0;
    for (int64_t s = 0; s < S; s++) acc += part[((int64_t)tz * S + s) * nw * ni + q];
    double *o = ImS + nw * ni * (int64_t)tidx[tz] + q;
    *o = accumulate ? *o + acc : acc;
}

template <int R>
int launch_R(epsx_ctx *ctx, const FmGrid &g, const FmDirectArgs &a, const FmTemps &tp, int64_t S,
             int64_t seg_len, int64_t nwt, const int64_t *star_ptr, double *part, int precision,
             size_t lds)
{
    dim3 grid((unsigned)(a.ni * S * nwt), (unsigned)tp.nT);
    // The LDS limit is raised outside graph capture; a captured launch repeats an eager one.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    EPSX_HIP(ctx, hipStreamIsCapturing(ctx->stream, &cap));
    const bool set_attr = cap == hipStreamCaptureStatusNone;
    if (precision == EPSX_PREC_FP32) {
        auto k = fm_direct_kernel<R, float>;
        if (set_attr)
            EPSX_HIP(ctx, hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k, grid, dim3(FM_THREADS), lds, ctx->stream, g, a, tp, S, seg_len, nwt,
                           star_ptr, part);
    } else {
        auto k = fm_direct_kernel<R, double>;
        if (set_attr)
            EPSX_HIP(ctx, hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k, grid, dim3(FM_THREADS), lds, ctx->stream, g, a, tp, S, seg_len, nwt,
                           star_ptr, part);
    }
    EPSX_HIP(ctx, hipGetLastError());
    return EPSX_OK;
}

}  // namespace

// Segments per (i, T) so that a launch holds enough blocks for the 256 CUs, unless the
// caller fixed the inner states per block (f_tile).
static void fm_split(int64_t nfs, int64_t ni, int64_t nT, int64_t nwt, int64_t f_tile,
                     int64_t &S, int64_t &seg_len)
{
    if (f_tile > 0) {
        seg_len = f_tile;
    } else {
        const int64_t blocks = ni * nT * nwt;
        int64_t want = (2048 + blocks - 1) / blocks;
        if (want < 1) want = 1;
        seg_len = (nfs + want - 1) / want;
        if (seg_len < 1) seg_len = 1;
    }
    S = (nfs + seg_len - 1) / seg_len;
    if (S < 1) S = 1;
}

size_t fm_part_elems(int64_t nfs, int64_t ni, int64_t nw, int64_t nT, int64_t f_tile)
{
    const int64_t nwt = (nw + FM_THREADS * 8 - 1) / (FM_THREADS * 8);
    int64_t S, seg;
    fm_split(nfs, ni, nT, nwt, f_tile, S, seg);
    return (size_t)(S * ni * nw * nT);
}

int fm_direct_launch_sp(epsx_ctx *ctx, const FmGrid &g, const FmDirectArgs &a, const FmTemps &tp,
                        const int64_t *d_star_ptr, int64_t f_tile, int precision, double *d_part,
                        int64_t part_cap, double *d_ImS, bool accumulate)
{
    if (a.nfs == 0 || a.ni == 0 || tp.nT == 0) return EPSX_OK;
    int R = (int)((g.nw + FM_THREADS - 1) / FM_THREADS);
    int64_t nwt = 1;
    if (R > 8) {
        nwt = (g.nw + FM_THREADS * 8 - 1) / (FM_THREADS * 8);
        R = 8;
    }
    int64_t S, seg_len;
    fm_split(a.nfs, a.ni, tp.nT, nwt, f_tile, S, seg_len);
    if ((int64_t)(S * a.ni * g.nw * tp.nT) > part_cap)
        return set_error(ctx, EPSX_ERR_STATE, "Fan-Migdal partial buffer too small");
    size_t lds = (((size_t)g.nw * 28 + 15) & ~(size_t)15) + sizeof(PairRec) * FM_CP +
                 (g.span <= LTAB_LDS_MAX ? sizeof(int32_t) * (size_t)g.span : 0);
    if (lds > 160 * 1024)
        return set_error(ctx, EPSX_ERR_ARG,
                         "the Fan-Migdal kernel needs %zu bytes of LDS for %lld frequencies and %lld "
                         "modes; at most 163840 exist", lds, (long long)g.nw, (long long)a.nm);
    int r;
    switch (R) {
    case 1: r = launch_R<1>(ctx, g, a, tp, S, seg_len, nwt, d_star_ptr, d_part, precision, lds); break;
    case 2: r = launch_R<2>(ctx, g, a, tp, S, seg_len, nwt, d_star_ptr, d_part, precision, lds); break;
    case 3: r = launch_R<3>(ctx, g, a, tp, S, seg_len, nwt, d_star_ptr, d_part, precision, lds); break;
    case 4: r = launch_R<4>(ctx, g, a, tp, S, seg_len, nwt, d_star_ptr, d_part, precision, lds); break;
    case 5: r = launch_R<5>(ctx, g, a, tp, S, seg_len, nwt, d_star_ptr, d_part, precision, lds); break;
    case 6: r = launch_R<6>(ctx, g, a, tp, S, seg_len, nwt, d_star_ptr, d_part, precision, lds); break;
    case 7: r = launch_R<7>(ctx, g, a, tp, S, seg_len, nwt, d_star_ptr, d_part, precision, lds); break;
    default: r = launch_R<8>(ctx, g, a, tp, S, seg_len, nwt, d_star_ptr, d_part, precision, lds); break;
    }
    if (r != EPSX_OK) return r;
    dim3 grid((unsigned)((g.nw * a.ni + 255) / 256), (unsigned)tp.nT);
    hipLaunchKernelGGL(fm_reduce_kernel, grid, dim3(256), 0, ctx->stream, g.nw, a.ni, S, tp.tidx,
                       d_part, d_ImS, accumulate ? 1 : 0, tp.act);
    EPSX_HIP(ctx, hipGetLastError());
    return EPSX_OK;
}

// Build the Fan-Migdal grid data of a FrequencyGrid on the host: relative offsets, the
// reciprocal cell widths in offset units and the cell of every relative offset.
int fm_grid_host(int64_t nw, const int64_t *offsets, std::vector<int32_t> &off,
                 std::vector<double> &rgap, std::vector<int32_t> &ltab, int64_t &span)
{
    off.resize(nw);
    rgap.resize(nw);
    span = offsets[nw - 1] - offsets[0];
    if (span >= (int64_t(1) << 30)) return -1;
    for (int64_t k = 0; k < nw; k++) off[k] = (int32_t)(offsets[k] - offsets[0]);
    for (int64_t k = 0; k + 1 < nw; k++) rgap[k] = 1.0 / (double)(off[k + 1] - off[k]);
    rgap[nw - 1] = 0;
    ltab.resize(span > 0 ? span : 1);
    int64_t k = 0;
    for (int64_t p = 0; p < span; p++) {
        while (k + 1 < nw - 1 && off[k + 1] <= p) k++;
        ltab[p] = (int32_t)k;
    }
    return 0;
}

// Star order of the inner states: stable sort by representative (sortperm(f_to_i)), and the
// start position of every representative's run.
void star_order(int64_t nf, int64_t ncol, const int32_t *f_to_i, std::vector<int32_t> &fsrc,
                std::vector<int64_t> &star_ptr)
{
    star_ptr.assign(ncol + 1, 0);
    for (int64_t f = 0; f < nf; f++) star_ptr[f_to_i[f]]++;  // f_to_i is 1-based
    for (int64_t c = 0; c < ncol; c++) star_ptr[c + 1] += star_ptr[c];
    std::vector<int64_t> fill(star_ptr.begin(), star_ptr.end() - 1);
    fsrc.resize(nf);
    for (int64_t f = 0; f < nf; f++) fsrc[fill[f_to_i[f] - 1]++] = (int32_t)f;
}

}  // namespace epsx

extern "C" int epsx_add_fan_migdal_imsigma(epsx_ctx *ctx, int64_t nm, int64_t ni, int64_t nf,
                                           const double *g2, const double *wq, const double *e_f,
                                           const int32_t *f_to_i, const double *w_f, int64_t nw,
                                           double w_origin, double dw_unit, const int64_t *offsets,
                                           int64_t ncol, const double *Sin, double mu, double T,
                                           double w_acoustic, int32_t precision, double *ImS)
{
    using namespace epsx;
    if (!ctx) return EPSX_ERR_ARG;
    if (nm < 1 || ni < 0 || nf < 0 || nw < 2 || ncol < 1 || !offsets || !Sin || !ImS)
        return set_error(ctx, EPSX_ERR_ARG, "add_fan_migdal_imsigma!: bad sizes or null arrays");
    if (!(T >= 0)) return set_error(ctx, EPSX_ERR_ARG, "Temperature must be positive");
    if (precision != EPSX_PREC_FP64 && precision != EPSX_PREC_FP32)
        return set_error(ctx, EPSX_ERR_ARG, "precision must be FP64 or FP32 for the direct kernel");
    for (int64_t f = 0; f < nf; f++)
        if (f_to_i[f] < 1 || f_to_i[f] > ncol)
            return set_error(ctx, EPSX_ERR_ARG, "f_to_i[%lld] = %d is outside 1:%lld", (long long)f + 1,
                             f_to_i[f], (long long)ncol);
    if (ni == 0 || nf == 0) return EPSX_OK;
    EPSX_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<int32_t> off, ltab, fsrc, fi(nf);
    std::vector<double> rgap, ws(nw), ef(nf), wf(nf), g2p((size_t)nm * nf * ni);
    std::vector<int64_t> sptr;
    int64_t span;
    if (fm_grid_host(nw, offsets, off, rgap, ltab, span))
        return set_error(ctx, EPSX_ERR_ARG, "grid span too large");
    grid_nodes(nw, w_origin, dw_unit, offsets, ws.data());
    star_order(nf, ncol, f_to_i, fsrc, sptr);
    for (int64_t p = 0; p < nf; p++) {
        const int64_t f = fsrc[p];
        ef[p] = e_f[f];
        wf[p] = w_f[f];
        fi[p] = f_to_i[f] - 1;
        for (int64_t i = 0; i < ni; i++)
            for (int64_t nu = 0; nu < nm; nu++)
                g2p[nu + nm * (p + nf * i)] = g2[nu + nm * (i + ni * f)];
    }
    std::vector<float> g2p32;
    if (precision == EPSX_PREC_FP32) {
        g2p32.resize(g2p.size());
        for (size_t q = 0; q < g2p.size(); q++) g2p32[q] = (float)g2p[q];
    }
    Tmp d_g2p, d_g2p32, d_wq, d_ef, d_wf, d_fi, d_fsrc, d_ws, d_off, d_rgap, d_ltab, d_S, d_sp,
        d_T, d_mu, d_tidx, d_part, d_ImS;
    EPSX_TRY(tmp_upload(ctx, d_g2p, g2p.data(), sizeof(double) * g2p.size()));
    if (precision == EPSX_PREC_FP32)
        EPSX_TRY(tmp_upload(ctx, d_g2p32, g2p32.data(), sizeof(float) * g2p32.size()));
    EPSX_TRY(tmp_upload(ctx, d_wq, wq, sizeof(double) * nm * ni * nf));
    EPSX_TRY(tmp_upload(ctx, d_ef, ef.data(), sizeof(double) * nf));
    EPSX_TRY(tmp_upload(ctx, d_wf, wf.data(), sizeof(double) * nf));
    EPSX_TRY(tmp_upload(ctx, d_fi, fi.data(), sizeof(int32_t) * nf));
    EPSX_TRY(tmp_upload(ctx, d_fsrc, fsrc.data(), sizeof(int32_t) * nf));
    EPSX_TRY(tmp_upload(ctx, d_ws, ws.data(), sizeof(double) * nw));
    EPSX_TRY(tmp_upload(ctx, d_off, off.data(), sizeof(int32_t) * nw));
    EPSX_TRY(tmp_upload(ctx, d_rgap, rgap.data(), sizeof(double) * nw));
    EPSX_TRY(tmp_upload(ctx, d_ltab, ltab.data(), sizeof(int32_t) * ltab.size()));
    EPSX_TRY(tmp_upload(ctx, d_S, Sin, sizeof(double) * 2 * nw * ncol));
    EPSX_TRY(tmp_upload(ctx, d_sp, sptr.data(), sizeof(int64_t) * sptr.size()));
    const int32_t zero = 0;
    EPSX_TRY(tmp_upload(ctx, d_T, &T, sizeof(double)));
    EPSX_TRY(tmp_upload(ctx, d_mu, &mu, sizeof(double)));
    EPSX_TRY(tmp_upload(ctx, d_tidx, &zero, sizeof(int32_t)));
    EPSX_TRY(tmp_upload(ctx, d_ImS, ImS, sizeof(double) * nw * ni));
    const size_t np = fm_part_elems(nf, ni, nw, 1, 0);
    EPSX_TRY(tmp_alloc(ctx, d_part.b, sizeof(double) * np));

    FmGrid g;
    g.nw = nw; g.span = span; g.ltab_len = (int64_t)ltab.size(); g.dw_unit = dw_unit;
    g.ws = d_ws.ptr<double>(); g.off = d_off.ptr<int32_t>(); g.rgap = d_rgap.ptr<double>();
    g.ltab = d_ltab.ptr<int32_t>();
    FmDirectArgs a{};
    a.nm = nm; a.ni = ni; a.nfs = nf;
    a.g2p = d_g2p.ptr<double>();
    a.g2p32 = precision == EPSX_PREC_FP32 ? d_g2p32.ptr<float>() : nullptr;
    a.ef = d_ef.ptr<double>(); a.wf = d_wf.ptr<double>();
    a.fi = d_fi.ptr<int32_t>(); a.fsrc = d_fsrc.ptr<int32_t>();
    a.wq_pair = d_wq.ptr<double>();
    a.S = d_S.ptr<double>(); a.ncol = ncol; a.w_acoustic = w_acoustic;
    FmTemps tp;
    tp.nT = 1; tp.T = d_T.ptr<double>(); tp.mu = d_mu.ptr<double>(); tp.eta_min = nullptr;
    tp.tidx = d_tidx.ptr<int32_t>();
    EPSX_TRY(fm_direct_launch_sp(ctx, g, a, tp, d_sp.ptr<int64_t>(), 0, precision,
                                 d_part.ptr<double>(), (int64_t)np, d_ImS.ptr<double>(), true));
    EPSX_HIP(ctx, hipMemcpyAsync(ImS, d_ImS.b.p, sizeof(double) * nw * ni, hipMemcpyDeviceToHost, ctx->stream));
    EPSX_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return EPSX_OK;
}
