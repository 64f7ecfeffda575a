/*
 * epsx_oracle_ladder.c -- TEST INFRASTRUCTURE ONLY: the CPU FP64 restatement of the ladder
 * (vertex-corrected Kubo conductivity) of EPSpectral.jl: src/ladder_kernel.jl (all) and the
 * application / solve loop of src/ladder.jl:163-379.  Checker of libepsx's ladder path and the
 * CPU baseline of its benchmark; nothing in the product links it.
 *
 * Layouts (column-major, as the reference): vertex J[ω, c12, cart, c0, i] complex, G[ω, c12, i]
 * complex, v[cart, i], R[a, b, f], B[m, f, i, type], B̃[m, b, ip, a, i, type],
 * X[j, c12, c0, ri, m, b, ip], C[(j, c12, c0, ri), (a, i, type)].
 *
 * Pinned by tests/test_oracle_ladder.py: the contour identities and the 2x2 attach
 * (test/test_ladder_bubble.jl:44-99), the spectral bubble (:101-122), the shift regrouping
 * (test/test_ladder_kernel.jl:56-81), B against the plain loop, bitwise (:101-132), one kernel
 * application against the naive loop over (i, f, ν, j) with exact hat interpolation (:134-172),
 * the vanishing Holstein current correction (test/test_ladder_holstein.jl:5-26), and a solved
 * J = J0 + L(J) (test/test_ladder_pb.jl:48-57).  The GMRES below restates Krylov.jl's gmres
 * (MGS Arnoldi, Givens QR, no restart); its iteration count and residual history are parity
 * unpinned beyond those invariants (Krylov.jl is not available here).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "epsx_oracle.h"

typedef double complex cplx;
static const double PI = 3.141592653589793;

/* Julia's inv(::ComplexF64) for finite normal inputs (Smith's algorithm) */
static cplx jinv(cplx w)
{
    double c = creal(w), d = cimag(w);
    if (fabs(d) <= fabs(c)) {
        double r = d / c, z = d * r + c;
        return CMPLX(1.0 / z, -r / z);
    } else {
        double r = c / d, z = (-c) * r + (-d);
        double q = 1.0 / z, p = -r / z;
        return CMPLX(p, q);
    }
}

/* ------------------------------------------------------------- ladder_kernel.jl:27-40 */
static cplx contour_green(double w, double e, cplx S, double mu, double T, int c12)
{
    cplx GR = jinv(w - e - S);
    cplx GA = conj(GR);
    cplx GK = (GR - GA) * (1 - 2 * orc_occ_fermion(w - mu, T));
    switch (c12) {
    case 1: return (GR + GA + GK) / 2;
    case 2: return (GR - GA + GK) / 2;
    case 3: return (-GR + GA + GK) / 2;
    default: return (-GR - GA + GK) / 2;
    }
}

void orc_contour_green(double w, double e, double Sre, double Sim, double mu, double T, int c12, double *out)
{
    cplx g = contour_green(w, e, CMPLX(Sre, Sim), mu, T, c12);
    out[0] = creal(g);
    out[1] = cimag(g);
}

/* ------------------------------------------------------------- ladder_kernel.jl:51-57 */
void orc_contour_greens(int64_t nw, int64_t ni, const double *ws, const double *e_i, const double *Sd_m,
                        const double *Sd_p, double Omega, double mu, double T, double *Gm_, double *Gp_)
{
    cplx *Gm = (cplx *)Gm_, *Gp = (cplx *)Gp_;
    const cplx *Sm = (const cplx *)Sd_m, *Sp = (const cplx *)Sd_p;
    for (int64_t i = 0; i < ni; i++)
        for (int c = 1; c <= 4; c++)
            for (int64_t w = 0; w < nw; w++) {
                Gm[w + nw * (c - 1 + 4 * i)] = contour_green(ws[w] - Omega / 2, e_i[i], Sm[w + nw * i], mu, T, c);
                Gp[w + nw * (c - 1 + 4 * i)] = contour_green(ws[w] + Omega / 2, e_i[i], Sp[w + nw * i], mu, T, c);
            }
}

#define GIDX(w, c, i) ((w) + nw * ((c) - 1 + 4 * (int64_t)(i)))
#define JIDX(w, c12, cart, c0, i) ((w) + nw * ((c12) - 1 + 4 * ((cart) - 1 + 3 * ((c0) - 1 + 2 * (int64_t)(i)))))

/* ------------------------------------------------------------- ladder_kernel.jl:61-89 */
void orc_attach_greens(int64_t nw, int64_t ni, const double *Gm_, const double *X_, const double *Gp_, double *J_)
{
    const cplx *Gm = (const cplx *)Gm_, *Gp = (const cplx *)Gp_, *X = (const cplx *)X_;
    cplx *J = (cplx *)J_;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < ni; i++)
        for (int c0 = 1; c0 <= 2; c0++)
            for (int cart = 1; cart <= 3; cart++)
                for (int c12 = 1; c12 <= 4; c12++) {
                    int c1 = (c12 - 1) % 2 + 1, c2 = (c12 - 1) / 2 + 1;
                    for (int64_t w = 0; w < nw; w++) {
                        cplx a = Gm[GIDX(w, c1, i)] * X[JIDX(w, 1, cart, c0, i)] * Gp[GIDX(w, 1 + 2 * (c2 - 1), i)];
                        cplx b = Gm[GIDX(w, c1 + 2, i)] * X[JIDX(w, 2, cart, c0, i)] * Gp[GIDX(w, 1 + 2 * (c2 - 1), i)];
                        cplx c = Gm[GIDX(w, c1, i)] * X[JIDX(w, 3, cart, c0, i)] * Gp[GIDX(w, 2 + 2 * (c2 - 1), i)];
                        cplx d = Gm[GIDX(w, c1 + 2, i)] * X[JIDX(w, 4, cart, c0, i)] * Gp[GIDX(w, 2 + 2 * (c2 - 1), i)];
                        J[JIDX(w, c12, cart, c0, i)] = ((a + b) + c) + d;
                    }
                }
}

/* ------------------------------------------------------------- ladder_kernel.jl:70-99 */
void orc_attach_greens_bare(int64_t nw, int64_t ni, const double *Gm_, const double *v, const double *Gp_,
                            double *J_)
{
    const cplx *Gm = (const cplx *)Gm_, *Gp = (const cplx *)Gp_;
    cplx *J = (cplx *)J_;
    for (int64_t i = 0; i < ni; i++)
        for (int c0 = 1; c0 <= 2; c0++)
            for (int cart = 1; cart <= 3; cart++)
                for (int c12 = 1; c12 <= 4; c12++) {
                    int c1 = (c12 - 1) % 2 + 1, c2 = (c12 - 1) / 2 + 1;
                    for (int64_t w = 0; w < nw; w++)
                        J[JIDX(w, c12, cart, c0, i)] = Gm[GIDX(w, c1 + 2 * (c0 - 1), i)] * v[cart - 1 + 3 * i] *
                                                       Gp[GIDX(w, c0 + 2 * (c2 - 1), i)];
                }
}

/* ------------------------------------------------------------- ladder_kernel.jl:116-122 */
void orc_conductivity_from_vertex(int64_t nw, int64_t ni, const double *J_, const double *v, const double *w_i,
                                  const double *Ww, double Omega, double T, double *sigma)
{
    const cplx *J = (const cplx *)J_;
    cplx F = Omega == 0 ? I / (4 * T) : -1.0 / (2 * I * Omega * (1 + 2 * orc_occ_boson(Omega, T)));
    cplx *Y = (cplx *)calloc((size_t)(3 * ni), sizeof(cplx));
    for (int64_t i = 0; i < ni; i++)
        for (int b = 1; b <= 3; b++) {
            cplx acc = 0;
            for (int64_t w = 0; w < nw; w++) acc += Ww[w] * (J[JIDX(w, 4, b, 1, i)] + J[JIDX(w, 1, b, 2, i)]);
            Y[b - 1 + 3 * i] = acc;
        }
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            cplx S = 0;
            for (int64_t i = 0; i < ni; i++) S += v[a + 3 * i] * w_i[i] * Y[b + 3 * i];
            sigma[a + 3 * b] = creal(F / (2 * PI * I) * S);
        }
    free(Y);
}

/* ------------------------------------------------------------- ladder_kernel.jl:158-176 */
void orc_vertex_shift_tables(int64_t nw, const int64_t *off, int64_t nshift, int32_t *klo, double *wlo, double *whi)
{
    for (int s = 0; s < 2; s++) {
        int64_t sign = s == 0 ? 1 : -1;
        for (int64_t m = 0; m < nshift; m++)
            for (int64_t j = 0; j < nw; j++) {
                int64_t p = off[j] + sign * m;
                int64_t lo = 0; /* searchsortedlast */
                for (int64_t q = 0; q < nw; q++)
                    if (off[q] <= p) lo = q + 1;
                int64_t k = lo < 1 ? 1 : (lo > nw - 1 ? nw - 1 : lo);
                double t = (double)(p - off[k - 1]) / (double)(off[k] - off[k - 1]);
                int inside = off[0] <= p && p <= off[nw - 1];
                size_t q = (size_t)(j + nw * (m + nshift * s));
                klo[q] = (int32_t)k;
                wlo[q] = inside ? 1 - t : 0;
                whi[q] = inside ? t : 0;
            }
    }
}

/* ---------------------------------------------------------------- the kernel application */
static int cmp_pair(const void *a, const void *b)
{
    const int64_t *x = (const int64_t *)a, *y = (const int64_t *)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    return (x[1] > y[1]) - (x[1] < y[1]);
}

/* apply_ladder_kernel!, src/ladder.jl:245-296 with the chunks of _ladder_workspace (:174-233) */
int orc_apply_ladder_kernel(const orc_ladder_setup *L, const double *J_, const double *Gm, const double *Gp,
                            double T, double w_ac, int64_t gpu_tile, double *LJ_, int nthreads)
{
    const int64_t nm = L->nm, ni = L->ni, nf = L->nf, nw = L->nw, ns = L->nshift;
    const cplx *J = (const cplx *)J_;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    /* stars: inner_states_by_star = sortperm(f_to_i), inner_star_ptr (src/ladder.jl:144-149) */
    int64_t *pairs = (int64_t *)malloc(sizeof(int64_t) * 2 * (size_t)nf);
    for (int64_t f = 0; f < nf; f++) { pairs[2 * f] = L->f_to_i[f]; pairs[2 * f + 1] = f; }
    qsort(pairs, (size_t)nf, 2 * sizeof(int64_t), cmp_pair);
    int64_t *by_star = (int64_t *)malloc(sizeof(int64_t) * (size_t)nf);
    int64_t *sptr = (int64_t *)calloc((size_t)(ni + 2), sizeof(int64_t));
    for (int64_t f = 0; f < nf; f++) { by_star[f] = pairs[2 * f + 1]; sptr[L->f_to_i[f]]++; }
    free(pairs);
    sptr[0] = 0; /* 0-based: star of i (0-based) is by_star[sptr[i] .. sptr[i+1]-1] */
    for (int64_t i = 0; i < ni; i++) sptr[i + 1] += sptr[i];
    /* shift tables (src/ladder.jl:151-152) */
    int32_t *klo = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nw * ns * 2));
    double *wlo = (double *)malloc(sizeof(double) * (size_t)(nw * ns * 2));
    double *whi = (double *)malloc(sizeof(double) * (size_t)(nw * ns * 2));
    orc_vertex_shift_tables(nw, L->offsets, ns, klo, wlo, whi);
    /* chunk ranges (src/ladder.jl:180-195) */
    const int64_t CHUNK = (int64_t)1 << 30;
    int64_t bf = 8 * (ns * ni * 2 > nm * ni ? ns * ni * 2 : nm * ni);
    int64_t bi = 8 * (ns * 9 * ni * 2 > nw * 16 * ns * 3 ? ns * 9 * ni * 2 : nw * 16 * ns * 3);
    int64_t max_f = gpu_tile > 0 ? gpu_tile : (CHUNK / bf > 1 ? CHUNK / bf : 1);
    int64_t max_i = CHUNK / bi > 1 ? CHUNK / bi : 1;
    const int64_t nrow = nw * 16, ncol = 6 * ni;
    double *Cp = (double *)calloc((size_t)(nrow * ncol), sizeof(double));
    double *Cm = (double *)calloc((size_t)(nrow * ncol), sizeof(double));
    int64_t i1 = 0;
    while (i1 < ni) {
        int64_t i2 = i1;
        while (i2 < ni - 1 && i2 - i1 + 2 <= max_i && sptr[i2 + 2] - sptr[i1] <= max_f) i2++;
        const int64_t nip = i2 - i1 + 1, f0 = sptr[i1], nfc = sptr[i2 + 1] - f0;
        /* ωq, nB, B over all outer states i (:257-278) */
        double *B = (double *)calloc((size_t)(ns * nfc * ni * 2), sizeof(double));
#pragma omp parallel for schedule(static)
        for (int64_t fc = 0; fc < nfc; fc++) {
            const int64_t f = by_star[f0 + fc];
            for (int64_t i = 0; i < ni; i++) {
                int64_t q = L->iq_kk[(L->iks_i[i] - 1) + L->nki * (int64_t)(L->iks_f[f] - 1)] - 1;
                for (int type = 1; type <= 2; type++) {
                    double *Bc = B + ns * (fc + nfc * (i + ni * (type - 1)));
                    for (int64_t nu = 0; nu < nm; nu++) {
                        double om = L->wph[nu + nm * q];
                        if (om < w_ac) continue;
                        double nB = orc_occ_boson(om, T);
                        double x = om / L->dw_unit;
                        int64_t lo = (int64_t)floor(x);
                        double g = L->g2[nu + nm * (i + ni * f)];
                        for (int64_t m = lo; m <= lo + 1; m++) {
                            double tent = 1 - fabs(x - (double)m);
                            double c = g * (tent > 0 ? tent : 0.0);
                            Bc[m] += type == 1 ? c : c * nB;
                        }
                    }
                    for (int64_t m = 0; m < ns; m++) Bc[m] *= L->w_f[f];
                }
            }
        }
        /* B̃[m, b, ip, a, i, type] (:240-264) */
        const int64_t K = ns * 3 * nip;
        double *Bt = (double *)malloc(sizeof(double) * (size_t)(K * ncol));
#pragma omp parallel for schedule(static)
        for (int64_t col = 0; col < ncol; col++) {
            const int64_t a = col % 3, i = (col / 3) % ni, type = col / (3 * ni);
            for (int64_t ip = 0; ip < nip; ip++)
                for (int b = 0; b < 3; b++)
                    for (int64_t m = 0; m < ns; m++) {
                        double acc = 0;
                        for (int64_t fs = sptr[i1 + ip]; fs < sptr[i1 + ip + 1]; fs++) {
                            const int64_t fc = fs - f0, f = by_star[fs];
                            acc += L->R[a + 3 * (b + 3 * f)] * B[m + ns * (fc + nfc * (i + ni * type))];
                        }
                        Bt[m + ns * (b + 3 * ip) + K * col] = acc;
                    }
        }
        free(B);
        /* X and C_s += X B̃ (:283-290) */
        double *X = (double *)malloc(sizeof(double) * (size_t)(nrow * K));
        for (int s = 0; s < 2; s++) {
            double *C = s == 0 ? Cp : Cm;
#pragma omp parallel for schedule(static)
            for (int64_t k = 0; k < K; k++) {
                const int64_t m = k % ns, b = (k / ns) % 3, ip = k / (3 * ns);
                const int64_t i = i1 + ip;
                for (int64_t r = 0; r < nrow; r++) {
                    const int64_t j = r % nw, c12 = (r / nw) % 4, c0 = (r / (4 * nw)) % 2, ri = r / (8 * nw);
                    const size_t tq = (size_t)(j + nw * (m + ns * s));
                    const int64_t kl = klo[tq] - 1;
                    cplx z = wlo[tq] * J[JIDX(kl, c12 + 1, b + 1, c0 + 1, i)] + whi[tq] * J[JIDX(kl + 1, c12 + 1, b + 1, c0 + 1, i)];
                    X[r + nrow * k] = ri == 0 ? creal(z) : cimag(z);
                }
            }
#pragma omp parallel for schedule(static)
            for (int64_t r = 0; r < nrow; r++)
                for (int64_t col = 0; col < ncol; col++) {
                    double acc = 0;
                    for (int64_t k = 0; k < K; k++) acc += X[r + nrow * k] * Bt[k + K * col];
                    C[r + nrow * col] += acc;
                }
        }
        free(X);
        free(Bt);
        i1 = i2 + 1;
    }
    /* apply_phonon_exchange_kernel! (ladder_kernel.jl:296-322) */
    const int64_t ny = nw * 4 * 2 * 2 * 3 * ni;  /* Y[ω, c12, c0, ri, a, i] */
    double *Y = (double *)malloc(sizeof(double) * (size_t)ny);
    double *KY = (double *)malloc(sizeof(double) * (size_t)ny);
    for (int64_t q = 0; q < ny; q++) Y[q] = (Cp[q] - Cm[q]) / 2;  /* type 1 block is the first ny */
    const int64_t nc = ny / nw;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; c++)
        for (int64_t r = 0; r < nw; r++) {
            double acc = 0;
            for (int64_t k = 0; k < nw; k++) acc += L->KK[r + nw * k] * Y[k + nw * c];
            KY[r + nw * c] = acc;
        }
    cplx *KJ = (cplx *)malloc(sizeof(cplx) * (size_t)(nw * 24 * ni));
#define CI(w, c12, c0, ri, a, i, type) ((w) + nw * ((c12) + 4 * ((c0) + 2 * ((ri) + 2 * ((a) + 3 * ((i) + ni * (int64_t)(type)))))))
    for (int64_t i = 0; i < ni; i++)
        for (int c0 = 0; c0 < 2; c0++)
            for (int a = 0; a < 3; a++)
                for (int c12 = 0; c12 < 4; c12++)
                    for (int64_t w = 0; w < nw; w++) {
                        cplx Up = CMPLX(Cp[CI(w, c12, c0, 0, a, i, 0)], Cp[CI(w, c12, c0, 1, a, i, 0)]);
                        cplx Um = CMPLX(Cm[CI(w, c12, c0, 0, a, i, 0)], Cm[CI(w, c12, c0, 1, a, i, 0)]);
                        cplx Vp = CMPLX(Cp[CI(w, c12, c0, 0, a, i, 1)], Cp[CI(w, c12, c0, 1, a, i, 1)]);
                        cplx Vm = CMPLX(Cm[CI(w, c12, c0, 0, a, i, 1)], Cm[CI(w, c12, c0, 1, a, i, 1)]);
                        cplx val;
                        if (c12 == 0 || c12 == 3) {
                            cplx ky = CMPLX(KY[CI(w, c12, c0, 0, a, i, 0)], KY[CI(w, c12, c0, 1, a, i, 0)]);
                            val = (c12 == 0 ? -I : I) * ky - (Up + Um) / 2 - (Vp + Vm);
                        } else if (c12 == 1) {
                            val = Vp + Um + Vm;
                        } else {
                            val = Up + Vp + Vm;
                        }
                        KJ[JIDX(w, c12 + 1, a + 1, c0 + 1, i)] = val;
                    }
#undef CI
    orc_attach_greens(nw, ni, Gm, (const double *)KJ, Gp, LJ_);
    cplx *LJ = (cplx *)LJ_;
    for (int64_t q = 0; q < nw * 24 * ni; q++) LJ[q] = -LJ[q];
    free(KJ); free(Y); free(KY); free(Cp); free(Cm); free(klo); free(wlo); free(whi); free(by_star); free(sptr);
    return 0;
}

/* ------------------------------------------------------------------- GMRES (Krylov.jl) */
static double vnorm(int64_t n, const cplx *x)
{
    double s = 0;
    for (int64_t q = 0; q < n; q++) s += creal(x[q]) * creal(x[q]) + cimag(x[q]) * cimag(x[q]);
    return sqrt(s);
}

static cplx vdot(int64_t n, const cplx *x, const cplx *y) /* x' y */
{
    cplx s = 0;
    for (int64_t q = 0; q < n; q++) s += conj(x[q]) * y[q];
    return s;
}

/* Krylov.jl's sym_givens for a complex a and a real b: c real, s complex, ρ with
 * [c s; conj(s) -c] [a; b] = [ρ; 0]. */
static void sym_givens(cplx a, double b, double *c, cplx *s, cplx *rho)
{
    double abs_a = cabs(a);
    if (abs_a == 0) { *c = 0; *s = 1; *rho = b; return; }
    if (b == 0) { *c = 1; *s = 0; *rho = a; return; }
    double nrm = hypot(abs_a, b);
    *c = abs_a / nrm;
    *s = (a / abs_a) * b / nrm;  /* s = sign(a) b / ‖(a, b)‖ */
    *rho = (a / abs_a) * nrm;
}

typedef struct {
    const orc_ladder_setup *L;
    const double *Gm, *Gp;
    double T, w_ac;
    int64_t gpu_tile;
    int nthreads;
    double *tmp;
} ladder_op;

static void apply_I_minus_L(ladder_op *op, const cplx *x, cplx *y, int64_t n)
{
    orc_apply_ladder_kernel(op->L, (const double *)x, op->Gm, op->Gp, op->T, op->w_ac, op->gpu_tile, op->tmp,
                            op->nthreads);
    const cplx *LJ = (const cplx *)op->tmp;
    for (int64_t q = 0; q < n; q++) y[q] = x[q] - LJ[q];
}

/* (I - L) x = b, x in/out (warm: x holds x0), atol; returns niter, solved */
static int64_t gmres(ladder_op *op, int64_t n, const cplx *b, cplx *x, int warm, double atol, int64_t itmax,
                     double *norms, int64_t *nnorms, int *solved)
{
    cplx *r = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    if (warm) {
        apply_I_minus_L(op, x, r, n);
        for (int64_t q = 0; q < n; q++) r[q] = b[q] - r[q];
    } else {
        for (int64_t q = 0; q < n; q++) { r[q] = b[q]; x[q] = 0; }
    }
    double beta = vnorm(n, r);
    int64_t h = 0;
    norms[h++] = beta;
    *solved = beta <= atol;
    if (beta == 0 || *solved) { *nnorms = h; free(r); return 0; }
    int64_t cap = itmax + 1;
    cplx *V = (cplx *)malloc(sizeof(cplx) * (size_t)(n * cap));
    cplx *Rm = (cplx *)calloc((size_t)(cap * cap), sizeof(cplx));
    double *cs = (double *)malloc(sizeof(double) * (size_t)cap);
    cplx *sn = (cplx *)malloc(sizeof(cplx) * (size_t)cap);
    cplx *z = (cplx *)calloc((size_t)(cap + 1), sizeof(cplx));
    cplx *q = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int64_t k = 0; k < n; k++) V[k] = r[k] / beta;
    z[0] = beta;
    int64_t iter = 0;
    const double btol = pow(2.220446049250313e-16, 0.75);
    int breakdown = 0;
    while (!(*solved || iter >= itmax || breakdown)) {
        const int64_t j = iter;
        iter++;
        apply_I_minus_L(op, V + n * j, q, n);
        for (int64_t i = 0; i <= j; i++) {
            cplx hij = vdot(n, V + n * i, q);
            Rm[i + cap * j] = hij;
            for (int64_t k = 0; k < n; k++) q[k] -= hij * V[k + n * i];
        }
        double Hbis = vnorm(n, q);
        for (int64_t i = 0; i < j; i++) {
            cplx a = Rm[i + cap * j], bb = Rm[i + 1 + cap * j];
            Rm[i + cap * j] = cs[i] * a + sn[i] * bb;
            Rm[i + 1 + cap * j] = conj(sn[i]) * a - cs[i] * bb;
        }
        cplx rho;
        sym_givens(Rm[j + cap * j], Hbis, &cs[j], &sn[j], &rho);
        Rm[j + cap * j] = rho;
        z[j + 1] = conj(sn[j]) * z[j];
        z[j] = cs[j] * z[j];
        double rn = cabs(z[j + 1]);
        norms[h++] = rn;
        breakdown = Hbis <= btol;
        *solved = rn <= atol;
        if (!(*solved || iter >= itmax || breakdown))
            for (int64_t k = 0; k < n; k++) V[k + n * (j + 1)] = q[k] / Hbis;
    }
    /* y = R \ z, x += V y */
    cplx *y = (cplx *)calloc((size_t)cap, sizeof(cplx));
    for (int64_t i = iter - 1; i >= 0; i--) {
        cplx s = z[i];
        for (int64_t k = i + 1; k < iter; k++) s -= Rm[i + cap * k] * y[k];
        y[i] = s / Rm[i + cap * i];
    }
    for (int64_t i = 0; i < iter; i++)
        for (int64_t k = 0; k < n; k++) x[k] += V[k + n * i] * y[i];
    *nnorms = h;
    free(y); free(V); free(Rm); free(cs); free(sn); free(z); free(q); free(r);
    return iter;
}

/* One (T, Ω) of _loop_ladder, src/ladder.jl:327-371 */
int orc_ladder_solve(const orc_ladder_setup *L, const double *Sigma_T, const double *e_i, const double *w_i,
                     const double *v, const double *Ww, double Omega, double mu, double T,
                     const orc_ladder_solve_params *p, double *J_, orc_ladder_solve_out *o)
{
    const int64_t nw = L->nw, ni = L->ni, n = nw * 24 * ni;
    /* Σ at ωs ∓ Ω/2 (:322-325, :332-333) */
    double *xs = (double *)malloc(sizeof(double) * (size_t)nw);
    int32_t *jt = (int32_t *)malloc(sizeof(int32_t) * (size_t)nw);
    double *tt = (double *)malloc(sizeof(double) * (size_t)nw);
    double *Sdm = (double *)malloc(sizeof(double) * 2 * (size_t)(nw * ni));
    double *Sdp = (double *)malloc(sizeof(double) * 2 * (size_t)(nw * ni));
    for (int64_t w = 0; w < nw; w++) xs[w] = L->ws[w] - Omega / 2;
    orc_linear_interpolation_table(nw, L->ws, nw, xs, jt, tt);
    orc_apply_interpolation_table(nw, nw, ni, Sigma_T, jt, tt, Sdm);
    for (int64_t w = 0; w < nw; w++) xs[w] = L->ws[w] + Omega / 2;
    orc_linear_interpolation_table(nw, L->ws, nw, xs, jt, tt);
    orc_apply_interpolation_table(nw, nw, ni, Sigma_T, jt, tt, Sdp);
    double *Gm = (double *)malloc(sizeof(double) * 2 * (size_t)(nw * 4 * ni));
    double *Gp = (double *)malloc(sizeof(double) * 2 * (size_t)(nw * 4 * ni));
    orc_contour_greens(nw, ni, L->ws, e_i, Sdm, Sdp, Omega, mu, T, Gm, Gp);
    cplx *J0 = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    orc_attach_greens_bare(nw, ni, Gm, v, Gp, (double *)J0);
    orc_conductivity_from_vertex(nw, ni, (const double *)J0, v, w_i, Ww, Omega, T, o->sigma_bubble);
    cplx *J = (cplx *)J_;
    ladder_op op = {L, Gm, Gp, T, p->w_acoustic, p->gpu_tile, p->nthreads,
                    (double *)malloc(sizeof(cplx) * (size_t)n)};
    o->converged = 0;
    o->niter = 0;
    o->n_norms = 0;
    if (p->solver == 0) {
        int solved;
        o->niter = gmres(&op, n, J0, J, p->warm, p->rtol * vnorm(n, J0), p->maxiter, o->norms, &o->n_norms, &solved);
        o->converged = solved;
    } else {
        memcpy(J, J0, sizeof(cplx) * (size_t)n);
        for (int64_t it = 1; it <= p->maxiter; it++) {
            orc_apply_ladder_kernel(L, (const double *)J, Gm, Gp, T, p->w_acoustic, p->gpu_tile, op.tmp, p->nthreads);
            const cplx *LJ = (const cplx *)op.tmp;
            for (int64_t q = 0; q < n; q++) J[q] = J0[q] + LJ[q];
            double *sg = o->sigma_iter + 9 * (it - 1);
            orc_conductivity_from_vertex(nw, ni, (const double *)J, v, w_i, Ww, Omega, T, sg);
            o->niter = it;
            if (it > 1) {
                double d = 0, nn = 0;
                for (int k = 0; k < 9; k++) {
                    d += (sg[k] - sg[k - 9]) * (sg[k] - sg[k - 9]);
                    nn += sg[k] * sg[k];
                }
                if (sqrt(d) < p->rtol * sqrt(nn)) { o->converged = 1; break; }
            }
        }
    }
    orc_conductivity_from_vertex(nw, ni, (const double *)J, v, w_i, Ww, Omega, T, o->sigma_ladder);
    free(op.tmp); free(J0); free(Gm); free(Gp); free(xs); free(jt); free(tt); free(Sdm); free(Sdp);
    return 0;
}
