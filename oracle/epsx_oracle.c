/*
 * epsx_oracle.c -- TEST INFRASTRUCTURE ONLY: the CPU FP64 restatement of EPSpectral.jl's
 * scGD0 path (grids, Fan-Migdal sum, Kramers-Kronig, multiplets, spectral DOS, μ solve,
 * the self-consistent loop).  The ladder part is epsx_oracle_ladder.c.
 *
 * It is the checker of the HIP library and the CPU baseline of bench.py; nothing in the
 * product links it.  Built by oracle/Makefile with gcc (no FMA contraction: x86-64 has no
 * FMA in its baseline ISA and -ffp-contract=off is set), OpenMP over chunks of outer
 * states as the reference threads over 2*nthreads chunks of i (src/scgd0_kernel.jl:62-66).
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *  - orc_grid_from_nodes, orc_graded_grid: the reference's known answers
 *    (test/test_frequency_grids.jl:9-74, :172-225) -- exact node lists, offsets, steps,
 *    error cases.
 *  - orc_cell_search_table, orc_interp_linear_grid: bitwise equal to a bisection lookup on
 *    nodes, their neighbouring floats, NaN and a 500x graded grid (:130-158); against
 *    numpy.interp (flat extrapolation) at 1e-13 (:160-166).
 *  - orc_linear_interpolation_table / apply: against numpy.interp at 1e-13, end values
 *    exact (:117-128).
 *  - orc_kramers_kronig_matrix: composite-Simpson subtracted PV on an irregular 7-node
 *    grid at 1e-8 (:84-109); single-pole transform at 0.1 (test/test_kramers_kronig.jl).
 *  - orc_trapezoid_weights: sum = span, exact for linear integrands
 *    (test/test_ladder_bubble.jl:62-69).
 *  - orc_add_fan_migdal / orc_fm_term: a naive numpy 4-loop with numpy.interp lookups on
 *    pseudorandom data, a graded grid, an acoustic mode with Inf g2 and a random f->i map,
 *    at 1e-13 (test/test_scgd0_kernel.jl:6-44).
 *  - orc_loop_scgd0: the closed-form Holstein one-shot at 1e-12
 *    (test/holstein_1d_reference.jl:9-17), the local scalar iteration at 1e-10 / 1e-9
 *    (:24-50), chunk invariance, two temperatures at once, restart in one iteration,
 *    maxiter / no-root / multiplet behaviour (test/test_scgd0_loop.jl), and the run record
 *    of examples/scgd0/run_scgd0_holstein_1d.jl:33-36 (11 iterations, the err sequence).
 *  - orc_ncarrier_spectral: the per-state formula of test/test_scgd0_mu.jl:7-25 at 1e-12,
 *    with the branch selection of :53-74.
 *  - orc_mu_bisection: μ pinned at 1e-9 through the local iteration above; the bisection
 *    bits of Roots.find_zero are parity unpinned (Roots is not available here).
 *  - orc_build_multiplets: hand-checked groups on constructed energy lists.
 */
#include "epsx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double SQRT_EPS = 1.4901161193847656e-08; /* sqrt(eps(Float64)) */
static const double PI = 3.141592653589793;

/* ---------------------------------------------------------------- utils.jl:14-38 */
double orc_occ_fermion(double e, double T)
{
    if (T > SQRT_EPS) return 1.0 / (exp(e / T) + 1.0);
    if (T >= 0) {
        double s = e > 0 ? 1.0 : (e < 0 ? -1.0 : 0.0);
        return (1.0 - s) / 2.0;
    }
    return NAN;
}

double orc_occ_boson(double e, double T)
{
    if (T > SQRT_EPS) return e == 0 ? -0.5 : 1.0 / expm1(e / T);
    if (T >= 0) {
        if (e > 0) return 0.0;
        if (e == 0) return -0.5;
        return -1.0;
    }
    return NAN;
}

/* ------------------------------------------------------ frequency_grids.jl:143-171 */
int orc_grid_from_nodes(int64_t n, const double *ws, double atol, double *origin,
                        double *unit, int64_t *offsets, double *dev_out, int64_t *bad_j,
                        double *best_unit)
{
    /* findmin(diff(ωs)): the first minimum, a NaN wins (Julia's findmin) */
    double dmin = ws[1] - ws[0];
    int64_t jmin = 1;
    for (int64_t j = 1; j < n - 1; j++) {
        double g = ws[j + 1] - ws[j];
        if (isnan(dmin)) break;
        if (isnan(g) || g < dmin) { dmin = g; jmin = j + 1; }
    }
    double span = ws[n - 1] - ws[0];
    if (!(dmin > 1e-8 * span)) { *bad_j = jmin; *dev_out = dmin; return 1; }
    double best_dev = INFINITY, bu = dmin;
    int64_t bj = 0;
    for (int k = 1; k <= 64; k++) {
        double u = span / rint(span / (dmin / k));
        /* findmax(abs.(p .- round.(p))): the first maximum, a NaN wins */
        double dev = -1;
        int64_t jdev = 1;
        for (int64_t j = 0; j < n; j++) {
            double p = (ws[j] - ws[0]) / u;
            double d = fabs(p - rint(p));
            if (isnan(dev)) break;
            if (isnan(d) || d > dev) { dev = d; jdev = j + 1; }
        }
        if (dev <= atol) {
            *origin = ws[0];
            *unit = u;
            for (int64_t j = 0; j < n; j++) offsets[j] = (int64_t)rint((ws[j] - ws[0]) / u);
            *dev_out = dev;
            return 0;
        }
        if (dev < best_dev) { best_dev = dev; bj = jdev; bu = u; }
    }
    *bad_j = bj;
    *dev_out = best_dev;
    *best_unit = bu;
    return 2;
}

/* ------------------------------------------------------ frequency_grids.jl:261-294 */
static int cmp_i64(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

int64_t orc_graded_grid(int64_t nwin, const double *win, double *origin, double *unit,
                        int64_t *ps, int64_t cap)
{
    if (nwin < 1) return -1;
    for (int64_t l = 0; l < nwin; l++)
        if (!(win[3 * l + 2] > 0) || !(win[3 * l] < win[3 * l + 1])) return -2;
    for (int64_t l = 1; l < nwin; l++)
        if (!(win[3 * l] <= win[3 * (l - 1)]) || !(win[3 * l + 1] >= win[3 * (l - 1) + 1]))
            return -3;
    double o = (win[0] + win[1]) / 2, dw1 = win[2], tol = 1e-8;
    int64_t n = 0;
    for (int64_t l = 0; l < nwin; l++) {
        double u = (win[3 * l] - o) / dw1, v = (win[3 * l + 1] - o) / dw1;
        double ratio = win[3 * l + 2] / dw1;
        if (!(fabs(ratio - rint(ratio)) <= tol)) return -4;
        int64_t r = (int64_t)rint(ratio);
        if (l == 0) {
            int64_t lo = (int64_t)ceil(u - tol), hi = (int64_t)floor(v + tol);
            for (int64_t p = lo; p <= hi; p++) { if (n >= cap) return -6; ps[n++] = p; }
            if (n < 2) return -5;
        } else {
            int64_t plo = ps[0], phi = ps[0];
            for (int64_t q = 0; q < n; q++) {
                if (ps[q] < plo) plo = ps[q];
                if (ps[q] > phi) phi = ps[q];
            }
            int64_t hi = (int64_t)floor(v + tol), lo = (int64_t)ceil(u - tol);
            for (int64_t p = phi + r; p <= hi; p += r) { if (n >= cap) return -6; ps[n++] = p; }
            for (int64_t p = plo - r; p >= lo; p -= r) { if (n >= cap) return -6; ps[n++] = p; }
        }
    }
    qsort(ps, (size_t)n, sizeof(int64_t), cmp_i64);
    *origin = o;
    *unit = dw1;
    return n;
}

/* ------------------------------------------------------ frequency_grids.jl:184-196 */
int64_t orc_cell_search_table(int64_t n, const double *ws, double dw_min, double *x0,
                              double *inv_h, int32_t *jtab, int64_t cap)
{
    double a = ws[0];
    double ih = 1.0 / ((1 - 1e-6) * dw_min);
    double nb_f = (ws[n - 1] - a) * ih;
    if (nb_f > 1099511627776.0) nb_f = 1099511627776.0; /* 2^40 */
    int64_t nbuckets = (int64_t)trunc(nb_f) + 1;
    *x0 = a;
    *inv_h = ih;
    if (nbuckets > (1 << 24)) return -1;
    int64_t *b = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    for (int64_t j = 0; j < n; j++) b[j] = (int64_t)((ws[j] - a) * ih);
    for (int64_t j = 1; j < n; j++)
        if (b[j] == b[j - 1]) { free(b); return -2; }
    int64_t len = b[n - 1] + 1;
    if (jtab) {
        if (len > cap) { free(b); return -3; }
        /* jtab[bkt + 1] = searchsortedfirst(buckets, bkt) - 1 = #buckets < bkt */
        int64_t q = 0;
        for (int64_t bkt = 0; bkt < len; bkt++) {
            while (q < n && b[q] < bkt) q++;
            jtab[bkt] = (int32_t)q;
        }
    }
    free(b);
    return len;
}

/* ------------------------------------------------------ frequency_grids.jl:207-219 */
void orc_interp_linear_grid(const double *ys, int64_t col, double x, int64_t n,
                            const double *ws, const int32_t *jtab, double x0, double inv_h,
                            double *out)
{
    const double *c = ys + 2 * n * col;
    if (!(x > ws[0])) { out[0] = c[0]; out[1] = c[1]; return; }
    if (x >= ws[n - 1]) { out[0] = c[2 * (n - 1)]; out[1] = c[2 * (n - 1) + 1]; return; }
    int64_t j = jtab[(int64_t)((x - x0) * inv_h)];   /* 1-based j = jtab[b + 1] */
    j += ws[j] <= x;                                    /* ωs[j + 1] <= x, 1-based */
    /* 1-based j -> 0-based cell j-1 */
    double t = (x - ws[j - 1]) / (ws[j] - ws[j - 1]);
    double a = 1 - t;
    out[0] = a * c[2 * (j - 1)] + t * c[2 * j];
    out[1] = a * c[2 * (j - 1) + 1] + t * c[2 * j + 1];
}

/* searchsortedlast on a sorted vector: the last index (1-based) with v[idx] <= x, 0 if none */
static int64_t ssl_d(int64_t n, const double *v, double x)
{
    int64_t lo = 0, hi = n; /* answer in [0, n] */
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) / 2;
        if (v[mid - 1] <= x) lo = mid; else hi = mid - 1;
    }
    return lo;
}

/* ------------------------------------------------------- frequency_grids.jl:65-84 */
void orc_linear_interpolation_table(int64_t n, const double *ws, int64_t m, const double *xs,
                                    int32_t *j, double *t)
{
    for (int64_t r = 0; r < m; r++) {
        double x = xs[r];
        if (x <= ws[0]) { j[r] = 1; t[r] = 0; }
        else if (x >= ws[n - 1]) { j[r] = (int32_t)(n - 1); t[r] = 1; }
        else {
            int64_t jr = ssl_d(n, ws, x);
            j[r] = (int32_t)jr;
            t[r] = (x - ws[jr - 1]) / (ws[jr] - ws[jr - 1]);
        }
    }
}

/* ------------------------------------------------------- frequency_grids.jl:93-96 */
void orc_apply_interpolation_table(int64_t m, int64_t n, int64_t ncol, const double *ys,
                                   const int32_t *j, const double *t, double *Y)
{
    for (int64_t c = 0; c < ncol; c++)
        for (int64_t r = 0; r < m; r++) {
            const double *y = ys + 2 * n * c;
            int64_t k = j[r] - 1;
            double tr = t[r];
            Y[2 * (r + m * c)] = (1 - tr) * y[2 * k] + tr * y[2 * (k + 1)];
            Y[2 * (r + m * c) + 1] = (1 - tr) * y[2 * k + 1] + tr * y[2 * (k + 1) + 1];
        }
}

/* ------------------------------------------------------- frequency_grids.jl:53-54 */
void orc_trapezoid_weights(int64_t n, const double *ws, double *W)
{
    for (int64_t j = 0; j < n; j++) {
        int64_t a = j + 1 < n ? j + 1 : n - 1, b = j - 1 >= 0 ? j - 1 : 0;
        W[j] = (ws[a] - ws[b]) / 2;
    }
}

/* ------------------------------------------------------- frequency_grids.jl:5-44 */
static double xlogx(double x) { return x == 0 ? 0.0 : x * log(fabs(x)); }

static double lskk(double x, double x1, double x2, double x3)
{
    return xlogx(x1 - x) / (x2 - x1) - xlogx(x2 - x) * (x3 - x1) / (x2 - x1) / (x3 - x2) +
           xlogx(x3 - x) / (x3 - x2);
}

void orc_kramers_kronig_matrix(int64_t n, const double *ws, double *K)
{
    double *xs = (double *)malloc(sizeof(double) * (size_t)(n + 2));
    xs[0] = ws[0] - (ws[1] - ws[0]);
    memcpy(xs + 1, ws, sizeof(double) * (size_t)n);
    xs[n + 1] = ws[n - 1] + (ws[n - 1] - ws[n - 2]);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; j++)
        for (int64_t i = 0; i < n; i++)
            K[i + n * j] = lskk(xs[i + 1], xs[j], xs[j + 1], xs[j + 2]) / PI;
    free(xs);
}

/* ------------------------------------------------------- scgd0_kernel.jl:27-39 */
double orc_fm_term(double w, double g2, double wq, double e_f, int64_t i_f, const double *Sin,
                   int64_t n, const double *ws, const int32_t *jtab, double x0, double inv_h,
                   double mu, double T, double w_f, double w_acoustic)
{
    if (wq < w_acoustic) return 0.0;
    double nq = orc_occ_boson(wq, T);
    double Sa[2], Se[2];
    orc_interp_linear_grid(Sin, i_f - 1, w + wq, n, ws, jtab, x0, inv_h, Sa);
    orc_interp_linear_grid(Sin, i_f - 1, w - wq, n, ws, jtab, x0, inv_h, Se);
    double za_re = w + wq - e_f - Sa[0], za_im = -Sa[1];
    double ze_re = w - wq - e_f - Se[0], ze_im = -Se[1];
    double ImGa = -za_im / (za_re * za_re + za_im * za_im);
    double ImGe = -ze_im / (ze_re * ze_re + ze_im * ze_im);
    return w_f * g2 *
           (ImGa * (nq + orc_occ_fermion(w + wq - mu, T)) +
            ImGe * (nq + 1 - orc_occ_fermion(w - wq - mu, T)));
}

/* ------------------------------------------------ scgd0_kernel.jl:59-86 + scgd0.jl:264-286
 * One f-chunk [f0, f1): per (ω, i) the chunk's terms are summed over νf (ν fastest) from
 * zero, then added to ImS -- `ImΣ_T .+= sum(bc; dims = 2, init = 0)`. */
void orc_add_fan_migdal(int64_t nm, int64_t ni, int64_t nf, int64_t f0, int64_t f1,
                        const double *g2, const double *wq, const double *e_f,
                        const int32_t *f_to_i, const double *w_f, int64_t n, const double *ws,
                        const int32_t *jtab, double x0, double inv_h, int64_t ncol,
                        const double *Sin, double mu, double T, double w_acoustic,
                        double *ImS, int nthreads)
{
    (void)ncol;
    int64_t nchunks = 2 * (nthreads > 0 ? nthreads : 1);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < nchunks; c++) {
        int64_t i0 = ni * c / nchunks, i1 = ni * (c + 1) / nchunks;
        for (int64_t i = i0; i < i1; i++)
            for (int64_t iw = 0; iw < n; iw++) {
                double acc = 0.0, w = ws[iw];
                for (int64_t f = f0; f < f1; f++)
                    for (int64_t nu = 0; nu < nm; nu++)
                        acc += orc_fm_term(w, g2[nu + nm * (i + ni * f)], wq[nu + nm * (f + nf * i)],
                                           e_f[f], f_to_i[f], Sin, n, ws, jtab, x0, inv_h, mu,
                                           T, w_f[f], w_acoustic);
                ImS[iw + n * i] += acc;
            }
    }
}

/* --------------------------------------------------------------- scgd0.jl:361-366 */
double orc_ncarrier_spectral(double mu, double T, int64_t nwa, const double *wsa,
                             const double *D, double W, int count_holes, double nstates_base,
                             double spin_degeneracy)
{
    double acc = 0;
    if (count_holes) {
        for (int64_t a = 0; a < nwa; a++) acc += D[a] * orc_occ_fermion(mu - wsa[a], T);
        acc = W - acc;
    } else {
        for (int64_t a = 0; a < nwa; a++) acc += D[a] * orc_occ_fermion(wsa[a] - mu, T);
    }
    return (acc + nstates_base) * spin_degeneracy;
}

/* Ordered-integer midpoint of two doubles of one sign (the Float64 bisection step of
 * Roots.jl's Bisection); 0.0 when the signs differ. */
static double bits_middle(double a, double b)
{
    if ((a < 0) != (b < 0) && a != 0 && b != 0) return 0.0;
    if (a == 0 || b == 0) {
        /* bisect between ±0 and the other end in the bit pattern of its sign */
        double o = a == 0 ? b : a;
        uint64_t uo;
        memcpy(&uo, &o, 8);
        uo &= 0x7fffffffffffffffULL;
        uint64_t m = uo / 2;
        double r;
        memcpy(&r, &m, 8);
        return o < 0 ? -r : r;
    }
    double sa = fabs(a), sb = fabs(b);
    uint64_t ua, ub;
    memcpy(&ua, &sa, 8);
    memcpy(&ub, &sb, 8);
    uint64_t m = ua / 2 + ub / 2 + (ua & ub & 1);
    double r;
    memcpy(&r, &m, 8);
    return a < 0 ? -r : r;
}

static double sgn(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }

double orc_mu_bisection(double a, double b, double T, int64_t nwa, const double *wsa,
                        const double *D, double W, int count_holes, double nstates_base,
                        double spin_degeneracy, double n_target, int *status)
{
#define EXCESS(m) (orc_ncarrier_spectral((m), T, nwa, wsa, D, W, count_holes, nstates_base, \
                                         spin_degeneracy) - n_target)
    double fa = EXCESS(a), fb = EXCESS(b);
    *status = 0;
    if (fa * fb > 0) { *status = 2; return NAN; }
    if (fa == 0) return a;
    if (fb == 0) return b;
    for (int it = 0; it < 200; it++) {
        double m = bits_middle(a, b);
        if (m == a || m == b) break;
        double fm = EXCESS(m);
        if (fm == 0) return m;
        if (sgn(fm) == sgn(fa)) { a = m; fa = fm; }
        else { b = m; fb = fm; }
    }
    return fabs(fa) <= fabs(fb) ? a : b;
#undef EXCESS
}

/* --------------------------------------------------------------- scgd0.jl:146-160 */
typedef struct { int32_t k; double e; int64_t i; } kei;
static int cmp_kei(const void *pa, const void *pb)
{
    const kei *a = (const kei *)pa, *b = (const kei *)pb;
    if (a->k != b->k) return a->k < b->k ? -1 : 1;
    if (a->e != b->e) return a->e < b->e ? -1 : 1;
    return (a->i > b->i) - (a->i < b->i); /* stable */
}

int64_t orc_build_multiplets(int64_t ni, const int32_t *iks, const double *es, double cutoff,
                             int64_t *mult_ptr, int64_t *mult_idx)
{
    kei *o = (kei *)malloc(sizeof(kei) * (size_t)ni);
    for (int64_t i = 0; i < ni; i++) { o[i].k = iks[i]; o[i].e = es[i]; o[i].i = i; }
    qsort(o, (size_t)ni, sizeof(kei), cmp_kei);
    int64_t nm = 0, nidx = 0;
    mult_ptr[0] = 0;
    int64_t g0 = 0; /* group = o[g0 .. q-1] */
    for (int64_t q = 1; q <= ni; q++) {
        int join = 0;
        if (q < ni) {
            const kei *cur = &o[q], *prev = &o[q - 1];
            join = cur->k == prev->k && es[cur->i] - es[prev->i] < cutoff;
        }
        if (!join) {
            if (q - g0 > 1) {
                for (int64_t r = g0; r < q; r++) mult_idx[nidx++] = o[r].i + 1;
                mult_ptr[++nm] = nidx;
            }
            g0 = q;
        }
    }
    free(o);
    return nm;
}

/* ------------------------------------------------------------- scgd0.jl:189-346 */
static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int orc_loop_scgd0(const orc_scgd0_setup *s, const orc_scgd0_params *p, orc_scgd0_out *o)
{
    const int64_t nm = s->nm, ni = s->ni, nf = s->nf, nw = s->nw, nwa = s->nwa, nT = p->nT;
    const int64_t maxiter = p->maxiter;
    const double *ws = s->ws, *wsa = s->wsa;
    double bracket[2] = {wsa[0] - 0.1, wsa[nwa - 1] + 0.1};
    int64_t nf_chunk = p->nf_chunk;
    if (nf_chunk <= 0) {
        int64_t c = ((int64_t)1 << 30) / (8 * nm * ni);
        nf_chunk = c < 1 ? 1 : (c > nf ? nf : c);
    } else if (nf_chunk > nf) nf_chunk = nf;

    /* grid maps of _setup_scgd0 (:162-163, :173-183) */
    double x0, inv_h;
    int64_t ntab = orc_cell_search_table(nw, ws, s->dw_min, &x0, &inv_h, NULL, 0);
    int32_t *jtab = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ntab > 0 ? ntab : 1));
    orc_cell_search_table(nw, ws, s->dw_min, &x0, &inv_h, jtab, ntab);
    int32_t *jsp = (int32_t *)malloc(sizeof(int32_t) * (size_t)nwa);
    double *tsp = (double *)malloc(sizeof(double) * (size_t)nwa);
    orc_linear_interpolation_table(nw, ws, nwa, wsa, jsp, tsp);
    double *Wa = (double *)malloc(sizeof(double) * (size_t)nwa);
    orc_trapezoid_weights(nwa, wsa, Wa);
    double *KK = (double *)malloc(sizeof(double) * (size_t)(nw * nw));
    orc_kramers_kronig_matrix(nw, ws, KK);

    size_t nS = (size_t)(nw * ni * nT);
    double *S = o->Sigma;
    if (p->Sigma_init) memcpy(S, p->Sigma_init, sizeof(double) * 2 * nS);
    else for (size_t q = 0; q < nS; q++) { S[2 * q] = 0.0; S[2 * q + 1] = -p->eta_init; }
    double *Snew = (double *)malloc(sizeof(double) * 2 * nS);
    memcpy(Snew, S, sizeof(double) * 2 * nS);
    double *Sin = (double *)malloc(sizeof(double) * 2 * nS);
    double *ImS = (double *)malloc(sizeof(double) * nS);
    double *ReS = (double *)malloc(sizeof(double) * (size_t)(nw * ni));
    double *Ssp = (double *)malloc(sizeof(double) * 2 * (size_t)(nwa * ni));
    double *A = (double *)malloc(sizeof(double) * (size_t)(nwa * ni));
    double *D = (double *)malloc(sizeof(double) * (size_t)nwa);
    double *wq = (double *)malloc(sizeof(double) * (size_t)(nm * nf_chunk * ni));
    int32_t *active = (int32_t *)malloc(sizeof(int32_t) * (size_t)nT);

    for (int64_t t = 0; t < nT; t++) {
        o->mu[t] = p->mu_start[t];
        active[t] = 1;
        o->iter_frozen[t] = 0;
    }
    for (int64_t q = 0; q < nT * maxiter; q++) o->hist_err[q] = o->hist_mu[q] = o->hist_minA[q] = NAN;
    o->status = 0;
    o->niter = 0;
    for (int64_t iter = 1; iter <= maxiter; iter++) {
        double t0 = now_s();
        /* 1. clamp (:257-260) */
        for (int64_t t = 0; t < nT; t++) {
            if (!active[t]) continue;
            for (int64_t q = 0; q < nw * ni; q++) {
                size_t k = (size_t)(q + nw * ni * t);
                Sin[2 * k] = S[2 * k];
                Sin[2 * k + 1] = fmin(S[2 * k + 1], -p->eta_min[t]);
            }
        }
        /* 2-3. Fan-Migdal over f-chunks (:263-286) */
        memset(ImS, 0, sizeof(double) * nS);
        for (int64_t fs = 0; fs < nf; fs += nf_chunk) {
            int64_t fe = fs + nf_chunk < nf ? fs + nf_chunk : nf, nc = fe - fs;
            /* ωq[ν, f, i] = ωph[ν, iq_kk[iks_i[i], iks_f[f]]] for the chunk */
            for (int64_t i = 0; i < ni; i++)
                for (int64_t f = 0; f < nc; f++) {
                    int64_t q = s->iq_kk[(s->iks_i[i] - 1) + s->nki * (int64_t)(s->iks_f[fs + f] - 1)] - 1;
                    for (int64_t nu = 0; nu < nm; nu++)
                        wq[nu + nm * (f + nc * i)] = s->wph[nu + nm * q];
                }
            for (int64_t t = 0; t < nT; t++) {
                if (!active[t]) continue;
                /* the chunk's wq block is [nm, nc, ni]; index it as a chunk-local f */
                orc_add_fan_migdal(nm, ni, nc, 0, nc, s->g2 + nm * ni * fs, wq, s->e_f + fs,
                                   s->f_to_i + fs, s->w_f + fs, nw, ws, jtab, x0, inv_h, ni,
                                   Sin + 2 * nw * ni * t, o->mu[t], p->Tlist[t], p->w_acoustic,
                                   ImS + nw * ni * t, p->nthreads);
            }
        }
        /* 4. Re Σ by KK (:289-292) */
        for (int64_t t = 0; t < nT; t++) {
            if (!active[t]) continue;
            const double *Im = ImS + nw * ni * t;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < ni; i++)
                for (int64_t r = 0; r < nw; r++) {
                    double acc = 0;
                    for (int64_t c = 0; c < nw; c++) acc += KK[r + nw * c] * Im[c + nw * i];
                    ReS[r + nw * i] = acc;
                }
            for (int64_t q = 0; q < nw * ni; q++) {
                size_t k = (size_t)(q + nw * ni * t);
                Snew[2 * k] = ReS[q];
                Snew[2 * k + 1] = Im[q];
            }
        }
        /* 5. multiplet average (:294-300) */
        for (int64_t t = 0; t < nT; t++) {
            if (!active[t]) continue;
            for (int64_t m = 0; m < s->n_mult; m++) {
                int64_t a = s->mult_ptr[m], b = s->mult_ptr[m + 1];
                for (int64_t r = 0; r < nw; r++) {
                    double sr = 0, si = 0;
                    for (int64_t q = a; q < b; q++) {
                        size_t k = (size_t)(r + nw * (s->mult_idx[q] - 1) + nw * ni * t);
                        sr += Snew[2 * k];
                        si += Snew[2 * k + 1];
                    }
                    sr /= (double)(b - a);
                    si /= (double)(b - a);
                    for (int64_t q = a; q < b; q++) {
                        size_t k = (size_t)(r + nw * (s->mult_idx[q] - 1) + nw * ni * t);
                        Snew[2 * k] = sr;
                        Snew[2 * k + 1] = si;
                    }
                }
            }
        }
        for (int64_t t = 0; t < nT; t++) {
            if (!active[t]) continue;
            const double *Sn = Snew + 2 * nw * ni * t;
            /* 6. spectral DOS (:304-308) */
            orc_apply_interpolation_table(nwa, nw, ni, Sn, jsp, tsp, Ssp);
            for (int64_t i = 0; i < ni; i++)
                for (int64_t a = 0; a < nwa; a++) {
                    /* -imag(inv(z)) / π with Julia's scaled complex inverse */
                    double c = wsa[a] - s->e_i[i] - Ssp[2 * (a + nwa * i)];
                    double d = -Ssp[2 * (a + nwa * i) + 1];
                    double q;
                    if (fabs(d) <= fabs(c)) {
                        double r = d / c, z = d * r + c;
                        q = -r / z;
                    } else {
                        double r = c / d, z = (-c) * r + (-d);
                        q = 1.0 / z;
                    }
                    A[a + nwa * i] = -q / PI;
                }
            for (int64_t a = 0; a < nwa; a++) {
                double acc = 0;
                for (int64_t i = 0; i < ni; i++) acc += A[a + nwa * i] * s->w_i[i];
                D[a] = acc * Wa[a];
            }
            double minA = INFINITY;
            for (int64_t i = 0; i < ni; i++) {
                double acc = 0;
                for (int64_t a = 0; a < nwa; a++) acc += A[a + nwa * i] * Wa[a];
                if (acc < minA) minA = acc;
            }
            o->hist_minA[t + nT * (iter - 1)] = minA;
            /* 7. μ (:310-324) */
            if (!p->fix_mu) {
                int holes = p->is_semiconductor && p->nlist[t] < 0;
                double n_target = p->nlist[t] + p->nelec;
                int st;
                double m = orc_mu_bisection(bracket[0], bracket[1], p->Tlist[t], nwa, wsa, D, s->W,
                                            holes, s->nstates_base, p->spin_degeneracy,
                                            n_target, &st);
                if (st) {
                    o->status = 2;
                    o->err_T = t;
                    o->err_n_target = n_target;
                    o->err_bracket[0] = bracket[0];
                    o->err_bracket[1] = bracket[1];
                    goto done;
                }
                o->mu[t] = m;
            }
            /* 8-9. convergence (:326-333) */
            double err = 0;
            double *St = S + 2 * nw * ni * t;
            for (int64_t q = 0; q < nw * ni; q++) {
                double e = hypot(Sn[2 * q] - St[2 * q], Sn[2 * q + 1] - St[2 * q + 1]);
                if (e > err || isnan(e)) err = e;
            }
            o->hist_err[t + nT * (iter - 1)] = err;
            o->hist_mu[t + nT * (iter - 1)] = o->mu[t];
            memcpy(St, Sn, sizeof(double) * 2 * (size_t)(nw * ni));
            if (err < p->tol) { active[t] = 0; o->iter_frozen[t] = iter; }
        }
        o->wall[iter - 1] = now_s() - t0;
        o->niter = iter;
        int any = 0;
        for (int64_t t = 0; t < nT; t++) any |= active[t];
        if (!any) break;
    }
    for (int64_t t = 0; t < nT; t++) o->converged[t] = !active[t];
done:
    free(jtab); free(jsp); free(tsp); free(Wa); free(KK); free(Snew); free(Sin); free(ImS);
    free(ReS); free(Ssp); free(A); free(D); free(wq); free(active);
    return o->status;
}
