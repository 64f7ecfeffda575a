/*
 * epsx_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU FP64 restatement of EPSpectral.jl's scGD0 and ladder algorithms, used as the
 * correctness checker of the HIP library (libepsx) and as the CPU baseline of the
 * benchmark.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path never links or calls it.
 *
 * Conventions follow the reference: arrays are column-major with the first index
 * fastest, state and q indices stored in integer arrays are 1-based (as Julia passes
 * them), complex numbers are interleaved (re, im) doubles (the ComplexF64 layout).
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * EPSpectral.jl repository root).  How each one is pinned (golden vectors, closed
 * forms, brute force) is listed in the header of epsx_oracle.c.
 */
#ifndef EPSX_ORACLE_H
#define EPSX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalars: src/utils.jl ---- */
double orc_occ_fermion(double e, double T);
double orc_occ_boson(double e, double T);

/* ---- grids: src/frequency_grids.jl ---- */
/* FrequencyGrid(ωs::AbstractVector; atol), :143-171.  Returns 0 on success, 1 for the
 * "apart" error (bad_j = jmin, 1-based), 2 for "not commensurate" (bad_j = worst node,
 * best_unit, best_dev as in the message). */
int orc_grid_from_nodes(int64_t n, const double *ws, double atol, double *origin,
                        double *unit, int64_t *offsets, double *dev_out, int64_t *bad_j,
                        double *best_unit);
/* graded_frequency_grid(windows), :261-294.  windows[3*l + {0,1,2}] = (wmin, wmax, dw).
 * Writes up to cap offsets; returns the node count, or a negative error code. */
int64_t orc_graded_grid(int64_t nwin, const double *windows, double *origin, double *unit,
                        int64_t *offsets, int64_t cap);
/* cell_search_table, :184-196.  Returns the table length, or -1 (too large) / -2
 * (two nodes in one bucket).  jtab may be NULL to query the length. */
int64_t orc_cell_search_table(int64_t n, const double *ws, double dw_min, double *x0,
                              double *inv_h, int32_t *jtab, int64_t cap);
/* _interp_linear_grid, :207-219, for a complex column. */
void orc_interp_linear_grid(const double *ys, int64_t col, double x, int64_t n,
                            const double *ws, const int32_t *jtab, double x0, double inv_h,
                            double *out);
/* linear_interpolation_table, :65-84 */
void orc_linear_interpolation_table(int64_t n, const double *ws, int64_t m, const double *xs,
                                    int32_t *j, double *t);
/* apply_interpolation_table!, :93-96, complex ys[n, ncol] -> Y[m, ncol] */
void orc_apply_interpolation_table(int64_t m, int64_t n, int64_t ncol, const double *ys,
                                   const int32_t *j, const double *t, double *Y);
/* trapezoid_weights, :53-54 */
void orc_trapezoid_weights(int64_t n, const double *ws, double *W);
/* kramers_kronig_matrix, :30-44 (K is n x n column-major) */
void orc_kramers_kronig_matrix(int64_t n, const double *ws, double *K);

/* ---- Fan-Migdal: src/scgd0_kernel.jl ---- */
/* _fan_migdal_imsigma_term, :27-39 (i_f 1-based, Σin complex [n, ncol]) */
double orc_fm_term(double w, double g2, double wq, double e_f, int64_t i_f, const double *Sin,
                   int64_t n, const double *ws, const int32_t *jtab, double x0, double inv_h,
                   double mu, double T, double w_f, double w_acoustic);
/* add_fan_migdal_imsigma! on the reference's [ν, i, f] inputs (the (νf, i) fusion of
 * src/scgd0.jl:264-279 done inside): ImS[w, i] += Σ_{f in fs, ν} term.  wq is
 * [nm, nf, ni] (the gathered pair table), f runs over [f0, f1) (0-based, the chunk). */
void orc_add_fan_migdal(int64_t nm, int64_t ni, int64_t nf, int64_t f0, int64_t f1,
                        const double *g2, const double *wq, const double *e_f,
                        const int32_t *f_to_i, const double *w_f, int64_t n, const double *ws,
                        const int32_t *jtab, double x0, double inv_h, int64_t ncol,
                        const double *Sin, double mu, double T, double w_acoustic,
                        double *ImS, int nthreads);

/* ---- scGD0 loop: src/scgd0.jl ---- */
typedef struct {
    int64_t nm, ni, nf, nw, nwa, nq, nki, nkf;
    const double *g2;        /* [nm, ni, nf] */
    const double *wph;       /* [nm, nq] */
    const int32_t *iq_kk;    /* [nki, nkf], 1-based q */
    const int32_t *iks_i;    /* [ni], 1-based */
    const int32_t *iks_f;    /* [nf], 1-based */
    const double *e_i, *w_i, *e_f, *w_f;
    const int32_t *f_to_i;   /* [nf], 1-based */
    double W, nstates_base;
    const double *ws;        /* [nw] output grid nodes */
    double dw_min;           /* ωs.dω_min */
    const double *wsa;       /* [nwa] spectral grid nodes */
    int64_t n_mult;          /* multiplets as CSR: members of m are */
    const int64_t *mult_ptr; /* mult_idx[mult_ptr[m] .. mult_ptr[m+1]-1] (0-based ptr) */
    const int64_t *mult_idx; /* 1-based outer states */
} orc_scgd0_setup;

typedef struct {
    int64_t nT;
    const double *Tlist, *nlist, *mu_start, *eta_min;
    double nelec, spin_degeneracy;
    int is_semiconductor;    /* occ.type === :Semiconductor */
    double eta_init, tol, w_acoustic;
    int64_t maxiter;
    int64_t nf_chunk;        /* 0: the CHUNK_BYTES default */
    int fix_mu;
    const double *Sigma_init;/* NULL or complex [nw, ni, nT] */
    int nthreads;
} orc_scgd0_params;

typedef struct {
    double *Sigma;           /* complex [nw, ni, nT] */
    double *mu;              /* [nT] */
    int32_t *converged;      /* [nT] */
    double *hist_err, *hist_mu, *hist_minA;  /* [nT, maxiter] */
    double *wall;            /* [maxiter] */
    int64_t *iter_frozen;    /* [nT] */
    int64_t niter;
    int status;              /* 0 ok, 2 no μ root (err_T = the temperature index) */
    int64_t err_T;
    double err_n_target, err_bracket[2];
} orc_scgd0_out;

/* _loop_scgd0, :189-346 */
int orc_loop_scgd0(const orc_scgd0_setup *s, const orc_scgd0_params *p, orc_scgd0_out *o);
/* compute_ncarrier_spectral, :361-366 */
double orc_ncarrier_spectral(double mu, double T, int64_t nwa, const double *wsa,
                             const double *D, double W, int count_holes, double nstates_base,
                             double spin_degeneracy);
/* The bracketed root of Roots.find_zero(f, (a, b)) with the default bisection, restated:
 * bisection on the ordered bit pattern until the bracket holds two adjacent floats. */
double orc_mu_bisection(double a, double b, double T, int64_t nwa, const double *wsa,
                        const double *D, double W, int count_holes, double nstates_base,
                        double spin_degeneracy, double n_target, int *status);
/* multiplets of _setup_scgd0, src/scgd0.jl:146-160: writes CSR (ptr 0-based, idx 1-based),
 * returns the number of multiplets. */
int64_t orc_build_multiplets(int64_t ni, const int32_t *iks, const double *es, double cutoff,
                             int64_t *mult_ptr, int64_t *mult_idx);

/* ---- ladder: src/ladder_kernel.jl, src/ladder.jl ---- */
void orc_contour_green(double w, double e, double Sre, double Sim, double mu, double T,
                       int c12, double *out);
void orc_contour_greens(int64_t nw, int64_t ni, const double *ws, const double *e_i,
                        const double *Sd_m, const double *Sd_p, double Omega, double mu,
                        double T, double *Gm, double *Gp);
void orc_attach_greens(int64_t nw, int64_t ni, const double *Gm, const double *X,
                       const double *Gp, double *J);
void orc_attach_greens_bare(int64_t nw, int64_t ni, const double *Gm, const double *v,
                            const double *Gp, double *J0);
void orc_conductivity_from_vertex(int64_t nw, int64_t ni, const double *J, const double *v,
                                  const double *w_i, const double *Ww, double Omega, double T,
                                  double *sigma);
void orc_vertex_shift_tables(int64_t nw, const int64_t *offsets, int64_t nshift, int32_t *klo,
                             double *wlo, double *whi);

typedef struct {
    int64_t nm, ni, nf, nw, nq, nki, nkf, nshift;
    const double *g2, *wph;
    const int32_t *iq_kk, *iks_i, *iks_f, *f_to_i;
    const double *w_f;
    const double *R;          /* [3, 3, nf] */
    const double *ws;         /* [nw] */
    const int64_t *offsets;   /* [nw] */
    double dw_unit;
    const double *KK;         /* [nw, nw] */
} orc_ladder_setup;

/* apply_ladder_kernel!, src/ladder.jl:245-296, with chunks of at most gpu_tile inner
 * states (0: one chunk per CHUNK_BYTES limit), LJ, J complex [nw, 4, 3, 2, ni]. */
int orc_apply_ladder_kernel(const orc_ladder_setup *L, const double *J, const double *Gm,
                            const double *Gp, double T, double w_acoustic, int64_t gpu_tile,
                            double *LJ, int nthreads);

/* _loop_ladder for one (T, Ω) with GMRES (src/ladder.jl:327-353) or the fixed point
 * (:354-369).  J is in/out: warm start when warm != 0.  Returns niter; norms gets the
 * residual history (cap maxiter + 1 entries), sigma_iter (fixed point) 9 per iteration. */
typedef struct {
    int solver;              /* 0 gmres, 1 fixed point */
    double rtol;
    int64_t maxiter;
    int warm;
    int64_t gpu_tile;
    double w_acoustic;
    int nthreads;
} orc_ladder_solve_params;

typedef struct {
    int32_t converged;
    int64_t niter;
    int64_t n_norms;
    double *norms;           /* maxiter + 1 */
    double *sigma_iter;      /* 9 * maxiter */
    double sigma_bubble[9], sigma_ladder[9];
} orc_ladder_solve_out;

int orc_ladder_solve(const orc_ladder_setup *L, const double *Sigma_T, const double *e_i,
                     const double *w_i, const double *v, const double *Ww, double Omega,
                     double mu, double T, const orc_ladder_solve_params *p, double *J,
                     orc_ladder_solve_out *o);

#ifdef __cplusplus
}
#endif
#endif
