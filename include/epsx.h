/*
 * epsx.h -- the C ABI of libepsx, the MI355X (gfx950) implementation of EPSpectral.jl's
 * scGD0 hot path.
 *
 * Every entry point takes plain pointers and sizes.  Arrays are column-major with the first
 * index fastest (Julia's layout), complex numbers are interleaved (re, im) doubles
 * (ComplexF64), and integer indices into state, q or grid arrays are 1-based, exactly as a
 * Julia `ccall` hands them over.  Host pointers are read during the call only; results go
 * into caller-allocated host arrays.
 *
 * Return codes: EPSX_OK, or one of the EPSX_ERR_* codes below.  After an error,
 * epsx_last_error(ctx) holds the message the reference throws or warns with (for a
 * Julia wrapper: EPSX_ERR_ARG and EPSX_ERR_NO_MU_ROOT become ArgumentError).
 *
 * Threading: one HIP stream per context; calls on one context are serialised by the
 * caller; separate contexts may be used from separate threads.
 *
 * The reference interface each entry replaces is cited as file:line of the EPSpectral.jl
 * repository.
 */
#ifndef EPSX_H
#define EPSX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EPSX_ABI_VERSION 2

#define EPSX_OK 0
#define EPSX_ERR_ARG 1          /* invalid argument: ArgumentError in the reference */
#define EPSX_ERR_NO_MU_ROOT 2   /* no μ in the bracket: src/scgd0.jl:317-322 */
#define EPSX_ERR_DEVICE 3       /* a HIP runtime error */
#define EPSX_ERR_OOM 4          /* device allocation failed */
#define EPSX_ERR_STATE 5        /* call order: e.g. run before set_problem */

typedef struct epsx_ctx epsx_ctx;

/* ---- context ------------------------------------------------------------------------ */
/* The device handle that replaces the `backend` keyword threaded through run_scgd0 /
 * run_scgd0_ladder (src/scgd0.jl:77; src/ladder.jl:67). */
int epsx_abi_version(void);
int epsx_open(int device, epsx_ctx **ctx);
void epsx_close(epsx_ctx *ctx);
const char *epsx_last_error(const epsx_ctx *ctx);
/* Seconds of the last run's kernels by class, measured with HIP events on the context's
 * stream, and the number of launches of each class.  Both arrays hold 8 entries.
 * After epsx_scgd0_run: [0] Fan-Migdal sum, [1] Kramers-Kronig, [2] spectral DOS and μ,
 * [3] convergence and copies, [4] the coupling reduction of the star method (once per run),
 * [5] the star method's main h·G̃ GEMM kernel (inside [0]; without its split-k sum; count =
 * launches), [6] count only: the temperature instances of those launches (of the Fan-Migdal
 * passes [0] with the Einstein factorisation), [7] count only: the kernel that ran the star
 * method's product in the timed iteration (0 none, 1 hipBLASLt, 2 + s the library's own FP64 GEMM
 * with tile shape s (11-13: the 16-byte-load kernel), 30 its low-precision kernel from stored
 * BF16 / FP32 operands, 31 the one that rounds FP64 operands as it stages them).  In the batched
 * loop the second iteration is the one timed.
 * After epsx_ladder_run / epsx_apply_ladder_kernel: [0] the shift-table GEMMs W = J·B̃,
 * [1] whole kernel applications (count = applications; a batched application counts once),
 * [2] B̃ builds, [3] count only: the (Ω, T) instances of those applications, [4..7] unused. */
int epsx_last_timings(const epsx_ctx *ctx, double *seconds8, int64_t *launches8);
/* Device memory the context holds (its buffers, the temporaries of the call in progress and the
 * released temporaries it keeps for reuse) and the part of it that is released and kept, in bytes.
 * No counterpart in the reference (its arrays are garbage-collected); a caller sizing problems
 * against HBM reads it. */
int epsx_device_memory(const epsx_ctx *ctx, int64_t *held, int64_t *pooled);

/* ---- grid maps (host, bit-exact with the reference) ---------------------------------- */
/* cell_search_table, src/frequency_grids.jl:184-196.  ws: the nodes, dw_min = ωs.dω_min.
 * jtab may be NULL to query *len.  Errors as the reference (table over 2^24 entries, two
 * nodes in one bucket). */
int epsx_cell_search_table(epsx_ctx *ctx, int64_t n, const double *ws, double dw_min,
                           int32_t *jtab, int64_t cap, int64_t *len, double *x0,
                           double *inv_h);
/* linear_interpolation_table, src/frequency_grids.jl:65-84 (j 1-based). */
int epsx_linear_interpolation_table(epsx_ctx *ctx, int64_t n, const double *ws, int64_t m,
                                    const double *xs, int32_t *j, double *t);
/* trapezoid_weights, src/frequency_grids.jl:53-54. */
int epsx_trapezoid_weights(epsx_ctx *ctx, int64_t n, const double *ws, double *W);
/* vertex_shift_tables, src/ladder_kernel.jl:158-176: klo (1-based), wlo, whi of size
 * [nw, nshift, 2] from the integer offsets of the grid. */
int epsx_vertex_shift_tables(epsx_ctx *ctx, int64_t nw, const int64_t *offsets,
                             int64_t nshift, int32_t *klo, double *wlo, double *whi);

/* ---- device operators on host buffers (the reference's unit-level functions) -------- */
/* kramers_kronig_matrix, src/frequency_grids.jl:30-44, built on the device. K: [n, n]. */
int epsx_kramers_kronig_matrix(epsx_ctx *ctx, int64_t n, const double *ws, double *K);
/* kramers_kronig(ωs, zs), src/frequency_grids.jl:227: ys = K * zs, zs [n, ncol] real,
 * the product on the device. */
int epsx_kramers_kronig(epsx_ctx *ctx, int64_t n, const double *ws, int64_t ncol,
                        const double *zs, double *ys);

/* No counterpart in the reference: epsx_kramers_kronig for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * ys = K * zs with ws [n], zs [n, ncol] and ys [n, ncol] in device memory; K and the product's
 * partial sums live in a fixed 40 MiB device workspace held per device until the process ends,
 * so n is at most 2048 (a larger n is EPSX_ERR_ARG).  Asynchronous on `stream` (a hipStream_t;
 * NULL: the context's stream).  No hipMalloc, hipFree or memset is issued.  All calls on one
 * device share the workspace: calls that overlap on one device (other streams, other contexts)
 * race. */
int epsx_kramers_kronig_noalloc(epsx_ctx *ctx, int64_t n, const double *d_ws, int64_t ncol,
                                const double *d_zs, double *d_ys, void *stream);

/* add_fan_migdal_imsigma!, src/scgd0_kernel.jl:59-72, with the per-pair inputs of the
 * reference's [ν, i, f] arrays (the (νf, i) fusion of src/scgd0.jl:264-279 is a layout
 * change the kernel does itself):
 *   ImS[ω, i] += Σ_{f, ν} _fan_migdal_imsigma_term(ωs[ω], g2[ν,i,f], wq[ν,i,f], e_f[f],
 *                                                 f_to_i[f], Sin, ...)
 * The grid is a FrequencyGrid (origin, unit, integer offsets).  Sin: complex [nw, ncol],
 * the clamped input self-energy; f_to_i in 1..ncol.  precision: 0 FP64, 1 FP32 (FP32
 * terms, compensated FP32 sum over f). */
int epsx_add_fan_migdal_imsigma(epsx_ctx *ctx, int64_t nm, int64_t ni, int64_t nf,
                                const double *g2, const double *wq, const double *e_f,
                                const int32_t *f_to_i, const double *w_f, int64_t nw,
                                double w_origin, double dw_unit, const int64_t *offsets,
                                int64_t ncol, const double *Sin, double mu, double T,
                                double w_acoustic, int32_t precision, double *ImS);

/* ---- scGD0 ---------------------------------------------------------------------------- */
/* The `setup` of _setup_scgd0 (src/scgd0.jl:173-183; built by hand at
 * test/holstein_1d_reference.jl:62-77), as plain arrays.  The library derives the grid
 * maps (cell and spectral tables, trapezoid weights, KK matrix) from the two grids itself.
 * g2 is copied to the device once, in the layout the Fan-Migdal kernel streams, and stays
 * resident. */
typedef struct {
    int64_t struct_size;      /* sizeof(epsx_scgd0_problem) */
    int64_t nm, ni, nf;       /* phonon modes, outer (IBZ) states, inner (full BZ) states */
    int64_t nq, nki, nkf;     /* phonon q points; rows and columns of iq_kk */
    const double *g2;         /* [nm, ni, nf] stored coupling */
    const double *wph;        /* [nm, nq] phonon frequencies */
    const int32_t *iq_kk;     /* [nki, nkf] 1-based q index of the pair (k_i, k_f) */
    const int32_t *iks_i;     /* [ni] 1-based row of iq_kk */
    const int32_t *iks_f;     /* [nf] 1-based column of iq_kk */
    const double *e_i, *w_i;  /* [ni] energies and weights of the outer states */
    const double *e_f, *w_f;  /* [nf] of the inner states */
    const int32_t *f_to_i;    /* [nf] 1-based IBZ representative */
    double W, nstates_base;   /* Σ w_i and the states below the window */
    int64_t nw;               /* ωs: ω_origin + offsets * dω_unit */
    double w_origin, dw_unit;
    const int64_t *w_offsets;
    int64_t nwa;              /* ωs_spectral */
    double wa_origin, dwa_unit;
    const int64_t *wa_offsets;
    int64_t n_mult;           /* multiplets (src/scgd0.jl:146-160) as CSR: */
    const int64_t *mult_ptr;  /* [n_mult + 1], 0-based */
    const int32_t *mult_idx;  /* 1-based outer states */
} epsx_scgd0_problem;

/* Fan-Migdal methods.  DIRECT is the reference's sum term by term (any phonon dispersion).
 * The factorised methods are exact rewrites behind a precondition the library checks:
 * EINSTEIN needs ωph[ν, :] constant per mode; STAR also needs e_f == e_{f_to_i[f]}
 * bitwise.  AUTO picks STAR, then EINSTEIN, then DIRECT. */
#define EPSX_FM_DIRECT 0
#define EPSX_FM_EINSTEIN 1
#define EPSX_FM_STAR 2
#define EPSX_FM_AUTO 3

/* Precision of the Fan-Migdal sum: FP64, or FP32 terms and couplings with a compensated
 * FP32 sum over inner states, or BF16 couplings with FP32 accumulation (factorised methods
 * only).  A non-FP64 run ends with FP64 sweeps until tol is met in FP64. */
#define EPSX_PREC_FP64 0
#define EPSX_PREC_FP32 1
#define EPSX_PREC_BF16 2

typedef struct {
    int64_t struct_size;      /* sizeof(epsx_scgd0_params) */
    int64_t nT;
    const double *Tlist;      /* [nT] temperatures (energy units) */
    const double *nlist;      /* [nT] doping per cell */
    const double *mu_start;   /* [nT] starting μ (occ.μlist or the bare μ) */
    const double *eta_min;    /* [nT] clamp of Im Σ */
    double nelec, spin_degeneracy;
    int32_t is_semiconductor; /* occ.type === :Semiconductor */
    int32_t fix_mu;
    double eta_init, tol, w_acoustic;
    int64_t maxiter;
    int64_t f_tile;           /* inner states per Fan-Migdal block (0: automatic); like
                                 gpu_tile it only reorders the sum over f */
    int32_t fm_method;        /* EPSX_FM_* */
    int32_t precision;        /* EPSX_PREC_* */
    int32_t batched;          /* 1: all active temperatures per launch */
    int32_t verbosity;
    const double *Sigma_init; /* NULL (start at -iη_init) or complex [nw, ni, nT] */
    int32_t reuse_star;       /* 1: keep the star method's G̃ from the previous run on this
                                 problem when it was built with the same w_acoustic (0: rebuild) */
    int32_t warm_start;       /* 0, or the precision (EPSX_PREC_BF16 / _FP32) of the context's
                                 network (epsx_net_*): no counterpart in the reference.  Σ_init is
                                 predicted on the device from the G0D0 self-energy of the first
                                 iteration; the second iteration runs the plain and the predicted
                                 start side by side, and each temperature goes on from the one with
                                 the smaller residual (fewer iterations from tol), so the loop always
                                 runs to the same tol.  Needs batched = 1, Sigma_init = NULL,
                                 maxiter >= 3 and a network over the nω nodes of ωs. */
} epsx_scgd0_params;

typedef struct {
    int64_t struct_size;      /* sizeof(epsx_scgd0_outputs) */
    double *Sigma;            /* complex [nw, ni, nT] */
    double *mu;               /* [nT] */
    int32_t *converged;       /* [nT] */
    double *hist_err;         /* [nT, maxiter], NaN after a temperature froze */
    double *hist_mu;          /* [nT, maxiter] */
    double *hist_minA;        /* [nT, maxiter] min_i ∫A_i */
    double *wall;             /* [maxiter] seconds per iteration */
    int64_t *iter_frozen;     /* [nT], 0 if not converged */
    int64_t niter;            /* out: iterations run */
    int32_t fm_method_used;   /* out: EPSX_FM_* actually run */
    int32_t n_low_precision;  /* out: iterations run below FP64 */
    int32_t *warm_kept;       /* NULL, or [nT]: 1 where the warm start was kept */
} epsx_scgd0_outputs;

/* Upload the problem (replaces the previous one). */
int epsx_scgd0_set_problem(epsx_ctx *ctx, const epsx_scgd0_problem *prob);
/* _loop_scgd0, src/scgd0.jl:189-346, on the uploaded problem.  Returns EPSX_ERR_NO_MU_ROOT
 * with the reference's message when the μ bracket holds no root; a run that reaches maxiter
 * unconverged returns EPSX_OK with converged[T] = 0 (the reference warns). */
int epsx_scgd0_run(epsx_ctx *ctx, const epsx_scgd0_params *par, epsx_scgd0_outputs *out);

/* ---- ladder ---------------------------------------------------------------------------- */
/* What _setup_ladder adds to the scGD0 setup (src/ladder.jl:112-158): the rotation R[:, :, f] of
 * each inner state's vertex and the band velocities v[cart, i].  The shift tables (step
 * dω_unit of ωs, nshift = ceil(max ωph / dω_unit) + 2), the inner-state stars and the trapezoid
 * weights are derived by the library.  Needs epsx_scgd0_set_problem first (same states, grid
 * and coupling). */
typedef struct {
    int64_t struct_size;      /* sizeof(epsx_ladder_problem) */
    const double *R;          /* [3, 3, nf] */
    const double *v;          /* [3, ni] */
} epsx_ladder_problem;

#define EPSX_SOLVER_GMRES 0
#define EPSX_SOLVER_FIXED_POINT 1

typedef struct {
    int64_t struct_size;      /* sizeof(epsx_ladder_params) */
    int64_t nT, nOmega;
    const double *Tlist;      /* [nT] */
    const double *mu;         /* [nT] */
    const double *Sigma;      /* complex [nw, ni, nT]: the scGD0 self-energy */
    const double *Omegas;     /* [nOmega], >= 0 */
    int32_t solver;           /* EPSX_SOLVER_* */
    int32_t precision;        /* GEMM operands: EPSX_PREC_FP64, _FP32 or _BF16 (FP32 sums) */
    double rtol;
    int64_t maxiter;
    double w_acoustic;        /* > 0 */
    int32_t store_vertex;
    int32_t verbosity;
    int32_t batched;          /* GMRES only.  0: the (T, Ω) solves in turn, each warm-started from the
                                 previous Ω at its T (the reference's loop).  1: the Ωs of each T as
                                 instances of one batched solve, each from J = 0, one stacked
                                 shift-table product per Arnoldi step, converged instances dropped */
    int32_t pad;
} epsx_ladder_params;

typedef struct {
    int64_t struct_size;      /* sizeof(epsx_ladder_outputs) */
    double *sigma_bubble;     /* [3, 3, nOmega, nT], unsymmetrised, without spin/volume */
    double *sigma_ladder;     /* [3, 3, nOmega, nT] */
    int32_t *converged;       /* [nOmega, nT] */
    int64_t *niter;           /* [nOmega, nT] */
    double *norms;            /* [maxiter + 1, nOmega, nT] GMRES residual norms, NaN-padded */
    double *sigma_iter;       /* [3, 3, maxiter, nOmega, nT] fixed-point iterates, NaN-padded */
    double *J;                /* NULL, or complex [nw, 4, 3, 2, ni, nOmega, nT] */
} epsx_ladder_outputs;

int epsx_ladder_set_problem(epsx_ctx *ctx, const epsx_ladder_problem *prob);
/* _loop_ladder, src/ladder.jl:303-379: per (T, Ω) the propagators, the bare vertex and
 * σ_bubble, then (I - L) J = J0 by GMRES (warm-started from the previous Ω, stopping at
 * rtol ||J0||) or the fixed point, and σ_ladder.  An unconverged solve returns EPSX_OK with
 * converged = 0 (the reference warns). */
int epsx_ladder_run(epsx_ctx *ctx, const epsx_ladder_params *par, epsx_ladder_outputs *out);
/* apply_ladder_kernel!, src/ladder.jl:245-296, on host buffers: LJ = L(J) at temperature T for
 * the propagators Gm, Gp (complex [nw, 4, ni]); J, LJ complex [nw, 4, 3, 2, ni]. */
int epsx_apply_ladder_kernel(epsx_ctx *ctx, const double *J, const double *Gm, const double *Gp,
                             double T, double w_acoustic, int32_t precision, double *LJ);

/* ---- learned Σ_init warm start ------------------------------------------------------ */
/* No counterpart in the reference: its hook is the Σ_init argument of _loop_scgd0
 * (src/scgd0.jl:66-67, :204-212), which these entry points fill.  A 1D convolutional network
 * over the frequency grid of one outer state: n_layers 'same'-padded convolutions of odd width
 * `kernel`, c_in -> c_hidden -> ... -> c_out channels, SiLU between layers.  Samples are
 * float [n, channels, length] (C order).  Parameters: per layer W [cout, cin, kernel] then
 * b [cout], FP32.  precision: EPSX_PREC_BF16 (operands rounded to BF16, FP32 accumulation)
 * or EPSX_PREC_FP32. */
typedef struct {
    int64_t struct_size;      /* sizeof(epsx_net_config) */
    int32_t n_layers;         /* 2..16 */
    int32_t kernel;           /* odd */
    int32_t c_in, c_hidden, c_out;
    int32_t pad;
    int64_t length;           /* grid points per sample */
} epsx_net_config;

typedef struct {
    int64_t struct_size;      /* sizeof(epsx_net_train_params) */
    int64_t epochs, batch;
    double lr, beta1, beta2, eps, weight_decay;   /* Adam with decoupled weight decay */
    int32_t precision;
    int32_t verbosity;
    uint64_t seed;            /* shuffling */
} epsx_net_train_params;

/* A new network with seeded Kaiming-uniform weights (last layer x 0.1) and zero biases. */
int epsx_net_create(epsx_ctx *ctx, const epsx_net_config *cfg, uint64_t seed);
int epsx_net_param_count(epsx_ctx *ctx, int64_t *count);
int epsx_net_get_params(epsx_ctx *ctx, float *params);
/* Sets the weights and clears the optimiser state. */
int epsx_net_set_params(epsx_ctx *ctx, const float *params);
/* y [n, c_out, length] = net(x [n, c_in, length]). */
int epsx_net_forward(epsx_ctx *ctx, const float *x, int64_t n, int32_t precision, float *y);
/* Mean squared error against target [n, c_out, length] and its gradient (param_count floats). */
int epsx_net_gradient(epsx_ctx *ctx, const float *x, const float *target, int64_t n, int32_t precision,
                      double *loss, float *grad);
/* Adam over shuffled mini-batches of the (device-resident) training set; epoch_loss[epochs]
 * receives the mean batch loss of each epoch. */
int epsx_net_train(epsx_ctx *ctx, const float *x, const float *target, int64_t n,
                   const epsx_net_train_params *params, double *epoch_loss);

#ifdef __cplusplus
}
#endif
#endif
