// Internal declarations of libepsx: the context, error reporting, device buffers and the
// kernels shared between translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <cstdio>
#include <string>
#include <vector>

#include "epsx.h"

namespace epsx {

// A device buffer owned by the context (freed with it or on reallocation).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    epsx_ctx *owner = nullptr;  // the context whose device memory count holds it
    bool tmp = false;           // a run's temporary: freeing returns it to the owner's pool
};

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
};

// Device-resident scGD0 problem (see epsx_scgd0_set_problem).
struct Scgd0Dev {
    bool ready = false;
    int64_t nm = 0, ni = 0, nf = 0, nq = 0, nki = 0, nkf = 0, nw = 0, nwa = 0, n_mult = 0;
    double W = 0, nstates_base = 0;
    // grid of ωs: relative integer offsets (from offsets[0]) and the node values
    double w_origin = 0, dw_unit = 0, dw_min = 0;
    int64_t span = 0;                 // offsets[n-1] - offsets[0]
    std::vector<double> ws_host, wsa_host, w_i_host;
    bool einstein = false;            // ωph[ν, :] constant for every ν
    bool star_ok = false;             // e_f == e_{f_to_i[f]} bitwise
    std::vector<double> wnu_host;     // [nm] mode frequency when einstein
    int64_t nstar = 0;                // number of outer states that have inner states
    // device arrays
    DevBuf g2p;        // [nm, nf (star-sorted), ni]  FP64
    DevBuf g2p32;      // same, FP32 (built on demand)
    DevBuf wph, iq_kk, iks_i, iks_f;  // phonon table and q map
    DevBuf ef_s, wf_s, fi_s, fsrc_s;  // per sorted inner state: ε, w, i_f (0-based), source f
    DevBuf e_i, w_i;
    DevBuf ws, off, rgap, ltab;       // output grid: nodes, rel. offsets (int32), 1/gap, cell table
    DevBuf wsa, Wa, jsp, tsp;         // spectral grid: nodes, trapezoid weights, table from ωs
    DevBuf KK;                        // [nw, nw]
    DevBuf KKo, KKd;                  // KK without its diagonal, and the diagonal
    DevBuf KKp;                       // KK with the even column pitch nw + (nw & 1), pad rows zero
    DevBuf mult_ptr, mult_idx;
    DevBuf star_ptr;                  // [ni + 1] sorted-position ranges of each star
    DevBuf Gt;                        // star-collapsed coupling G̃ [(ν, i'), i] (built on demand)
    bool gt_ready = false;            // G̃ holds the current problem's reduction ...
    double gt_wac = 0;                // ... with modes below this ω_acoustic skipped
    // Column classes of the star product: one per multiplet, one per other outer state (nc of
    // them; 0: none, the product runs over all ni columns).  G̃c holds G̃ with the columns of each
    // multiplet averaged, so the product's columns are the multiplet averages themselves.
    int64_t nc = 0;
    DevBuf cls, cls_ptr, cls_idx;     // class of each state; members of each class (CSR, 0-based)
    DevBuf Gtc;                       // [(ν, i'), class]
    DevBuf Gb, Gf;                    // G̃ (G̃c) in BF16 / FP32, k-contiguous rows of pitch kp (on demand)
    bool gb_ready = false, gf_ready = false;   // they hold the current G̃
    DevBuf KKb, KKf;                  // off-diagonal KK transposed in BF16 / FP32, pitch nw rounded to 8
    bool kkb_ready = false, kkf_ready = false;
    DevBuf chunks, skip, wnu;         // star chunks of the reduction, acoustic flags, ω_ν
    DevBuf hk, ht;                    // cells and weights of ω ± ω_ν (Einstein modes)
    int64_t nchunks = 0;
    int sr_cap = 0;                   // stage size of the reduction (doubles; 0: none)
    int64_t ltab_len = 0;
};

// Device-resident ladder problem on top of the scGD0 problem (epsx_ladder_set_problem).
struct LadderDev {
    bool ready = false;
    int64_t ns = 0;                   // shift tables of step dω_unit
    double h = 0;
    std::vector<double> v_host;       // [3, ni]
    DevBuf Rs, v, Ww, klo, wlo, whi;  // rotations in star order, velocities, weights, shift tables
    DevBuf Bt;                        // B̃ of the current temperature
    double bt_T = -1, bt_wac = -1;
    DevBuf Btl;                       // B̃ regrouped by column blocks (every shift of a block contiguous), in
                                      // FP64 or BF16 / FP32 with rows of pitch 3 ni rounded to 8 (on demand)
    int32_t btl_prec = -1;            // the precision Btl holds for the current B̃ (-1: none) ...
    int64_t btl_cb = 0;               // ... and its block width
    // propagators and small work buffers (the application's large buffers belong to each call)
    DevBuf Gm, Gp, S, jm, tm, jp, tp, Yc, red, zero_idx;
};

constexpr int NET_MAX_LAYERS = 16;

// The warm-start network: configuration, FP32 master weights, Adam moments and activations.
struct NetDev {
    bool ready = false;
    int layers = 0, kernel = 0, c_in = 0, c_hidden = 0, c_out = 0;
    int64_t length = 0, nparam = 0, cap_nb = 0, step = 0;
    DevBuf params, grad, m, v, Wt, loss, bsum, part;
    DevBuf X[NET_MAX_LAYERS + 1], Z[NET_MAX_LAYERS];
    DevBuf Y, dA, dB, T, idx;
    DevBuf Wp;   // BF16 weights of the inference convolution, tap-major, per layer
    bool wp_ok = false;   // Wp holds the current params (cleared wherever params change)
};

}  // namespace epsx

struct epsx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    epsx::Scgd0Dev s;
    epsx::LadderDev l;
    epsx::NetDev net;
    std::vector<epsx::DevBuf *> owned;
    epsx::DevBuf gemm_ws;   // split-k partial products
    void *blas = nullptr;   // rocBLAS handle on `stream` (created on first use), for plain GEMMs
    void *blaslt = nullptr; // hipBLASLt handle and its chosen algorithms (epsx_gemm.hip)
    void *pin = nullptr;    // pinned host staging of read_back (kept for the next call)
    size_t pin_bytes = 0;
    hipStream_t side = nullptr;   // a second stream (default priority) for copies beside the work
    void *pin_side = nullptr;     // pinned host memory the side stream copies into
    // the replayed loop's host slots, kept for the next run: mask words and err (mapped, coherent:
    // the convergence kernel writes them) and the staged precision levels
    void *pin_slots = nullptr;
    size_t pin_slots_bytes = 0;
    void *pin_prec = nullptr;
    size_t pin_prec_bytes = 0;
    void *pin_book = nullptr;     // the loop's bookkeeping (frozen, histories, μ) staged both ways
    size_t pin_book_bytes = 0;
    hipEvent_t run_ev[8] = {};    // the scGD0 run's timing events (created once, kept)
    size_t pin_side_bytes = 0;
    // device buffers a run released, by size, for the next run's temporaries (work on the
    // context's one stream is ordered, so a buffer may be handed out while a released user's
    // kernels are still queued ahead)
    std::multimap<size_t, void *> pool;
    size_t pool_bytes = 0;
    // device memory the context holds (its buffers, temporaries and pool), and an optional cap on
    // it (EPSX_ALLOC_CAP_MB, read at epsx_open; 0: none) under which an allocation fails as if
    // the device were full, so that the out-of-memory path can be exercised on any device
    size_t dev_bytes = 0;
    size_t alloc_cap = 0;
    // this device's copy of the fixed workspace of epsx_kramers_kronig_noalloc (epsx_kk.hip),
    // looked up when the context opens
    double *kk_scratch = nullptr;
    std::map<std::array<int64_t, 5>, int> gemm_tune;   // (m, n, k, instances, caller workspace or -1) -> tile shape
    int last_gemm_shape = -1;   // the tile shape the last FP64 dgemm launched (after any fallback)
    // kernel time of the last run by class: 0 Fan-Migdal, 1 Kramers-Kronig, 2 DOS + μ,
    // 3 convergence + copies, 4 coupling reduction (star method), 5 star GEMMs, 6 their instances,
    // 7 the kernel of the star method's product in the timed iteration (count only)
    double t_class[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t n_class[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

namespace epsx {

int set_error(epsx_ctx *ctx, int code, const char *fmt, ...);
void blas_destroy(epsx_ctx *ctx);   // epsx_ladder.hip
void blaslt_destroy(epsx_ctx *ctx); // epsx_gemm.hip
int hip_fail(epsx_ctx *ctx, hipError_t e, const char *what, const char *file, int line);
int dev_alloc(epsx_ctx *ctx, DevBuf &b, size_t bytes);   // context-owned, freed at close
void dev_free(DevBuf &b);   // hipFree, or back to its context's pool (a run's temporary)
void pool_clear(epsx_ctx *ctx);
// hipMalloc for the context: when it fails (or the cap is reached) and the pool holds released
// buffers, the work on both streams is waited for, the pool is emptied and the allocation tried
// once more.
int ctx_malloc(epsx_ctx *ctx, void **p, size_t bytes);
int tmp_alloc(epsx_ctx *ctx, DevBuf &b, size_t bytes);   // caller-owned (see Tmp)

// A scoped device temporary.
struct Tmp {
    DevBuf b;
    Tmp() = default;
    Tmp(const Tmp &) = delete;
    Tmp &operator=(const Tmp &) = delete;
    ~Tmp() { dev_free(b); }
    template <class T> T *ptr() { return static_cast<T *>(b.p); }
};
int tmp_upload(epsx_ctx *ctx, Tmp &t, const void *host, size_t bytes);

#define EPSX_HIP(ctx, call)                                                    \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) return epsx::hip_fail((ctx), e_, #call, __FILE__, __LINE__); \
    } while (0)

#define EPSX_TRY(call)              \
    do {                            \
        int r_ = (call);            \
        if (r_ != EPSX_OK) return r_; \
    } while (0)

template <class T>
inline T *dptr(DevBuf &b) { return static_cast<T *>(b.p); }

// Upload a host array into a (re)allocated device buffer.
int upload(epsx_ctx *ctx, DevBuf &b, const void *host, size_t bytes);

// Copy device memory into any host array after the work on the context's stream: through a pinned
// staging buffer (at most 64 MiB per pass, kept in the context), then host threads copy out of it
// (a pageable destination is copied at host memory speed, its first-touch page faults spread over
// the threads).
int read_back(epsx_ctx *ctx, void *dst, const void *src, size_t bytes);
// The reverse of read_back: any host array into device memory through the pinned staging buffer
// (host threads copy into it), ordered after the work already on the context's stream.
int write_from_host(epsx_ctx *ctx, void *dst, const void *src, size_t bytes);
// memcpy by up to 8 host threads (slices of at least 1 MiB).
void parallel_copy(void *dst, const void *src, size_t n);

// Host-side grid maps (epsx_grids.cpp), shared by the C ABI and the loop set-up.
int64_t cell_search_table(int64_t n, const double *ws, double dw_min, double *x0,
                          double *inv_h, int32_t *jtab, int64_t cap);
void linear_interpolation_table(int64_t n, const double *ws, int64_t m, const double *xs,
                                int32_t *j, double *t);
void trapezoid_weights(int64_t n, const double *ws, double *W);
void grid_nodes(int64_t n, double origin, double unit, const int64_t *offsets, double *ws);

// Kernel launchers (epsx_kk.hip)
int kk_build(epsx_ctx *ctx, int64_t n, const double *d_ws, double *d_K);
int kk_split(epsx_ctx *ctx, int64_t n, const double *d_K, double *d_Ko, double *d_Kd);
// the device address of the no-alloc workspace on the context's device (called by epsx_open)
int kk_scratch_init(epsx_ctx *ctx);
// the off-diagonal KK transposed into a BF16 / FP32 k-contiguous operand of row pitch np
int kk_lp_build(epsx_ctx *ctx, int precision, int64_t n, int64_t np, const double *d_Ko, void *out);
// columns of the gated slots in BF16 / FP32 with rows padded to np (zero beyond nrow)
int lp_cols(epsx_ctx *ctx, int precision, int64_t nrow, int64_t np, int64_t ncol, const double *src, int64_t ld,
            int64_t sstride, void *dst, int64_t dstride, int nslots, const int32_t *tidx, const int32_t *gate);

// Kernel time of a stretch of the stream added to ctx->t_class[k] (HIP events; stop() waits for
// the end event).  Not for use under graph capture.
struct ClassTimer {
    epsx_ctx *ctx;
    int k;
    hipEvent_t a = nullptr, b = nullptr;
    ClassTimer(epsx_ctx *c, int cls) : ctx(c), k(cls)
    {
        if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) (void)hipEventRecord(a, c->stream);
    }
    int stop(int64_t launches)
    {
        float ms = 0;
        EPSX_HIP(ctx, hipEventRecord(b, ctx->stream));
        EPSX_HIP(ctx, hipEventSynchronize(b));
        EPSX_HIP(ctx, hipEventElapsedTime(&ms, a, b));
        ctx->t_class[k] += ms * 1e-3;
        ctx->n_class[k] += launches;
        return EPSX_OK;
    }
    ~ClassTimer()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// A batch of GEMM instances (epsx_gemm.hip).  Instance z uses A + sA a, B + sB b, C + sC c,
// where a, b, c are z, or idx[z] for the operands flagged in imap (bit 0 A, bit 1 B, bit 2 C).
// gate (device, indexed by idx[z] or z): 0 skips the instance.  Z/Zim: write
// Z = complex(result, Zim) at C's positions instead of C.
struct GemmBatch {
    int64_t n_inst = 1;
    int64_t sA = 0, sB = 0, sC = 0;
    const int32_t *idx = nullptr;
    int imap = 0;
    const int32_t *gate = nullptr;
    double2 *Z = nullptr;
    const double *Zim = nullptr;
    int precision = EPSX_PREC_FP64;   // FP32 / BF16: operands rounded when staged, FP32 accumulation
    const double *diag = nullptr;     // with Z: add diag[i] Zim[i, j] in FP64
    bool tune = true;                 // the autotuner may time this product (its gate is not hiding work)
    hipEvent_t mark = nullptr;        // recorded after the main kernel (before the split-k sum)
    double *part = nullptr;           // split-k partials in this caller-provided device buffer of
    int64_t part_cap = 0;             // part_cap doubles (none: the context's workspace); a split
                                      // whose partials do not fit runs unsplit
    float *Cf = nullptr;              // lpgemm_tn only: FP32 output in place of C (no accumulate, no Z)
};
// C[m, n] = A[m, k] * B[k, n] (+ C if accumulate), FP64, column-major, on ctx->stream
int dgemm(epsx_ctx *ctx, int64_t m, int64_t n, int64_t k, const double *A, int64_t lda, const double *B,
          int64_t ldb, double *C, int64_t ldc, bool accumulate, const GemmBatch &bt);
// The same product from low-precision operands stored k-contiguous (precision EPSX_PREC_BF16:
// uint16 BF16 bits, EPSX_PREC_FP32: float; AT[k, m] and B[k, n], row pitches, strides and k
// multiples of 8, zero-padded beyond the true k), FP32 accumulation, FP64 output.
// C_z = op(A_z) B_z for z < batch through hipBLASLt with a workspace-free algorithm (no split-k
// workspace, no atomics: repeats give the same bits).  FP64: C = A B, column-major (strides in
// elements, sA or sB may be 0).  BF16 / FP32: C (FP32) = Aᵀ B from k-contiguous operands (A k x m,
// B k x n), FP32 accumulation.  blaslt_prepare, outside graph capture, times the heuristic's
// workspace-free candidates on the given buffers (C overwritten) for batch counts 1..max_batch and
// keeps the fastest per problem and batch count in the context; blaslt_gemm then only launches,
// splitting its batch into the chunks those timings price cheapest.
int blaslt_prepare(epsx_ctx *ctx, int prec, int64_t m, int64_t n, int64_t k, const void *A, int64_t lda, int64_t sA,
                   const void *B, int64_t ldb, int64_t sB, void *C, int64_t ldc, int64_t sC, int64_t max_batch, bool *ok);
int blaslt_gemm(epsx_ctx *ctx, int prec, int64_t m, int64_t n, int64_t k, const void *A, int64_t lda, int64_t sA,
                const void *B, int64_t ldb, int64_t sB, void *C, int64_t ldc, int64_t sC, int64_t batch);
int lpgemm_tn(epsx_ctx *ctx, int precision, int64_t m, int64_t n, int64_t k, const void *AT, int64_t lda,
              const void *B, int64_t ldb, double *C, int64_t ldc, bool accumulate, const GemmBatch &bt);
int dgemm_nn(epsx_ctx *ctx, int64_t m, int64_t ncol, int64_t k, const double *A, int64_t lda,
             const double *B, int64_t ldb, double *C, int64_t ldc, bool accumulate,
             const int32_t *gate = nullptr);   // gate: device flag, 0 = skip the product

// Grid data the Fan-Migdal kernels read (device pointers).
struct FmGrid {
    int64_t nw = 0, span = 0, ltab_len = 0;
    double dw_unit = 0;
    const double *ws = nullptr;     // nodes
    const int32_t *off = nullptr;   // relative offsets
    const double *rgap = nullptr;   // 1 / (off[k+1] - off[k])
    const int32_t *ltab = nullptr;  // cell of each relative offset in [0, span]
};

// Inputs of one direct Fan-Migdal launch (device pointers).
struct FmDirectArgs {
    int64_t nm, ni, nfs;            // modes, outer states, sorted inner states
    const double *g2p;              // [nm, nfs, ni] coupling, inner states star-sorted
    const float *g2p32;             // FP32 copy (precision 1) or null
    const double *ef, *wf;          // [nfs] per sorted inner state
    const int32_t *fi;              // [nfs] 0-based column of Sin (IBZ representative)
    const int32_t *fsrc;            // [nfs] 0-based inner state (for the q gather)
    // phonon frequency of (ν, i, f): wq_pair[ν + nm (i + ni f)] if non-null, else
    // wph[ν + nm (iq_kk[(iks_i[i]-1) + nki (iks_f[f]-1)] - 1)]
    const double *wq_pair;
    const double *wph;
    const int32_t *iq_kk, *iks_i, *iks_f;
    int64_t nki;
    const double *S;                // complex [nw, ncol, nT] unclamped self-energy
    int64_t ncol;
    double w_acoustic;
};

// The temperatures of one launch: slot z of the launch runs temperature index tidx[z]; T, mu
// and eta_min are indexed by the temperature index.
struct FmTemps {
    int nT;                         // launched slots
    const double *T;                // device [nT_all]
    const double *mu;               // device [nT_all]
    const double *eta_min;          // device [nT_all] (clamp); may be null (no clamp)
    const int32_t *tidx;            // device [nT] temperature index of each slot
    const int32_t *act = nullptr;   // device [nT_all] convergence mask (null: all active)
};

// One direct Fan-Migdal pass over the launched temperatures: ImS[:, :, tidx[z]] (+)= the sum.
int fm_direct_launch_sp(epsx_ctx *ctx, const FmGrid &g, const FmDirectArgs &a, const FmTemps &tp,
                        const int64_t *d_star_ptr, int64_t f_tile, int precision, double *d_part,
                        int64_t part_cap, double *d_ImS, bool accumulate);
// factorised (star) Fan-Migdal method, epsx_fm_star.hip
int star_stage_cap(int64_t nm, const std::vector<int64_t> &star_ptr);
void star_chunks(int64_t nm, int cap, const std::vector<int64_t> &star_ptr, std::vector<int32_t> &chunk);
int star_reduce_launch(epsx_ctx *ctx, int cap, int64_t nm, int64_t ni, int64_t nfs, const double *g2p,
                       const double *wf, const int64_t *d_star_ptr, const int32_t *d_chunks, int64_t nchunks,
                       const uint8_t *d_skip, double *Gt);
void star_lookup_host(int64_t nw, int64_t nm, const double *wnu, double h, const std::vector<int32_t> &off,
                      const std::vector<int32_t> &ltab, const std::vector<double> &rgap, int64_t span,
                      std::vector<int32_t> &hk, std::vector<double> &ht);
int h_build_launch(epsx_ctx *ctx, int64_t nw, int64_t nm, int64_t ni, const double *ws, const int32_t *hk,
                   const double *ht, const double *wnu, const double *e_i, const double *S, const double *T,
                   const double *mu, const double *eta, const int32_t *tidx, int nslots, double w_acoustic,
                   double *occ, double *hout, const int32_t *act, const int32_t *act_h, uint16_t *hb,
                   const int32_t *act_b, float *hf, const int32_t *act_f, int64_t kp, int64_t ldh);
int gt_lp_launch(epsx_ctx *ctx, int precision, int64_t K, int64_t ldg, int64_t kp, int64_t ni, const double *Gt, void *gl);
// the Einstein factorisation's H'[ω, ν + nm (p - p0)] = w_f H_ν[ω, f] for the sorted inner states
// [p0, p0 + np) of every slot (occ from h_build_launch), and the check that no mode below
// ω_acoustic holds a non-finite coupling
int einstein_h_launch(epsx_ctx *ctx, int64_t nw, int64_t nm, int64_t ni, const double *ws, const int32_t *hk, const double *ht,
                      const double *wnu, const double *S, const double *eta, const int32_t *tidx, int nslots,
                      double w_acoustic, const double *occ, int64_t p0, int64_t np, const double *ef, const double *wf,
                      const int32_t *fi, double *hout, int64_t ldh, const int32_t *act);
int acoustic_finite_check(epsx_ctx *ctx, int64_t nm, int64_t n, const double *g2p, const uint8_t *skip, bool *ok);
// the warm-start network's Σ[:, :, tdst[z]] predicted from the G0D0 Σ[:, :, tsrc[z]], z < nz (epsx_net.hip)
int net_predict_sigma(epsx_ctx *ctx, int64_t nw, int64_t ni, int nz, const int32_t *tsrc, const int32_t *tdst,
                      double2 *S, const double *ws, const double *e_i, const double *mu, const double *T, int precision);
size_t fm_part_elems(int64_t nfs, int64_t ni, int64_t nw, int64_t nT, int64_t f_tile);
int fm_grid_host(int64_t nw, const int64_t *offsets, std::vector<int32_t> &off,
                 std::vector<double> &rgap, std::vector<int32_t> &ltab, int64_t &span);
void star_order(int64_t nf, int64_t ncol, const int32_t *f_to_i, std::vector<int32_t> &fsrc,
                std::vector<int64_t> &star_ptr);

}  // namespace epsx
