// Host-side grid maps of libepsx, bit-exact with the reference's builders (integer tables
// and the floating-point expressions evaluated in the reference's order, no contraction).
#include <cmath>
#include <cstdlib>
#include <vector>

#include "epsx_common.h"

#pragma clang fp contract(off)

namespace epsx {

// Node j of a FrequencyGrid: ω_origin + offsets[j] * dω_unit (src/frequency_grids.jl:138-141).
void grid_nodes(int64_t n, double origin, double unit, const int64_t *offsets, double *ws)
{
    for (int64_t j = 0; j < n; j++) ws[j] = origin + (double)offsets[j] * unit;
}

// cell_search_table (src/frequency_grids.jl:184-196): the bucket of x is
// trunc((x - x0) * inv_h) with inv_h = 1 / ((1 - 1e-6) dω_min); jtab[b] counts the nodes in
// earlier buckets.  Returns the table length, -1 above 2^24 entries, -2 for two nodes in one
// bucket, -3 if cap is too small.
int64_t cell_search_table(int64_t n, const double *ws, double dw_min, double *x0, double *inv_h,
                          int32_t *jtab, int64_t cap)
{
    const double a = ws[0];
    const double ih = 1.0 / ((1 - 1e-6) * dw_min);
    double nb = (ws[n - 1] - a) * ih;
    if (nb > 1099511627776.0) nb = 1099511627776.0;  // min(., 2^40)
    const int64_t nbuckets = (int64_t)std::trunc(nb) + 1;
    *x0 = a;
    *inv_h = ih;
    if (nbuckets > (int64_t(1) << 24)) return -1;
    std::vector<int64_t> b(n);
    for (int64_t j = 0; j < n; j++) b[j] = (int64_t)((ws[j] - a) * ih);
    for (int64_t j = 1; j < n; j++)
        if (b[j] <= b[j - 1]) return -2;
    const int64_t len = b[n - 1] + 1;
    if (!jtab) return len;
    if (len > cap) return -3;
    int64_t q = 0;
    for (int64_t bkt = 0; bkt < len; bkt++) {
        while (q < n && b[q] < bkt) q++;
        jtab[bkt] = (int32_t)q;
    }
    return len;
}

// searchsortedlast: the number of entries <= x
static int64_t count_le(int64_t n, const double *v, double x)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo + 1) / 2;
        if (v[mid - 1] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// linear_interpolation_table (src/frequency_grids.jl:65-84).
void linear_interpolation_table(int64_t n, const double *ws, int64_t m, const double *xs,
                                int32_t *j, double *t)
{
    for (int64_t r = 0; r < m; r++) {
        const double x = xs[r];
        if (x <= ws[0]) {
            j[r] = 1;
            t[r] = 0;
        } else if (x >= ws[n - 1]) {
            j[r] = (int32_t)(n - 1);
            t[r] = 1;
        } else {
            const int64_t jr = count_le(n, ws, x);
            j[r] = (int32_t)jr;
            t[r] = (x - ws[jr - 1]) / (ws[jr] - ws[jr - 1]);
        }
    }
}

// trapezoid_weights (src/frequency_grids.jl:53-54).
void trapezoid_weights(int64_t n, const double *ws, double *W)
{
    for (int64_t j = 0; j < n; j++) {
        const int64_t a = j + 1 < n ? j + 1 : n - 1, b = j > 0 ? j - 1 : 0;
        W[j] = (ws[a] - ws[b]) / 2;
    }
}

}  // namespace epsx

extern "C" {

int epsx_cell_search_table(epsx_ctx *ctx, int64_t n, const double *ws, double dw_min,
                           int32_t *jtab, int64_t cap, int64_t *len, double *x0, double *inv_h)
{
    if (n < 2 || !ws || !len || !x0 || !inv_h)
        return epsx::set_error(ctx, EPSX_ERR_ARG, "cell_search_table needs at least two nodes");
    int64_t r = epsx::cell_search_table(n, ws, dw_min, x0, inv_h, jtab, cap);
    if (r == -1) {
        double nb = (ws[n - 1] - ws[0]) * *inv_h;
        return epsx::set_error(ctx, EPSX_ERR_ARG,
                               "cell search table of %lld entries for a grid with minimum spacing "
                               "%.17g over a range %.17g; use a grid without a near-coincident "
                               "pair of nodes",
                               (long long)((int64_t)std::trunc(std::fmin(nb, 1099511627776.0)) + 1),
                               dw_min, ws[n - 1] - ws[0]);
    }
    if (r == -2) return epsx::set_error(ctx, EPSX_ERR_ARG, "two grid nodes fall in one bucket");
    if (r == -3) return epsx::set_error(ctx, EPSX_ERR_ARG, "jtab capacity too small");
    *len = r;
    return EPSX_OK;
}

int epsx_linear_interpolation_table(epsx_ctx *ctx, int64_t n, const double *ws, int64_t m,
                                    const double *xs, int32_t *j, double *t)
{
    if (n < 2 || m < 0 || !ws || (m && (!xs || !j || !t)))
        return epsx::set_error(ctx, EPSX_ERR_ARG, "linear_interpolation_table: bad arguments");
    epsx::linear_interpolation_table(n, ws, m, xs, j, t);
    return EPSX_OK;
}

int epsx_trapezoid_weights(epsx_ctx *ctx, int64_t n, const double *ws, double *W)
{
    if (n < 1 || !ws || !W) return epsx::set_error(ctx, EPSX_ERR_ARG, "trapezoid_weights: bad arguments");
    epsx::trapezoid_weights(n, ws, W);
    return EPSX_OK;
}

// vertex_shift_tables (src/ladder_kernel.jl:158-176): for s = up (+1), down (-1), shift m-1
// steps of the integer offsets, the clamped cell of p and the exact-rational weight.
int epsx_vertex_shift_tables(epsx_ctx *ctx, int64_t nw, const int64_t *off, int64_t nshift,
                             int32_t *klo, double *wlo, double *whi)
{
    if (nw < 2 || nshift < 1 || !off || !klo || !wlo || !whi)
        return epsx::set_error(ctx, EPSX_ERR_ARG, "vertex_shift_tables: bad arguments");
    for (int s = 0; s < 2; s++) {
        const int64_t sign = s == 0 ? 1 : -1;
        for (int64_t m = 0; m < nshift; m++)
            for (int64_t j = 0; j < nw; j++) {
                const int64_t p = off[j] + sign * m;
                // searchsortedlast(offsets, p) clamped to 1..nw-1
                int64_t lo = 0, hi = nw;
                while (lo < hi) {
                    int64_t mid = lo + (hi - lo + 1) / 2;
                    if (off[mid - 1] <= p) lo = mid;
                    else hi = mid - 1;
                }
                int64_t k = lo < 1 ? 1 : (lo > nw - 1 ? nw - 1 : lo);
                const double t = (double)(p - off[k - 1]) / (double)(off[k] - off[k - 1]);
                const bool inside = off[0] <= p && p <= off[nw - 1];
                const size_t q = (size_t)(j + nw * (m + nshift * s));
                klo[q] = (int32_t)k;
                wlo[q] = inside ? 1 - t : 0;
                whi[q] = inside ? t : 0;
            }
    }
    return EPSX_OK;
}

}  // extern "C"
