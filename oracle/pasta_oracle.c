/*
 * pasta_oracle.c -- CPU oracle for the PASTA model-evaluation path.
 *
 * TEST INFRASTRUCTURE, NOT THE PRODUCT.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product path (the HIP
 * library under pasta_amd/) never links or calls this file.
 *
 * A plain-C restatement of the reference Rust crate (paths relative to the
 * reference repository).  Build: oracle/Makefile (gcc -O2 -ffp-contract=off,
 * so no multiply-add is contracted, like rustc's default).
 *
 * Parity status (what pins each part; the tests named live in tests/):
 *   grid + neighbours + Lagrange weights ....... pinned: closed-form Lagrange
 *       weights, partition of unity, exact reproduction of cubic/quadratic
 *       polynomials, one-hot at nodes, brute-force staircase neighbours
 *       (test_oracle_grid.py).
 *   interpolation incl. tail quirk ............. pinned: polynomial-in-label
 *       grid reproduced exactly; tail pixels scaled by the extra 1/65535
 *       (test_oracle_interp.py).  gemv accumulation order restated from
 *       nalgebra 0.32.6 gemv/axcpy (column by column): order unpinned at ULP.
 *   rotation kernel, Gaussian kernel ........... pinned: closed-form profile,
 *       unit sum, symmetry, tap-count formula.
 *   overlap-add FFT convolution ................ pinned against direct FP64
 *       convolution (numpy.convolve) to 1e-6 relative; the FFT's own rounding
 *       (realfft 3.3 / rustfft 6.2) is unpinned at ULP: this file uses its own
 *       radix-2 f32 FFT.
 *   RV shift + resample (f32-quantised weight) . pinned: closed-form linear
 *       interpolation, weight quantisation reproduced exactly.
 *   chunk / linear / fixed / constant continuum  pinned: LS projection against
 *       numpy.linalg.lstsq (FP64), chunk bounds and blend factors against a
 *       numpy restatement, chi2 closed form.  nalgebra's f32 SVD solve is
 *       unpinned at ULP: this file solves the same least-squares problem by
 *       FP64 Householder QR and rounds the solution to f32.
 *   PSO + bounds ............................... pinned by invariants (bounds,
 *       monotone global best, evaluation count, population injection,
 *       convergence on a quadratic bowl).  The reference seeds its RNG from
 *       entropy (particleswarm.rs:65), so the random stream here is this
 *       build's own counter-based SplitMix64, shared bit-for-bit with the GPU
 *       path.
 *   Brent root / uncertainty ................... pinned by closed-form roots
 *       (test_oracle_uncertainty.py).
 *   linear interpolation (interpolate_linear) .. pinned: trilinear-in-label
 *       grid reproduced exactly, one-hot at nodes, weights interpolate the
 *       labels themselves (test_oracle_binary.py).
 *   binary model / costs ........................ pinned: identical stars or a
 *       zero light ratio give produce_model / the single-star cost; FP64 numpy
 *       composition with np.interp resampling; the time-series cost is the sum
 *       of binary costs.  nalgebra's DVector::mean is restated as a sequential
 *       f32 fold and one division: unpinned at ULP.
 *   binary searches (11 / 9 / 2 labels) ......... pinned by the PSO invariants
 *       above on the binary bounds.
 */
#include "pasta_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define FFTSIZE 2048     /* convolve_rv.rs:433 */
#define BATCH_SIZE 128   /* interpolate.rs:917 */
#define C_KMS 299792.0   /* convolve_rv.rs:440,547 */
#define LN_10 2.302585092994045684017991454684364208

/* ======================================================================
 * Grid geometry -- interpolate.rs:58-670
 * ====================================================================== */

/* Range::get_right_index (interpolate.rs:73-75): binary search; returns
 * index >= 0 on exact match, or -(insertion point)-1 when absent. */
static int right_index(const double *v, int n, double x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (v[mid] < x) lo = mid + 1;
        else if (v[mid] > x) hi = mid;
        else return mid;
    }
    return -lo - 1;
}

static int insert_sorted(double *v, int n, double x) {
    for (int i = 0; i < n; i++)
        if (v[i] == x) return n;
    int pos = right_index(v, n, x);
    pos = -pos - 1;
    memmove(v + pos + 1, v + pos, sizeof(double) * (size_t)(n - pos));
    v[pos] = x;
    return n + 1;
}

/* Grid::new (interpolate.rs:267-331) */
or_grid *or_grid_from_labels(const double *labels, int n) {
    if (n <= 0) return NULL;
    double *t = malloc(sizeof(double) * n), *m = malloc(sizeof(double) * n),
           *l = malloc(sizeof(double) * n);
    int nt = 0, nm = 0, nl = 0;
    for (int i = 0; i < n; i++) {
        nt = insert_sorted(t, nt, labels[3 * i + 0]);
        nm = insert_sorted(m, nm, labels[3 * i + 1]);
        nl = insert_sorted(l, nl, labels[3 * i + 2]);
    }
    or_grid *g = calloc(1, sizeof(or_grid));
    g->n_teff = nt; g->n_m = nm; g->n_logg = nl;
    g->teff = t; g->m = m; g->logg = l;
    g->logg_lo = malloc(sizeof(int) * nt);
    g->logg_hi = malloc(sizeof(int) * nt);
    g->cum = malloc(sizeof(int) * nt);
    for (int i = 0; i < nt; i++) { g->logg_lo[i] = nl - 1; g->logg_hi[i] = 0; }
    for (int q = 0; q < n; q++) {
        int i = right_index(t, nt, labels[3 * q + 0]);
        int j = right_index(l, nl, labels[3 * q + 2]);
        if (j < g->logg_lo[i]) g->logg_lo[i] = j;
        if (j > g->logg_hi[i]) g->logg_hi[i] = j;
    }
    int acc = 0;
    for (int i = 0; i < nt; i++) {
        if (g->logg_lo[i] > g->logg_hi[i]) { or_grid_free(g); return NULL; }
        g->cum[i] = acc;
        acc += g->logg_hi[i] - g->logg_lo[i] + 1;
    }
    g->n_models = acc * nm;
    return g;
}

void or_grid_free(or_grid *g) {
    if (!g) return;
    free(g->teff); free(g->m); free(g->logg);
    free(g->logg_lo); free(g->logg_hi); free(g->cum);
    free(g);
}

/* InMemFetcher::find_spectrum index formula (model_fetchers.rs:234-236) */
int or_grid_model_index(const or_grid *g, int i, int j, int k) {
    return (g->cum[i] + k - g->logg_lo[i]) * g->n_m + j;
}

/* Grid::list_gridpoints (interpolate.rs:334-348): teff-major, logg, then m */
int or_grid_list_gridpoints(const or_grid *g, double *out) {
    int q = 0;
    for (int i = 0; i < g->n_teff; i++)
        for (int k = g->logg_lo[i]; k <= g->logg_hi[i]; k++)
            for (int j = 0; j < g->n_m; j++) {
                out[3 * q + 0] = g->teff[i];
                out[3 * q + 1] = g->m[j];
                out[3 * q + 2] = g->logg[k];
                q++;
            }
    return q;
}

/* Grid::is_teff_logg_between_bounds (interpolate.rs:351-369) */
int or_is_teff_logg_between_bounds(const or_grid *g, double teff, double logg) {
    int r = right_index(g->teff, g->n_teff, teff);
    if (r >= 0) {
        return logg >= g->logg[g->logg_lo[r]] && logg <= g->logg[g->logg_hi[r]];
    }
    int i = -r - 1;
    if (i == 0 || i == g->n_teff) return 0;
    return logg >= g->logg[g->logg_lo[i - 1]] && logg <= g->logg[g->logg_hi[i - 1]] &&
           logg >= g->logg[g->logg_lo[i]] && logg <= g->logg[g->logg_hi[i]];
}

/* Grid::get_logg_index_limits_at (interpolate.rs:378-398) */
int or_get_logg_index_limits_at(const or_grid *g, double teff, int *lo, int *hi) {
    int r = right_index(g->teff, g->n_teff, teff);
    if (r >= 0) {
        if (r == g->n_teff - 1) { *lo = g->logg_lo[r]; *hi = g->logg_hi[r]; return OR_OK; }
        int a = g->logg_lo[r], b = g->logg_lo[r + 1];
        int c = g->logg_hi[r], d = g->logg_hi[r + 1];
        *lo = a > b ? a : b; *hi = c < d ? c : d;
        return OR_OK;
    }
    int i = -r - 1;
    if (i == 0 || i == g->n_teff) return OR_ERR_BOUNDS;
    int a = g->logg_lo[i - 1], b = g->logg_lo[i];
    int c = g->logg_hi[i - 1], d = g->logg_hi[i];
    *lo = a > b ? a : b; *hi = c < d ? c : d;
    return OR_OK;
}

/* Grid::get_teff_index_limits_at (interpolate.rs:402-421) */
int or_get_teff_index_limits_at(const or_grid *g, double logg, int *lo, int *hi) {
    int first = -1, last = -1;
    for (int i = 0; i < g->n_teff; i++) {
        if (logg >= g->logg[g->logg_lo[i]] && logg <= g->logg[g->logg_hi[i]]) {
            if (first < 0) first = i;
            last = i;
        }
    }
    if (first < 0) return OR_ERR_BOUNDS;
    *lo = first; *hi = last;
    return OR_OK;
}

/* Range::find_neighbors (interpolate.rs:105-159) */
int or_find_neighbors(const double *v, int n, double x, int left, int right, int out[4]) {
    int r = right_index(v, n, x);
    /* some[q]: the reference returns Some(index) in slot q (None slots are -1) */
    int some[4] = {1, 1, 1, 1};
    if (r >= 0) {
        int i = r;
        if (i < left || i > right) return OR_ERR_BOUNDS;
        if (i == left) { some[0] = 0; out[1] = i; out[2] = i + 1; out[3] = i + 2; }
        else if (i == right) { out[0] = i - 2; out[1] = i - 1; out[2] = i; some[3] = 0; }
        else if (i == right - 1) { out[0] = i - 2; out[1] = i - 1; out[2] = i; out[3] = i + 1; }
        else { out[0] = i - 1; out[1] = i; out[2] = i + 1; out[3] = i + 2; }
    } else {
        int i = -r - 1;
        if (i <= left || i > right) return OR_ERR_BOUNDS;
        if (i == left + 1) { some[0] = 0; out[1] = i - 1; out[2] = i; out[3] = i + 1; }
        else if (i == right) { out[0] = i - 2; out[1] = i - 1; out[2] = i; some[3] = 0; }
        else { out[0] = i - 2; out[1] = i - 1; out[2] = i; out[3] = i + 1; }
    }
    /* Some(index) outside the axis: the reference's value lookup panics
       (interpolate.rs:483-530; i - 2 below zero wraps past the end first) */
    for (int q = 0; q < 4; q++) {
        if (!some[q]) { out[q] = -1; continue; }
        if (out[q] < 0 || out[q] >= n) return OR_ERR_NEIGHBORS;
    }
    return OR_OK;
}

/* ---------------- Lagrange factors: cubic.rs:12-91,143-157 ---------------- */
static void factors_quadratic(double x, double x0, double x1, double x2, int start, double f[4]) {
    double xsq = x * x;
    double d0 = x0 * x0 - x0 * x1 - x0 * x2 + x1 * x2;
    double d1 = -x1 * x1 + x0 * x1 - x0 * x2 + x1 * x2;
    double d2 = x2 * x2 + x0 * x1 - x0 * x2 - x1 * x2;
    double f0 = (xsq - (x1 + x2) * x + x1 * x2) / d0;
    double f1 = (-xsq + (x0 + x2) * x - x0 * x2) / d1;
    double f2 = (xsq - (x0 + x1) * x + x0 * x1) / d2;
    if (start) { f[0] = 0.0; f[1] = f0; f[2] = f1; f[3] = f2; }
    else { f[0] = f0; f[1] = f1; f[2] = f2; f[3] = 0.0; }
}

static void factors_cubic(double x, double x0, double x1, double x2, double x3, double f[4]) {
    double xsq = x * x;
    double xcu = xsq * x;
    double d0 = x0 * x0 * x0 - x0 * x0 * x1 - x0 * x0 * x2 - x0 * x0 * x3 + x0 * x1 * x2 +
                x0 * x1 * x3 + x0 * x2 * x3 - x1 * x2 * x3;
    double d1 = -x1 * x1 * x1 + x1 * x1 * x0 + x1 * x1 * x2 + x1 * x1 * x3 - x0 * x1 * x2 -
                x0 * x1 * x3 + x0 * x2 * x3 - x1 * x2 * x3;
    double d2 = x2 * x2 * x2 - x2 * x2 * x0 - x2 * x2 * x1 - x2 * x2 * x3 + x0 * x1 * x2 -
                x0 * x1 * x3 + x0 * x2 * x3 + x1 * x2 * x3;
    double d3 = -x3 * x3 * x3 + x3 * x3 * x0 + x3 * x3 * x1 + x3 * x3 * x2 + x0 * x1 * x2 -
                x0 * x1 * x3 - x0 * x2 * x3 - x1 * x2 * x3;
    f[0] = (xcu - (x1 + x2 + x3) * xsq + (x1 * x2 + x1 * x3 + x2 * x3) * x - (x1 * x2 * x3)) / d0;
    f[1] = (-xcu + (x0 + x2 + x3) * xsq - (x0 * x2 + x0 * x3 + x2 * x3) * x + (x0 * x2 * x3)) / d1;
    f[2] = (xcu - (x0 + x1 + x3) * xsq + (x0 * x1 + x0 * x3 + x1 * x3) * x - (x0 * x1 * x3)) / d2;
    f[3] = (-xcu + (x0 + x1 + x2) * xsq - (x0 * x1 + x0 * x2 + x1 * x2) * x + (x0 * x1 * x2)) / d3;
}

/* coefficients_from_neighbors (cubic.rs:143-157); nb value valid where has[i] */
static int coeffs_from_neighbors(double x, const double nb[4], const int has[4], double f[4]) {
    if (has[0] && has[1] && has[2] && has[3]) {
        double s = nb[3] - nb[0];
        factors_cubic((x - nb[0]) / s, 0.0, (nb[1] - nb[0]) / s, (nb[2] - nb[0]) / s, 1.0, f);
        return OR_OK;
    }
    if (has[0] && has[1] && has[2] && !has[3]) {
        double s = nb[2] - nb[0];
        factors_quadratic((x - nb[0]) / s, 0.0, (nb[1] - nb[0]) / s, 1.0, 0, f);
        return OR_OK;
    }
    if (!has[0] && has[1] && has[2] && has[3]) {
        double s = nb[3] - nb[1];
        factors_quadratic((x - nb[1]) / s, 0.0, (nb[2] - nb[1]) / s, 1.0, 1, f);
        return OR_OK;
    }
    if (!has[0] && !has[1] && !has[2] && !has[3]) {
        f[0] = f[1] = f[2] = f[3] = 0.0;
        return OR_OK;
    }
    return OR_ERR_NEIGHBORS;
}

/* Grid::get_local_grid (interpolate.rs:424-533) + calculate_interpolation_coefficients
 * (cubic.rs:159-243).  Output order is the reference's neighbour order:
 * column-major over the 4x4 (logg row r, teff column c) matrix, then m:
 * q = (c*4 + r)*4 + mi. */
int or_local_weights(const or_grid *g, double teff, double m, double logg, int idx[64],
                     float w32[64], double w64[64], double ft_out[16], double fl_out[16],
                     double fm_out[4]) {
    if (!or_is_teff_logg_between_bounds(g, teff, logg)) return OR_ERR_BOUNDS;
    if (!(m >= g->m[0] && m <= g->m[g->n_m - 1])) return OR_ERR_BOUNDS;
    int tlo, thi, llo, lhi, rc;
    if ((rc = or_get_teff_index_limits_at(g, logg, &tlo, &thi))) return rc;
    if ((rc = or_get_logg_index_limits_at(g, teff, &llo, &lhi))) return rc;
    int tn[4], mn[4], ln[4];
    if ((rc = or_find_neighbors(g->teff, g->n_teff, teff, tlo, thi, tn))) return rc;
    if ((rc = or_find_neighbors(g->m, g->n_m, m, 0, g->n_m - 1, mn))) return rc;
    if ((rc = or_find_neighbors(g->logg, g->n_logg, logg, llo, lhi, ln))) return rc;
    /* has[r][c]: matrix entry (teff nb c, logg nb r) present (interpolate.rs:460-479) */
    int has[4][4];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) {
            has[r][c] = 0;
            if (tn[c] >= 0 && ln[r] >= 0) {
                int i = tn[c];
                if (ln[r] >= g->logg_lo[i] && ln[r] <= g->logg_hi[i]) has[r][c] = 1;
            }
        }
    double FT[4][4], FL[4][4], FM[4];
    for (int r = 0; r < 4; r++) { /* map_rows: teff factors per logg row */
        double nb[4]; int h[4];
        for (int c = 0; c < 4; c++) { h[c] = has[r][c]; nb[c] = h[c] ? g->teff[tn[c]] : 0.0; }
        double f[4];
        if ((rc = coeffs_from_neighbors(teff, nb, h, f))) return rc;
        for (int c = 0; c < 4; c++) FT[r][c] = f[c];
    }
    for (int c = 0; c < 4; c++) { /* map_columns: logg factors per teff column */
        double nb[4]; int h[4];
        for (int r = 0; r < 4; r++) { h[r] = has[r][c]; nb[r] = h[r] ? g->logg[ln[r]] : 0.0; }
        double f[4];
        if ((rc = coeffs_from_neighbors(logg, nb, h, f))) return rc;
        for (int r = 0; r < 4; r++) FL[r][c] = f[r];
    }
    {
        double nb[4]; int h[4];
        for (int q = 0; q < 4; q++) { h[q] = mn[q] >= 0; nb[q] = h[q] ? g->m[mn[q]] : 0.0; }
        if ((rc = coeffs_from_neighbors(m, nb, h, FM))) return rc;
    }
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) {
            double tl = FT[r][c] * FL[r][c];
            for (int mi = 0; mi < 4; mi++) {
                int q = (c * 4 + r) * 4 + mi;
                double w = tl * FM[mi];
                if (w64) w64[q] = w;
                if (w32) w32[q] = (float)w;
                if (idx) idx[q] = (has[r][c] && mn[mi] >= 0)
                                      ? or_grid_model_index(g, tn[c], mn[mi], ln[r])
                                      : -1;
            }
        }
    if (ft_out) for (int r = 0; r < 4; r++) for (int c = 0; c < 4; c++) ft_out[r * 4 + c] = FT[r][c];
    if (fl_out) for (int r = 0; r < 4; r++) for (int c = 0; c < 4; c++) fl_out[r * 4 + c] = FL[r][c];
    if (fm_out) for (int q = 0; q < 4; q++) fm_out[q] = FM[q];
    return OR_OK;
}

/* Grid::clamp_1d / get_limits_at (interpolate.rs:625-664) */
static int grid_limits_at(const or_grid *g, const double p[3], int dim, double *lo, double *hi) {
    int a, b, rc;
    if (dim == 0) {
        if ((rc = or_get_teff_index_limits_at(g, p[2], &a, &b))) return rc;
        *lo = g->teff[a]; *hi = g->teff[b];
    } else if (dim == 1) {
        *lo = g->m[0]; *hi = g->m[g->n_m - 1];
    } else {
        if ((rc = or_get_logg_index_limits_at(g, p[0], &a, &b))) return rc;
        *lo = g->logg[a]; *hi = g->logg[b];
    }
    return OR_OK;
}

static double clampd(double x, double lo, double hi) {
    /* f64::clamp: NaN passes through */
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

int or_clamp_1d(const or_grid *g, const double p[3], int dim, double *out) {
    double lo, hi;
    int rc = grid_limits_at(g, p, dim, &lo, &hi);
    if (rc) return rc;
    *out = clampd(p[dim], lo, hi);
    return OR_OK;
}

/* ======================================================================
 * Interpolation -- interpolate.rs:930-983
 * ====================================================================== */
int or_interpolate_f32(const or_grid *g, const uint16_t *spectra, const float *factors, int W,
                       double teff, double m, double logg, float *out) {
    int idx[64];
    float w[64];
    int rc = or_local_weights(g, teff, m, logg, idx, w, NULL, NULL, NULL, NULL);
    if (rc) return rc;
    float fac[64];
    for (int j = 0; j < 64; j++)
        fac[j] = (idx[j] >= 0 && factors) ? factors[idx[j]] : (float)(1.0 / 65535.0);
    int nb = (W / BATCH_SIZE) * BATCH_SIZE;
    for (int x = 0; x < W; x++) {
        /* gemv column by column (nalgebra gemv + axcpy): y = M0*w0; y += Mj*wj */
        float y = 0.0f;
        for (int j = 0; j < 64; j++) {
            float u = idx[j] >= 0 ? (float)spectra[(size_t)idx[j] * W + x] : 0.0f;
            float Mv = (x < nb) ? u * fac[j]                      /* interpolate.rs:962 */
                                : u / 65535.0f * fac[j];          /* tail quirk :978 */
            float t = Mv * w[j];
            y = (j == 0) ? t : t + y;
        }
        out[x] = y;
    }
    return OR_OK;
}

/* f32w != 0: the 64 weights rounded to f32 as the reference stores them (cubic.rs:241) */
static int interpolate_f64_impl(const or_grid *g, const uint16_t *spectra, const float *factors,
                                int W, double teff, double m, double logg, double *out, int f32w) {
    int idx[64];
    double w[64];
    int rc = or_local_weights(g, teff, m, logg, idx, NULL, w, NULL, NULL, NULL);
    if (rc) return rc;
    if (f32w)
        for (int j = 0; j < 64; j++) w[j] = (double)(float)w[j];
    double fac[64];
    for (int j = 0; j < 64; j++)
        fac[j] = (idx[j] >= 0 && factors) ? (double)factors[idx[j]] : 1.0 / 65535.0;
    int nb = (W / BATCH_SIZE) * BATCH_SIZE;
    for (int x = 0; x < W; x++) {
        double y = 0.0;
        for (int j = 0; j < 64; j++) {
            double u = idx[j] >= 0 ? (double)spectra[(size_t)idx[j] * W + x] : 0.0;
            double Mv = (x < nb) ? u * fac[j] : u / 65535.0 * fac[j];
            y += Mv * w[j];
        }
        out[x] = y;
    }
    return OR_OK;
}

int or_interpolate_f64(const or_grid *g, const uint16_t *spectra, const float *factors, int W,
                       double teff, double m, double logg, double *out) {
    return interpolate_f64_impl(g, spectra, factors, W, teff, m, logg, out, 0);
}

/* ======================================================================
 * Convolutions -- convolve_rv.rs
 * ====================================================================== */

/* build_rotation_kernel (convolve_rv.rs:438-475); returns length or error */
static int rotation_kernel_len(double vsini, double dvelo) {
    double epsilon = 0.6;
    double vrot = vsini / C_KMS;
    double a = 4.0 * (epsilon - 1.0) / (M_PI * epsilon);
    double v = ceil(2.0 * vrot / dvelo * sqrt(1.0 + a * a));
    long nme = (long)v + 2;
    if (!(v >= 0.0)) nme = 2; /* `as usize` saturates negative/NaN to 0 */
    if (nme < 1) nme = 1;
    if (nme > (1L << 30)) return -1;
    long n = nme + 1 - nme % 2;
    return (int)n;
}

int or_rotation_kernel(double vsini, double dvelo, float *out, int max_len) {
    double epsilon = 0.6;
    double vrot = vsini / C_KMS;
    int n = rotation_kernel_len(vsini, dvelo);
    if (n < 0 || n > max_len) return n < 0 ? OR_ERR_KERNEL : -(n + 1000);
    double mid = ((double)(n - 1) * dvelo) / 2.0;
    float sum = 0.0f;
    for (int i = 0; i < n; i++) {
        double vk = (double)i * dvelo - mid;
        double q = vk / vrot;
        double y = 1.0 - q * q;
        float k = 0.0f;
        if (y > 0.0)
            k = (float)((2.0 * (1.0 - epsilon) * sqrt(y) + M_PI * epsilon / 2.0 * y) /
                        (M_PI * vrot * (1.0 - epsilon / 3.0)));
        out[i] = k;
        sum = sum + k;
    }
    for (int i = 0; i < n; i++) out[i] = out[i] / sum;
    return n;
}

int or_rotation_kernel_f64(double vsini, double dvelo, double *out, int max_len) {
    double epsilon = 0.6;
    double vrot = vsini / C_KMS;
    int n = rotation_kernel_len(vsini, dvelo);
    if (n < 0 || n > max_len) return n < 0 ? OR_ERR_KERNEL : -(n + 1000);
    double mid = ((double)(n - 1) * dvelo) / 2.0;
    double sum = 0.0;
    for (int i = 0; i < n; i++) {
        double vk = (double)i * dvelo - mid;
        double q = vk / vrot;
        double y = 1.0 - q * q;
        double k = 0.0;
        if (y > 0.0)
            k = (2.0 * (1.0 - epsilon) * sqrt(y) + M_PI * epsilon / 2.0 * y) /
                (M_PI * vrot * (1.0 - epsilon / 3.0));
        out[i] = k;
        sum += k;
    }
    for (int i = 0; i < n; i++) out[i] /= sum;
    return n;
}

/* ---- radix-2 f32 complex FFT of size FFTSIZE (stands in for realfft) ---- */
static float tw_re[FFTSIZE / 2], tw_im[FFTSIZE / 2];
static int bitrev[FFTSIZE];
static pthread_once_t fft_once = PTHREAD_ONCE_INIT;

static void fft_init(void) {
    for (int k = 0; k < FFTSIZE / 2; k++) {
        double a = -2.0 * M_PI * (double)k / (double)FFTSIZE;
        tw_re[k] = (float)cos(a);
        tw_im[k] = (float)sin(a);
    }
    int lg = 0;
    while ((1 << lg) < FFTSIZE) lg++;
    for (int i = 0; i < FFTSIZE; i++) {
        int r = 0;
        for (int b = 0; b < lg; b++)
            if (i & (1 << b)) r |= 1 << (lg - 1 - b);
        bitrev[i] = r;
    }
}

static void fft_c(float *re, float *im, int inverse) {
    for (int i = 0; i < FFTSIZE; i++) {
        int j = bitrev[i];
        if (j > i) {
            float t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= FFTSIZE; len <<= 1) {
        int half = len >> 1, stride = FFTSIZE / len;
        for (int s = 0; s < FFTSIZE; s += len)
            for (int k = 0; k < half; k++) {
                float wr = tw_re[k * stride];
                float wi = inverse ? -tw_im[k * stride] : tw_im[k * stride];
                int a = s + k, b = s + k + half;
                float xr = re[b] * wr - im[b] * wi;
                float xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr; im[b] = im[a] - xi;
                re[a] = re[a] + xr; im[a] = im[a] + xi;
            }
    }
}

int or_oa_out_len(int l, int m) {
    int step = FFTSIZE - (m - 1);
    int nx = ((l + step - 1) / step) * step;
    return nx + m - 1;
}

/* oa_convolve (convolve_rv.rs:121-157) */
int or_oa_convolve(const float *x, int l, const float *h, int m, float *y, int y_cap) {
    pthread_once(&fft_once, fft_init);
    if (m < 1 || m > FFTSIZE) return OR_ERR_KERNEL;
    int step = FFTSIZE - (m - 1);
    int nx = ((l + step - 1) / step) * step;
    int ylen = nx + m - 1;
    if (ylen > y_cap) return OR_ERR_ARG;
    float *hr = calloc(FFTSIZE, sizeof(float)), *hi = calloc(FFTSIZE, sizeof(float));
    float *xr = malloc(sizeof(float) * FFTSIZE), *xi = malloc(sizeof(float) * FFTSIZE);
    memcpy(hr, h, sizeof(float) * m);
    fft_c(hr, hi, 0);
    memset(y, 0, sizeof(float) * ylen);
    for (int pos = 0; pos < nx; pos += step) {
        int cnt = (pos + step < l) ? step : l - pos;
        memset(xr, 0, sizeof(float) * FFTSIZE);
        memset(xi, 0, sizeof(float) * FFTSIZE);
        if (cnt > 0) memcpy(xr, x + pos, sizeof(float) * cnt);
        fft_c(xr, xi, 0);
        for (int k = 0; k < FFTSIZE; k++) {
            float a = xr[k] * hr[k] - xi[k] * hi[k];
            float b = xr[k] * hi[k] + xi[k] * hr[k];
            xr[k] = a; xi[k] = b;
        }
        fft_c(xr, xi, 1);
        for (int k = 0; k < FFTSIZE; k++) y[pos + k] = y[pos + k] + xr[k] / (float)FFTSIZE;
    }
    free(hr); free(hi); free(xr); free(xi);
    return ylen;
}

void or_direct_convolve_f64(const double *x, int l, const double *h, int m, double *y) {
    for (int n = 0; n < l + m - 1; n++) {
        double s = 0.0;
        int k0 = n - (l - 1);
        if (k0 < 0) k0 = 0;
        int k1 = n < m - 1 ? n : m - 1;
        for (int k = k0; k <= k1; k++) s += h[k] * x[n - k];
        y[n] = s;
    }
}

/* make_gaussian (convolve_rv.rs:107-118) */
void or_make_gaussian(float *arr, int n, float sigma) {
    int mid = (n - 1) / 2;
    for (int i = 0; i < n; i++) {
        float x = ((float)i - (float)mid) / sigma;
        arr[i] = expf(-0.5f * x * x);
    }
    float sum = 0.0f;
    for (int i = 0; i < n; i++) sum = sum + arr[i];
    for (int i = 0; i < n; i++) arr[i] = arr[i] / sum;
}

/* FixedTargetDispersionLogSpace::new kernel (convolve_rv.rs:169-195) */
int or_fixed_gauss_kernel(double resolution, double log_step, float *out, int max_len) {
    if (!(resolution > 0.0)) return OR_ERR_ARG;
    double sigma = 1.0 / (LN_10 * 2.355 * resolution * log_step);
    long nme = (long)ceil(6.0 * sigma);
    long n = nme + 1 - nme % 2;
    if (n > max_len) return OR_ERR_ARG;
    or_make_gaussian(out, (int)n, (float)sigma);
    return (int)n;
}

/* nalgebra 0.32 Matrix::dotx: 8 partial accumulators, then a scalar tail */
static float dot_na(const float *a, const float *b, int n) {
    float res = 0.0f;
    int i = 0;
    float c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0, c6 = 0, c7 = 0;
    while (n - i >= 8) {
        c0 = c0 + a[i + 0] * b[i + 0]; c1 = c1 + a[i + 1] * b[i + 1];
        c2 = c2 + a[i + 2] * b[i + 2]; c3 = c3 + a[i + 3] * b[i + 3];
        c4 = c4 + a[i + 4] * b[i + 4]; c5 = c5 + a[i + 5] * b[i + 5];
        c6 = c6 + a[i + 6] * b[i + 6]; c7 = c7 + a[i + 7] * b[i + 7];
        i += 8;
    }
    res = res + (c0 + c4);
    res = res + (c1 + c5);
    res = res + (c2 + c6);
    res = res + (c3 + c7);
    for (; i < n; i++) res = res + a[i] * b[i];
    return res;
}

/* ======================================================================
 * Dispersion objects -- convolve_rv.rs:71-431
 * ====================================================================== */
static double wl_float_index(const or_wlgrid *sg, double wl) {
    return (log10(wl) - sg->first) / sg->step; /* interpolate.rs:708 */
}

/* resample_observed_to_synth (convolve_rv.rs:303-331), Logspace synthetic grid */
static void resample_obs_to_synth(const double *wl_obs, int n_obs, const or_wlgrid *sg,
                                  const float *y, float *out) {
    int j = 0;
    for (int i = 0; i < sg->n; i++) {
        double x = pow(10.0, sg->first + sg->step * (double)i);
        while (j < n_obs - 1 && wl_obs[j + 1] < x) j++;
        if (j == n_obs - 1) {
            out[i] = y[j];
        } else {
            float w = (float)((x - wl_obs[j]) / (wl_obs[j + 1] - wl_obs[j]));
            out[i] = (1.0f - w) * y[j] + w * y[j + 1];
        }
    }
}

or_disp *or_disp_new(int kind, const double *wl, int n, const or_wlgrid *sg, double resolution,
                     const float *disp) {
    or_disp *d = calloc(1, sizeof(or_disp));
    d->kind = kind;
    d->n_obs = n;
    d->W = sg->n;
    d->wl = malloc(sizeof(double) * n);
    memcpy(d->wl, wl, sizeof(double) * n);
    d->fidx = malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) d->fidx[i] = wl_float_index(sg, wl[i]);
    if (kind == OR_DISP_FIXED) {
        d->kernel = malloc(sizeof(float) * 4096);
        int K = or_fixed_gauss_kernel(resolution, sg->step, d->kernel, 4096);
        if (K < 0) { or_disp_free(d); return NULL; }
        d->K = K;
        d->sigma = 1.0 / (LN_10 * 2.355 * resolution * sg->step);
    } else if (kind == OR_DISP_VARIABLE) {
        /* VariableTargetDispersion::new (convolve_rv.rs:350-402) */
        float *dr = malloc(sizeof(float) * sg->n);
        resample_obs_to_synth(wl, n, sg, disp, dr);
        float *sig = malloc(sizeof(float) * sg->n);
        float largest = -INFINITY;
        for (int i = 0; i < sg->n; i++) {
            double wlx = pow(10.0, sg->first + sg->step * (double)i);
            sig[i] = dr[i] / (float)(sg->step * LN_10 * wlx);
            if (sig[i] > largest || isnan(sig[i])) largest = sig[i];
        }
        long kme = (long)ceilf(largest * 6.0f);
        long K = kme + 1 - kme % 2;
        d->K = (int)K;
        d->kernels = malloc(sizeof(float) * (size_t)K * sg->n);
        for (int i = 0; i < sg->n; i++) or_make_gaussian(d->kernels + (size_t)i * K, (int)K, sig[i]);
        free(dr); free(sig);
    }
    return d;
}

void or_disp_free(or_disp *d) {
    if (!d) return;
    free(d->wl); free(d->fidx); free(d->kernel); free(d->kernels);
    free(d);
}

int or_disp_kernel_len(const or_disp *d) { return d->K; }

/* fixed: the K taps; variable: W rows of K taps (row i = synthetic pixel i's kernel) */
long or_disp_kernels(const or_disp *d, float *out, long cap) {
    if (d->kind == OR_DISP_FIXED) {
        if (cap < d->K) return OR_ERR_ARG;
        memcpy(out, d->kernel, sizeof(float) * d->K);
        return d->K;
    }
    if (d->kind != OR_DISP_VARIABLE) return 0;
    long n = (long)d->K * d->W;
    if (cap < n) return OR_ERR_ARG;
    memcpy(out, d->kernels, sizeof(float) * (size_t)n);
    return n;
}

/* ======================================================================
 * produce_model -- interpolate.rs:1068-1090 composed of
 * convolve_rotation (convolve_rv.rs:484-529), convolve_segment, and
 * shift_and_resample (convolve_rv.rs:532-604)
 * ====================================================================== */

/* A segment of an array (ArraySegment, convolve_rv.rs:23-67): indexing is
 * start+i with no check against len; reading past the array is an error
 * (the reference panics). */
typedef struct { float *a; int alen; int start; int len; } seg32;
typedef struct { double *a; int alen; int start; int len; } seg64;

static int rot_conv32(const or_model *M, const float *in, double vsini, seg32 *out) {
    int W = M->sg.n;
    if (vsini < 0.0) return OR_ERR_VSINI;
    double dvelo = LN_10 * M->sg.step;
    int n = rotation_kernel_len(vsini, dvelo);
    if (n < 0) return OR_ERR_KERNEL;
    if (vsini == 0.0 || n <= 2) {
        out->a = malloc(sizeof(float) * W);
        memcpy(out->a, in, sizeof(float) * W);
        out->alen = W; out->start = 0; out->len = W;
        return OR_OK;
    }
    if (n > FFTSIZE) return OR_ERR_KERNEL;
    float *k = malloc(sizeof(float) * n);
    or_rotation_kernel(vsini, dvelo, k, n);
    int ylen = or_oa_out_len(W, n);
    out->a = malloc(sizeof(float) * ylen);
    or_oa_convolve(in, W, k, n, out->a, ylen);
    out->alen = ylen; out->start = n / 2; out->len = W;
    free(k);
    return OR_OK;
}

static int disp_conv32(const or_disp *d, seg32 *s) {
    int W = s->len;
    if (d->kind == OR_DISP_NONE) return OR_OK; /* segment passes through (:102-104) */
    float *flat = malloc(sizeof(float) * W);
    memcpy(flat, s->a + s->start, sizeof(float) * W);
    if (d->kind == OR_DISP_FIXED) {
        int ylen = or_oa_out_len(W, d->K);
        float *y = malloc(sizeof(float) * ylen);
        or_oa_convolve(flat, W, d->kernel, d->K, y, ylen);
        free(s->a);
        s->a = malloc(sizeof(float) * W);
        memcpy(s->a, y + (d->K - 1) / 2, sizeof(float) * W); /* rows(n, len) (:207-210) */
        free(y);
    } else {
        if (d->W != W) { free(flat); return OR_ERR_ARG; }
        int K = d->K, h = (K - 1) / 2;
        free(s->a);
        s->a = malloc(sizeof(float) * W);
        for (int i = 0; i < W; i++) { /* convolve_rv.rs:422-429 */
            if (i < h || i >= W - h) s->a[i] = flat[i];
            else s->a[i] = dot_na(d->kernels + (size_t)i * K, flat + i - h, K);
        }
    }
    free(flat);
    s->alen = W; s->start = 0; s->len = W;
    return OR_OK;
}

static int range_check(const or_model *M, const or_disp *d, double rv) {
    double shift = 1.0 - rv / C_KMS;
    double first = pow(10.0, M->sg.first);
    double last = pow(10.0, M->sg.first + M->sg.step * (double)M->sg.n);
    if (d->wl[0] * shift < first) return OR_ERR_RANGE;
    if (d->wl[d->n_obs - 1] * shift > last) return OR_ERR_RANGE;
    return OR_OK;
}

static int resample32(const or_model *M, const or_disp *d, const seg32 *s, double rv, float *out) {
    int rc = range_check(M, d, rv);
    if (rc) return rc;
    double shift = 1.0 - rv / C_KMS;
    double shift_log = log10(shift) / M->sg.step;
    for (int i = 0; i < d->n_obs; i++) {
        double index = d->fidx[i] + shift_log;
        double fl = floor(index);
        long j = fl > 0 ? (long)fl : 0;
        float w = (float)index - (float)j;                     /* convolve_rv.rs:571 */
        long a = s->start + j, b = s->start + j + 1;
        if (b >= s->alen) return OR_ERR_INDEX;
        out[i] = (1.0f - w) * s->a[a] + w * s->a[b];
    }
    return OR_OK;
}

int or_interpolate_and_convolve(const or_model *M, const or_disp *d, const double lab[4],
                                float *out) {
    int W = M->sg.n;
    float *interp = malloc(sizeof(float) * W);
    int rc = or_interpolate_f32(M->g, M->spectra, M->factors, W, lab[0], lab[1], lab[2], interp);
    if (rc) { free(interp); return rc; }
    seg32 s = {0};
    rc = rot_conv32(M, interp, lab[3], &s);
    free(interp);
    if (rc) return rc;
    rc = disp_conv32(d, &s);
    if (!rc) memcpy(out, s.a + s.start, sizeof(float) * W);
    free(s.a);
    return rc;
}

int or_produce_model(const or_model *M, const or_disp *d, const double lab[5], float *out) {
    int W = M->sg.n;
    float *interp = malloc(sizeof(float) * W);
    int rc = or_interpolate_f32(M->g, M->spectra, M->factors, W, lab[0], lab[1], lab[2], interp);
    if (rc) { free(interp); return rc; }
    seg32 s = {0};
    rc = rot_conv32(M, interp, lab[3], &s);
    free(interp);
    if (rc) return rc;
    rc = disp_conv32(d, &s);
    if (!rc) rc = resample32(M, d, &s, lab[4], out);
    free(s.a);
    return rc;
}

/* ---------------- FP64 mode ---------------- */
/* qweight != 0: FP64 arithmetic over the values the reference stores in f32: the 64
 * interpolation weights (cubic.rs:241), the rotation and instrumental taps
 * (convolve_rv.rs:107-118,468-474) and the resample weight f32(index) - f32(j) (:570-572).
 * This separates the f32 arithmetic error of a path from those quantisations, which the
 * reference's own result carries (the resample weight alone up to 7.8e-3 of a pixel step
 * for indices >= 131072). */
static int produce_model_f64_impl(const or_model *M, const or_disp *d, const double lab[5],
                                  double *out, int qweight) {
    int W = M->sg.n;
    double *interp = malloc(sizeof(double) * W);
    int rc = interpolate_f64_impl(M->g, M->spectra, M->factors, W, lab[0], lab[1], lab[2], interp,
                                  qweight);
    if (rc) { free(interp); return rc; }
    double vsini = lab[3];
    if (vsini < 0.0) { free(interp); return OR_ERR_VSINI; }
    double dvelo = LN_10 * M->sg.step;
    int n = rotation_kernel_len(vsini, dvelo);
    seg64 s;
    if (vsini == 0.0 || n <= 2) {
        s.a = interp; s.alen = W; s.start = 0; s.len = W;
    } else {
        if (n > FFTSIZE) { free(interp); return OR_ERR_KERNEL; }
        double *k = malloc(sizeof(double) * n);
        if (qweight) {  /* the reference's f32 taps, normalised by their f32 sum */
            float *k32 = malloc(sizeof(float) * n);
            or_rotation_kernel(vsini, dvelo, k32, n);
            for (int i = 0; i < n; i++) k[i] = (double)k32[i];
            free(k32);
        } else {
            or_rotation_kernel_f64(vsini, dvelo, k, n);
        }
        int ylen = or_oa_out_len(W, n);
        double *y = calloc((size_t)ylen, sizeof(double));
        or_direct_convolve_f64(interp, W, k, n, y); /* W+n-1 <= ylen; rest zero */
        free(k); free(interp);
        s.a = y; s.alen = ylen; s.start = n / 2; s.len = W;
    }
    if (d->kind != OR_DISP_NONE) {
        double *flat = malloc(sizeof(double) * W);
        memcpy(flat, s.a + s.start, sizeof(double) * W);
        double *res = malloc(sizeof(double) * W);
        if (d->kind == OR_DISP_FIXED) {
            int K = d->K, h = (K - 1) / 2;
            double *kk = malloc(sizeof(double) * K), ks = 0.0;
            if (qweight) {  /* the reference's f32 taps (make_gaussian) */
                for (int i = 0; i < K; i++) kk[i] = (double)d->kernel[i];
            } else {
                for (int i = 0; i < K; i++) {
                    double x = ((double)i - (double)h) / d->sigma;
                    kk[i] = exp(-0.5 * x * x);
                    ks += kk[i];
                }
                for (int i = 0; i < K; i++) kk[i] /= ks;
            }
            double *y = malloc(sizeof(double) * (W + K - 1));
            or_direct_convolve_f64(flat, W, kk, K, y);
            memcpy(res, y + h, sizeof(double) * W);
            free(y); free(kk);
        } else {
            int K = d->K, h = (K - 1) / 2;
            for (int i = 0; i < W; i++) {
                if (i < h || i >= W - h) { res[i] = flat[i]; continue; }
                double acc = 0.0;
                const float *kc = d->kernels + (size_t)i * K;
                for (int q = 0; q < K; q++) acc += (double)kc[q] * flat[i - h + q];
                res[i] = acc;
            }
        }
        free(flat); free(s.a);
        s.a = res; s.alen = W; s.start = 0; s.len = W;
    }
    int rc2 = range_check(M, d, lab[4]);
    if (rc2) { free(s.a); return rc2; }
    double shift = 1.0 - lab[4] / C_KMS;
    double shift_log = log10(shift) / M->sg.step;
    for (int i = 0; i < d->n_obs; i++) {
        double index = d->fidx[i] + shift_log;
        double fl = floor(index);
        long j = fl > 0 ? (long)fl : 0;
        double w = qweight ? (double)((float)index - (float)j) : index - (double)j;
        long a = s.start + j, b = a + 1;
        if (b >= s.alen) { free(s.a); return OR_ERR_INDEX; }
        out[i] = (1.0 - w) * s.a[a] + w * s.a[b];
    }
    free(s.a);
    return OR_OK;
}

int or_produce_model_f64(const or_model *M, const or_disp *d, const double lab[5], double *out) {
    return produce_model_f64_impl(M, d, lab, out, 0);
}

int or_produce_model_f64q(const or_model *M, const or_disp *d, const double lab[5], double *out) {
    return produce_model_f64_impl(M, d, lab, out, 1);
}

/* ======================================================================
 * Least squares: FP64 Householder QR (stands in for nalgebra's f32 SVD
 * solve with eps 1e-14, continuum_fitting.rs:59-65,223)
 * ====================================================================== */
typedef struct { int n, k; double *V; double *beta; double *R; } qr_t;

static qr_t *qr_new(const double *A, int n, int k) {
    qr_t *q = calloc(1, sizeof(qr_t));
    q->n = n; q->k = k;
    q->V = malloc(sizeof(double) * (size_t)n * k);
    memcpy(q->V, A, sizeof(double) * (size_t)n * k);
    q->beta = calloc(k, sizeof(double));
    q->R = calloc((size_t)k * k, sizeof(double));
    double *V = q->V;
    for (int j = 0; j < k; j++) {
        double nrm = 0.0;
        for (int i = j; i < n; i++) nrm += V[(size_t)i * k + j] * V[(size_t)i * k + j];
        nrm = sqrt(nrm);
        double x0 = V[(size_t)j * k + j];
        double alpha = x0 > 0 ? -nrm : nrm;
        V[(size_t)j * k + j] = x0 - alpha;
        double b = 0.0;
        for (int i = j; i < n; i++) b += V[(size_t)i * k + j] * V[(size_t)i * k + j];
        q->beta[j] = b;
        for (int c = j + 1; c < k; c++) {
            double s = 0.0;
            for (int i = j; i < n; i++) s += V[(size_t)i * k + j] * V[(size_t)i * k + c];
            double f = b > 0 ? 2.0 * s / b : 0.0;
            for (int i = j; i < n; i++) V[(size_t)i * k + c] -= f * V[(size_t)i * k + j];
        }
        q->R[j * k + j] = alpha;
        for (int c = j + 1; c < k; c++) q->R[j * k + c] = V[(size_t)j * k + c];
    }
    return q;
}

static void qr_solve(const qr_t *q, const double *y, double *x) {
    int n = q->n, k = q->k;
    double *t = malloc(sizeof(double) * n);
    memcpy(t, y, sizeof(double) * n);
    for (int j = 0; j < k; j++) {
        double s = 0.0;
        for (int i = j; i < n; i++) s += q->V[(size_t)i * k + j] * t[i];
        double f = q->beta[j] > 0 ? 2.0 * s / q->beta[j] : 0.0;
        for (int i = j; i < n; i++) t[i] -= f * q->V[(size_t)i * k + j];
    }
    for (int j = k - 1; j >= 0; j--) {
        double s = t[j];
        for (int c = j + 1; c < k; c++) s -= q->R[j * k + c] * x[c];
        x[j] = s / q->R[j * k + j];
    }
    free(t);
}

static void qr_free(qr_t *q) {
    if (!q) return;
    free(q->V); free(q->beta); free(q->R); free(q);
}

/* ======================================================================
 * Continuum fitters -- continuum_fitting.rs
 * ====================================================================== */
typedef struct {
    qr_t **qr;       /* per chunk, on the f32 design values */
    qr_t **qr64;     /* per chunk, on the f64 design values (FP64 mode) */
    double **A64;    /* per chunk f64 design */
    qr_t *lin;       /* linear */
} cont_priv;

static double map_range_d(double x, double fl, double fh, double tl, double th, int clamp) {
    double v = (x - fl) / (fh - fl) * (th - tl) + tl;
    if (clamp) { v = fmax(v, tl); v = fmin(v, th); }
    return v;
}
static float map_range_f(float x, float fl, float fh, float tl, float th) {
    return (x - fl) / (fh - fl) * (th - tl) + tl;
}

/* __powidf2 (exponentiation by squaring) -- what f64::powi lowers to */
static double powi_d(double a, int b) {
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return r;
}

/* poly_blend (continuum_fitting.rs:119-123) */
static float poly_blend(float x) {
    float xc = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
    float s = sinf(xc * (float)M_PI / 2.0f);
    return s * s;
}

/* ChunkFitter::new (continuum_fitting.rs:140-214) */
or_cont *or_cont_new_chunk(const double *wl, int n, int n_chunks, int p_order, double overlap) {
    if (n_chunks < 2 || n < 2) return NULL;
    or_cont *c = calloc(1, sizeof(or_cont));
    cont_priv *pv = calloc(1, sizeof(cont_priv));
    c->kind = OR_CONT_CHUNK; c->n = n; c->n_chunks = n_chunks; c->p_order = p_order;
    c->wl = malloc(sizeof(double) * n);
    memcpy(c->wl, wl, sizeof(double) * n);
    c->startstop = malloc(sizeof(int) * 2 * n_chunks);
    for (int i = 0; i < n_chunks; i++) {
        double start = map_range_d((double)i - overlap, 0.0, (double)n_chunks, wl[0], wl[n - 1], 1);
        double stop = map_range_d((double)i + 1.0 + overlap, 0.0, (double)n_chunks, wl[0],
                                  wl[n - 1], 1);
        int a = -1, b = -1;
        for (int x = 0; x < n; x++) if (wl[x] > start) { a = x; break; }
        for (int x = n - 1; x >= 0; x--) if (wl[x] < stop) { b = x; break; }
        if (a < 0 || b < 0) { or_cont_free(c); free(pv); return NULL; }
        c->startstop[2 * i] = a;
        c->startstop[2 * i + 1] = b + 1;
    }
    c->startstop[0] = 0;
    c->startstop[2 * (n_chunks - 1) + 1] = n;
    int k = p_order + 1;
    c->A = calloc(n_chunks, sizeof(float *));
    pv->qr = calloc(n_chunks, sizeof(qr_t *));
    pv->qr64 = calloc(n_chunks, sizeof(qr_t *));
    pv->A64 = calloc(n_chunks, sizeof(double *));
    for (int ch = 0; ch < n_chunks; ch++) {
        int s = c->startstop[2 * ch], e = c->startstop[2 * ch + 1];
        int len = e - s;
        c->A[ch] = malloc(sizeof(float) * (size_t)len * k);
        pv->A64[ch] = malloc(sizeof(double) * (size_t)len * k);
        double *tmp = malloc(sizeof(double) * (size_t)len * k);
        for (int x = 0; x < len; x++) {
            double xr = map_range_d(wl[s + x], wl[s], wl[e - 1], -1.0, 1.0, 0);
            for (int i = 0; i < k; i++) {
                double pw = powi_d(xr, i);
                c->A[ch][(size_t)x * k + i] = (float)pw;
                tmp[(size_t)x * k + i] = (double)(float)pw;
                pv->A64[ch][(size_t)x * k + i] = pw;
            }
        }
        pv->qr[ch] = qr_new(tmp, len, k);
        pv->qr64[ch] = qr_new(pv->A64[ch], len, k);
        free(tmp);
    }
    c->fac = calloc(n_chunks, sizeof(float *));
    for (int ch = 0; ch < n_chunks - 1; ch++) {
        int stop = c->startstop[2 * ch + 1];
        int sn = c->startstop[2 * (ch + 1)];
        int len = stop - sn;
        if (len < 0) len = 0;
        c->fac[ch] = malloc(sizeof(float) * (len > 0 ? len : 1));
        for (int x = 0; x < len; x++)
            c->fac[ch][x] = poly_blend(map_range_f((float)wl[sn + x], (float)wl[sn],
                                                   (float)wl[stop - 1], 1.0f, 0.0f));
    }
    c->design = (float *)pv; /* private data rides in an unused slot */
    return c;
}

or_cont *or_cont_new_linear(const float *design, int n, int ncols) {
    or_cont *c = calloc(1, sizeof(or_cont));
    c->kind = OR_CONT_LINEAR; c->n = n; c->ncols = ncols;
    float *dm = malloc(sizeof(float) * (size_t)n * ncols);
    memcpy(dm, design, sizeof(float) * (size_t)n * ncols);
    double *tmp = malloc(sizeof(double) * (size_t)n * ncols);
    for (size_t i = 0; i < (size_t)n * ncols; i++) tmp[i] = design[i];
    cont_priv *pv = calloc(1, sizeof(cont_priv));
    pv->lin = qr_new(tmp, n, ncols);
    free(tmp);
    c->A = calloc(2, sizeof(float *));
    c->A[0] = dm;
    c->A[1] = (float *)pv;
    return c;
}

or_cont *or_cont_new_fixed(const float *cont, int n, int ignore) {
    or_cont *c = calloc(1, sizeof(or_cont));
    c->kind = OR_CONT_FIXED; c->n = n; c->ignore = ignore;
    c->fixed = malloc(sizeof(float) * n);
    memcpy(c->fixed, cont, sizeof(float) * n);
    return c;
}

or_cont *or_cont_new_constant(int n) {
    or_cont *c = calloc(1, sizeof(or_cont));
    c->kind = OR_CONT_CONSTANT; c->n = n;
    return c;
}

void or_cont_free(or_cont *c) {
    if (!c) return;
    if (c->kind == OR_CONT_CHUNK) {
        cont_priv *pv = (cont_priv *)c->design;
        for (int i = 0; i < c->n_chunks; i++) {
            if (c->A) free(c->A[i]);
            if (c->fac) free(c->fac[i]);
            if (pv) { qr_free(pv->qr[i]); qr_free(pv->qr64[i]); free(pv->A64[i]); }
        }
        if (pv) { free(pv->qr); free(pv->qr64); free(pv->A64); free(pv); }
        free(c->A); free(c->fac); free(c->startstop); free(c->wl);
    } else if (c->kind == OR_CONT_LINEAR) {
        cont_priv *pv = (cont_priv *)c->A[1];
        qr_free(pv->lin); free(pv);
        free(c->A[0]); free(c->A);
    }
    free(c->fixed);
    free(c);
}

int or_cont_n_params(const or_cont *c) {
    switch (c->kind) {
    case OR_CONT_CHUNK: return c->n_chunks * (c->p_order + 1);
    case OR_CONT_LINEAR: return c->ncols;
    default: return 1;
    }
}

int or_cont_startstop(const or_cont *c, int *out) {
    if (c->kind != OR_CONT_CHUNK) return OR_ERR_ARG;
    memcpy(out, c->startstop, sizeof(int) * 2 * c->n_chunks);
    return c->n_chunks;
}

int or_cont_design_row(const or_cont *c, int chunk, int row, float *out) {
    if (c->kind != OR_CONT_CHUNK) return OR_ERR_ARG;
    int k = c->p_order + 1;
    memcpy(out, c->A[chunk] + (size_t)row * k, sizeof(float) * k);
    return k;
}

int or_cont_fac(const or_cont *c, int boundary, float *out, int cap) {
    if (c->kind != OR_CONT_CHUNK || boundary >= c->n_chunks - 1) return OR_ERR_ARG;
    int len = c->startstop[2 * boundary + 1] - c->startstop[2 * (boundary + 1)];
    if (len > cap) return OR_ERR_ARG;
    memcpy(out, c->fac[boundary], sizeof(float) * len);
    return len;
}

static float dot_k(const float *a, const float *b, int k) { return dot_na(a, b, k); }

/* ChunkFitter::fit_continuum -> _fit_chunks + compute_chi2_from_chunks
 * (continuum_fitting.rs:216-226,280-360,382-396) */
static int chunk_fit32(const or_cont *c, const float *flux, const float *var, const float *synth,
                       float *params, double *chi2) {
    cont_priv *pv = (cont_priv *)c->design;
    int n = c->n, k = c->p_order + 1, nc = c->n_chunks;
    float *y = malloc(sizeof(float) * n);
    for (int i = 0; i < n; i++) y[i] = flux[i] / synth[i];
    float *pf = malloc(sizeof(float) * nc * k);
    double *yd = malloc(sizeof(double) * n), *xs = malloc(sizeof(double) * k);
    for (int ch = 0; ch < nc; ch++) {
        int s = c->startstop[2 * ch], e = c->startstop[2 * ch + 1];
        for (int i = s; i < e; i++) yd[i - s] = y[i];
        qr_solve(pv->qr[ch], yd, xs);
        for (int i = 0; i < k; i++) pf[ch * k + i] = (float)xs[i];
    }
    free(yd); free(xs);
    if (params) memcpy(params, pf, sizeof(float) * nc * k);
    const int *ss = c->startstop;
    float total = 0.0f;
    {
        int sn = ss[2];
        float seg = 0.0f;
        for (int i = 50; i < sn; i++) {
            float p = dot_k(c->A[0] + (size_t)i * k, pf, k);
            float r = flux[i] - p * synth[i];
            seg = seg + (r * r) / var[i];
        }
        total = total + seg;
    }
    for (int ch = 1; ch < nc; ch++) {
        int start = ss[2 * ch], stop_prev = ss[2 * (ch - 1) + 1], start_prev = ss[2 * (ch - 1)];
        int start_next = (ch == nc - 1) ? n - 50 : ss[2 * (ch + 1)];
        float seg = 0.0f;
        for (int i = start; i < stop_prev; i++) {
            float fac = c->fac[ch - 1][i - start];
            float prev = dot_k(c->A[ch - 1] + (size_t)(i - start_prev) * k, pf + (ch - 1) * k, k);
            float cur = dot_k(c->A[ch] + (size_t)(i - start) * k, pf + ch * k, k);
            float p = prev * fac + cur * (1.0f - fac);
            float r = flux[i] - p * synth[i];
            seg = seg + (r * r) / var[i];
        }
        total = total + seg;
        seg = 0.0f;
        for (int i = stop_prev; i < start_next; i++) {
            float p = dot_k(c->A[ch] + (size_t)(i - start) * k, pf + ch * k, k);
            float r = flux[i] - p * synth[i];
            seg = seg + (r * r) / var[i];
        }
        total = total + seg;
    }
    *chi2 = (double)total / (double)(n - 100);
    free(y); free(pf);
    return OR_OK;
}

static int chunk_fit64(const or_cont *c, const float *flux, const float *var, const double *synth,
                       double *params, double *chi2) {
    cont_priv *pv = (cont_priv *)c->design;
    int n = c->n, k = c->p_order + 1, nc = c->n_chunks;
    double *y = malloc(sizeof(double) * n);
    for (int i = 0; i < n; i++) y[i] = (double)flux[i] / synth[i];
    double *pf = malloc(sizeof(double) * nc * k);
    for (int ch = 0; ch < nc; ch++) {
        int s = c->startstop[2 * ch];
        qr_solve(pv->qr64[ch], y + s, pf + ch * k);
    }
    if (params) memcpy(params, pf, sizeof(double) * nc * k);
    const int *ss = c->startstop;
    const double *wl = c->wl;
    double total = 0.0;
    for (int i = 50; i < ss[2]; i++) {
        double p = 0.0;
        for (int q = 0; q < k; q++) p += pv->A64[0][(size_t)i * k + q] * pf[q];
        double r = flux[i] - p * synth[i];
        total += r * r / var[i];
    }
    for (int ch = 1; ch < nc; ch++) {
        int start = ss[2 * ch], stop_prev = ss[2 * (ch - 1) + 1], start_prev = ss[2 * (ch - 1)];
        int start_next = (ch == nc - 1) ? n - 50 : ss[2 * (ch + 1)];
        for (int i = start; i < start_next; i++) {
            double cur = 0.0;
            for (int q = 0; q < k; q++) cur += pv->A64[ch][(size_t)(i - start) * k + q] * pf[ch * k + q];
            double p = cur;
            if (i < stop_prev) {
                double prev = 0.0;
                for (int q = 0; q < k; q++)
                    prev += pv->A64[ch - 1][(size_t)(i - start_prev) * k + q] * pf[(ch - 1) * k + q];
                double xx = (wl[i] - wl[start]) / (wl[stop_prev - 1] - wl[start]) * (0.0 - 1.0) + 1.0;
                xx = xx < 0 ? 0 : (xx > 1 ? 1 : xx);
                double sn = sin(xx * M_PI / 2.0);
                double fac = sn * sn;
                p = prev * fac + cur * (1.0 - fac);
            }
            double r = flux[i] - p * synth[i];
            total += r * r / var[i];
        }
    }
    *chi2 = total / (double)(n - 100);
    free(y); free(pf);
    return OR_OK;
}

/* LinearModelFitter::fit_continuum (continuum_fitting.rs:69-95) */
static int linear_fit32(const or_cont *c, const float *flux, const float *var, const float *synth,
                        float *params, double *chi2) {
    cont_priv *pv = (cont_priv *)c->A[1];
    const float *A = c->A[0];
    int n = c->n, k = c->ncols;
    float *b = malloc(sizeof(float) * n);
    double *bd = malloc(sizeof(double) * n), *xs = malloc(sizeof(double) * k);
    for (int i = 0; i < n; i++) { b[i] = flux[i] / synth[i]; bd[i] = b[i]; }
    qr_solve(pv->lin, bd, xs);
    float *sol = malloc(sizeof(float) * k);
    for (int i = 0; i < k; i++) sol[i] = (float)xs[i];
    if (params) memcpy(params, sol, sizeof(float) * k);
    /* model = A*sol (gemv column by column), residual = model - b */
    float *r = malloc(sizeof(float) * n);
    for (int i = 0; i < n; i++) {
        float y = 0.0f;
        for (int j = 0; j < k; j++) {
            float t = A[(size_t)i * k + j] * sol[j];
            y = (j == 0) ? t : t + y;
        }
        r[i] = y - b[i];
    }
    int len = n - 100;
    float *rv = malloc(sizeof(float) * len);
    for (int i = 0; i < len; i++) rv[i] = r[50 + i] / var[50 + i];
    float d = dot_na(rv, r + 50, len);
    *chi2 = (double)d / (double)n;
    free(b); free(bd); free(xs); free(sol); free(r); free(rv);
    return OR_OK;
}

static int linear_fit64(const or_cont *c, const float *flux, const float *var, const double *synth,
                        double *params, double *chi2) {
    cont_priv *pv = (cont_priv *)c->A[1];
    const float *A = c->A[0];
    int n = c->n, k = c->ncols;
    double *b = malloc(sizeof(double) * n), *xs = malloc(sizeof(double) * k);
    for (int i = 0; i < n; i++) b[i] = flux[i] / synth[i];
    qr_solve(pv->lin, b, xs);
    if (params) memcpy(params, xs, sizeof(double) * k);
    double acc = 0.0;
    for (int i = 50; i < n - 50; i++) {
        double y = 0.0;
        for (int j = 0; j < k; j++) y += (double)A[(size_t)i * k + j] * xs[j];
        double r = y - b[i];
        acc += r * r / var[i];
    }
    *chi2 = acc / (double)n;
    free(b); free(xs);
    return OR_OK;
}

/* FixedContinuum::fit_continuum (continuum_fitting.rs:442-474) */
static int fixed_fit32(const or_cont *c, const float *flux, const float *var, const float *synth,
                       float *params, double *chi2) {
    int n = c->n, ig = c->ignore, len = n - 2 * ig;
    float *r = malloc(sizeof(float) * n), *rv = malloc(sizeof(float) * (len > 0 ? len : 1));
    for (int i = 0; i < n; i++) r[i] = flux[i] - synth[i] * c->fixed[i];
    for (int i = 0; i < len; i++) rv[i] = r[ig + i] / var[ig + i];
    float d = dot_na(rv, r + ig, len);
    *chi2 = (double)d / (double)len;
    if (params) params[0] = 0.0f;
    free(r); free(rv);
    return OR_OK;
}

/* ConstantContinuum::fit_continuum (continuum_fitting.rs:485-500) */
static int constant_fit32(const or_cont *c, const float *flux, const float *var,
                          const float *synth, float *params, double *chi2) {
    int n = c->n;
    float sum = 0.0f;
    for (int i = 0; i < n; i++) sum = sum + flux[i] / synth[i];
    float mean = sum / (float)n;
    float *r = malloc(sizeof(float) * n), *rv = malloc(sizeof(float) * n);
    for (int i = 0; i < n; i++) r[i] = flux[i] - synth[i] * mean;
    for (int i = 0; i < n; i++) rv[i] = r[i] / var[i];
    float d = dot_na(rv, r, n);
    *chi2 = (double)d / (double)n;
    if (params) params[0] = mean;
    free(r); free(rv);
    return OR_OK;
}

int or_fit_continuum(const or_cont *c, const float *flux, const float *var, const float *synth,
                     float *params, double *chi2) {
    switch (c->kind) {
    case OR_CONT_CHUNK: return chunk_fit32(c, flux, var, synth, params, chi2);
    case OR_CONT_LINEAR: return linear_fit32(c, flux, var, synth, params, chi2);
    case OR_CONT_FIXED: return fixed_fit32(c, flux, var, synth, params, chi2);
    case OR_CONT_CONSTANT: return constant_fit32(c, flux, var, synth, params, chi2);
    }
    return OR_ERR_ARG;
}

int or_fit_continuum_f64(const or_cont *c, const float *flux, const float *var,
                         const double *synth, double *params, double *chi2) {
    int n = c->n;
    switch (c->kind) {
    case OR_CONT_CHUNK: return chunk_fit64(c, flux, var, synth, params, chi2);
    case OR_CONT_LINEAR: return linear_fit64(c, flux, var, synth, params, chi2);
    case OR_CONT_FIXED: {
        int ig = c->ignore;
        double acc = 0.0;
        for (int i = ig; i < n - ig; i++) {
            double r = flux[i] - synth[i] * c->fixed[i];
            acc += r * r / var[i];
        }
        *chi2 = acc / (double)(n - 2 * ig);
        if (params) params[0] = 0.0;
        return OR_OK;
    }
    case OR_CONT_CONSTANT: {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += flux[i] / synth[i];
        double mean = s / n, acc = 0.0;
        for (int i = 0; i < n; i++) {
            double r = flux[i] - synth[i] * mean;
            acc += r * r / var[i];
        }
        *chi2 = acc / (double)n;
        if (params) params[0] = mean;
        return OR_OK;
    }
    }
    return OR_ERR_ARG;
}

/* ChunkFitter::build_continuum_from_chunks (continuum_fitting.rs:228-278) and
 * the other modes' build_continuum */
int or_build_continuum(const or_cont *c, const float *params, float *out) {
    int n = c->n;
    if (c->kind == OR_CONT_CONSTANT) { out[0] = params[0]; return 1; }
    if (c->kind == OR_CONT_FIXED) { memcpy(out, c->fixed, sizeof(float) * n); return n; }
    if (c->kind == OR_CONT_LINEAR) {
        int k = c->ncols;
        for (int i = 0; i < n; i++) {
            float y = 0.0f;
            for (int j = 0; j < k; j++) {
                float t = c->A[0][(size_t)i * k + j] * params[j];
                y = (j == 0) ? t : t + y;
            }
            out[i] = y;
        }
        return n;
    }
    int k = c->p_order + 1, nc = c->n_chunks;
    memset(out, 0, sizeof(float) * n);
    for (int ch = 0; ch < nc; ch++) {
        int s = c->startstop[2 * ch], e = c->startstop[2 * ch + 1];
        for (int i = s; i < e; i++) {
            float poly = 0.0f;
            for (int j = 0; j < k; j++) {
                float t = c->A[ch][(size_t)(i - s) * k + j] * params[ch * k + j];
                poly = (j == 0) ? t : t + poly;
            }
            if (ch == 0) { out[i] = poly; continue; }
            int stop_prev = c->startstop[2 * (ch - 1) + 1];
            float fac = poly_blend(map_range_f((float)c->wl[i], (float)c->wl[stop_prev - 1],
                                               (float)c->wl[s], 0.0f, 1.0f));
            out[i] = out[i] * fac + poly * (1.0f - fac);
        }
    }
    return n;
}

/* CostFunction::cost (fitting.rs:313-326) */
int or_cost(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
            const float *var, const double lab[5], double *chi2, float *params) {
    float *synth = malloc(sizeof(float) * d->n_obs);
    int rc = or_produce_model(M, d, lab, synth);
    if (!rc) rc = or_fit_continuum(c, flux, var, synth, params, chi2);
    free(synth);
    return rc;
}

static int cost_f64_impl(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
                         const float *var, const double lab[5], double *chi2, double *params,
                         int qweight) {
    double *synth = malloc(sizeof(double) * d->n_obs);
    int rc = produce_model_f64_impl(M, d, lab, synth, qweight);
    if (!rc) rc = or_fit_continuum_f64(c, flux, var, synth, params, chi2);
    free(synth);
    return rc;
}

int or_cost_f64(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
                const float *var, const double lab[5], double *chi2) {
    return cost_f64_impl(M, d, c, flux, var, lab, chi2, NULL, 0);
}

int or_cost_f64q(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
                 const float *var, const double lab[5], double *chi2, double *params) {
    return cost_f64_impl(M, d, c, flux, var, lab, chi2, params, 1);
}

/* ======================================================================
 * PSO -- particleswarm.rs:57-307, bounds.rs:9-186
 * ====================================================================== */

/* Counter-based SplitMix64: output n of a SplitMix64 stream seeded with key. */
uint64_t or_rng_u64(uint64_t key, uint64_t ctr) {
    uint64_t z = key + (ctr + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double or_rng_uniform(uint64_t key, uint64_t ctr) {
    return (double)(or_rng_u64(key, ctr) >> 11) * 0x1.0p-53;
}

#define KEY_POS 0x51A5A0000000001ULL
#define KEY_VEL 0x51A5A0000000002ULL
#define KEY_ITER 0x51A5A0000000003ULL

/* random_uniform (particleswarm.rs:153-161) with this build's stream */
static double rand_range(double a, double b, double u) {
    if (a == b) return a;
    if (a < b) return a + (b - a) * u;
    return b + (a - b) * u;
}

/* SingleBounds::outer_limits + ConstrainedPSOBounds (bounds.rs:107-110,150-166) */
int or_bounds_outer(const or_bounds *b, double lo[5], double hi[5]) {
    const or_grid *g = b->g;
    lo[0] = g->teff[0]; hi[0] = g->teff[g->n_teff - 1];
    lo[1] = g->m[0]; hi[1] = g->m[g->n_m - 1];
    lo[2] = g->logg[0]; hi[2] = g->logg[g->n_logg - 1];
    lo[3] = b->vsini_lo; hi[3] = b->vsini_hi;
    lo[4] = b->rv_lo; hi[4] = b->rv_hi;
    for (int d = 0; d < 5; d++) {
        if (b->ckind[d] == 1) { lo[d] = b->ca[d]; hi[d] = b->ca[d]; }
        else if (b->ckind[d] == 2) { lo[d] = fmax(lo[d], b->ca[d]); hi[d] = fmin(hi[d], b->cb[d]); }
    }
    return OR_OK;
}

/* SingleBounds::is_within_bounds + constraints (bounds.rs:117-129,178-185) */
int or_bounds_within(const or_bounds *b, const double p[5]) {
    const or_grid *g = b->g;
    int ok = (p[1] >= g->m[0] && p[1] <= g->m[g->n_m - 1]) &
             or_is_teff_logg_between_bounds(g, p[0], p[2]);
    ok = ok && p[3] >= b->vsini_lo && p[3] <= b->vsini_hi && p[4] >= b->rv_lo && p[4] <= b->rv_hi;
    if (!ok) return 0;
    for (int d = 0; d < 5; d++) {
        if (b->ckind[d] == 1 && !(fabs(p[d] - b->ca[d]) <= 2.220446049250313e-16)) return 0;
        if (b->ckind[d] == 2 && !(p[d] >= b->ca[d] && p[d] <= b->cb[d])) return 0;
    }
    return 1;
}

/* PSOBounds::clamp_1d (bounds.rs:13-16,112-115,168-176) */
int or_bounds_clamp(const or_bounds *b, const double p[5], int d, double *out) {
    double lo, hi;
    if (d == 3) { lo = b->vsini_lo; hi = b->vsini_hi; }
    else if (d == 4) { lo = b->rv_lo; hi = b->rv_hi; }
    else {
        int rc = grid_limits_at(b->g, p, d, &lo, &hi);
        if (rc) return rc;
    }
    if (b->ckind[d] == 1) { lo = b->ca[d]; hi = b->ca[d]; }
    else if (b->ckind[d] == 2) { lo = fmax(lo, b->ca[d]); hi = fmin(hi, b->cb[d]); }
    if (lo > hi) return OR_ERR_PSO; /* f64::clamp panics on min > max */
    *out = clampd(p[d], lo, hi);
    return OR_OK;
}

/* initialize_positions_and_velocities (particleswarm.rs:137-150) with the
 * rejection sampler generate_random_within_bounds (bounds.rs:21-35) */
int or_pso_init_positions(const or_bounds *b, int P, uint64_t seed, double *pos, double *vel) {
    double lo[5], hi[5];
    or_bounds_outer(b, lo, hi);
    uint64_t kp = or_rng_u64(seed, KEY_POS), kv = or_rng_u64(seed, KEY_VEL);
    uint64_t attempt = 0;
    int got = 0;
    while (got < P) {
        double p[5];
        for (int d = 0; d < 5; d++)
            p[d] = rand_range(lo[d], hi[d], or_rng_uniform(kp, attempt * 5 + d));
        attempt++;
        if (attempt > 100000000ULL) return OR_ERR_PSO;
        if (or_bounds_within(b, p)) {
            memcpy(pos + 5 * got, p, sizeof(p));
            got++;
        }
    }
    for (int q = 0; q < P; q++)
        for (int d = 0; d < 5; d++) {
            double delta = hi[d] - lo[d];
            vel[5 * q + d] = rand_range(-delta, delta, or_rng_uniform(kv, (uint64_t)q * 5 + d));
        }
    return OR_OK;
}

typedef struct { double pos[5], vel[5], cost, bpos[5], bcost; } particle_t;

/* stable sort by cost; NaN compares as equal (particleswarm.rs:127-131) */
static void sort_particles(particle_t *p, int n) {
    for (int i = 1; i < n; i++) {
        particle_t t = p[i];
        int j = i - 1;
        while (j >= 0 && p[j].cost > t.cost) { p[j + 1] = p[j]; j--; }
        p[j + 1] = t;
    }
}

int or_pso_run(const or_bounds *b, const or_pso_settings *s, uint64_t seed, const double *init_pos,
               const double *init_vel, or_cost_fn fn, void *ctx, double best[5], double *best_cost,
               double *trace_best, long *n_evals) {
    int P = s->num_particles;
    double *pos = malloc(sizeof(double) * 5 * P), *vel = malloc(sizeof(double) * 5 * P);
    double *cost = malloc(sizeof(double) * P);
    int rc = or_pso_init_positions(b, P, seed, pos, vel);
    if (rc) goto done;
    if (init_pos) memcpy(pos, init_pos, sizeof(double) * 5 * P);
    if (init_vel) memcpy(vel, init_vel, sizeof(double) * 5 * P);
    long evals = 0;
    rc = fn(ctx, P, pos, cost);
    evals += P;
    if (rc) goto done;
    particle_t *pt = malloc(sizeof(particle_t) * P);
    for (int q = 0; q < P; q++) {
        memcpy(pt[q].pos, pos + 5 * q, sizeof(double) * 5);
        memcpy(pt[q].vel, vel + 5 * q, sizeof(double) * 5);
        memcpy(pt[q].bpos, pos + 5 * q, sizeof(double) * 5);
        pt[q].cost = pt[q].bcost = cost[q];
    }
    sort_particles(pt, P);
    double g[5], gc = pt[0].cost;
    memcpy(g, pt[0].pos, sizeof(g));
    if (trace_best) { trace_best[0] = gc; memcpy(trace_best + 1, g, sizeof(g)); }
    uint64_t ki = or_rng_u64(seed, KEY_ITER);
    for (int it = 1; it <= s->max_iters && !rc; it++) {
        for (int q = 0; q < P && !rc; q++) {
            for (int d = 0; d < 5; d++) {
                uint64_t c0 = (((uint64_t)it * P + q) * 5 + d) * 2;
                double momentum = pt[q].vel[d] * s->inertia;
                double to_opt = pt[q].bpos[d] - pt[q].pos[d];
                double pull = rand_range(0.0, to_opt, or_rng_uniform(ki, c0)) * s->cognitive;
                double to_g = g[d] - pt[q].pos[d];
                double pullg = rand_range(0.0, to_g, or_rng_uniform(ki, c0 + 1)) * s->social;
                pt[q].vel[d] = momentum + pull + pullg;
                double np[5];
                memcpy(np, pt[q].pos, sizeof(np));
                np[d] += pt[q].vel[d];
                double cl;
                rc = or_bounds_clamp(b, np, d, &cl);
                if (rc) break;
                pt[q].pos[d] = cl;
                if (!or_bounds_within(b, pt[q].pos)) { rc = OR_ERR_PSO; break; }
            }
        }
        if (rc) break;
        for (int q = 0; q < P; q++) memcpy(pos + 5 * q, pt[q].pos, sizeof(double) * 5);
        rc = fn(ctx, P, pos, cost);
        evals += P;
        if (rc) break;
        for (int q = 0; q < P; q++) {
            pt[q].cost = cost[q];
            if (pt[q].cost < pt[q].bcost) {
                memcpy(pt[q].bpos, pt[q].pos, sizeof(double) * 5);
                pt[q].bcost = pt[q].cost;
                if (pt[q].cost < gc) {
                    memcpy(g, pt[q].pos, sizeof(g));
                    gc = pt[q].cost;
                }
            }
        }
        if (trace_best) { trace_best[6 * it] = gc; memcpy(trace_best + 6 * it + 1, g, sizeof(g)); }
    }
    if (!rc) { memcpy(best, g, sizeof(g)); *best_cost = gc; }
    if (n_evals) *n_evals = evals;
    free(pt);
done:
    free(pos); free(vel); free(cost);
    return rc;
}

/* ---- model cost over a population, optionally threaded (rayon stand-in) ---- */
typedef struct {
    const or_model *M; const or_disp *d; const or_cont *c;
    const float *flux, *var;
    int n_threads;
    int f64q;  /* 1: FP64 over the reference's f32-stored values (or_cost_f64q) */
} fit_ctx;

typedef struct { fit_ctx *fc; const double *pos; double *cost; int lo, hi, rc; } job_t;

static void *cost_worker(void *arg) {
    job_t *j = arg;
    for (int q = j->lo; q < j->hi; q++) {
        int rc = j->fc->f64q
                     ? or_cost_f64q(j->fc->M, j->fc->d, j->fc->c, j->fc->flux, j->fc->var,
                                    j->pos + 5 * q, j->cost + q, NULL)
                     : or_cost(j->fc->M, j->fc->d, j->fc->c, j->fc->flux, j->fc->var,
                               j->pos + 5 * q, j->cost + q, NULL);
        if (rc && !j->rc) j->rc = rc;
    }
    return NULL;
}

static int model_cost_fn(void *ctx, int n, const double *pos, double *cost) {
    fit_ctx *fc = ctx;
    int T = fc->n_threads < 1 ? 1 : fc->n_threads;
    if (T > n) T = n;
    pthread_t th[256];
    job_t jobs[256];
    if (T > 256) T = 256;
    for (int t = 0; t < T; t++) {
        jobs[t].fc = fc; jobs[t].pos = pos; jobs[t].cost = cost; jobs[t].rc = 0;
        jobs[t].lo = (int)((long)n * t / T); jobs[t].hi = (int)((long)n * (t + 1) / T);
        if (T > 1) pthread_create(&th[t], NULL, cost_worker, &jobs[t]);
        else cost_worker(&jobs[t]);
    }
    int rc = 0;
    for (int t = 0; t < T; t++) {
        if (T > 1) pthread_join(th[t], NULL);
        if (jobs[t].rc && !rc) rc = jobs[t].rc;
    }
    return rc;
}

/* SingleFitter::fit (fitting.rs:399-458) */
int or_fit(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
           const float *var, const or_bounds *b, const or_pso_settings *s, uint64_t seed,
           const double *init_pos, int n_threads, double best[5], double *best_cost,
           float *params, long *n_evals) {
    fit_ctx fc = {M, d, c, flux, var, n_threads, 0};
    int rc = or_pso_run(b, s, seed, init_pos, NULL, model_cost_fn, &fc, best, best_cost, NULL,
                        n_evals);
    if (rc) return rc;
    double chi2;
    rc = or_cost(M, d, c, flux, var, best, &chi2, params);
    if (n_evals) *n_evals += 1;
    return rc;
}

/* CostFunction::cost over many label vectors on n_threads threads (the
 * reference's rayon par_iter over particles, src/fitting.rs:328-330). */
int or_cost_batch(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
                  const float *var, const double *labels, int n, int n_threads, double *chi2) {
    fit_ctx fc = {M, d, c, flux, var, n_threads, 0};
    return model_cost_fn(&fc, n, labels, chi2);
}

int or_cost_batch_f64q(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
                       const float *var, const double *labels, int n, int n_threads,
                       double *chi2) {
    fit_ctx fc = {M, d, c, flux, var, n_threads, 1};
    return model_cost_fn(&fc, n, labels, chi2);
}

/* ======================================================================
 * Uncertainty -- fitting.rs:333-372 (UncertaintyCostFunction), 482-568
 * (compute_uncertainty); the root finder is argmin 0.10.0's BrentRoot.
 * ====================================================================== */

int or_brent_root(or_scalar_fn fn, void *ctx, double min, double max, double tol, double *root,
                  int *n_evals) {
    const double eps = 2.220446049250313e-16;
    double a = min, b = max, c = max, d = NAN, e = NAN, fa, fb, fc;
    int rc, n = 0;
    if ((rc = fn(ctx, a, &fa))) return rc;
    n++;
    if ((rc = fn(ctx, b, &fb))) return rc;
    n++;
    if (fa * fb > 0.0) return OR_ERR_SIGN;
    fc = fb;
    double best = b, best_cost = fabs(fb);
    for (;;) {
        if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) {
            c = a;
            fc = fa;
            d = b - a;
            e = d;
        }
        if (fabs(fc) < fabs(fb)) {
            a = b; b = c; c = a;
            fa = fb; fb = fc; fc = fa;
        }
        const double tol1 = 2.0 * eps * fabs(b) + 0.5 * tol;
        const double xm = 0.5 * (c - b);
        if (fabs(xm) <= tol1 || fb == 0.0) {
            if (fabs(fb) < best_cost) { best = b; best_cost = fabs(fb); }
            break;
        }
        if (fabs(e) >= tol1 && fabs(fa) > fabs(fb)) {
            const double s = fb / fa;
            double p, q;
            if (a == c) {
                p = 2.0 * xm * s;
                q = 1.0 - s;
            } else {
                const double qq = fa / fc, r = fb / fc;
                p = s * (2.0 * xm * qq * (qq - r) - (b - a) * (r - 1.0));
                q = (qq - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0.0) q = -q;
            p = fabs(p);
            const double min1 = 3.0 * xm * q - fabs(tol1 * q), min2 = fabs(e * q);
            if (2.0 * p < (min1 < min2 ? min1 : min2)) {
                e = d;
                d = p / q;
            } else {
                d = xm;
                e = d;
            }
        } else {
            d = xm;
            e = d;
        }
        a = b;
        fa = fb;
        b += fabs(d) > tol1 ? d : (xm > 0.0 ? tol1 : -tol1);
        if ((rc = fn(ctx, b, &fb))) return rc;
        n++;
        if (fabs(fb) < best_cost) { best = b; best_cost = fabs(fb); }
    }
    *root = best;
    if (n_evals) *n_evals = n;
    return OR_OK;
}

typedef struct {
    const or_model *M;
    const or_disp *d;
    const or_cont *c;
    const float *flux, *var;
    double label[5];
    int index;
    double target, radius;
    float *synth;
} unc_ctx;

/* UncertaintyCostFunction::cost (fitting.rs:350-371) */
static int unc_cost(void *vctx, double x, double *f) {
    unc_ctx *u = (unc_ctx *)vctx;
    double lab[5];
    memcpy(lab, u->label, sizeof(lab));
    lab[u->index] += x * u->radius;
    double chi2;
    int rc = or_produce_model(u->M, u->d, lab, u->synth);
    if (!rc) rc = or_fit_continuum(u->c, u->flux, u->var, u->synth, NULL, &chi2);
    if (rc) return rc;
    *f = chi2 - u->target;
    return OR_OK;
}

int or_compute_uncertainty(const or_model *M, const or_disp *d, const or_cont *c,
                           const float *flux, const float *var, double spec_res,
                           const double label[5], const double *search_radius, double out[10],
                           int status[10]) {
    static const double def[5] = {2000.0, 0.3, 0.3, 40.0, 40.0};
    double sr[5];
    memcpy(sr, search_radius ? search_radius : def, sizeof(sr));
    sr[0] *= label[0];
    sr[3] = fmax(label[3] * sr[3], 100.0);
    unc_ctx u = {M, d, c, flux, var, {0}, 0, 0.0, 0.0, NULL};
    memcpy(u.label, label, sizeof(u.label));
    u.synth = malloc(sizeof(float) * d->n_obs);
    double best_chi2;
    int rc = or_produce_model(M, d, label, u.synth);
    if (!rc) rc = or_fit_continuum(c, flux, var, u.synth, NULL, &best_chi2);
    if (rc) { free(u.synth); return rc; }
    const double first = d->wl[0], last = d->wl[d->n_obs - 1];
    const double n = 4.0 * spec_res * (last - first) / (first + last);
    u.target = best_chi2 * (1.0 + sqrt(2.0 / n));
    /* SingleBounds over the grid with vsini (0, 1e4) and RV (-1e3, 1e3) (fitting.rs:519) */
    or_bounds b = {M->g, 0.0, 1e4, -1e3, 1e3, {0, 0, 0, 0, 0}, {0}, {0}};
    for (int i = 0; i < 5; i++) {
        u.index = i;
        u.radius = sr[i];
        for (int right = 0; right < 2; right++) {
            double bl[5], bound, root;
            memcpy(bl, label, sizeof(bl));
            bl[i] += right ? sr[i] : -sr[i];
            int r = or_bounds_clamp(&b, bl, i, &bound);
            if (!r) {
                const double resc = (bound - label[i]) / sr[i];
                r = right ? or_brent_root(unc_cost, &u, 0.0, fmin(resc, 1.0), 1e-3, &root, NULL)
                          : or_brent_root(unc_cost, &u, resc, 0.0, 1e-3, &root, NULL);
            }
            status[2 * i + right] = r;
            out[2 * i + right] = r ? NAN : (right ? root * sr[i] : -root * sr[i]);
        }
    }
    free(u.synth);
    return OR_OK;
}

/* ======================================================================
 * Binary stars -- interpolate.rs:168-212,536-620,796-890,985-1030,
 * cubic.rs:6-10,246-262, convolve_rv.rs:606-703, fitting.rs:571-728,
 * bounds.rs:188-276
 * ====================================================================== */

/* Range::find_neighbors_linear (interpolate.rs:168-212) */
static int find_neighbors_linear(const double *v, int n, double x, int left, int right,
                                 int out[2]) {
    int r = right_index(v, n, x);
    if (r >= 0) {
        if (r < left || r > right) return OR_ERR_BOUNDS;
        if (r == right) { out[0] = r - 1; out[1] = r; }
        else { out[0] = r; out[1] = r + 1; }
        return OR_OK;
    }
    int i = -r - 1;
    if (i <= left || i > right) return OR_ERR_BOUNDS;
    out[0] = i - 1; out[1] = i;
    return OR_OK;
}

/* calculate_factors_linear (cubic.rs:6-10) */
static void factors_linear(double x, double x0, double x1, double f[2]) {
    f[0] = (x - x1) / (x0 - x1);
    f[1] = (x - x0) / (x1 - x0);
}

int or_local_weights_linear(const or_grid *g, double teff, double m, double logg, int idx[8],
                            float w[8], double w64[8]) {
    /* get_local_grid_linear (interpolate.rs:536-620) */
    if (!or_is_teff_logg_between_bounds(g, teff, logg)) return OR_ERR_BOUNDS;
    if (!(m >= g->m[0] && m <= g->m[g->n_m - 1])) return OR_ERR_BOUNDS;
    int tl, th, ll, lh, rc;
    if ((rc = or_get_teff_index_limits_at(g, logg, &tl, &th))) return rc;
    if ((rc = or_get_logg_index_limits_at(g, teff, &ll, &lh))) return rc;
    int ti[2], mi[2], li[2];
    if ((rc = find_neighbors_linear(g->teff, g->n_teff, teff, tl, th, ti))) return rc;
    if ((rc = find_neighbors_linear(g->m, g->n_m, m, 0, g->n_m - 1, mi))) return rc;
    if ((rc = find_neighbors_linear(g->logg, g->n_logg, logg, ll, lh, li))) return rc;
    /* calculate_interpolation_coefficients_linear (cubic.rs:246-262): the cartesian
     * product teff x logg x m, (a * b) * c in f64, then f32 */
    double ft[2], fl[2], fm[2];
    factors_linear(teff, g->teff[ti[0]], g->teff[ti[1]], ft);
    factors_linear(logg, g->logg[li[0]], g->logg[li[1]], fl);
    factors_linear(m, g->m[mi[0]], g->m[mi[1]], fm);
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
                int q = (a * 2 + b) * 2 + c;
                /* find_spectrum(teff, m, logg) (model_fetchers.rs:234-236); usize
                 * arithmetic: a logg index below the column's first panics */
                if (li[b] < g->logg_lo[ti[a]]) return OR_ERR_INDEX;
                int id = or_grid_model_index(g, ti[a], mi[c], li[b]);
                if (id >= g->n_models) return OR_ERR_INDEX;
                idx[q] = id;
                double p = (ft[a] * fl[b]) * fm[c];
                if (w) w[q] = (float)p;
                if (w64) w64[q] = p;
            }
    return OR_OK;
}

/* interpolate_linear (interpolate.rs:985-1030): every pixel is x / 65535 * factor (no
 * 128-pixel batch distinction), then the 8-column gemv in column order */
int or_interpolate_linear_f32(const or_grid *g, const uint16_t *spectra, const float *factors,
                              int W, double teff, double m, double logg, float *out) {
    int idx[8];
    float w[8];
    int rc = or_local_weights_linear(g, teff, m, logg, idx, w, NULL);
    if (rc) return rc;
    float fac[8];
    for (int j = 0; j < 8; j++) fac[j] = factors ? factors[idx[j]] : (float)(1.0 / 65535.0);
    for (int x = 0; x < W; x++) {
        float y = 0.0f;
        for (int j = 0; j < 8; j++) {
            float Mv = (float)spectra[(size_t)idx[j] * W + x] / 65535.0f * fac[j];
            float t = Mv * w[j];
            y = (j == 0) ? t : t + y;
        }
        out[x] = y;
    }
    return OR_OK;
}

int or_interpolate_linear_f64(const or_grid *g, const uint16_t *spectra, const float *factors,
                              int W, double teff, double m, double logg, double *out) {
    int idx[8];
    double w[8];
    int rc = or_local_weights_linear(g, teff, m, logg, idx, NULL, w);
    if (rc) return rc;
    for (int x = 0; x < W; x++) {
        double y = 0.0;
        for (int j = 0; j < 8; j++) {
            double f = factors ? (double)factors[idx[j]] : 1.0 / 65535.0;
            y += (double)spectra[(size_t)idx[j] * W + x] / 65535.0 * f * w[j];
        }
        out[x] = y;
    }
    return OR_OK;
}

/* DVector::mean of nalgebra 0.32.6, restated as a sequential f32 fold and one division
 * (ULP behaviour unpinned: the crate is absent from the reference tree) */
static float mean_f32(const float *v, int n) {
    float s = 0.0f;
    for (int i = 0; i < n; i++) s = s + v[i];
    return s / (float)n;
}

/* produce_binary_components + shift_resample_and_add_binary_components
 * (interpolate.rs:822-890, convolve_rv.rs:606-690, Logspace synthetic grid) */
int or_binary_model_norm(const or_model *M, const or_model *CM, const or_disp *d,
                         const double lab1[5], const double lab2[5], float light_ratio,
                         float *out) {
    int W = M->sg.n, rc;
    if (CM->sg.n != W) return OR_ERR_ARG;
    float *interp = malloc(sizeof(float) * W);
    float *c1 = malloc(sizeof(float) * W), *c2 = malloc(sizeof(float) * W);
    seg32 s1 = {0}, s2 = {0};
    /* components in the reference's order: model 1, continuum 1, model 2, continuum 2 */
    rc = or_interpolate_f32(M->g, M->spectra, M->factors, W, lab1[0], lab1[1], lab1[2], interp);
    if (!rc) rc = rot_conv32(M, interp, lab1[3], &s1);
    if (!rc) rc = disp_conv32(d, &s1);
    if (!rc) rc = or_interpolate_linear_f32(CM->g, CM->spectra, CM->factors, W, lab1[0], lab1[1],
                                            lab1[2], c1);
    if (!rc) rc = or_interpolate_f32(M->g, M->spectra, M->factors, W, lab2[0], lab2[1], lab2[2],
                                     interp);
    if (!rc) rc = rot_conv32(M, interp, lab2[3], &s2);
    if (!rc) rc = disp_conv32(d, &s2);
    if (!rc) rc = or_interpolate_linear_f32(CM->g, CM->spectra, CM->factors, W, lab2[0], lab2[1],
                                            lab2[2], c2);
    if (rc) goto done;
    {
        float lr = light_ratio * mean_f32(c1, W) / mean_f32(c2, W);
        double sf1 = 1.0 - lab1[4] / C_KMS, sf2 = 1.0 - lab2[4] / C_KMS;
        double first = pow(10.0, M->sg.first);
        double last = pow(10.0, M->sg.first + M->sg.step * (double)W);
        if (d->wl[0] * fmin(sf1, sf2) < first) { rc = OR_ERR_RANGE; goto done; }
        if (d->wl[d->n_obs - 1] * fmax(sf1, sf2) > last) { rc = OR_ERR_RANGE; goto done; }
        double sl1 = log10(sf1) / M->sg.step, sl2 = log10(sf2) / M->sg.step;
        for (int i = 0; i < d->n_obs; i++) {
            double x1 = d->fidx[i] + sl1, x2 = d->fidx[i] + sl2;
            double f1 = floor(x1), f2 = floor(x2);
            long j1 = f1 > 0 ? (long)f1 : 0, j2 = f2 > 0 ? (long)f2 : 0;
            if (s1.start + j1 + 1 >= s1.alen || s2.start + j2 + 1 >= s2.alen || j1 + 1 >= W ||
                j2 + 1 >= W) { rc = OR_ERR_INDEX; goto done; }
            float w1 = (float)x1 - (float)j1, w2 = (float)x2 - (float)j2;
            float n1 = (1.0f - w1) * s1.a[s1.start + j1] + w1 * s1.a[s1.start + j1 + 1];
            float k1 = (1.0f - w1) * c1[j1] + w1 * c1[j1 + 1];
            float n2 = (1.0f - w2) * s2.a[s2.start + j2] + w2 * s2.a[s2.start + j2 + 1];
            float k2 = (1.0f - w2) * c2[j2] + w2 * c2[j2 + 1];
            out[i] = (n1 * k1 + lr * n2 * k2) / (k1 + lr * k2);
        }
    }
done:
    free(interp); free(c1); free(c2); free(s1.a); free(s2.a);
    return rc;
}

/* BinaryCostFunction::cost (fitting.rs:581-604): labels (star 1, star 2, light ratio) */
int or_binary_cost(const or_model *M, const or_model *CM, const or_disp *d, const or_cont *c,
                   const float *flux, const float *var, const double lab[11], double *chi2,
                   float *params) {
    float *synth = malloc(sizeof(float) * d->n_obs);
    int rc = or_binary_model_norm(M, CM, d, lab, lab + 5, (float)lab[10], synth);
    if (!rc) rc = or_fit_continuum(c, flux, var, synth, params, chi2);
    free(synth);
    return rc;
}

/* BinaryBounds (bounds.rs:188-276) + ConstrainedPSOBounds (bounds.rs:64-132) */
static int binb_limits(const or_binbounds *b, const double p[11], int i, double *lo, double *hi) {
    int rc = OR_OK;
    switch (i) {
    case 3: case 8: *lo = b->vsini_lo; *hi = b->vsini_hi; break;
    case 4: case 9: *lo = b->rv_lo; *hi = b->rv_hi; break;
    case 10: *lo = b->lr_lo; *hi = b->lr_hi; break;
    default:
        rc = i < 5 ? grid_limits_at(b->g, p, i, lo, hi) : grid_limits_at(b->g, p + 5, i - 5, lo, hi);
    }
    if (rc) return rc;
    if (b->ckind[i] == 1) { *lo = b->ca[i]; *hi = b->ca[i]; }
    else if (b->ckind[i] == 2) { *lo = fmax(*lo, b->ca[i]); *hi = fmin(*hi, b->cb[i]); }
    return OR_OK;
}

int or_binbounds_outer(const or_binbounds *b, double lo[11], double hi[11]) {
    const or_grid *g = b->g;
    for (int s = 0; s < 2; s++) {
        lo[5 * s] = g->teff[0]; hi[5 * s] = g->teff[g->n_teff - 1];
        lo[5 * s + 1] = g->m[0]; hi[5 * s + 1] = g->m[g->n_m - 1];
        lo[5 * s + 2] = g->logg[0]; hi[5 * s + 2] = g->logg[g->n_logg - 1];
        lo[5 * s + 3] = b->vsini_lo; hi[5 * s + 3] = b->vsini_hi;
        lo[5 * s + 4] = b->rv_lo; hi[5 * s + 4] = b->rv_hi;
    }
    lo[10] = b->lr_lo; hi[10] = b->lr_hi;
    for (int d = 0; d < 11; d++) {
        if (b->ckind[d] == 1) { lo[d] = b->ca[d]; hi[d] = b->ca[d]; }
        else if (b->ckind[d] == 2) { lo[d] = fmax(lo[d], b->ca[d]); hi[d] = fmin(hi[d], b->cb[d]); }
    }
    return OR_OK;
}

int or_binbounds_within(const or_binbounds *b, const double p[11]) {
    const or_grid *g = b->g;
    int ok = 1;
    for (int s = 0; s < 2 && ok; s++) {
        const double *q = p + 5 * s;
        ok = or_is_teff_logg_between_bounds(g, q[0], q[2]) && q[1] >= g->m[0] &&
             q[1] <= g->m[g->n_m - 1];
    }
    ok = ok && p[3] >= b->vsini_lo && p[3] <= b->vsini_hi && p[4] >= b->rv_lo &&
         p[4] <= b->rv_hi && p[8] >= b->vsini_lo && p[8] <= b->vsini_hi && p[9] >= b->rv_lo &&
         p[9] <= b->rv_hi && p[10] >= b->lr_lo && p[10] <= b->lr_hi;
    if (!ok) return 0;
    for (int d = 0; d < 11; d++) {
        if (b->ckind[d] == 1 && !(fabs(p[d] - b->ca[d]) <= 2.220446049250313e-16)) return 0;
        if (b->ckind[d] == 2 && !(p[d] >= b->ca[d] && p[d] <= b->cb[d])) return 0;
    }
    return 1;
}

int or_binbounds_clamp(const or_binbounds *b, const double p[11], int d, double *out) {
    double lo, hi;
    int rc = binb_limits(b, p, d, &lo, &hi);
    if (rc) return OR_ERR_PSO;
    if (lo > hi) return OR_ERR_PSO;
    *out = clampd(p[d], lo, hi);
    return OR_OK;
}

/* BinaryBoundsWithoutRV (bounds.rs:278-345): (teff, m, logg, vsini) x 2 + light ratio */
static int binb9_limits(const or_binbounds *b, const double p[9], int i, double *lo, double *hi) {
    int rc = OR_OK;
    switch (i) {
    case 3: case 7: *lo = b->vsini_lo; *hi = b->vsini_hi; break;
    case 8: *lo = b->lr_lo; *hi = b->lr_hi; break;
    default:
        rc = i < 3 ? grid_limits_at(b->g, p, i, lo, hi) : grid_limits_at(b->g, p + 4, i - 4, lo, hi);
    }
    if (rc) return rc;
    if (b->ckind[i] == 1) { *lo = b->ca[i]; *hi = b->ca[i]; }
    else if (b->ckind[i] == 2) { *lo = fmax(*lo, b->ca[i]); *hi = fmin(*hi, b->cb[i]); }
    return OR_OK;
}

static void binb9_outer(const or_binbounds *b, double lo[9], double hi[9]) {
    const or_grid *g = b->g;
    for (int s = 0; s < 2; s++) {
        lo[4 * s] = g->teff[0]; hi[4 * s] = g->teff[g->n_teff - 1];
        lo[4 * s + 1] = g->m[0]; hi[4 * s + 1] = g->m[g->n_m - 1];
        lo[4 * s + 2] = g->logg[0]; hi[4 * s + 2] = g->logg[g->n_logg - 1];
        lo[4 * s + 3] = b->vsini_lo; hi[4 * s + 3] = b->vsini_hi;
    }
    lo[8] = b->lr_lo; hi[8] = b->lr_hi;
    for (int d = 0; d < 9; d++) {
        if (b->ckind[d] == 1) { lo[d] = b->ca[d]; hi[d] = b->ca[d]; }
        else if (b->ckind[d] == 2) { lo[d] = fmax(lo[d], b->ca[d]); hi[d] = fmin(hi[d], b->cb[d]); }
    }
}

static int binb9_within(const or_binbounds *b, const double p[9]) {
    const or_grid *g = b->g;
    int ok = 1;
    for (int s = 0; s < 2 && ok; s++) {
        const double *q = p + 4 * s;
        ok = or_is_teff_logg_between_bounds(g, q[0], q[2]) && q[1] >= g->m[0] &&
             q[1] <= g->m[g->n_m - 1];
    }
    ok = ok && p[3] >= b->vsini_lo && p[3] <= b->vsini_hi && p[7] >= b->vsini_lo &&
         p[7] <= b->vsini_hi && p[8] >= b->lr_lo && p[8] <= b->lr_hi;
    if (!ok) return 0;
    for (int d = 0; d < 9; d++) {
        if (b->ckind[d] == 1 && !(fabs(p[d] - b->ca[d]) <= 2.220446049250313e-16)) return 0;
        if (b->ckind[d] == 2 && !(p[d] >= b->ca[d] && p[d] <= b->cb[d])) return 0;
    }
    return 1;
}

/* BinaryRVBounds (bounds.rs:347-372): two RVs in rv_range */
static void binrv_outer(const or_binbounds *b, double lo[2], double hi[2]) {
    for (int d = 0; d < 2; d++) {
        lo[d] = b->rv_lo; hi[d] = b->rv_hi;
        if (b->ckind[d] == 1) { lo[d] = b->ca[d]; hi[d] = b->ca[d]; }
        else if (b->ckind[d] == 2) { lo[d] = fmax(lo[d], b->ca[d]); hi[d] = fmin(hi[d], b->cb[d]); }
    }
}

static int binrv_within(const or_binbounds *b, const double p[2]) {
    if (!(p[0] >= b->rv_lo && p[0] <= b->rv_hi && p[1] >= b->rv_lo && p[1] <= b->rv_hi)) return 0;
    for (int d = 0; d < 2; d++) {
        if (b->ckind[d] == 1 && !(fabs(p[d] - b->ca[d]) <= 2.220446049250313e-16)) return 0;
        if (b->ckind[d] == 2 && !(p[d] >= b->ca[d] && p[d] <= b->cb[d])) return 0;
    }
    return 1;
}

/* the three binary searches: D = 11 (BinaryBounds), 9 (without RV), 2 (the RVs) */
static void bn_outer(const or_binbounds *b, int D, double *lo, double *hi) {
    if (D == 11) or_binbounds_outer(b, lo, hi);
    else if (D == 9) binb9_outer(b, lo, hi);
    else binrv_outer(b, lo, hi);
}

static int bn_within(const or_binbounds *b, int D, const double *p) {
    return D == 11 ? or_binbounds_within(b, p) : D == 9 ? binb9_within(b, p) : binrv_within(b, p);
}

static int bn_clamp(const or_binbounds *b, int D, const double *p, int d, double *out) {
    double lo, hi;
    int rc;
    if (D == 11) rc = binb_limits(b, p, d, &lo, &hi);
    else if (D == 9) rc = binb9_limits(b, p, d, &lo, &hi);
    else {
        lo = b->rv_lo; hi = b->rv_hi; rc = OR_OK;
        if (b->ckind[d] == 1) { lo = b->ca[d]; hi = b->ca[d]; }
        else if (b->ckind[d] == 2) { lo = fmax(lo, b->ca[d]); hi = fmin(hi, b->cb[d]); }
    }
    if (rc || lo > hi) return OR_ERR_PSO;
    *out = clampd(p[d], lo, hi);
    return OR_OK;
}

typedef struct { double pos[11], vel[11], cost, bpos[11], bcost; } particle11_t;

static void sort_particles11(particle11_t *p, int n) {
    for (int i = 1; i < n; i++) {
        particle11_t t = p[i];
        int j = i - 1;
        while (j >= 0 && p[j].cost > t.cost) { p[j + 1] = p[j]; j--; }
        p[j + 1] = t;
    }
}

/* The PSO of particleswarm.rs over D = 11, 9 or 2 binary parameters (BinaryFitter,
 * BinaryTimeriesKnownRVFitter, BinaryRVFitter); the same draws and update order as
 * or_pso_run with D parameters */
int or_pso_run_d(const or_binbounds *b, int D, const or_pso_settings *s, uint64_t seed,
                 const double *init_pos, or_cost_fn fn, void *ctx, double *best,
                 double *best_cost, double *trace_best, long *n_evals) {
    if (D != 11 && D != 9 && D != 2) return OR_ERR_ARG;
    const int P = s->num_particles;
    double lo[11], hi[11];
    bn_outer(b, D, lo, hi);
    double *pos = malloc(sizeof(double) * D * P), *vel = malloc(sizeof(double) * D * P);
    double *cost = malloc(sizeof(double) * P);
    particle11_t *pt = malloc(sizeof(particle11_t) * P);
    int rc = OR_OK;
    uint64_t kp = or_rng_u64(seed, KEY_POS), kv = or_rng_u64(seed, KEY_VEL);
    uint64_t attempt = 0;
    for (int got = 0; got < P;) {
        double p[11];
        for (int d = 0; d < D; d++)
            p[d] = rand_range(lo[d], hi[d], or_rng_uniform(kp, attempt * D + d));
        attempt++;
        if (attempt > 100000000ULL) { rc = OR_ERR_PSO; goto done; }
        if (bn_within(b, D, p)) { memcpy(pos + D * got, p, sizeof(double) * D); got++; }
    }
    for (int q = 0; q < P; q++)
        for (int d = 0; d < D; d++) {
            double delta = hi[d] - lo[d];
            vel[D * q + d] = rand_range(-delta, delta, or_rng_uniform(kv, (uint64_t)q * D + d));
        }
    if (init_pos) memcpy(pos, init_pos, sizeof(double) * D * P);
    long evals = 0;
    rc = fn(ctx, P, pos, cost);
    evals += P;
    if (rc) goto done;
    for (int q = 0; q < P; q++) {
        memcpy(pt[q].pos, pos + D * q, sizeof(double) * D);
        memcpy(pt[q].vel, vel + D * q, sizeof(double) * D);
        memcpy(pt[q].bpos, pos + D * q, sizeof(double) * D);
        pt[q].cost = pt[q].bcost = cost[q];
    }
    sort_particles11(pt, P);
    double g[11], gc = pt[0].cost;
    memcpy(g, pt[0].pos, sizeof(double) * D);
    if (trace_best) { trace_best[0] = gc; memcpy(trace_best + 1, g, sizeof(double) * D); }
    uint64_t ki = or_rng_u64(seed, KEY_ITER);
    for (int it = 1; it <= s->max_iters && !rc; it++) {
        for (int q = 0; q < P && !rc; q++) {
            for (int d = 0; d < D; d++) {
                uint64_t c0 = (((uint64_t)it * P + q) * D + d) * 2;
                double momentum = pt[q].vel[d] * s->inertia;
                double to_opt = pt[q].bpos[d] - pt[q].pos[d];
                double pull = rand_range(0.0, to_opt, or_rng_uniform(ki, c0)) * s->cognitive;
                double to_g = g[d] - pt[q].pos[d];
                double pullg = rand_range(0.0, to_g, or_rng_uniform(ki, c0 + 1)) * s->social;
                pt[q].vel[d] = momentum + pull + pullg;
                double np[11];
                memcpy(np, pt[q].pos, sizeof(double) * D);
                np[d] += pt[q].vel[d];
                double cl;
                rc = bn_clamp(b, D, np, d, &cl);
                if (rc) break;
                pt[q].pos[d] = cl;
                if (!bn_within(b, D, pt[q].pos)) { rc = OR_ERR_PSO; break; }
            }
        }
        if (rc) break;
        for (int q = 0; q < P; q++) memcpy(pos + D * q, pt[q].pos, sizeof(double) * D);
        rc = fn(ctx, P, pos, cost);
        evals += P;
        if (rc) break;
        for (int q = 0; q < P; q++) {
            pt[q].cost = cost[q];
            if (pt[q].cost < pt[q].bcost) {
                memcpy(pt[q].bpos, pt[q].pos, sizeof(double) * D);
                pt[q].bcost = pt[q].cost;
                if (pt[q].cost < gc) {
                    memcpy(g, pt[q].pos, sizeof(double) * D);
                    gc = pt[q].cost;
                }
            }
        }
        if (trace_best) {
            trace_best[(D + 1) * it] = gc;
            memcpy(trace_best + (D + 1) * it + 1, g, sizeof(double) * D);
        }
    }
    if (!rc) { memcpy(best, g, sizeof(double) * D); *best_cost = gc; }
    if (n_evals) *n_evals = evals;
done:
    free(pos); free(vel); free(cost); free(pt);
    return rc;
}

int or_pso_run_binary(const or_binbounds *b, const or_pso_settings *s, uint64_t seed,
                      const double *init_pos, or_cost_fn fn, void *ctx, double best[11],
                      double *best_cost, double *trace_best, long *n_evals) {
    return or_pso_run_d(b, 11, s, seed, init_pos, fn, ctx, best, best_cost, trace_best, n_evals);
}

/* RVCostFunction::cost (fitting.rs:736-781): the components are synthetic-grid arrays
 * (ArraySegments from vectors); each is shifted and resampled on its own, then
 * lr = light_ratio * mean(c2') / mean(c1'), model = (s1' c1' + s2' c2' lr) / (c1' + c2' lr) */
int or_binary_rv_cost(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
                      const float *var, const float *model1, const float *model2,
                      const float *cont1, const float *cont2, double light_ratio, double rv1,
                      double rv2, double *chi2) {
    int W = M->sg.n, n = d->n_obs, rc;
    float *s1 = malloc(sizeof(float) * n), *s2 = malloc(sizeof(float) * n);
    float *k1 = malloc(sizeof(float) * n), *k2 = malloc(sizeof(float) * n);
    seg32 a1 = {(float *)model1, W, 0, W}, a2 = {(float *)model2, W, 0, W};
    seg32 b1 = {(float *)cont1, W, 0, W}, b2 = {(float *)cont2, W, 0, W};
    rc = resample32(M, d, &a1, rv1, s1);
    if (!rc) rc = resample32(M, d, &a2, rv2, s2);
    if (!rc) rc = resample32(M, d, &b1, rv1, k1);
    if (!rc) rc = resample32(M, d, &b2, rv2, k2);
    if (!rc) {
        float lr = (float)light_ratio * mean_f32(k2, n) / mean_f32(k1, n);
        for (int i = 0; i < n; i++) {
            float synth = s1[i] * k1[i] + s2[i] * k2[i] * lr;
            float cont = k1[i] + k2[i] * lr;
            s1[i] = synth / cont;
        }
        rc = or_fit_continuum(c, flux, var, s1, NULL, chi2);
    }
    free(s1); free(s2); free(k1); free(k2);
    return rc;
}

/* BinaryTimeseriesKnownRVCostFunction::cost (fitting.rs:954-985): components once from
 * lab9 = (teff, m, logg, vsini) x 2 + light ratio, then per observed spectrum k (same
 * wavelengths, RVs rvs[2k], rvs[2k+1]) the binary model and its continuum-fit chi2,
 * summed in spectrum order */
int or_binary_ts_cost(const or_model *M, const or_model *CM, const or_disp *d, const or_cont *c,
                      const float *fluxes, const float *vars, const double *rvs, int S,
                      const double lab9[9], double *chi2) {
    int n = d->n_obs, rc = OR_OK;
    float *synth = malloc(sizeof(float) * n);
    double total = 0.0;
    for (int k = 0; k < S && !rc; k++) {
        double l1[5] = {lab9[0], lab9[1], lab9[2], lab9[3], rvs[2 * k]};
        double l2[5] = {lab9[4], lab9[5], lab9[6], lab9[7], rvs[2 * k + 1]};
        rc = or_binary_model_norm(M, CM, d, l1, l2, (float)lab9[8], synth);
        double ck = 0.0;
        if (!rc) rc = or_fit_continuum(c, fluxes + (size_t)k * n, vars + (size_t)k * n, synth,
                                       NULL, &ck);
        total += ck;
    }
    *chi2 = total;
    free(synth);
    return rc;
}
