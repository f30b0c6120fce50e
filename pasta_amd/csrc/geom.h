// geom.h -- grid geometry, Lagrange weights, bounds, PSO step and RNG.
//
// Compiled for both the host (C++ driver) and the device (gfx950 kernels), so
// the plain host-driven fit and the device-driven batched fit take identical
// decisions.  Built with -ffp-contract=off: every f64 expression below rounds
// exactly as written, like rustc's default (no multiply-add contraction).
//
// Reference (SimonVanSchuylenbergh/PASTA) sources restated here:
//   Range::get_right_index / find_neighbors      src/interpolate.rs:73-75,105-159
//   Grid staircase limits + get_local_grid       src/interpolate.rs:351-533
//   Lagrange factors + 64-weight product         src/cubic.rs:12-91,143-243
//   InMemFetcher::find_spectrum index            src/model_fetchers.rs:234-236
//   bounds / clamp_1d / constraints               src/bounds.rs:9-186, src/interpolate.rs:625-669
//   PSO velocity/position update                  src/particleswarm.rs:153-161,234-298
//   linear neighbours + trilinear weights         src/interpolate.rs:168-212,536-620,
//                                                 src/cubic.rs:6-10,246-262
//   binary bounds (11 and 9 parameters)           src/bounds.rs:188-345
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define PHD __host__ __device__ inline
#else
#define PHD inline
#endif

#define PASTA_MAX_AXIS 128

namespace pasta {

enum Status : int {
    ST_OK = 0,
    ST_BOUNDS = 1,     // label outside the grid (interpolate.rs:425-430)
    ST_NEIGHBORS = 2,  // not enough neighbours (cubic.rs:155)
    ST_VSINI = 3,      // vsini < 0 (convolve_rv.rs:489-491)
    ST_KERNEL = 4,     // rotation kernel longer than 2048 taps (convolve_rv.rs:512-519)
    ST_RANGE = 5,      // RV shift leaves the synthetic range (convolve_rv.rs:549-567)
    ST_INDEX = 6,      // resample reads past the model segment (the reference panics)
    ST_PSO = 7,        // particle left the bounds (particleswarm.rs:267-275)
    ST_SOLVE = 8,      // continuum normal equations not positive definite
};

// Grid axes + staircase.  Lives in device memory (one per grid) and on the host.
struct GridGeom {
    int n_teff, n_m, n_logg, n_models;
    double teff[PASTA_MAX_AXIS], m[PASTA_MAX_AXIS], logg[PASTA_MAX_AXIS];
    int logg_lo[PASTA_MAX_AXIS], logg_hi[PASTA_MAX_AXIS], cum[PASTA_MAX_AXIS];
};

// binary search: >= 0 exact index, else -(insertion point)-1
PHD int right_index(const double* v, int n, double x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (v[mid] < x) lo = mid + 1;
        else if (v[mid] > x) hi = mid;
        else return mid;
    }
    return -lo - 1;
}

PHD int model_index(const GridGeom& g, int i, int j, int k) {
    return (g.cum[i] + k - g.logg_lo[i]) * g.n_m + j;
}

PHD bool teff_logg_between_bounds(const GridGeom& g, double teff, double logg) {
    int r = right_index(g.teff, g.n_teff, teff);
    if (r >= 0) return logg >= g.logg[g.logg_lo[r]] && logg <= g.logg[g.logg_hi[r]];
    int i = -r - 1;
    if (i == 0 || i == g.n_teff) return false;
    return logg >= g.logg[g.logg_lo[i - 1]] && logg <= g.logg[g.logg_hi[i - 1]] &&
           logg >= g.logg[g.logg_lo[i]] && logg <= g.logg[g.logg_hi[i]];
}

PHD bool m_between_bounds(const GridGeom& g, double m) {
    return m >= g.m[0] && m <= g.m[g.n_m - 1];
}

PHD int logg_index_limits_at(const GridGeom& g, double teff, int* lo, int* hi) {
    int r = right_index(g.teff, g.n_teff, teff);
    int a, b;
    if (r >= 0) {
        if (r == g.n_teff - 1) { *lo = g.logg_lo[r]; *hi = g.logg_hi[r]; return ST_OK; }
        a = r; b = r + 1;
    } else {
        int i = -r - 1;
        if (i == 0 || i == g.n_teff) return ST_BOUNDS;
        a = i - 1; b = i;
    }
    *lo = g.logg_lo[a] > g.logg_lo[b] ? g.logg_lo[a] : g.logg_lo[b];
    *hi = g.logg_hi[a] < g.logg_hi[b] ? g.logg_hi[a] : g.logg_hi[b];
    return ST_OK;
}

PHD int teff_index_limits_at(const GridGeom& g, double logg, int* lo, int* hi) {
    int first = -1, last = -1;
    for (int i = 0; i < g.n_teff; i++) {
        if (logg >= g.logg[g.logg_lo[i]] && logg <= g.logg[g.logg_hi[i]]) {
            if (first < 0) first = i;
            last = i;
        }
    }
    if (first < 0) return ST_BOUNDS;
    *lo = first; *hi = last;
    return ST_OK;
}

// out[q] = -1 where the reference has None.  Where the limits span fewer than three
// points the rules can name an index outside the axis (i + 2 past the end, i - 2 below
// 0); the reference then panics on the value lookup (interpolate.rs:483-530), here that
// is ST_NEIGHBORS.
PHD int find_neighbors(const double* v, int n, double x, int left, int right, int out[4]) {
    int r = right_index(v, n, x);
    bool none0 = false, none3 = false;
    if (r >= 0) {
        int i = r;
        if (i < left || i > right) return ST_BOUNDS;
        if (i == left) { none0 = true; out[1] = i; out[2] = i + 1; out[3] = i + 2; }
        else if (i == right) { out[0] = i - 2; out[1] = i - 1; out[2] = i; none3 = true; }
        else if (i == right - 1) { out[0] = i - 2; out[1] = i - 1; out[2] = i; out[3] = i + 1; }
        else { out[0] = i - 1; out[1] = i; out[2] = i + 1; out[3] = i + 2; }
    } else {
        int i = -r - 1;
        if (i <= left || i > right) return ST_BOUNDS;
        if (i == left + 1) { none0 = true; out[1] = i - 1; out[2] = i; out[3] = i + 1; }
        else if (i == right) { out[0] = i - 2; out[1] = i - 1; out[2] = i; none3 = true; }
        else { out[0] = i - 2; out[1] = i - 1; out[2] = i; out[3] = i + 1; }
    }
    for (int q = none0 ? 1 : 0; q < (none3 ? 3 : 4); q++)
        if (out[q] < 0 || out[q] >= n) return ST_NEIGHBORS;
    if (none0) out[0] = -1;
    if (none3) out[3] = -1;
    return ST_OK;
}

PHD void lagrange_quadratic(double x, double x0, double x1, double x2, bool start, double f[4]) {
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

PHD void lagrange_cubic(double x, double x0, double x1, double x2, double x3, double f[4]) {
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

// which neighbours exist -> cubic / quadratic / zero / error
PHD int coeffs_from_neighbors(double x, const double nb[4], const bool has[4], double f[4]) {
    if (has[0] && has[1] && has[2] && has[3]) {
        double s = nb[3] - nb[0];
        lagrange_cubic((x - nb[0]) / s, 0.0, (nb[1] - nb[0]) / s, (nb[2] - nb[0]) / s, 1.0, f);
        return ST_OK;
    }
    if (has[0] && has[1] && has[2] && !has[3]) {
        double s = nb[2] - nb[0];
        lagrange_quadratic((x - nb[0]) / s, 0.0, (nb[1] - nb[0]) / s, 1.0, false, f);
        return ST_OK;
    }
    if (!has[0] && has[1] && has[2] && has[3]) {
        double s = nb[3] - nb[1];
        lagrange_quadratic((x - nb[1]) / s, 0.0, (nb[2] - nb[1]) / s, 1.0, true, f);
        return ST_OK;
    }
    if (!has[0] && !has[1] && !has[2] && !has[3]) {
        f[0] = f[1] = f[2] = f[3] = 0.0;
        return ST_OK;
    }
    return ST_NEIGHBORS;
}

// The 64 neighbour model indices (-1 = absent) and f32 weights, in the
// reference's order q = (c*4 + r)*4 + mi (c: teff neighbour, r: logg
// neighbour, mi: [M/H] neighbour).
PHD int local_weights(const GridGeom& g, double teff, double m, double logg, int idx[64],
                      float w[64]) {
    if (!teff_logg_between_bounds(g, teff, logg)) return ST_BOUNDS;
    if (!m_between_bounds(g, m)) return ST_BOUNDS;
    int tlo, thi, llo, lhi, rc;
    if ((rc = teff_index_limits_at(g, logg, &tlo, &thi))) return rc;
    if ((rc = logg_index_limits_at(g, teff, &llo, &lhi))) return rc;
    int tn[4], mn[4], ln[4];
    if ((rc = find_neighbors(g.teff, g.n_teff, teff, tlo, thi, tn))) return rc;
    if ((rc = find_neighbors(g.m, g.n_m, m, 0, g.n_m - 1, mn))) return rc;
    if ((rc = find_neighbors(g.logg, g.n_logg, logg, llo, lhi, ln))) return rc;
    bool has[4][4];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) {
            has[r][c] = false;
            if (tn[c] >= 0 && ln[r] >= 0) {
                int i = tn[c];
                has[r][c] = ln[r] >= g.logg_lo[i] && ln[r] <= g.logg_hi[i];
            }
        }
    double FT[4][4], FL[4][4], FM[4];
    for (int r = 0; r < 4; r++) {
        double nb[4]; bool h[4]; double f[4];
        for (int c = 0; c < 4; c++) { h[c] = has[r][c]; nb[c] = h[c] ? g.teff[tn[c]] : 0.0; }
        if ((rc = coeffs_from_neighbors(teff, nb, h, f))) return rc;
        for (int c = 0; c < 4; c++) FT[r][c] = f[c];
    }
    for (int c = 0; c < 4; c++) {
        double nb[4]; bool h[4]; double f[4];
        for (int r = 0; r < 4; r++) { h[r] = has[r][c]; nb[r] = h[r] ? g.logg[ln[r]] : 0.0; }
        if ((rc = coeffs_from_neighbors(logg, nb, h, f))) return rc;
        for (int r = 0; r < 4; r++) FL[r][c] = f[r];
    }
    {
        double nb[4]; bool h[4];
        for (int q = 0; q < 4; q++) { h[q] = mn[q] >= 0; nb[q] = h[q] ? g.m[mn[q]] : 0.0; }
        if ((rc = coeffs_from_neighbors(m, nb, h, FM))) return rc;
    }
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) {
            double tl = FT[r][c] * FL[r][c];
            for (int mi = 0; mi < 4; mi++) {
                int q = (c * 4 + r) * 4 + mi;
                w[q] = (float)(tl * FM[mi]);
                idx[q] = (has[r][c] && mn[mi] >= 0) ? model_index(g, tn[c], mn[mi], ln[r]) : -1;
            }
        }
    return ST_OK;
}

// Range::find_neighbors_linear (interpolate.rs:168-212): the two bracketing indices
PHD int find_neighbors_linear(const double* v, int n, double x, int left, int right, int out[2]) {
    int r = right_index(v, n, x);
    if (r >= 0) {
        if (r < left || r > right) return ST_BOUNDS;
        if (r == right) { out[0] = r - 1; out[1] = r; }
        else { out[0] = r; out[1] = r + 1; }
        return ST_OK;
    }
    int i = -r - 1;
    if (i <= left || i > right) return ST_BOUNDS;
    out[0] = i - 1;
    out[1] = i;
    return ST_OK;
}

// get_local_grid_linear + calculate_interpolation_coefficients_linear + the neighbour
// order of interpolate_linear (interpolate.rs:536-620,985-995, cubic.rs:6-10,246-262):
// q = (a*2 + b)*2 + c for teff neighbour a, logg neighbour b, [M/H] neighbour c; weight
// f32((ft[a] * fl[b]) * fm[c]).  Every neighbour is fetched with the storage index
// formula (model_fetchers.rs:234-236) whether or not it lies inside its Teff column's
// logg range, as the reference does; an index the reference could not fetch (logg below
// the column's first, or past the last model) is ST_INDEX.
PHD int local_weights_linear(const GridGeom& g, double teff, double m, double logg, int idx[8],
                             float w[8]) {
    if (!teff_logg_between_bounds(g, teff, logg)) return ST_BOUNDS;
    if (!m_between_bounds(g, m)) return ST_BOUNDS;
    int tlo, thi, llo, lhi, rc;
    if ((rc = teff_index_limits_at(g, logg, &tlo, &thi))) return rc;
    if ((rc = logg_index_limits_at(g, teff, &llo, &lhi))) return rc;
    int tn[2], mn[2], ln[2];
    if ((rc = find_neighbors_linear(g.teff, g.n_teff, teff, tlo, thi, tn))) return rc;
    if ((rc = find_neighbors_linear(g.m, g.n_m, m, 0, g.n_m - 1, mn))) return rc;
    if ((rc = find_neighbors_linear(g.logg, g.n_logg, logg, llo, lhi, ln))) return rc;
    double ft[2], fl[2], fm[2];
    ft[0] = (teff - g.teff[tn[1]]) / (g.teff[tn[0]] - g.teff[tn[1]]);
    ft[1] = (teff - g.teff[tn[0]]) / (g.teff[tn[1]] - g.teff[tn[0]]);
    fl[0] = (logg - g.logg[ln[1]]) / (g.logg[ln[0]] - g.logg[ln[1]]);
    fl[1] = (logg - g.logg[ln[0]]) / (g.logg[ln[1]] - g.logg[ln[0]]);
    fm[0] = (m - g.m[mn[1]]) / (g.m[mn[0]] - g.m[mn[1]]);
    fm[1] = (m - g.m[mn[0]]) / (g.m[mn[1]] - g.m[mn[0]]);
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
                const int q = (a * 2 + b) * 2 + c;
                const int i = tn[a], k = ln[b];
                if (k < g.logg_lo[i]) return ST_INDEX;
                const int id = model_index(g, i, mn[c], k);
                if (id >= g.n_models) return ST_INDEX;
                idx[q] = id;
                w[q] = (float)((ft[a] * fl[b]) * fm[c]);
            }
    return ST_OK;
}

// ---------------- bounds (bounds.rs:9-186) ----------------
struct Bounds {
    double vsini_lo, vsini_hi, rv_lo, rv_hi;
    int ckind[5];  // 0 none, 1 fixed, 2 range
    double ca[5], cb[5];
};

PHD void bounds_outer(const GridGeom& g, const Bounds& b, double lo[5], double hi[5]) {
    lo[0] = g.teff[0]; hi[0] = g.teff[g.n_teff - 1];
    lo[1] = g.m[0]; hi[1] = g.m[g.n_m - 1];
    lo[2] = g.logg[0]; hi[2] = g.logg[g.n_logg - 1];
    lo[3] = b.vsini_lo; hi[3] = b.vsini_hi;
    lo[4] = b.rv_lo; hi[4] = b.rv_hi;
    for (int d = 0; d < 5; d++) {
        if (b.ckind[d] == 1) { lo[d] = b.ca[d]; hi[d] = b.ca[d]; }
        else if (b.ckind[d] == 2) { lo[d] = fmax(lo[d], b.ca[d]); hi[d] = fmin(hi[d], b.cb[d]); }
    }
}

PHD bool bounds_within(const GridGeom& g, const Bounds& b, const double p[5]) {
    bool ok = m_between_bounds(g, p[1]) & teff_logg_between_bounds(g, p[0], p[2]);
    ok = ok && p[3] >= b.vsini_lo && p[3] <= b.vsini_hi && p[4] >= b.rv_lo && p[4] <= b.rv_hi;
    if (!ok) return false;
    for (int d = 0; d < 5; d++) {
        if (b.ckind[d] == 1 && !(fabs(p[d] - b.ca[d]) <= 2.220446049250313e-16)) return false;
        if (b.ckind[d] == 2 && !(p[d] >= b.ca[d] && p[d] <= b.cb[d])) return false;
    }
    return true;
}

PHD int bounds_clamp(const GridGeom& g, const Bounds& b, const double p[5], int d, double* out) {
    double lo, hi;
    int a, c, rc;
    if (d == 3) { lo = b.vsini_lo; hi = b.vsini_hi; }
    else if (d == 4) { lo = b.rv_lo; hi = b.rv_hi; }
    else if (d == 0) {
        if ((rc = teff_index_limits_at(g, p[2], &a, &c))) return ST_PSO;
        lo = g.teff[a]; hi = g.teff[c];
    } else if (d == 1) {
        lo = g.m[0]; hi = g.m[g.n_m - 1];
    } else {
        if ((rc = logg_index_limits_at(g, p[0], &a, &c))) return ST_PSO;
        lo = g.logg[a]; hi = g.logg[c];
    }
    if (b.ckind[d] == 1) { lo = b.ca[d]; hi = b.ca[d]; }
    else if (b.ckind[d] == 2) { lo = fmax(lo, b.ca[d]); hi = fmin(hi, b.cb[d]); }
    if (lo > hi) return ST_PSO;
    double x = p[d];
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    *out = x;
    return ST_OK;
}

// ---------------- RNG: counter-based SplitMix64 ----------------
PHD uint64_t rng_u64(uint64_t key, uint64_t ctr) {
    uint64_t z = key + (ctr + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
PHD double rng_uniform(uint64_t key, uint64_t ctr) {
    return (double)(rng_u64(key, ctr) >> 11) * 0x1.0p-53;
}
static const uint64_t KEY_POS = 0x51A5A0000000001ULL;
static const uint64_t KEY_VEL = 0x51A5A0000000002ULL;
static const uint64_t KEY_ITER = 0x51A5A0000000003ULL;

PHD double rand_range(double a, double b, double u) {
    if (a == b) return a;
    if (a < b) return a + (b - a) * u;
    return b + (a - b) * u;
}

struct PsoParams {
    double inertia, cognitive, social;
};

// One particle's update for iteration `it` (particleswarm.rs:234-276).
// pos/vel updated in place; bpos = personal best, g = global best.
PHD int pso_update_particle(const GridGeom& g, const Bounds& b, const PsoParams& s,
                            uint64_t key_iter, int it, int P, int q, double pos[5], double vel[5],
                            const double bpos[5], const double gbest[5]) {
    for (int d = 0; d < 5; d++) {
        uint64_t c0 = (((uint64_t)it * (uint64_t)P + (uint64_t)q) * 5 + (uint64_t)d) * 2;
        double momentum = vel[d] * s.inertia;
        double to_opt = bpos[d] - pos[d];
        double pull = rand_range(0.0, to_opt, rng_uniform(key_iter, c0)) * s.cognitive;
        double to_g = gbest[d] - pos[d];
        double pullg = rand_range(0.0, to_g, rng_uniform(key_iter, c0 + 1)) * s.social;
        vel[d] = momentum + pull + pullg;
        double np[5];
        for (int e = 0; e < 5; e++) np[e] = pos[e];
        np[d] += vel[d];
        double cl;
        int rc = bounds_clamp(g, b, np, d, &cl);
        if (rc) return rc;
        pos[d] = cl;
        if (!bounds_within(g, b, pos)) return ST_PSO;
    }
    return ST_OK;
}

// ---------------- bounds of the binary searches (bounds.rs:188-345) ----------------
// A D-parameter search over one or two stars: parameter d is a grid axis (teff, [M/H],
// logg) of the star whose Teff sits at base[d], or a plain range [lo, hi] (vsini, RV,
// light ratio); constraints as in ConstrainedPSOBounds (bounds.rs:64-132).
#define PASTA_MAX_D 11
enum BoundKind : int { BK_TEFF = 0, BK_M = 1, BK_LOGG = 2, BK_RANGE = 3 };
struct BoundsN {
    int D;
    int kind[PASTA_MAX_D];
    int base[PASTA_MAX_D];
    double lo[PASTA_MAX_D], hi[PASTA_MAX_D];
    int ckind[PASTA_MAX_D];  // 0 none, 1 fixed, 2 range
    double ca[PASTA_MAX_D], cb[PASTA_MAX_D];
};

// BinaryBounds (11: star 1, star 2, light ratio) or BinaryBoundsWithoutRV (9)
PHD BoundsN binary_bounds(bool with_rv, double vs_lo, double vs_hi, double rv_lo, double rv_hi,
                          double lr_lo, double lr_hi) {
    BoundsN b{};
    const int per = with_rv ? 5 : 4;
    b.D = 2 * per + 1;
    for (int s = 0; s < 2; s++) {
        const int o = s * per;
        b.kind[o] = BK_TEFF; b.kind[o + 1] = BK_M; b.kind[o + 2] = BK_LOGG;
        b.base[o] = b.base[o + 1] = b.base[o + 2] = o;
        b.kind[o + 3] = BK_RANGE; b.lo[o + 3] = vs_lo; b.hi[o + 3] = vs_hi;
        if (with_rv) { b.kind[o + 4] = BK_RANGE; b.lo[o + 4] = rv_lo; b.hi[o + 4] = rv_hi; }
    }
    b.kind[b.D - 1] = BK_RANGE;
    b.lo[b.D - 1] = lr_lo;
    b.hi[b.D - 1] = lr_hi;
    return b;
}

PHD void boundsn_outer(const GridGeom& g, const BoundsN& b, double* lo, double* hi) {
    for (int d = 0; d < b.D; d++) {
        switch (b.kind[d]) {
        case BK_TEFF: lo[d] = g.teff[0]; hi[d] = g.teff[g.n_teff - 1]; break;
        case BK_M: lo[d] = g.m[0]; hi[d] = g.m[g.n_m - 1]; break;
        case BK_LOGG: lo[d] = g.logg[0]; hi[d] = g.logg[g.n_logg - 1]; break;
        default: lo[d] = b.lo[d]; hi[d] = b.hi[d];
        }
        if (b.ckind[d] == 1) { lo[d] = b.ca[d]; hi[d] = b.ca[d]; }
        else if (b.ckind[d] == 2) { lo[d] = fmax(lo[d], b.ca[d]); hi[d] = fmin(hi[d], b.cb[d]); }
    }
}

PHD bool boundsn_within(const GridGeom& g, const BoundsN& b, const double* p) {
    for (int d = 0; d < b.D; d++)
        if (b.kind[d] == BK_TEFF) {
            const int o = b.base[d];
            if (!(teff_logg_between_bounds(g, p[o], p[o + 2]) && m_between_bounds(g, p[o + 1])))
                return false;
        }
    for (int d = 0; d < b.D; d++)
        if (b.kind[d] == BK_RANGE && !(p[d] >= b.lo[d] && p[d] <= b.hi[d])) return false;
    for (int d = 0; d < b.D; d++) {
        if (b.ckind[d] == 1 && !(fabs(p[d] - b.ca[d]) <= 2.220446049250313e-16)) return false;
        if (b.ckind[d] == 2 && !(p[d] >= b.ca[d] && p[d] <= b.cb[d])) return false;
    }
    return true;
}

PHD int boundsn_clamp(const GridGeom& g, const BoundsN& b, const double* p, int d, double* out) {
    double lo, hi;
    int a, c;
    const int o = b.base[d];
    switch (b.kind[d]) {
    case BK_TEFF:
        if (teff_index_limits_at(g, p[o + 2], &a, &c)) return ST_PSO;
        lo = g.teff[a]; hi = g.teff[c];
        break;
    case BK_M: lo = g.m[0]; hi = g.m[g.n_m - 1]; break;
    case BK_LOGG:
        if (logg_index_limits_at(g, p[o], &a, &c)) return ST_PSO;
        lo = g.logg[a]; hi = g.logg[c];
        break;
    default: lo = b.lo[d]; hi = b.hi[d];
    }
    if (b.ckind[d] == 1) { lo = b.ca[d]; hi = b.ca[d]; }
    else if (b.ckind[d] == 2) { lo = fmax(lo, b.ca[d]); hi = fmin(hi, b.cb[d]); }
    if (lo > hi) return ST_PSO;
    double x = p[d];
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    *out = x;
    return ST_OK;
}

// pso_update_particle for D parameters (the same draws for D = 5)
PHD int pso_update_particle_n(const GridGeom& g, const BoundsN& b, const PsoParams& s,
                              uint64_t key_iter, int it, int P, int q, double* pos, double* vel,
                              const double* bpos, const double* gbest) {
    const int D = b.D;
    for (int d = 0; d < D; d++) {
        uint64_t c0 = (((uint64_t)it * (uint64_t)P + (uint64_t)q) * (uint64_t)D + (uint64_t)d) * 2;
        double momentum = vel[d] * s.inertia;
        double to_opt = bpos[d] - pos[d];
        double pull = rand_range(0.0, to_opt, rng_uniform(key_iter, c0)) * s.cognitive;
        double to_g = gbest[d] - pos[d];
        double pullg = rand_range(0.0, to_g, rng_uniform(key_iter, c0 + 1)) * s.social;
        vel[d] = momentum + pull + pullg;
        double np[PASTA_MAX_D];
        for (int e = 0; e < D; e++) np[e] = pos[e];
        np[d] += vel[d];
        double cl;
        int rc = boundsn_clamp(g, b, np, d, &cl);
        if (rc) return rc;
        pos[d] = cl;
        if (!boundsn_within(g, b, pos)) return ST_PSO;
    }
    return ST_OK;
}

// f64::powi lowers to __powidf2: exponentiation by squaring.
PHD double powi_d(double a, int b) {
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return r;
}

// map_range(x, lo, hi, -1, 1) of continuum_fitting.rs:109-116,172-174
PHD double chunk_xhat(double x, double lo, double hi) {
    return (x - lo) / (hi - lo) * (1.0 - -1.0) + -1.0;
}

}  // namespace pasta
