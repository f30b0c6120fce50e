/*
 * CPU oracle for the PASTA model-evaluation path.
 *
 * TEST INFRASTRUCTURE, NOT THE PRODUCT.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library.  It is a plain-C
 * restatement of the reference Rust code (SimonVanSchuylenbergh/PASTA); every
 * function cites the reference file:line it follows.  Two arithmetic modes:
 *   - "f32-faithful": the reference's f32 casts, operation order and quirks;
 *   - "fp64": the same algorithm with every intermediate in double precision
 *     (exact Lagrange weights, direct convolution, unquantised resample weight).
 *
 * Parity status per function is documented in pasta_oracle.c's header.
 */
#ifndef PASTA_ORACLE_H
#define PASTA_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes (negative) */
#define OR_OK 0
#define OR_ERR_BOUNDS -1      /* label outside grid (interpolate.rs:425-430) */
#define OR_ERR_NEIGHBORS -2   /* not enough neighbours (cubic.rs:155) */
#define OR_ERR_VSINI -3       /* vsini < 0 (convolve_rv.rs:489-491) */
#define OR_ERR_KERNEL -4      /* rotation kernel > FFTSIZE (convolve_rv.rs:512-519) */
#define OR_ERR_RANGE -5       /* RV shift leaves the synthetic range (convolve_rv.rs:549-567) */
#define OR_ERR_INDEX -6       /* resample reads past the end of the segment (reference panics) */
#define OR_ERR_ARG -7         /* bad argument / length mismatch */
#define OR_ERR_PSO -8         /* particle left bounds (particleswarm.rs:267-275) */
#define OR_ERR_SIGN -9        /* BrentRoot: f(min) and f(max) have the same sign */

/* ---------------- grid geometry (interpolate.rs:58-670) ---------------- */
typedef struct or_grid {
    int n_teff, n_m, n_logg;
    double *teff, *m, *logg;
    int *logg_lo, *logg_hi; /* per teff index: first/last logg index present */
    int *cum;               /* per teff index: (teff,logg) pairs at lower teff */
    int n_models;
} or_grid;

or_grid *or_grid_from_labels(const double *labels /* n x 3 (teff,m,logg) */, int n);
void or_grid_free(or_grid *g);
int or_grid_model_index(const or_grid *g, int i_teff, int j_m, int k_logg);
int or_grid_list_gridpoints(const or_grid *g, double *out /* n_models x 3 */);
int or_is_teff_logg_between_bounds(const or_grid *g, double teff, double logg);
int or_get_logg_index_limits_at(const or_grid *g, double teff, int *lo, int *hi);
int or_get_teff_index_limits_at(const or_grid *g, double logg, int *lo, int *hi);
int or_find_neighbors(const double *values, int n, double x, int left, int right,
                      int out[4] /* -1 = None */);
/* local grid + 64 weights.  idx[64]: model index or -1; w32[64] f32 weights
 * (cubic.rs:232-243); w64[64] the same products before the f32 cast. */
int or_local_weights(const or_grid *g, double teff, double m, double logg, int idx[64],
                     float w32[64], double w64[64], double ft[16], double fl[16],
                     double fm[4]);
int or_clamp_1d(const or_grid *g, const double p[3], int dim, double *out);

/* ---------------- interpolation (interpolate.rs:930-983) ---------------- */
/* spectra: n_models x W uint16, one contiguous spectrum per model.
 * factors: per-model f32 (already divided by 65535) or NULL (1/65535). */
int or_interpolate_f32(const or_grid *g, const uint16_t *spectra, const float *factors,
                       int W, double teff, double m, double logg, float *out);
int or_interpolate_f64(const or_grid *g, const uint16_t *spectra, const float *factors,
                       int W, double teff, double m, double logg, double *out);

/* ---------------- convolution (convolve_rv.rs) ---------------- */
int or_rotation_kernel(double vsini, double dvelo, float *out, int max_len);
int or_rotation_kernel_f64(double vsini, double dvelo, double *out, int max_len);
/* overlap-add with 2048-point FFT; y must hold next_multiple(l, step)+m-1 */
int or_oa_convolve(const float *x, int l, const float *h, int m, float *y, int y_cap);
int or_oa_out_len(int l, int m);
void or_make_gaussian(float *arr, int n, float sigma);
int or_fixed_gauss_kernel(double resolution, double log_step, float *out, int max_len);
/* direct full convolution in f64, y length l+m-1 */
void or_direct_convolve_f64(const double *x, int l, const double *h, int m, double *y);

/* ---------------- star description (dispersion + continuum) ---------------- */
enum { OR_DISP_NONE = 0, OR_DISP_FIXED = 1, OR_DISP_VARIABLE = 2 };
enum { OR_CONT_CHUNK = 0, OR_CONT_LINEAR = 1, OR_CONT_FIXED = 2, OR_CONT_CONSTANT = 3 };

typedef struct or_wlgrid {
    double first, step; /* Logspace: log10(first), log10 step */
    int n;
} or_wlgrid;

typedef struct or_disp {
    int kind;
    int n_obs;
    double *wl;      /* observed wavelengths */
    double *fidx;    /* float index of each observed wavelength in the synthetic grid */
    float *kernel;   /* fixed: K taps */
    double sigma;    /* fixed: Gaussian sigma in synthetic pixels (f64) */
    int K;
    float *kernels;  /* variable: K x W, column per synthetic pixel */
    int W;
} or_disp;

or_disp *or_disp_new(int kind, const double *wl, int n, const or_wlgrid *sg, double resolution,
                     const float *disp /* variable: per observed pixel */);
void or_disp_free(or_disp *d);
int or_disp_kernel_len(const or_disp *d);
/* fixed: the K taps; variable: W rows of K taps (row i = synthetic pixel i); returns the
 * number of values written */
long or_disp_kernels(const or_disp *d, float *out, long cap);

typedef struct or_cont {
    int kind;
    int n;
    /* chunk */
    int n_chunks, p_order;
    int *startstop;      /* n_chunks x 2 */
    float **A;           /* per chunk: (stop-start) x (p+1), row-major */
    float **fac;         /* per chunk boundary c: factor array over [start_{c+1}, stop_c) */
    double *wl;
    /* linear */
    int ncols;
    float *design;       /* n x ncols row-major */
    /* fixed */
    float *fixed;
    int ignore;
} or_cont;

or_cont *or_cont_new_chunk(const double *wl, int n, int n_chunks, int p_order, double overlap);
or_cont *or_cont_new_linear(const float *design, int n, int ncols);
or_cont *or_cont_new_fixed(const float *cont, int n, int ignore);
or_cont *or_cont_new_constant(int n);
void or_cont_free(or_cont *c);
int or_cont_n_params(const or_cont *c);
int or_cont_startstop(const or_cont *c, int *out);
int or_cont_design_row(const or_cont *c, int chunk, int row, float *out);
int or_cont_fac(const or_cont *c, int boundary, float *out, int cap);

/* fit_continuum: params (f32) and chi2 (continuum_fitting.rs:382-396 and friends) */
int or_fit_continuum(const or_cont *c, const float *flux, const float *var, const float *synth,
                     float *params, double *chi2);
int or_fit_continuum_f64(const or_cont *c, const float *flux, const float *var,
                         const double *synth, double *params, double *chi2);
int or_build_continuum(const or_cont *c, const float *params, float *out);

/* ---------------- model production (interpolate.rs:1068-1090) ---------------- */
typedef struct or_model {
    const or_grid *g;
    const uint16_t *spectra;
    const float *factors;
    or_wlgrid sg;
} or_model;

/* produce_model in f32-faithful mode; out has n_obs floats */
int or_produce_model(const or_model *M, const or_disp *d, const double lab[5], float *out);
/* FP64 mode: interpolation with f64 weights, direct f64 convolutions, exact resample weight */
int or_produce_model_f64(const or_model *M, const or_disp *d, const double lab[5], double *out);
/* FP64 mode with the reference's f32-quantised resample weight (convolve_rv.rs:570-572) */
int or_produce_model_f64q(const or_model *M, const or_disp *d, const double lab[5], double *out);
/* rotation+dispersion only, the synthetic-grid segment (interpolate_and_convolve) */
int or_interpolate_and_convolve(const or_model *M, const or_disp *d, const double lab[4],
                                float *out /* W */);
int or_cost(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
            const float *var, const double lab[5], double *chi2, float *params);
int or_cost_batch(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
                  const float *var, const double *labels, int n, int n_threads, double *chi2);
/* or_cost_batch in the FP64 mode over the reference's f32-stored values (or_cost_f64q) */
int or_cost_batch_f64q(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
                       const float *var, const double *labels, int n, int n_threads,
                       double *chi2);
int or_cost_f64(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
                const float *var, const double lab[5], double *chi2);
/* or_cost_f64 with the quantised resample weight; params (nullable) the f64 continuum
 * parameters */
int or_cost_f64q(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
                 const float *var, const double lab[5], double *chi2, double *params);

/* ---------------- PSO (particleswarm.rs, bounds.rs, fitting.rs) ---------------- */
typedef struct or_pso_settings {
    int num_particles;
    int max_iters;
    double inertia, cognitive, social, delta;
} or_pso_settings;

typedef struct or_bounds {
    const or_grid *g;
    double vsini_lo, vsini_hi, rv_lo, rv_hi;
    int ckind[5]; /* 0 none, 1 fixed, 2 range */
    double ca[5], cb[5];
} or_bounds;

uint64_t or_rng_u64(uint64_t key, uint64_t ctr);
double or_rng_uniform(uint64_t key, uint64_t ctr);
int or_bounds_outer(const or_bounds *b, double lo[5], double hi[5]);
int or_bounds_within(const or_bounds *b, const double p[5]);
int or_bounds_clamp(const or_bounds *b, const double p[5], int d, double *out);
int or_pso_init_positions(const or_bounds *b, int P, uint64_t seed, double *pos, double *vel);

typedef int (*or_cost_fn)(void *ctx, int n, const double *pos /* n x 5 */, double *cost);
/* Runs the PSO; positions: optional injected initial positions (P x 5) else drawn.
 * trace_best: (max_iters+1) x 6 (cost + position) of the global best after each iteration. */
int or_pso_run(const or_bounds *b, const or_pso_settings *s, uint64_t seed, const double *init_pos,
               const double *init_vel, or_cost_fn fn, void *ctx, double best[5], double *best_cost,
               double *trace_best, long *n_evals);
/* PSO with the model cost (stack C of the reference, one thread) */
int or_fit(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
           const float *var, const or_bounds *b, const or_pso_settings *s, uint64_t seed,
           const double *init_pos, int n_threads, double best[5], double *best_cost,
           float *params, long *n_evals);

/* ---------------- binary stars (interpolate.rs:796-890,985-1030, convolve_rv.rs:606-703,
 * fitting.rs:571-728, bounds.rs:188-276) ---------------- */
/* get_local_grid_linear + calculate_interpolation_coefficients_linear: the 8 neighbours
 * (teff x logg x [M/H] order) and weights (w f32 and/or w64 f64, either may be NULL) */
int or_local_weights_linear(const or_grid *g, double teff, double m, double logg, int idx[8],
                            float w[8], double w64[8]);
int or_interpolate_linear_f32(const or_grid *g, const uint16_t *spectra, const float *factors,
                              int W, double teff, double m, double logg, float *out);
int or_interpolate_linear_f64(const or_grid *g, const uint16_t *spectra, const float *factors,
                              int W, double teff, double m, double logg, double *out);
/* produce_binary_model_norm: CM is the continuum (unnormalised) grid; out has n_obs floats */
int or_binary_model_norm(const or_model *M, const or_model *CM, const or_disp *d,
                         const double lab1[5], const double lab2[5], float light_ratio,
                         float *out);
/* BinaryCostFunction::cost: lab = star 1 (5), star 2 (5), light ratio */
int or_binary_cost(const or_model *M, const or_model *CM, const or_disp *d, const or_cont *c,
                   const float *flux, const float *var, const double lab[11], double *chi2,
                   float *params);
typedef struct or_binbounds {
    const or_grid *g;
    double vsini_lo, vsini_hi, rv_lo, rv_hi, lr_lo, lr_hi;
    int ckind[11]; /* 0 none, 1 fixed, 2 range */
    double ca[11], cb[11];
} or_binbounds;
int or_binbounds_outer(const or_binbounds *b, double lo[11], double hi[11]);
int or_binbounds_within(const or_binbounds *b, const double p[11]);
int or_binbounds_clamp(const or_binbounds *b, const double p[11], int d, double *out);
/* the PSO over D = 11 (BinaryBounds), 9 (BinaryBoundsWithoutRV: star labels without RV
 * at 0-3 and 4-7, light ratio at 8) or 2 (BinaryRVBounds) parameters; fn receives
 * positions n x D; trace_best (max_iters+1) x (D+1) */
int or_pso_run_d(const or_binbounds *b, int D, const or_pso_settings *s, uint64_t seed,
                 const double *init_pos, or_cost_fn fn, void *ctx, double *best,
                 double *best_cost, double *trace_best, long *n_evals);
/* RVCostFunction::cost: components are synthetic-grid arrays (W each) */
int or_binary_rv_cost(const or_model *M, const or_disp *d, const or_cont *c, const float *flux,
                      const float *var, const float *model1, const float *model2,
                      const float *cont1, const float *cont2, double light_ratio, double rv1,
                      double rv2, double *chi2);
/* BinaryTimeseriesKnownRVCostFunction::cost: S spectra (S x n fluxes/vars) on the
 * dispersion's wavelengths, rvs S x 2, lab9 = (teff, m, logg, vsini) x 2 + light ratio */
int or_binary_ts_cost(const or_model *M, const or_model *CM, const or_disp *d, const or_cont *c,
                      const float *fluxes, const float *vars, const double *rvs, int S,
                      const double lab9[9], double *chi2);
/* the PSO over 11 parameters; fn receives positions n x 11; trace_best (max_iters+1) x 12 */
int or_pso_run_binary(const or_binbounds *b, const or_pso_settings *s, uint64_t seed,
                      const double *init_pos, or_cost_fn fn, void *ctx, double best[11],
                      double *best_cost, double *trace_best, long *n_evals);

/* ---------------- uncertainty (fitting.rs:333-372,482-568) ---------------- */
/* BrentRoot of argmin 0.10.0 (a dependency absent from the reference tree), restated from
 * the Brent / zbrent root finder it implements: brackets [min, max], tolerance tol, returns
 * the iterate with the smallest |f| (argmin's best parameter).  Parity of the iterate
 * sequence with argmin is unpinned; roots are pinned by closed forms. */
typedef int (*or_scalar_fn)(void *ctx, double x, double *f);
int or_brent_root(or_scalar_fn fn, void *ctx, double min, double max, double tol, double *root,
                  int *n_evals);
/* compute_uncertainty: out[2*i] = lower, out[2*i+1] = upper distance of label i to the
 * chi2 = best * (1 + sqrt(2/n)) crossing; status[2*i+k] = 0 or an OR_ERR_* code (the
 * reference's Err, None in Python).  search_radius NULL = the Python defaults
 * (2000, 0.3, 0.3, 40, 40) (lib.rs:1554).  Returns an error if the best model fails. */
int or_compute_uncertainty(const or_model *M, const or_disp *d, const or_cont *c,
                           const float *flux, const float *var, double spec_res,
                           const double label[5], const double *search_radius, double out[10],
                           int status[10]);

#ifdef __cplusplus
}
#endif
#endif
