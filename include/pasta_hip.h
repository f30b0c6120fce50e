/*
 * pasta_hip.h -- C ABI of libpasta_hip.so, the MI355X (gfx950) model-evaluation
 * path of PASTA.
 *
 * Plain pointers and sizes only.  Every function returns an int status:
 * PASTA_OK (0) on success, a negative PASTA_ERR_* otherwise, with a message in
 * pasta_last_error() (thread-local).  Handles are immutable after creation and
 * may be shared between threads; evaluation calls on one handle serialise on
 * an internal lock.
 *
 * Each entry point names the reference (SimonVanSchuylenbergh/PASTA) interface
 * it replaces.  Labels are always (Teff, [M/H], logg, vsini, RV), f64.
 */
#ifndef PASTA_HIP_H
#define PASTA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PASTA_OK 0
#define PASTA_ERR_ARG -1      /* bad argument / length mismatch */
#define PASTA_ERR_HIP -2      /* HIP runtime error */
#define PASTA_ERR_NOMEM -3    /* device allocation failed / admission refused */
#define PASTA_ERR_MODEL -4    /* model evaluation failed (per-item status holds the reason) */
#define PASTA_ERR_FIT -5      /* a fit failed (particle out of bounds, evaluation error) */
#define PASTA_ERR_STATE -6    /* request unknown / not finished / cancelled */

/* Per-evaluation status codes (status_out arrays). */
#define PASTA_ST_OK 0
#define PASTA_ST_BOUNDS 1     /* src/interpolate.rs:425-430 (labels outside the grid) */
#define PASTA_ST_NEIGHBORS 2  /* src/cubic.rs:155 (not enough neighbours) */
#define PASTA_ST_VSINI 3      /* src/convolve_rv.rs:489-491 (vsini < 0) */
#define PASTA_ST_KERNEL 4     /* src/convolve_rv.rs:512-519 (rotation kernel > 2048 taps) */
#define PASTA_ST_RANGE 5      /* src/convolve_rv.rs:549-567 (RV shift outside synthetic range) */
#define PASTA_ST_INDEX 6      /* resample reads past the model segment (the reference panics) */
#define PASTA_ST_PSO 7        /* src/particleswarm.rs:267-275 (particle left the bounds) */
#define PASTA_ST_SOLVE 8      /* continuum normal equations not positive definite */
#define PASTA_ST_SIGN 9       /* compute_uncertainty: chi2 - target has one sign at both ends of
                                 the search bracket (argmin BrentRoot's error, fitting.rs:538-545) */
#define PASTA_ST_CLAMP 10     /* compute_uncertainty: the search bound cannot be clamped into the
                                 grid (bounds.rs:13-16 via fitting.rs:532) */

const char *pasta_last_error(void);
int pasta_version(void);
/* Free device memory in bytes (for admission decisions). */
int pasta_device_mem_info(uint64_t *free_bytes, uint64_t *total_bytes);

/* ------------------------------------------------------------------ grid --
 * Replaces InMemFetcher::new + GridInterpolator::new (src/model_fetchers.rs:136-166,
 * src/interpolate.rs:267-331,905-910; Python InMemInterpolator, src/lib.rs:1888-1894).
 * labels: n_models x 3 (Teff, [M/H], logg) of each row of `spectra` (any order);
 * the grid geometry is built from them like Grid::new and the rows are stored
 * in HBM in list_gridpoints order (src/interpolate.rs:334-348).
 * spectra: n_models x n_pix uint16, or NULL to allocate only (see
 * pasta_grid_device_rows).  factors: per-row f32 max/65535 (includes_factor=True,
 * src/model_fetchers.rs:53-63) or NULL (1/65535).
 * The synthetic wavelength grid is WlGrid::Logspace(log10_first, log10_step, n_pix).
 */
typedef struct pasta_grid pasta_grid;
int pasta_grid_create(const uint16_t *spectra, const float *factors, const double *labels,
                      int n_models, int n_pix, double log10_first, double log10_step,
                      pasta_grid **out);
int pasta_grid_destroy(pasta_grid *g);
/* n_teff, n_m, n_logg, n_models, n_pix */
int pasta_grid_shape(const pasta_grid *g, int out[5]);
/* axes (Python teffs/ms/loggs/logg_limits/cumulative_grid_size, src/lib.rs:1436-1457) */
int pasta_grid_axes(const pasta_grid *g, double *teff, double *m, double *logg, int *logg_lo,
                    int *logg_hi, int *cum);
/* Device rows in storage order (list_gridpoints): pointer to row 0 and the row
 * pitch in elements, so a device-side producer can fill the grid in place. */
int pasta_grid_device_rows(const pasta_grid *g, uint16_t **rows, int64_t *pitch);
/* Copy rows [row0, row0+nrows) (storage order, n_pix each) back to the host. */
int pasta_grid_read_rows(const pasta_grid *g, int row0, int nrows, uint16_t *out);
/* Upload per-row factors after creation (storage order), or NULL to clear. */
int pasta_grid_set_factors(pasta_grid *g, const float *factors_storage_order);
/* Rotation-stage crossover (no counterpart in the reference, which always convolves by
 * FFT, src/convolve_rv.rs:484-529): kernels of at least min_taps taps are convolved by a
 * 4096-point FFT where their tiles fit one transform, shorter ones directly on the matrix
 * cores.  min_taps <= 0 restores the default; a value above 2048 turns the FFT path off.
 * Applies to evaluations launched (and graphs captured) after the call. */
int pasta_grid_set_fft_taps(pasta_grid *g, int min_taps);
/* get_local_grid + calculate_interpolation_coefficients (src/interpolate.rs:424-533,
 * src/cubic.rs:232-243): storage row of each of the 64 neighbours (-1 absent) and
 * the f32 weights, in the reference's neighbour order. */
int pasta_grid_local_weights(const pasta_grid *g, double teff, double m, double logg,
                             int idx[64], float w[64]);
/* Grid::clamp_1d (src/interpolate.rs:625-628) */
int pasta_grid_clamp_1d(const pasta_grid *g, double teff, double m, double logg, int index,
                        double *out);
int pasta_grid_is_teff_logg_between_bounds(const pasta_grid *g, double teff, double logg);

/* ------------------------------------------------------------ dispersion --
 * Replaces NoConvolutionDispersion / FixedResolutionDispersion (log-spaced) /
 * VariableResolutionDispersion (src/lib.rs:762-797; src/convolve_rv.rs:77-86,
 * 168-196, 346-402).  wl: observed wavelengths (n).  The synthetic grid is the
 * grid's WlGrid::Logspace. */
typedef struct pasta_disp pasta_disp;
int pasta_disp_create_none(const pasta_grid *g, const double *wl, int n, pasta_disp **out);
int pasta_disp_create_fixed(const pasta_grid *g, const double *wl, int n, double resolution,
                            pasta_disp **out);
int pasta_disp_create_variable(const pasta_grid *g, const double *wl, int n, const float *disp,
                               pasta_disp **out);
int pasta_disp_destroy(pasta_disp *d);
/* kernel length (0 for none) and, if out != NULL, the taps (fixed) or the
 * K x n_pix kernel matrix, row k = tap k (variable, get_kernels src/lib.rs:1043-1060) */
int pasta_disp_kernel(const pasta_disp *d, float *out, int64_t cap, int *k_len);

/* ------------------------------------------------------------- continuum --
 * Replaces ChunkContinuumFitter / LinearModelContinuumFitter / FixedContinuum /
 * ConstantContinuum (src/lib.rs:944-983; src/continuum_fitting.rs:46-107,127-423,
 * 427-505).  design: n x ncols row-major. */
typedef struct pasta_cont pasta_cont;
int pasta_cont_create_chunk(const double *wl, int n, int n_chunks, int p_order, double overlap,
                            pasta_cont **out);
int pasta_cont_create_linear(const float *design, int n, int ncols, pasta_cont **out);
int pasta_cont_create_fixed(const float *continuum, int n, int ignore_first_and_last,
                            pasta_cont **out);
int pasta_cont_create_constant(int n, pasta_cont **out);
int pasta_cont_destroy(pasta_cont *c);
int pasta_cont_n_params(const pasta_cont *c);
/* chunk start/stop indices (n_chunks x 2) */
int pasta_cont_chunks(const pasta_cont *c, int *startstop, int cap);
/* build_continuum (src/continuum_fitting.rs:228-278,398-413) on the host */
int pasta_cont_build(const pasta_cont *c, const float *params, int n_params, float *out,
                     int64_t cap);

/* ------------------------------------------------------------------ star --
 * An observed spectrum bound to a grid, dispersion and continuum model, resident
 * in HBM (ObservedSpectrum + CostFunction fields, src/fitting.rs:22-38,299-305).
 * flux/var: n (= dispersion and continuum length). */
typedef struct pasta_star pasta_star;
int pasta_star_create(pasta_grid *g, const pasta_disp *d, const pasta_cont *c, const float *flux,
                      const float *var, int n, pasta_star **out);
int pasta_star_destroy(pasta_star *s);
/* bytes of HBM one evaluation of this star needs (model buffer + partials) */
int pasta_star_eval_bytes(const pasta_star *s, int64_t *bytes);

/* ----------------------------------------------------------- evaluation --
 * Batched CostFunction::cost (src/fitting.rs:313-326) = SingleFitter::chi2 /
 * Python chi2 (src/lib.rs:1603-1627): labels P x 5 -> chi2 (NaN + status on
 * failure, the allow_nan behaviour).  params (nullable): P x n_params continuum
 * parameters.  Returns PASTA_ERR_MODEL if any status is non-zero. */
int pasta_eval_chi2(pasta_star *s, const double *labels, int P, double *chi2, int *status,
                    float *params);
/* No counterpart in the reference.  pasta_eval_chi2 for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): same arguments, same kernels and bit-identical results, but every device
 * buffer beyond the star (itself created beforehand) and the caller's host arrays comes
 * from one 64 MiB scratch array per GPU, held until the process ends.  An evaluation set
 * needing more scratch (about P x N > 15 million, e.g. more than 64 particles at
 * N = 200,000) is rejected with PASTA_ERR_ARG.  Calls in one process are serialised; the
 * array keeps the last call's contents.  Runs on the null stream. */
int pasta_eval_chi2_noalloc(pasta_star *s, const double *labels, int P, double *chi2,
                            int *status, float *params);
/* chi2 of n stars with one label vector each (labels n x 5), in one batched launch:
 * the evaluation behind chi2_bulk (src/lib.rs:1297-1338), which builds each star from
 * its own wavelengths with NoConvolution and ChunkFitter(5, 8, 0.2).  chi2[k] is NaN
 * and status[k] the PASTA_ST_* reason for a failed star; returns PASTA_ERR_MODEL then. */
int pasta_eval_chi2_multi(pasta_star *const *stars, int n, const double *labels, double *chi2,
                          int *status);
/* produce_model / produce_model_bulk (src/interpolate.rs:1068-1090, src/lib.rs:993-1011,
 * 1064-1089): labels P x 5 -> out P x n_obs */
int pasta_produce_model(pasta_grid *g, const pasta_disp *d, const double *labels, int P,
                        float *out, int *status);
/* interpolate (src/interpolate.rs:930-983, src/lib.rs:1111-1123): labels P x 3 -> P x n_pix */
int pasta_interpolate(pasta_grid *g, const double *labels, int P, float *out, int *status);
/* interpolate_and_convolve (src/interpolate.rs:1030-1047, src/lib.rs:1141-1157):
 * labels P x 4 -> P x n_pix */
int pasta_interpolate_and_convolve(pasta_grid *g, const pasta_disp *d, const double *labels,
                                   int P, float *out, int *status);
/* get_vsini_kernel (src/lib.rs:1962-1969): returns the tap count in *len */
int pasta_vsini_kernel(const pasta_grid *g, double vsini, float *out, int cap, int *len);

/* ------------------------------------------------------------------ fit --
 * SingleFitter::fit (src/fitting.rs:399-458) with the PSO of src/particleswarm.rs
 * and the bounds of src/bounds.rs.  The reference seeds its RNG from entropy
 * (src/particleswarm.rs:65); here `seed` selects a counter-based SplitMix64
 * stream, so a fit is reproducible.  constraint kinds: 0 none, 1 fixed (a),
 * 2 range [a, b] (FixConstraint / RangeConstraint, src/lib.rs:817-835). */
typedef struct pasta_fit_settings {
    int num_particles;
    int max_iters;
    double inertia_factor;    /* default 0.7213475204 (src/lib.rs:66) */
    double cognitive_factor;  /* default 1.1931471806 */
    double social_factor;     /* default 1.1931471806 */
    double delta;             /* default 1e-7 (stored, unused, as in the reference) */
    double vsini_range[2];
    double rv_range[2];
    int constraint_kind[5];
    double constraint_a[5];
    double constraint_b[5];
} pasta_fit_settings;

typedef struct pasta_fit_result {
    double label[5];
    double chi2;
    int iterations;
    int status;               /* PASTA_ST_* of the failure, 0 on success */
    double time_s;            /* wall time of the fit */
    int n_params;
    float continuum_params[512];
} pasta_fit_result;

/* The plain path: one star, one PSO iteration per host round trip.
 * init_pos: optional P x 5 starting positions (population injection,
 * src/particleswarm.rs:186-195).  trace: optional (max_iters+1) x 6 of
 * (global-best cost, position) per iteration. */
int pasta_fit(pasta_star *s, const pasta_fit_settings *st, uint64_t seed, const double *init_pos,
              pasta_fit_result *out, double *trace);

/* pasta_fit with the reference's PSOobserver (src/fitting.rs:217-256; trace_directory of
 * Python fit, src/lib.rs:1476-1507,838-848): population receives, for every iteration
 * it = 0 .. max_iters-1, the swarm after that iteration as P x 6 doubles (position,
 * cost) in particle order, the content of the reference's it_{it}.json. */
int pasta_fit_observed(pasta_star *s, const pasta_fit_settings *st, uint64_t seed,
                       const double *init_pos, pasta_fit_result *out, double *trace,
                       double *population);

/* The batched path (replaces fit_bulk, src/lib.rs:1511-1540): every particle of
 * every star in one device-driven, hipGraph-captured PSO iteration. */
int pasta_fit_bulk(pasta_star *const *stars, int n_stars, const pasta_fit_settings *st,
                   const uint64_t *seeds, pasta_fit_result *out, double *trace);

/* ----------------------------------------------------------- uncertainty --
 * SingleFitter::compute_uncertainty (src/fitting.rs:482-568; Python entry and
 * default search radius (2000, 0.3, 0.3, 40, 40) at src/lib.rs:1545-1568): for every
 * label i, the distances below (out[2i]) and above (out[2i+1]) `label` at which
 * chi2 crosses best * (1 + sqrt(2/n)), n = 4 R (wl_last - wl_first) / (wl_first +
 * wl_last), found by Brent root searches (argmin BrentRoot, tolerance 1e-3 of the
 * search radius).  status[j] = 0, or the PASTA_ST_* reason that side has no value
 * (the reference's None).  search_radius NULL = the defaults.  Returns
 * PASTA_ERR_MODEL when the model at `label` itself fails (the reference raises). */
int pasta_compute_uncertainty(pasta_star *s, double spec_res, const double label[5],
                              const double search_radius[5], double out[10], int status[10]);
/* compute_uncertainty_bulk (src/lib.rs:1570-1600): n stars, labels n x 5, out and
 * status n x 10; all searches advance together, one batched evaluation per round.  A
 * star whose best model fails gets that status on all ten sides. */
int pasta_compute_uncertainty_bulk(pasta_star *const *stars, int n, double spec_res,
                                   const double *labels, const double *search_radius,
                                   double *out, int *status);

/* ---------------------------------------------------------------- binary --
 * Binary stars.  Replaces interpolate_linear (src/interpolate.rs:985-1030, Python
 * src/lib.rs:1127-1139), produce_binary_model_norm (src/interpolate.rs:822-890,
 * src/convolve_rv.rs:606-690, Python src/lib.rs:1013-1035) and BinaryFitter
 * (src/fitting.rs:571-728 with BinaryBounds, src/bounds.rs:188-276; Python
 * get_binary_fitter / fit / chi2, src/lib.rs:1177-1193,1631-1700).  Binary labels are 11
 * values: (Teff, [M/H], logg, vsini, RV) of star 1, the same of star 2, then the light
 * ratio.  The continuum grid `cg` holds unnormalised models on the same synthetic
 * wavelengths; its linear interpolation weights the two normalised components. */
/* interpolate_linear: labels P x 3 -> out P x n_pix (NaN rows + status on failure) */
int pasta_interpolate_linear(pasta_grid *g, const double *labels, int P, float *out,
                             int *status);
typedef struct pasta_binary pasta_binary;
/* An observed binary (the BinaryCostFunction fields, src/fitting.rs:571-578); c, flux and
 * var may all be NULL for model-only use (pasta_binary_model). */
int pasta_binary_create(pasta_grid *g, pasta_grid *cg, const pasta_disp *d, const pasta_cont *c,
                        const float *flux, const float *var, int n, pasta_binary **out);
int pasta_binary_destroy(pasta_binary *b);
/* produce_binary_model_norm for labels P x 11 -> out P x n (light ratio as f32) */
int pasta_binary_model(pasta_binary *b, const double *labels, int P, float *out, int *status);
/* BinaryCostFunction::cost = BinaryFitter::chi2 (src/fitting.rs:581-604,708-728): labels
 * P x 11 -> chi2 (NaN + status on failure); params (nullable) P x n_params. */
int pasta_binary_eval_chi2(pasta_binary *b, const double *labels, int P, double *chi2,
                           int *status, float *params);
typedef struct pasta_binary_settings {
    int num_particles;
    int max_iters;
    double inertia_factor;
    double cognitive_factor;
    double social_factor;
    double delta;
    double vsini_range[2];
    double rv_range[2];
    double light_ratio_range[2]; /* (0, 1) in BinaryFitter::fit (src/fitting.rs:656-661) */
    int constraint_kind[11];     /* per binary label: 0 none, 1 fixed (a), 2 range [a, b] */
    double constraint_a[11];
    double constraint_b[11];
} pasta_binary_settings;
typedef struct pasta_binary_result {
    double label1[5];
    double label2[5];
    double light_ratio;
    double chi2;
    int iterations;
    int status;
    double time_s;
    int n_params;
    float continuum_params[512];
} pasta_binary_result;
/* BinaryFitter::fit (src/fitting.rs:635-706): the PSO over the 11 labels, one batched
 * evaluation per iteration; seed and init_pos (P x 11) as pasta_fit; trace optional
 * (max_iters+1) x 12 of (global-best cost, position). */
int pasta_binary_fit(pasta_binary *b, const pasta_binary_settings *st, uint64_t seed,
                     const double *init_pos, pasta_binary_result *out, double *trace);

/* BinaryTimeriesKnownRVFitter (src/fitting.rs:882-1078; Python
 * get_timeseries_known_rv_fitter, src/lib.rs:1212-1227,1781-1855): S spectra of one binary
 * on the dispersion's wavelengths (fluxes, vars S x n) with known RVs (rvs S x 2); labels
 * are 9 values, (Teff, [M/H], logg, vsini) of each star and the light ratio, and the
 * chi2 of a label is the spectrum-ordered sum (BinaryBoundsWithoutRV, src/bounds.rs:278-345). */
int pasta_binary_ts_create(pasta_grid *g, pasta_grid *cg, const pasta_disp *d,
                           const pasta_cont *c, const float *fluxes, const float *vars,
                           const double *rvs, int S, int n, pasta_binary **out);
int pasta_binary_ts_eval_chi2(pasta_binary *b, const double *labels, int P, double *chi2,
                              int *status);
typedef struct pasta_ts_result {
    double label1[4];
    double label2[4];
    double light_ratio;
    double chi2;
    int iterations;
    int status;
    double time_s;
    int n_params;   /* per spectrum */
    int n_spectra;
} pasta_ts_result;
/* continuum_params (nullable) S x n_params and chis (nullable) S of the best labels;
 * trace optional (max_iters+1) x 10 */
int pasta_binary_ts_fit(pasta_binary *b, const pasta_binary_settings *st, uint64_t seed,
                        const double *init_pos, pasta_ts_result *out, float *continuum_params,
                        double *chis, double *trace);

/* BinaryRVFitter (src/fitting.rs:730-880 with RVCostFunction and BinaryRVBounds; Python
 * get_binary_rv_fitter, src/lib.rs:1195-1210,1702-1779): the two components and their
 * continua are fixed synthetic-grid arrays (n_pix each) and only (rv1, rv2) are fitted;
 * the settings' rv_range and constraints 0-1 apply. */
typedef struct pasta_binary_rv pasta_binary_rv;
int pasta_binary_rv_create(pasta_grid *g, const pasta_disp *d, const pasta_cont *c,
                           const float *flux, const float *var, int n, const float *model1,
                           const float *model2, const float *continuum1,
                           const float *continuum2, double light_ratio, pasta_binary_rv **out);
int pasta_binary_rv_destroy(pasta_binary_rv *r);
/* RVCostFunction::cost for rvs P x 2 */
int pasta_binary_rv_eval_chi2(pasta_binary_rv *r, const double *rvs, int P, double *chi2,
                              int *status);
typedef struct pasta_rv_result {
    double rv1;
    double rv2;
    double chi2;      /* the swapped pair's chi2 when lower (src/fitting.rs:846-863) */
    int iterations;
    int status;
    double time_s;
} pasta_rv_result;
/* trace optional (max_iters+1) x 3 */
int pasta_binary_rv_fit(pasta_binary_rv *r, const pasta_binary_settings *st, uint64_t seed,
                        const double *init_pos, pasta_rv_result *out, double *trace);

/* ---------------------------------------------------------------- server --
 * Continuous batching of fit requests over a resident grid: requests are
 * admitted into free slots (and free HBM) between PSO iterations, every
 * admitted star's particles share each iteration's launches, and a finished
 * star leaves its slot at once.  Results do not depend on the mix: each star
 * has its own RNG stream and no reduction crosses stars. */
typedef struct pasta_server pasta_server;
typedef struct pasta_server_config {
    int max_stars;            /* concurrent stars (slots) */
    int max_particles;        /* particles per star (upper bound) */
    int iters_per_step;       /* PSO iterations per pasta_server_step */
    int64_t hbm_budget_bytes; /* 0 = use free memory minus 4 GiB */
    int timing;               /* 1 = run iterations as eager launches timed with HIP events */
} pasta_server_config;
int pasta_server_create(pasta_grid *g, const pasta_server_config *cfg, pasta_server **out);
int pasta_server_destroy(pasta_server *sv);
/* The server takes ownership of `s` on success: its buffers are freed when the request
 * finishes or is cancelled (the result stays readable through pasta_server_poll). */
int pasta_server_submit(pasta_server *sv, pasta_star *s, const pasta_fit_settings *st,
                        uint64_t seed, uint64_t *request_id);
/* Admit waiting requests, run iters_per_step iterations, retire finished stars.
 * *active = stars in flight after the step, *waiting = queued. */
int pasta_server_step(pasta_server *sv, int *active, int *waiting);
/* done: 1 finished (result filled), 0 pending; PASTA_ERR_STATE if unknown. */
int pasta_server_poll(pasta_server *sv, uint64_t request_id, int *done, pasta_fit_result *out);
int pasta_server_cancel(pasta_server *sv, uint64_t request_id);
/* counters: [0] evals done, [1] requests finished, [2] iterations run,
 * [3] HBM bytes in use by in-flight stars (their evaluation buffers, PSO state and the
 * stars themselves), [4] device-busy microseconds (first evaluations of admitted stars and
 * PSO iterations, HIP events on the server's stream), [5] batch rebuilds (admissions and
 * retirements; the PSO state of stars in flight moves between batches on the device),
 * [6] host microseconds spent in those rebuilds */
int pasta_server_stats(const pasta_server *sv, int64_t out[7]);
/* With timing on: [0] last iteration's k_model ms, [1] sum of k_model ms, [2] last
 * iteration ms (PSO update through best update), [3] sum of iteration ms,
 * [4] iterations summed, [5] sum of the timed k_model launches' algorithmic FP32 flops,
 * [6] and of their algorithmic bytes (per evaluation, as the reference's algorithm
 * counts them: 128 W + 2 (K_rot + K_instr) W + 6 N flops, 128 W + 8 N bytes, W the
 * synthetic span, N the observed pixels, K_rot from the evaluation's vsini).
 * reset != 0 clears the sums after reading. */
int pasta_server_timing(pasta_server *sv, int reset, double out[7]);
/* Switch between graph replays (0, the default) and eager launches timed with
 * HIP events (1); both run the same kernels on the server's stream. */
int pasta_server_set_timing(pasta_server *sv, int on);

#ifdef __cplusplus
}
#endif
#endif
