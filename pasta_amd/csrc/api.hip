// api.hip -- host driver behind the C ABI in include/pasta_hip.h.
//
// Host-side set-up restated from the reference (SimonVanSchuylenbergh/PASTA):
//   Grid::new / list_gridpoints              src/interpolate.rs:267-348
//   WlGrid::get_float_index_of_wl (Logspace)  src/interpolate.rs:705-708
//   FixedTargetDispersionLogSpace::new        src/convolve_rv.rs:169-195
//   VariableTargetDispersion::new             src/convolve_rv.rs:303-331,350-402
//   ChunkFitter::new                          src/continuum_fitting.rs:140-214
//   PSO init / loop / best update             src/particleswarm.rs:111-150,180-307
//   SingleFitter::fit                         src/fitting.rs:399-458
// The least-squares solve of each chunk (an f32 SVD in the reference,
// src/continuum_fitting.rs:181,223) is done as FP64 normal equations: the Gram
// matrix AᵀA of the f32 design values is fixed per star, so its Cholesky factor
// is computed here once and the device only forms Aᵀy per evaluation.
#include "host.h"

static thread_local std::string g_err;
int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

const char* pasta_last_error(void) { return g_err.c_str(); }

std::shared_mutex& capture_mutex() {
    static std::shared_mutex m;
    return m;
}
int pasta_version(void) { return 100; }

int pasta_device_mem_info(uint64_t* free_b, uint64_t* total_b) {
    size_t f = 0, t = 0;
    HIPCHK(hipMemGetInfo(&f, &t));
    if (free_b) *free_b = f;
    if (total_b) *total_b = t;
    return PASTA_OK;
}

// ---------------------------------------------------------------------------
// grid
// ---------------------------------------------------------------------------
static int insert_unique_sorted(std::vector<double>& v, double x) {
    auto it = std::lower_bound(v.begin(), v.end(), x);
    if (it != v.end() && *it == x) return 0;
    v.insert(it, x);
    return 1;
}

int pasta_grid_create(const uint16_t* spectra, const float* factors, const double* labels,
                      int n_models, int n_pix, double log10_first, double log10_step,
                      pasta_grid** out) {
    if (!labels || n_models <= 0 || n_pix <= 1 || !out || !(log10_step > 0))
        return fail(PASTA_ERR_ARG, "pasta_grid_create: bad arguments");
    std::vector<double> t, m, l;
    for (int i = 0; i < n_models; i++) {
        insert_unique_sorted(t, labels[3 * i]);
        insert_unique_sorted(m, labels[3 * i + 1]);
        insert_unique_sorted(l, labels[3 * i + 2]);
    }
    if ((int)t.size() > PASTA_MAX_AXIS || (int)m.size() > PASTA_MAX_AXIS ||
        (int)l.size() > PASTA_MAX_AXIS)
        return fail(PASTA_ERR_ARG, "grid axis longer than 128 values");
    auto g = std::make_unique<pasta_grid>();
    GridGeom& G = g->geom;
    G.n_teff = (int)t.size(); G.n_m = (int)m.size(); G.n_logg = (int)l.size();
    std::copy(t.begin(), t.end(), G.teff);
    std::copy(m.begin(), m.end(), G.m);
    std::copy(l.begin(), l.end(), G.logg);
    for (int i = 0; i < G.n_teff; i++) { G.logg_lo[i] = G.n_logg - 1; G.logg_hi[i] = 0; }
    for (int q = 0; q < n_models; q++) {
        int i = right_index(G.teff, G.n_teff, labels[3 * q]);
        int k = right_index(G.logg, G.n_logg, labels[3 * q + 2]);
        G.logg_lo[i] = std::min(G.logg_lo[i], k);
        G.logg_hi[i] = std::max(G.logg_hi[i], k);
    }
    int acc = 0;
    for (int i = 0; i < G.n_teff; i++) {
        if (G.logg_lo[i] > G.logg_hi[i]) return fail(PASTA_ERR_ARG, "Invalid logg limits");
        G.cum[i] = acc;
        acc += G.logg_hi[i] - G.logg_lo[i] + 1;
    }
    G.n_models = acc * G.n_m;
    if (G.n_models != n_models)
        return fail(PASTA_ERR_ARG, "grid labels do not form a staircase grid (" +
                                       std::to_string(n_models) + " rows, " +
                                       std::to_string(G.n_models) + " grid points)");
    // storage row of each input row; every grid point exactly once
    std::vector<int> row_of(n_models), seen(n_models, 0);
    for (int q = 0; q < n_models; q++) {
        int i = right_index(G.teff, G.n_teff, labels[3 * q]);
        int j = right_index(G.m, G.n_m, labels[3 * q + 1]);
        int k = right_index(G.logg, G.n_logg, labels[3 * q + 2]);
        int r = model_index(G, i, j, k);
        if (r < 0 || r >= n_models || seen[r]) return fail(PASTA_ERR_ARG, "duplicate grid label");
        seen[r] = 1;
        row_of[q] = r;
    }
    g->W = n_pix;
    g->pitch = ((long long)n_pix + 7) & ~7LL;
    g->first = log10_first;
    g->step = log10_step;
    // WlGrid::get_first_and_last (src/interpolate.rs:695-698): one step past the last pixel
    g->first_wl = std::pow(10.0, log10_first);
    g->last_wl = std::pow(10.0, log10_first + log10_step * (double)n_pix);
    CHK(g->d_geom.upload(&G, 1));
    CHK(g->d_spectra.alloc((size_t)n_models * g->pitch));
    HIPCHK(hipMemset(g->d_spectra.p, 0, (size_t)n_models * g->pitch * sizeof(uint16_t)));
    if (spectra) {
        for (int q = 0; q < n_models; q++)
            HIPCHK(hipMemcpy(g->d_spectra.p + (size_t)row_of[q] * g->pitch,
                             spectra + (size_t)q * n_pix, (size_t)n_pix * sizeof(uint16_t),
                             hipMemcpyHostToDevice));
    }
    if (factors) {
        std::vector<float> f(n_models);
        for (int q = 0; q < n_models; q++) f[row_of[q]] = factors[q];
        CHK(g->d_factors.upload(f));
        g->has_factors = true;
    }
    HIPCHK(hipDeviceSynchronize());
    *out = g.release();
    return PASTA_OK;
}

int pasta_grid_destroy(pasta_grid* g) {
    delete g;
    return PASTA_OK;
}

int pasta_grid_shape(const pasta_grid* g, int out[5]) {
    if (!g) return fail(PASTA_ERR_ARG, "null grid");
    out[0] = g->geom.n_teff; out[1] = g->geom.n_m; out[2] = g->geom.n_logg;
    out[3] = g->geom.n_models; out[4] = g->W;
    return PASTA_OK;
}

int pasta_grid_axes(const pasta_grid* g, double* teff, double* m, double* logg, int* lo, int* hi,
                    int* cum) {
    if (!g) return fail(PASTA_ERR_ARG, "null grid");
    const GridGeom& G = g->geom;
    for (int i = 0; i < G.n_teff; i++) {
        if (teff) teff[i] = G.teff[i];
        if (lo) lo[i] = G.logg_lo[i];
        if (hi) hi[i] = G.logg_hi[i];
        if (cum) cum[i] = G.cum[i];
    }
    for (int i = 0; i < G.n_m; i++) if (m) m[i] = G.m[i];
    for (int i = 0; i < G.n_logg; i++) if (logg) logg[i] = G.logg[i];
    return PASTA_OK;
}

int pasta_grid_device_rows(const pasta_grid* g, uint16_t** rows, int64_t* pitch) {
    if (!g) return fail(PASTA_ERR_ARG, "null grid");
    *rows = g->d_spectra.p;
    *pitch = g->pitch;
    return PASTA_OK;
}

int pasta_grid_read_rows(const pasta_grid* g, int row0, int nrows, uint16_t* out) {
    if (!g || !out || row0 < 0 || nrows < 0 || row0 + nrows > g->geom.n_models)
        return fail(PASTA_ERR_ARG, "bad row range");
    HIPCHK(hipMemcpy2D(out, (size_t)g->W * sizeof(uint16_t),
                       g->d_spectra.p + (size_t)row0 * g->pitch, (size_t)g->pitch * sizeof(uint16_t),
                       (size_t)g->W * sizeof(uint16_t), (size_t)nrows, hipMemcpyDeviceToHost));
    return PASTA_OK;
}

int pasta_grid_set_fft_taps(pasta_grid* g, int min_taps) {
    if (!g) return fail(PASTA_ERR_ARG, "null grid");
    g->fft_min_taps = min_taps <= 0 ? FFT_MIN_TAPS : min_taps;
    return PASTA_OK;
}

int pasta_grid_set_factors(pasta_grid* g, const float* f) {
    if (!g) return fail(PASTA_ERR_ARG, "null grid");
    if (!f) { g->has_factors = false; return PASTA_OK; }
    CHK(g->d_factors.upload(f, g->geom.n_models));
    g->has_factors = true;
    return PASTA_OK;
}

int pasta_grid_local_weights(const pasta_grid* g, double teff, double m, double logg, int idx[64],
                             float w[64]) {
    if (!g) return fail(PASTA_ERR_ARG, "null grid");
    int rc = local_weights(g->geom, teff, m, logg, idx, w);
    if (rc) return fail(PASTA_ERR_MODEL, "labels outside the grid / not enough neighbours");
    return PASTA_OK;
}

int pasta_grid_clamp_1d(const pasta_grid* g, double teff, double m, double logg, int index,
                        double* out) {
    if (!g || index < 0 || index > 2) return fail(PASTA_ERR_ARG, "bad clamp_1d arguments");
    Bounds b{};
    b.vsini_lo = 0; b.vsini_hi = 0; b.rv_lo = 0; b.rv_hi = 0;
    double p[5] = {teff, m, logg, 0, 0};
    if (bounds_clamp(g->geom, b, p, index, out)) return fail(PASTA_ERR_MODEL, "Cannot clamp");
    return PASTA_OK;
}

int pasta_grid_is_teff_logg_between_bounds(const pasta_grid* g, double teff, double logg) {
    return g && teff_logg_between_bounds(g->geom, teff, logg) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// dispersion
// ---------------------------------------------------------------------------
// make_gaussian (src/convolve_rv.rs:107-118)
static void make_gaussian(float* arr, int n, float sigma) {
    int mid = (n - 1) / 2;
    for (int i = 0; i < n; i++) {
        float x = ((float)i - (float)mid) / sigma;
        arr[i] = expf(-0.5f * x * x);
    }
    float s = 0.0f;
    for (int i = 0; i < n; i++) s = s + arr[i];
    for (int i = 0; i < n; i++) arr[i] = arr[i] / s;
}

static int disp_common(const pasta_grid* g, const double* wl, int n, pasta_disp* d) {
    if (!g || !wl || n < 2) return fail(PASTA_ERR_ARG, "dispersion: bad arguments");
    d->wl.assign(wl, wl + n);
    d->fidx.resize(n);
    for (int i = 0; i < n; i++) d->fidx[i] = (std::log10(wl[i]) - g->first) / g->step;
    for (int i = 1; i < n; i++)
        if (!(d->fidx[i] >= d->fidx[i - 1]))
            return fail(PASTA_ERR_ARG, "observed wavelengths must be sorted ascending");
    d->W = g->W;
    d->first = g->first;
    d->step = g->step;
    return PASTA_OK;
}

int pasta_disp_create_none(const pasta_grid* g, const double* wl, int n, pasta_disp** out) {
    auto d = std::make_unique<pasta_disp>();
    CHK(disp_common(g, wl, n, d.get()));
    d->kind = DISP_NONE;
    *out = d.release();
    return PASTA_OK;
}

int pasta_disp_create_fixed(const pasta_grid* g, const double* wl, int n, double resolution,
                            pasta_disp** out) {
    if (!(resolution > 0.0)) return fail(PASTA_ERR_ARG, "Resolution must be positive");
    auto d = std::make_unique<pasta_disp>();
    CHK(disp_common(g, wl, n, d.get()));
    d->kind = DISP_FIXED;
    double sigma = 1.0 / (LN_10 * 2.355 * resolution * g->step);
    long long nme = (long long)std::ceil(6.0 * sigma);
    long long K = nme + 1 - nme % 2;
    if (K > 2 * HD_MAX + 1)
        return fail(PASTA_ERR_ARG, "instrumental kernel longer than 255 taps (resolution too low)");
    d->K = (int)K;
    d->kern.resize(K);
    make_gaussian(d->kern.data(), (int)K, (float)sigma);
    *out = d.release();
    return PASTA_OK;
}

int pasta_disp_create_variable(const pasta_grid* g, const double* wl, int n, const float* disp,
                               pasta_disp** out) {
    if (!disp) return fail(PASTA_ERR_ARG, "null dispersion array");
    auto d = std::make_unique<pasta_disp>();
    CHK(disp_common(g, wl, n, d.get()));
    d->kind = DISP_VARIABLE;
    const int W = g->W;
    // resample_observed_to_synth (src/convolve_rv.rs:303-331)
    std::vector<float> dr(W), sig(W);
    int j = 0;
    for (int i = 0; i < W; i++) {
        double x = std::pow(10.0, g->first + g->step * (double)i);
        while (j < n - 1 && wl[j + 1] < x) j++;
        if (j == n - 1) {
            dr[i] = disp[j];
        } else {
            float w = (float)((x - wl[j]) / (wl[j + 1] - wl[j]));
            dr[i] = (1.0f - w) * disp[j] + w * disp[j + 1];
        }
    }
    float largest = -INFINITY;
    for (int i = 0; i < W; i++) {
        double x = std::pow(10.0, g->first + g->step * (double)i);
        sig[i] = dr[i] / (float)(g->step * LN_10 * x);
        if (std::isnan(sig[i])) return fail(PASTA_ERR_ARG, "NaN dispersion");
        largest = std::max(largest, sig[i]);
    }
    long long kme = (long long)std::ceil(largest * 6.0f);
    long long K = kme + 1 - kme % 2;
    if (K < 1 || K > 2 * HD_MAX + 1)
        return fail(PASTA_ERR_ARG, "variable instrumental kernel longer than 255 taps");
    d->K = (int)K;
    d->kern.resize((size_t)W * K);
    for (int i = 0; i < W; i++) make_gaussian(d->kern.data() + (size_t)i * K, (int)K, sig[i]);
    *out = d.release();
    return PASTA_OK;
}

int pasta_disp_destroy(pasta_disp* d) {
    delete d;
    return PASTA_OK;
}

int pasta_disp_kernel(const pasta_disp* d, float* out, int64_t cap, int* k_len) {
    if (!d) return fail(PASTA_ERR_ARG, "null dispersion");
    if (k_len) *k_len = d->K;
    if (!out) return PASTA_OK;
    if (d->kind == DISP_FIXED) {
        if (cap < d->K) return fail(PASTA_ERR_ARG, "buffer too small");
        std::copy(d->kern.begin(), d->kern.end(), out);
    } else if (d->kind == DISP_VARIABLE) {
        if (cap < (int64_t)d->kern.size()) return fail(PASTA_ERR_ARG, "buffer too small");
        for (int k = 0; k < d->K; k++)
            for (int i = 0; i < d->W; i++) out[(size_t)k * d->W + i] = d->kern[(size_t)i * d->K + k];
    }
    return PASTA_OK;
}

// ---------------------------------------------------------------------------
// continuum
// ---------------------------------------------------------------------------
static double map_range_d(double x, double fl, double fh, double tl, double th, bool clamp) {
    double v = (x - fl) / (fh - fl) * (th - tl) + tl;
    if (clamp) v = std::min(std::max(v, tl), th);
    return v;
}
static float map_range_f(float x, float fl, float fh, float tl, float th) {
    return (x - fl) / (fh - fl) * (th - tl) + tl;
}
static float poly_blend(float x) {  // src/continuum_fitting.rs:119-123
    float xc = x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x);
    float s = sinf(xc * (float)M_PI / 2.0f);
    return s * s;
}

static int cholesky(std::vector<double>& A, int n) {  // in place, lower
    for (int j = 0; j < n; j++) {
        double s = A[j * n + j];
        for (int k = 0; k < j; k++) s -= A[j * n + k] * A[j * n + k];
        if (!(s > 0.0)) return -1;
        double d = std::sqrt(s);
        A[j * n + j] = d;
        for (int i = j + 1; i < n; i++) {
            double t = A[i * n + j];
            for (int k = 0; k < j; k++) t -= A[i * n + k] * A[j * n + k];
            A[i * n + j] = t / d;
        }
        for (int i = 0; i < j; i++) A[i * n + j] = 0.0;
    }
    return 0;
}

// chi2 blocks: (kind, chunk, i0, i1)
static void push_blocks(std::vector<int4>& out, int kind, int chunk, int i0, int i1) {
    for (int s = i0; s < i1; s += BLK) out.push_back(make_int4(kind, chunk, s, std::min(s + BLK, i1)));
}
// moment blocks (chunk, i0, i1, second chunk or -1), each with a partial slot per chunk:
// slot_a is the next slot; slot_b is assigned when the second chunk's slots are laid out
static void push_mblocks(pasta_cont& c, int chunk, int i0, int i1, int chunk2, int& slot) {
    for (int s = i0; s < i1; s += BLK) {
        c.mblocks.push_back(make_int4(chunk, s, std::min(s + BLK, i1), chunk2));
        c.mslot.push_back(make_int2(slot++, -1));
    }
}

// Moment blocks of a chunk continuum.  The pixels of the overlap of chunks ch and ch + 1
// feed both chunks' moments; their blocks carry both chunks, so that one read of the
// ratios serves the two products.  A chunk's partial slots are contiguous and in pixel
// order: the overlap with the previous chunk, its own pixels, the overlap with the next.
// Two-chunk blocks need a pixel in at most two chunks and the two parameter sets in one
// 32-column matrix-core product (p1 <= 16); otherwise every chunk gets blocks of its own.
static void chunk_mblocks(pasta_cont& c) {
    const int nc = c.n_chunks;
    const int* ss = c.startstop.data();
    bool shared = c.p1 <= 16;
    for (int ch = 1; ch + 1 < nc && shared; ch++)
        if (ss[2 * (ch + 1)] < ss[2 * (ch - 1) + 1]) shared = false;  // a pixel in 3 chunks
    c.mblocks.clear();
    c.mslot.clear();
    c.mb_off.assign(1, 0);
    int slot = 0;
    int right0 = 0, right1 = 0;  // blocks of the previous chunk's overlap with this one
    for (int ch = 0; ch < nc; ch++) {
        const int start = ss[2 * ch], stop = ss[2 * ch + 1];
        if (!shared) {
            push_mblocks(c, ch, start, stop, -1, slot);
            c.mb_off.push_back(slot);
            continue;
        }
        for (int b = right0; b < right1; b++) c.mslot[b].y = slot++;
        const int lo = ch > 0 ? std::max(start, ss[2 * (ch - 1) + 1]) : start;
        const int next = ch + 1 < nc ? ss[2 * (ch + 1)] : stop;
        const int hi = std::min(stop, next);
        push_mblocks(c, ch, lo, hi, -1, slot);
        right0 = right1 = (int)c.mblocks.size();
        if (ch + 1 < nc && next < stop) {
            push_mblocks(c, ch, next, stop, ch + 1, slot);
            right1 = (int)c.mblocks.size();
        }
        c.mb_off.push_back(slot);
    }
}

static void host_design_row(double xh, int p1, float* a) {
    double sq[5];
    sq[0] = xh;
    for (int t = 1; t < 5; t++) sq[t] = sq[t - 1] * sq[t - 1];
    for (int k = 0; k < p1; k++) {
        double r = 1.0;
        for (int t = 0; t < 5; t++)
            if (k & (1 << t)) r *= sq[t];
        a[k] = (float)r;
    }
}

int pasta_cont_create_chunk(const double* wl, int n, int n_chunks, int p_order, double overlap,
                            pasta_cont** out) {
    if (!wl || n < 200 || n_chunks < 2 || n_chunks > CH_MAX || p_order < 0 || p_order + 1 > 31)
        return fail(PASTA_ERR_ARG, "chunk continuum: bad arguments (n >= 200, 2..32 chunks, "
                                   "degree <= 30)");
    auto c = std::make_unique<pasta_cont>();
    c->kind = CONT_CHUNK;
    c->n = n;
    c->n_chunks = n_chunks;
    c->p1 = p_order + 1;
    c->n_params = n_chunks * c->p1;
    c->wl.assign(wl, wl + n);
    c->startstop.resize(2 * n_chunks);
    for (int i = 0; i < n_chunks; i++) {
        double start = map_range_d((double)i - overlap, 0.0, (double)n_chunks, wl[0], wl[n - 1], true);
        double stop = map_range_d((double)i + 1.0 + overlap, 0.0, (double)n_chunks, wl[0],
                                  wl[n - 1], true);
        int a = -1, b = -1;
        for (int x = 0; x < n; x++) if (wl[x] > start) { a = x; break; }
        for (int x = n - 1; x >= 0; x--) if (wl[x] < stop) { b = x; break; }
        if (a < 0 || b < 0) return fail(PASTA_ERR_ARG, "chunk boundary not found");
        c->startstop[2 * i] = a;
        c->startstop[2 * i + 1] = b + 1;
    }
    c->startstop[0] = 0;
    c->startstop[2 * n_chunks - 1] = n;
    const int p1 = c->p1;
    c->lohi.resize(2 * n_chunks);
    c->chol.assign((size_t)n_chunks * p1 * p1, 0.0);
    std::vector<float> a(p1);
    for (int ch = 0; ch < n_chunks; ch++) {
        int s = c->startstop[2 * ch], e = c->startstop[2 * ch + 1];
        if (e - s < p1) return fail(PASTA_ERR_ARG, "chunk shorter than the polynomial order");
        c->lohi[2 * ch] = wl[s];
        c->lohi[2 * ch + 1] = wl[e - 1];
        // design rows for the device: p1 rows from pixel s & ~3, stride rounded up to 4
        const int b0 = s & ~3, stride = (e - b0 + 3) & ~3;
        const size_t off = c->cdesign.size();
        c->cdoff.push_back((int)off);
        c->cdesign.resize(off + (size_t)p1 * stride, 0.0f);
        std::vector<double> G((size_t)p1 * p1, 0.0);
        for (int i = s; i < e; i++) {
            host_design_row(chunk_xhat(wl[i], wl[s], wl[e - 1]), p1, a.data());
            for (int k = 0; k < p1; k++) c->cdesign[off + (size_t)k * stride + (i - b0)] = a[k];
            for (int k = 0; k < p1; k++)
                for (int l = 0; l <= k; l++) G[k * p1 + l] += (double)a[k] * (double)a[l];
        }
        for (int k = 0; k < p1; k++)
            for (int l = k + 1; l < p1; l++) G[k * p1 + l] = G[l * p1 + k];
        if (cholesky(G, p1)) return fail(PASTA_ERR_ARG, "chunk design matrix is rank deficient");
        std::copy(G.begin(), G.end(), c->chol.begin() + (size_t)ch * p1 * p1);
    }
    c->fac.assign(n, 0.0f);
    for (int ch = 0; ch + 1 < n_chunks; ch++) {
        int stop = c->startstop[2 * ch + 1], sn = c->startstop[2 * (ch + 1)];
        for (int i = sn; i < stop; i++)
            c->fac[i] = poly_blend(map_range_f((float)wl[i], (float)wl[sn], (float)wl[stop - 1], 1.0f, 0.0f));
    }
    chunk_mblocks(*c);
    // chi2 segments (src/continuum_fitting.rs:300-358)
    const int* ss = c->startstop.data();
    if (ss[2] < 50) return fail(PASTA_ERR_ARG, "first chunk shorter than the 50-pixel trim");
    push_blocks(c->cblocks, 0, 0, 50, ss[2]);
    for (int ch = 1; ch < n_chunks; ch++) {
        int start = ss[2 * ch], stop_prev = ss[2 * (ch - 1) + 1];
        int start_next = (ch == n_chunks - 1) ? n - 50 : ss[2 * (ch + 1)];
        if (stop_prev < start || start_next < stop_prev)
            return fail(PASTA_ERR_ARG, "chunk overlaps do not form consecutive segments");
        push_blocks(c->cblocks, 1, ch, start, stop_prev);
        push_blocks(c->cblocks, 0, ch, stop_prev, start_next);
    }
    c->chi2_div = (double)(n - 100);
    *out = c.release();
    return PASTA_OK;
}

int pasta_cont_create_linear(const float* design, int n, int ncols, pasta_cont** out) {
    if (!design || n < 101 || ncols < 1 || ncols > P1_MAX)
        return fail(PASTA_ERR_ARG, "linear continuum: bad arguments (1..32 columns)");
    auto c = std::make_unique<pasta_cont>();
    c->kind = CONT_LINEAR;
    c->n = n;
    c->p1 = ncols;
    c->n_params = ncols;
    c->design_rm.assign(design, design + (size_t)n * ncols);
    // device layout: PT = 16 or 32 zero-padded columns of n4 = n rounded up to 4 values
    // (the matrix-core moments read 4 pixels of one column per lane, the chi2 kernel one
    // pixel of each column per thread)
    const int pt = ncols <= 16 ? 16 : 32;
    const size_t n4 = (size_t)(n + 3) & ~(size_t)3;
    c->design_q.assign(n4 * pt, 0.0f);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < ncols; k++) c->design_q[(size_t)k * n4 + i] = design[(size_t)i * ncols + k];
    std::vector<double> G((size_t)ncols * ncols, 0.0);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < ncols; k++)
            for (int l = 0; l <= k; l++)
                G[k * ncols + l] += (double)design[(size_t)i * ncols + k] * (double)design[(size_t)i * ncols + l];
    for (int k = 0; k < ncols; k++)
        for (int l = k + 1; l < ncols; l++) G[k * ncols + l] = G[l * ncols + k];
    if (cholesky(G, ncols)) return fail(PASTA_ERR_ARG, "design matrix is rank deficient");
    c->chol = G;
    {
        int slot = 0;
        push_mblocks(*c, 0, 0, n, -1, slot);
        c->mb_off = {0, slot};
    }
    push_blocks(c->cblocks, 0, 0, 50, n - 50);
    c->chi2_div = (double)n;
    *out = c.release();
    return PASTA_OK;
}

int pasta_cont_create_fixed(const float* continuum, int n, int ignore, pasta_cont** out) {
    if (!continuum || n < 1 || ignore < 0 || 2 * ignore >= n)
        return fail(PASTA_ERR_ARG, "fixed continuum: bad arguments");
    auto c = std::make_unique<pasta_cont>();
    c->kind = CONT_FIXED;
    c->n = n;
    c->ignore = ignore;
    c->fixedc.assign(continuum, continuum + n);
    push_blocks(c->cblocks, 0, 0, ignore, n - ignore);
    c->chi2_div = (double)(n - 2 * ignore);
    *out = c.release();
    return PASTA_OK;
}

int pasta_cont_create_constant(int n, pasta_cont** out) {
    if (n < 1) return fail(PASTA_ERR_ARG, "constant continuum: bad arguments");
    auto c = std::make_unique<pasta_cont>();
    c->kind = CONT_CONSTANT;
    c->n = n;
    // one moment block over all pixels: the mean is one sequential fold (k_const_fold)
    c->mblocks.push_back(make_int4(0, 0, n, -1));
    c->mslot.push_back(make_int2(0, -1));
    c->mb_off = {0, 1};
    push_blocks(c->cblocks, 0, 0, 0, n);
    c->chi2_div = (double)n;
    *out = c.release();
    return PASTA_OK;
}

int pasta_cont_destroy(pasta_cont* c) {
    delete c;
    return PASTA_OK;
}

int pasta_cont_n_params(const pasta_cont* c) { return c ? c->n_params : 0; }

int pasta_cont_chunks(const pasta_cont* c, int* ss, int cap) {
    if (!c || c->kind != CONT_CHUNK) return fail(PASTA_ERR_ARG, "not a chunk continuum");
    if (cap < 2 * c->n_chunks) return fail(PASTA_ERR_ARG, "buffer too small");
    std::copy(c->startstop.begin(), c->startstop.end(), ss);
    return c->n_chunks;
}

// build_continuum (src/continuum_fitting.rs:97-106,228-278,398-413,476-478,502-504)
int pasta_cont_build(const pasta_cont* c, const float* params, int n_params, float* out,
                     int64_t cap) {
    if (!c) return fail(PASTA_ERR_ARG, "null continuum");
    const int n = c->n;
    if (c->kind == CONT_CONSTANT) {
        if (n_params < 1 || cap < 1) return fail(PASTA_ERR_ARG, "bad parameters");
        out[0] = params[0];
        return 1;
    }
    if (cap < n) return fail(PASTA_ERR_ARG, "buffer too small");
    if (c->kind == CONT_FIXED) {
        std::copy(c->fixedc.begin(), c->fixedc.end(), out);
        return n;
    }
    if (n_params != c->n_params)
        return fail(PASTA_ERR_ARG, "Incorrect number of continuum parameters");
    if (c->kind == CONT_LINEAR) {
        for (int i = 0; i < n; i++) {
            float y = 0.0f;
            for (int k = 0; k < c->p1; k++) {
                float t = c->design_rm[(size_t)i * c->p1 + k] * params[k];
                y = (k == 0) ? t : t + y;
            }
            out[i] = y;
        }
        return n;
    }
    const int p1 = c->p1;
    std::vector<float> a(p1);
    std::fill(out, out + n, 0.0f);
    for (int ch = 0; ch < c->n_chunks; ch++) {
        int s = c->startstop[2 * ch], e = c->startstop[2 * ch + 1];
        for (int i = s; i < e; i++) {
            host_design_row(chunk_xhat(c->wl[i], c->wl[s], c->wl[e - 1]), p1, a.data());
            float poly = 0.0f;
            for (int k = 0; k < p1; k++) {
                float t = a[k] * params[ch * p1 + k];
                poly = (k == 0) ? t : t + poly;
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

// ---------------------------------------------------------------------------
// star
// ---------------------------------------------------------------------------
// Each host thread uploads stars on a stream of its own (default priority, non-blocking):
// a copy then waits only for itself, not for the batches running on other streams.
namespace {
struct UploadStream {
    hipStream_t s = nullptr;
    ~UploadStream() {
        if (s) (void)hipStreamDestroy(s);
    }
};
thread_local UploadStream t_upload;
}  // namespace

template <typename T>
static int dev_copy(pasta_star* s, const T* h, size_t n, const T** out) {
    if (n == 0) { *out = nullptr; return PASTA_OK; }
    void* p = nullptr;
    std::shared_lock<std::shared_mutex> lk(capture_mutex());
    if (!t_upload.s) HIPCHK(hipStreamCreateWithFlags(&t_upload.s, hipStreamNonBlocking));
    HIPCHK(hipMalloc(&p, n * sizeof(T)));
    s->allocs.push_back(p);
    s->resident_bytes += (int64_t)(n * sizeof(T));
    HIPCHK(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, t_upload.s));
    HIPCHK(hipStreamSynchronize(t_upload.s));
    *out = (const T*)p;
    return PASTA_OK;
}

pasta_star::~pasta_star() {
    cache.reset();
    std::shared_lock<std::shared_mutex> lk(capture_mutex());
    for (void* p : allocs) (void)hipFree(p);
}

// Star without continuum (produce_model / interpolate paths)
int star_build(pasta_grid* g, const pasta_disp* d, const pasta_cont* c, const float* flux,
                      const float* var, int n, std::unique_ptr<pasta_star>& out) {
    auto s = std::make_unique<pasta_star>();
    s->grid = g;
    s->N = n;
    DevStar& D = s->dev;
    D.N = n;
    D.disp_kind = d ? d->kind : DISP_NONE;
    D.Kd = d ? d->K : 0;
    if (d) {
        CHK(dev_copy(s.get(), d->fidx.data(), d->fidx.size(), &D.fidx));
        CHK(dev_copy(s.get(), d->wl.data(), d->wl.size(), &D.wl));
        CHK(dev_copy(s.get(), d->kern.data(), d->kern.size(), &D.dkern));
        D.wl0 = d->wl.front();
        D.wlN = d->wl.back();
        double f0 = d->fidx.front(), fN = d->fidx.back();
        D.n_tiles = (int)std::floor((fN - f0 + 1.0) / TILE) + 2;
        s->span = (int)(std::floor(fN) - std::floor(f0)) + 2;
    } else {
        D.n_tiles = (g->W + TILE - 1) / TILE;
        s->span = g->W;
    }
    if (flux) CHK(dev_copy(s.get(), flux, n, &D.flux));
    if (var) {  // chi2 terms are r * r * (1 / var): one multiply per evaluation and pixel
        std::vector<float> iv(n);
        for (int i = 0; i < n; i++) iv[i] = 1.0f / var[i];
        CHK(dev_copy(s.get(), iv.data(), n, &D.ivar));
    }
    if (c) {
        D.cont_kind = c->kind;
        D.n_chunks = c->n_chunks;
        D.p1 = c->p1;
        D.ignore = c->ignore;
        D.n_params = c->n_params;
        D.n_mblocks = (int)c->mblocks.size();
        D.n_mslots = c->mb_off.empty() ? 0 : c->mb_off.back();
        D.n_cblocks = (int)c->cblocks.size();
        D.chi2_div = c->chi2_div;
        CHK(dev_copy(s.get(), c->lohi.data(), c->lohi.size(), &D.chunk_lohi));
        CHK(dev_copy(s.get(), c->startstop.data(), c->startstop.size(), &D.startstop));
        CHK(dev_copy(s.get(), c->chol.data(), c->chol.size(), &D.chol));
        CHK(dev_copy(s.get(), c->fac.data(), c->fac.size(), &D.fac));
        CHK(dev_copy(s.get(), c->design_q.data(), c->design_q.size(), &D.design));
        CHK(dev_copy(s.get(), c->cdesign.data(), c->cdesign.size(), &D.cdesign));
        CHK(dev_copy(s.get(), c->cdoff.data(), c->cdoff.size(), &D.cdoff));
        CHK(dev_copy(s.get(), c->fixedc.data(), c->fixedc.size(), &D.fixedc));
        CHK(dev_copy(s.get(), c->mblocks.data(), c->mblocks.size(), &D.mblocks));
        CHK(dev_copy(s.get(), c->mslot.data(), c->mslot.size(), &D.mslot));
        CHK(dev_copy(s.get(), c->mb_off.data(), c->mb_off.size(), &D.mb_off));
        CHK(dev_copy(s.get(), c->cblocks.data(), c->cblocks.size(), &D.cblocks));
        s->n_params = c->n_params;
        s->p1 = c->p1;
        s->cont_kind = c->kind;
        s->n_chunks = c->n_chunks;
    } else {
        D.cont_kind = CONT_FIXED;
        D.n_params = 0;
        D.chi2_div = 1.0;
        s->n_params = 0;
        s->cont_kind = -1;
    }
    out = std::move(s);
    return PASTA_OK;
}

int pasta_star_create(pasta_grid* g, const pasta_disp* d, const pasta_cont* c, const float* flux,
                      const float* var, int n, pasta_star** out) {
    if (!g || !d || !c || !flux || !var || !out) return fail(PASTA_ERR_ARG, "null argument");
    if ((int)d->wl.size() != n)
        return fail(PASTA_ERR_ARG, "Observed spectrum and target dispersion must have the same length");
    if (c->n != n)
        return fail(PASTA_ERR_ARG, "Length of observed and synthetic spectrum are different");
    if (d->W != g->W || d->step != g->step || d->first != g->first)
        return fail(PASTA_ERR_ARG, "dispersion was built for another synthetic grid");
    std::unique_ptr<pasta_star> s;
    CHK(star_build(g, d, c, flux, var, n, s));
    *out = s.release();
    return PASTA_OK;
}

int pasta_star_destroy(pasta_star* s) {
    delete s;
    return PASTA_OK;
}

static int64_t star_eval_bytes(const pasta_star* s) {
    const DevStar& D = s->dev;
    int64_t b = (int64_t)D.N * 4 + sizeof(DevEval) + ROT_STRIDE * 4 + 5 * 8 * 4;
    b += (int64_t)D.n_mslots * std::max(1, D.p1) * 8 + (int64_t)D.n_cblocks * 8 + D.n_params * 4;
    b += (int64_t)(D.n_tiles + D.n_mblocks + D.n_cblocks + std::max(1, D.n_chunks)) * 8;
    return b;
}

int pasta_star_eval_bytes(const pasta_star* s, int64_t* bytes) {
    if (!s) return fail(PASTA_ERR_ARG, "null star");
    *bytes = star_eval_bytes(s);
    return PASTA_OK;
}

// ---------------------------------------------------------------------------
// batch of evaluations
// ---------------------------------------------------------------------------
int batch_build(Batch& B, pasta_grid* g, const std::vector<pasta_star*>& stars,
                       const std::vector<int>& counts, int mode) {
    B.mode = mode;
    B.grid = g;
    B.stars = stars;
    B.n_stars = (int)stars.size();
    B.count = counts;
    B.first_eval.assign(B.n_stars, 0);
    int ne = 0;
    for (int s = 0; s < B.n_stars; s++) {
        B.first_eval[s] = ne;
        ne += counts[s];
    }
    B.n_evals = ne;
    B.h_evals.assign(ne, EvalInit{});
    int64_t mo = 0, mp = 0, cp = 0, pf = 0;
    std::vector<int4> bmom, bchi2;
    std::vector<int2> solve;
    int max_tiles = 0;
    for (int s = 0; s < B.n_stars; s++) max_tiles = std::max(max_tiles, stars[s]->dev.n_tiles);
    for (int s = 0; s < B.n_stars; s++) {
        const DevStar& D = stars[s]->dev;
        const int outlen = mode == 0 ? D.N : g->W;
        for (int q = 0; q < counts[s]; q++) {
            int e = B.first_eval[s] + q;
            EvalInit& ev = B.h_evals[e];
            ev.star = s;
            ev.active = 1;
            if (D.cont_kind == CONT_LINEAR || D.cont_kind == CONT_CHUNK)
                mo = (mo + 3) & ~(int64_t)3;  // float4 reads of the matrix-core moments
            ev.m_off = mo;
            mo += outlen;
            ev.mp_off = mp;
            mp += (int64_t)D.n_mslots * std::max(1, D.p1);
            ev.cp_off = cp;
            cp += D.n_cblocks;
            ev.pf_off = pf;
            pf += std::max(1, D.n_params);
            if (D.cont_kind == CONT_CHUNK)
                for (int c = 0; c < D.n_chunks; c++) solve.push_back(make_int2(e, c));
            else if (D.cont_kind == CONT_LINEAR)
                solve.push_back(make_int2(e, 0));
        }
    }
    // continuum blocks, per class of stars (continuum kind, parameter count): groups of
    // evaluations of one star share each block of pixels
    B.cls.clear();
    std::vector<int> cls_of(B.n_stars);
    for (int s = 0; s < B.n_stars; s++) {
        const DevStar& D = stars[s]->dev;
        const int p1 = D.cont_kind == CONT_CONSTANT ? 1 : (D.cont_kind == CONT_FIXED ? 0 : D.p1);
        int c = 0;
        while (c < (int)B.cls.size() && !(B.cls[c].kind == D.cont_kind && B.cls[c].p1 == p1)) c++;
        if (c == (int)B.cls.size()) B.cls.push_back(ContClass{D.cont_kind, p1, 0, 0, 0, 0});
        cls_of[s] = c;
    }
    for (int c = 0; c < (int)B.cls.size(); c++) {
        ContClass& K = B.cls[c];
        const int gm = mom_variant(K.kind, K.p1).e, gc = chi2_variant(K.kind, K.p1).e;
        K.mom_off = (int)bmom.size();
        K.chi2_off = (int)bchi2.size();
        for (int s = 0; s < B.n_stars; s++) {
            if (cls_of[s] != c) continue;
            const DevStar& D = stars[s]->dev;
            const int e0 = B.first_eval[s], e1 = e0 + counts[s];
            for (int e = e0; e < e1; e += gm)
                for (int b = 0; b < D.n_mblocks; b++)
                    bmom.push_back(make_int4(e, std::min(gm, e1 - e), b, 0));
            for (int e = e0; e < e1; e += gc)
                for (int b = 0; b < D.n_cblocks; b++)
                    bchi2.push_back(make_int4(e, std::min(gc, e1 - e), b, 0));
        }
        K.n_mom = (int)bmom.size() - K.mom_off;
        K.n_chi2 = (int)bchi2.size() - K.chi2_off;
    }
    // model blocks: tile-major over the evaluations (sorted on the device each launch)
    if ((int64_t)max_tiles * ne > 0x7fffffff) return fail(PASTA_ERR_ARG, "batch too large");
    B.model_len = mo; B.mpart_len = mp; B.cpart_len = cp; B.pf_len = pf;
    B.n_bm = max_tiles * ne; B.n_bmom = (int)bmom.size();
    B.n_solve = (int)solve.size(); B.n_bchi2 = (int)bchi2.size();
    std::vector<DevStar> ds(B.n_stars);
    for (int s = 0; s < B.n_stars; s++) ds[s] = stars[s]->dev;
    CHK(B.d_stars.upload(ds));
    if (!B.stream) HIPCHK(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking));
    CHK(B.d_init.upload(B.h_evals));
    CHK(B.d_evals.alloc((size_t)std::max(ne, 1)));
    HIPCHK(launch_eval_init(B.d_init.p, ne, B.d_evals.p, B.stream));
    CHK(B.d_labels.alloc((size_t)ne * 5));
    CHK(B.d_rot.alloc((size_t)ne * ROT_STRIDE));
    CHK(B.d_model.alloc((size_t)std::max<int64_t>(mo, 1)));
    CHK(B.d_mpart.alloc((size_t)std::max<int64_t>(mp, 1)));
    CHK(B.d_cpart.alloc((size_t)std::max<int64_t>(cp, 1)));
    CHK(B.d_pf.alloc((size_t)std::max<int64_t>(pf, 1)));
    CHK(B.d_chi2.alloc(ne));
    CHK(B.d_status.alloc(ne));
    B.tb_stride = max_tiles + 1;
    CHK(B.d_tbound.alloc((size_t)ne * B.tb_stride));
    CHK(B.d_hist.alloc((size_t)g->geom.n_models + 1));
    CHK(B.d_perm.alloc(ne));
    CHK(B.d_pair_eq.alloc(ne));
    CHK(B.d_items.alloc(ne));
    CHK(B.d_union.alloc(ne));
    CHK(B.d_nitems.alloc(1));
    CHK(B.d_bmom.upload(bmom));
    CHK(B.d_solve.upload(solve));
    CHK(B.d_bchi2.upload(bchi2));
    HIPCHK(hipStreamSynchronize(B.stream));  // the records are in place for any stream
    return PASTA_OK;
}

// Whether any evaluation of vsini <= max_vsini on these stars can reach the grid's FFT
// crossover: the rotation kernel's taps, plus the fixed instrumental kernel's when it is
// composed into them.  Launches without that possibility run the kernels built without the
// FFT stage (the path choice is then made with the crossover off, so it never depends on
// this bound being tight).
bool may_fft(const pasta_grid* g, double max_vsini, int extra_taps) {
    if (g->fft_min_taps > KROT_MAX) return false;
    if (!(max_vsini < 1e30)) return true;
    if (!(max_vsini > 0.0)) return false;
    const double eps = 0.6, dvelo = LN_10 * g->step, vrot = max_vsini / C_KMS;
    const double a = 4.0 * (eps - 1.0) / (M_PI * eps);
    const double v = std::ceil(2.0 * vrot / dvelo * std::sqrt(1.0 + a * a));
    if (!(v < 1e6)) return true;
    return (long long)v + 3 + extra_taps >= g->fft_min_taps;
}

// The most composed taps an evaluation of vsini <= max_vsini forms (the rotation kernel's
// length as k_prep computes it, src/convolve_rv.rs:421-431, plus the composed instrumental
// taps), with a margin; KROT_MAX when unbounded.  Sizes k_prep's LDS.
int prep_taps_bound(const pasta_grid* g, double max_vsini, int extra_taps) {
    if (!(max_vsini < 1e30)) return KROT_MAX;
    if (!(max_vsini > 0.0)) return std::min(KROT_MAX, 16 + extra_taps);
    const double eps = 0.6, dvelo = LN_10 * g->step, vrot = max_vsini / C_KMS;
    const double a = 4.0 * (eps - 1.0) / (M_PI * eps);
    const double v = std::ceil(2.0 * vrot / dvelo * std::sqrt(1.0 + a * a));
    if (!(v < 1e6)) return KROT_MAX;
    return (int)std::min<long long>(KROT_MAX, (long long)v + 3 + extra_taps + 16);
}

int composed_extra_taps(const pasta_star* s) {
    return s->dev.disp_kind == DISP_FIXED ? s->dev.Kd - 1 : 0;
}

bool may_fft(const pasta_grid* g, double max_vsini, const std::vector<pasta_star*>& stars) {
    int extra = 0;
    for (const pasta_star* s : stars) extra = std::max(extra, composed_extra_taps(s));
    return may_fft(g, max_vsini, extra);
}

double max_label_vsini(const double* labels5, int n) {
    double m = 0.0;
    for (int i = 0; i < n; i++) {
        const double v = labels5[5 * (size_t)i + 3];
        if (!(v <= m)) m = std::isnan(v) ? HUGE_VAL : v;
    }
    return m;
}

int batch_launch(Batch& B) {
    DevGrid G = B.grid->dev();
    G.stream_models = B.model_len * (long long)sizeof(float) > MALL_BYTES;
    const bool fft = may_fft(B.grid, B.max_vsini, B.stars);
    {
        int extra = 0;
        for (const pasta_star* s : B.stars) extra = std::max(extra, composed_extra_taps(s));
        G.prep_taps = prep_taps_bound(B.grid, B.max_vsini, extra);
    }
    HIPCHK(launch_eval(G, B.d_stars.p, B.d_evals.p, B.n_evals, B.d_labels.p, B.d_rot.p,
                       B.d_tbound.p, B.tb_stride, B.d_hist.p, B.d_perm.p, B.d_pair_eq.p, B.d_items.p,
                       B.d_nitems.p, B.d_union.p, B.n_bm, B.d_bmom.p,
                       B.d_bchi2.p, B.cls.data(), (int)B.cls.size(), B.d_solve.p, B.n_solve,
                       B.d_model.p, B.d_mpart.p, B.d_pf.p, B.d_cpart.p, B.d_chi2.p,
                       B.d_status.p, B.mode, fft, B.eager_events ? B.ev[0] : nullptr,
                       B.eager_events ? B.ev[1] : nullptr, B.stream));
    return PASTA_OK;
}

int batch_eval(Batch& B, const double* labels, double* chi2, int* status) {
    HIPCHK(hipMemcpyAsync(B.d_labels.p, labels, sizeof(double) * 5 * B.n_evals,
                          hipMemcpyHostToDevice, B.stream));
    B.max_vsini = max_label_vsini(labels, B.n_evals);
    CHK(batch_launch(B));
    if (chi2)
        HIPCHK(hipMemcpyAsync(chi2, B.d_chi2.p, sizeof(double) * B.n_evals, hipMemcpyDeviceToHost,
                              B.stream));
    if (status)
        HIPCHK(hipMemcpyAsync(status, B.d_status.p, sizeof(int) * B.n_evals, hipMemcpyDeviceToHost,
                              B.stream));
    HIPCHK(hipStreamSynchronize(B.stream));
    return PASTA_OK;
}

const char* status_text(int st) {
    switch (st) {
    case ST_BOUNDS: return "labels out of the grid bounds";
    case ST_NEIGHBORS: return "not enough neighbours";
    case ST_VSINI: return "vsini must be positive";
    case ST_KERNEL: return "rotation kernel larger than FFTSIZE";
    case ST_RANGE: return "observed wavelength outside the synthetic range after the RV shift";
    case ST_INDEX: return "resample index past the end of the model";
    case ST_PSO: return "particle is not within bounds";
    case ST_SOLVE: return "continuum solve failed";
    default: return "evaluation failed";
    }
}

static int get_cached_batch(pasta_star* s, int P, Batch** out) {
    if (!s->cache || s->cache_P != P) {
        s->cache.reset(new Batch());
        CHK(batch_build(*s->cache, s->grid, {s}, {P}, 0));
        s->cache_P = P;
    }
    *out = s->cache.get();
    return PASTA_OK;
}

int pasta_eval_chi2(pasta_star* s, const double* labels, int P, double* chi2, int* status,
                    float* params) {
    if (!s || !labels || P <= 0 || !chi2) return fail(PASTA_ERR_ARG, "bad arguments");
    if (s->cont_kind < 0) return fail(PASTA_ERR_ARG, "star has no continuum model");
    std::lock_guard<std::mutex> lk(s->mu);
    Batch* B;
    CHK(get_cached_batch(s, P, &B));
    std::vector<int> st(P);
    CHK(batch_eval(*B, labels, chi2, st.data()));
    if (params) {
        const int np = std::max(1, s->n_params);
        if (s->cont_kind == CONT_FIXED) {
            for (int i = 0; i < P * np; i++) params[i] = 0.0f;
        } else {
            HIPCHK(hipMemcpy(params, B->d_pf.p, sizeof(float) * (size_t)P * np, hipMemcpyDeviceToHost));
        }
    }
    int bad = 0;
    for (int i = 0; i < P; i++) {
        if (status) status[i] = st[i];
        if (st[i] && !bad) bad = st[i];
    }
    if (bad) return fail(PASTA_ERR_MODEL, status_text(bad));
    return PASTA_OK;
}

int pasta_eval_chi2_multi(pasta_star* const* stars, int n, const double* labels, double* chi2,
                          int* status) {
    if (!stars || n <= 0 || !labels || !chi2) return fail(PASTA_ERR_ARG, "bad arguments");
    pasta_grid* g = stars[0]->grid;
    for (int k = 0; k < n; k++) {
        if (!stars[k] || stars[k]->grid != g) return fail(PASTA_ERR_ARG, "stars must share one grid");
        if (stars[k]->cont_kind < 0) return fail(PASTA_ERR_ARG, "star has no continuum model");
    }
    std::lock_guard<std::mutex> lk(g->mu);
    Batch B;
    CHK(batch_build(B, g, std::vector<pasta_star*>(stars, stars + n), std::vector<int>(n, 1), 0));
    std::vector<int> st(n);
    CHK(batch_eval(B, labels, chi2, st.data()));
    int bad = 0;
    for (int k = 0; k < n; k++) {
        if (status) status[k] = st[k];
        if (st[k] && !bad) bad = st[k];
    }
    if (bad) return fail(PASTA_ERR_MODEL, status_text(bad));
    return PASTA_OK;
}

static int model_only(pasta_grid* g, const pasta_disp* d, const double* labels5, int P,
                      float* out, int* status, int mode) {
    if (!g || !labels5 || P <= 0 || !out) return fail(PASTA_ERR_ARG, "bad arguments");
    pasta_disp none;
    if (!d) {  // synthetic-grid output without instrumental broadening
        none.kind = DISP_NONE;
        none.W = g->W;
        none.wl = {std::pow(10.0, g->first), std::pow(10.0, g->first + g->step * (g->W - 1))};
        none.fidx = {0.0, (double)(g->W - 1)};
        d = &none;
    }
    std::unique_ptr<pasta_star> s;
    int n = mode == 0 ? (int)d->wl.size() : g->W;
    CHK(star_build(g, d, nullptr, nullptr, nullptr, n, s));
    if (mode == 1) s->dev.n_tiles = (g->W + TILE - 1) / TILE;
    std::lock_guard<std::mutex> lk(g->mu);
    Batch B;
    CHK(batch_build(B, g, {s.get()}, {P}, mode));
    std::vector<int> st(P);
    CHK(batch_eval(B, labels5, nullptr, st.data()));
    HIPCHK(hipMemcpy(out, B.d_model.p, sizeof(float) * (size_t)P * n, hipMemcpyDeviceToHost));
    int bad = 0;
    for (int i = 0; i < P; i++) {
        if (status) status[i] = st[i];
        if (st[i]) {
            if (!bad) bad = st[i];
            for (int k = 0; k < n; k++) out[(size_t)i * n + k] = NAN;
        }
    }
    if (bad) return fail(PASTA_ERR_MODEL, status_text(bad));
    return PASTA_OK;
}

int pasta_produce_model(pasta_grid* g, const pasta_disp* d, const double* labels, int P, float* out,
                        int* status) {
    if (!d) return fail(PASTA_ERR_ARG, "null dispersion");
    return model_only(g, d, labels, P, out, status, 0);
}

int pasta_interpolate(pasta_grid* g, const double* labels, int P, float* out, int* status) {
    if (!labels || P <= 0) return fail(PASTA_ERR_ARG, "bad arguments");
    std::vector<double> l5((size_t)P * 5, 0.0);
    for (int i = 0; i < P; i++)
        for (int d = 0; d < 3; d++) l5[5 * i + d] = labels[3 * i + d];
    return model_only(g, nullptr, l5.data(), P, out, status, 1);
}

int pasta_interpolate_and_convolve(pasta_grid* g, const pasta_disp* d, const double* labels, int P,
                                   float* out, int* status) {
    if (!labels || P <= 0 || !d) return fail(PASTA_ERR_ARG, "bad arguments");
    std::vector<double> l5((size_t)P * 5, 0.0);
    for (int i = 0; i < P; i++)
        for (int k = 0; k < 4; k++) l5[5 * i + k] = labels[4 * i + k];
    pasta_disp dd = *d;  // synthetic-grid output ignores the observed sampling
    return model_only(g, &dd, l5.data(), P, out, status, 1);
}

// taps of build_rotation_kernel (src/convolve_rv.rs:438-449); 0 for the pass-through of
// vsini == 0 (:509-511)
static int rotation_taps(double vsini, double step) {
    if (!(vsini > 0.0)) return 0;
    const double eps = 0.6, dvelo = LN_10 * step, vrot = vsini / C_KMS;
    const double a = 4.0 * (eps - 1.0) / (M_PI * eps);
    const double v = std::ceil(2.0 * vrot / dvelo * std::sqrt(1.0 + a * a));
    long long nme = (v >= 0.0 && v < 1e9) ? (long long)v + 2 : 2;
    return (int)(nme + 1 - nme % 2);
}

int pasta_vsini_kernel(const pasta_grid* g, double vsini, float* out, int cap, int* len) {
    if (!g) return fail(PASTA_ERR_ARG, "null grid");
    // build_rotation_kernel (src/convolve_rv.rs:438-475)
    const double eps = 0.6, dvelo = LN_10 * g->step, vrot = vsini / C_KMS;
    double a = 4.0 * (eps - 1.0) / (M_PI * eps);
    double v = std::ceil(2.0 * vrot / dvelo * std::sqrt(1.0 + a * a));
    long long nme = (v >= 0.0 && v < 1e9) ? (long long)v + 2 : 2;
    if (nme < 1) nme = 1;
    int n = (int)(nme + 1 - nme % 2);
    if (len) *len = n;
    if (!out) return PASTA_OK;
    if (cap < n) return fail(PASTA_ERR_ARG, "buffer too small");
    double mid = ((double)(n - 1) * dvelo) / 2.0;
    float s = 0.0f;
    for (int i = 0; i < n; i++) {
        double vk = (double)i * dvelo - mid;
        double q = vk / vrot;
        double y = 1.0 - q * q;
        float k = 0.0f;
        if (y > 0.0)
            k = (float)((2.0 * (1.0 - eps) * std::sqrt(y) + M_PI * eps / 2.0 * y) /
                        (M_PI * vrot * (1.0 - eps / 3.0)));
        out[i] = k;
        s = s + k;
    }
    for (int i = 0; i < n; i++) out[i] = out[i] / s;
    return PASTA_OK;
}

// ---------------------------------------------------------------------------
// PSO set-up shared by the plain and batched fits
// ---------------------------------------------------------------------------
int settings_check(const pasta_fit_settings* st) {
    if (!st || st->num_particles < 1 || st->max_iters < 0)
        return fail(PASTA_ERR_ARG, "bad PSO settings");
    if (st->social_factor < 0.0) return fail(PASTA_ERR_ARG, "social factor must be >=0.");
    if (st->delta < 0.0) return fail(PASTA_ERR_ARG, "delta must be >=0.");
    return PASTA_OK;
}

static Bounds make_bounds(const pasta_fit_settings* st) {
    Bounds b{};
    b.vsini_lo = st->vsini_range[0]; b.vsini_hi = st->vsini_range[1];
    b.rv_lo = st->rv_range[0]; b.rv_hi = st->rv_range[1];
    for (int d = 0; d < 5; d++) {
        b.ckind[d] = st->constraint_kind[d];
        b.ca[d] = st->constraint_a[d];
        b.cb[d] = st->constraint_b[d];
    }
    return b;
}

// initialize_positions_and_velocities with the rejection sampler
// (src/particleswarm.rs:137-150, src/bounds.rs:21-35)
static int pso_init_positions(const GridGeom& g, const Bounds& b, int P, uint64_t seed, double* pos,
                              double* vel) {
    double lo[5], hi[5];
    bounds_outer(g, b, lo, hi);
    const uint64_t kp = rng_u64(seed, KEY_POS), kv = rng_u64(seed, KEY_VEL);
    uint64_t attempt = 0;
    int got = 0;
    while (got < P) {
        double p[5];
        for (int d = 0; d < 5; d++) p[d] = rand_range(lo[d], hi[d], rng_uniform(kp, attempt * 5 + d));
        attempt++;
        if (attempt > 100000000ULL)
            return fail(PASTA_ERR_FIT, "cannot draw particles inside the bounds");
        if (bounds_within(g, b, p)) {
            std::memcpy(pos + 5 * got, p, sizeof(p));
            got++;
        }
    }
    for (int q = 0; q < P; q++)
        for (int d = 0; d < 5; d++) {
            double delta = hi[d] - lo[d];
            vel[5 * q + d] = rand_range(-delta, delta, rng_uniform(kv, (uint64_t)q * 5 + d));
        }
    return PASTA_OK;
}

static int final_params(pasta_star* s, const double best[5], pasta_fit_result* out) {
    // one more evaluation at the best labels for the continuum parameters (src/fitting.rs:436-440)
    Batch B;
    CHK(batch_build(B, s->grid, {s}, {1}, 0));
    double c;
    int st;
    CHK(batch_eval(B, best, &c, &st));
    if (st) return fail(PASTA_ERR_FIT, status_text(st));
    int np = std::min(512, std::max(1, s->n_params));
    out->n_params = np;
    if (s->cont_kind == CONT_FIXED) {
        for (int k = 0; k < np; k++) out->continuum_params[k] = 0.0f;
    } else {
        HIPCHK(hipMemcpy(out->continuum_params, B.d_pf.p, sizeof(float) * np, hipMemcpyDeviceToHost));
    }
    return PASTA_OK;
}

// ---------------------------------------------------------------------------
// plain fit: host PSO, one device round trip per iteration
// ---------------------------------------------------------------------------
static int fit_impl(pasta_star* s, const pasta_fit_settings* st, uint64_t seed,
                    const double* init_pos, pasta_fit_result* out, double* trace,
                    double* population) {
    if (!s || !out) return fail(PASTA_ERR_ARG, "null argument");
    if (s->cont_kind < 0) return fail(PASTA_ERR_ARG, "star has no continuum model");
    CHK(settings_check(st));
    std::memset(out, 0, sizeof(*out));
    auto t_start = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    const GridGeom& g = s->grid->geom;
    const Bounds b = make_bounds(st);
    const PsoParams ps{st->inertia_factor, st->cognitive_factor, st->social_factor};
    const int P = st->num_particles;
    std::vector<double> pos(5 * P), vel(5 * P), bpos(5 * P), cost(P), bcost(P);
    std::vector<int> sts(P);
    CHK(pso_init_positions(g, b, P, seed, pos.data(), vel.data()));
    if (init_pos) std::memcpy(pos.data(), init_pos, sizeof(double) * 5 * P);
    Batch* B;
    CHK(get_cached_batch(s, P, &B));
    auto evaluate = [&]() -> int {
        CHK(batch_eval(*B, pos.data(), cost.data(), sts.data()));
        for (int q = 0; q < P; q++)
            if (sts[q]) {
                out->status = sts[q];
                return fail(PASTA_ERR_FIT, std::string("Error computing cost: ") + status_text(sts[q]));
            }
        return PASTA_OK;
    };
    CHK(evaluate());
    // sort by cost, stable (src/particleswarm.rs:127-131)
    std::vector<int> order(P);
    for (int q = 0; q < P; q++) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] < cost[c]; });
    {
        std::vector<double> p2(5 * P), v2(5 * P), c2(P);
        for (int q = 0; q < P; q++) {
            std::memcpy(&p2[5 * q], &pos[5 * order[q]], 40);
            std::memcpy(&v2[5 * q], &vel[5 * order[q]], 40);
            c2[q] = cost[order[q]];
        }
        pos.swap(p2); vel.swap(v2); cost.swap(c2);
    }
    bpos = pos;
    bcost = cost;
    double gb[5], gc = cost[0];
    std::memcpy(gb, pos.data(), sizeof(gb));
    if (trace) { trace[0] = gc; std::memcpy(trace + 1, gb, sizeof(gb)); }
    const uint64_t ki = rng_u64(seed, KEY_ITER);
    for (int it = 1; it <= st->max_iters; it++) {
        for (int q = 0; q < P; q++) {
            int rc = pso_update_particle(g, b, ps, ki, it, P, q, &pos[5 * q], &vel[5 * q],
                                         &bpos[5 * q], gb);
            if (rc) {
                out->status = rc;
                return fail(PASTA_ERR_FIT, "Particle is not within bounds");
            }
        }
        CHK(evaluate());
        for (int q = 0; q < P; q++) {
            if (cost[q] < bcost[q]) {
                std::memcpy(&bpos[5 * q], &pos[5 * q], 40);
                bcost[q] = cost[q];
                if (cost[q] < gc) {
                    std::memcpy(gb, &pos[5 * q], 40);
                    gc = cost[q];
                }
            }
        }
        if (trace) { trace[6 * it] = gc; std::memcpy(trace + 6 * it + 1, gb, sizeof(gb)); }
        if (population)  // the population after iteration it: (position, cost) per particle
            for (int q = 0; q < P; q++) {
                double* o = population + ((size_t)(it - 1) * P + q) * 6;
                std::memcpy(o, &pos[5 * q], 40);
                o[5] = cost[q];
            }
    }
    std::memcpy(out->label, gb, sizeof(gb));
    out->chi2 = gc;
    out->iterations = st->max_iters;
    CHK(final_params(s, gb, out));
    out->time_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return PASTA_OK;
}

int pasta_fit(pasta_star* s, const pasta_fit_settings* st, uint64_t seed, const double* init_pos,
              pasta_fit_result* out, double* trace) {
    return fit_impl(s, st, seed, init_pos, out, trace, nullptr);
}

int pasta_fit_observed(pasta_star* s, const pasta_fit_settings* st, uint64_t seed,
                       const double* init_pos, pasta_fit_result* out, double* trace,
                       double* population) {
    if (!population) return fail(PASTA_ERR_ARG, "null population buffer");
    return fit_impl(s, st, seed, init_pos, out, trace, population);
}

// ---------------------------------------------------------------------------
// batched fit: device PSO, graph-captured iterations
// ---------------------------------------------------------------------------
struct PsoBatch {
    Batch B;
    std::vector<DevPso> h_pso;
    std::vector<double> h_pos, h_vel;  // initial positions / velocities of every evaluation
};

static int pso_batch_setup(PsoBatch& PB, pasta_grid* g, const std::vector<pasta_star*>& stars,
                           const std::vector<const pasta_fit_settings*>& sts,
                           const std::vector<uint64_t>& seeds, bool with_trace, int* max_it) {
    Batch& B = PB.B;
    std::vector<int> counts;
    for (auto* st : sts) counts.push_back(st->num_particles);
    CHK(batch_build(B, g, stars, counts, 0));
    const int ns = (int)stars.size();
    std::vector<double>& pos = PB.h_pos;
    std::vector<double>& vel = PB.h_vel;
    pos.assign((size_t)B.n_evals * 5, 0.0);
    vel.assign((size_t)B.n_evals * 5, 0.0);
    PB.h_pso.assign(ns, DevPso{});
    *max_it = 0;
    B.max_vsini = 0.0;
    for (int s = 0; s < ns; s++) {
        const pasta_fit_settings* st = sts[s];
        DevPso& P = PB.h_pso[s];
        P.b = make_bounds(st);
        double lo[5], hi[5];
        bounds_outer(g->geom, P.b, lo, hi);  // the PSO keeps its particles inside these
        B.max_vsini = std::max(B.max_vsini, std::isnan(hi[3]) ? HUGE_VAL : hi[3]);
        P.s = PsoParams{st->inertia_factor, st->cognitive_factor, st->social_factor};
        P.key_iter = rng_u64(seeds[s], KEY_ITER);
        P.P = st->num_particles;
        P.first_eval = B.first_eval[s];
        P.iter = 0;
        P.max_iters = st->max_iters;
        *max_it = std::max(*max_it, st->max_iters);
        CHK(pso_init_positions(g->geom, P.b, P.P, seeds[s], &pos[5 * (size_t)P.first_eval],
                               &vel[5 * (size_t)P.first_eval]));
    }
    CHK(B.d_labels.upload(pos));
    CHK(B.d_vel.upload(vel));
    CHK(B.d_bpos.alloc(pos.size()));
    CHK(B.d_bcost.alloc(B.n_evals));
    CHK(B.d_pso.upload(PB.h_pso));
    if (with_trace) CHK(B.d_trace.alloc((size_t)ns * (*max_it + 1) * 6));
    return PASTA_OK;
}

static int pso_first_eval(PsoBatch& PB) {
    Batch& B = PB.B;
    CHK(batch_launch(B));
    HIPCHK(launch_pso_init(B.d_pso.p, B.n_stars, B.d_evals.p, B.d_labels.p, B.d_vel.p, B.d_bpos.p,
                           B.d_bcost.p, B.d_chi2.p, B.d_trace.p, B.stream));
    return PASTA_OK;
}

static int pso_iteration_launches(PsoBatch& PB) {
    Batch& B = PB.B;
    if (B.eager_events) HIPCHK(hipEventRecord(B.ev[2], B.stream));
    HIPCHK(launch_pso_update(B.grid->d_geom.p, B.d_pso.p, B.d_evals.p, B.n_evals, B.d_labels.p,
                             B.d_vel.p, B.d_bpos.p, B.stream));
    CHK(batch_launch(B));
    HIPCHK(launch_pso_best(B.d_pso.p, B.n_stars, B.d_evals.p, B.d_labels.p, B.d_bpos.p,
                           B.d_bcost.p, B.d_chi2.p, B.d_trace.p, B.stream));
    if (B.eager_events) HIPCHK(hipEventRecord(B.ev[3], B.stream));
    return PASTA_OK;
}

static int pso_capture(PsoBatch& PB) {
    Batch& B = PB.B;
    hipGraph_t gr;
    int rc;
    hipError_t e;
    {
        std::unique_lock<std::shared_mutex> lk(capture_mutex());
        HIPCHK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
        rc = pso_iteration_launches(PB);
        e = hipStreamEndCapture(B.stream, &gr);
    }
    if (rc) return rc;
    HIPCHK(e);
    HIPCHK(hipGraphInstantiate(&B.graph, gr, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(gr));
    return PASTA_OK;
}

int pasta_fit_bulk(pasta_star* const* stars, int n_stars, const pasta_fit_settings* st,
                   const uint64_t* seeds, pasta_fit_result* out, double* trace) {
    if (!stars || n_stars <= 0 || !seeds || !out) return fail(PASTA_ERR_ARG, "bad arguments");
    CHK(settings_check(st));
    auto t_start = std::chrono::steady_clock::now();
    pasta_grid* g = stars[0]->grid;
    std::vector<pasta_star*> sv(stars, stars + n_stars);
    for (auto* s : sv) {
        if (!s || s->grid != g) return fail(PASTA_ERR_ARG, "stars must share one grid");
        if (s->cont_kind < 0) return fail(PASTA_ERR_ARG, "star has no continuum model");
    }
    std::lock_guard<std::mutex> lk(g->mu);
    PsoBatch PB;
    int max_it = 0;
    std::vector<const pasta_fit_settings*> sts(n_stars, st);
    CHK(pso_batch_setup(PB, g, sv, sts, std::vector<uint64_t>(seeds, seeds + n_stars),
                        trace != nullptr, &max_it));
    CHK(pso_first_eval(PB));
    if (max_it > 0) {
        CHK(pso_capture(PB));
        for (int it = 0; it < max_it; it++) HIPCHK(hipGraphLaunch(PB.B.graph, PB.B.stream));
    }
    HIPCHK(hipMemcpyAsync(PB.h_pso.data(), PB.B.d_pso.p, sizeof(DevPso) * n_stars,
                          hipMemcpyDeviceToHost, PB.B.stream));
    if (trace)
        HIPCHK(hipMemcpyAsync(trace, PB.B.d_trace.p, sizeof(double) * n_stars * (max_it + 1) * 6,
                              hipMemcpyDeviceToHost, PB.B.stream));
    HIPCHK(hipStreamSynchronize(PB.B.stream));
    double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    int first_bad = 0;
    for (int s = 0; s < n_stars; s++) {
        pasta_fit_result& r = out[s];
        std::memset(&r, 0, sizeof(r));
        const DevPso& P = PB.h_pso[s];
        r.status = P.status;
        r.iterations = st->max_iters;
        r.time_s = elapsed;
        std::memcpy(r.label, P.gbest, sizeof(r.label));
        r.chi2 = P.gcost;
        if (P.status) {
            if (!first_bad) first_bad = P.status;
            continue;
        }
    }
    // continuum parameters at the best labels: one evaluation per star
    {
        Batch F;
        CHK(batch_build(F, g, sv, std::vector<int>(n_stars, 1), 0));
        std::vector<double> lab((size_t)n_stars * 5);
        for (int s = 0; s < n_stars; s++) std::memcpy(&lab[5 * s], out[s].label, 40);
        std::vector<double> c(n_stars);
        std::vector<int> fs(n_stars);
        CHK(batch_eval(F, lab.data(), c.data(), fs.data()));
        std::vector<float> pf((size_t)std::max<int64_t>(F.pf_len, 1));
        HIPCHK(hipMemcpy(pf.data(), F.d_pf.p, sizeof(float) * pf.size(), hipMemcpyDeviceToHost));
        for (int s = 0; s < n_stars; s++) {
            if (out[s].status) continue;
            int np = std::min(512, std::max(1, sv[s]->n_params));
            out[s].n_params = np;
            for (int k = 0; k < np; k++)
                out[s].continuum_params[k] =
                    sv[s]->cont_kind == CONT_FIXED ? 0.0f : pf[F.h_evals[s].pf_off + k];
        }
    }
    if (first_bad) return fail(PASTA_ERR_FIT, std::string("a fit failed: ") + status_text(first_bad));
    return PASTA_OK;
}

// ---------------------------------------------------------------------------
// compute_uncertainty (src/fitting.rs:333-372,482-568): Brent root searches of
// chi2(label + x * radius) - target on both sides of every label, all searches of
// all stars advanced together, one batched chi2 evaluation per round.
// ---------------------------------------------------------------------------

// argmin 0.10.0's BrentRoot (Brent / zbrent): state machine fed one function value
// at a time.  best = iterate with the smallest |f| (argmin's best parameter).
struct BrentSearch {
    double a = 0, b = 0, c = 0, d = NAN, e = NAN, fa = 0, fb = 0, fc = 0, tol = 1e-3;
    double best = 0, best_cost = 0, x = 0;
    int state = 0;  // 0: wants f(a), 1: wants f(b), 2: wants f(x), 3: done, 4: failed
    int err = 0;
    void start(double lo, double hi, double t) {
        a = lo; b = hi; c = hi; d = e = NAN; tol = t; state = 0; x = a;
    }
    void fail_with(int code) { state = 4; err = code; }
    void feed(double f) {
        if (state == 0) { fa = f; state = 1; x = b; return; }
        if (state == 1) {
            fb = f;
            if (fa * fb > 0.0) { fail_with(PASTA_ST_SIGN); return; }
            fc = fb;
            best = b; best_cost = std::fabs(fb);
            state = 2;
            plan();
            return;
        }
        fb = f;
        if (std::fabs(fb) < best_cost) { best = b; best_cost = std::fabs(fb); }
        plan();
    }
    // the part of an iteration before its evaluation: converge, or choose the next b
    void plan() {
        const double eps = 2.220446049250313e-16;
        if ((fb > 0.0 && fc > 0.0) || (fb < 0.0 && fc < 0.0)) {
            c = a; fc = fa; d = b - a; e = d;
        }
        if (std::fabs(fc) < std::fabs(fb)) {
            a = b; b = c; c = a;
            fa = fb; fb = fc; fc = fa;
        }
        const double tol1 = 2.0 * eps * std::fabs(b) + 0.5 * tol;
        const double xm = 0.5 * (c - b);
        if (std::fabs(xm) <= tol1 || fb == 0.0) {
            if (std::fabs(fb) < best_cost) { best = b; best_cost = std::fabs(fb); }
            state = 3;
            return;
        }
        if (std::fabs(e) >= tol1 && std::fabs(fa) > std::fabs(fb)) {
            const double sg = fb / fa;
            double p, q;
            if (a == c) {
                p = 2.0 * xm * sg;
                q = 1.0 - sg;
            } else {
                const double qq = fa / fc, r = fb / fc;
                p = sg * (2.0 * xm * qq * (qq - r) - (b - a) * (r - 1.0));
                q = (qq - 1.0) * (r - 1.0) * (sg - 1.0);
            }
            if (p > 0.0) q = -q;
            p = std::fabs(p);
            const double min1 = 3.0 * xm * q - std::fabs(tol1 * q), min2 = std::fabs(e * q);
            if (2.0 * p < std::min(min1, min2)) { e = d; d = p / q; }
            else { d = xm; e = d; }
        } else {
            d = xm; e = d;
        }
        a = b;
        fa = fb;
        b += std::fabs(d) > tol1 ? d : (xm > 0.0 ? tol1 : -tol1);
        x = b;
    }
    bool wants() const { return state <= 2; }
};

static int uncertainty_impl(pasta_star* const* stars, int n, double spec_res, const double* labels,
                            const double* search_radius, double* out, int* status,
                            std::vector<int>& bst) {
    if (!stars || n <= 0 || !labels || !out || !status) return fail(PASTA_ERR_ARG, "bad arguments");
    if (!(spec_res > 0.0)) return fail(PASTA_ERR_ARG, "spec_res must be positive");
    pasta_grid* g = stars[0]->grid;
    for (int k = 0; k < n; k++) {
        if (!stars[k] || stars[k]->grid != g) return fail(PASTA_ERR_ARG, "stars must share one grid");
        if (stars[k]->cont_kind < 0) return fail(PASTA_ERR_ARG, "star has no continuum model");
    }
    static const double def_radius[5] = {2000.0, 0.3, 0.3, 40.0, 40.0};  // src/lib.rs:1554
    const double* r0 = search_radius ? search_radius : def_radius;
    std::lock_guard<std::mutex> lk(g->mu);
    Batch B;
    // best chi2 of every star at its label
    std::vector<pasta_star*> all(stars, stars + n);
    CHK(batch_build(B, g, all, std::vector<int>(n, 1), 0));
    std::vector<double> best(n);
    bst.assign(n, 0);
    CHK(batch_eval(B, labels, best.data(), bst.data()));
    // SingleBounds over the grid, vsini (0, 1e4), RV (-1e3, 1e3) (src/fitting.rs:519)
    Bounds bounds{};
    bounds.vsini_lo = 0.0; bounds.vsini_hi = 1e4; bounds.rv_lo = -1e3; bounds.rv_hi = 1e3;
    std::vector<BrentSearch> bs((size_t)n * 10);
    std::vector<double> target(n), radius((size_t)n * 5);
    for (int k = 0; k < n; k++) {
        const double* lab = labels + 5 * (size_t)k;
        double* sr = &radius[5 * (size_t)k];
        for (int i = 0; i < 5; i++) sr[i] = r0[i];
        sr[0] *= lab[0];
        sr[3] = std::max(lab[3] * sr[3], 100.0);
        const DevStar& D = stars[k]->dev;
        const double nn = 4.0 * spec_res * (D.wlN - D.wl0) / (D.wl0 + D.wlN);
        target[k] = best[k] * (1.0 + std::sqrt(2.0 / nn));
        for (int i = 0; i < 5; i++)
            for (int right = 0; right < 2; right++) {
                BrentSearch& S = bs[10 * (size_t)k + 2 * i + right];
                if (bst[k]) { S.fail_with(bst[k]); continue; }
                double bl[5], bound;
                std::memcpy(bl, lab, sizeof(bl));
                bl[i] += right ? sr[i] : -sr[i];
                if (bounds_clamp(g->geom, bounds, bl, i, &bound)) { S.fail_with(PASTA_ST_CLAMP); continue; }
                const double resc = (bound - lab[i]) / sr[i];
                if (right) S.start(0.0, std::min(resc, 1.0), 1e-3);
                else S.start(resc, 0.0, 1e-3);
            }
    }
    // rounds: every search that wants a value gets it from one batched evaluation
    for (int round = 0; round < 10000; round++) {
        std::vector<pasta_star*> rs;
        std::vector<int> cnt;
        std::vector<double> lab;
        std::vector<size_t> who;
        for (int k = 0; k < n; k++) {
            int c = 0;
            for (int j = 0; j < 10; j++) {
                const BrentSearch& S = bs[10 * (size_t)k + j];
                if (!S.wants()) continue;
                const int i = j / 2;
                double l[5];
                std::memcpy(l, labels + 5 * (size_t)k, sizeof(l));
                l[i] += S.x * radius[5 * (size_t)k + i];
                lab.insert(lab.end(), l, l + 5);
                who.push_back(10 * (size_t)k + j);
                c++;
            }
            if (c) { rs.push_back(stars[k]); cnt.push_back(c); }
        }
        if (who.empty()) break;
        CHK(batch_build(B, g, rs, cnt, 0));
        std::vector<double> c2(who.size());
        std::vector<int> st(who.size());
        CHK(batch_eval(B, lab.data(), c2.data(), st.data()));
        for (size_t q = 0; q < who.size(); q++) {
            BrentSearch& S = bs[who[q]];
            const int k = (int)(who[q] / 10);
            if (st[q]) S.fail_with(st[q]);
            else S.feed(c2[q] - target[k]);
        }
    }
    for (int k = 0; k < n; k++)
        for (int j = 0; j < 10; j++) {
            const BrentSearch& S = bs[10 * (size_t)k + j];
            const double sr = radius[5 * (size_t)k + j / 2];
            const bool ok = S.state == 3;
            status[10 * (size_t)k + j] = ok ? 0 : (S.state == 4 ? S.err : PASTA_ST_SIGN);
            out[10 * (size_t)k + j] = ok ? ((j & 1) ? S.best * sr : -S.best * sr) : NAN;
        }
    return PASTA_OK;
}

int pasta_compute_uncertainty_bulk(pasta_star* const* stars, int n, double spec_res,
                                   const double* labels, const double* search_radius,
                                   double* out, int* status) {
    std::vector<int> bst;
    return uncertainty_impl(stars, n, spec_res, labels, search_radius, out, status, bst);
}

int pasta_compute_uncertainty(pasta_star* s, double spec_res, const double label[5],
                              const double search_radius[5], double out[10], int status[10]) {
    if (!s || !label) return fail(PASTA_ERR_ARG, "bad arguments");
    pasta_star* const arr[1] = {s};
    std::vector<int> bst;
    CHK(uncertainty_impl(arr, 1, spec_res, label, search_radius, out, status, bst));
    if (bst[0]) return fail(PASTA_ERR_MODEL, status_text(bst[0]));  // the reference raises
    return PASTA_OK;
}

// ---------------------------------------------------------------------------
// server: continuous batching of fit requests over the resident grid
// ---------------------------------------------------------------------------
#include <deque>
#include <map>

struct Request {
    uint64_t id = 0;
    pasta_star* star = nullptr;  // owned; freed when the request finishes or is cancelled
    pasta_fit_settings st{};
    uint64_t seed = 0;
    int state = 0;  // 0 waiting, 1 running, 2 finished, 3 cancelled
    int64_t bytes = 0;  // HBM while in flight: evaluation buffers, PSO state, the star itself
    std::chrono::steady_clock::time_point t_submit;
    pasta_fit_result res{};
    bool started = false;  // its PSO state lives in the current batch, star slot `slot`
    int slot = -1;
};

// Two batches alternate: a rebuild fills the idle one and moves the PSO state of the stars
// that stay in flight from the current one on the device (k_pso_remap), so admission and
// retirement never copy running stars' state through the host.
struct pasta_server {
    pasta_grid* g = nullptr;
    pasta_server_config cfg{};
    int64_t budget = 0, in_use = 0;
    uint64_t next_id = 1;
    std::deque<Request*> waiting;
    std::vector<Request*> running;     // in flight (admitted, not finished or cancelled)
    std::vector<Request*> batch_reqs;  // the requests the current batch holds, by star slot
    std::map<uint64_t, std::unique_ptr<Request>> all;
    std::unique_ptr<PsoBatch> pb[2];   // kept across rebuilds: buffers keep their capacity
    int cur = 0;                       // pb[cur] is the current batch
    DBuf<int> d_src_eval, d_src_star;  // remap tables of the last rebuild
    Batch fin;                         // evaluation of finished stars' best labels (reused)
    bool active = false;               // pb[cur] holds a captured iteration of batch_reqs
    bool dirty = false;                // running differs from batch_reqs
    int64_t evals_done = 0, finished = 0, iters_run = 0, rebuilds = 0;
    double busy_ms = 0;                // device time of first evaluations and iterations
    double rebuild_ms = 0;             // host time in batch rebuilds (build, remap, capture)
    hipEvent_t ev_busy[4] = {nullptr, nullptr, nullptr, nullptr};
    bool first_eval_timed = false;     // ev_busy[2..3] bracket a first evaluation in flight
    double t_model_last = 0, t_model_sum = 0, t_iter_last = 0, t_iter_sum = 0, t_n = 0;
    double t_flops_sum = 0, t_bytes_sum = 0;  // algorithmic work of the timed k_model launches
    std::vector<double> lab_host;
    std::mutex mu;
    ~pasta_server() {
        pb[0].reset();
        pb[1].reset();
        for (auto& e : ev_busy)
            if (e) (void)hipEventDestroy(e);
        for (auto& kv : all) delete kv.second->star;
    }
};

static void request_release_star(Request* r) {
    delete r->star;
    r->star = nullptr;
}

int pasta_server_create(pasta_grid* g, const pasta_server_config* cfg, pasta_server** out) {
    if (!g || !cfg || !out || cfg->max_stars < 1 || cfg->max_particles < 1 || cfg->iters_per_step < 1)
        return fail(PASTA_ERR_ARG, "bad server configuration");
    auto sv = std::make_unique<pasta_server>();
    sv->g = g;
    sv->cfg = *cfg;
    if (cfg->hbm_budget_bytes > 0) {
        sv->budget = cfg->hbm_budget_bytes;
    } else {
        size_t f = 0, t = 0;
        HIPCHK(hipMemGetInfo(&f, &t));
        sv->budget = std::max<int64_t>((int64_t)f - (4LL << 30), 1LL << 28);
    }
    for (auto& e : sv->ev_busy) HIPCHK(hipEventCreate(&e));
    *out = sv.release();
    return PASTA_OK;
}

int pasta_server_destroy(pasta_server* sv) {
    delete sv;
    return PASTA_OK;
}

int pasta_server_submit(pasta_server* sv, pasta_star* s, const pasta_fit_settings* st, uint64_t seed,
                        uint64_t* id) {
    if (!sv || !s || !id) return fail(PASTA_ERR_ARG, "null argument");
    if (s->grid != sv->g) return fail(PASTA_ERR_ARG, "star belongs to another grid");
    if (s->cont_kind < 0) return fail(PASTA_ERR_ARG, "star has no continuum model");
    CHK(settings_check(st));
    if (st->num_particles > sv->cfg.max_particles)
        return fail(PASTA_ERR_ARG, "more particles than the server's max_particles");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto r = std::make_unique<Request>();
    r->id = sv->next_id++;
    r->st = *st;
    r->seed = seed;
    r->bytes = star_eval_bytes(s) * st->num_particles + (int64_t)sizeof(double) * 16 * st->num_particles +
               (int64_t)sizeof(DevPso) + s->resident_bytes;
    if (r->bytes > sv->budget) return fail(PASTA_ERR_NOMEM, "request larger than the HBM budget");
    r->star = s;  // owned from here on
    r->t_submit = std::chrono::steady_clock::now();
    *id = r->id;
    sv->waiting.push_back(r.get());
    sv->all[r->id] = std::move(r);
    return PASTA_OK;
}

// Build the idle batch for the running set.  Stars already in the current batch resume
// there from their device-resident PSO state (moved by k_pso_remap); new stars get their
// initial population evaluated and sorted.
static int server_rebuild(pasta_server* sv) {
    PsoBatch* old = sv->active ? sv->pb[sv->cur].get() : nullptr;
    const int nxt = old ? 1 - sv->cur : sv->cur;
    sv->active = false;
    sv->batch_reqs.clear();
    if (sv->running.empty()) return PASTA_OK;
    if (!sv->pb[nxt]) sv->pb[nxt] = std::make_unique<PsoBatch>();
    PsoBatch& PB = *sv->pb[nxt];
    Batch& B = PB.B;
    if (B.graph) {
        HIPCHK(hipGraphExecDestroy(B.graph));
        B.graph = nullptr;
    }
    std::vector<pasta_star*> stars;
    std::vector<const pasta_fit_settings*> sts;
    std::vector<uint64_t> seeds;
    for (Request* r : sv->running) {
        stars.push_back(r->star);
        sts.push_back(&r->st);
        seeds.push_back(r->seed);
    }
    int max_it = 0;
    CHK(pso_batch_setup(PB, sv->g, stars, sts, seeds, false, &max_it));
    std::vector<int> src_eval(B.n_evals, -1), src_star(B.n_stars, -1);
    bool any_new = false, any_started = false;
    for (int s = 0; s < B.n_stars; s++) {
        Request* r = sv->running[s];
        const int f0 = B.first_eval[s], P = B.count[s];
        if (r->started) {
            if (!old || r->slot < 0 || r->slot >= old->B.n_stars)
                return fail(PASTA_ERR_STATE, "server lost a running star's state");
            src_star[s] = r->slot;
            for (int q = 0; q < P; q++) src_eval[f0 + q] = old->B.first_eval[r->slot] + q;
            any_started = true;
        } else {
            any_new = true;
        }
    }
    if (any_started) {
        // the old batch's work has finished: every step ends with a stream synchronisation
        CHK(sv->d_src_eval.upload(src_eval));
        CHK(sv->d_src_star.upload(src_star));
        const Batch& O = old->B;
        HIPCHK(launch_pso_remap(sv->d_src_eval.p, B.n_evals, B.d_evals.p, sv->d_src_star.p,
                                B.n_stars, O.d_pso.p, B.d_pso.p, O.d_labels.p, B.d_labels.p,
                                O.d_vel.p, B.d_vel.p, O.d_bpos.p, B.d_bpos.p, O.d_bcost.p,
                                B.d_bcost.p, B.stream));
    }
    if (any_new) {
        HIPCHK(hipEventRecord(sv->ev_busy[2], B.stream));
        CHK(pso_first_eval(PB));
        HIPCHK(hipEventRecord(sv->ev_busy[3], B.stream));
        sv->first_eval_timed = true;
    }
    for (auto& e : B.ev)
        if (!e) HIPCHK(hipEventCreate(&e));
    CHK(pso_capture(PB));
    for (int s = 0; s < B.n_stars; s++) {
        sv->running[s]->slot = s;
        sv->running[s]->started = true;
    }
    sv->batch_reqs = sv->running;
    sv->cur = nxt;
    sv->active = true;
    sv->rebuilds++;
    return PASTA_OK;
}

// Finished stars (all iterations done, or failed) leave the batch: their continuum
// parameters come from one more evaluation at the best labels, then their stars are freed.
static int server_retire(pasta_server* sv) {
    Batch& B = sv->pb[sv->cur]->B;
    std::vector<DevPso> ps(B.n_stars);
    HIPCHK(hipMemcpy(ps.data(), B.d_pso.p, sizeof(DevPso) * B.n_stars, hipMemcpyDeviceToHost));
    std::vector<int> done_idx;
    for (int s = 0; s < B.n_stars; s++)
        if (ps[s].done || ps[s].status) done_idx.push_back(s);
    if (done_idx.empty()) return PASTA_OK;
    std::vector<pasta_star*> fs;
    std::vector<double> lab;
    for (int s : done_idx) {
        fs.push_back(sv->batch_reqs[s]->star);
        for (int d = 0; d < 5; d++) lab.push_back(ps[s].gbest[d]);
    }
    Batch& F = sv->fin;
    CHK(batch_build(F, sv->g, fs, std::vector<int>(fs.size(), 1), 0));
    std::vector<double> c(fs.size());
    std::vector<int> st(fs.size());
    CHK(batch_eval(F, lab.data(), c.data(), st.data()));
    std::vector<float> pf((size_t)std::max<int64_t>(F.pf_len, 1));
    HIPCHK(hipMemcpy(pf.data(), F.d_pf.p, sizeof(float) * pf.size(), hipMemcpyDeviceToHost));
    auto now = std::chrono::steady_clock::now();
    for (int k = 0; k < (int)done_idx.size(); k++) {
        const int s = done_idx[k];
        Request* r = sv->batch_reqs[s];
        pasta_fit_result& res = r->res;
        std::memset(&res, 0, sizeof(res));
        res.status = ps[s].status;
        res.iterations = r->st.max_iters;
        res.chi2 = ps[s].gcost;
        std::memcpy(res.label, ps[s].gbest, 40);
        res.time_s = std::chrono::duration<double>(now - r->t_submit).count();
        const int np = std::min(512, std::max(1, r->star->n_params));
        res.n_params = np;
        for (int q = 0; q < np; q++)
            res.continuum_params[q] =
                r->star->cont_kind == CONT_FIXED ? 0.0f : pf[F.h_evals[k].pf_off + q];
        r->state = 2;
        sv->in_use -= r->bytes;
        sv->finished++;
        // neither the current batch nor `fin` runs again before the next rebuild
        request_release_star(r);
    }
    std::vector<Request*> keep;
    for (Request* r : sv->running)
        if (r->state == 1) keep.push_back(r);
    sv->running.swap(keep);
    sv->dirty = true;
    return PASTA_OK;
}

// Algorithmic work of one k_model launch over the batch's current labels, counted as the
// reference's algorithm does it for each evaluation (synthetic span W, observed pixels N):
// interpolation 64 products and sums per pixel (128 W flops, 64 u16 reads = 128 W bytes),
// the rotation and instrumental "same" convolutions with their own tap counts
// (2 (K_rot + K_instr) W flops, the rotation taps from the evaluation's vsini), the
// resample (6 N flops, an 8-byte float index read per observed pixel).
static int server_algorithmic_work(pasta_server* sv, Batch& B, double* flops, double* bytes) {
    sv->lab_host.resize((size_t)B.n_evals * 5);
    HIPCHK(hipMemcpy(sv->lab_host.data(), B.d_labels.p, sizeof(double) * sv->lab_host.size(),
                     hipMemcpyDeviceToHost));
    double f = 0.0, b = 0.0;
    for (int s = 0; s < B.n_stars; s++) {
        const pasta_star* st = B.stars[s];
        const double W = st->span, N = st->N;
        const int kg = st->dev.disp_kind == DISP_NONE ? 0 : st->dev.Kd;
        for (int q = 0; q < B.count[s]; q++) {
            const double vsini = sv->lab_host[5 * (size_t)(B.first_eval[s] + q) + 3];
            const int kr = rotation_taps(vsini, sv->g->step);
            f += 128.0 * W + 2.0 * (double)(kr + kg) * W + 6.0 * N;
            b += 128.0 * W + 8.0 * N;
        }
    }
    *flops = f;
    *bytes = b;
    return PASTA_OK;
}

int pasta_server_step(pasta_server* sv, int* active, int* waiting) {
    if (!sv) return fail(PASTA_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    std::lock_guard<std::mutex> lg(sv->g->mu);
    // admission: free slots and HBM budget
    while (!sv->waiting.empty() && (int)sv->running.size() < sv->cfg.max_stars) {
        Request* r = sv->waiting.front();
        if (r->state == 3) { sv->waiting.pop_front(); continue; }
        if (sv->in_use + r->bytes > sv->budget && !sv->running.empty()) break;
        sv->waiting.pop_front();
        r->state = 1;
        sv->in_use += r->bytes;
        sv->running.push_back(r);
        sv->dirty = true;
    }
    if (sv->dirty) {
        const auto t0 = std::chrono::steady_clock::now();
        CHK(server_rebuild(sv));
        sv->rebuild_ms +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        sv->dirty = false;
    }
    if (sv->active) {
        Batch& B = sv->pb[sv->cur]->B;
        HIPCHK(hipEventRecord(sv->ev_busy[0], B.stream));
        for (int it = 0; it < sv->cfg.iters_per_step; it++) {
            if (!sv->cfg.timing) {
                HIPCHK(hipGraphLaunch(B.graph, B.stream));
            } else {  // the same launches without the graph, bracketed by HIP events
                B.eager_events = true;
                int rc = pso_iteration_launches(*sv->pb[sv->cur]);
                B.eager_events = false;
                CHK(rc);
                HIPCHK(hipStreamSynchronize(B.stream));
                float m = 0.0f, t = 0.0f;
                HIPCHK(hipEventElapsedTime(&m, B.ev[0], B.ev[1]));
                HIPCHK(hipEventElapsedTime(&t, B.ev[2], B.ev[3]));
                sv->t_model_last = m;
                sv->t_model_sum += m;
                sv->t_iter_last = t;
                sv->t_iter_sum += t;
                sv->t_n += 1;
                double fl = 0.0, by = 0.0;
                CHK(server_algorithmic_work(sv, B, &fl, &by));
                sv->t_flops_sum += fl;
                sv->t_bytes_sum += by;
            }
        }
        HIPCHK(hipEventRecord(sv->ev_busy[1], B.stream));
        HIPCHK(hipStreamSynchronize(B.stream));
        float busy = 0.0f;
        HIPCHK(hipEventElapsedTime(&busy, sv->ev_busy[0], sv->ev_busy[1]));
        if (sv->first_eval_timed) {
            float fe = 0.0f;
            HIPCHK(hipEventElapsedTime(&fe, sv->ev_busy[2], sv->ev_busy[3]));
            busy += fe;
            sv->first_eval_timed = false;
        }
        sv->busy_ms += busy;
        sv->iters_run += sv->cfg.iters_per_step;
        sv->evals_done += (int64_t)B.n_evals * sv->cfg.iters_per_step;
        CHK(server_retire(sv));
        if (sv->dirty && sv->running.empty()) {
            sv->active = false;
            sv->dirty = false;
            sv->batch_reqs.clear();
        }
    }
    if (active) *active = (int)sv->running.size();
    if (waiting) *waiting = (int)sv->waiting.size();
    return PASTA_OK;
}

int pasta_server_poll(pasta_server* sv, uint64_t id, int* done, pasta_fit_result* out) {
    if (!sv || !done) return fail(PASTA_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->all.find(id);
    if (it == sv->all.end()) return fail(PASTA_ERR_STATE, "unknown request");
    Request* r = it->second.get();
    if (r->state == 3) return fail(PASTA_ERR_STATE, "request was cancelled");
    *done = r->state == 2;
    if (*done && out) *out = r->res;
    return PASTA_OK;
}

int pasta_server_cancel(pasta_server* sv, uint64_t id) {
    if (!sv) return fail(PASTA_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->all.find(id);
    if (it == sv->all.end()) return fail(PASTA_ERR_STATE, "unknown request");
    Request* r = it->second.get();
    if (r->state == 0) {  // still queued: admission skips it
        r->state = 3;
        request_release_star(r);
        return PASTA_OK;
    }
    if (r->state == 1) {
        // the star leaves the running set; the current batch, which still holds it, is
        // rebuilt before the next launch, so its buffers can go now
        sv->running.erase(std::find(sv->running.begin(), sv->running.end(), r));
        sv->in_use -= r->bytes;
        r->state = 3;
        request_release_star(r);
        sv->dirty = true;
        return PASTA_OK;
    }
    if (r->state == 3) return fail(PASTA_ERR_STATE, "request was cancelled");
    return fail(PASTA_ERR_STATE, "request already finished");
}

int pasta_server_stats(const pasta_server* sv, int64_t out[7]) {
    if (!sv || !out) return fail(PASTA_ERR_ARG, "null argument");
    out[6] = (int64_t)(sv->rebuild_ms * 1e3);
    out[0] = sv->evals_done;
    out[1] = sv->finished;
    out[2] = sv->iters_run;
    out[3] = sv->in_use;
    out[4] = (int64_t)(sv->busy_ms * 1e3);
    out[5] = sv->rebuilds;
    return PASTA_OK;
}

int pasta_server_timing(pasta_server* sv, int reset, double out[7]) {
    if (!sv || !out) return fail(PASTA_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    out[0] = sv->t_model_last;
    out[1] = sv->t_model_sum;
    out[2] = sv->t_iter_last;
    out[3] = sv->t_iter_sum;
    out[4] = sv->t_n;
    out[5] = sv->t_flops_sum;
    out[6] = sv->t_bytes_sum;
    if (reset) sv->t_model_sum = sv->t_iter_sum = sv->t_n = sv->t_flops_sum = sv->t_bytes_sum = 0;
    return PASTA_OK;
}

int pasta_server_set_timing(pasta_server* sv, int on) {
    if (!sv) return fail(PASTA_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->cfg.timing = on ? 1 : 0;
    return PASTA_OK;
}
