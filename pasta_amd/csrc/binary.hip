// binary.hip -- binary-star model, chi2 and fit behind the C ABI (include/pasta_hip.h).
//
// Restated from the reference (SimonVanSchuylenbergh/PASTA):
//   interpolate_linear                         src/interpolate.rs:985-1030
//   produce_binary_components / _model_norm    src/interpolate.rs:822-890
//   shift_resample_and_add_binary_components   src/convolve_rv.rs:606-690
//   BinaryCostFunction / BinaryFitter          src/fitting.rs:571-728
//   RVCostFunction / BinaryRVFitter            src/fitting.rs:730-880
//   BinaryTimeseriesKnownRVCostFunction /
//     BinaryTimeriesKnownRVFitter              src/fitting.rs:882-1078
//   BinaryBounds, BinaryBoundsWithoutRV,
//     BinaryRVBounds                           src/bounds.rs:188-372
// One evaluation of P binary labels: the 2P component models on the synthetic grid
// (k_prep ... k_model in synthetic-grid mode) and, on a second stream beside them, the 2P
// linear continua of the continuum grid and their means (k_interp_lin, k_seq_mean); then
// the combination at the observed pixels (k_binary_combine) and the continuum fit and chi2
// of the P models (k_moments ... k_final), with one copy back at the end.
#include "host.h"

namespace pasta {
hipError_t launch_cont(const DevStar* stars, DevEval* evals, int n_evals, const int4* bmom,
                       const int4* bchi2, const ContClass* cls, int n_cls,
                       const int2* solve_items, int n_solve, const float* model, double* mpart,
                       float* pf, double* cpart, double* chi2, int* status, hipStream_t st);
hipError_t launch_binary_cont(const DevGrid& CG, const double* clabels, int ncomp, int* cstat,
                              float* cont, float* cmean, hipStream_t st);
hipError_t launch_binary(const DevGrid& G, const DevStar* stars, DevEval* oevals,
                         int n_obs_evals, const double* labels11, const DevEval* cevals, int ncomp,
                         const float* comp, const int* cstat, const float* cont,
                         const float* cmean, int n_obs_max, float* model, hipStream_t st);
hipError_t launch_binary_rv(const DevGrid& G, const DevStar* stars, DevEval* oevals, int n,
                            const double* rv, const float* model1, const float* model2,
                            const float* cont1, const float* cont2, float light_ratio, int n_obs,
                            float* kc, int* kst, float* kmean, float* model, hipStream_t st);
hipError_t launch_interp_lin(const DevGrid& G, const double* labels, int lstride, int n,
                             float* out, int* status, hipStream_t st);
}  // namespace pasta

// A binary observed in S spectra on one wavelength grid.  D = 11: one spectrum, the RVs
// are labels (BinaryFitter); D = 9: S spectra with known RVs (the time-series fitter).
struct pasta_binary {
    pasta_grid* g = nullptr;
    pasta_grid* cg = nullptr;
    int D = 11, S = 1;
    std::vector<double> rvs;  // S x 2 (D = 9)
    std::vector<std::unique_ptr<pasta_star>> obs;  // observed pixels (+ continuum, flux, var)
    std::unique_ptr<pasta_star> synth;  // the same dispersion on the synthetic grid
    bool has_chi2 = false;
    int N = 0, n_params = 0, cont_kind = -1;
    int P = -1;  // batch size the buffers below are built for
    std::unique_ptr<Batch> Bc, Bo;  // components (2P, synthetic grid), observed (S x P)
    DBuf<double> d_lab11;
    DBuf<float> d_cont, d_cmean;
    DBuf<int> d_cstat;
    std::vector<double> h_comp, h_lab11;
    // the continua and their (sequential) means run on a second stream beside the component
    // models: fork after the label upload, join before the combination
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    std::mutex mu;
    ~pasta_binary() {
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (side) (void)hipStreamDestroy(side);
    }
};

static int binary_prepare(pasta_binary* b, int P) {
    if (b->P == P) return PASTA_OK;
    b->P = -1;
    b->Bc.reset(new Batch());
    b->Bo.reset(new Batch());
    CHK(batch_build(*b->Bc, b->g, {b->synth.get()}, {2 * P}, 1));
    std::vector<pasta_star*> os;
    for (auto& o : b->obs) os.push_back(o.get());
    CHK(batch_build(*b->Bo, b->g, os, std::vector<int>(b->S, P), 0));
    CHK(b->d_lab11.alloc((size_t)11 * P * b->S));
    CHK(b->d_cont.alloc((size_t)2 * P * b->g->W));
    CHK(b->d_cmean.alloc((size_t)2 * P));
    CHK(b->d_cstat.alloc((size_t)2 * P));
    b->h_comp.assign((size_t)10 * P, 0.0);
    b->h_lab11.assign((size_t)11 * P * b->S, 0.0);
    if (!b->side) HIPCHK(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
    if (!b->fork) HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    if (!b->join) HIPCHK(hipEventCreateWithFlags(&b->join, hipEventDisableTiming));
    b->P = P;
    return PASTA_OK;
}

// the whole evaluation of P label vectors (D values each), enqueued on the observed
// batch's stream; the observed records are S x P, spectrum-major
static int binary_launch(pasta_binary* b, const double* labels, int P, bool chi2) {
    CHK(binary_prepare(b, P));
    Batch& Bc = *b->Bc;
    Batch& Bo = *b->Bo;
    hipStream_t st = Bo.stream;
    const int D = b->D, per = D == 11 ? 5 : 4;
    for (int q = 0; q < P; q++) {
        const double* l = labels + (size_t)D * q;
        for (int s = 0; s < 2; s++) {
            double* c = &b->h_comp[(size_t)5 * (2 * q + s)];
            for (int k = 0; k < 4; k++) c[k] = l[per * s + k];
            c[4] = 0.0;  // components stay on the synthetic grid (interpolate_and_convolve)
        }
        for (int k = 0; k < b->S; k++) {
            double* o = &b->h_lab11[(size_t)11 * (k * P + q)];
            if (D == 11) {
                std::memcpy(o, l, sizeof(double) * 11);
            } else {
                for (int s = 0; s < 2; s++) {
                    for (int j = 0; j < 4; j++) o[5 * s + j] = l[4 * s + j];
                    o[5 * s + 4] = b->rvs[2 * k + s];
                }
                o[10] = l[8];
            }
        }
    }
    HIPCHK(hipMemcpyAsync(Bc.d_labels.p, b->h_comp.data(), sizeof(double) * 10 * P,
                          hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->d_lab11.p, b->h_lab11.data(), sizeof(double) * 11 * P * b->S,
                          hipMemcpyHostToDevice, st));
    DevGrid G = b->g->dev(), CG = b->cg->dev();
    HIPCHK(hipEventRecord(b->fork, st));
    HIPCHK(hipStreamWaitEvent(b->side, b->fork, 0));
    HIPCHK(launch_binary_cont(CG, Bc.d_labels.p, P, b->d_cstat.p, b->d_cont.p, b->d_cmean.p,
                              b->side));
    HIPCHK(hipEventRecord(b->join, b->side));
    HIPCHK(launch_eval(G, Bc.d_stars.p, Bc.d_evals.p, Bc.n_evals, Bc.d_labels.p, Bc.d_rot.p,
                       Bc.d_tbound.p, Bc.tb_stride, Bc.d_hist.p, Bc.d_perm.p, Bc.d_pair_eq.p,
                       Bc.d_items.p, Bc.d_nitems.p, Bc.d_union.p, Bc.n_bm, Bc.d_bmom.p, Bc.d_bchi2.p,
                       Bc.cls.data(), 0, Bc.d_solve.p, 0, Bc.d_model.p, Bc.d_mpart.p, Bc.d_pf.p,
                       Bc.d_cpart.p, Bc.d_chi2.p, Bc.d_status.p, 1,
                       may_fft(b->g, HUGE_VAL, Bc.stars), nullptr, nullptr, st));
    HIPCHK(hipStreamWaitEvent(st, b->join, 0));
    HIPCHK(launch_binary(G, Bo.d_stars.p, Bo.d_evals.p, P * b->S, b->d_lab11.p, Bc.d_evals.p, P,
                         Bc.d_model.p, b->d_cstat.p, b->d_cont.p, b->d_cmean.p, b->N,
                         Bo.d_model.p, st));
    if (chi2)
        HIPCHK(launch_cont(Bo.d_stars.p, Bo.d_evals.p, Bo.n_evals, Bo.d_bmom.p, Bo.d_bchi2.p,
                           Bo.cls.data(), (int)Bo.cls.size(), Bo.d_solve.p, Bo.n_solve,
                           Bo.d_model.p, Bo.d_mpart.p, Bo.d_pf.p, Bo.d_cpart.p, Bo.d_chi2.p,
                           Bo.d_status.p, st));
    return PASTA_OK;
}

// status of each observed record (written by k_binary_combine)
static int binary_status(pasta_binary* b, int n, int* status) {
    Batch& Bo = *b->Bo;
    std::vector<DevEval> ev(n);
    HIPCHK(hipMemcpyAsync(ev.data(), Bo.d_evals.p, sizeof(DevEval) * n, hipMemcpyDeviceToHost,
                          Bo.stream));
    HIPCHK(hipStreamSynchronize(Bo.stream));
    for (int q = 0; q < n; q++) status[q] = ev[q].status;
    return PASTA_OK;
}

static int binary_common(pasta_grid* g, pasta_grid* cg, const pasta_disp* d, int n) {
    if (!g || !cg || !d) return fail(PASTA_ERR_ARG, "null argument");
    if ((int)d->wl.size() != n)
        return fail(PASTA_ERR_ARG, "Observed spectrum and target dispersion must have the same length");
    if (d->W != g->W || d->step != g->step || d->first != g->first)
        return fail(PASTA_ERR_ARG, "dispersion was built for another synthetic grid");
    if (cg->W != g->W || cg->step != g->step || cg->first != g->first)
        return fail(PASTA_ERR_ARG, "the continuum grid has other synthetic wavelengths");
    return PASTA_OK;
}

static int binary_new(pasta_grid* g, pasta_grid* cg, const pasta_disp* d, const pasta_cont* c,
                      const float* fluxes, const float* vars, int S, int n, int D,
                      std::unique_ptr<pasta_binary>& out) {
    CHK(binary_common(g, cg, d, n));
    const bool chi2 = c && fluxes && vars;
    if (chi2 && c->n != n)
        return fail(PASTA_ERR_ARG, "Length of observed and synthetic spectrum are different");
    auto b = std::make_unique<pasta_binary>();
    b->g = g;
    b->cg = cg;
    b->D = D;
    b->S = S;
    b->N = n;
    b->has_chi2 = chi2;
    b->obs.resize(S);
    for (int k = 0; k < S; k++)
        CHK(star_build(g, d, chi2 ? c : nullptr, chi2 ? fluxes + (size_t)k * n : nullptr,
                       chi2 ? vars + (size_t)k * n : nullptr, n, b->obs[k]));
    CHK(star_build(g, d, nullptr, nullptr, nullptr, g->W, b->synth));
    b->synth->dev.n_tiles = (g->W + TILE - 1) / TILE;
    b->n_params = chi2 ? c->n_params : 0;
    b->cont_kind = chi2 ? c->kind : -1;
    out = std::move(b);
    return PASTA_OK;
}

int pasta_binary_create(pasta_grid* g, pasta_grid* cg, const pasta_disp* d, const pasta_cont* c,
                        const float* flux, const float* var, int n, pasta_binary** out) {
    if (!out) return fail(PASTA_ERR_ARG, "null argument");
    std::unique_ptr<pasta_binary> b;
    CHK(binary_new(g, cg, d, c, flux, var, 1, n, 11, b));
    *out = b.release();
    return PASTA_OK;
}

int pasta_binary_ts_create(pasta_grid* g, pasta_grid* cg, const pasta_disp* d, const pasta_cont* c,
                           const float* fluxes, const float* vars, const double* rvs, int S, int n,
                           pasta_binary** out) {
    if (!out || !c || !fluxes || !vars || !rvs || S < 1) return fail(PASTA_ERR_ARG, "bad arguments");
    std::unique_ptr<pasta_binary> b;
    CHK(binary_new(g, cg, d, c, fluxes, vars, S, n, 9, b));
    b->rvs.assign(rvs, rvs + 2 * S);
    *out = b.release();
    return PASTA_OK;
}

int pasta_binary_destroy(pasta_binary* b) {
    delete b;
    return PASTA_OK;
}

int pasta_binary_model(pasta_binary* b, const double* labels, int P, float* out, int* status) {
    if (!b || !labels || P <= 0 || !out) return fail(PASTA_ERR_ARG, "bad arguments");
    if (b->D != 11) return fail(PASTA_ERR_ARG, "binary models take 11 labels");
    std::lock_guard<std::mutex> lk(b->mu);
    std::lock_guard<std::mutex> lg(b->g->mu);
    CHK(binary_launch(b, labels, P, false));
    std::vector<int> st(P);
    CHK(binary_status(b, P, st.data()));
    const std::vector<EvalInit>& he = b->Bo->h_evals;
    int bad = 0;
    for (int q = 0; q < P; q++) {
        if (status) status[q] = st[q];
        float* o = out + (size_t)q * b->N;
        if (st[q]) {
            if (!bad) bad = st[q];
            for (int i = 0; i < b->N; i++) o[i] = NAN;
        } else {
            HIPCHK(hipMemcpy(o, b->Bo->d_model.p + he[q].m_off, sizeof(float) * b->N,
                             hipMemcpyDeviceToHost));
        }
    }
    if (bad) return fail(PASTA_ERR_MODEL, status_text(bad));
    return PASTA_OK;
}

// chi2 of the S x P observed records; per label, the spectrum-ordered sum (the reference's
// Iterator::sum over Result, which stops at the first failing spectrum); params (nullable)
// S x P x n_params
static int binary_chi2(pasta_binary* b, const double* labels, int P, double* chi2, int* status,
                       double* per_spectrum, float* params) {
    if (!b || !labels || P <= 0 || !chi2) return fail(PASTA_ERR_ARG, "bad arguments");
    if (!b->has_chi2) return fail(PASTA_ERR_ARG, "binary built without an observed spectrum");
    std::lock_guard<std::mutex> lk(b->mu);
    std::lock_guard<std::mutex> lg(b->g->mu);
    CHK(binary_launch(b, labels, P, true));
    Batch& Bo = *b->Bo;
    const int n = P * b->S;
    std::vector<int> st(n);
    std::vector<double> c(n);
    HIPCHK(hipMemcpyAsync(c.data(), Bo.d_chi2.p, sizeof(double) * n, hipMemcpyDeviceToHost, Bo.stream));
    HIPCHK(hipMemcpyAsync(st.data(), Bo.d_status.p, sizeof(int) * n, hipMemcpyDeviceToHost,
                          Bo.stream));
    HIPCHK(hipStreamSynchronize(Bo.stream));
    int bad = 0;
    for (int q = 0; q < P; q++) {
        double tot = 0.0;
        int s = 0;
        for (int k = 0; k < b->S && !s; k++) {
            s = st[(size_t)k * P + q];
            tot += c[(size_t)k * P + q];
        }
        if (status) status[q] = s;
        chi2[q] = s ? NAN : tot;
        if (s && !bad) bad = s;
    }
    if (per_spectrum)
        for (int e = 0; e < n; e++) per_spectrum[e] = st[e] ? NAN : c[e];
    const int np = b->n_params;
    if (params && np > 0) {
        // every record has the same parameter count, so the parameters are one block
        if (b->cont_kind != CONT_FIXED)
            HIPCHK(hipMemcpy(params, Bo.d_pf.p, sizeof(float) * (size_t)n * np,
                             hipMemcpyDeviceToHost));
        for (int e = 0; e < n; e++)
            if (st[e] || b->cont_kind == CONT_FIXED)
                for (int k = 0; k < np; k++) params[(size_t)e * np + k] = st[e] ? NAN : 0.0f;
    }
    if (bad) return fail(PASTA_ERR_MODEL, status_text(bad));
    return PASTA_OK;
}

int pasta_binary_eval_chi2(pasta_binary* b, const double* labels, int P, double* chi2, int* status,
                           float* params) {
    if (b && b->D != 11) return fail(PASTA_ERR_ARG, "binary chi2 takes 11 labels");
    return binary_chi2(b, labels, P, chi2, status, nullptr, params);
}

int pasta_binary_ts_eval_chi2(pasta_binary* b, const double* labels, int P, double* chi2,
                              int* status) {
    if (b && b->D != 9) return fail(PASTA_ERR_ARG, "time-series chi2 takes 9 labels");
    return binary_chi2(b, labels, P, chi2, status, nullptr, nullptr);
}

int pasta_interpolate_linear(pasta_grid* g, const double* labels, int P, float* out, int* status) {
    if (!g || !labels || P <= 0 || !out) return fail(PASTA_ERR_ARG, "bad arguments");
    std::lock_guard<std::mutex> lg(g->mu);
    DBuf<double> dl;
    DBuf<float> dout;
    DBuf<int> dst;
    CHK(dl.upload(labels, (size_t)3 * P));
    CHK(dout.alloc((size_t)P * g->W));
    CHK(dst.alloc(P));
    HIPCHK(launch_interp_lin(g->dev(), dl.p, 3, P, dout.p, dst.p, nullptr));
    std::vector<int> st(P);
    HIPCHK(hipMemcpy(st.data(), dst.p, sizeof(int) * P, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(out, dout.p, sizeof(float) * P * g->W, hipMemcpyDeviceToHost));
    int bad = 0;
    for (int q = 0; q < P; q++) {
        if (status) status[q] = st[q];
        if (st[q]) {
            if (!bad) bad = st[q];
            for (int i = 0; i < g->W; i++) out[(size_t)q * g->W + i] = NAN;
        }
    }
    if (bad) return fail(PASTA_ERR_MODEL, status_text(bad));
    return PASTA_OK;
}

// ---------------------------------------------------------------------------
// BinaryRVFitter: fixed components, the two RVs fitted (src/fitting.rs:730-880)
// ---------------------------------------------------------------------------
struct pasta_binary_rv {
    pasta_grid* g = nullptr;
    std::unique_ptr<pasta_star> obs;
    DBuf<float> d_comp;  // model1, model2, continuum1, continuum2 (W each)
    float light_ratio = 0.0f;
    int N = 0, P = -1;
    std::unique_ptr<Batch> Bo;
    DBuf<double> d_rv;
    DBuf<float> d_kc, d_kmean;
    DBuf<int> d_kst;
    std::mutex mu;
};

int pasta_binary_rv_create(pasta_grid* g, const pasta_disp* d, const pasta_cont* c,
                           const float* flux, const float* var, int n, const float* model1,
                           const float* model2, const float* continuum1, const float* continuum2,
                           double light_ratio, pasta_binary_rv** out) {
    if (!g || !d || !c || !flux || !var || !model1 || !model2 || !continuum1 || !continuum2 || !out)
        return fail(PASTA_ERR_ARG, "null argument");
    if ((int)d->wl.size() != n)
        return fail(PASTA_ERR_ARG, "Observed spectrum and target dispersion must have the same length");
    if (d->W != g->W || d->step != g->step || d->first != g->first)
        return fail(PASTA_ERR_ARG, "dispersion was built for another synthetic grid");
    if (c->n != n) return fail(PASTA_ERR_ARG, "Length of observed and synthetic spectrum are different");
    auto r = std::make_unique<pasta_binary_rv>();
    r->g = g;
    r->N = n;
    r->light_ratio = (float)light_ratio;  // `self.light_ratio as FluxFloat` (fitting.rs:767-768)
    CHK(star_build(g, d, c, flux, var, n, r->obs));
    const size_t W = g->W;
    std::vector<float> comp(4 * W);
    std::memcpy(&comp[0], model1, sizeof(float) * W);
    std::memcpy(&comp[W], model2, sizeof(float) * W);
    std::memcpy(&comp[2 * W], continuum1, sizeof(float) * W);
    std::memcpy(&comp[3 * W], continuum2, sizeof(float) * W);
    CHK(r->d_comp.upload(comp));
    *out = r.release();
    return PASTA_OK;
}

int pasta_binary_rv_destroy(pasta_binary_rv* r) {
    delete r;
    return PASTA_OK;
}

int pasta_binary_rv_eval_chi2(pasta_binary_rv* r, const double* rvs, int P, double* chi2,
                              int* status) {
    if (!r || !rvs || P <= 0 || !chi2) return fail(PASTA_ERR_ARG, "bad arguments");
    std::lock_guard<std::mutex> lk(r->mu);
    if (r->P != P) {
        r->P = -1;
        r->Bo.reset(new Batch());
        CHK(batch_build(*r->Bo, r->g, {r->obs.get()}, {P}, 0));
        CHK(r->d_rv.alloc((size_t)2 * P));
        CHK(r->d_kc.alloc((size_t)2 * P * r->N));
        CHK(r->d_kmean.alloc((size_t)2 * P));
        CHK(r->d_kst.alloc((size_t)2 * P));
        r->P = P;
    }
    Batch& Bo = *r->Bo;
    hipStream_t st = Bo.stream;
    HIPCHK(hipMemcpyAsync(r->d_rv.p, rvs, sizeof(double) * 2 * P, hipMemcpyHostToDevice, st));
    const size_t W = r->g->W;
    const float* cp = r->d_comp.p;
    HIPCHK(launch_binary_rv(r->g->dev(), Bo.d_stars.p, Bo.d_evals.p, P, r->d_rv.p, cp, cp + W,
                            cp + 2 * W, cp + 3 * W, r->light_ratio, r->N, r->d_kc.p, r->d_kst.p,
                            r->d_kmean.p, Bo.d_model.p, st));
    HIPCHK(launch_cont(Bo.d_stars.p, Bo.d_evals.p, Bo.n_evals, Bo.d_bmom.p, Bo.d_bchi2.p,
                       Bo.cls.data(), (int)Bo.cls.size(), Bo.d_solve.p, Bo.n_solve, Bo.d_model.p,
                       Bo.d_mpart.p, Bo.d_pf.p, Bo.d_cpart.p, Bo.d_chi2.p, Bo.d_status.p, st));
    std::vector<int> sts(P);
    HIPCHK(hipMemcpyAsync(chi2, Bo.d_chi2.p, sizeof(double) * P, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(sts.data(), Bo.d_status.p, sizeof(int) * P, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    int bad = 0;
    for (int q = 0; q < P; q++) {
        if (status) status[q] = sts[q];
        if (sts[q]) {
            chi2[q] = NAN;
            if (!bad) bad = sts[q];
        }
    }
    if (bad) return fail(PASTA_ERR_MODEL, status_text(bad));
    return PASTA_OK;
}

// ---------------------------------------------------------------------------
// the PSO of src/particleswarm.rs over D parameters with one batched evaluation per
// iteration (BinaryFitter::fit, BinaryRVFitter::fit, BinaryTimeriesKnownRVFitter::fit)
// ---------------------------------------------------------------------------
static int binary_settings_check(const pasta_binary_settings* st) {
    if (!st || st->num_particles < 1 || st->max_iters < 0) return fail(PASTA_ERR_ARG, "bad PSO settings");
    if (st->social_factor < 0.0) return fail(PASTA_ERR_ARG, "social factor must be >=0.");
    if (st->delta < 0.0) return fail(PASTA_ERR_ARG, "delta must be >=0.");
    return PASTA_OK;
}

static BoundsN binary_settings_bounds(const pasta_binary_settings* st, int D) {
    BoundsN bn{};
    if (D == 2) {  // BinaryRVBounds (bounds.rs:347-372)
        bn.D = 2;
        for (int d = 0; d < 2; d++) {
            bn.kind[d] = BK_RANGE;
            bn.lo[d] = st->rv_range[0];
            bn.hi[d] = st->rv_range[1];
        }
    } else {
        bn = binary_bounds(D == 11, st->vsini_range[0], st->vsini_range[1], st->rv_range[0],
                           st->rv_range[1], st->light_ratio_range[0], st->light_ratio_range[1]);
    }
    for (int d = 0; d < D; d++) {
        bn.ckind[d] = st->constraint_kind[d];
        bn.ca[d] = st->constraint_a[d];
        bn.cb[d] = st->constraint_b[d];
    }
    return bn;
}

// initialize_positions_and_velocities with the rejection sampler for D parameters
// (src/particleswarm.rs:137-150, src/bounds.rs:21-35)
static int pso_init_positions_n(const GridGeom& g, const BoundsN& b, int P, uint64_t seed,
                                double* pos, double* vel) {
    const int D = b.D;
    double lo[PASTA_MAX_D], hi[PASTA_MAX_D];
    boundsn_outer(g, b, lo, hi);
    const uint64_t kp = rng_u64(seed, KEY_POS), kv = rng_u64(seed, KEY_VEL);
    uint64_t attempt = 0;
    for (int got = 0; got < P;) {
        double p[PASTA_MAX_D];
        for (int d = 0; d < D; d++) p[d] = rand_range(lo[d], hi[d], rng_uniform(kp, attempt * D + d));
        attempt++;
        if (attempt > 100000000ULL) return fail(PASTA_ERR_FIT, "cannot draw particles inside the bounds");
        if (boundsn_within(g, b, p)) {
            std::memcpy(pos + (size_t)D * got, p, sizeof(double) * D);
            got++;
        }
    }
    for (int q = 0; q < P; q++)
        for (int d = 0; d < D; d++) {
            const double delta = hi[d] - lo[d];
            vel[(size_t)D * q + d] = rand_range(-delta, delta, rng_uniform(kv, (uint64_t)q * D + d));
        }
    return PASTA_OK;
}

// eval(pos P x D, cost P, status P) -> PASTA_OK / PASTA_ERR_MODEL / other error
template <typename Eval>
static int pso_fit_n(const GridGeom& g, const BoundsN& bn, const pasta_binary_settings* st,
                     uint64_t seed, const double* init_pos, Eval eval, double* gbest,
                     double* gcost, double* trace, int* status) {
    const PsoParams ps{st->inertia_factor, st->cognitive_factor, st->social_factor};
    const int P = st->num_particles, D = bn.D;
    std::vector<double> pos((size_t)D * P), vel((size_t)D * P), bpos, cost(P), bcost;
    std::vector<int> sts(P);
    CHK(pso_init_positions_n(g, bn, P, seed, pos.data(), vel.data()));
    if (init_pos) std::memcpy(pos.data(), init_pos, sizeof(double) * D * P);
    auto evaluate = [&]() -> int {
        const int rc = eval(pos.data(), cost.data(), sts.data());
        if (rc == PASTA_ERR_MODEL)
            for (int q = 0; q < P; q++)
                if (sts[q]) {
                    *status = sts[q];
                    return fail(PASTA_ERR_FIT, std::string("Error computing cost: ") + status_text(sts[q]));
                }
        return rc;
    };
    CHK(evaluate());
    std::vector<int> order(P);
    for (int q = 0; q < P; q++) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return cost[a] < cost[c]; });
    {
        std::vector<double> p2(pos.size()), v2(vel.size()), c2(P);
        for (int q = 0; q < P; q++) {
            std::memcpy(&p2[(size_t)D * q], &pos[(size_t)D * order[q]], sizeof(double) * D);
            std::memcpy(&v2[(size_t)D * q], &vel[(size_t)D * order[q]], sizeof(double) * D);
            c2[q] = cost[order[q]];
        }
        pos.swap(p2); vel.swap(v2); cost.swap(c2);
    }
    bpos = pos;
    bcost = cost;
    double gc = cost[0];
    std::memcpy(gbest, pos.data(), sizeof(double) * D);
    if (trace) { trace[0] = gc; std::memcpy(trace + 1, gbest, sizeof(double) * D); }
    const uint64_t ki = rng_u64(seed, KEY_ITER);
    for (int it = 1; it <= st->max_iters; it++) {
        for (int q = 0; q < P; q++) {
            const int rc = pso_update_particle_n(g, bn, ps, ki, it, P, q, &pos[(size_t)D * q],
                                                 &vel[(size_t)D * q], &bpos[(size_t)D * q], gbest);
            if (rc) {
                *status = rc;
                return fail(PASTA_ERR_FIT, "Particle is not within bounds");
            }
        }
        CHK(evaluate());
        for (int q = 0; q < P; q++)
            if (cost[q] < bcost[q]) {
                std::memcpy(&bpos[(size_t)D * q], &pos[(size_t)D * q], sizeof(double) * D);
                bcost[q] = cost[q];
                if (cost[q] < gc) {
                    std::memcpy(gbest, &pos[(size_t)D * q], sizeof(double) * D);
                    gc = cost[q];
                }
            }
        if (trace) {
            trace[(size_t)(D + 1) * it] = gc;
            std::memcpy(trace + (size_t)(D + 1) * it + 1, gbest, sizeof(double) * D);
        }
    }
    *gcost = gc;
    return PASTA_OK;
}

int pasta_binary_fit(pasta_binary* b, const pasta_binary_settings* st, uint64_t seed,
                     const double* init_pos, pasta_binary_result* out, double* trace) {
    if (!b || !out) return fail(PASTA_ERR_ARG, "null argument");
    if (b->D != 11) return fail(PASTA_ERR_ARG, "binary fits take 11 labels");
    if (!b->has_chi2) return fail(PASTA_ERR_ARG, "binary built without an observed spectrum");
    CHK(binary_settings_check(st));
    std::memset(out, 0, sizeof(*out));
    const auto t0 = std::chrono::steady_clock::now();
    const int P = st->num_particles;
    double gb[11], gc;
    CHK(pso_fit_n(b->g->geom, binary_settings_bounds(st, 11), st, seed, init_pos,
                  [&](const double* pos, double* cost, int* sts) {
                      return pasta_binary_eval_chi2(b, pos, P, cost, sts, nullptr);
                  },
                  gb, &gc, trace, &out->status));
    std::memcpy(out->label1, gb, sizeof(out->label1));
    std::memcpy(out->label2, gb + 5, sizeof(out->label2));
    out->light_ratio = gb[10];
    out->chi2 = gc;
    out->iterations = st->max_iters;
    // continuum parameters of the best model (src/fitting.rs:687-696)
    std::vector<float> prm((size_t)std::max(1, b->n_params));
    double c;
    int s1 = 0;
    const int rc = pasta_binary_eval_chi2(b, gb, 1, &c, &s1, prm.data());
    if (rc) {
        out->status = s1;
        return rc == PASTA_ERR_MODEL ? fail(PASTA_ERR_FIT, status_text(s1)) : rc;
    }
    out->n_params = std::min(512, b->n_params);
    for (int k = 0; k < out->n_params; k++) out->continuum_params[k] = prm[k];
    out->time_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return PASTA_OK;
}

int pasta_binary_rv_fit(pasta_binary_rv* r, const pasta_binary_settings* st, uint64_t seed,
                        const double* init_pos, pasta_rv_result* out, double* trace) {
    if (!r || !out) return fail(PASTA_ERR_ARG, "null argument");
    CHK(binary_settings_check(st));
    std::memset(out, 0, sizeof(*out));
    const auto t0 = std::chrono::steady_clock::now();
    const int P = st->num_particles;
    double gb[2], gc;
    CHK(pso_fit_n(r->g->geom, binary_settings_bounds(st, 2), st, seed, init_pos,
                  [&](const double* pos, double* cost, int* sts) {
                      return pasta_binary_rv_eval_chi2(r, pos, P, cost, sts);
                  },
                  gb, &gc, trace, &out->status));
    // the swapped pair (src/fitting.rs:846-863): its chi2 is reported when lower, with the
    // RVs as found
    const double sw[2] = {gb[1], gb[0]};
    double cs;
    int ss = 0;
    const int rc = pasta_binary_rv_eval_chi2(r, sw, 1, &cs, &ss);
    if (rc) {
        out->status = ss;
        return rc == PASTA_ERR_MODEL ? fail(PASTA_ERR_FIT, status_text(ss)) : rc;
    }
    out->rv1 = gb[0];
    out->rv2 = gb[1];
    out->chi2 = cs < gc ? cs : gc;
    out->iterations = st->max_iters;
    out->time_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return PASTA_OK;
}

int pasta_binary_ts_fit(pasta_binary* b, const pasta_binary_settings* st, uint64_t seed,
                        const double* init_pos, pasta_ts_result* out, float* continuum_params,
                        double* chis, double* trace) {
    if (!b || !out) return fail(PASTA_ERR_ARG, "null argument");
    if (b->D != 9) return fail(PASTA_ERR_ARG, "time-series fits take 9 labels");
    CHK(binary_settings_check(st));
    std::memset(out, 0, sizeof(*out));
    const auto t0 = std::chrono::steady_clock::now();
    const int P = st->num_particles;
    double gb[9], gc;
    CHK(pso_fit_n(b->g->geom, binary_settings_bounds(st, 9), st, seed, init_pos,
                  [&](const double* pos, double* cost, int* sts) {
                      return pasta_binary_ts_eval_chi2(b, pos, P, cost, sts);
                  },
                  gb, &gc, trace, &out->status));
    std::memcpy(out->label1, gb, sizeof(out->label1));
    std::memcpy(out->label2, gb + 4, sizeof(out->label2));
    out->light_ratio = gb[8];
    out->chi2 = gc;
    out->iterations = st->max_iters;
    // per-spectrum continuum parameters and chi2 of the best model (src/fitting.rs:1025-1039)
    std::vector<float> prm((size_t)b->S * std::max(1, b->n_params));
    std::vector<double> per(b->S);
    double c;
    int s1 = 0;
    const int rc = binary_chi2(b, gb, 1, &c, &s1, per.data(), prm.data());
    if (rc) {
        out->status = s1;
        return rc == PASTA_ERR_MODEL ? fail(PASTA_ERR_FIT, status_text(s1)) : rc;
    }
    out->n_params = b->n_params;
    out->n_spectra = b->S;
    if (continuum_params)
        std::memcpy(continuum_params, prm.data(), sizeof(float) * b->S * b->n_params);
    if (chis) std::memcpy(chis, per.data(), sizeof(double) * b->S);
    out->time_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return PASTA_OK;
}
