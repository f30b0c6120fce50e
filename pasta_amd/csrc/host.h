// host.h -- host-side handles and helpers shared by the C-ABI sources (api.hip,
// binary.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

#include "engine.h"
#include "../../include/pasta_hip.h"

namespace pasta {
hipError_t launch_eval(const DevGrid& G, const DevStar* stars, DevEval* evals, int n_evals,
                       const double* labels, float* rot, int* tbound, int tb_stride,
                       int* sort_hist, int* perm, int* pair_eq, int4* items, int* n_items,
                       DevUnion* unions, int n_model,
                       const int4* bmom, const int4* bchi2, const ContClass* cls, int n_cls,
                       const int2* solve_items, int n_solve, float* model, double* mpart, float* pf,
                       double* cpart, double* chi2, int* status, int mode, bool fft,
                       hipEvent_t ev_model0, hipEvent_t ev_model1, hipStream_t st);
hipError_t launch_pso_update(const GridGeom* geom, DevPso* pso, DevEval* evals, int n, double* pos,
                             double* vel, const double* bpos, hipStream_t st);
hipError_t launch_pso_best(DevPso* pso, int n_stars, const DevEval* evals, const double* pos,
                           double* bpos, double* bcost, const double* cost, double* trace,
                           hipStream_t st);
hipError_t launch_eval_init(const EvalInit* in, int n, DevEval* out, hipStream_t st);
hipError_t launch_pso_remap(const int* src_eval, int n_evals, DevEval* evals, const int* src_star,
                            int n_stars, const DevPso* pso_old, DevPso* pso_new, const double* pos_old,
                            double* pos_new, const double* vel_old, double* vel_new,
                            const double* bpos_old, double* bpos_new, const double* bcost_old,
                            double* bcost_new, hipStream_t st);
hipError_t launch_pso_init(DevPso* pso, int n_stars, const DevEval* evals, double* pos,
                           double* vel, double* bpos, double* bcost, double* cost, double* trace,
                           hipStream_t st);
}  // namespace pasta

using namespace pasta;

// sets the thread-local message of pasta_last_error and returns code
int fail(int code, const std::string& msg);

// Stream capture (the batched and served PSO iteration graphs) holds this exclusively;
// device allocations, frees and synchronous copies hold it shared, so that another
// thread's memory call never lands inside a capture (which HIP then invalidates).
std::shared_mutex& capture_mutex();

#define HIPCHK(x)                                                                         \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess)                                                             \
            return fail(e_ == hipErrorOutOfMemory ? PASTA_ERR_NOMEM : PASTA_ERR_HIP,      \
                        std::string(#x) + ": " + hipGetErrorString(e_));                  \
    } while (0)
#define CHK(x)                  \
    do {                        \
        int r_ = (x);           \
        if (r_) return r_;      \
    } while (0)

static const double LN_10 = 2.302585092994045684017991454684364208;
static const double C_KMS = 299792.0;

// ---------------------------------------------------------------------------
// device buffer
// ---------------------------------------------------------------------------
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) {
            std::shared_lock<std::shared_mutex> lk(capture_mutex());
            (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        if (count <= n && p) return PASTA_OK;
        release();
        size_t c = count ? count : 1;
        std::shared_lock<std::shared_mutex> lk(capture_mutex());
        HIPCHK(hipMalloc((void**)&p, c * sizeof(T)));
        n = c;
        return PASTA_OK;
    }
    int upload(const T* h, size_t count) {
        CHK(alloc(count));
        std::shared_lock<std::shared_mutex> lk(capture_mutex());
        if (count) HIPCHK(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
        return PASTA_OK;
    }
    int upload(const std::vector<T>& v) { return upload(v.data(), v.size()); }
};

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct pasta_grid {
    GridGeom geom{};
    DBuf<GridGeom> d_geom;
    DBuf<uint16_t> d_spectra;
    DBuf<float> d_factors;
    bool has_factors = false;
    long long pitch = 0;
    int W = 0;
    double first = 0, step = 0, first_wl = 0, last_wl = 0;
    int fft_min_taps = FFT_MIN_TAPS;
    std::mutex mu;
    DevGrid dev() const {
        DevGrid g;
        g.spectra = d_spectra.p;
        g.factors = has_factors ? d_factors.p : nullptr;
        g.geom = d_geom.p;
        g.pitch = pitch;
        g.W = W;
        g.Wq = (W / 128) * 128;  // BATCH_SIZE (src/interpolate.rs:917,969)
        g.sg_first = first;
        g.sg_step = step;
        g.synth_first_wl = first_wl;
        g.synth_last_wl = last_wl;
        g.n_models_bins = geom.n_models + 1;
        g.fft_min_taps = fft_min_taps;
        g.prep_taps = KROT_MAX;
        g.stream_models = 0;
        return g;
    }
};

struct pasta_disp {
    int kind = DISP_NONE;
    std::vector<double> wl, fidx;
    std::vector<float> kern;  // fixed: K; variable: W x K
    int K = 0;
    int W = 0;
    double first = 0, step = 0;
};

struct pasta_cont {
    int kind = CONT_CONSTANT;
    int n = 0;
    int n_chunks = 0, p1 = 1, ignore = 0, n_params = 1;
    std::vector<int> startstop;
    std::vector<double> lohi, chol, wl;
    std::vector<float> fac, design_q, design_rm, fixedc;
    std::vector<float> cdesign;  // chunk design rows for the matrix-core moments
    std::vector<int> cdoff;
    std::vector<int4> mblocks, cblocks;
    std::vector<int2> mslot;  // per moment block: partial slots of its chunk(s)
    std::vector<int> mb_off;  // chunk c owns partial slots [mb_off[c], mb_off[c+1])
    double chi2_div = 1.0;
};

struct Batch;

struct pasta_star {
    pasta_grid* grid = nullptr;
    int N = 0;
    int n_params = 1, p1 = 1, cont_kind = CONT_CONSTANT, n_chunks = 0;
    DevStar dev{};
    std::vector<void*> allocs;
    int64_t resident_bytes = 0;  // device memory the star's own buffers hold
    int span = 0;                // synthetic pixels the observed range spans
    std::unique_ptr<Batch> cache;
    int cache_P = -1;
    std::mutex mu;
    ~pasta_star();
};

struct Batch {
    int mode = 0;  // 0: resample to the observed pixels; 1: synthetic-grid output
    // largest vsini of the batch's evaluations (from the labels or the PSO bounds; infinite
    // when unknown): selects whether its launches carry the FFT rotation path
    double max_vsini = HUGE_VAL;
    int n_evals = 0, n_stars = 0;
    std::vector<ContClass> cls;
    pasta_grid* grid = nullptr;
    std::vector<pasta_star*> stars;
    std::vector<int> first_eval, count;
    std::vector<EvalInit> h_evals;  // host part of the evaluation records
    DBuf<EvalInit> d_init;
    int64_t model_len = 0, mpart_len = 0, cpart_len = 0, pf_len = 0;
    DBuf<DevStar> d_stars;
    DBuf<DevEval> d_evals;
    DBuf<double> d_labels, d_mpart, d_cpart, d_chi2;
    DBuf<float> d_rot, d_model, d_pf;
    DBuf<int> d_status, d_hist, d_perm, d_tbound, d_pair_eq, d_nitems;
    DBuf<int4> d_items;
    DBuf<DevUnion> d_union;
    int tb_stride = 1;
    DBuf<int4> d_bmom, d_bchi2;
    DBuf<int2> d_solve;
    int n_bm = 0, n_bmom = 0, n_solve = 0, n_bchi2 = 0;
    // batched PSO state
    DBuf<DevPso> d_pso;
    DBuf<double> d_vel, d_bpos, d_bcost, d_trace;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // model begin/end, iteration begin/end
    bool eager_events = false;  // record ev[] around eager (non-graph) launches
    ~Batch() {
        if (graph) (void)hipGraphExecDestroy(graph);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Star on a grid: dispersion (observed sampling) and, for chi2, continuum + flux/var
int star_build(pasta_grid* g, const pasta_disp* d, const pasta_cont* c, const float* flux,
               const float* var, int n, std::unique_ptr<pasta_star>& out);
// Evaluation batch over stars with counts[s] evaluations each; mode 0 resamples to the
// observed pixels, mode 1 keeps the synthetic grid
int batch_build(Batch& B, pasta_grid* g, const std::vector<pasta_star*>& stars,
                const std::vector<int>& counts, int mode);
// whether evaluations of vsini <= max_vsini on these stars can reach the FFT crossover
bool may_fft(const pasta_grid* g, double max_vsini, const std::vector<pasta_star*>& stars);
bool may_fft(const pasta_grid* g, double max_vsini, int extra_taps);
int composed_extra_taps(const pasta_star* s);  // taps a composed instrumental kernel adds
double max_label_vsini(const double* labels5, int n);  // NaN counts as unbounded
int batch_launch(Batch& B);
int batch_eval(Batch& B, const double* labels, double* chi2, int* status);
const char* status_text(int st);
int settings_check(const pasta_fit_settings* st);
