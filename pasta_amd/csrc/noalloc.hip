// noalloc.hip -- pasta_eval_chi2_noalloc: the batched chi2 evaluation of pasta_eval_chi2
// for callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// Every device buffer the evaluation needs beyond its inputs and outputs (the star
// descriptor copy, the labels, the evaluation records, rotation taps, tile bounds, sort
// and grouping tables, resampled models, continuum partials and parameters, chi2 and
// status) is carved out of one file-scope device array, `scratch`, of
// PASTA_NOALLOC_SCRATCH_BYTES.  Its device address is taken once per device with
// hipGetSymbolAddress and kept per device; host code never passes the array's name to a
// kernel.  The evaluation records and block tables, which pasta_eval_chi2 builds on the
// host, are written here by a layout kernel, so no host or device memory is allocated.
// The kernels are the ones pasta_eval_chi2 launches, over the same layout, so results
// are bit-identical to it.
#include "host.h"

#define PASTA_NOALLOC_SCRATCH_BYTES (64ull << 20)
#define PASTA_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[PASTA_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;  // serialises calls within the process (the array is shared)
void* g_scratch_ptr[PASTA_NOALLOC_MAX_DEVICES] = {};

// the device address of `scratch` on the current device, looked up on the first call
// made with that device current
int scratch_address(void** out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= PASTA_NOALLOC_MAX_DEVICES)
        return fail(PASTA_ERR_ARG, "device index beyond the scratch table");
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(PASTA_ERR_HIP, "hipGetSymbolAddress returned no address for scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PASTA_OK;
}

// sub-buffers of the scratch array, 256-byte aligned; with base == nullptr only sizes
struct Carve {
    char* base;
    size_t off = 0;
    template <typename T>
    T* take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T* p = base ? (T*)(base + off) : nullptr;
        off += (n ? n : 1) * sizeof(T);
        return p;
    }
};

// The layout of one star's P evaluations, as batch_build makes it for a single star
// in mode 0 (resample to the observed pixels).
struct Layout {
    int P;
    int64_t m_stride;   // floats between consecutive evaluations' models
    int64_t mp_stride;  // moment partials (doubles) per evaluation
    int64_t cp_stride;  // chi2 partials (doubles) per evaluation
    int64_t pf_stride;  // continuum parameters (floats) per evaluation
    int gm, gc;         // evaluations per moment / chi2 block
    int n_mblocks, n_cblocks;
    int solve_per_eval;  // chunk: n_chunks (e, c) items; linear: 1; constant, fixed: 0
    int n_bmom, n_bchi2, n_solve;
};

Layout make_layout(const DevStar& D, int P) {
    Layout L{};
    L.P = P;
    const bool mfma = D.cont_kind == CONT_LINEAR || D.cont_kind == CONT_CHUNK;
    L.m_stride = mfma ? (((int64_t)D.N + 3) & ~(int64_t)3) : (int64_t)D.N;
    L.mp_stride = (int64_t)D.n_mslots * std::max(1, D.p1);
    L.cp_stride = D.n_cblocks;
    L.pf_stride = std::max(1, D.n_params);
    const int p1 = D.cont_kind == CONT_CONSTANT ? 1 : (D.cont_kind == CONT_FIXED ? 0 : D.p1);
    L.gm = mom_variant(D.cont_kind, p1).e;
    L.gc = chi2_variant(D.cont_kind, p1).e;
    L.n_mblocks = D.n_mblocks;
    L.n_cblocks = D.n_cblocks;
    L.solve_per_eval = D.cont_kind == CONT_CHUNK ? D.n_chunks : D.cont_kind == CONT_LINEAR ? 1 : 0;
    L.n_bmom = ((P + L.gm - 1) / L.gm) * D.n_mblocks;
    L.n_bchi2 = ((P + L.gc - 1) / L.gc) * D.n_cblocks;
    L.n_solve = P * L.solve_per_eval;
    return L;
}

// One workgroup per evaluation: its record (zeroed, then the fields batch_build sets),
// and the blocks and solve items it heads.
__global__ __launch_bounds__(64) void k_noalloc_layout(Layout L, DevEval* __restrict__ evals,
                                                       int4* __restrict__ bmom,
                                                       int4* __restrict__ bchi2,
                                                       int2* __restrict__ solve) {
    const int e = blockIdx.x;
    int* w = (int*)(evals + e);
    for (int i = threadIdx.x; i < (int)(sizeof(DevEval) / sizeof(int)); i += blockDim.x) w[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        DevEval& ev = evals[e];
        ev.star = 0;
        ev.active = 1;
        ev.m_off = (long long)e * L.m_stride;
        ev.mp_off = (long long)e * L.mp_stride;
        ev.cp_off = (long long)e * L.cp_stride;
        ev.pf_off = (long long)e * L.pf_stride;
    }
    for (int c = threadIdx.x; c < L.solve_per_eval; c += blockDim.x)
        solve[(size_t)e * L.solve_per_eval + c] = make_int2(e, c);
    if (e % L.gm == 0) {
        const int g = e / L.gm, n = min(L.gm, L.P - e);
        for (int b = threadIdx.x; b < L.n_mblocks; b += blockDim.x)
            bmom[(size_t)g * L.n_mblocks + b] = make_int4(e, n, b, 0);
    }
    if (e % L.gc == 0) {
        const int g = e / L.gc, n = min(L.gc, L.P - e);
        for (int b = threadIdx.x; b < L.n_cblocks; b += blockDim.x)
            bchi2[(size_t)g * L.n_cblocks + b] = make_int4(e, n, b, 0);
    }
}

struct Bufs {
    DevStar* star;
    DevEval* evals;
    double *labels, *mpart, *cpart, *chi2;
    float *rot, *model, *pf;
    int *status, *hist, *perm, *tbound, *pair_eq, *nitems;
    int4 *items, *bmom, *bchi2;
    DevUnion* unions;
    int2* solve;
};

size_t carve(Carve& c, const Layout& L, const pasta_star* s, Bufs& b) {
    const DevStar& D = s->dev;
    const size_t P = L.P;
    b.star = c.take<DevStar>(1);
    b.evals = c.take<DevEval>(P);
    b.labels = c.take<double>(P * 5);
    b.rot = c.take<float>(P * ROT_STRIDE);
    b.tbound = c.take<int>(P * (size_t)(D.n_tiles + 1));
    b.hist = c.take<int>((size_t)s->grid->geom.n_models + 1);
    b.perm = c.take<int>(P);
    b.pair_eq = c.take<int>(P);
    b.items = c.take<int4>(P);
    b.unions = c.take<DevUnion>(P);
    b.nitems = c.take<int>(1);
    b.model = c.take<float>(P * (size_t)L.m_stride);
    b.mpart = c.take<double>(P * (size_t)L.mp_stride);
    b.cpart = c.take<double>(P * (size_t)L.cp_stride);
    b.pf = c.take<float>(P * (size_t)L.pf_stride);
    b.chi2 = c.take<double>(P);
    b.status = c.take<int>(P);
    b.bmom = c.take<int4>(L.n_bmom);
    b.bchi2 = c.take<int4>(L.n_bchi2);
    b.solve = c.take<int2>(L.n_solve);
    return c.off;
}

}  // namespace

int pasta_eval_chi2_noalloc(pasta_star* s, const double* labels, int P, double* chi2, int* status,
                            float* params) {
    if (!s || !labels || P <= 0 || !chi2) return fail(PASTA_ERR_ARG, "bad arguments");
    if (s->cont_kind < 0) return fail(PASTA_ERR_ARG, "star has no continuum model");
    const DevStar& D = s->dev;
    const Layout L = make_layout(D, P);
    Bufs b{};
    Carve sizing{nullptr};
    const size_t need = carve(sizing, L, s, b);
    if (need > PASTA_NOALLOC_SCRATCH_BYTES)
        return fail(PASTA_ERR_ARG,
                    "evaluation needs " + std::to_string(need) + " bytes of scratch, more than the " +
                        std::to_string(PASTA_NOALLOC_SCRATCH_BYTES) +
                        " the allocation-free path holds (fewer particles or pixels per call)");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    // null-stream work must not land inside another thread's stream capture
    std::shared_lock<std::shared_mutex> cap(capture_mutex());
    void* base = nullptr;
    CHK(scratch_address(&base));
    Carve c{(char*)base};
    carve(c, L, s, b);
    hipStream_t st = nullptr;  // the null stream: no stream is created
    HIPCHK(hipMemcpyAsync(b.star, &D, sizeof(DevStar), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b.labels, labels, sizeof(double) * 5 * (size_t)P, hipMemcpyHostToDevice,
                          st));
    hipLaunchKernelGGL(k_noalloc_layout, dim3(P), dim3(64), 0, st, L, b.evals, b.bmom, b.bchi2,
                       b.solve);
    HIPCHK(hipGetLastError());
    const int p1 = D.cont_kind == CONT_CONSTANT ? 1 : (D.cont_kind == CONT_FIXED ? 0 : D.p1);
    const ContClass cls{D.cont_kind, p1, 0, L.n_bmom, 0, L.n_bchi2};
    const DevGrid G = s->grid->dev();
    HIPCHK(launch_eval(G, b.star, b.evals, P, b.labels, b.rot, b.tbound, D.n_tiles + 1, b.hist,
                       b.perm, b.pair_eq, b.items, b.nitems, b.unions, D.n_tiles * P, b.bmom, b.bchi2, &cls, 1,
                       b.solve, L.n_solve, b.model, b.mpart, b.pf, b.cpart, b.chi2, b.status, 0,
                       may_fft(s->grid, max_label_vsini(labels, P), composed_extra_taps(s)), nullptr, nullptr, st));
    HIPCHK(hipMemcpyAsync(chi2, b.chi2, sizeof(double) * (size_t)P, hipMemcpyDeviceToHost, st));
    if (status)
        HIPCHK(hipMemcpyAsync(status, b.status, sizeof(int) * (size_t)P, hipMemcpyDeviceToHost, st));
    if (params && s->cont_kind != CONT_FIXED)
        HIPCHK(hipMemcpyAsync(params, b.pf, sizeof(float) * (size_t)P * L.pf_stride,
                              hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (params && s->cont_kind == CONT_FIXED)
        for (size_t i = 0; i < (size_t)P * L.pf_stride; i++) params[i] = 0.0f;
    // the first failed evaluation's status names the error (statuses read through a
    // fixed host window when the caller passes none)
    int window[256];
    for (int i0 = 0; i0 < P; i0 += 256) {
        const int n = std::min(256, P - i0);
        const int* st_h = status ? status + i0 : window;
        if (!status)
            HIPCHK(hipMemcpy(window, b.status + i0, sizeof(int) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++)
            if (st_h[i]) return fail(PASTA_ERR_MODEL, status_text(st_h[i]));
    }
    return PASTA_OK;
}
