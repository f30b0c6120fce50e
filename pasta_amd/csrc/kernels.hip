// kernels.hip -- gfx950 kernels of the model-evaluation path.
//
// One chi2 evaluation (reference: CostFunction::cost, src/fitting.rs:313-326) is
//   k_prep     neighbour indices + 64 Lagrange weights, rotation taps, RV shift
//              (src/cubic.rs:232-243, src/interpolate.rs:424-533,
//               src/convolve_rv.rs:438-475,532-567)
//   k_model    per (evaluation, synthetic tile of TILE pixels), all in LDS:
//              interpolate (src/interpolate.rs:930-983, incl. the n%128 tail
//              quirk) -> rotation convolution (zero-padded "same",
//              src/convolve_rv.rs:484-529) -> instrumental convolution
//              (src/convolve_rv.rs:160-211 / 405-431; a fixed instrumental kernel is
//              folded into the rotation taps away from the grid edges) -> RV shift +
//              linear resample with the f32-quantised weight (src/convolve_rv.rs:569-595)
//   k_moments  per-chunk Aᵀ(f/m) in f64 (src/continuum_fitting.rs:216-226)
//   k_solve    per (evaluation, chunk) f64 Cholesky solve of the normal equations
//   k_chi2     blended-polynomial chi2 terms, f64 accumulation
//              (src/continuum_fitting.rs:280-360, 69-95, 442-500)
//   k_final    fixed-order sum of the chi2 partials
// All reductions use fixed trees, so results are bit-reproducible run to run
// and independent of which other evaluations share the launch.
//
// Built with -ffp-contract=off: only the explicit fmaf() calls fuse.
#include "engine.h"

namespace pasta {

// Interpolated tile for the rotation stage: logical index k lives in slot k + 4*(k/16)
// (4 pad floats after every 16), so that the 16-float-strided MFMA operand rows fall
// in distinct LDS banks.  Logical extent <= 7 + (TILE + 1 + 2*HD_MAX) + KROT_MAX + 55.
constexpr int ILOG = (7 + TILE + 1 + 2 * HD_MAX + KROT_MAX + 55 + 31) & ~31;  // 5696 at 3312
constexpr int ILDS = ILOG + ILOG / 4;
// For the f16 rotation the tile is kept as two f16 arrays (hi, lo) of ILOGB elements
// (logical extent <= 7 + (TILE + 1 + 2*HD_MAX) + KB_MAX + 71).
constexpr int ILOGB = (7 + TILE + 1 + 2 * HD_MAX + KB_MAX + 71 + 7) & ~7;     // 5184 at 3312
// The rotation output R (TILE + 2 hd + 2 values, 16 zeros after) and the instrumental
// stage's output C (<= TILE + 1 values) share k_model's one LDS buffer with the rotation's
// operands: R overwrites the interpolated tile at the buffer's start once every wavefront
// has read its operands (for the FFT path, whose transform buffer that is, R follows the
// transform buffer), and C follows R (for the FFT path, C takes the transform buffer).
constexpr int RLDS = TILE + 1 + 2 * HD_MAX + 8 + 16 + 8;
constexpr int LDS_ROT = ILDS + KROT_MAX + 48;  // floats of k_model's buffer
constexpr int C_OFF = (RLDS + 3) & ~3;         // C after R
// k_model's LDS: 4 workgroups per CU share its 160 KB
static_assert((LDS_ROT + 2 * HD_MAX + 1 + 8) * 4 <= 40960,
              "k_model LDS beyond a quarter of the CU's");
static_assert(C_OFF + TILE + 8 <= LDS_ROT, "R and C within the buffer");
constexpr double C_KMS = 299792.0;   // src/convolve_rv.rs:440,547
constexpr double LN10 = 2.302585092994045684017991454684364208;

typedef const __attribute__((address_space(1))) double GlobalD;
typedef const __attribute__((address_space(1))) float GlobalF;

__device__ __forceinline__ long long jfloor(double x) {
    double f = floor(x);
    return f > 0.0 ? (long long)f : 0;  // `as usize` saturates at 0
}

// Workgroup i is dispatched to XCD i % 8: give each XCD a contiguous range of
// logical blocks so neighbouring tiles (shared halos, shared neighbour rows)
// meet in the same L2.
__device__ __forceinline__ int xcd_remap(int i, int n) {
    int q = n >> 3, r = n & 7;
    int x = i & 7, k = i >> 3;
    return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ int rotation_len(double vsini, double dvelo) {
    const double eps = 0.6;
    double vrot = vsini / C_KMS;
    double a = 4.0 * (eps - 1.0) / (M_PI * eps);
    double v = ceil(2.0 * vrot / dvelo * sqrt(1.0 + a * a));
    long long nme = (v >= 0.0) ? (long long)v + 2 : 2;
    if (v > 1e9) return 1 << 30;
    if (nme < 1) nme = 1;
    return (int)(nme + 1 - nme % 2);
}

// first observed pixel whose resample index floor(fidx + shift) >= t
__device__ __forceinline__ int obs_lower_bound(const double* __restrict__ fidx, int N, double sl,
                                               long long t) {
    if (t <= 0) return 0;
    int lo = 0, hi = N;
    const double tt = (double)t;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (fidx[mid] + sl >= tt) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Two-way f16 split x = hi + lo (each round-to-nearest; x - hi is exact in f32): for
// 2^-3 <= |x| < 2^14 both parts are normal f16 and hi + lo is x to 2^-22 relative; below,
// lo's absolute error is at most 2^-25.  Products of two split values keep ~2^-21 relative
// accuracy when the smallest cross term lo*lo (<= 2^-22 relative) is dropped.  Callers scale
// x by a power of 2 first (exact), so that their largest |x| is just below 2^14.
__device__ __forceinline__ void split2(float x, unsigned short& h, unsigned short& l) {
    const _Float16 fh = (_Float16)x;
    const _Float16 fl = (_Float16)(x - (float)fh);
    h = __builtin_bit_cast(unsigned short, fh);
    l = __builtin_bit_cast(unsigned short, fl);
}

// 2^(14 - e) for a positive bound b in [2^(e-1), 2^e): b * scale < 2^14
__device__ __forceinline__ float pow2_below_2_14(float b) {
    if (!(b > 0.0f)) return 1.0f;
    int e;
    frexpf(b, &e);
    return ldexpf(1.0f, 14 - e);
}

// native 16-byte vector (HIP's uint4 struct is copied member-wise into 4 dword loads)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// k_prep: one wavefront per evaluation
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_prep(DevGrid G, const DevStar* __restrict__ stars,
                                             const double* __restrict__ labels,
                                             DevEval* __restrict__ evals, float* __restrict__ rot,
                                             int* __restrict__ tbound, int tb_stride, int mode) {
    const int e = blockIdx.x;
    const int lane = threadIdx.x;
    DevEval& ev = evals[e];
    __shared__ int s_K;
    __shared__ float s_sum;
    const double* lab = labels + 5 * (size_t)e;
    const double vsini = lab[3];
    const double dvelo = LN10 * G.sg_step;
    __shared__ int s_st;
    __shared__ int s_comp;
    __shared__ int s_fft;
    // rotation taps and composed taps, G.prep_taps + 4 floats each (dynamic LDS: a launch
    // whose kernels are short holds several times as many evaluations per CU)
    extern __shared__ __align__(16) float s_dyn[];
    const int cap = G.prep_taps;
    float* const s_taps[2] = {s_dyn, s_dyn + cap + 4};
    __shared__ float s_g[2 * HD_MAX + 1];  // instrumental taps
    const bool active = ev.active;
    if (lane == 0) s_st = active ? local_weights(*G.geom, lab[0], lab[1], lab[2], ev.idx, ev.w) : -1;
    __syncthreads();
    // neighbour j = lane.  fw = factor * weight: the reference forms f32(u) * factor, then
    // the gemv multiplies by the weight (interpolate.rs:962,966); one product per neighbour
    // differs from that by an f32 rounding per term
    const int ij = ev.idx[lane];
    float fwj = 0.0f;
    if (active) {
        // the present neighbours' rows and weights first, in the neighbour order (an absent
        // neighbour's weight is 0, so leaving it out of the sum changes nothing), then rows of
        // weight 0: the interpolation gathers only the present rows, rounded up to 8
        const float f = (ij >= 0 && G.factors) ? G.factors[ij] : 1.0f / 65535.0f;
        fwj = ij >= 0 ? f * ev.w[lane] : 0.0f;
        const unsigned long long pm = __ballot(ij >= 0);
        const int np = __popcll(pm);
        const int pos = __popcll(pm & ((1ull << lane) - 1ull));
        if (ij >= 0) {
            ev.fw[pos] = fwj;
            ev.rowp[pos] = G.spectra + (long long)ij * G.pitch;
        }
        if (lane >= np) {
            ev.fw[lane] = 0.0f;
            ev.rowp[lane] = G.spectra;
        }
        if (lane == 0) ev.nrow = np;
    }
    // scale of the f16-split interpolated values: |sum fw u| <= 65535 sum |fw|
    float rsc = 1.0f;
    if (active) {
        float a = fabsf(fwj);
        for (int o = 32; o; o >>= 1) a += __shfl_xor(a, o, 64);
        rsc = pow2_below_2_14(65535.0f * a * 1.0001f);
    }
    unsigned int hsh = 2166136261u;  // FNV-1a over the neighbour rows: fast reject in pairing
    for (int j = 0; j < 64; j++) hsh = (hsh ^ (unsigned int)__builtin_amdgcn_readlane(ij, j)) * 16777619u;
    const unsigned long long present = __ballot(ij >= 0);
    const int first = present ? __builtin_amdgcn_readlane(ij, __builtin_ctzll(present)) : 0;
    if (lane == 0) {
        int st = s_st, krot = 0;
        bool comp = false;
        if (active) {
            const DevStar& S = stars[ev.star];
            if (!st && vsini < 0.0) st = ST_VSINI;
            if (!st) {
                int K = rotation_len(vsini, dvelo);
                if (vsini == 0.0 || K <= 2) krot = 0;
                else if (K > KROT_MAX) st = ST_KERNEL;
                else if (K > cap) st = ST_KERNEL;  // cap bounds every K of the launch: unreached
                else krot = K;
            }
            if (!st && mode == 1) {  // synthetic-grid output: no RV shift / resample
                ev.shift_log = 0.0;
                ev.cmax = G.W;
                ev.tile_base = 0;
            } else if (!st) {
                double shift = 1.0 - lab[4] / C_KMS;
                if (S.wl0 * shift < G.synth_first_wl) st = ST_RANGE;
                else if (S.wlN * shift > G.synth_last_wl) st = ST_RANGE;
                double sl = log10(shift) / G.sg_step;
                ev.shift_log = sl;
                int cmax = G.W;
                if (S.disp_kind == DISP_NONE && krot > 0) {
                    long long step = KROT_MAX - (krot - 1);
                    long long nx = ((G.W + step - 1) / step) * step;
                    cmax = (int)(nx + krot - 1 - krot / 2);
                }
                ev.cmax = cmax;
                long long jl = jfloor(S.fidx[S.N - 1] + sl);
                if (!st && jl + 1 >= cmax) st = ST_INDEX;
                const long long j0 = jfloor(S.fidx[0] + sl);
                ev.tile_base = (int)(j0 / TILE);
                // Rotation then fixed instrumental broadening, both zero-padded "same"
                // convolutions, equal one "same" convolution with the composed kernel
                // wherever the rotation output the instrumental stage reads lies inside
                // [0, W) (no zero padding between the stages): true for every tile when
                // the convolved range [j0, jl + 1] stays hd pixels inside the grid.  (Both
                // kernels have odd lengths, so the composed one is centred like its parts.)
                const int hd = (S.Kd - 1) / 2;
                comp = !st && S.disp_kind == DISP_FIXED && krot > 0 && (S.Kd & 1) &&
                       krot + S.Kd - 1 <= cap && j0 - hd >= 0 && jl + 1 + hd <= G.W - 1;
            }
        }
        ev.status = st;
        ev.comp = comp;
        s_comp = comp;
        ev.krot = krot;
        // rotation path: the FFT for kernels of at least fft_min_taps taps that fit one
        // transform with their tiles, the matrix cores otherwise
        bool fft = false;
        if (st == ST_OK && krot > 0) {
            const DevStar& S = stars[ev.star];
            const int Kt = comp ? krot + S.Kd - 1 : krot;
            const int hd = (comp || S.disp_kind == DISP_NONE) ? 0 : (S.Kd - 1) / 2;
            fft = Kt >= G.fft_min_taps && fft_fits(Kt, hd);
        }
        ev.fft = fft;
        s_fft = fft;
        ev.rows_hash = (int)hsh;
        ev.rsc = rsc;
        // storage row of the first present neighbour
        ev.key = (st == ST_OK && present) ? first : G.geom->n_models;
        s_K = (st == ST_OK) ? krot : 0;
    }
    __syncthreads();
    // observed-pixel range of every synthetic tile of this evaluation, 64 searches at a time
    if (mode == 0 && ev.status == ST_OK) {
        const DevStar& S = stars[ev.star];
        const double sl = ev.shift_log;
        const long long tb0 = (long long)ev.tile_base * TILE;
        int* out = tbound + (size_t)e * tb_stride;
        for (int k = lane; k <= S.n_tiles; k += 64)
            out[k] = obs_lower_bound(S.fidx, S.N, sl, tb0 + (long long)k * TILE);
    }
    const int K = s_K;
    if (K > 0) {
        // the taps are formed in LDS (the sequential sum and the composition read them back
        // there, not from global memory), then stored to the evaluation's tap slot
        float* h = rot + (size_t)e * ROT_STRIDE;
        float* hs = s_taps[0];
        const double eps = 0.6;
        const double vrot = vsini / C_KMS;
        const double mid = ((double)(K - 1) * dvelo) / 2.0;
        for (int i = lane; i < K; i += 64) {
            double vk = (double)i * dvelo - mid;
            double q = vk / vrot;
            double y = 1.0 - q * q;
            float k = 0.0f;
            if (y > 0.0)
                k = (float)((2.0 * (1.0 - eps) * sqrt(y) + M_PI * eps / 2.0 * y) /
                            (M_PI * vrot * (1.0 - eps / 3.0)));
            hs[i] = k;
        }
        if (lane < 4) hs[K + lane] = 0.0f;  // the sum's last quad: + 0 leaves it unchanged
        __syncthreads();
        if (lane == 0) {  // f32 sequential sum, as kernel.iter().sum() (convolve_rv.rs:470)
            float s = 0.0f;
            for (int i = 0; i < K; i += 4) {
                const f4 v = *(const f4*)(hs + i);
                s = s + v.x;
                s = s + v.y;
                s = s + v.z;
                s = s + v.w;
            }
            s_sum = s;
        }
        __syncthreads();
        const float s = s_sum;
        for (int i = lane; i < K; i += 64) hs[i] = hs[i] / s;
        int Kt = K;
        if (s_comp) {
            // composed taps c[k] = sum_u g[u] h[k - u] of the normalised rotation taps h and
            // the instrumental taps g, summed in f64 and rounded once to f32
            const DevStar& S = stars[ev.star];
            const int Kd = S.Kd;
            Kt = K + Kd - 1;
            float* cs = s_taps[1];
            for (int u = lane; u < Kd; u += 64) s_g[u] = S.dkern[u];
            __syncthreads();
            for (int k = lane; k < Kt; k += 64) {
                const int u0 = k - (K - 1) > 0 ? k - (K - 1) : 0;
                const int u1 = k < Kd - 1 ? k : Kd - 1;
                double a = 0.0;
                for (int u = u0; u <= u1; u++) a += (double)s_g[u] * (double)hs[k - u];
                cs[k] = (float)a;
            }
            hs = cs;
            if (lane == 0) ev.krot = Kt;
        }
        __syncthreads();
        for (int k = lane; k < Kt; k += 64) h[k] = hs[k];  // the taps in the evaluation's slot
        if (Kt <= KB_MAX && !s_fft) {
            // reversed, zero-padded taps tpad[16 + u] = h[Kt - 1 - u], scaled by a power of
            // 2 (largest tap just below 2^14) and split into f16 hi/lo, as pairs
            // tw[part][parity][q] = (tpad[2q + parity], tpad[2q + parity + 1])
            float tmax = 0.0f;
            for (int k = lane; k < Kt; k += 64) tmax = fmaxf(tmax, hs[k]);
            for (int o = 32; o; o >>= 1) tmax = fmaxf(tmax, __shfl_xor(tmax, o, 64));
            const float tsc = pow2_below_2_14(tmax);
            if (lane == 0) ev.rinv = 1.0f / (rsc * tsc);  // a power of 2: exact
            unsigned int* tw = (unsigned int*)(h + KROT_MAX);
            const int nq = 16 * ((Kt + 46) >> 5) + 8;  // every pair the MFMA groups read
            for (int i = lane; i < 2 * nq; i += 64) {
                const int par = i >= nq, q = i - par * nq;
                unsigned short v[2][2];
#pragma unroll
                for (int z = 0; z < 2; z++) {
                    const int u = 2 * q + par + z - 16;
                    split2((u >= 0 && u < Kt) ? hs[Kt - 1 - u] * tsc : 0.0f, v[z][0], v[z][1]);
                }
#pragma unroll
                for (int part = 0; part < 2; part++)
                    tw[(2 * part + par) * TW_MAX + q] =
                        (unsigned int)v[0][part] | ((unsigned int)v[1][part] << 16);
            }
        }
    }
}


__device__ __forceinline__ void unpack8(const u32x4 v, float u[8]) {
    u[0] = (float)(v.x & 0xffffu); u[1] = (float)(v.x >> 16);
    u[2] = (float)(v.y & 0xffffu); u[3] = (float)(v.y >> 16);
    u[4] = (float)(v.z & 0xffffu); u[5] = (float)(v.z >> 16);
    u[6] = (float)(v.w & 0xffffu); u[7] = (float)(v.w >> 16);
}

// Interpolated synthetic flux of 8-pixel groups aligned to 8 in absolute pixel
// coordinates: group gg covers pixels [g0 + 8gg, g0 + 8gg + 8) (g0 a multiple of 8)
// and lands at logical index 8gg of dst, in the padded slot layout when PAD.  Pixels
// outside [0, W) are zero.  The 64 neighbour rows are addressed through one buffer
// descriptor per row built from scalar registers (row index and factor*weight are
// uniform over the workgroup), so a 16-byte gather costs no vector address
// arithmetic; a group outside the row is given an offset past the descriptor's
// extent and reads zeros.  16 rows are in flight per thread at a time.  Absent
// neighbours read row 0 with weight 0 (their weight is 0 in the reference too).
constexpr int OOB_OFFSET = 0x7ffffff0;

template <bool FAST>
__device__ __forceinline__ void acc8(float (&acc)[8], const u32x4 v, float fw, long long x0, int Wq) {
    float u[8];
    unpack8(v, u);
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const float q = (FAST || x0 + t < Wq) ? u[t] : u[t] / 65535.0f;  // tail quirk :978
        acc[t] = fmaf(q, fw, acc[t]);
    }
}

template <bool FAST, int RF>
__device__ __forceinline__ void gather_batch(const DevGrid& G, const DevEval& ev, int j0, int voff,
                                             long long x0, float (&acc)[8]) {
    const int nrec = (int)(G.pitch * 2);
    u32x4 v[RF];
#pragma unroll
    for (int jj = 0; jj < RF; jj++) {
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void*)ev.rowp[j0 + jj], 0, nrec, 0x00020000);
        v[jj] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, 0, 0));
    }
#pragma unroll
    for (int jj = 0; jj < RF; jj++) {
        const float fw = __builtin_bit_cast(
            float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, ev.fw[j0 + jj])));
        acc8<FAST>(acc, v[jj], fw, x0, G.Wq);
    }
}

// the present rows (rounded up to 8) in batches of 16 rows in flight, an 8-row batch last
template <bool FAST>
__device__ __forceinline__ void gather_rows(const DevGrid& G, const DevEval& ev, int voff, long long x0,
                                            float (&acc)[8]) {
    const int n8 = (__builtin_amdgcn_readfirstlane(ev.nrow) + 7) & ~7;
    int j0 = 0;
#pragma unroll 1
    for (; j0 + 16 <= n8; j0 += 16) gather_batch<FAST, 16>(G, ev, j0, voff, x0, acc);
    if (j0 < n8) gather_batch<FAST, 8>(G, ev, j0, voff, x0, acc);
}

// Workgroup barrier that orders LDS accesses only: k_model's threads share nothing through
// global memory, so outstanding global loads and stores (the model's resampled pixels) need
// not drain at every barrier, as the fence of __syncthreads() would make them.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ int islot(int k) { return k + ((k >> 4) << 2); }

typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// out[q] = sum_u tr[u] * in[q + u] for q in [0, n), u over the taps zero-padded to a
// multiple of 8 in order (sequential fmaf).  One output per thread and round: the lanes of
// a wavefront read consecutive LDS words, which no two lanes share a bank for (a sliding
// window of 8 outputs per thread puts the lanes 32 bytes apart: 16 lanes per bank).  Up
// to 8 taps stay in registers; longer kernels read them as LDS broadcasts.  `in` is
// finite up to index n - 1 + (padded tap count) - 1.
__device__ __forceinline__ void conv_tile(const float* in, const float* tr, int K, int n,
                                          float* out, bool mask, long long x0, int W) {
    const int Kp = (K + 7) & ~7;
    if (Kp == 8) {
        float h[8];
#pragma unroll
        for (int u = 0; u < 8; u++) h[u] = tr[u];
        for (int q = threadIdx.x; q < n; q += NT) {
            float a = 0.0f;
#pragma unroll
            for (int u = 0; u < 8; u++) a = fmaf(h[u], in[q + u], a);
            const long long x = x0 + q;
            out[q] = (mask && (x < 0 || x >= W)) ? 0.0f : a;
        }
        return;
    }
    for (int q = threadIdx.x; q < n; q += NT) {
        float a = 0.0f;
        for (int u0 = 0; u0 < Kp; u0 += 8) {
#pragma unroll
            for (int u = 0; u < 8; u++) a = fmaf(tr[u0 + u], in[q + u0 + u], a);
        }
        const long long x = x0 + q;
        out[q] = (mask && (x < 0 || x >= W)) ? 0.0f : a;
    }
}

// Rotational broadening on the matrix cores: R[p] = sum_u tr[u] * I[p + u] for
// p < np, as the Toeplitz product D(blocks x 16) = A(blocks x J) * T(J x 16) with
// A[a][j] = I[16a + j], T[j][n] = tr[j - n] and R[16a + n] = D[a][n], on
// v_mfma_f32_16x16x4f32 (fp32 products, fp32 accumulation).  Operand layout of the
// instruction (checked on gfx950): A lane l = A[l%16][l/16], B lane l = B[l/16][l%16],
// D lane l register i = D[4(l/16)+i][l%16].  The contraction index is dealt out so
// that lane (m, kk) of step r in group c takes j = 16c + 4kk + r: its four A operands
// of a group are I[16(a+c) + 4kk .. +3], one 16-byte LDS read in the padded layout
// (slot 20(a+c) + 4kk), conflict-free across the 16 rows.  hpad[16 + u] = tr[u] for
// u < K and zero elsewhere in [0, 16C + 16).  Each wave owns 16-block tiles w and w+4
// (and w + 8, w + 12); rows past the last block read the last block (their results are
// not stored).  Outputs are stored to out[p - off] for off <= p < np, zero outside [0, W)
// when mask; out may be the operand tile (the stores follow a barrier).
// A wavefront's output tiles (NTR 16 x 16-block tiles per round, 4 apart: t, t + 4, ...;
// rounds 4 NTR tiles apart) are held in registers until every wavefront has read its
// operands, then stored over the operand tile: ROT_ROUNDS rounds cover any tile and
// instrumental halo.
constexpr int ROT_ROUNDS = 2;
constexpr int ROT_TILES_MAX = (((7 + TILE + 2 + 2 * HD_MAX + 15) >> 4) + 15) >> 4;
static_assert(ROT_TILES_MAX <= 8 * ROT_ROUNDS, "rotation outputs within two rounds per wavefront");

template <int NTR, bool SCALE>
__device__ __forceinline__ void store_rot(const f4 (&c)[ROT_ROUNDS][NTR], int tiles, int np,
                                          int off, float* out, bool mask, long long x0, int W,
                                          float sc) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = lane & 15, kk = lane >> 4;
    lds_barrier();  // every wavefront has read its operands
#pragma unroll
    for (int it = 0; it < ROT_ROUNDS; it++) {
#pragma unroll
        for (int j = 0; j < NTR; j++) {
            const int t = wv + 4 * j + 4 * NTR * it;
            if (4 * j + 4 * NTR * it >= ROT_TILES_MAX || t >= tiles) continue;  // (no such tile)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int p = 16 * (16 * t + 4 * kk + i) + m;
                if (p >= off && p < np) {
                    const long long x = x0 + (p - off);
                    out[p - off] =
                        (mask && (x < 0 || x >= W)) ? 0.0f : (SCALE ? c[it][j][i] * sc : c[it][j][i]);
                }
            }
        }
    }
}

__device__ __forceinline__ void conv_rot_mfma(const float* Ipad, const float* hpad, int K, int np,
                                              int off, float* out, bool mask, long long x0, int W) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nblk = (np + 15) >> 4;
    const int tiles = (nblk + 15) >> 4;
    const int C = (K + 15 + 15) >> 4;
    const int m = lane & 15, kk = lane >> 4;
    const float* hb = hpad + 16 + 4 * kk - m;
    f4 k[ROT_ROUNDS][2];
#pragma unroll
    for (int it = 0; it < ROT_ROUNDS; it++) {
        const int t0 = wv + 8 * it, t1 = t0 + 4;
        if (8 * it >= ROT_TILES_MAX || t0 >= tiles) break;
        const bool has1 = 8 * it + 4 < ROT_TILES_MAX && t1 < tiles;
        const int a0 = min(16 * t0 + m, nblk - 1), a1 = min(16 * t1 + m, nblk - 1);
        const f4* ia0 = (const f4*)(Ipad + 20 * a0 + 4 * kk);
        const f4* ia1 = (const f4*)(Ipad + 20 * a1 + 4 * kk);
        f4 c0 = {0.0f, 0.0f, 0.0f, 0.0f}, c1 = {0.0f, 0.0f, 0.0f, 0.0f};
        if (has1) {
#pragma unroll 2
            for (int c = 0; c < C; c++) {
                const f4 x0v = ia0[5 * c], x1v = ia1[5 * c];
                const float* h = hb + 16 * c;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float bv = h[r];
                    c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0v[r], bv, c0, 0, 0, 0);
                    c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x1v[r], bv, c1, 0, 0, 0);
                }
            }
        } else {
#pragma unroll 2
            for (int c = 0; c < C; c++) {
                const f4 x0v = ia0[5 * c];
                const float* h = hb + 16 * c;
#pragma unroll
                for (int r = 0; r < 4; r++)
                    c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x0v[r], h[r], c0, 0, 0, 0);
            }
        }
        k[it][0] = c0;
        k[it][1] = c1;
    }
    store_rot<2, false>(k, tiles, np, off, out, mask, x0, W, 1.0f);
}

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4b __attribute__((ext_vector_type(4)));

// Rotational broadening for K <= KB_MAX on v_mfma_f32_16x16x32_f16, the same Toeplitz
// product as conv_rot_mfma with 32 contraction steps per instruction.  Both operands are
// scaled by powers of 2 (interpolated values by the evaluation's rsc, taps so that the
// largest is just below 2^14) and split into f16 hi + lo; the three products hi*hi, hi*lo,
// lo*hi are accumulated in f32 and the output is scaled back by rinv (exact).  Relative
// error per product ~3 * 2^-22 (the representations and the dropped lo*lo).  Operand
// layout (checked on gfx950 for the bf16 instruction of the same shape): A lane l holds
// A[l%16][8(l/16)+t], B lane l holds B[8(l/16)+t][l%16], t < 8, D as for the f32
// instruction.  Lane (m, kk) of group c takes j = 32c + 8kk + t: its A operands are
// I[16a + 32c + 8kk .. +7] (one 16-byte read per split part), its B operands the padded
// taps tpad[32c + 8kk + 16 - n .. +7] (tpad[16 + u] = tr[u]), read as four packed pairs
// from the copy whose pairs start at the parity of that index: tw[part][p][q] =
// (tpad[2q + p], tpad[2q + p + 1]).
__device__ __forceinline__ void load_a2(const unsigned short* p, f16x8& h, f16x8& l) {
    h = __builtin_bit_cast(f16x8, *(const u16x8*)p);
    l = __builtin_bit_cast(f16x8, *(const u16x8*)(p + ILOGB));
}

__device__ __forceinline__ void load_b2(const unsigned int* q, f16x8& h, f16x8& l) {
    const u32x4b vh = {q[0], q[1], q[2], q[3]};
    const u32x4b vl = {q[2 * TW_MAX], q[2 * TW_MAX + 1], q[2 * TW_MAX + 2], q[2 * TW_MAX + 3]};
    h = __builtin_bit_cast(f16x8, vh);
    l = __builtin_bit_cast(f16x8, vl);
}

// c += the three split products of one 32-step group (smallest terms first)
__device__ __forceinline__ f4 mfma3(const f16x8& ah, const f16x8& al, const f16x8& bh,
                                    const f16x8& bl, f4 c) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, c, 0, 0, 0);
    return c;
}

__device__ __forceinline__ void conv_rot_f16(const unsigned short* Ib2, const unsigned int* tw,
                                             int K, int np, int off, float* out, bool mask,
                                             long long x0, int W, float rinv) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nblk = (np + 15) >> 4;
    const int tiles = (nblk + 15) >> 4;
    const int C = (K + 15 + 31) >> 5;
    const int m = lane & 15, kk = lane >> 4;
    const int s0 = 8 * kk + 16 - m;  // tpad index of this lane's first B operand, group 0
    const unsigned int* twp = tw + (s0 & 1) * TW_MAX + (s0 >> 1);
    // three output tiles per round (t, t + 4, t + 8) share each group's tap operands
    f4 k[ROT_ROUNDS][3];
#pragma unroll
    for (int it = 0; it < ROT_ROUNDS; it++) {
        const int t0 = wv + 12 * it, t1 = t0 + 4, t2 = t0 + 8;
        if (12 * it >= ROT_TILES_MAX || t0 >= tiles) break;
        // (tiles no tile length reaches are left out at compile time)
        const bool has1 = 12 * it + 4 < ROT_TILES_MAX && t1 < tiles;
        const bool has2 = 12 * it + 8 < ROT_TILES_MAX && t2 < tiles;
        const int a0 = min(16 * t0 + m, nblk - 1), a1 = min(16 * t1 + m, nblk - 1);
        const int a2 = min(16 * t2 + m, nblk - 1);
        const unsigned short* ia0 = Ib2 + 16 * a0 + 8 * kk;
        const unsigned short* ia1 = Ib2 + 16 * a1 + 8 * kk;
        const unsigned short* ia2 = Ib2 + 16 * a2 + 8 * kk;
        f4 c0 = {0.0f, 0.0f, 0.0f, 0.0f}, c1 = c0, c2 = c0;
        if (has2) {
            for (int c = 0; c < C; c++) {
                f16x8 bh, bl, ah, al, xh, xl, yh, yl;
                load_b2(twp + 16 * c, bh, bl);
                load_a2(ia0 + 32 * c, ah, al);
                load_a2(ia1 + 32 * c, xh, xl);
                load_a2(ia2 + 32 * c, yh, yl);
                c0 = mfma3(ah, al, bh, bl, c0);
                c1 = mfma3(xh, xl, bh, bl, c1);
                c2 = mfma3(yh, yl, bh, bl, c2);
            }
        } else if (has1) {
            for (int c = 0; c < C; c++) {
                f16x8 bh, bl, ah, al, xh, xl;
                load_b2(twp + 16 * c, bh, bl);
                load_a2(ia0 + 32 * c, ah, al);
                load_a2(ia1 + 32 * c, xh, xl);
                c0 = mfma3(ah, al, bh, bl, c0);
                c1 = mfma3(xh, xl, bh, bl, c1);
            }
        } else {
            for (int c = 0; c < C; c++) {
                f16x8 bh, bl, ah, al;
                load_b2(twp + 16 * c, bh, bl);
                load_a2(ia0 + 32 * c, ah, al);
                c0 = mfma3(ah, al, bh, bl, c0);
            }
        }
        k[it][0] = c0;
        k[it][1] = c1;
        k[it][2] = c2;
    }
    store_rot<3, true>(k, tiles, np, off, out, mask, x0, W, rinv);
}

// ---------------------------------------------------------------------------
// FFT convolution for wide rotation kernels (the reference convolves with an f32 FFT,
// src/convolve_rv.rs:121-157,484-529).  The linear convolution y = x * a of a tile's
// interpolated range x (at most FFT_N samples) with the kernel a (K taps) is read off a
// circular one of length FFT_N wherever no index wraps: y[n] for K - 1 <= n < FFT_N.  The
// real transforms of length FFT_N run as complex transforms of FFT_M = FFT_N / 2 points
// z[n] = x[2n] + i x[2n+1] (the standard split into even and odd samples), in place in LDS
// on NT = 256 threads: a Stockham transform with radices 8, 8, 8, 4, every thread holding
// one radix-8 (or two radix-4) butterflies per pass, the passes separated by LDS barriers.
// Twiddles come from sincospif (correctly rounded argument scaling), the higher powers of
// a pass's twiddle from one or two products of accurate ones.  The kernel's spectrum is
// computed once per evaluation (k_rot_spectrum) into its tap slot.
// ---------------------------------------------------------------------------
constexpr int FFT_SLOTS = FFT_M + FFT_M / 16;  // padded complex slots
static_assert(2 * FFT_SLOTS <= ILDS, "FFT buffer within the rotation stage's LDS");
static_assert(2 * FFT_SLOTS + RLDS <= LDS_ROT && TILE + 1 <= 2 * FFT_SLOTS,
              "FFT path: R behind the transform buffer, C in it");
static_assert(2 * (FFT_M + 1) <= ROT_STRIDE, "kernel spectrum within an evaluation's tap slot");
static_assert(NT == 256 && FFT_M == 8 * NT, "one radix-8 butterfly per thread and pass");

struct cf {
    float x, y;
};
__device__ __forceinline__ cf cadd(cf a, cf b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf csub(cf a, cf b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cf cmul(cf a, cf b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cf cconj(cf a) { return {a.x, -a.y}; }
__device__ __forceinline__ cf cmi(cf a) { return {a.y, -a.x}; }  // a * (-i)
__device__ __forceinline__ cf cscale(cf a, float s) { return {a.x * s, a.y * s}; }
// e^{i pi t}
__device__ __forceinline__ cf cexpi(float t) {
    float s, c;
    sincospif(t, &s, &c);
    return {c, s};
}
// padded slot of complex index i: one pad slot after every 16 spreads the strided
// butterfly stores over the LDS banks
__device__ __forceinline__ int fpad(int i) { return i + (i >> 4); }

__device__ __forceinline__ void dft2(cf& a, cf& b) {
    const cf t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

// forward 8-point DFT in place (decimation in frequency), natural-order output
__device__ __forceinline__ void dft8(cf (&v)[8]) {
    const float r = 0.707106781186547524f;
    dft2(v[0], v[4]); dft2(v[1], v[5]); dft2(v[2], v[6]); dft2(v[3], v[7]);
    v[5] = {(v[5].x + v[5].y) * r, (v[5].y - v[5].x) * r};     // * e^{-i pi/4}
    v[6] = cmi(v[6]);                                          // * e^{-i pi/2}
    v[7] = {(v[7].y - v[7].x) * r, -(v[7].x + v[7].y) * r};    // * e^{-3i pi/4}
    dft2(v[0], v[2]); dft2(v[1], v[3]); dft2(v[4], v[6]); dft2(v[5], v[7]);
    v[3] = cmi(v[3]);
    v[7] = cmi(v[7]);
    dft2(v[0], v[1]); dft2(v[2], v[3]); dft2(v[4], v[5]); dft2(v[6], v[7]);
    const cf o1 = v[4], o2 = v[2], o3 = v[6], o4 = v[1], o6 = v[3];
    v[1] = o1; v[2] = o2; v[3] = o3; v[4] = o4; v[6] = o6;
}

__device__ __forceinline__ void dft4(cf (&v)[4]) {
    dft2(v[0], v[2]); dft2(v[1], v[3]);
    v[3] = cmi(v[3]);
    dft2(v[0], v[1]); dft2(v[2], v[3]);
    const cf t = v[1];
    v[1] = v[2];
    v[2] = t;
}

// Forward DFT of FFT_M points in place: input buf[i] (buf[fpad(i)] when PADDED_IN), output
// buf[fpad(k)], natural order.  Passes (Ns, R) = (1, 8), (8, 8), (64, 8), (512, 4): thread j
// reads in[j + r FFT_M / R], applies e^{-2 pi i (j mod Ns) r / (Ns R)}, transforms, and
// writes out[(j / Ns) Ns R + j mod Ns + r Ns].  Starts and ends with an LDS barrier.
template <bool PADDED_IN>
__device__ void fft_m(cf* buf) {
    const int j = threadIdx.x;
    cf v[8];
    lds_barrier();
#pragma unroll
    for (int r = 0; r < 8; r++) v[r] = buf[PADDED_IN ? fpad(j + NT * r) : j + NT * r];
    dft8(v);
    lds_barrier();
#pragma unroll
    for (int r = 0; r < 8; r++) buf[fpad(8 * j + r)] = v[r];
#pragma unroll 1
    for (int Ns = 8; Ns <= 64; Ns *= 8) {
        lds_barrier();
#pragma unroll
        for (int r = 0; r < 8; r++) v[r] = buf[fpad(j + NT * r)];
        const int jm = j & (Ns - 1);
        const float t = -2.0f * (float)jm / (float)(8 * Ns);  // exact: a dyadic rational
        const cf w1 = cexpi(t), w2 = cexpi(2.0f * t), w4 = cexpi(4.0f * t);
        const cf w3 = cmul(w1, w2);
        v[1] = cmul(v[1], w1);
        v[2] = cmul(v[2], w2);
        v[3] = cmul(v[3], w3);
        v[4] = cmul(v[4], w4);
        v[5] = cmul(v[5], cmul(w1, w4));
        v[6] = cmul(v[6], cmul(w2, w4));
        v[7] = cmul(v[7], cmul(w3, w4));
        dft8(v);
        lds_barrier();
        const int d = (j / Ns) * Ns * 8 + jm;
#pragma unroll
        for (int r = 0; r < 8; r++) buf[fpad(d + Ns * r)] = v[r];
    }
    lds_barrier();
    cf u[2][4];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int jj = j + NT * h;  // Ns = 512 > jj: no index expansion
#pragma unroll
        for (int r = 0; r < 4; r++) u[h][r] = buf[fpad(jj + 512 * r)];
        const float t = -2.0f * (float)jj / 2048.0f;
        const cf w1 = cexpi(t), w2 = cexpi(2.0f * t);
        u[h][1] = cmul(u[h][1], w1);
        u[h][2] = cmul(u[h][2], w2);
        u[h][3] = cmul(u[h][3], cmul(w1, w2));
        dft4(u[h]);
    }
    lds_barrier();
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int r = 0; r < 4; r++) buf[fpad(j + NT * h + 512 * r)] = u[h][r];
    lds_barrier();
}

// The even/odd split of a real transform.  From Z = DFT_M(z), the spectrum of x at k and
// M - k (0 < k < M): E = (Z[k] + conj Z[M-k]) / 2, O = (Z[k] - conj Z[M-k]) / 2i,
// X[k] = E + W^k O, X[M-k] = conj(E - W^k O), W = e^{-2 pi i / FFT_N}.
__device__ __forceinline__ void real_split(cf zk, cf zm, cf w, cf& xk, cf& xm) {
    const cf e = cscale(cadd(zk, cconj(zm)), 0.5f);
    const cf o = cscale(cmi(csub(zk, cconj(zm))), 0.5f);
    const cf wo = cmul(w, o);
    xk = cadd(e, wo);
    xm = cconj(csub(e, wo));
}

// Its inverse for the half spectrum Y of a real sequence y: Z'[k] = (Y[k] + conj Y[M-k])
// + i (Y[k] - conj Y[M-k]) W^-k has IDFT_M(Z')[n] = y[2n] + i y[2n+1]; stored conjugated,
// so that the forward transform computes the inverse.
__device__ __forceinline__ void real_join(cf yk, cf ym, cf w, cf& zk, cf& zm) {
    const cf a = cadd(yk, cconj(ym));
    const cf b = cmul(csub(yk, cconj(ym)), cconj(w));
    zk = cconj(cadd(a, {-b.y, b.x}));                // a + i b
    // the same at M - k: W^-(M-k) = -W^k
    const cf a2 = cadd(ym, cconj(yk));
    const cf b2 = cmul(csub(ym, cconj(yk)), {-w.x, -w.y});
    zm = cconj(cadd(a2, {-b2.y, b2.x}));
}

// Frequencies of thread t in the split steps: k = t + NT i (i < 4) paired with M - k, and
// k = M / 2 on thread 0; W^k = e^{-i pi k / M}.
__device__ __forceinline__ cf split_twiddle(int t, int i) {
    return cexpi(-(float)(t + NT * i) / (float)FFT_M);
}

// Spectrum A[k], k <= M, of the K taps h (zero-padded to FFT_N) into spec (global).
__device__ void kernel_spectrum(const float* __restrict__ h, int K, cf* buf, cf* spec) {
    float* xb = (float*)buf;
    for (int i = threadIdx.x; i < FFT_N; i += NT) xb[i] = i < K ? h[i] : 0.0f;
    fft_m<false>(buf);
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int k = t + NT * i;
        const cf zk = buf[fpad(k)], zm = buf[fpad((FFT_M - k) & (FFT_M - 1))];
        if (k == 0) {
            spec[0] = {zk.x + zk.y, 0.0f};
            spec[FFT_M] = {zk.x - zk.y, 0.0f};
        } else {
            cf xk, xm;
            real_split(zk, zm, split_twiddle(t, i), xk, xm);
            spec[k] = xk;
            spec[FFT_M - k] = xm;
        }
    }
    if (t == 0) {
        const cf z = buf[fpad(FFT_M / 2)];
        cf xk, xm;
        real_split(z, z, cexpi(-0.5f), xk, xm);
        spec[FFT_M / 2] = xk;
    }
}

// out[q] = sum_u h[u] x[off + q + K - 1 - u] (the rotation stage's "same" convolution of the
// tile, x at logical index 0 = the first interpolated pixel) for q < nR, zero outside
// [0, W) when mask.  xb: the tile as FFT_N floats (zero past the interpolated range), which
// the transform overwrites; spec: the kernel's spectrum.
__device__ void conv_fft(float* xb, const cf* __restrict__ spec, int K, int off, int nR,
                         float* out, bool mask, long long x0, int W) {
    cf* buf = (cf*)xb;
    const int t = threadIdx.x;
    // the spectrum's values for this thread's frequencies, read now and used after the
    // forward transform (the LDS barriers do not wait for them)
    cf ak[4], am[4], ah = {0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int k = t + NT * i;
        ak[i] = spec[k];
        am[i] = spec[FFT_M - k];
    }
    if (t == 0) ah = spec[FFT_M / 2];
    fft_m<false>(buf);
    const float s = 1.0f / (float)FFT_N;  // the inverse transform's scale, exact
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int k = t + NT * i;
        const int km = (FFT_M - k) & (FFT_M - 1);
        const cf zk = buf[fpad(k)], zm = buf[fpad(km)];
        if (k == 0) {
            // X[0] and X[M] are real: Re Z[0] +- Im Z[0]
            const float y0 = (zk.x + zk.y) * ak[i].x * s, yM = (zk.x - zk.y) * am[i].x * s;
            buf[fpad(0)] = cconj({y0 + yM, y0 - yM});
        } else {
            const cf w = split_twiddle(t, i);
            cf xk, xm, zk2, zm2;
            real_split(zk, zm, w, xk, xm);
            real_join(cscale(cmul(xk, ak[i]), s), cscale(cmul(xm, am[i]), s), w, zk2, zm2);
            buf[fpad(k)] = zk2;
            buf[fpad(km)] = zm2;
        }
    }
    if (t == 0) {
        const cf z = buf[fpad(FFT_M / 2)];
        const cf w = cexpi(-0.5f);
        cf xk, xm, zk2, zm2;
        real_split(z, z, w, xk, xm);
        const cf y = cscale(cmul(xk, ah), s);
        real_join(y, y, w, zk2, zm2);
        buf[fpad(FFT_M / 2)] = zk2;
    }
    fft_m<true>(buf);
    // IDFT_M(Z')[n] = conj(DFT_M(conj Z'))[n] = y[2n] + i y[2n+1]
    const int base = off + K - 1;
    for (int q = t; q < nR; q += NT) {
        const int n = base + q;
        const cf z = buf[fpad(n >> 1)];
        const float y = (n & 1) ? -z.y : z.x;
        const long long x = x0 + q;
        out[q] = (mask && (x < 0 || x >= W)) ? 0.0f : y;
    }
}

// the kernel spectra of the evaluations on the FFT path, into their tap slots
__global__ __launch_bounds__(NT) void k_rot_spectrum(const DevEval* __restrict__ evals,
                                                     float* __restrict__ rot) {
    const DevEval& ev = evals[blockIdx.x];
    if (ev.status != ST_OK || !ev.fft) return;
    __shared__ __align__(16) cf buf[FFT_SLOTS];
    float* h = rot + (size_t)blockIdx.x * ROT_STRIDE;
    kernel_spectrum(h, ev.krot, buf, (cf*)h);
}

// ---------------------------------------------------------------------------
// k_model: one workgroup per (work item, synthetic tile).  A work item is one
// evaluation, or two or three evaluations that gather the same 64 neighbour rows: their
// interpolation shares every 16-byte gather and u16 conversion (one conversion and one
// FMA per evaluation per pixel-row), then each runs its own rotation, instrumental
// broadening and resample.  Each evaluation's arithmetic is the same as alone.
// ---------------------------------------------------------------------------
struct TileJob {
    const DevEval* ev;
    const DevStar* S;
    int e;           // evaluation
    int ilo, ihi;    // observed pixels of the tile (mode 0)
    long long clo;   // first synthetic pixel of the convolved range C
    long long rlo;   // first pixel of the rotation output R (clo - hd)
    int nC, hd, K, nR;
    bool mask;
    int kind;        // 0 no rotation (the tile is R), 1 f16 rotation, 2 f32 rotation, 3 FFT
    int inst;        // instrumental stage: DISP_* of the star, DISP_NONE when composed
    long long g0;    // first interpolated 8-pixel group (multiple of 8)
    int off, ng;     // logical index of rlo - K/2 (or rlo); groups interpolated
};

__device__ __forceinline__ bool tile_job(const DevGrid& G, const DevStar* stars, const DevEval* evals,
                                         int e, int tile, const int* tbound, int tb_stride, int mode,
                                         TileJob& J) {
    // branch-free: every read is issued before the job's validity is decided, so the reads of
    // several jobs overlap (a failed evaluation's star index and tile-bound row exist; its
    // bounds are clamped into the star's pixels)
    const DevEval& ev = evals[e];
    const DevStar& S = stars[ev.star];
    bool ok = ev.status == ST_OK && tile < S.n_tiles;
    J.ev = &ev;
    J.S = &S;
    J.e = e;
    const double sl = ev.shift_log;
    const long long t0 = (long long)(ev.tile_base + tile) * TILE;
    long long chi;
    if (mode == 0) {
        const int* tb = tbound + (size_t)e * tb_stride + tile;
        J.ilo = __builtin_amdgcn_readfirstlane(tb[0]);
        J.ihi = __builtin_amdgcn_readfirstlane(tb[1]);
        const int nm = S.N - 1;
        const int i0 = min(max(J.ilo, 0), nm), i1 = min(max(J.ihi - 1, 0), nm);
        const double f0 = S.fidx[i0], f1 = S.fidx[i1];
        ok = ok && J.ilo < J.ihi;
        J.clo = jfloor(f0 + sl);
        chi = jfloor(f1 + sl) + 1;
    } else {
        ok = ok && t0 < G.W;
        J.ilo = J.ihi = 0;
        J.clo = t0;
        chi = (t0 + TILE < G.W ? t0 + TILE : G.W) - 1;
    }
    J.nC = (int)(chi - J.clo + 1);
    J.inst = ev.comp ? (int)DISP_NONE : S.disp_kind;
    J.hd = (J.inst == DISP_NONE) ? 0 : (S.Kd - 1) / 2;
    J.K = ev.krot;
    const int hr = J.K > 0 ? J.K / 2 : 0;
    J.rlo = J.clo - J.hd;
    J.nR = J.nC + 2 * J.hd;
    J.mask = J.inst != DISP_NONE;
    J.kind = J.K == 0 ? 0 : (ev.fft ? 3 : (J.K <= KB_MAX ? 1 : 2));
    const long long first = J.rlo - hr;
    J.g0 = first & ~7LL;
    J.off = (int)(first - J.g0);
    J.ng = (J.off + J.nR + 2 * hr + 7) >> 3;
    return ok;
}

// store one interpolated 8-pixel group (local group index gl) in the job's tile layout
__device__ __forceinline__ void store_group(const TileJob& J, int gl, const float (&acc)[8],
                                            float* Ib) {
    if (J.kind == 1) {
        const float sc = J.ev->rsc;
        u16x8 h, l;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            unsigned short a, b;
            split2(acc[t] * sc, a, b);
            h[t] = a;
            l[t] = b;
        }
        unsigned short* d = (unsigned short*)Ib + 8 * gl;
        *(u16x8*)d = h;
        *(u16x8*)(d + ILOGB) = l;
    } else {
        float* d = J.kind == 2 ? Ib + islot(8 * gl) : Ib + 8 * gl;
        *(f4*)d = f4{acc[0], acc[1], acc[2], acc[3]};
        *(f4*)(d + 4) = f4{acc[4], acc[5], acc[6], acc[7]};
    }
}

// interpolated values of the group at pixel x0 for one evaluation
__device__ __forceinline__ void gather_one(const DevGrid& G, const DevEval& ev, long long x0,
                                           float (&acc)[8]) {
    const bool inside = x0 >= 0 && x0 < G.W;
    const int voff = inside ? (int)(2 * x0) : OOB_OFFSET;
#pragma unroll
    for (int t = 0; t < 8; t++) acc[t] = 0.0f;
    // all 8 pixels of every lane's group before the n % 128 tail: no quirk handling
    if (__all(x0 + 7 < G.Wq)) gather_rows<true>(G, ev, voff, x0, acc);
    else gather_rows<false>(G, ev, voff, x0, acc);
    if (x0 + 8 > G.W) {
#pragma unroll
        for (int t = 0; t < 8; t++)
            if (x0 + t >= G.W) acc[t] = 0.0f;
    }
}

// GN evaluations over the union of their rows: one gather and conversion per union row, GN
// weight sets (0 where a row is not an evaluation's neighbour: fmaf(q, 0, acc) == acc, as no
// partial sum is -0); each evaluation's nonzero terms run in the same order as alone
template <int GN, bool FAST>
__device__ __forceinline__ void gather_rows_g(const DevGrid& G, const DevUnion& U, int voff,
                                              long long x0, float (&acc)[GN][8]) {
    const int nrec = (int)(G.pitch * 2);
    constexpr int RF = 8;  // rows in flight (registers)
    const int n8 = __builtin_amdgcn_readfirstlane(U.n);  // a multiple of 8
#pragma unroll 1
    for (int j0 = 0; j0 < n8; j0 += RF) {
        u32x4 v[RF];
#pragma unroll
        for (int jj = 0; jj < RF; jj++) {
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                (void*)U.rowp[j0 + jj], 0, nrec, 0x00020000);
            v[jj] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, 0, 0));
        }
#pragma unroll
        for (int jj = 0; jj < RF; jj++) {
            float f[GN];
#pragma unroll
            for (int g = 0; g < GN; g++)
                f[g] = __builtin_bit_cast(
                    float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, U.fw[g][j0 + jj])));
            float u[8];
            unpack8(v[jj], u);
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const float q = (FAST || x0 + t < G.Wq) ? u[t] : u[t] / 65535.0f;  // tail quirk :978
#pragma unroll
                for (int g = 0; g < GN; g++) acc[g][t] = fmaf(q, f[g], acc[g][t]);
            }
        }
    }
}

// taps of a job into LDS (and the zero padding the MFMA steps read past its tile); issued
// before the interpolation so that their latency overlaps it
// 16-byte quads of each of the four split-tap arrays that the f16 MFMA groups of a K-tap
// kernel read: words [0, 16 * ceil((K + 15) / 32) + 8), from the highest B operand index
// 20 + 16 (C - 1) + 3 of conv_rot_f16
__device__ __forceinline__ int tap_quads(int K) { return (16 * ((K + 46) >> 5) + 8) >> 2; }

// the used quads of a K-tap kernel's split taps, as (array, quad) -> index in the arrays
__device__ __forceinline__ int tap_quad_index(int i, int nq4) {
    const int a = i / nq4;
    return a * (TW_MAX / 4) + (i - a * nq4);
}

__device__ __forceinline__ void load_taps(const TileJob& J, const float* rot, float* Ib,
                                          float* hrev, unsigned int* tw, float* grev,
                                          bool with_tw = true) {
    const int tid = threadIdx.x;
    const DevStar& S = *J.S;
    if (J.kind == 1) {  // split taps prepared per evaluation by k_prep
        if (with_tw) {
            const u32x4* src = (const u32x4*)(rot + (size_t)J.e * ROT_STRIDE + KROT_MAX);
            const int nq4 = tap_quads(J.K);
            for (int i = tid; i < 4 * nq4; i += NT) {
                const int q = tap_quad_index(i, nq4);
                ((u32x4*)tw)[q] = src[q];
            }
        }
        unsigned short* Ib2 = (unsigned short*)Ib;
        if (tid < 64) {  // the MFMA groups read past the tile
            Ib2[8 * J.ng + tid] = 0;
            Ib2[ILOGB + 8 * J.ng + tid] = 0;
        }
    } else if (J.kind == 2) {
        const float* h = rot + (size_t)J.e * ROT_STRIDE;
        const int K = J.K;
        const int Kp = 16 * ((K + 30) >> 4) + 16;  // every index the MFMA steps read
        for (int i = tid; i < Kp; i += NT) {
            const int u = i - 16;
            hrev[i] = (u >= 0 && u < K) ? h[K - 1 - u] : 0.0f;
        }
        if (tid < 48) Ib[islot(8 * J.ng + tid)] = 0.0f;  // the MFMA steps read past the tile
    }
    if (J.inst == DISP_FIXED) {
        const int Kp = (S.Kd + 7) & ~7;
        for (int i = tid; i < Kp; i += NT) grev[i] = i < S.Kd ? S.dkern[S.Kd - 1 - i] : 0.0f;
    }
}

// RV shift + resample (f32-quantised weight, convolve_rv.rs:569-573) of the job's observed
// pixels from the convolved tile Cb.  floor(idx) clamped at 0 (`as usize`) is an integer below
// 2^31: exact in f64 and in a 32-bit int, and (float) of it rounds the same integer as (float)
// of a 64-bit index.  Four float indices in flight per thread (global loads, not waited on by
// LDS reads).
template <bool STREAM>
__device__ __forceinline__ void resample(const TileJob& J, const DevStar& S, const float* Cb,
                                         float* out) {
    const int tid = threadIdx.x;
    const double sl = J.ev->shift_log;
    const int clo = (int)J.clo, ihi = J.ihi;
    const GlobalD* fx = (const GlobalD*)S.fidx;
    for (int i0 = J.ilo + tid; i0 < ihi; i0 += 4 * NT) {
        double idx[4];
#pragma unroll
        for (int u = 0; u < 4; u++) idx[u] = i0 + u * NT < ihi ? fx[i0 + u * NT] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * NT;
            if (i < ihi) {
                const double x = idx[u] + sl;
                const double f = floor(x);
                const double jf = f > 0.0 ? f : 0.0;
                const int j = (int)jf - clo;
                const float w = (float)x - (float)jf;
                const float a = Cb[j], c = Cb[j + 1];
                const float v = (1.0f - w) * a + w * c;
                if (STREAM) __builtin_nontemporal_store(v, out + i);
                else out[i] = v;
            }
        }
    }
}

// rotation, instrumental broadening and resample of a job whose interpolated tile and
// taps are in LDS (stored but not yet synchronised)
template <bool FFT = false>
__device__ __forceinline__ void finish_job(const DevGrid& G, const TileJob& J, const float* rot,
                                           float* model, int mode, float* Ib, float* hrev,
                                           unsigned int* tw, float* grev) {
    const int tid = threadIdx.x;
    const DevStar& S = *J.S;
    lds_barrier();
    // rotational broadening, outputs p = off + q for R index q (stored over the operands)
    float* R = Ib;
    if (J.kind == 1)
        conv_rot_f16((const unsigned short*)Ib, tw, J.K, J.off + J.nR, J.off, R, J.mask, J.rlo, G.W,
                     J.ev->rinv);
    else if (J.kind == 2)
        conv_rot_mfma(Ib, hrev, J.K, J.off + J.nR, J.off, R, J.mask, J.rlo, G.W);
    else if (FFT && J.kind == 3)
        R = Ib + 2 * FFT_SLOTS;  // formed piece by piece (fft_pieces) behind the transform buffer
    else
        R = Ib + J.off;  // no rotation: the interpolated tile itself
    if (tid < 16) R[J.nR + tid] = 0.0f;
    lds_barrier();

    // instrumental broadening -> C (behind R, or in the FFT path's transform buffer)
    const int nC = J.nC, hd = J.hd;
    const float* Cb = R;
    float* const Cw = (FFT && J.kind == 3) ? Ib : Ib + C_OFF;
    if (J.inst == DISP_FIXED) {
        conv_tile(R, grev, S.Kd, nC, Cw, false, 0, 0);
        Cb = Cw;
        lds_barrier();
    } else if (J.inst == DISP_VARIABLE) {
        const int Kd = S.Kd;
        for (int q = tid; q < nC; q += NT) {
            const long long x = J.clo + q;
            float c;
            if (x < hd || x >= G.W - hd) {
                c = R[q + hd];
            } else {
                const float* kr = S.dkern + x * Kd;
                float a = 0.0f;
                for (int u = 0; u < Kd; u++) a = fmaf(kr[u], R[q + u], a);
                c = a;
            }
            Cw[q] = c;
        }
        Cb = Cw;
        lds_barrier();
    }

    // RV shift + resample (f32-quantised weight, convolve_rv.rs:569-573)
    float* out = model + J.ev->m_off;
    if (mode == 1) {
        for (int q = tid; q < nC; q += NT) out[J.clo + q] = Cb[q];
        return;
    }
    // Models larger than the MALL are read back from HBM anyway: streamed past the caches
    // (non-temporal stores), they do not evict the grid rows the interpolation re-reads.
    if (G.stream_models) resample<true>(J, S, Cb, out);
    else resample<false>(J, S, Cb, out);
}

// The FFT path's rotation output R behind the transform buffer: the tile's output range in
// one transform, or in two pieces (fft_piece) when the range and the kernel do not fit one.
// A piece starting at output q0 (a multiple of 8) transforms the interpolated samples from
// q0 on: shifting the input shifts the output, so its outputs are R[q0 ..].  Each piece
// gathers its own input (an FFT evaluation's tile is interpolated once more where the
// pieces overlap, by the kernel's width).
__device__ __forceinline__ void fft_pieces(const DevGrid& G, const TileJob& J, const float* rot,
                                           float* Ib) {
    const int tid = threadIdx.x;
    const int P = J.off + J.nR + J.K - 1 <= FFT_N ? J.nR : fft_piece(J.nR);
    const cf* spec = (const cf*)(rot + (size_t)J.e * ROT_STRIDE);
    float* R = Ib + 2 * FFT_SLOTS;
    for (int q0 = 0; q0 < J.nR; q0 += P) {
        const int nq = min(P, J.nR - q0);
        const int gq = q0 >> 3;
        const int ngq = min(J.ng - gq, (J.off + nq + J.K - 1 + 7) >> 3);
        if (q0 > 0) lds_barrier();  // the previous piece's transform has been read
        for (int gg = tid; gg < ngq; gg += NT) {
            float acc[8];
            gather_one(G, *J.ev, J.g0 + 8LL * (gq + gg), acc);
            *(f4*)(Ib + 8 * gg) = f4{acc[0], acc[1], acc[2], acc[3]};
            *(f4*)(Ib + 8 * gg + 4) = f4{acc[4], acc[5], acc[6], acc[7]};
        }
        for (int i = 8 * ngq + tid; i < FFT_N; i += NT) Ib[i] = 0.0f;  // the transform's input
        lds_barrier();
        conv_fft(Ib, spec, J.K, J.off, nq, R + q0, J.mask, J.rlo + q0, G.W);
    }
}

// evaluation e alone on this tile (its job exists; the LDS is free)
template <bool FFT = false>
__device__ __forceinline__ void run_solo(const DevGrid& G, const DevStar* stars, const DevEval* evals,
                                         const float* rot, int e, int tile, const int* tbound,
                                         int tb_stride, int mode, float* model, float* Ib,
                                         float* hrev, unsigned int* tw, float* grev) {
    TileJob B;
    tile_job(G, stars, evals, e, tile, tbound, tb_stride, mode, B);
    load_taps(B, rot, Ib, hrev, tw, grev);
    if (FFT && B.kind == 3) {
        fft_pieces(G, B, rot, Ib);
    } else {
        for (int gg = threadIdx.x; gg < B.ng; gg += NT) {
            float acc[8];
            gather_one(G, *B.ev, B.g0 + 8LL * gg, acc);
            store_group(B, gg, acc, Ib);
        }
    }
    finish_job<FFT>(G, B, rot, model, mode, Ib, hrev, tw, grev);
}

// A work item of GN (2 or 3) evaluations whose rows are the union U (k_union) and whose tiles
// span at most 2 NT 8-pixel groups together: one shared interpolation, A's groups to LDS at once,
// the others' kept in registers (and for a pair the second one's split taps, read now), then each
// evaluation's rotation, instrumental broadening and resample in turn.  A's taps are in LDS.
template <int GN>
__device__ __forceinline__ void run_group(const DevGrid& G, const DevStar* stars, const DevEval* evals,
                                          const float* rot, int tile, const int* tbound,
                                          int tb_stride, int mode, float* model, const TileJob& A,
                                          int e1, int e2, const DevUnion& U, long long pg0, int png,
                                          float* Ib, float* hrev, unsigned int* tw,
                                          float* grev) {
    // with three evaluations the registers go to the kept interpolations instead
    const DevEval& eb = evals[e1];
    // (kernels of up to 640 taps: their split taps are 2 rounds of 16-byte loads)
    constexpr int KB_PREFETCH = 640;
    const bool btw = GN == 2 && eb.krot > 0 && eb.krot <= KB_PREFETCH && !eb.fft;
    const int bq4 = btw ? tap_quads(eb.krot) : 0;
    u32x4 tkeep[2];
    static_assert(4 * ((16 * ((KB_PREFETCH + 46) >> 5) + 8) >> 2) <= 2 * NT, "split taps in 2 rounds");
    if (GN == 2) {
        const u32x4* src = (const u32x4*)(rot + (size_t)e1 * ROT_STRIDE + KROT_MAX);
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const int i = threadIdx.x + r * NT;
            tkeep[r] = i < 4 * bq4 ? src[tap_quad_index(i, bq4)] : u32x4{0u, 0u, 0u, 0u};
        }
    }
    float keep[GN - 1][2][8];
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int gg = threadIdx.x + r * NT;
        float acc[GN][8];
#pragma unroll
        for (int g = 0; g < GN; g++)
#pragma unroll
            for (int t = 0; t < 8; t++) acc[g][t] = 0.0f;
        if (gg < png) {
            const long long x0 = pg0 + 8LL * gg;
            const bool inside = x0 >= 0 && x0 < G.W;
            const int voff = inside ? (int)(2 * x0) : OOB_OFFSET;
            if (__all(x0 + 7 < G.Wq)) gather_rows_g<GN, true>(G, U, voff, x0, acc);
            else gather_rows_g<GN, false>(G, U, voff, x0, acc);
            if (x0 + 8 > G.W) {
#pragma unroll
                for (int t = 0; t < 8; t++)
                    if (x0 + t >= G.W)
#pragma unroll
                        for (int g = 0; g < GN; g++) acc[g][t] = 0.0f;
            }
            const long long ga = (x0 - A.g0) >> 3;
            if (ga >= 0 && ga < A.ng) store_group(A, (int)ga, acc[0], Ib);
        }
#pragma unroll
        for (int g = 1; g < GN; g++)
#pragma unroll
            for (int t = 0; t < 8; t++) keep[g - 1][r][t] = acc[g][t];
    }
    finish_job(G, A, rot, model, mode, Ib, hrev, tw, grev);
#pragma unroll
    for (int g = 1; g < GN; g++) {
        lds_barrier();
        TileJob B;
        tile_job(G, stars, evals, g == 1 ? e1 : e2, tile, tbound, tb_stride, mode, B);
        if (g == 1 && btw) {  // the second evaluation's split taps, read before A's rotation
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const int i = threadIdx.x + r * NT;
                if (i < 4 * bq4) ((u32x4*)tw)[tap_quad_index(i, bq4)] = tkeep[r];
            }
            load_taps(B, rot, Ib, hrev, tw, grev, false);
        } else {
            load_taps(B, rot, Ib, hrev, tw, grev);
        }
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const int gg = threadIdx.x + r * NT;
            const long long gb = (pg0 + 8LL * gg - B.g0) >> 3;
            if (gg < png && gb >= 0 && gb < B.ng) store_group(B, (int)gb, keep[g - 1][r], Ib);
        }
        finish_job(G, B, rot, model, mode, Ib, hrev, tw, grev);
    }
}

template <bool FFT>
__global__ __launch_bounds__(NT, 4) void k_model(DevGrid G, const DevStar* __restrict__ stars,
                                              const DevEval* __restrict__ evals,
                                              const float* __restrict__ rot,
                                              const int4* __restrict__ items,
                                              const int* __restrict__ n_items,
                                              const DevUnion* __restrict__ unions, int n_evals,
                                              const int* __restrict__ tbound, int tb_stride,
                                              float* __restrict__ model, int mode) {
    // Rotation stage storage, shared by its two forms: f32 (padded tile Ib, reversed taps
    // hrev: zero, then the taps from index 16, zeros after) or f16 x 2 (tile arrays,
    // split tap pairs tw).  Ib also holds the instrumental-broadening output.
    static_assert(2 * ILOGB * 2 + 4 * TW_MAX * 4 <= (ILDS + KROT_MAX + 48) * 4, "LDS union");
    __shared__ __align__(16) float lds_rot[LDS_ROT];
    __shared__ float grev[2 * HD_MAX + 1 + 8];
    float* const Ib = lds_rot;
    float* const hrev = lds_rot + ILDS;
    unsigned int* const tw = (unsigned int*)(lds_rot + ILOGB);

    // tile-major over the work items (in neighbour-row order): the workgroups resident on
    // one XCD work on one tile of evaluations that gather overlapping rows
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int tile = b / n_evals;
    const int k = b - tile * n_evals;
    if (k >= __builtin_amdgcn_readfirstlane(*n_items)) return;
    // the work item's evaluations are uniform: keep them (and so their weight loads) scalar
    const int ids[3] = {__builtin_amdgcn_readfirstlane(items[k].x),
                        __builtin_amdgcn_readfirstlane(items[k].y),
                        __builtin_amdgcn_readfirstlane(items[k].z)};
    TileJob A;
    const bool hasA = tile_job(G, stars, evals, ids[0], tile, tbound, tb_stride, mode, A);
    // the other evaluations' interpolated ranges (their jobs are formed again when they run)
    bool hasB = false, hasC = false;
    long long bg0 = 0, cg0 = 0;
    int bng = 0, cng = 0;
    {  // formed unconditionally (an absent evaluation repeats the first) so their reads overlap
        TileJob T;
        hasB = tile_job(G, stars, evals, ids[1] >= 0 ? ids[1] : ids[0], tile, tbound, tb_stride,
                        mode, T) && ids[1] >= 0;
        bg0 = T.g0;
        bng = T.ng;
        TileJob T2;
        hasC = tile_job(G, stars, evals, ids[2] >= 0 ? ids[2] : ids[0], tile, tbound, tb_stride,
                        mode, T2) && ids[2] >= 0;
        cg0 = T2.g0;
        cng = T2.ng;
    }
    // A work item with an evaluation on the FFT path runs its evaluations one at a time
    // through the one copy of the FFT stage (its registers would not fit beside a group's)
    if (FFT && ((hasA && A.kind == 3) || (hasB && evals[ids[1]].fft) || (hasC && evals[ids[2]].fft))) {
        bool first = true;
        for (int g = 0; g < 3; g++) {
            if (!(g == 0 ? hasA : g == 1 ? hasB : hasC)) continue;
            if (!first) lds_barrier();
            first = false;
            run_solo<true>(G, stars, evals, rot, ids[g], tile, tbound, tb_stride, mode, model, Ib,
                           hrev, tw, grev);
        }
        return;
    }
    if (hasA) load_taps(A, rot, Ib, hrev, tw, grev);
    const DevUnion& U = unions[k];
    if (hasA && hasB && __builtin_amdgcn_readfirstlane(U.n) >= 0) {
        long long pg0 = min(A.g0, bg0), pg1 = max(A.g0 + 8LL * A.ng, bg0 + 8LL * bng);
        if (hasC) {
            const long long q0 = min(pg0, cg0), q1 = max(pg1, cg0 + 8LL * cng);
            if (((q1 - q0) >> 3) <= 2 * NT) {
                run_group<3>(G, stars, evals, rot, tile, tbound, tb_stride, mode, model, A, ids[1],
                             ids[2], U, q0, (int)((q1 - q0) >> 3), Ib, hrev, tw, grev);
                return;
            }
        } else if (((pg1 - pg0) >> 3) <= 2 * NT) {
            run_group<2>(G, stars, evals, rot, tile, tbound, tb_stride, mode, model, A, ids[1], -1,
                         U, pg0, (int)((pg1 - pg0) >> 3), Ib, hrev, tw, grev);
            return;
        }
    }
    // one evaluation at a time
    if (hasA) {
        for (int gg = threadIdx.x; gg < A.ng; gg += NT) {
            float acc[8];
            gather_one(G, *A.ev, A.g0 + 8LL * gg, acc);
            store_group(A, gg, acc, Ib);
        }
        finish_job(G, A, rot, model, mode, Ib, hrev, tw, grev);
    }
    if (hasB) {
        if (hasA) lds_barrier();
        run_solo(G, stars, evals, rot, ids[1], tile, tbound, tb_stride, mode, model, Ib, hrev,
                 tw, grev);
    }
    if (hasC) {
        if (hasA || hasB) lds_barrier();
        run_solo(G, stars, evals, rot, ids[2], tile, tbound, tb_stride, mode, model, Ib, hrev,
                 tw, grev);
    }
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// Sums v[0..N) over the 64 lanes of a wave by recursive halving: at each step a lane
// keeps one half of its values and adds its partner's copy of that half, so N values
// cost about N exchanges instead of 6N.  Afterwards v[0] of lane l holds the wave sum
// of value halve_index<N>(l) (an index >= N marks a lane that holds only padding).
template <int N, int O>
struct Halve {
    __device__ static __forceinline__ void run(double* v, int lane) {
        constexpr int H = (N + 1) / 2;
        const bool up = (lane & O) != 0;
#pragma unroll
        for (int i = 0; i < H; i++) {
            const double hi = (H + i < N) ? v[H + i] : 0.0;
            const double keep = up ? hi : v[i];
            const double give = up ? v[i] : hi;
            v[i] = keep + __shfl_xor(give, O, 64);
        }
        Halve<H, O / 2>::run(v, lane);
    }
};
template <int N>
struct Halve<N, 0> {
    __device__ static __forceinline__ void run(double*, int) {}
};

template <int N>
__device__ __forceinline__ int halve_index(int lane) {
    int idx = 0, n = N, real = N;  // real: how many of the lane's n values are not padding
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int h = (n + 1) / 2;
        if (lane & o) {
            idx += h;
            real = real > h ? real - h : 0;
        } else {
            real = real < h ? real : h;
        }
        n = h;
    }
    return real > 0 ? idx : N;
}

// Workgroup sums of v[0..N) (N <= 64): the result for value t is returned to thread t < N.
template <int N>
__device__ __forceinline__ double group_sums(double (&v)[N], double* red /* 4*N */) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    Halve<N, 32>::run(v, lane);
    const int idx = halve_index<N>(lane);
    if (idx < N) red[wv * N + idx] = v[0];
    __syncthreads();
    const int t = threadIdx.x;
    return t < N ? ((red[t] + red[N + t]) + red[2 * N + t]) + red[3 * N + t] : 0.0;
}

// f32 design-matrix row of a chunk: A[i,k] = f32(powi(xhat, k)), with the
// squares shared between k exactly as __powidf2 forms them.
template <int PM>
__device__ __forceinline__ void design_row(double xh, int p1, float (&a)[PM]) {
    double sq[5];
    sq[0] = xh;
#pragma unroll
    for (int t = 1; t < 5; t++) sq[t] = sq[t - 1] * sq[t - 1];
#pragma unroll
    for (int k = 0; k < PM; k++) {
        if (k < p1) {
            double r = 1.0;
#pragma unroll
            for (int t = 0; t < 5; t++)
                if (k & (1 << t)) r *= sq[t];
            a[k] = (float)r;
        }
    }
}

// nalgebra 0.32 dot order: 8 partial accumulators, pairwise merge, scalar tail.  The
// accumulators start at -0.0 rather than +0.0: -0 + x == x for every x, so the first add of
// each is exact and the compiler drops it; the result can differ from a +0.0 start only in
// the sign of a zero (a +0.0 start turns -0 into +0), which no use of it can tell apart.
template <int PM>
__device__ __forceinline__ float dot_na(const float (&a)[PM], const float* b, int n) {
    float c[8];
#pragma unroll
    for (int t = 0; t < 8; t++) c[t] = -0.0f;
    const int nb = (n >> 3) << 3;
#pragma unroll
    for (int k = 0; k < PM; k++)
        if (k < nb) c[k & 7] = c[k & 7] + a[k] * b[k];
    float r = -0.0f;
    r = r + (c[0] + c[4]);
    r = r + (c[1] + c[5]);
    r = r + (c[2] + c[6]);
    r = r + (c[3] + c[7]);
#pragma unroll
    for (int k = 0; k < PM; k++)
        if (k >= nb && k < n) r = r + a[k] * b[k];
    return r;
}

// The constant continuum's parameter as the reference forms it
// (continuum_fitting.rs:490-492): the f32 quotients f/m, added one after another in pixel
// order (`target.iter().sum::<f32>()`, a sequential fold), divided by n in f32.  A
// workgroup covers the E <= 16 evaluations of one star in block descriptor (first
// evaluation, count, 0, 0).  The star's pixels go in chunks of 192, one per thread of
// waves 1-3: the thread reads its pixel's flux and the E models' values there, forms the E
// quotients (IEEE divisions) and stores them in a double-buffered LDS tile, one row per
// evaluation; lane e < E of wave 0 adds row e's 192 values in order while the next chunk
// is being formed.  A thread's reads of a chunk (1 + E) are issued four chunks before its
// divisions, unconditionally (at a clamped pixel past the star's end), so no read waits on
// another and the chain of additions sets the pace; the barriers order the LDS tile only.
// Past the last pixel the tile holds zeros: adding +0 leaves a sum unchanged (but for the
// sign of a zero sum).
constexpr int FOLD_CH = NT - 64;  // pixels per chunk: one per producer thread
constexpr int FOLD_SETS = 4;     // register sets of reads in flight per producer thread
constexpr int FOLD_ROW = 196;     // LDS row stride: 16-byte rows, 196 = 4 mod 64 banks, so
                                  // 16 lanes' 16-byte reads cover the 64 banks once
template <int E>
__global__ __launch_bounds__(NT) void k_const_fold(const DevStar* __restrict__ stars,
                                                   const DevEval* __restrict__ evals,
                                                   const int4* __restrict__ blocks,
                                                   const float* __restrict__ model,
                                                   float* __restrict__ pf) {
    static_assert(E <= 16, "one lane of wave 0 per evaluation, <= 16 for conflict-free rows");
    static_assert(FOLD_CH % 4 == 0 && FOLD_ROW >= FOLD_CH, "float4 rows");
    __shared__ __align__(16) float q[2][E][FOLD_ROW];
    __shared__ long long s_off[E];
    __shared__ int s_ok[E];
    const int4 bd = blocks[blockIdx.x];
    const DevStar& S = stars[evals[bd.x].star];
    const int N = S.N, cnt = bd.y;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n_ch = (N + FOLD_CH - 1) / FOLD_CH;
    if (threadIdx.x < E) {
        const int e = threadIdx.x;
        const bool ok = e < cnt && evals[bd.x + e].status == ST_OK;
        s_ok[e] = ok;
        s_off[e] = evals[bd.x + min(e, cnt - 1)].m_off;  // every evaluation owns N values
    }
    __syncthreads();
    GlobalF* flux = (GlobalF*)S.flux;
    GlobalF* gmodel = (GlobalF*)model;
    const int t = threadIdx.x - 64;  // producer thread: pixel t of each chunk
    unsigned okm[E];                  // all ones for an evaluation whose quotients count
    long long offv[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        okm[e] = s_ok[e] ? ~0u : 0u;
        offv[e] = s_off[e];
    }
    // chunk c's reads go to register set c % FOLD_SETS, issued FOLD_SETS steps before its
    // quotients are formed (at most (FOLD_SETS - 1) x 17 newer reads in flight, below the 63
    // the wait counter tracks)
    float f[FOLD_SETS], m[FOLD_SETS][E];
    auto load = [&](int c, float& fv, float (&mv)[E]) {
        const int pc = min(c * FOLD_CH + t, N - 1);
        fv = flux[pc];
#pragma unroll
        for (int e = 0; e < E; e++) mv[e] = gmodel[offv[e] + pc];
    };
    auto quot = [&](int c, float fv, const float (&mv)[E]) {
        // every quotient is formed (a clamped read past the star's end divides too) and
        // then masked, so no branch separates the divisions from their reads
        const unsigned in = c * FOLD_CH + t < N ? ~0u : 0u;
        float* row = &q[c & 1][0][t];
#pragma unroll
        for (int e = 0; e < E; e++)
            row[e * FOLD_ROW] = __uint_as_float(__float_as_uint(fv / mv[e]) & in & okm[e]);
    };
    auto fold = [&](int buf, float s) {
        if (lane < cnt) {
            const float4* r = (const float4*)&q[buf][lane][0];
#pragma unroll 16
            for (int i = 0; i < FOLD_CH / 4; i++) {
                const float4 v = r[i];
                s = s + v.x;
                s = s + v.y;
                s = s + v.z;
                s = s + v.w;
            }
        }
        return s;
    };
    const bool prod = wv > 0 && N > 0;
    if (prod) {
#pragma unroll
        for (int j = 0; j < FOLD_SETS; j++) load(j, f[j], m[j]);
        quot(0, f[0], m[0]);
        load(FOLD_SETS, f[0], m[0]);
    }
    __syncthreads();
    float s = -0.0f;
    for (int k0 = 0; k0 < n_ch; k0 += FOLD_SETS) {
#pragma unroll
        for (int j = 0; j < FOLD_SETS; j++) {
            // step k: wave 0 adds chunk k; the producers form chunk k + 1 and issue the
            // reads of chunk k + 1 + FOLD_SETS into the set it frees
            const int k = k0 + j;
            if (k >= n_ch) break;  // uniform over the workgroup
            const int nx = (j + 1) % FOLD_SETS;
            if (wv == 0) s = fold(k & 1, s);
            else if (prod && k + 1 < n_ch) {
                quot(k + 1, f[nx], m[nx]);
                load(k + 1 + FOLD_SETS, f[nx], m[nx]);
            }
            lds_barrier();
        }
    }
    if (wv == 0 && lane < cnt && s_ok[lane]) pf[evals[bd.x + lane].pf_off] = s / (float)N;
}

// Continuum moments Aᵀ(f/m) of the chunk and linear continua on f64 matrix cores
// (v_mfma_f64_16x16x4f64): the product of a block's design rows (n x p1, f32) with the
// n x E ratios y = f/m of E <= 64 evaluations of one star.  A workgroup covers one moment
// block (a chunk's pixels for the chunk continuum) and 16 evaluations per wavefront; a
// wavefront takes 16 pixels per round as 4 k-steps of 4.  Lane (j, g) = (lane % 16,
// lane / 16) holds in k-step q the design value of parameter 16t + j and the ratio of
// evaluation j, both at pixel a0 + 4g + q, with a0 a multiple of 4 so that the model, flux
// and design quads are float4 reads; pixels outside the block contribute zero.  The
// design sits in rows of `stride` values starting at pixel b0 (a multiple of 4): the
// chunk continuum stores the chunk's f32(powi(x^, k)) rows (b0 = the chunk start rounded
// down to 4), the linear continuum PT = 16 * NTL zero-padded columns (b0 = 0).  Products
// of two f32 values are exact in f64 and accumulate in f64, so the sums agree with a
// sequential f64 sum to f64 rounding.  The kernel streams the models: the occupancy hint
// (4 workgroups per CU) keeps it at <= 128 registers, i.e. >= 4 wavefronts per SIMD to
// keep enough reads in flight.
template <int NTL>
__global__ __launch_bounds__(NT, 4) void k_moments_mfma(const DevStar* __restrict__ stars,
                                                        const DevEval* __restrict__ evals,
                                                        const int4* __restrict__ blocks,
                                                        const float* __restrict__ model,
                                                        double* __restrict__ mpart) {
    typedef double d4 __attribute__((ext_vector_type(4)));
    const int4 bd = blocks[blockIdx.x];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int e0 = wv * 16;
    if (e0 >= bd.y) return;  // wavefront-uniform; the kernel has no barrier
    const DevStar& S = stars[evals[bd.x].star];
    const int4 mb = S.mblocks[bd.z];  // (chunk, i0, i1, second chunk or -1)
    const int2 sl = S.mslot[bd.z];     // partial slots of the chunk(s)
    const int j = lane & 15, g = lane >> 4, e = e0 + j;
    const bool ok = e < bd.y && evals[bd.x + e].status == ST_OK;
    const float* m = model + (ok ? evals[bd.x + e].m_off : 0);
    const int np = S.p1;
    const float* dbase;
    int stride, b0;
    if (S.cont_kind == CONT_CHUNK) {
        b0 = S.startstop[2 * mb.x] & ~3;
        stride = (S.startstop[2 * mb.x + 1] - b0 + 3) & ~3;
        dbase = S.cdesign + S.cdoff[mb.x];
    } else {
        b0 = 0;
        stride = (S.N + 3) & ~3;
        dbase = S.design;
    }
    d4 acc[4][NTL];  // one accumulator per k-step: independent matrix-core chains
#pragma unroll
    for (int h = 0; h < 4; h++)
#pragma unroll
        for (int t = 0; t < NTL; t++) acc[h][t] = d4{0.0, 0.0, 0.0, 0.0};
    typedef float v4 __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(1))) const v4 gf4;  // global-memory quads
    // design rows beyond p1 only feed output rows that are never stored, and an absent
    // evaluation only its own output column: such lanes read any valid row
    const float* drow = dbase + (size_t)min(j, np - 1) * stride - b0;
    // the second 16-column tile: the second chunk's parameters 0..15 for a block in the
    // overlap of two chunks, else parameters 16..31 of the same design (p1 > 16)
    const bool two = mb.w >= 0;
    const float* drow2;
    if (two) {
        const int b0b = S.startstop[2 * mb.w] & ~3;
        const int strideb = (S.startstop[2 * mb.w + 1] - b0b + 3) & ~3;
        drow2 = S.cdesign + S.cdoff[mb.w] + (size_t)min(j, np - 1) * strideb - b0b;
    } else {
        drow2 = dbase + (size_t)min(16 + j, np - 1) * stride - b0;
    }
    const bool tile2 = NTL > 1 && (two || np > 16);  // uniform over the workgroup
    auto mfma_round = [&](const v4& d0, const v4& d1, const float (&y)[4]) {
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)d0.x, (double)y[0], acc[0][0], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)d0.y, (double)y[1], acc[1][0], 0, 0, 0);
        acc[2][0] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)d0.z, (double)y[2], acc[2][0], 0, 0, 0);
        acc[3][0] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)d0.w, (double)y[3], acc[3][0], 0, 0, 0);
        if (NTL > 1 && tile2) {
            acc[0][NTL - 1] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)d1.x, (double)y[0], acc[0][NTL - 1], 0, 0, 0);
            acc[1][NTL - 1] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)d1.y, (double)y[1], acc[1][NTL - 1], 0, 0, 0);
            acc[2][NTL - 1] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)d1.z, (double)y[2], acc[2][NTL - 1], 0, 0, 0);
            acc[3][NTL - 1] = __builtin_amdgcn_mfma_f64_16x16x4f64((double)d1.w, (double)y[3], acc[3][NTL - 1], 0, 0, 0);
        }
    };
    const v4 dz = {0.0f, 0.0f, 0.0f, 0.0f};
    auto round = [&](int a, const float (&y)[4]) {
        const v4 d0 = *(gf4*)(drow + a);
        v4 d1 = dz;
        if (NTL > 1 && tile2) d1 = *(gf4*)(drow2 + a);
        mfma_round(d0, d1, y);
    };
    // interior rounds: all 16 pixels inside the block, no per-pixel condition; the next
    // round's flux, model and design quads are read before this round's products
    const int s_in = (mb.y + 3) & ~3;
    int s0 = s_in;
    if (s0 + 16 <= mb.z) {
        v4 f = *(gf4*)(S.flux + s0 + 4 * g);
        v4 mv = *(gf4*)(m + s0 + 4 * g);
        v4 d0 = *(gf4*)(drow + s0 + 4 * g);
        v4 d1 = dz;
        if (NTL > 1 && tile2) d1 = *(gf4*)(drow2 + s0 + 4 * g);
        for (; s0 + 16 <= mb.z; s0 += 16) {
            const int a = s0 + 4 * g;
            const bool more = s0 + 32 <= mb.z;
            const v4 fn = more ? *(gf4*)(S.flux + a + 16) : f;
            const v4 mn = more ? *(gf4*)(m + a + 16) : mv;
            const v4 dn0 = more ? *(gf4*)(drow + a + 16) : d0;
            v4 dn1 = d1;
            if (more && NTL > 1 && tile2) dn1 = *(gf4*)(drow2 + a + 16);
            const float y[4] = {f.x * __builtin_amdgcn_rcpf(mv.x), f.y * __builtin_amdgcn_rcpf(mv.y),
                                f.z * __builtin_amdgcn_rcpf(mv.z), f.w * __builtin_amdgcn_rcpf(mv.w)};
            mfma_round(d0, d1, y);
            f = fn;
            mv = mn;
            d0 = dn0;
            d1 = dn1;
        }
    }
    // the pixels before s_in (fewer than 4) and after the last whole round, masked
    for (int r0 = -16; r0 < 16 * ((mb.z - s0 + 15) / 16); r0 += 16) {
        const int base = r0 < 0 ? s_in - 16 : s0 + r0;
        const int a = base + 4 * g;
        float y[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        bool any = false;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = a + q;
            const bool in = r0 < 0 ? (i >= mb.y && i < s_in) : (i >= s0 && i < mb.z);
            if (in) y[q] = S.flux[i] * __builtin_amdgcn_rcpf(m[i]);
            any |= in;
        }
        // a quad with no pixel inside reads no design (its y are zero); a partial quad's
        // design row is in the stored range because the row starts at the chunk start
        // rounded down to 4 and its length is rounded up to 4
        if (!__any(any)) continue;
        const int ac = any ? a : (mb.y & ~3);
        round(ac, y);
    }
    if (!ok) return;
    double* out = mpart + evals[bd.x + e].mp_off;
    // output lane layout of the f64 16x16 product: lane (j, g) register r holds row 4r + g
    // of column j (evaluation j): parameter 16t + 4r + g of the block's chunk, or for the
    // second tile of a two-chunk block parameter 4r + g of the second chunk
#pragma unroll
    for (int t = 0; t < NTL; t++) {
        if (t == 1 && !tile2) break;
        const d4 s = (acc[0][t] + acc[1][t]) + (acc[2][t] + acc[3][t]);
        const bool second = t == 1 && two;
        double* o = out + (size_t)(second ? sl.y : sl.x) * np;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int k = (second ? 0 : 16 * t) + 4 * r + g;
            if (k < np) o[k] = s[r];
        }
    }
}

// one thread per (evaluation, chunk)
__global__ void k_solve(const DevStar* __restrict__ stars, const DevEval* __restrict__ evals,
                        const int2* __restrict__ items, int n_items,
                        const double* __restrict__ mpart, float* __restrict__ pf) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_items) return;
    const int2 it = items[t];
    DevEval& ev = const_cast<DevEval&>(evals[it.x]);
    if (ev.status != ST_OK) return;
    const DevStar& S = stars[ev.star];
    const int c = it.y, p1 = S.p1;
    int mb0, mb1;
    if (S.cont_kind == CONT_CHUNK) { mb0 = S.mb_off[c]; mb1 = S.mb_off[c + 1]; }
    else { mb0 = 0; mb1 = S.n_mblocks; }
    double bvec[P1_MAX], z[P1_MAX];
    for (int k = 0; k < p1; k++) {
        double s = 0.0;
        for (int mb = mb0; mb < mb1; mb++) s += mpart[ev.mp_off + (size_t)mb * p1 + k];
        bvec[k] = s;
    }
    const double* L = S.chol + (size_t)c * p1 * p1;
    for (int i = 0; i < p1; i++) {  // L z = b
        double s = bvec[i];
        for (int j = 0; j < i; j++) s -= L[i * p1 + j] * z[j];
        z[i] = s / L[i * p1 + i];
    }
    for (int i = p1 - 1; i >= 0; i--) {  // Lᵀ x = z
        double s = z[i];
        for (int j = i + 1; j < p1; j++) s -= L[j * p1 + i] * bvec[j];
        bvec[i] = s / L[i * p1 + i];
    }
    for (int k = 0; k < p1; k++) pf[ev.pf_off + (size_t)c * p1 + k] = (float)bvec[k];
}

// chi2 terms of E evaluations of one star over one block of observed pixels: block
// descriptor (first evaluation, count, chi2 block, 0); the star's blocks are (kind,
// chunk, i0, i1) with kind 1 where two chunk polynomials blend.  Design rows are formed
// once per pixel; each evaluation's continuum parameters are read from LDS.
template <int PM, int E, int P1F>
__global__ __launch_bounds__(NT) void k_chi2(const DevStar* __restrict__ stars,
                                             const DevEval* __restrict__ evals,
                                             const int4* __restrict__ blocks,
                                             const float* __restrict__ model,
                                             const float* __restrict__ pf,
                                             double* __restrict__ cpart) {
    // the 30-column linear continuum's variant (16 evaluations per block, so each design
    // column is read once per 16 evaluations) serves that continuum only
    constexpr bool LIN = PM == 32 && E == 16;
    __shared__ double red[4 * E];
    __shared__ __align__(16) float spf[E][2 * PM];
    const int4 bd = blocks[blockIdx.x];
    const DevStar& S = stars[evals[bd.x].star];
    const int4 cb = S.cblocks[bd.z];  // (kind, chunk, i0, i1)
    const float* m[E];
    bool ok[E];
    bool any = false;
#pragma unroll
    for (int e = 0; e < E; e++) {
        ok[e] = e < bd.y && evals[bd.x + e].status == ST_OK;
        m[e] = model + (ok[e] ? evals[bd.x + e].m_off : 0);
        any |= ok[e];
    }
    if (!any) return;
    double acc[E];
#pragma unroll
    for (int e = 0; e < E; e++) acc[e] = 0.0;
    const int p1 = P1F > 0 ? P1F : S.p1;
    if (S.cont_kind == CONT_CHUNK || S.cont_kind == CONT_LINEAR || S.cont_kind == CONT_CONSTANT) {
        const int c = S.cont_kind == CONT_CHUNK ? cb.y : 0;
        const int np = S.cont_kind == CONT_CONSTANT ? 1 : p1;
        for (int t = threadIdx.x; t < E * PM; t += NT) {
            const int e = t / PM, k = t - e * PM;
            float v = 0.0f, vp = 0.0f;
            if (e < bd.y && k < np) {
                const float* P = pf + evals[bd.x + e].pf_off;
                v = P[c * np + k];
                vp = c > 0 ? P[(c - 1) * np + k] : 0.0f;
            }
            spf[e][k] = v;
            spf[e][PM + k] = vp;
        }
        __syncthreads();
    }
    if (!LIN && S.cont_kind == CONT_CHUNK) {
        const int c = cb.y;
        const double lo = S.chunk_lohi[2 * c], hi = S.chunk_lohi[2 * c + 1];
        const double plo = c > 0 ? S.chunk_lohi[2 * c - 2] : 0.0;
        const double phi = c > 0 ? S.chunk_lohi[2 * c - 1] : 1.0;
        const bool blend = cb.x == 1;
        constexpr int U = 2;  // pixels per thread and round, loads issued first
        for (int i0 = cb.z + threadIdx.x; i0 < cb.w; i0 += U * NT) {
            // the parameters stay in LDS (broadcast reads) rather than being hoisted into
            // E * 2 * PM registers
            __asm__ volatile("" ::: "memory");
            double wl[U];
            float fl[U], vr[U], fc[U], mv[U][E];  // vr: 1 / variance
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u * NT;
                const bool in = i < cb.w;
                wl[u] = in ? S.wl[i] : lo;
                fl[u] = in ? S.flux[i] : 0.0f;
                vr[u] = in ? S.ivar[i] : 1.0f;
                fc[u] = (in && blend) ? S.fac[i] : 0.0f;
#pragma unroll
                for (int e = 0; e < E; e++) mv[u][e] = (in && ok[e]) ? m[e][i] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (i0 + u * NT >= cb.w) break;
                float a[PM], ap[PM];
                design_row<PM>(chunk_xhat(wl[u], lo, hi), p1, a);
                if (blend) design_row<PM>(chunk_xhat(wl[u], plo, phi), p1, ap);
#pragma unroll
                for (int e = 0; e < E; e++) {
                    if (!ok[e]) continue;
                    float p = dot_na<PM>(a, spf[e], p1);
                    if (blend) {
                        const float prev = dot_na<PM>(ap, spf[e] + PM, p1);
                        p = prev * fc[u] + p * (1.0f - fc[u]);
                    }
                    const float r = fl[u] - p * mv[u][e];
                    acc[e] += (double)((r * r) * vr[u]);
                }
            }
        }
    } else if (S.cont_kind == CONT_LINEAR) {
        const size_t n4 = (size_t)(S.N + 3) & ~(size_t)3;  // design column stride
        GlobalF* des = (GlobalF*)S.design;
        GlobalF* flx = (GlobalF*)S.flux;
        GlobalF* ivr = (GlobalF*)S.ivar;
        for (int i = cb.z + threadIdx.x; i < cb.w; i += NT) {
            __asm__ volatile("" ::: "memory");
            float d[PM], mv[E];
#pragma unroll
            for (int k = 0; k < PM; k++) d[k] = k < p1 ? des[(size_t)k * n4 + i] : 0.0f;
            const float fl = flx[i], iv = ivr[i];
#pragma unroll
            for (int e = 0; e < E; e++) {  // every m[e] points at N values
                const float x = ((GlobalF*)m[e])[i];
                mv[e] = ok[e] ? x : 1.0f;
            }
#pragma unroll
            for (int e = 0; e < E; e++) {
                if (!ok[e]) continue;
                // one evaluation's parameters at a time (LDS broadcasts), not E * PM registers
                if (LIN) __asm__ volatile("" ::: "memory");
                const float bq = fl / mv[e];
                float y = 0.0f;
#pragma unroll
                for (int k = 0; k < PM; k++) {
                    if (k < p1) {
                        const float t = d[k] * spf[e][k];
                        y = (k == 0) ? t : t + y;
                    }
                }
                const float r = y - bq;
                acc[e] += (double)((r * iv) * r);
            }
        }
    } else if (!LIN && S.cont_kind == CONT_FIXED) {
        for (int i = cb.z + threadIdx.x; i < cb.w; i += NT) {
            const float fl = S.flux[i], iv = S.ivar[i], fc = S.fixedc[i];
#pragma unroll
            for (int e = 0; e < E; e++) {
                if (!ok[e]) continue;
                const float r = fl - m[e][i] * fc;
                acc[e] += (double)((r * iv) * r);
            }
        }
    } else if (!LIN) {
        for (int i = cb.z + threadIdx.x; i < cb.w; i += NT) {
            const float fl = S.flux[i], iv = S.ivar[i];
#pragma unroll
            for (int e = 0; e < E; e++) {
                if (!ok[e]) continue;
                const float r = fl - m[e][i] * spf[e][0];
                acc[e] += (double)((r * iv) * r);
            }
        }
    }
    const double tot = group_sums<E>(acc, red);
    const int t = threadIdx.x;
    if (t < E && t < bd.y) {
        const DevEval& ev = evals[bd.x + t];
        if (ev.status == ST_OK) cpart[ev.cp_off + bd.z] = tot;
    }
}

__global__ void k_final(const DevStar* __restrict__ stars, const DevEval* __restrict__ evals,
                        int n, const double* __restrict__ cpart, double* __restrict__ chi2,
                        int* __restrict__ status) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const DevEval& ev = evals[e];
    if (status) status[e] = ev.status;
    if (ev.status != ST_OK) {
        if (ev.status > 0) chi2[e] = __longlong_as_double(0x7ff8000000000000LL);
        return;
    }
    const DevStar& S = stars[ev.star];
    double s = 0.0;
    for (int q = 0; q < S.n_cblocks; q++) s += cpart[ev.cp_off + q];
    chi2[e] = s / S.chi2_div;
}

// ---------------------------------------------------------------------------
// device-side PSO (batched path)
// ---------------------------------------------------------------------------
// thread per particle: velocity/position update (particleswarm.rs:234-276)
__global__ void k_pso_update(const GridGeom* __restrict__ geom, DevPso* __restrict__ pso,
                             DevEval* __restrict__ evals, int n, double* __restrict__ pos,
                             double* __restrict__ vel, const double* __restrict__ bpos) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    DevEval& ev = evals[e];
    DevPso& ps = pso[ev.star];
    if (ps.done || ps.status) { ev.active = 0; return; }
    ev.active = 1;
    const int q = e - ps.first_eval;
    double gb[5];
    for (int d = 0; d < 5; d++) gb[d] = ps.gbest[d];
    int rc = pso_update_particle(*geom, ps.b, ps.s, ps.key_iter, ps.iter, ps.P, q, pos + 5 * (size_t)e,
                                 vel + 5 * (size_t)e, bpos + 5 * (size_t)e, gb);
    if (rc) atomicCAS(&ps.status, 0, rc);
}

// thread per star: personal / global best update (particleswarm.rs:283-298)
__global__ void k_pso_best(DevPso* __restrict__ pso, int n_stars, const DevEval* __restrict__ evals,
                           const double* __restrict__ pos, double* __restrict__ bpos,
                           double* __restrict__ bcost, const double* __restrict__ cost,
                           double* __restrict__ trace) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_stars) return;
    DevPso& ps = pso[s];
    if (ps.done || ps.status) return;
    for (int q = 0; q < ps.P; q++) {
        const int e = ps.first_eval + q;
        if (evals[e].status != ST_OK) { ps.status = evals[e].status; return; }
    }
    for (int q = 0; q < ps.P; q++) {
        const int e = ps.first_eval + q;
        const double c = cost[e];
        if (c < bcost[e]) {
            for (int d = 0; d < 5; d++) bpos[5 * e + d] = pos[5 * e + d];
            bcost[e] = c;
            if (c < ps.gcost) {
                for (int d = 0; d < 5; d++) ps.gbest[d] = pos[5 * e + d];
                ps.gcost = c;
            }
        }
    }
    if (trace) {
        double* tr = trace + ((size_t)s * (ps.max_iters + 1) + ps.iter) * 6;
        tr[0] = ps.gcost;
        for (int d = 0; d < 5; d++) tr[1 + d] = ps.gbest[d];
    }
    ps.iter += 1;
    if (ps.iter > ps.max_iters) ps.done = 1;
}

// thread per star after the first evaluation: stable sort by cost, bests
// (particleswarm.rs:117-131, 209-215)
__global__ void k_pso_init(DevPso* __restrict__ pso, int n_stars, const DevEval* __restrict__ evals,
                           double* __restrict__ pos, double* __restrict__ vel,
                           double* __restrict__ bpos, double* __restrict__ bcost,
                           double* __restrict__ cost, double* __restrict__ trace) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_stars) return;
    DevPso& ps = pso[s];
    if (ps.done || ps.status || ps.iter != 0) return;  // only stars not started yet
    const int f0 = ps.first_eval, P = ps.P;
    for (int q = 0; q < P; q++)
        if (evals[f0 + q].status != ST_OK) { ps.status = evals[f0 + q].status; return; }
    for (int i = 1; i < P; i++) {  // stable insertion sort
        double c = cost[f0 + i], p[5], v[5];
        for (int d = 0; d < 5; d++) { p[d] = pos[5 * (f0 + i) + d]; v[d] = vel[5 * (f0 + i) + d]; }
        int j = i - 1;
        while (j >= 0 && cost[f0 + j] > c) {
            cost[f0 + j + 1] = cost[f0 + j];
            for (int d = 0; d < 5; d++) {
                pos[5 * (f0 + j + 1) + d] = pos[5 * (f0 + j) + d];
                vel[5 * (f0 + j + 1) + d] = vel[5 * (f0 + j) + d];
            }
            j--;
        }
        cost[f0 + j + 1] = c;
        for (int d = 0; d < 5; d++) { pos[5 * (f0 + j + 1) + d] = p[d]; vel[5 * (f0 + j + 1) + d] = v[d]; }
    }
    for (int q = 0; q < P; q++) {
        for (int d = 0; d < 5; d++) bpos[5 * (f0 + q) + d] = pos[5 * (f0 + q) + d];
        bcost[f0 + q] = cost[f0 + q];
    }
    ps.gcost = cost[f0];
    for (int d = 0; d < 5; d++) ps.gbest[d] = pos[5 * f0 + d];
    if (trace) {
        double* tr = trace + (size_t)s * (ps.max_iters + 1) * 6;
        tr[0] = ps.gcost;
        for (int d = 0; d < 5; d++) tr[1 + d] = ps.gbest[d];
    }
    ps.iter = 1;
    if (ps.max_iters < 1) ps.done = 1;
}

// Evaluation records from their host part: one workgroup per record zeroes it and sets the
// star, the active flag and the buffer offsets.
__global__ __launch_bounds__(64) void k_eval_init(const EvalInit* __restrict__ in, int n,
                                                  DevEval* __restrict__ out) {
    const int e = blockIdx.x;
    if (e >= n) return;
    int* w = (int*)(out + e);
    for (int i = threadIdx.x; i < (int)(sizeof(DevEval) / sizeof(int)); i += blockDim.x) w[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        const EvalInit v = in[e];
        DevEval& ev = out[e];
        ev.star = v.star;
        ev.active = v.active;
        ev.m_off = v.m_off;
        ev.mp_off = v.mp_off;
        ev.cp_off = v.cp_off;
        ev.pf_off = v.pf_off;
    }
}

// Continuous batching: the PSO state of stars that stay in flight moves from the previous
// batch into the rebuilt one on the device.  Thread per evaluation of the new batch:
// src_eval[e] is its particle's evaluation in the old batch (-1: a new star, left as
// initialised); thread s < n_stars also moves the star record (src_star[s] >= 0),
// keeping the new batch's first_eval.
__global__ void k_pso_remap(const int* __restrict__ src_eval, int n_evals,
                            DevEval* __restrict__ evals, const int* __restrict__ src_star,
                            int n_stars,
                            const DevPso* __restrict__ pso_old, DevPso* __restrict__ pso_new,
                            const double* __restrict__ pos_old, double* __restrict__ pos_new,
                            const double* __restrict__ vel_old, double* __restrict__ vel_new,
                            const double* __restrict__ bpos_old, double* __restrict__ bpos_new,
                            const double* __restrict__ bcost_old, double* __restrict__ bcost_new) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_evals) {
        const int o = src_eval[t];
        if (o >= 0) {
            evals[t].active = 0;  // resumed: not part of the first-evaluation pass
            for (int d = 0; d < 5; d++) {
                pos_new[5 * (size_t)t + d] = pos_old[5 * (size_t)o + d];
                vel_new[5 * (size_t)t + d] = vel_old[5 * (size_t)o + d];
                bpos_new[5 * (size_t)t + d] = bpos_old[5 * (size_t)o + d];
            }
            bcost_new[t] = bcost_old[o];
        }
    }
    if (t < n_stars) {
        const int o = src_star[t];
        if (o >= 0) {
            const int f0 = pso_new[t].first_eval;
            pso_new[t] = pso_old[o];
            pso_new[t].first_eval = f0;
        }
    }
}

// ---------------------------------------------------------------------------
// counting sort of evaluations by key (order only: results do not depend on it)
// ---------------------------------------------------------------------------
__global__ void k_sort_zero(int* __restrict__ hist, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) hist[i] = 0;
}

__global__ void k_sort_hist(const DevEval* __restrict__ evals, int n, int* __restrict__ hist) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) atomicAdd(&hist[evals[e].key], 1);
}

// exclusive scan of hist[0..n) in place, one workgroup of 1024 threads
__global__ __launch_bounds__(1024) void k_sort_scan(int* __restrict__ hist, int n) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int lo = t * per, hi = min(n, lo + per);
    int s = 0;
    for (int i = lo; i < hi; i++) s += hist[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t > 0 ? part[t - 1] : 0;
    for (int i = lo; i < hi; i++) {
        int c = hist[i];
        hist[i] = run;
        run += c;
    }
}

__global__ void k_sort_scatter(const DevEval* __restrict__ evals, int n, int* __restrict__ offs,
                               int* __restrict__ perm) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) perm[atomicAdd(&offs[evals[e].key], 1)] = e;
}

// ---------------------------------------------------------------------------
// host-side launchers (called from the driver in api.hip)
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// pairing: adjacent evaluations in neighbour-row order that gather the same 64 rows
// for the same synthetic tiles become one work item of k_model
// ---------------------------------------------------------------------------
// link[i] = 1: evaluations perm[i] and perm[i + 1] (both valid, same synthetic tiles) have at
// most LINK_MAX distinct present rows together.  One wavefront per pair: lane l holds the l-th
// present row of each (ascending lists), finds its row of the first among the second's by a
// binary search over the lanes, and |A u B| = |A| + |B| - |A n B|.
__global__ __launch_bounds__(64) void k_pair_eq(const DevEval* __restrict__ evals,
                                                const int* __restrict__ perm, int n,
                                                int* __restrict__ eq) {
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    if (i >= n) return;
    int r = 0;
    if (i + 1 < n) {
        const DevEval& a = evals[perm[i]];
        const DevEval& b = evals[perm[i + 1]];
        if (a.status == ST_OK && b.status == ST_OK && a.tile_base == b.tile_base) {
            const int na = a.nrow, nb = b.nrow;
            const unsigned long long pa = lane < na ? (unsigned long long)a.rowp[lane] : ~0ull;
            const unsigned long long pb = lane < nb ? (unsigned long long)b.rowp[lane] : ~0ull;
            // first lane of b whose row is >= pa (b's rows ascending, padded with ~0)
            int lo = 0;
#pragma unroll
            for (int step = 32; step; step >>= 1) {
                const unsigned long long v = __shfl(pb, lo + step - 1, 64);
                if (v < pa) lo += step;
            }
            const bool found = lane < na && lo < nb && __shfl(pb, lo, 64) == pa;
            const int inter = __popcll(__ballot(found));
            r = na + nb - inter <= LINK_MAX;
        }
    }
    if (lane == 0) eq[i] = r;
}

// The union of the present rows of every work item of two or three evaluations, one wavefront
// per item: the rows (storage indices) of all its evaluations set bits of a bitmap in LDS over
// [first row, first row + 2048); a row's union position is the number of set bits below it
// (prefix popcounts over the 64 words); every evaluation writes its weight at its rows'
// positions, 0 elsewhere; rows of weight 0 pad the list to a multiple of 8.
__global__ __launch_bounds__(64) void k_union(const DevEval* __restrict__ evals,
                                              const int4* __restrict__ items,
                                              const int* __restrict__ n_items,
                                              DevUnion* __restrict__ unions,
                                              const uint16_t* spectra, long long pitch) {
    constexpr int SPAN = 2048;
    __shared__ unsigned int bits[SPAN / 32];
    const int k = blockIdx.x;
    const int lane = threadIdx.x;
    if (k >= *n_items) return;
    const int4 it = items[k];
    if (it.y < 0) return;
    const int ids[3] = {it.x, it.y, it.z};
    const int ng = it.z >= 0 ? 3 : 2;
    int row[3];
    int lo = 0x7fffffff;
#pragma unroll
    for (int g = 0; g < 3; g++) {
        row[g] = -1;
        if (g < ng && lane < evals[ids[g]].nrow)
            row[g] = (int)((evals[ids[g]].rowp[lane] - spectra) / pitch);
        if (row[g] >= 0) lo = min(lo, row[g]);
    }
    for (int o = 32; o; o >>= 1) lo = min(lo, __shfl_xor(lo, o, 64));
    bits[lane] = 0u;
    __syncthreads();
    bool wide = false;
#pragma unroll
    for (int g = 0; g < 3; g++)
        if (row[g] >= 0) {
            const int x = row[g] - lo;
            if (x >= SPAN) wide = true;
            else atomicOr(&bits[x >> 5], 1u << (x & 31));
        }
    DevUnion& U = unions[k];
    if (__any(wide)) {  // rows too far apart for the bitmap: one evaluation at a time
        if (lane == 0) U.n = -1;
        return;
    }
    __syncthreads();
    const unsigned int w = bits[lane];
    int pre = __popc(w);  // inclusive prefix of the set bits per word
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(pre, o, 64);
        if (lane >= o) pre += t;
    }
    const int m = __shfl(pre, 63, 64);
    if (m > UNION_MAX) {
        if (lane == 0) U.n = -1;
        return;
    }
    const int m8 = (m + 7) & ~7;
    for (int q = lane; q < UNION_MAX; q += 64) {
#pragma unroll
        for (int g = 0; g < 3; g++) U.fw[g][q] = 0.0f;
        if (q >= m) U.rowp[q] = spectra;
    }
    __syncthreads();  // the zeros land before the weights (the barrier waits for the stores)
    __shared__ int wpre[SPAN / 32];  // exclusive prefix per word
    wpre[lane] = pre - __popc(w);
    __syncthreads();
#pragma unroll
    for (int g = 0; g < 3; g++)
        if (row[g] >= 0) {
            const int x = row[g] - lo;
            const int pos = wpre[x >> 5] + __popc(bits[x >> 5] & ((1u << (x & 31)) - 1u));
            U.rowp[pos] = evals[ids[g]].rowp[lane];
            U.fw[g][pos] = evals[ids[g]].fw[lane];
        }
    if (lane == 0) U.n = m8;
}

// inclusive block scans over 1024 threads (op: max or sum)
template <bool MAX>
__device__ __forceinline__ int block_scan(int v, int* wsum /* 16 */) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v = MAX ? max(v, u) : v + u;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    if (threadIdx.x < 64) {
        int x = lane < 16 ? wsum[lane] : (MAX ? -1 : 0);
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int u = __shfl_up(x, o, 64);
            if (lane >= o) x = MAX ? max(x, u) : x + u;
        }
        if (lane < 16) wsum[lane] = x;
    }
    __syncthreads();
    const int prev = w > 0 ? wsum[w - 1] : (MAX ? -1 : 0);
    const int r = MAX ? max(v, prev) : v + prev;
    __syncthreads();
    return r;
}

// Greedy grouping along each run of linked evaluations (k_pair_eq): positions 0-2, 3-5, ... of
// a run share their gathers; items[k] = (first evaluation, second or -1, third or -1, 0), in
// neighbour-row order.
__global__ __launch_bounds__(1024) void k_pair_items(const int* __restrict__ perm,
                                                     const int* __restrict__ eq, int n,
                                                     int4* __restrict__ items,
                                                     int* __restrict__ n_items) {
    __shared__ int wsum[16];
    __shared__ int s_carry[2];  // run start of the chunk's last element, items so far
    const int tid = threadIdx.x;
    if (tid == 0) { s_carry[0] = 0; s_carry[1] = 0; }
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const int i = base + tid;
        const bool valid = i < n;
        const int mark = (valid && (i == 0 || !eq[i - 1])) ? i : -1;
        const int start = max(block_scan<true>(mark, wsum), s_carry[0]);
        const int emit = (valid && (i - start) % 3 == 0) ? 1 : 0;
        const int incl = block_scan<false>(emit, wsum);
        const int k = s_carry[1] + incl - emit;
        if (emit) {
            const bool two = eq[i] != 0;
            const bool three = two && eq[i + 1] != 0;  // eq[i] != 0 implies i + 1 < n
            items[k] = make_int4(perm[i], two ? perm[i + 1] : -1, three ? perm[i + 2] : -1, 0);
        }
        __syncthreads();
        if (tid == 1023) {
            s_carry[0] = start;
            s_carry[1] = k + emit;
        }
        __syncthreads();
    }
    if (tid == 0) *n_items = s_carry[1];
}

hipError_t launch_cont(const DevStar* stars, DevEval* evals, int n_evals, const int4* bmom,
                       const int4* bchi2, const ContClass* cls, int n_cls,
                       const int2* solve_items, int n_solve, const float* model, double* mpart,
                       float* pf, double* cpart, double* chi2, int* status, hipStream_t st);

hipError_t launch_eval(const DevGrid& G, const DevStar* stars, DevEval* evals, int n_evals,
                       const double* labels, float* rot, int* tbound, int tb_stride,
                       int* sort_hist, int* perm, int* pair_eq, int4* items, int* n_items,
                       DevUnion* unions, int n_model,
                       const int4* bmom, const int4* bchi2, const ContClass* cls, int n_cls,
                       const int2* solve_items, int n_solve, float* model, double* mpart, float* pf,
                       double* cpart, double* chi2, int* status, int mode, bool fft,
                       hipEvent_t ev_model0, hipEvent_t ev_model1, hipStream_t st) {
    if (n_evals <= 0) return hipSuccess;
    // without the FFT path the crossover is off for k_prep, so no evaluation selects it
    DevGrid Gp = G;
    if (!fft) Gp.fft_min_taps = 1 << 30;
    if (Gp.prep_taps < 8 || Gp.prep_taps > KROT_MAX) Gp.prep_taps = KROT_MAX;
    hipLaunchKernelGGL(k_prep, dim3(n_evals), dim3(64),
                       2 * (Gp.prep_taps + 4) * sizeof(float), st, Gp, stars, labels, evals, rot,
                       tbound, tb_stride, mode);
    if (fft) hipLaunchKernelGGL(k_rot_spectrum, dim3(n_evals), dim3(NT), 0, st, evals, rot);
    const int nb = G.n_models_bins;
    hipLaunchKernelGGL(k_sort_zero, dim3((nb + 255) / 256), dim3(256), 0, st, sort_hist, nb);
    hipLaunchKernelGGL(k_sort_hist, dim3((n_evals + 255) / 256), dim3(256), 0, st, evals, n_evals,
                       sort_hist);
    hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, st, sort_hist, nb);
    hipLaunchKernelGGL(k_sort_scatter, dim3((n_evals + 255) / 256), dim3(256), 0, st, evals,
                       n_evals, sort_hist, perm);
    hipLaunchKernelGGL(k_pair_eq, dim3(n_evals), dim3(64), 0, st, evals, perm, n_evals, pair_eq);
    hipLaunchKernelGGL(k_pair_items, dim3(1), dim3(1024), 0, st, perm, pair_eq, n_evals, items,
                       n_items);
    hipLaunchKernelGGL(k_union, dim3(n_evals), dim3(64), 0, st, evals, items, n_items, unions,
                       G.spectra, G.pitch);
    if (ev_model0) (void)hipEventRecord(ev_model0, st);
    if (n_model > 0 && fft)
        hipLaunchKernelGGL(k_model<true>, dim3(n_model), dim3(NT), 0, st, G, stars, evals, rot,
                           items, n_items, unions, n_evals, tbound, tb_stride, model, mode);
    else if (n_model > 0)
        hipLaunchKernelGGL(k_model<false>, dim3(n_model), dim3(NT), 0, st, G, stars, evals, rot,
                           items, n_items, unions, n_evals, tbound, tb_stride, model, mode);
    if (ev_model1) (void)hipEventRecord(ev_model1, st);
    return launch_cont(stars, evals, n_evals, bmom, bchi2, cls, n_cls, solve_items, n_solve,
                       model, mpart, pf, cpart, chi2, status, st);
}

// continuum fit + chi2 of evaluations whose models are in `model` and whose status is set
hipError_t launch_cont(const DevStar* stars, DevEval* evals, int n_evals, const int4* bmom,
                       const int4* bchi2, const ContClass* cls, int n_cls,
                       const int2* solve_items, int n_solve, const float* model, double* mpart,
                       float* pf, double* cpart, double* chi2, int* status, hipStream_t st) {
    if (n_evals <= 0) return hipSuccess;
    for (int c = 0; c < n_cls; c++) {
        const ContClass& K = cls[c];
        if (K.n_mom <= 0) continue;
        const int4* blk = bmom + K.mom_off;
        if (K.kind == CONT_LINEAR || K.kind == CONT_CHUNK) {
            // chunk continua: two-tile kernel (blocks in the overlap of two chunks carry both)
            if (K.kind == CONT_LINEAR && K.p1 <= 16)
                hipLaunchKernelGGL((k_moments_mfma<1>), dim3(K.n_mom), dim3(NT), 0, st, stars,
                                   evals, blk, model, mpart);
            else
                hipLaunchKernelGGL((k_moments_mfma<2>), dim3(K.n_mom), dim3(NT), 0, st, stars,
                                   evals, blk, model, mpart);
            continue;
        }
        if (K.kind == CONT_CONSTANT)  // the parameter itself: no solve
            hipLaunchKernelGGL((k_const_fold<16>), dim3(K.n_mom), dim3(NT), 0, st, stars, evals,
                               blk, model, pf);
    }
    if (n_solve > 0)
        hipLaunchKernelGGL(k_solve, dim3((n_solve + 127) / 128), dim3(128), 0, st, stars, evals,
                           solve_items, n_solve, mpart, pf);
    for (int c = 0; c < n_cls; c++) {
        const ContClass& K = cls[c];
        if (K.n_chi2 <= 0) continue;
        const KVariant v = chi2_variant(K.kind, K.p1);
        const int4* blk = bchi2 + K.chi2_off;
#define PASTA_CHI2(PM, E, F)                                                                  \
    if (v.pm == PM && v.e == E && v.f == F)                                                    \
        hipLaunchKernelGGL((k_chi2<PM, E, F>), dim3(K.n_chi2), dim3(NT), 0, st, stars, evals,   \
                           blk, model, pf, cpart);
        PASTA_CHI2(1, 8, 0)
        PASTA_CHI2(16, CHI2_GROUP, 5)
        PASTA_CHI2(16, CHI2_GROUP, 9)
        PASTA_CHI2(16, CHI2_GROUP, 11)
        PASTA_CHI2(16, CHI2_GROUP, 0)
        PASTA_CHI2(32, 4, 0)
        PASTA_CHI2(32, 16, 0)
#undef PASTA_CHI2
    }
    hipLaunchKernelGGL(k_final, dim3((n_evals + 127) / 128), dim3(128), 0, st, stars, evals,
                       n_evals, cpart, chi2, status);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// binary stars
// ---------------------------------------------------------------------------

// interpolate_linear (src/interpolate.rs:985-1030) of evaluation e = blockIdx.y at the
// labels labels[e * lstride .. +3): every pixel is x / 65535 * factor, then the gemv
// over the 8 neighbours in column order, rounded exactly as the reference's f32 steps;
// block x covers pixels [8 NT x, 8 NT (x + 1)).  status[e]: 0 or the weights' failure.
__global__ __launch_bounds__(NT) void k_interp_lin(DevGrid G, const double* __restrict__ labels,
                                                   int lstride, float* __restrict__ out,
                                                   int* __restrict__ status) {
    __shared__ int s_idx[8];
    __shared__ float s_w[8], s_f[8];
    __shared__ int s_st;
    const int e = blockIdx.y;
    if (threadIdx.x == 0) {
        const double* l = labels + (size_t)e * lstride;
        int idx[8];
        float w[8];
        const int st = local_weights_linear(*G.geom, l[0], l[1], l[2], idx, w);
        s_st = st;
        for (int j = 0; j < 8; j++) {
            const int id = st ? 0 : idx[j];
            s_idx[j] = id;
            s_w[j] = st ? 0.0f : w[j];
            s_f[j] = G.factors ? G.factors[id] : (float)(1.0 / 65535.0);
        }
        if (blockIdx.x == 0) status[e] = st;
    }
    __syncthreads();
    if (s_st) return;
    const long long x0 = (long long)blockIdx.x * NT * 8 + threadIdx.x;
    float* o = out + (size_t)e * G.W;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const long long x = x0 + (long long)k * NT;
        if (x >= G.W) break;
        float y = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const float u = (float)G.spectra[(size_t)s_idx[j] * G.pitch + x];
            const float t = (u / 65535.0f * s_f[j]) * s_w[j];
            y = (j == 0) ? t : t + y;
        }
        o[x] = y;
    }
}

// DVector::mean of each row of `in` (W values): a sequential f32 fold in pixel order and
// one division.  Two wavefronts per row: wavefront 1 stages chunks of 4,096 values into a
// double-buffered LDS ring while lane 0 of wavefront 0 adds the previous chunk in order,
// 64 values per step after one batch of LDS reads.  Rows with a non-zero status are
// skipped.
constexpr int SEQ_CHUNK = 4096;
__global__ __launch_bounds__(128) void k_seq_mean(const float* __restrict__ in, int W,
                                                  const int* __restrict__ status,
                                                  float* __restrict__ mean) {
    __shared__ float4 ring[2][SEQ_CHUNK / 4];
    const int e = blockIdx.x;
    if (status[e]) return;
    const float* r = in + (size_t)e * W;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nch = (W + SEQ_CHUNK - 1) / SEQ_CHUNK;
    float s = 0.0f;
    // chunk c is staged in round c and summed in round c + 1
    for (int c = 0; c <= nch; c++) {
        if (wv == 1 && c < nch) {  // all 64 loads of a lane in flight before the stores
            float* dst = reinterpret_cast<float*>(ring[c & 1]);
            const int x0 = c * SEQ_CHUNK + lane;
            float v[SEQ_CHUNK / 64];
#pragma unroll
            for (int k = 0; k < SEQ_CHUNK / 64; k++) v[k] = x0 + 64 * k < W ? r[x0 + 64 * k] : 0.0f;
#pragma unroll
            for (int k = 0; k < SEQ_CHUNK / 64; k++) dst[lane + 64 * k] = v[k];
        }
        if (wv == 0 && lane == 0 && c > 0) {
            const float4* q = ring[(c - 1) & 1];
            const int m = W - (c - 1) * SEQ_CHUNK < SEQ_CHUNK ? W - (c - 1) * SEQ_CHUNK : SEQ_CHUNK;
            // 64 values per step: 16 LDS reads issued together, then 64 adds in order
            const int m64 = m >> 6;
            for (int k = 0; k < m64; k++) {
                float4 a[16];
#pragma unroll
                for (int t = 0; t < 16; t++) a[t] = q[16 * k + t];
#pragma unroll
                for (int t = 0; t < 16; t++) {
                    s = s + a[t].x;
                    s = s + a[t].y;
                    s = s + a[t].z;
                    s = s + a[t].w;
                }
            }
            const float* qf = reinterpret_cast<const float*>(q);
            for (int k = m64 << 6; k < m; k++) s = s + qf[k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) mean[e] = s / (float)W;
}

// shift_resample_and_add_binary_components (src/convolve_rv.rs:606-690, Logspace grid)
// for binary evaluation b = blockIdx.y with the components of c = b % ncomp (the
// evaluations of several spectra share one set): components 2c (star 1) and 2c+1 (star 2) are the
// convolved models `comp` (synthetic grid, W per component, status in cevals), their
// linear continua `cont` (status cstat) and the continua's means.  Writes the normalised
// model at obs evaluation b's m_off and its status: the first failure in the reference's
// order (model 1, continuum 1, model 2, continuum 2, the wavelength range of both shifts,
// then a resample index past the continuum's end, checked at the first and last pixel of
// the increasing observed wavelengths).
__global__ __launch_bounds__(NT) void k_binary_combine(DevGrid G, const DevStar* __restrict__ stars,
                                                       DevEval* __restrict__ oevals,
                                                       const double* __restrict__ labels11,
                                                       const DevEval* __restrict__ cevals,
                                                       const float* __restrict__ comp,
                                                       const int* __restrict__ cstat,
                                                       const float* __restrict__ cont,
                                                       const float* __restrict__ cmean,
                                                       int ncomp, float* __restrict__ model) {
    const int b = blockIdx.y, cb = b % ncomp;
    const DevStar& S = stars[oevals[b].star];
    const double* l = labels11 + (size_t)11 * b;
    int st = cevals[2 * cb].status;
    if (!st) st = cstat[2 * cb];
    if (!st) st = cevals[2 * cb + 1].status;
    if (!st) st = cstat[2 * cb + 1];
    const double sf1 = 1.0 - l[4] / C_KMS, sf2 = 1.0 - l[9] / C_KMS;
    if (!st && (S.wl0 * fmin(sf1, sf2) < G.synth_first_wl || S.wlN * fmax(sf1, sf2) > G.synth_last_wl))
        st = ST_RANGE;
    const double sl1 = log10(sf1) / G.sg_step, sl2 = log10(sf2) / G.sg_step;
    if (!st) {
        const double f0 = S.fidx[0], f1 = S.fidx[S.N - 1];
        if (jfloor(f0 + sl1) + 1 >= G.W || jfloor(f1 + sl1) + 1 >= G.W ||
            jfloor(f0 + sl2) + 1 >= G.W || jfloor(f1 + sl2) + 1 >= G.W)
            st = ST_INDEX;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) oevals[b].status = st;
    if (st) return;
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= S.N) return;
    const float lr = (float)l[10] * cmean[2 * cb] / cmean[2 * cb + 1];
    const float* n1 = comp + (size_t)(2 * cb) * G.W;
    const float* n2 = n1 + G.W;
    const float* c1 = cont + (size_t)(2 * cb) * G.W;
    const float* c2 = c1 + G.W;
    const double x1 = S.fidx[i] + sl1, x2 = S.fidx[i] + sl2;
    const long long j1 = jfloor(x1), j2 = jfloor(x2);
    const float w1 = (float)x1 - (float)j1, w2 = (float)x2 - (float)j2;
    const float a1 = (1.0f - w1) * n1[j1] + w1 * n1[j1 + 1];
    const float k1 = (1.0f - w1) * c1[j1] + w1 * c1[j1 + 1];
    const float a2 = (1.0f - w2) * n2[j2] + w2 * n2[j2 + 1];
    const float k2 = (1.0f - w2) * c2[j2] + w2 * c2[j2 + 1];
    model[oevals[b].m_off + i] = (a1 * k1 + lr * a2 * k2) / (k1 + lr * k2);
}

// the linear continua of ncomp label pairs (labels clabels with stride 5) and their means
hipError_t launch_binary_cont(const DevGrid& CG, const double* clabels, int ncomp, int* cstat,
                              float* cont, float* cmean, hipStream_t st) {
    if (ncomp <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_interp_lin, dim3((CG.W + 8 * NT - 1) / (8 * NT), 2 * ncomp), dim3(NT), 0,
                       st, CG, clabels, 5, cont, cstat);
    hipLaunchKernelGGL(k_seq_mean, dim3(2 * ncomp), dim3(128), 0, st, cont, CG.W, cstat, cmean);
    return hipGetLastError();
}

// binary models of n_obs_evals evaluations (observed records oevals, labels11) from the
// components of ncomp label pairs (records cevals) and their continua
hipError_t launch_binary(const DevGrid& G, const DevStar* stars, DevEval* oevals,
                         int n_obs_evals, const double* labels11, const DevEval* cevals, int ncomp,
                         const float* comp, const int* cstat, const float* cont,
                         const float* cmean, int n_obs_max, float* model, hipStream_t st) {
    if (n_obs_evals <= 0 || ncomp <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_binary_combine, dim3((n_obs_max + NT - 1) / NT, n_obs_evals), dim3(NT), 0,
                       st, G, stars, oevals, labels11, cevals, comp, cstat, cont, cmean, ncomp,
                       model);
    return hipGetLastError();
}

// RVCostFunction (src/fitting.rs:736-781), first pass: the continua of evaluation b =
// blockIdx.y shifted by its two RVs (rv[2b], rv[2b+1]) into kc rows 2b, 2b+1 (N each) and
// its status: the first failure in the reference's order (model 1 and continuum 1 at
// rv1: range, then resample index; model 2 and continuum 2 at rv2), written to the
// observed record and to both rows' status.  Components: W-value synthetic-grid arrays.
__global__ __launch_bounds__(NT) void k_rv_shift(DevGrid G, const DevStar* __restrict__ stars,
                                                 DevEval* __restrict__ oevals,
                                                 const double* __restrict__ rv,
                                                 const float* __restrict__ cont1,
                                                 const float* __restrict__ cont2,
                                                 float* __restrict__ kc, int* __restrict__ kst) {
    const int b = blockIdx.y;
    const DevStar& S = stars[oevals[b].star];
    int st = ST_OK;
    double sl[2];
    for (int c = 0; c < 2 && !st; c++) {
        const double sf = 1.0 - rv[2 * b + c] / C_KMS;
        if (S.wl0 * sf < G.synth_first_wl || S.wlN * sf > G.synth_last_wl) st = ST_RANGE;
        sl[c] = log10(sf) / G.sg_step;
        if (!st && (jfloor(S.fidx[0] + sl[c]) + 1 >= G.W || jfloor(S.fidx[S.N - 1] + sl[c]) + 1 >= G.W))
            st = ST_INDEX;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        oevals[b].status = st;
        kst[2 * b] = st;
        kst[2 * b + 1] = st;
    }
    if (st) return;
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= S.N) return;
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const float* a = c ? cont2 : cont1;
        const double x = S.fidx[i] + sl[c];
        const long long j = jfloor(x);
        const float w = (float)x - (float)j;
        kc[(size_t)(2 * b + c) * S.N + i] = (1.0f - w) * a[j] + w * a[j + 1];
    }
}

// second pass: lr = light_ratio * mean(c2') / mean(c1'), model = (s1' c1' + s2' c2' lr) /
// (c1' + c2' lr) at the observed pixels, s' the shifted models
__global__ __launch_bounds__(NT) void k_rv_combine(DevGrid G, const DevStar* __restrict__ stars,
                                                   const DevEval* __restrict__ oevals,
                                                   const double* __restrict__ rv,
                                                   const float* __restrict__ model1,
                                                   const float* __restrict__ model2,
                                                   const float* __restrict__ kc,
                                                   const float* __restrict__ kmean,
                                                   float light_ratio, float* __restrict__ model) {
    const int b = blockIdx.y;
    const DevEval& ev = oevals[b];
    if (ev.status) return;
    const DevStar& S = stars[ev.star];
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= S.N) return;
    float s[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const float* a = c ? model2 : model1;
        const double x = S.fidx[i] + log10(1.0 - rv[2 * b + c] / C_KMS) / G.sg_step;
        const long long j = jfloor(x);
        const float w = (float)x - (float)j;
        s[c] = (1.0f - w) * a[j] + w * a[j + 1];
    }
    const float lr = light_ratio * kmean[2 * b + 1] / kmean[2 * b];
    const float k1 = kc[(size_t)(2 * b) * S.N + i], k2 = kc[(size_t)(2 * b + 1) * S.N + i];
    const float synth = s[0] * k1 + s[1] * k2 * lr;
    const float cont = k1 + k2 * lr;
    model[ev.m_off + i] = synth / cont;
}

hipError_t launch_binary_rv(const DevGrid& G, const DevStar* stars, DevEval* oevals, int n,
                            const double* rv, const float* model1, const float* model2,
                            const float* cont1, const float* cont2, float light_ratio, int n_obs,
                            float* kc, int* kst, float* kmean, float* model, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const dim3 grid((n_obs + NT - 1) / NT, n);
    hipLaunchKernelGGL(k_rv_shift, grid, dim3(NT), 0, st, G, stars, oevals, rv, cont1, cont2, kc,
                       kst);
    hipLaunchKernelGGL(k_seq_mean, dim3(2 * n), dim3(128), 0, st, kc, n_obs, kst, kmean);
    hipLaunchKernelGGL(k_rv_combine, grid, dim3(NT), 0, st, G, stars, oevals, rv, model1, model2,
                       kc, kmean, light_ratio, model);
    return hipGetLastError();
}

hipError_t launch_interp_lin(const DevGrid& G, const double* labels, int lstride, int n,
                             float* out, int* status, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_interp_lin, dim3((G.W + 8 * NT - 1) / (8 * NT), n), dim3(NT), 0, st, G,
                       labels, lstride, out, status);
    return hipGetLastError();
}

hipError_t launch_pso_update(const GridGeom* geom, DevPso* pso, DevEval* evals, int n, double* pos,
                             double* vel, const double* bpos, hipStream_t st) {
    hipLaunchKernelGGL(k_pso_update, dim3((n + 127) / 128), dim3(128), 0, st, geom, pso, evals, n,
                       pos, vel, bpos);
    return hipGetLastError();
}

hipError_t launch_pso_best(DevPso* pso, int n_stars, const DevEval* evals, const double* pos,
                           double* bpos, double* bcost, const double* cost, double* trace,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_pso_best, dim3((n_stars + 63) / 64), dim3(64), 0, st, pso, n_stars, evals,
                       pos, bpos, bcost, cost, trace);
    return hipGetLastError();
}

hipError_t launch_eval_init(const EvalInit* in, int n, DevEval* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_eval_init, dim3(n), dim3(64), 0, st, in, n, out);
    return hipGetLastError();
}

hipError_t launch_pso_remap(const int* src_eval, int n_evals, DevEval* evals, const int* src_star,
                            int n_stars, const DevPso* pso_old, DevPso* pso_new,
                            const double* pos_old, double* pos_new, const double* vel_old,
                            double* vel_new, const double* bpos_old, double* bpos_new,
                            const double* bcost_old, double* bcost_new, hipStream_t st) {
    const int n = n_evals > n_stars ? n_evals : n_stars;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pso_remap, dim3((n + 255) / 256), dim3(256), 0, st, src_eval, n_evals,
                       evals, src_star, n_stars, pso_old, pso_new, pos_old, pos_new, vel_old,
                       vel_new, bpos_old, bpos_new, bcost_old, bcost_new);
    return hipGetLastError();
}

hipError_t launch_pso_init(DevPso* pso, int n_stars, const DevEval* evals, double* pos, double* vel,
                           double* bpos, double* bcost, double* cost, double* trace,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_pso_init, dim3((n_stars + 63) / 64), dim3(64), 0, st, pso, n_stars, evals,
                       pos, vel, bpos, bcost, cost, trace);
    return hipGetLastError();
}

}  // namespace pasta
