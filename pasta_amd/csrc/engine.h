// engine.h -- device-visible descriptors shared by the kernels and the host driver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "geom.h"

namespace pasta {

// Tile and block shapes are build-time parameters (pasta_amd/tune.py rebuilds the library
// with other values and times the batched iteration with each).
#ifndef PASTA_TILE
#define PASTA_TILE 3312
#endif
#ifndef PASTA_CONT_BLOCK
#define PASTA_CONT_BLOCK 4096
#endif
#ifndef PASTA_CHI2_GROUP
#define PASTA_CHI2_GROUP 8
#endif
constexpr int TILE = PASTA_TILE;  // synthetic pixels per model-kernel workgroup: the tile
                                  // and a 7-tap instrumental halo are whole 256-output MFMA
                                  // tiles of the rotation
static_assert(TILE % 256 == 240 && TILE >= 752 && TILE <= 3824,
              "tile = 256 k - 16 pixels, 3 <= k <= 15 (k_model's LDS fits 4 workgroups per CU)");
constexpr int NT = 256;           // threads per workgroup (4 wavefronts)
constexpr int HD_MAX = 127;       // max half-width of the instrumental kernel
constexpr int KROT_MAX = 2048;    // FFTSIZE of the reference (convolve_rv.rs:433)
// Rotation kernels of up to KB_MAX taps run on f16 matrix cores with a two-way operand
// split; their taps are kept as 2 split parts x 2 alignments x TW_MAX packed f16 pairs.
constexpr int KB_MAX = 1536;
constexpr int TW_MAX = (32 * ((KB_MAX + 15 + 31) / 32) + 32) / 2;  // 800
// per-evaluation rotation buffer: f32 taps, then the split pairs (or the kernel's spectrum,
// 2049 complex values, for the FFT path)
constexpr int ROT_STRIDE = KROT_MAX + 4 * TW_MAX;
// Default crossover of the rotation stage: kernels of at least this many taps are convolved
// by FFT (if the tile and kernel fit one 4096-point transform), shorter ones directly on the
// matrix cores.  Measured on the batched workload (128 stars x 64 particles, full grid): the
// FFT stage costs about the same at any size, the f16 x 2 products stay cheaper up to
// KB_MAX taps, and the f32 products above KB_MAX cost more than the FFT.  Set per grid with
// pasta_grid_set_fft_taps.
constexpr int FFT_MIN_TAPS = KB_MAX + 1;
constexpr int FFT_N = 4096;       // real points of the FFT path's transforms
constexpr int FFT_M = FFT_N / 2;  // their complex half-length
// The FFT path convolves a tile's rotation output range (nR <= TILE + 1 + 2 hd values) in
// one transform, or in two pieces of fft_piece(nR) values when the range and the kernel do
// not fit one; a piece's interpolated input (<= 7 + piece + K - 1 samples) stays within
// FFT_N samples, so no output index of the circular convolution wraps.
__host__ __device__ constexpr int fft_piece(int nR) { return ((nR + 1) / 2 + 7) & ~7; }
__host__ __device__ constexpr bool fft_fits(int K, int hd) {
    return 7 + fft_piece(TILE + 1 + 2 * hd) + K - 1 <= FFT_N;
}
constexpr int CH_MAX = 32;        // max chunks
constexpr int P1_MAX = 32;        // max continuum parameters per chunk / linear columns
constexpr int BLK = PASTA_CONT_BLOCK;  // observed pixels per reduction block
static_assert(BLK % 1024 == 0 && BLK >= 1024, "reduction blocks of whole 1024-pixel units");
// evaluations of one star per chi2 workgroup (polynomial continua of up to 16 parameters)
constexpr int CHI2_GROUP = PASTA_CHI2_GROUP;
static_assert(CHI2_GROUP == 4 || CHI2_GROUP == 8 || CHI2_GROUP == 16, "chi2 group of 4, 8 or 16");

enum DispKind : int { DISP_NONE = 0, DISP_FIXED = 1, DISP_VARIABLE = 2 };
enum ContKind : int { CONT_CHUNK = 0, CONT_LINEAR = 1, CONT_FIXED = 2, CONT_CONSTANT = 3 };

// Continuum-stage kernel variant for one class of stars (one continuum kind and
// parameter count): pm = register slots per design row, e = evaluations of one star
// per workgroup (they share its design rows), f = parameter count fixed at compile
// time (0: read from the star).
struct KVariant {
    int pm, e, f;
};
__host__ __device__ constexpr KVariant mom_variant(int kind, int p1) {
    // chunk and linear: the matrix-core kernel, 64 evaluations per workgroup; constant: the
    // sequential fold of the quotients, 16 evaluations per workgroup
    return (kind == CONT_LINEAR || kind == CONT_CHUNK) ? KVariant{p1 <= 16 ? 16 : 32, 64, 0}
           : kind == CONT_CONSTANT                      ? KVariant{1, 16, 0}
                                                        : KVariant{1, 8, 0};
}
__host__ __device__ constexpr KVariant chi2_variant(int kind, int p1) {
    return (kind == CONT_CONSTANT || kind == CONT_FIXED) ? KVariant{1, 8, 0}
           : (p1 == 5 || p1 == 9 || p1 == 11)           ? KVariant{16, CHI2_GROUP, p1}
           : p1 <= 16                                    ? KVariant{16, CHI2_GROUP, 0}
           : kind == CONT_LINEAR                         ? KVariant{32, 16, 0}
                                                         : KVariant{32, 4, 0};
}

// Stars of one (continuum kind, parameter count) class in a batch: their moment and chi2
// blocks are [mom_off, mom_off + n_mom) and [chi2_off, chi2_off + n_chi2) of the tables.
struct ContClass {
    int kind, p1, mom_off, n_mom, chi2_off, n_chi2;
};

struct DevGrid {
    const uint16_t* spectra;  // n_models rows of `pitch` uint16 (pitch % 8 == 0)
    const float* factors;     // per-model f32 factor (max/65535) or nullptr
    const GridGeom* geom;
    long long pitch;
    int W, Wq;                // pixels; first pixel of the n % 128 tail
    double sg_first, sg_step; // log10 of the first wavelength, log10 step
    double synth_first_wl, synth_last_wl;  // 10^first, 10^(first + step*W)
    int n_models_bins;        // n_models + 1 (sort bins; the last holds failed evaluations)
    int fft_min_taps;         // rotation kernels of at least this many taps use the FFT path
    int prep_taps;            // k_prep's LDS holds kernels of up to this many composed taps (a
                              // bound from the launch's largest vsini; KROT_MAX when unknown)
    int stream_models;        // 1: k_model writes the resampled models past the caches (a
                              // launch whose models exceed the MALL; per launch)
};
// the MI355X's last-level (MALL) cache: launches with more model bytes stream them
constexpr long long MALL_BYTES = 256LL << 20;

struct DevStar {
    int N, disp_kind, Kd, cont_kind;
    int n_chunks, p1, ignore, n_params;
    int n_mblocks, n_cblocks, n_tiles, n_mslots;  // moment blocks, chi2 blocks, tiles,
                                                  // moment partial slots
    double chi2_div, wl0, wlN;
    const float* flux;
    const float* ivar;   // 1 / variance (f32), the chi2 weight
    const double* fidx;       // float index of each observed pixel in the synthetic grid
    const double* wl;         // observed wavelengths
    const float* dkern;       // fixed: Kd taps; variable: W rows of Kd taps
    const double* chunk_lohi; // per chunk: wl[start], wl[stop-1]
    const int* startstop;     // per chunk: start, stop
    const double* chol;       // per chunk (or linear): p1 x p1 lower Cholesky factor of AᵀA
    const float* fac;         // per observed pixel: overlap blend factor
    const float* design;      // linear: PT = 16 (p1 <= 16) or 32 columns, column k at
                              // k * N4 (N4 = N rounded up to 4), zero padded
    const float* fixedc;      // fixed continuum
    const float* cdesign;     // chunk: per chunk c, p1 rows of f32(powi(x^, k)) from pixel
                              // startstop[2c] & ~3, row stride rounded up to 4, zero padded
    const int* cdoff;         // chunk c's design at cdesign + cdoff[c]
    const int4* mblocks;      // moment blocks (chunk, i0, i1, second chunk or -1): a block
                              // in the overlap of two chunks carries both
    const int2* mslot;        // per moment block: the partial slots of its chunk(s)
    const int* mb_off;        // chunk c owns partial slots [mb_off[c], mb_off[c+1])
    const int4* cblocks;      // chi2 blocks (kind, chunk, i0, i1); kind 1 = overlap
};

// Per-evaluation record, written by k_prep.
struct DevEval {
    int star;
    int status;
    int krot;        // rotation taps (composed taps when comp); 0 = pass-through
    int tile_base;   // first synthetic tile touched
    int cmax;        // exclusive upper bound of readable model-segment indices
    int active;      // 0 = skip (finished / failed star in the batched path)
    int key;         // storage row of the first present neighbour: groups evaluations
                     // that gather the same rows (n_models for failed ones)
    int rows_hash;   // hash of idx[]: evaluations with equal rows can share their gathers
    int comp;        // 1: the krot taps are the rotation kernel convolved with the fixed
                     // instrumental kernel, so k_model has no instrumental stage
    int fft;         // 1: the rotation runs as an FFT convolution; the tap slot holds the
                     // kernel's spectrum (k_rot_spectrum)
    float rsc;       // f16 rotation: power of 2 the interpolated values are scaled by before
                     // their split (|value| * rsc < 2^14)
    float rinv;      // and 1 / (rsc * tap scale): undoes both scales on the rotation's output
    int nrow;        // present neighbours: rowp/fw hold them first, in the neighbour order,
                     // then rows of weight 0 (row 0) up to 64
    double shift_log;
    long long m_off;  // offset of this evaluation's resampled model in the model buffer
    long long mp_off; // offset of its continuum-moment partials (doubles)
    long long cp_off; // offset of its chi2 partials (doubles)
    long long pf_off; // offset of its continuum parameters (floats)
    int idx[64];
    float w[64];
    float fw[64];     // factor * weight of the present neighbours (nrow), then 0
    const uint16_t* rowp[64];  // their grid rows, then row 0
};

// Rows of a k_model work item of two or three evaluations (k_union): the union of their
// present rows in storage order (= each one's neighbour order), every evaluation's weight
// on each (0 where the row is not its neighbour), rows of weight 0 (row 0) up to a multiple
// of 8.  n < 0: the union exceeds UNION_MAX rows; the evaluations then run one at a time.
constexpr int UNION_MAX = 128;
// consecutive evaluations (in neighbour-row order) share a work item when their present rows
// have at most this many distinct rows together: the same cell, or cells one node apart
#ifndef PASTA_LINK_MAX
#define PASTA_LINK_MAX 80
#endif
constexpr int LINK_MAX = PASTA_LINK_MAX;
struct DevUnion {
    int n;
    int pad[3];
    const uint16_t* rowp[UNION_MAX];
    float fw[3][UNION_MAX];
};

// What the host sets in an evaluation record (the rest is written by k_prep); expanded
// into DevEval on the device by k_eval_init, so a batch uploads 40 bytes per evaluation.
struct EvalInit {
    int star;
    int active;
    long long m_off, mp_off, cp_off, pf_off;
};

// Batched PSO state per star (device).
struct DevPso {
    Bounds b;
    PsoParams s;
    uint64_t key_iter;
    int P;            // particles
    int first_eval;   // index of particle 0's evaluation
    int iter;         // next iteration number (1-based)
    int max_iters;
    int status;       // 0 running, ST_* error
    int done;
    double gcost;
    double gbest[5];
};

}  // namespace pasta
