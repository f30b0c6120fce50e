// synth.hip -- synthetic workload generator (libpasta_synth.so), used by bench.py and
// the full-size tests to fill a resident grid in place.  Not part of the evaluation
// path.  Same recipe as pasta_amd/synth.py (f64 flux, lines applied in ascending
// centre order, a line touches pixels with centre-8w <= lambda < centre+8w,
// uint16(flux * 65535) by truncation).
#include <hip/hip_runtime.h>
#include <stdint.h>

__global__ void k_synth_grid(uint16_t* __restrict__ rows, long long pitch, int W, double first,
                             double step, const double* __restrict__ unit_labels, int n_lines,
                             const double* __restrict__ centre, const double* __restrict__ width,
                             const double* __restrict__ a, const double* __restrict__ b,
                             double max_reach) {
    const int row = blockIdx.y;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= W) return;
    const double wl = pow(10.0, first + step * (double)x);
    const double u0 = unit_labels[3 * row], u1 = unit_labels[3 * row + 1], u2 = unit_labels[3 * row + 2];
    int lo = 0, hi = n_lines;  // first line with centre >= wl - max_reach
    const double t = wl - max_reach;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (centre[mid] < t) lo = mid + 1;
        else hi = mid;
    }
    double flux = 1.0;
    for (int l = lo; l < n_lines && centre[l] <= wl + max_reach; l++) {
        const double c = centre[l], w = width[l];
        if (!(wl >= c - 8.0 * w && wl < c + 8.0 * w)) continue;
        const double z = u0 * a[3 * l] + u1 * a[3 * l + 1] + u2 * a[3 * l + 2] + b[l];
        const double d = 1.0 / (1.0 + exp(-z));
        const double q = (wl - c) / w;
        flux *= 1.0 - d * exp(-0.5 * (q * q));
    }
    rows[(long long)row * pitch + x] = (uint16_t)(flux * 65535.0);
}

__global__ void k_copy(float4* __restrict__ dst, const float4* __restrict__ src, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

extern "C" {

// Fills rows[0..n_models) (row pitch `pitch` elements) of a device grid.
int pasta_synth_fill_grid(uint16_t* rows, int64_t pitch, int n_models, int W, double log10_first,
                          double log10_step, const double* unit_labels_host, int n_lines,
                          const double* centre_host, const double* width_host,
                          const double* a_host, const double* b_host) {
    double *ul, *c, *w, *a, *b;
    if (hipMalloc(&ul, sizeof(double) * 3 * n_models) != hipSuccess) return -1;
    if (hipMalloc(&c, sizeof(double) * n_lines) != hipSuccess) return -1;
    if (hipMalloc(&w, sizeof(double) * n_lines) != hipSuccess) return -1;
    if (hipMalloc(&a, sizeof(double) * 3 * n_lines) != hipSuccess) return -1;
    if (hipMalloc(&b, sizeof(double) * n_lines) != hipSuccess) return -1;
    (void)hipMemcpy(ul, unit_labels_host, sizeof(double) * 3 * n_models, hipMemcpyHostToDevice);
    (void)hipMemcpy(c, centre_host, sizeof(double) * n_lines, hipMemcpyHostToDevice);
    (void)hipMemcpy(w, width_host, sizeof(double) * n_lines, hipMemcpyHostToDevice);
    (void)hipMemcpy(a, a_host, sizeof(double) * 3 * n_lines, hipMemcpyHostToDevice);
    (void)hipMemcpy(b, b_host, sizeof(double) * n_lines, hipMemcpyHostToDevice);
    double max_w = 0.0;
    for (int l = 0; l < n_lines; l++) max_w = width_host[l] > max_w ? width_host[l] : max_w;
    dim3 grid((W + 255) / 256, n_models);
    hipLaunchKernelGGL(k_synth_grid, grid, dim3(256), 0, 0, rows, (long long)pitch, W, log10_first,
                       log10_step, ul, n_lines, c, w, a, b, 8.0 * max_w);
    hipError_t e = hipDeviceSynchronize();
    (void)hipFree(ul); (void)hipFree(c); (void)hipFree(w); (void)hipFree(a); (void)hipFree(b);
    return e == hipSuccess ? 0 : -2;
}

// One plain copy of `bytes` (multiple of 16) between two fresh buffers: the
// reference point for scaling FETCH_SIZE / WRITE_SIZE counters.
int pasta_synth_copy_probe(int64_t bytes) {
    void *s, *d;
    if (bytes % 16) return -1;
    if (hipMalloc(&s, bytes) != hipSuccess) return -1;
    if (hipMalloc(&d, bytes) != hipSuccess) { (void)hipFree(s); return -1; }
    (void)hipMemset(s, 1, bytes);
    (void)hipDeviceSynchronize();
    long long n = bytes / 16;
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, (float4*)d, (const float4*)s, n);
    hipError_t e = hipDeviceSynchronize();
    (void)hipFree(s);
    (void)hipFree(d);
    return e == hipSuccess ? 0 : -2;
}
}
