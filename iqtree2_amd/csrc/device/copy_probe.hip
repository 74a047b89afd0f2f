// Plain device-to-device copy of a known size, used to calibrate HBM byte
// counters (rocprofv3 FETCH_SIZE / WRITE_SIZE) against the bytes a kernel
// must move: each launch reads and writes exactly `bytes`.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) k_copy(const double4* __restrict__ src, double4* __restrict__ dst,
                                              size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int main(int argc, char** argv) {
    const size_t bytes = (argc > 1 ? strtoull(argv[1], nullptr, 10) : 256ull) << 20;   // MiB
    const int reps = argc > 2 ? atoi(argv[2]) : 10;
    const size_t n = bytes / sizeof(double4);
    double4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * sizeof(double4)) != hipSuccess || hipMalloc(&b, n * sizeof(double4)) != hipSuccess) {
        fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    if (hipMemset(a, 0, n * sizeof(double4)) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    k_copy<<<4096, 256>>>(a, b, n);
    if (hipEventRecord(e0, 0) != hipSuccess) return 1;
    for (int r = 0; r < reps; r++) k_copy<<<4096, 256>>>(a, b, n);
    if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
    printf("copy_probe: %zu bytes read + %zu bytes written per launch, %d launches, %.3f ms avg, %.1f GB/s\n",
           n * sizeof(double4), n * sizeof(double4), reps, ms / reps,
           2.0 * n * sizeof(double4) * reps / (ms * 1e-3) / 1e9);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
