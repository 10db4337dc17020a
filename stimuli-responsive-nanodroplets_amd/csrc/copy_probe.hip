// copy_probe — calibration for the FETCH_SIZE / WRITE_SIZE counters.
//
// Copies 512 MiB (twice the 256 MB last-level cache, so the traffic reaches
// HBM) three times with a plain grid-stride kernel.  Each dispatch moves
// exactly 512 MiB in and 512 MiB out; the per-dispatch counter values of
//   rocprofv3 --pmc FETCH_SIZE --kernel-trace -- bin/copy_probe
//   rocprofv3 --pmc WRITE_SIZE --kernel-trace -- bin/copy_probe
// divided into 512 MiB give the factors that turn the counters (KiB per
// dispatch) into bytes for the solver's kernels.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void __launch_bounds__(256) k_copy(const double4 *__restrict__ a, double4 *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

int main()
{
    const size_t bytes = size_t(512) << 20;
    const size_t n = bytes / sizeof(double4);
    double4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        std::printf("copy_probe: allocation failed\n");
        return 1;
    }
    if (hipMemset(a, 0, bytes) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    float best = 1e30f;
    for (int r = 0; r < 3; ++r) {
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) {
            std::printf("copy_probe: kernel failed\n");
            return 1;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) best = ms;
    }
    std::printf("copy_probe: %zu bytes in + %zu bytes out per dispatch, best %.3f ms (%.0f GB/s)\n", bytes, bytes,
                best, 2.0 * bytes / (best * 1e-3) / 1e9);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
