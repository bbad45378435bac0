// Calibration program for HBM traffic counters: one plain copy kernel that reads and writes a
// known number of bytes (argument: bytes per buffer, default 256 MiB), run a few times. Under
// rocprofv3 --pmc FETCH_SIZE (or WRITE_SIZE) the counter's reading per dispatch, divided by the
// bytes moved, is the scale factor applied to the same counter on the engine's kernels.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void k_copy(const float4* __restrict__ a, float4* __restrict__ b, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

int main(int argc, char** argv) {
    const long bytes = argc > 1 ? atol(argv[1]) : (256L << 20);
    const long n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc((void**)&a, n * 16) != hipSuccess || hipMalloc((void**)&b, n * 16) != hipSuccess) return 1;
    if (hipMemset(a, 0, n * 16) != hipSuccess) return 1;
    for (int r = 0; r < 5; ++r) hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, 0, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    printf("copy_probe: %ld bytes read and %ld bytes written per dispatch, 5 dispatches\n", n * 16, n * 16);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
