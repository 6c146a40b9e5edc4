// Calibration of the HBM byte counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE): a plain copy
// of 1 GiB, three times. Each copy reads and writes exactly 1 GiB, so the counters' readings for
// it give the factors that turn the slot kernel's readings into bytes.
#include <hip/hip_runtime.h>
#include <cstdio>

__global__ void hbm_copy_kernel(const int4 *__restrict__ a, int4 *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

int main()
{
    const size_t bytes = 1ull << 30, n = bytes / sizeof(int4);
    int4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess)
        return 1;
    if (hipMemset(a, 1, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess)
        return 1;
    for (int r = 0; r < 3; r++)
        hipLaunchKernelGGL(hbm_copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n);
    if (hipDeviceSynchronize() != hipSuccess)
        return 1;
    printf("copied 3 x %zu bytes\n", bytes);
    return 0;
}
