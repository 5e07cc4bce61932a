// Bandwidth calibration probe (measurement infrastructure, not part of the search): a plain device-to-device copy
// of a known number of bytes, so that hardware byte counters (FETCH_SIZE / WRITE_SIZE) of the search kernels can be
// scaled by what the same counters read for a kernel whose traffic is exactly known.
#include <hip/hip_runtime.h>
#include <cstdint>

__global__ void lbp_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n16) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n16; e += (int64_t)gridDim.x * blockDim.x)
        dst[e] = src[e];
}

extern "C" int lbp_copy(int64_t nbytes, int iters, double *ms_per_copy) {
    if (nbytes <= 0 || nbytes % 16 || iters <= 0) return -1;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, nbytes) != hipSuccess) return -2;
    if (hipMalloc(&b, nbytes) != hipSuccess) { hipFree(a); return -2; }
    hipMemset(a, 1, nbytes);
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    const int64_t n16 = nbytes / 16;
    hipLaunchKernelGGL(lbp_copy_kernel, dim3(4096), dim3(256), 0, 0, (const uint4 *)a, (uint4 *)b, n16);
    hipEventRecord(e0, 0);
    for (int it = 0; it < iters; it++)
        hipLaunchKernelGGL(lbp_copy_kernel, dim3(4096), dim3(256), 0, 0, (const uint4 *)a, (uint4 *)b, n16);
    hipEventRecord(e1, 0);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (ms_per_copy) *ms_per_copy = ms / iters;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}
