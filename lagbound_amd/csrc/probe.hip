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

// Row-gather probe (measurement infrastructure): the gathers of an evaluation -- one 16*S-byte V row per index, S
// lanes of 16 bytes per row, EB rows in flight per slot -- without any arithmetic around them, in 1024-thread
// workgroups. With H > 0 the first H rows (the highest-degree vertices under the degree numbering) are staged in LDS
// once per workgroup and gathered from there. It bounds what a gather-bound kernel can reach on a given index
// stream, and what staging the hub rows in LDS can add to it.
template <int EB, int S, bool LDS>
__global__ __launch_bounds__(1024) void lbp_gather_kernel(const uint4 *__restrict__ V, const int *__restrict__ idx,
                                                          int64_t m, int H, float *__restrict__ out) {
    extern __shared__ uint4 hub[];
    if (LDS) {
        for (int e = threadIdx.x; e < H * S; e += 1024) hub[e] = V[e];
        __syncthreads();
    }
    const int slot = threadIdx.x / S, q = threadIdx.x % S;
    const int64_t nsl = (int64_t)gridDim.x * (1024 / S);
    float acc = 0.f;
    for (int64_t e0 = (int64_t)blockIdx.x * (1024 / S) + slot; e0 < m; e0 += nsl * EB) {
        int jj[EB];
#pragma unroll
        for (int t = 0; t < EB; t++) {
            const int64_t e = e0 + t * nsl;
            jj[t] = e < m ? idx[e] : 0;
        }
        uint4 u[EB];
#pragma unroll
        for (int t = 0; t < EB; t++) {
            if (LDS && jj[t] < H) u[t] = hub[jj[t] * S + q];
            else u[t] = V[(int64_t)jj[t] * S + q];
        }
#pragma unroll
        for (int t = 0; t < EB; t++)
            acc += __uint_as_float(u[t].x) + __uint_as_float(u[t].y) + __uint_as_float(u[t].z) + __uint_as_float(u[t].w);
    }
    out[(int64_t)blockIdx.x * 1024 + threadIdx.x] = acc;
}

template <int S, bool LDS>
static void lbp_gather_launch(int eb, int grid, size_t lds, const uint4 *V, const int *idx, int64_t m, int H, float *out) {
    if (eb == 4) hipLaunchKernelGGL((lbp_gather_kernel<4, S, LDS>), dim3(grid), dim3(1024), lds, 0, V, idx, m, H, out);
    else hipLaunchKernelGGL((lbp_gather_kernel<8, S, LDS>), dim3(grid), dim3(1024), lds, 0, V, idx, m, H, out);
}

// Times the gathers of idx[0..m) into n rows of 16*S bytes (S = 2 or 4), EB = 4 or 8 in flight per slot, `grid`
// workgroups (at most 4096), with the first H rows in LDS (H = 0: none; H * S * 16 bytes must fit a workgroup's LDS).
// Returns 0 and the mean ms per pass over `iters` passes.
extern "C" int lbp_gather(const int *idx_host, int64_t m, int64_t n, int S, int eb, int grid, int H, int iters,
                          double *ms_per_pass) {
    if (m <= 0 || n <= 0 || (S != 2 && S != 4) || (eb != 4 && eb != 8) || grid < 1 || grid > 4096 || H < 0 ||
        H > n || iters <= 0)
        return -1;
    for (int64_t e = 0; e < m; e++)
        if (idx_host[e] < 0 || idx_host[e] >= n) return -1;
    const size_t lds = (size_t)H * S * 16;
    int dev = 0, maxlds = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (lds > (size_t)maxlds) return -1;
    uint4 *V = nullptr;
    int *idx = nullptr;
    float *out = nullptr;
    if (hipMalloc(&V, (size_t)n * S * 16) != hipSuccess) return -2;
    if (hipMalloc(&idx, sizeof(int) * (size_t)m) != hipSuccess) { hipFree(V); return -2; }
    if (hipMalloc(&out, sizeof(float) * 1024 * (size_t)grid) != hipSuccess) { hipFree(V); hipFree(idx); return -2; }
    hipMemset(V, 0, (size_t)n * S * 16);
    hipMemcpy(idx, idx_host, sizeof(int) * (size_t)m, hipMemcpyHostToDevice);
    if (H) {
        if (S == 4) {
            hipFuncSetAttribute((const void *)lbp_gather_kernel<4, 4, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipFuncSetAttribute((const void *)lbp_gather_kernel<8, 4, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        } else {
            hipFuncSetAttribute((const void *)lbp_gather_kernel<4, 2, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            hipFuncSetAttribute((const void *)lbp_gather_kernel<8, 2, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        }
    }
    auto launch = [&]() {
        if (S == 4) { if (H) lbp_gather_launch<4, true>(eb, grid, lds, V, idx, m, H, out); else lbp_gather_launch<4, false>(eb, grid, 0, V, idx, m, H, out); }
        else { if (H) lbp_gather_launch<2, true>(eb, grid, lds, V, idx, m, H, out); else lbp_gather_launch<2, false>(eb, grid, 0, V, idx, m, H, out); }
    };
    launch();
    hipDeviceSynchronize();
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, 0);
    for (int it = 0; it < iters; it++) launch();
    hipEventRecord(e1, 0);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (ms_per_pass) *ms_per_pass = ms / iters;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(V);
    hipFree(idx);
    hipFree(out);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}
