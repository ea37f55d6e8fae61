// diag.hip -- HBM bandwidth probe: a plain streaming copy of known size.
// Gives the practical bandwidth ceiling next to the 8 TB/s published peak,
// and the reference point that scales rocprof FETCH_SIZE / WRITE_SIZE
// counters to bytes.
#include <hip/hip_runtime.h>

#include "gpe_internal.hpp"
#include "../../include/vaikkari_gpe_diag.h"

namespace {

__global__ void __launch_bounds__(256) k_copy_stream(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                     long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" gpe_status gpe_diag_copy_bandwidth(uint64_t bytes, uint32_t iters, double *ms_per_copy) {
    if (!ms_per_copy || bytes < 4096 || iters == 0) {
        gpe::set_error("gpe_diag_copy_bandwidth: invalid argument");
        return GPE_EINVAL;
    }
    const long long n = (long long)(bytes / sizeof(float4));
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    gpe_status rc = GPE_OK;
    float ms = 0.0f;
    if (hipMalloc(&a, n * sizeof(float4)) != hipSuccess || hipMalloc(&b, n * sizeof(float4)) != hipSuccess) {
        gpe::set_error("gpe_diag_copy_bandwidth: cannot allocate 2 x %llu bytes", (unsigned long long)bytes);
        rc = GPE_ENOMEM;
    }
    if (rc == GPE_OK && (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
                         hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
                         hipMemsetAsync(a, 0, n * sizeof(float4), st) != hipSuccess)) {
        gpe::set_error("gpe_diag_copy_bandwidth: HIP setup failed");
        rc = GPE_EDEVICE;
    }
    if (rc == GPE_OK) {
        const int blocks = 256 * 8 * 4;  // 4 blocks of 256 threads per CU slot, 8 XCDs x 32 CUs
        hipLaunchKernelGGL(k_copy_stream, dim3(blocks), dim3(256), 0, st, (const float4 *)a, (float4 *)b, n);
        (void)hipEventRecord(e0, st);
        for (uint32_t i = 0; i < iters; i++)
            hipLaunchKernelGGL(k_copy_stream, dim3(blocks), dim3(256), 0, st, (const float4 *)a, (float4 *)b, n);
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess ||
            hipGetLastError() != hipSuccess) {
            gpe::set_error("gpe_diag_copy_bandwidth: kernel failed");
            rc = GPE_EDEVICE;
        }
        *ms_per_copy = ms / iters;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
