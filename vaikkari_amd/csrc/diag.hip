// diag.hip -- HBM bandwidth probes: a plain streaming copy of known size
// (also the reference point that scales rocprof FETCH_SIZE / WRITE_SIZE
// counters to bytes), the stencil's read + read-modify-write pattern without
// the stencil, and a read-only stream.  They give the practical bandwidth
// ceilings next to the 8 TB/s published peak.
#include <hip/hip_runtime.h>

#include "gpe_internal.hpp"
#include "../../include/vaikkari_gpe_diag.h"

namespace {

__global__ void __launch_bounds__(256) k_copy_stream(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                     long long n) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// The stencil's memory pattern without the stencil: read prev, read-modify-
// write next with non-temporal access, 16 B (one FP64 complex) per lane,
// one element per thread (as the gather pass launches one block per tile).
__global__ void __launch_bounds__(256) k_rmw_stream(const double2 *__restrict__ prev, double2 *__restrict__ next,
                                                    long long n) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= n) return;
    const double2 p = prev[i];
    double2 q;
    q.x = __builtin_nontemporal_load(&next[i].x);
    q.y = __builtin_nontemporal_load(&next[i].y);
    q.x = q.x + p.x;
    q.y = q.y + p.y;
    __builtin_nontemporal_store(q.x, &next[i].x);
    __builtin_nontemporal_store(q.y, &next[i].y);
}

// Read-only stream; the buffer holds zeros, so the store never happens but
// keeps the loads alive.
__global__ void __launch_bounds__(256) k_read_stream(const double2 *__restrict__ src, double *__restrict__ sink,
                                                     long long n) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= n) return;
    const double2 v = src[i];
    if (v.x + v.y == 1.0) sink[0] = v.x;
}

}  // namespace

extern "C" gpe_status gpe_diag_stream_bandwidth(uint32_t kind, uint64_t bytes, uint32_t iters, double *ms_per_pass) {
    if (!ms_per_pass || bytes < 4096 || iters == 0 || kind > GPE_DIAG_READ) {
        gpe::set_error("gpe_diag_stream_bandwidth: invalid argument");
        return GPE_EINVAL;
    }
    if (kind == GPE_DIAG_COPY) return gpe_diag_copy_bandwidth(bytes, iters, ms_per_pass);
    const long long n = (long long)(bytes / sizeof(double2));
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    gpe_status rc = GPE_OK;
    float ms = 0.0f;
    if (hipMalloc(&a, n * sizeof(double2)) != hipSuccess || hipMalloc(&b, n * sizeof(double2)) != hipSuccess) {
        gpe::set_error("gpe_diag_stream_bandwidth: cannot allocate 2 x %llu bytes", (unsigned long long)bytes);
        rc = GPE_ENOMEM;
    }
    if (rc == GPE_OK && (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
                         hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
                         hipMemsetAsync(a, 0, n * sizeof(double2), st) != hipSuccess ||
                         hipMemsetAsync(b, 0, n * sizeof(double2), st) != hipSuccess)) {
        gpe::set_error("gpe_diag_stream_bandwidth: HIP setup failed");
        rc = GPE_EDEVICE;
    }
    if (rc == GPE_OK) {
        const unsigned blocks = (unsigned)((n + 255) / 256);
        auto pass = [&] {
            if (kind == GPE_DIAG_RMW)
                hipLaunchKernelGGL(k_rmw_stream, dim3(blocks), dim3(256), 0, st, (const double2 *)a, (double2 *)b, n);
            else
                hipLaunchKernelGGL(k_read_stream, dim3(blocks), dim3(256), 0, st, (const double2 *)a, (double *)b, n);
        };
        pass();
        (void)hipEventRecord(e0, st);
        for (uint32_t i = 0; i < iters; i++) pass();
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess ||
            hipGetLastError() != hipSuccess) {
            gpe::set_error("gpe_diag_stream_bandwidth: kernel failed");
            rc = GPE_EDEVICE;
        }
        *ms_per_pass = ms / iters;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

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
