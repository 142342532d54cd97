// k_probe.hip -- HBM calibration kernels: a plain copy, a read-only sum and a
// write-only fill over a buffer of known size.  They give the achievable
// bandwidth the collide-stream kernels are compared with, and the reference
// point that scales the FETCH_SIZE / WRITE_SIZE counters of rocprofv3.
#include <string.h>

#include "internal.hpp"

namespace dnsg {

// one 16-byte element per thread, one 4 KB chunk per workgroup (the fastest
// plain form measured on MI355X; grid-stride loops and larger chunks are slower)
__global__ __launch_bounds__(256) void k_probe_copy(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                     size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_probe_read(const double2 *__restrict__ src, double *__restrict__ out,
                                                     size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double2 v = src[i];
        if (v.x + v.y == 12345.678) out[0] = v.x;  // keeps the load alive
    }
}

__global__ __launch_bounds__(256) void k_probe_fill(double2 *__restrict__ dst, size_t n, double v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = make_double2(v, v);
}

}  // namespace dnsg

using namespace dnsg;

extern "C" int dnsg_hbm_probe(size_t bytes, int kind, int reps, double *gbs) {
    if (bytes < (1 << 20) || reps < 1 || !gbs || kind < 0 || kind > 2) return DNSG_ERR_ARG;
    const size_t n = bytes / sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    double *o = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = DNSG_OK;
    if (hipMalloc(&a, n * sizeof(double2)) != hipSuccess || hipMalloc(&b, n * sizeof(double2)) != hipSuccess ||
        hipMalloc(&o, 8) != hipSuccess || hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = DNSG_ERR_HIP;
    } else {
        const dim3 grd((unsigned)((n + 255) / 256)), blk(256);
        hipLaunchKernelGGL(k_probe_fill, grd, blk, 0, s, a, n, 1.0);
        hipLaunchKernelGGL(k_probe_fill, grd, blk, 0, s, b, n, 2.0);
        float best = 1e30f;
        for (int r = 0; r < reps + 1 && rc == DNSG_OK; r++) {
            (void)hipEventRecord(e0, s);
            if (kind == 0) hipLaunchKernelGGL(k_probe_copy, grd, blk, 0, s, a, b, n);
            else if (kind == 1) hipLaunchKernelGGL(k_probe_read, grd, blk, 0, s, a, o, n);
            else hipLaunchKernelGGL(k_probe_fill, grd, blk, 0, s, b, n, 3.0);
            (void)hipEventRecord(e1, s);
            if (hipEventSynchronize(e1) != hipSuccess) rc = DNSG_ERR_HIP;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (r > 0 && ms < best) best = ms;  // the first launch warms up
        }
        const double moved = (kind == 0 ? 2.0 : 1.0) * (double)(n * sizeof(double2));
        *gbs = moved / (best * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (o) (void)hipFree(o);
    return rc;
}
