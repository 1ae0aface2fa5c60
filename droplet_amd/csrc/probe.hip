// probe.hip — stream-copy bandwidth probe: the achievable HBM rate of a plain
// read+write kernel, the yardstick next to which the step kernel's algorithmic
// bandwidth is reported, and the known-size kernel that calibrates the
// FETCH_SIZE / WRITE_SIZE counters.
#include <hip/hip_runtime.h>

#include "../../include/droplet.h"
#include "lattice_impl.h"

namespace {

__global__ __launch_bounds__(256) void stream_copy(const double2* __restrict__ src, double2* __restrict__ dst,
                                                   int64_t n2) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2; i += stride) dst[i] = src[i];
}

// K independent SoA streams copied at once, x fastest with one double per
// lane, non-temporal stores: the access pattern of the step kernel (15 + 5
// arrays in, 15 + 5 out) without its arithmetic
template <int K>
__global__ __launch_bounds__(256) void stream_copy_soa(const double* __restrict__ src, double* __restrict__ dst,
                                                       int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        double v[K];
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = src[k * n + i];
#pragma unroll
        for (int k = 0; k < K; k++) __builtin_nontemporal_store(v[k], &dst[k * n + i]);
    }
}

}  // namespace

extern "C" int drp_copy_bandwidth_soa(size_t bytes_per_stream, int streams, int reps, double* gbs) {
    if (!gbs || bytes_per_stream < 4096 || reps < 1 || (streams != 10 && streams != 20 && streams != 40))
        return drp_fail(DRP_ERR_ARG, "bad probe arguments (streams must be 10, 20 or 40)");
    const int64_t n = (int64_t)(bytes_per_stream / 8);
    const size_t total = (size_t)n * 8 * streams;
    double *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = DRP_OK;
    float ms = 0.0f;
    const int blocks = 256 * 8 * 4;
    auto launch = [&]() {
        if (streams == 10) hipLaunchKernelGGL(stream_copy_soa<10>, dim3(blocks), dim3(256), 0, st, a, b, n);
        else if (streams == 20) hipLaunchKernelGGL(stream_copy_soa<20>, dim3(blocks), dim3(256), 0, st, a, b, n);
        else hipLaunchKernelGGL(stream_copy_soa<40>, dim3(blocks), dim3(256), 0, st, a, b, n);
    };
    if (hipMalloc((void**)&a, total) != hipSuccess || hipMalloc((void**)&b, total) != hipSuccess) {
        rc = drp_fail(DRP_ERR_ALLOC, "probe allocation of 2 x %zu bytes failed", total);
        goto done;
    }
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(a, 0, total, st) != hipSuccess) {
        rc = drp_fail(DRP_ERR_HIP, "probe setup failed");
        goto done;
    }
    launch();
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; r++) launch();
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = drp_fail(DRP_ERR_HIP, "probe timing failed");
        goto done;
    }
    *gbs = 2.0 * (double)total * reps / (ms * 1e-3) / 1e9;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

extern "C" int drp_copy_bandwidth(size_t bytes, int reps, double* gbs) {
    if (!gbs || bytes < 4096 || reps < 1) return drp_fail(DRP_ERR_ARG, "bad probe arguments");
    bytes &= ~(size_t)15;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = DRP_OK;
    float ms = 0.0f;
    const int64_t n2 = (int64_t)(bytes / 16);
    const int blocks = 256 * 8 * 4;  // 8 XCDs x 32 CUs x 4 workgroups
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        rc = drp_fail(DRP_ERR_ALLOC, "probe allocation of 2 x %zu bytes failed", bytes);
        goto done;
    }
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(a, 0, bytes, st) != hipSuccess) {
        rc = drp_fail(DRP_ERR_HIP, "probe setup failed");
        goto done;
    }
    hipLaunchKernelGGL(stream_copy, dim3(blocks), dim3(256), 0, st, (const double2*)a, (double2*)b, n2);
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(stream_copy, dim3(blocks), dim3(256), 0, st, (const double2*)a, (double2*)b, n2);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = drp_fail(DRP_ERR_HIP, "probe timing failed");
        goto done;
    }
    *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
