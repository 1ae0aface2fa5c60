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
template <int K, bool NTL = false>
__global__ __launch_bounds__(256) void stream_copy_soa(const double* __restrict__ src, double* __restrict__ dst,
                                                       int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        double v[K];
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = NTL ? __builtin_nontemporal_load(&src[k * n + i]) : src[k * n + i];
#pragma unroll
        for (int k = 0; k < K; k++) __builtin_nontemporal_store(v[k], &dst[k * n + i]);
    }
}

// 16-byte clang vector: the non-temporal builtins take native vector types
typedef double dvec2 __attribute__((ext_vector_type(2)));

// One pass, no grid stride: each lane moves U 16-byte items spaced one
// workgroup apart, all U loads issued before the first store.  NT selects
// non-temporal loads and stores.
template <int U, bool NT>
__global__ __launch_bounds__(256) void stream_copy_unrolled(const dvec2* __restrict__ src, dvec2* __restrict__ dst,
                                                            int64_t n2) {
    const int64_t base = (int64_t)blockIdx.x * (256 * U) + threadIdx.x;
    dvec2 v[U];
#pragma unroll
    for (int k = 0; k < U; k++) {
        const int64_t i = base + k * 256;
        if (i < n2) v[k] = NT ? __builtin_nontemporal_load(&src[i]) : src[i];
    }
#pragma unroll
    for (int k = 0; k < U; k++) {
        const int64_t i = base + k * 256;
        if (i < n2) {
            if (NT) __builtin_nontemporal_store(v[k], &dst[i]);
            else dst[i] = v[k];
        }
    }
}

// read-only and write-only passes of the same shape: the read side sums what it
// reads and stores one value per lane only if the sum is exactly -1 (never, for
// the zero-filled source), so that the loads cannot be removed
template <int U>
__global__ __launch_bounds__(256) void stream_read(const dvec2* __restrict__ src, double* __restrict__ sink,
                                                   int64_t n2) {
    const int64_t base = (int64_t)blockIdx.x * (256 * U) + threadIdx.x;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < U; k++) {
        const int64_t i = base + k * 256;
        if (i < n2) {
            const dvec2 v = __builtin_nontemporal_load(&src[i]);
            acc += v.x + v.y;
        }
    }
    if (acc == -1.0) sink[blockIdx.x * 256 + threadIdx.x] = acc;
}

template <int U>
__global__ __launch_bounds__(256) void stream_write(dvec2* __restrict__ dst, int64_t n2) {
    const int64_t base = (int64_t)blockIdx.x * (256 * U) + threadIdx.x;
#pragma unroll
    for (int k = 0; k < U; k++) {
        const int64_t i = base + k * 256;
        if (i < n2) __builtin_nontemporal_store(dvec2{0.0, 1.0}, &dst[i]);
    }
}

}  // namespace

extern "C" int drp_copy_bandwidth_kind(size_t bytes, int kind, int reps, double* gbs) {
    if (!gbs || bytes < 4096 || reps < 1 || kind < 0 || kind > 6)
        return drp_fail(DRP_ERR_ARG, "bad probe arguments (kind 0..6)");
    bytes &= ~(size_t)15;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = DRP_OK;
    float ms = 0.0f;
    const int64_t n2 = (int64_t)(bytes / 16);
    constexpr int U = 4;
    const int64_t blocks_u = (n2 + 256 * U - 1) / (256 * U);
    double moved = 2.0 * (double)bytes;
    auto launch = [&]() -> hipError_t {
        switch (kind) {
            case 0:
                hipLaunchKernelGGL(stream_copy, dim3(256 * 8 * 4), dim3(256), 0, st, (const double2*)a, (double2*)b, n2);
                break;
            case 1: return hipMemcpyDtoDAsync(b, a, bytes, st);
            case 2:
                hipLaunchKernelGGL((stream_copy_unrolled<U, true>), dim3(blocks_u), dim3(256), 0, st, (const dvec2*)a,
                                   (dvec2*)b, n2);
                break;
            case 3:
                hipLaunchKernelGGL((stream_copy_unrolled<U, false>), dim3(blocks_u), dim3(256), 0, st,
                                   (const dvec2*)a, (dvec2*)b, n2);
                break;
            case 4:
                hipLaunchKernelGGL((stream_read<U>), dim3(blocks_u), dim3(256), 0, st, (const dvec2*)a, (double*)b, n2);
                break;
            case 6:  // 20 structure-of-arrays streams of bytes / 20, one double per lane
                hipLaunchKernelGGL((stream_copy_soa<20, true>), dim3(256 * 8 * 4), dim3(256), 0, st, (const double*)a,
                                   (double*)b, (int64_t)(bytes / 160));
                break;
            default: hipLaunchKernelGGL((stream_write<U>), dim3(blocks_u), dim3(256), 0, st, (dvec2*)b, n2); break;
        }
        return hipGetLastError();
    };
    if (kind == 4 || kind == 5) moved = (double)bytes;
    if (kind == 6) moved = 2.0 * (double)(bytes / 160 * 160);
    // the read-only kernel's sink (never written for a zero source) spans
    // blocks_u * 256 doubles <= bytes / 8 bytes, inside b
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        rc = drp_fail(DRP_ERR_ALLOC, "probe allocation of 2 x %zu bytes failed", bytes);
        goto done;
    }
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(a, 0, bytes, st) != hipSuccess) {
        rc = drp_fail(DRP_ERR_HIP, "probe setup failed");
        goto done;
    }
    if (launch() != hipSuccess) {
        rc = drp_fail(DRP_ERR_HIP, "probe launch failed");
        goto done;
    }
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; r++) (void)launch();
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = drp_fail(DRP_ERR_HIP, "probe timing failed");
        goto done;
    }
    *gbs = moved * reps / (ms * 1e-3) / 1e9;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

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
