// noalloc.hip — drp_reduce_noalloc: the diagnostics reductions of drp_reduce
// without any device allocation, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// drp_reduce keeps its per-workgroup partial sums in a buffer the lattice
// allocates on first use and finishes the compensated sum on the host.  Here
// the partial sums, the final sums and the phi-range words live in one
// file-scope device array, `scratch`, which every GPU the library runs on holds
// until the process ends (REDUCE_NOALLOC_MAX_NODES / 256 * NPART + 16 doubles,
// 10 MiB + 128 B).  A second kernel forms the same Neumaier-compensated sums in
// the same order as drp_reduce's host loop, so both entry points return
// identical statistics.  Calls that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/droplet.h"
#include "d3q15.h"
#include "kernels.h"
#include "lattice_impl.h"

using namespace drp;

namespace {

constexpr int64_t REDUCE_NOALLOC_MAX_NODES = DRP_REDUCE_NOALLOC_MAX_NODES;  // 2^25, e.g. 512 x 256 x 256
constexpr int64_t MAX_BLOCKS = REDUCE_NOALLOC_MAX_NODES / 256;
constexpr int64_t SCRATCH_DOUBLES = MAX_BLOCKS * NPART + 16;
static_assert(REDUCE_NOALLOC_MAX_NODES % 256 == 0, "whole reduce workgroups");

// partial sums [MAX_BLOCKS][NPART] | final sums [NPART] | phi range slot (2 words) | pad
__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

// Neumaier-compensated sums of the nblocks x NPART partials, block order,
// component k by thread k (the order of drp_reduce's host loop); the chunk of
// partials is staged through LDS by the whole workgroup.  Also copies the phi
// range slot next to the sums, so one copy returns everything.
constexpr int CHUNK = 384;
__global__ __launch_bounds__(256) void reduce_final(const double* __restrict__ partial, int nblocks,
                                                    const uint64_t* __restrict__ range_slot, double* __restrict__ out) {
    __shared__ double sh[CHUNK * NPART];
    double s = 0.0, comp = 0.0;
    const int k = threadIdx.x;
    for (int b0 = 0; b0 < nblocks; b0 += CHUNK) {
        const int nb = min(CHUNK, nblocks - b0);
        __syncthreads();
        for (int i = threadIdx.x; i < nb * NPART; i += blockDim.x) sh[i] = partial[(int64_t)b0 * NPART + i];
        __syncthreads();
        if (k < NPART) {
            for (int b = 0; b < nb; b++) {
                const double v = sh[b * NPART + k], t = s + v;
                comp += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
                s = t;
            }
        }
    }
    if (k < NPART) out[k] = s + comp;
    if (k == 0) {
        uint64_t* w = reinterpret_cast<uint64_t*>(out + NPART);
        w[0] = range_slot[0];
        w[1] = range_slot[1];
    }
}

// device address of `scratch` per device, looked up once per device with that
// device current (the host-side name is not a device address)
std::mutex g_mu;
double* g_scratch[64] = {};

int scratch_for(int dev, double** out) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (dev < 0 || dev >= 64) return drp_fail(DRP_ERR_ARG, "device %d out of range", dev);
    if (!g_scratch[dev]) {
        int prev = -1;
        if (hipGetDevice(&prev) != hipSuccess) return drp_fail(DRP_ERR_HIP, "hipGetDevice failed");
        if (prev != dev && hipSetDevice(dev) != hipSuccess) return drp_fail(DRP_ERR_HIP, "hipSetDevice(%d) failed", dev);
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != dev) (void)hipSetDevice(prev);
        if (e != hipSuccess || !p)
            return drp_fail(DRP_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch[dev] = (double*)p;
    }
    *out = g_scratch[dev];
    return DRP_OK;
}

}  // namespace

extern "C" int drp_reduce_noalloc(drp_lattice* L, drp_stats* st) {
    if (!L || !st) return drp_fail(DRP_ERR_ARG, "null argument");
    if (L->n > REDUCE_NOALLOC_MAX_NODES)
        return drp_fail(DRP_ERR_ARG, "drp_reduce_noalloc: %lld nodes exceed its maximum of %lld",
                        (long long)L->n, (long long)REDUCE_NOALLOC_MAX_NODES);
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    if (L->cfg.precision != DRP_FP64)
        return drp_fail(DRP_ERR_UNSUPPORTED, "drp_reduce_noalloc: FP64 lattices only (the FP32 FP64 view allocates)");
    double* ptr = nullptr;
    int rc = scratch_for(L->dev, &ptr);
    if (rc) return rc;
    double* partial = ptr;
    double* final_ = ptr + MAX_BLOCKS * NPART;
    // launch_reduce_fp64 runs one 256-node workgroup per partial-sum record
    if ((L->n + 255) / 256 > MAX_BLOCKS) return drp_fail(DRP_ERR_ARG, "drp_reduce_noalloc: too many partial sums");
    int nblocks = 0;
    hipError_t e = launch_reduce_fp64(L->cfg.nx, L->cfg.ny, L->cfg.nz, L->phys, (const double*)L->mac[L->cur],
                                      partial, &nblocks, L->stream);
    if (e != hipSuccess) return drp_fail(DRP_ERR_HIP, "reduce launch failed: %s", hipGetErrorString(e));
    hipLaunchKernelGGL(reduce_final, dim3(1), dim3(256), 0, L->stream, (const double*)partial, nblocks,
                       (const uint64_t*)(L->range + 2 * (L->step % 3)), final_);
    if ((e = hipGetLastError()) != hipSuccess)
        return drp_fail(DRP_ERR_HIP, "reduce_final launch failed: %s", hipGetErrorString(e));
    double host[NPART + 2];
    if ((e = hipMemcpyAsync(host, final_, sizeof host, hipMemcpyDeviceToHost, L->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(L->stream)) != hipSuccess)
        return drp_fail(DRP_ERR_HIP, "drp_reduce_noalloc copy failed: %s", hipGetErrorString(e));
    uint64_t rg[2];
    static_assert(sizeof(double) == sizeof(uint64_t), "range words");
    __builtin_memcpy(rg, host + NPART, sizeof rg);
    stats_from_sums(L, host, rg, st);
    return DRP_OK;
}
