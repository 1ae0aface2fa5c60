// census.hip — droplet census on the GPU: connected clusters of nodes with
// phi >= threshold, their count and volumes.  The reference does this offline
// on gathered field files with a sequential Hoshen-Kopelman scan
// (program/percolation.f90:103-258, volumes and count :288-315).  Here the
// clusters are found in parallel with a lock-free union-find over the face
// neighbours, periodic in all three directions:
//   cc_init     label[i] = i for a droplet node, -1 otherwise
//   cc_union    every droplet node links to its +x, +y, +z droplet neighbours
//               (union by atomicMin: a root only ever gains a smaller parent,
//               so every root is the smallest node index of its cluster and
//               parent chains strictly decrease)
//   cc_flatten  label[i] = root, volume[root] += 1 (one atomic per root and wavefront)
//   cc_collect  one (root, volume) record per cluster, in any order; the host
//               sorts them by root
// Work buffer (int32, 4 n + 4): labels | volumes | records (2 per cluster) |
// counter.  Clusters are pairwise non-adjacent, so one node of each forms an
// independent set of the periodic grid graph, which has at most (n + 1) / 2
// nodes (every side is >= 3, the graph has a matching missing at most one node).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/droplet.h"
#include "lattice_impl.h"

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return drp_fail(DRP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace {

__device__ __forceinline__ int find_root(const int* __restrict__ lab, int i) {
    int p = __atomic_load_n(&lab[i], __ATOMIC_RELAXED);
    while (p != i) {
        i = p;
        p = __atomic_load_n(&lab[i], __ATOMIC_RELAXED);
    }
    return i;
}

// Join the clusters of droplet nodes a and b (Playne & Hawick 2018, IEEE TPDS
// 29(6), union of their "label equivalence" method): repeatedly hang the larger
// root under the smaller one with atomicMin; if another thread changed that
// root's parent meanwhile, continue from the parent it set.  Labels only
// decrease, so the loop ends.
__device__ void unite(int* lab, int a, int b) {
    for (;;) {
        a = find_root(lab, a);
        b = find_root(lab, b);
        if (a == b) return;
        if (a < b) {
            const int old = atomicMin(&lab[b], a);
            if (old == b) return;
            b = old;
        } else {
            const int old = atomicMin(&lab[a], b);
            if (old == a) return;
            a = old;
        }
    }
}

__global__ __launch_bounds__(256) void cc_init(const double* __restrict__ phi, double thr, int n,
                                               int* __restrict__ lab, int* __restrict__ vol) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    lab[i] = phi[i] >= thr ? i : -1;
    vol[i] = 0;
}

__global__ __launch_bounds__(256) void cc_union(int nx, int ny, int nz, int* lab) {
    const int n = nx * ny * nz;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || lab[i] < 0) return;
    const int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
    const int jx = (x + 1 == nx) ? i - x : i + 1;
    const int jy = (y + 1 == ny) ? i - y * nx : i + nx;
    const int jz = (z + 1 == nz) ? i - z * nx * ny : i + nx * ny;
    if (lab[jx] >= 0) unite(lab, i, jx);
    if (lab[jy] >= 0) unite(lab, i, jy);
    if (lab[jz] >= 0) unite(lab, i, jz);
}

__global__ __launch_bounds__(256) void cc_flatten(int n, int* lab, int* __restrict__ vol) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int r = -1;
    if (i < n && lab[i] >= 0) {
        r = find_root(lab, i);
        // only the own entry is written; other threads read it as a valid
        // parent (r is an ancestor of i), so their finds stay correct
        __atomic_store_n(&lab[i], r, __ATOMIC_RELAXED);
    }
    // volume counts: one atomic per distinct root of the wavefront (neighbouring
    // nodes mostly share a root), not one per node
    const int lane = threadIdx.x & 63;
    bool pending = r >= 0;
    for (;;) {
        const unsigned long long m = __ballot(pending);
        if (m == 0) break;
        const int leader = __ffsll((long long)m) - 1;
        const int lr = __shfl(r, leader, 64);
        const bool same = pending && r == lr;
        const unsigned long long sm = __ballot(same);
        if (lane == leader) atomicAdd(&vol[lr], __popcll(sm));
        pending = pending && !same;
    }
}

__global__ __launch_bounds__(256) void cc_collect(int n, const int* __restrict__ lab, const int* __restrict__ vol,
                                                  int* __restrict__ rec, int* __restrict__ counter) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || lab[i] != i) return;
    const int k = atomicAdd(counter, 1);
    rec[2 * k] = i;          // k < (n + 1) / 2 <= n / 2 + 1

    rec[2 * k + 1] = vol[i];
}

__global__ void cc_zero(int* p) { *p = 0; }

inline dim3 blocks(int n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

extern "C" int drp_droplet_census(drp_lattice* L, double threshold, int64_t* count, int64_t* volumes, int64_t* first,
                                  int64_t cap, int32_t* labels, size_t n) {
    if (!L || !count) return drp_fail(DRP_ERR_ARG, "null argument");
    if (cap < 0 || (cap > 0 && (!volumes || !first))) return drp_fail(DRP_ERR_ARG, "bad output arrays");
    if (labels && n != (size_t)L->n) return drp_fail(DRP_ERR_ARG, "labels size %zu != %lld nodes", n, (long long)L->n);
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    if (L->n >= ((int64_t)1 << 30)) return drp_fail(DRP_ERR_ARG, "census needs fewer than 2^30 nodes");
    const int N = (int)L->n;
    const double* phi = nullptr;
    if (L->cfg.precision == DRP_FP64) {
        phi = (const double*)L->mac[L->cur];
    } else {
        const double* m = nullptr;
        int rc = fp32_macro_fp64_view(L, &m);
        if (rc) return rc;
        phi = m;
    }
    if (!L->census) HIP_TRY(hipMalloc((void**)&L->census, (4 * (size_t)N + 4) * sizeof(int)));
    int* lab = L->census;
    int* vol = lab + N;
    int* rec = vol + N;          // 2 (N / 2 + 1) ints
    int* counter = rec + 2 * (N / 2 + 1);
    const hipStream_t st = L->stream;
    hipLaunchKernelGGL(cc_init, blocks(N), dim3(256), 0, st, phi, threshold, N, lab, vol);
    hipLaunchKernelGGL(cc_union, blocks(N), dim3(256), 0, st, L->cfg.nx, L->cfg.ny, L->cfg.nz, lab);
    hipLaunchKernelGGL(cc_flatten, blocks(N), dim3(256), 0, st, N, lab, vol);
    hipLaunchKernelGGL(cc_zero, dim3(1), dim3(1), 0, st, counter);
    hipLaunchKernelGGL(cc_collect, blocks(N), dim3(256), 0, st, N, lab, vol, rec, counter);
    HIP_TRY(hipGetLastError());
    int nc = 0;
    HIP_TRY(hipMemcpyAsync(&nc, counter, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<int> r(2 * (size_t)nc);
    if (nc > 0) HIP_TRY(hipMemcpyAsync(r.data(), rec, r.size() * sizeof(int), hipMemcpyDeviceToHost, st));
    if (labels) HIP_TRY(hipMemcpyAsync(labels, lab, (size_t)N * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<std::pair<int, int>> c((size_t)nc);
    for (int k = 0; k < nc; k++) c[k] = {r[2 * k], r[2 * k + 1]};
    std::sort(c.begin(), c.end());
    *count = nc;
    for (int64_t k = 0; k < std::min<int64_t>(cap, nc); k++) {
        first[k] = c[k].first;
        volumes[k] = c[k].second;
    }
    return DRP_OK;
}

// Droplet size distribution of the current state (program/mkdsd.f90 over the
// radii program/percolation.f90 derives from cluster volumes, :298):
// r = (3 V / (4 * 3.14))^(1/3) (the reference's 3.14), d = 2 r / scale; droplets
// with d > d_max are dropped (mkdsd.f90 :68-72); counts[i] = number with
// int(d / bin) == i for i < nbins (d > 0).  *total receives the number counted.
extern "C" int drp_droplet_size_histogram(drp_lattice* L, double threshold, double scale, double bin, double d_max,
                                          int64_t* counts, int64_t nbins, int64_t* total) {
    if (!L || !counts || !total || nbins < 1 || !(scale > 0.0) || !(bin > 0.0))
        return drp_fail(DRP_ERR_ARG, "bad histogram arguments");
    int64_t nc = 0, cap = 1 << 16;
    std::vector<int64_t> vol((size_t)cap), first((size_t)cap);
    int rc = drp_droplet_census(L, threshold, &nc, vol.data(), first.data(), cap, nullptr, 0);
    if (!rc && nc > cap) {  // more droplets than the first guess: once more with room for all
        cap = nc;
        vol.resize((size_t)cap);
        first.resize((size_t)cap);
        rc = drp_droplet_census(L, threshold, &nc, vol.data(), first.data(), cap, nullptr, 0);
    }
    if (rc) return rc;
    for (int64_t i = 0; i < nbins; i++) counts[i] = 0;
    int64_t t = 0;
    for (int64_t k = 0; k < nc; k++) {
        const double r = std::pow(3.0 * (double)vol[k] / (4.0 * 3.14), 1.0 / 3.0);
        const double d = 2.0 * r / scale;
        if (d > d_max || !(d > 0.0)) continue;
        const int64_t i = (int64_t)(d / bin);
        if (i < nbins) {
            counts[i]++;
            t++;
        }
    }
    *total = t;
    return DRP_OK;
}
