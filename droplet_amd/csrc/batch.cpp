// batch.cpp — many independent lattices advanced by one kernel launch per step.
//
// The reference runs ensembles as MPI sub-communicators, one group per
// parameter set (prototype_program/two-phase/multidrop2.f90:235-260,
// program/LBM.f90:215-232), or as a sequential case loop
// (program/lbm_drop_iht_xy_many.f90).  Here every member is an ordinary
// drp_lattice (own buffers, own configuration, any grid size) living on the
// batch's stream; one step of all members is a single launch of the batched
// step kernel, whose workgroups find their member by binary search over the
// members' first-tile offsets.
//
// The per-step arguments of a member depend on the step only through the
// ping-pong parity (period 2), the phi-range slot (period 3) and whether the
// walls / body force are switched on yet.  Once every member is past its
// switch-on steps, six consecutive steps repeat exactly, so those six launches
// are captured once into a hipGraph and replayed.  The graph holds kernel nodes
// only (no memset or memcpy nodes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/droplet.h"
#include "kernels.h"
#include "lattice_impl.h"

using namespace drp;

#define HIP_TRYB(expr)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return drp_fail(DRP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

struct drp_batch {
    std::vector<drp_lattice*> members;
    hipStream_t stream = nullptr;
    int precision = DRP_FP64;
    int shape = SHAPE_32x4x2;
    size_t arg_size = 0;
    int total_tiles = 0;
    int* d_tile_start = nullptr;
    void* d_single = nullptr;     // arguments of one step (direct launches)
    void* d_cycle = nullptr;      // arguments of six consecutive steps (graph)
    int64_t cycle_first = -1;     // first step of the captured cycle
    int cycle_regime = -1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool use_graph = true;
};

static int64_t common_step(const drp_batch* B) { return B->members.empty() ? 0 : B->members[0]->step; }

// number of switch-on steps (walls, body force) already reached at step n, summed over members
static int regime(const drp_batch* B, int64_t n) {
    int r = 0;
    for (const drp_lattice* L : B->members) {
        if (L->cfg.wall_axis != DRP_WALL_NONE && n >= L->cfg.wall_start) r++;
        if (L->cfg.force_kind != DRP_FORCE_NONE && n >= L->cfg.force_start) r++;
    }
    return r;
}

// host image of the members' arguments for step n (members' buffers as they will be then)
static void build_args(drp_batch* B, int64_t n, std::vector<char>& out) {
    const int64_t now = common_step(B);
    const int* dims = SHAPE_DIMS[B->shape];
    out.resize(B->arg_size * B->members.size());
    for (size_t i = 0; i < B->members.size(); i++) {
        drp_lattice* L = B->members[i];
        const int cur = L->cur;
        L->cur = (int)((cur + (n - 1 - now)) & 1);  // buffer parity at the start of step n
        if (B->precision == DRP_FP64) {
            StepArgs a = make_step_args(L, n);
            set_tiles(a, dims[0], dims[1], dims[2]);
            std::memcpy(out.data() + i * B->arg_size, &a, sizeof a);
        } else {
            StepArgs32 a = make_step_args32(L, n);
            set_tiles(a, dims[0], dims[1], dims[2]);
            std::memcpy(out.data() + i * B->arg_size, &a, sizeof a);
        }
        L->cur = cur;
    }
}

static int set_layout(drp_batch* B) {
    const int* dims = SHAPE_DIMS[B->shape];
    std::vector<int> start(B->members.size());
    int64_t total = 0;
    for (size_t i = 0; i < B->members.size(); i++) {
        const drp_config& c = B->members[i]->cfg;
        start[i] = (int)total;
        total += (int64_t)((c.nx + dims[0] - 1) / dims[0]) * ((c.ny + dims[1] - 1) / dims[1]) *
                 ((c.nz + dims[2] - 1) / dims[2]);
    }
    if (total >= (int64_t)1 << 31) return drp_fail(DRP_ERR_ARG, "batch has too many tiles");
    B->total_tiles = (int)total;
    HIP_TRYB(hipMemcpyAsync(B->d_tile_start, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice,
                            B->stream));
    HIP_TRYB(hipStreamSynchronize(B->stream));
    B->cycle_first = -1;  // arguments and graph depend on the tile shape
    return DRP_OK;
}

static void drop_graph(drp_batch* B) {
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    B->exec = nullptr;
    B->graph = nullptr;
}

static hipError_t launch(drp_batch* B, const void* d_args) {
    BatchLaunch bl{d_args, B->d_tile_start, (int)B->members.size(), B->total_tiles};
    // the representative arguments select the kernel variant; all members agree on it
    const drp_lattice* L0 = B->members[0];
    if (B->precision == DRP_FP64) {
        StepArgs a = make_step_args(const_cast<drp_lattice*>(L0), L0->step + 1);
        return launch_step_fp64(a, B->shape, B->stream, &bl);
    }
    StepArgs32 a = make_step_args32(const_cast<drp_lattice*>(L0), L0->step + 1);
    return launch_step_fp32(a, B->shape, B->stream, &bl);
}

static void advance(drp_batch* B, int k) {
    for (drp_lattice* L : B->members) {
        L->step += k;
        L->cur = (L->cur + k) & 1;
    }
}

static int single_step(drp_batch* B) {
    std::vector<char> host;
    build_args(B, common_step(B) + 1, host);
    HIP_TRYB(hipMemcpyAsync(B->d_single, host.data(), host.size(), hipMemcpyHostToDevice, B->stream));
    HIP_TRYB(launch(B, B->d_single));
    advance(B, 1);
    return DRP_OK;
}

// capture steps n .. n+5 into a graph (arguments of each in its own slot)
static int build_cycle(drp_batch* B, int64_t n) {
    drop_graph(B);
    std::vector<char> host, all;
    for (int k = 0; k < 6; k++) {
        build_args(B, n + k, host);
        all.insert(all.end(), host.begin(), host.end());
    }
    HIP_TRYB(hipMemcpyAsync(B->d_cycle, all.data(), all.size(), hipMemcpyHostToDevice, B->stream));
    HIP_TRYB(hipStreamSynchronize(B->stream));
    const size_t slot = B->arg_size * B->members.size();
    std::vector<int64_t> saved_step;
    std::vector<int> saved_cur;
    for (drp_lattice* L : B->members) {
        saved_step.push_back(L->step);
        saved_cur.push_back(L->cur);
    }
    HIP_TRYB(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int k = 0; k < 6 && e == hipSuccess; k++) {
        e = launch(B, (const char*)B->d_cycle + k * slot);
        advance(B, 1);
    }
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(B->stream, &g);
    for (size_t i = 0; i < B->members.size(); i++) {
        B->members[i]->step = saved_step[i];
        B->members[i]->cur = saved_cur[i];
    }
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return drp_fail(DRP_ERR_HIP, "capture launch failed: %s", hipGetErrorString(e));
    }
    if (e2 != hipSuccess) return drp_fail(DRP_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e2));
    B->graph = g;
    HIP_TRYB(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
    B->cycle_first = n;
    B->cycle_regime = regime(B, n);
    return DRP_OK;
}

extern "C" int drp_batch_create(const drp_config* cfgs, int count, drp_batch** out) {
    if (!out || !cfgs || count < 1) return drp_fail(DRP_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    for (int i = 1; i < count; i++) {
        const drp_config &a = cfgs[0], &b = cfgs[i];
        const bool les_a = a.les_cs != 0.0 && a.les_delta != 0.0, les_b = b.les_cs != 0.0 && b.les_delta != 0.0;
        if (a.precision != b.precision || (a.wall_axis != 0) != (b.wall_axis != 0) || a.force_kind != b.force_kind ||
            les_a != les_b || (a.nu1 == a.nu2) != (b.nu1 == b.nu2) ||
            (a.force_kind != DRP_FORCE_NONE && a.force_start != b.force_start))
            return drp_fail(DRP_ERR_UNSUPPORTED,
                            "batch members must share precision, wall/force/LES switches and viscosity-ratio kind");
    }
    if (cfgs[0].les_cs != 0.0 && cfgs[0].les_delta != 0.0)
        return drp_fail(DRP_ERR_UNSUPPORTED, "LES members are not supported in a batch");
    for (int i = 0; i < count; i++)
        if (cfgs[i].model != DRP_MODEL_FREE_ENERGY || cfgs[i].advect_k_high != 0)
            return drp_fail(DRP_ERR_UNSUPPORTED, "batch members must use the free-energy model without "
                                                 "scale-separated advection");
    if (cfgs[0].force_kind == DRP_FORCE_SPECTRAL)
        return drp_fail(DRP_ERR_UNSUPPORTED, "spectral forcing is not supported in a batch");
    drp_batch* B = new (std::nothrow) drp_batch();
    if (!B) return drp_fail(DRP_ERR_ALLOC, "host allocation failed");
    B->precision = cfgs[0].precision;
    B->shape = default_shape(B->precision);
    B->arg_size = B->precision == DRP_FP64 ? sizeof(StepArgs) : sizeof(StepArgs32);
    int rc = DRP_OK;
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess)
        rc = drp_fail(DRP_ERR_HIP, "stream creation failed");
    for (int i = 0; i < count && !rc; i++) {
        drp_lattice* L = nullptr;
        rc = lattice_create(&cfgs[i], B->stream, &L);
        if (!rc) B->members.push_back(L);
    }
    if (!rc && (hipMalloc((void**)&B->d_tile_start, count * sizeof(int)) != hipSuccess ||
                hipMalloc(&B->d_single, count * B->arg_size) != hipSuccess ||
                hipMalloc(&B->d_cycle, 6 * count * B->arg_size) != hipSuccess))
        rc = drp_fail(DRP_ERR_ALLOC, "batch argument buffers");
    if (!rc) rc = set_layout(B);
    if (rc) {
        drp_batch_destroy(B);
        return rc;
    }
    *out = B;
    return DRP_OK;
}

extern "C" void drp_batch_destroy(drp_batch* B) {
    if (!B) return;
    if (B->stream) (void)hipStreamSynchronize(B->stream);
    drop_graph(B);
    for (drp_lattice* L : B->members) drp_destroy(L);
    if (B->d_tile_start) (void)hipFree(B->d_tile_start);
    if (B->d_single) (void)hipFree(B->d_single);
    if (B->d_cycle) (void)hipFree(B->d_cycle);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
}

extern "C" int drp_batch_size(const drp_batch* B) { return B ? (int)B->members.size() : -1; }

extern "C" drp_lattice* drp_batch_member(drp_batch* B, int i) {
    if (!B || i < 0 || i >= (int)B->members.size()) {
        drp_fail(DRP_ERR_ARG, "batch member %d out of range", i);
        return nullptr;
    }
    B->cycle_first = -1;  // the caller may re-initialise members
    return B->members[i];
}

extern "C" int drp_batch_set_tile_shape(drp_batch* B, int shape) {
    if (!B) return drp_fail(DRP_ERR_ARG, "null batch");
    if (shape < 0 || shape >= SHAPE_COUNT) return drp_fail(DRP_ERR_ARG, "tile shape %d out of range", shape);
    B->shape = shape;
    return set_layout(B);
}

extern "C" int drp_batch_use_graph(drp_batch* B, int enable) {
    if (!B) return drp_fail(DRP_ERR_ARG, "null batch");
    B->use_graph = enable != 0;
    return DRP_OK;
}

extern "C" int drp_batch_step(drp_batch* B, int nsteps) {
    if (!B) return drp_fail(DRP_ERR_ARG, "null batch");
    if (nsteps < 0) return drp_fail(DRP_ERR_ARG, "nsteps < 0");
    const int64_t s0 = common_step(B);
    for (drp_lattice* L : B->members) {
        if (!L->initialized) return drp_fail(DRP_ERR_STATE, "batch member not initialised");
        if (L->step != s0 || L->cur != B->members[0]->cur)
            return drp_fail(DRP_ERR_STATE, "batch members must be at the same step");
    }
    int k = 0;
    while (k < nsteps) {
        const int64_t n = common_step(B) + 1;
        const bool stable = regime(B, n) == regime(B, n + 6);
        if (B->use_graph && nsteps - k >= 6 && stable) {
            const bool aligned = B->cycle_first > 0 && (n - B->cycle_first) % 6 == 0 &&
                                 B->cycle_regime == regime(B, n);
            if (!aligned) {
                int rc = build_cycle(B, n);
                if (rc) return rc;
            }
            HIP_TRYB(hipGraphLaunch(B->exec, B->stream));
            advance(B, 6);
            k += 6;
        } else {
            int rc = single_step(B);
            if (rc) return rc;
            k += 1;
        }
    }
    return DRP_OK;
}

// Time every tile shape on a temporary batch of the same members (each with a
// centred droplet of diameter 0.4 nx), nsteps steps per shape after 2 warm-up
// steps, and return the fastest shape.
extern "C" int drp_batch_autotune(const drp_config* cfgs, int count, int nsteps, int* best, double* ms_out) {
    if (!cfgs || count < 1 || !best || nsteps < 1) return drp_fail(DRP_ERR_ARG, "bad batch autotune arguments");
    drp_batch* B = nullptr;
    int rc = drp_batch_create(cfgs, count, &B);
    if (rc) return rc;
    for (int i = 0; i < count && !rc; i++) {
        const drp_config& c = cfgs[i];
        const drp_drop d{0.5 * (c.nx - 1), 0.5 * (c.ny - 1), 0.5 * (c.nz - 1), 0.4 * c.nx};
        rc = drp_init_droplets(B->members[i], &d, 1, 1.0 / 3.0);
    }
    double bestms = 1e300;
    for (int s = 0; s < SHAPE_COUNT && !rc; s++) {
        rc = drp_batch_set_tile_shape(B, s);
        double ms = 0.0;
        if (!rc) rc = drp_batch_time_steps(B, 2, &ms);
        if (!rc) rc = drp_batch_time_steps(B, nsteps, &ms);
        if (rc) break;
        ms /= nsteps;
        if (ms_out) ms_out[s] = ms;
        if (ms < bestms) {
            bestms = ms;
            *best = s;
        }
    }
    drp_batch_destroy(B);
    return rc;
}

extern "C" int drp_batch_sync(drp_batch* B) {
    if (!B) return drp_fail(DRP_ERR_ARG, "null batch");
    HIP_TRYB(hipStreamSynchronize(B->stream));
    return DRP_OK;
}

extern "C" int drp_batch_time_steps(drp_batch* B, int nsteps, double* ms) {
    if (!B || !ms) return drp_fail(DRP_ERR_ARG, "null argument");
    hipEvent_t e0, e1;
    HIP_TRYB(hipEventCreate(&e0));
    HIP_TRYB(hipEventCreate(&e1));
    HIP_TRYB(hipEventRecord(e0, B->stream));
    int rc = drp_batch_step(B, nsteps);
    HIP_TRYB(hipEventRecord(e1, B->stream));
    HIP_TRYB(hipEventSynchronize(e1));
    float t = 0.0f;
    HIP_TRYB(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t;
    return rc;
}
