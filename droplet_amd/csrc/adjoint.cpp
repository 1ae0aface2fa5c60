// adjoint.cpp — checkpointed reverse sweep over a recorded FP64 run.
//
// A tape keeps the full state (f, g, phi, p, u) of its lattice at the tape's
// first step and every K steps after it.  The reverse sweep walks the K-step
// segments from last to first: each segment is recomputed forward from its
// checkpoint into a segment buffer (the step kernel is deterministic, so the
// recomputed states are bit-identical to the recorded run), then the adjoint
// kernels (adjoint.hip) run over the segment's steps in reverse order.
// Memory: (T/K + 1) checkpoints + (K - 1) segment states of 35 doubles per node,
// plus two adjoint state sets; at 96^3, T = 400, K = 20 about 10.5 GB.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/droplet.h"
#include "kernels.h"
#include "lattice_impl.h"

using namespace drp;

#define HIP_TRYA(expr)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return drp_fail(DRP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

struct drp_tape {
    drp_lattice* L = nullptr;
    int K = 1;
    int64_t start = 0;             // lattice step at which recording began
    int64_t recorded = 0;          // steps recorded
    bool start_eq = false;         // the first state is the equilibrium of constant fields
    std::vector<double*> ckpt;     // state at start + j K: f | g (30n) | phi p u (5n)
    double* seg = nullptr;         // K - 1 recomputed states
    double* adj = nullptr;         // two adjoint sets: F-part (15n) | G (15n)
    double* phin = nullptr;        // two stencil parts of the phi adjoint (n each)
    double* work = nullptr;        // lap / grad adjoints (4n)
    double* acc = nullptr;         // per-workgroup parameter sensitivities
    double* plane = nullptr;       // deformation seed on its plane
    uint64_t* range = nullptr;     // phi-range slots written by recomputation
    int nblk = 0;
    int64_t bytes = 0;
};

static size_t state_doubles(const drp_lattice* L) { return 35 * (size_t)L->n; }

static int tape_alloc(drp_tape* T, void** p, size_t bytes) {
    HIP_TRYA(hipMalloc(p, bytes));
    T->bytes += (int64_t)bytes;
    return DRP_OK;
}

static int save_checkpoint(drp_tape* T) {
    drp_lattice* L = T->L;
    const size_t n = (size_t)L->n;
    double* c = nullptr;
    int rc = tape_alloc(T, (void**)&c, state_doubles(L) * sizeof(double));
    if (rc) return rc;
    T->ckpt.push_back(c);
    HIP_TRYA(hipMemcpyAsync(c, L->pop[L->cur], 30 * n * sizeof(double), hipMemcpyDeviceToDevice, L->stream));
    HIP_TRYA(hipMemcpyAsync(c + 30 * n, L->mac[L->cur], 5 * n * sizeof(double), hipMemcpyDeviceToDevice, L->stream));
    return DRP_OK;
}

extern "C" int drp_tape_create(drp_lattice* L, int K, drp_tape** out) {
    if (!L || !out) return drp_fail(DRP_ERR_ARG, "null argument");
    *out = nullptr;
    if (K < 1) return drp_fail(DRP_ERR_ARG, "checkpoint interval %d < 1", K);
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    if (L->cfg.precision != DRP_FP64) return drp_fail(DRP_ERR_ARG, "the adjoint needs an FP64 lattice");
    if (L->phys.dnu != 0.0) return drp_fail(DRP_ERR_ARG, "the adjoint needs nu1 == nu2");
    if (L->phys.les) return drp_fail(DRP_ERR_ARG, "the adjoint does not cover the Smagorinsky model");
    if (L->cfg.model != DRP_MODEL_FREE_ENERGY || L->cfg.advect_k_high != 0)
        return drp_fail(DRP_ERR_ARG, "the adjoint covers the free-energy model without scale-separated advection only");
    if (L->cfg.force_kind == DRP_FORCE_SPECTRAL)
        return drp_fail(DRP_ERR_ARG, "the adjoint does not cover the velocity-dependent spectral forcing");
    if (L->n >= (int64_t)1 << 32) return drp_fail(DRP_ERR_ARG, "the adjoint supports grids of fewer than 2^32 nodes");
    const int wn = L->cfg.wall_axis == DRP_WALL_Y ? L->cfg.ny : L->cfg.nz;
    if (L->cfg.wall_axis != DRP_WALL_NONE && wn < 3)
        return drp_fail(DRP_ERR_ARG, "the adjoint needs at least 3 layers between the walls");
    drp_tape* T = new drp_tape();
    T->L = L;
    T->K = K;
    T->start = L->step;
    T->start_eq = L->eq_init && L->step == 0;
    int rc = save_checkpoint(T);
    if (!rc) rc = hipStreamSynchronize(L->stream) == hipSuccess ? DRP_OK : drp_fail(DRP_ERR_HIP, "checkpoint copy");
    if (rc) {
        drp_tape_destroy(T);
        return rc;
    }
    *out = T;
    return DRP_OK;
}

extern "C" int drp_tape_step(drp_tape* T, int nsteps) {
    if (!T) return drp_fail(DRP_ERR_ARG, "null tape");
    if (nsteps < 0) return drp_fail(DRP_ERR_ARG, "nsteps < 0");
    drp_lattice* L = T->L;
    if (L->step != T->start + T->recorded) return drp_fail(DRP_ERR_STATE, "lattice was stepped outside its tape");
    for (int k = 0; k < nsteps; k++) {
        int rc = drp_step(L, 1);
        if (rc) return rc;
        T->recorded++;
        if (T->recorded % T->K == 0) {
            rc = save_checkpoint(T);
            if (rc) return rc;
        }
    }
    return DRP_OK;
}

extern "C" int drp_tape_loss(drp_tape* T, const double w[2], double* loss, double* D, double* area) {
    if (!T || !w || !loss) return drp_fail(DRP_ERR_ARG, "null argument");
    drp_lattice* L = T->L;
    double Lo, Bo, Dv = 0.0;
    int rc = DRP_OK;
    if (w[0] != 0.0 || D) rc = drp_deformation(L, 2, &Lo, &Bo, &Dv);
    if (rc) return rc;
    drp_stats st;
    rc = drp_reduce(L, &st);
    if (rc) return rc;
    *loss = w[0] * Dv + w[1] * st.interface_area;
    if (D) *D = Dv;
    if (area) *area = st.interface_area;
    return DRP_OK;
}

static int ensure_work(drp_tape* T) {
    drp_lattice* L = T->L;
    const size_t n = (size_t)L->n;
    int rc = DRP_OK;
    if (!T->seg && T->K > 1) rc = tape_alloc(T, (void**)&T->seg, (size_t)(T->K - 1) * state_doubles(L) * sizeof(double));
    if (!rc && !T->adj) rc = tape_alloc(T, (void**)&T->adj, 2 * 30 * n * sizeof(double));
    if (!rc && !T->phin) rc = tape_alloc(T, (void**)&T->phin, 2 * n * sizeof(double));
    if (!rc && !T->work) rc = tape_alloc(T, (void**)&T->work, 4 * n * sizeof(double));
    T->nblk = adj_block_count((int64_t)n);
    if (!rc && !T->acc) rc = tape_alloc(T, (void**)&T->acc, (size_t)T->nblk * 3 * sizeof(double));
    if (!rc && !T->plane) rc = tape_alloc(T, (void**)&T->plane, (size_t)L->cfg.nx * std::max(L->cfg.ny, L->cfg.nz) * sizeof(double));
    if (!rc && !T->range) {
        rc = tape_alloc(T, (void**)&T->range, 6 * sizeof(uint64_t));
        for (int k = 0; k < 3 && !rc; k++)
            if (launch_range_set(T->range, k, L->cfg.phi1, L->cfg.phi2, L->stream) != hipSuccess)
                rc = drp_fail(DRP_ERR_HIP, "range initialisation failed");
    }
    return rc;
}

// state j of the current segment (0 = its checkpoint)
static double* seg_state(drp_tape* T, int j, int seg_index) {
    return j == 0 ? T->ckpt[seg_index] : T->seg + (size_t)(j - 1) * state_doubles(T->L);
}

extern "C" int drp_tape_backward(drp_tape* T, const double w[2], double* loss, double grad[DRP_GRAD_COUNT]) {
    if (!T || !w || !grad) return drp_fail(DRP_ERR_ARG, "null argument");
    drp_lattice* L = T->L;
    if (L->step != T->start + T->recorded) return drp_fail(DRP_ERR_STATE, "lattice was stepped outside its tape");
    int rc = ensure_work(T);
    if (rc) return rc;
    const drp_config& c = L->cfg;
    const int64_t n = L->n;
    hipStream_t st = L->stream;

    // ---- loss and its seed dJ/dphi at the final state
    double lossv = 0.0;
    rc = drp_tape_loss(T, w, &lossv, nullptr, nullptr);
    if (rc) return rc;
    if (loss) *loss = lossv;
    int cur = 0;
    double* Fa[2] = {T->adj, T->adj + 30 * n};
    double* Pa[2] = {T->phin, T->phin + n};
    double* lapA = T->work;
    double* gradA = T->work + n;
    HIP_TRYA(hipMemsetAsync(Fa[cur], 0, 30 * n * sizeof(double), st));
    HIP_TRYA(hipMemsetAsync(T->acc, 0, (size_t)T->nblk * 3 * sizeof(double), st));
    HIP_TRYA(launch_adj_seed_area_fp64(c.nx, c.ny, c.nz, (const double*)L->mac[L->cur], w[1], lapA, gradA, st));
    HIP_TRYA(launch_adj_stencil_fp64(c.nx, c.ny, c.nz, lapA, gradA, Pa[cur], st));
    if (w[0] != 0.0) {
        double D;
        int n1 = 0, n2 = 0, third = 0;
        std::vector<double> g((size_t)c.nx * std::max(c.ny, c.nz));
        rc = deformation_inertia_grad(L, &D, g.data(), &n1, &n2, &third);
        if (rc) return rc;
        for (int k = 0; k < n1 * n2; k++) g[k] *= w[0];
        HIP_TRYA(hipMemcpyAsync(T->plane, g.data(), (size_t)n1 * n2 * sizeof(double), hipMemcpyHostToDevice, st));
        HIP_TRYA(launch_adj_add_plane(c.nx, c.ny, n1, n2, c.wall_axis == DRP_WALL_Y, third, T->plane, Pa[cur], st));
        HIP_TRYA(hipStreamSynchronize(st));  // g is a host temporary
    }

    // ---- reverse sweep over segments
    const int64_t Ttot = T->recorded;
    const int nseg = (int)((Ttot + T->K - 1) / T->K);
    for (int sgi = nseg - 1; sgi >= 0; sgi--) {
        const int64_t s0 = (int64_t)sgi * T->K;                 // first recorded step index of the segment
        const int len = (int)std::min<int64_t>(T->K, Ttot - s0);  // reverse steps in this segment
        // recompute states 1 .. len-1 of the segment
        for (int j = 0; j + 1 < len; j++) {
            const int64_t nstep = T->start + s0 + j + 1;  // reference loop counter of this forward step
            StepArgs a = make_step_args(L, nstep);
            const double* in = seg_state(T, j, sgi);
            double* outp = seg_state(T, j + 1, sgi);
            a.f_in = in;
            a.g_in = in + 15 * n;
            a.mac_in = in + 30 * n;
            a.f_out = outp;
            a.g_out = outp + 15 * n;
            a.mac_out = outp + 30 * n;
            a.range = T->range;
            a.slot_read = 0;
            a.slot_write = 1;
            a.slot_reset = 2;
            HIP_TRYA(launch_step_fp64(a, L->shape, st));
        }
        for (int j = len - 1; j >= 0; j--) {
            const int64_t rel = s0 + j;                     // recorded step index of state n
            const int64_t nstep = T->start + rel + 1;       // forward step n -> n+1
            const double* state = seg_state(T, j, sgi);
            AdjArgs a{};
            a.nx = c.nx;
            a.ny = c.ny;
            a.nz = c.nz;
            a.n = n;
            a.f = state;
            a.mac = state + 30 * n;
            a.Fin = Fa[cur];
            a.Gin = Fa[cur] + 15 * n;
            a.phin_in = Pa[cur];
            a.Fout = Fa[1 - cur];
            a.Gout = Fa[1 - cur] + 15 * n;
            a.lapA = lapA;
            a.gradA = gradA;
            a.acc = T->acc;
            a.tau_f = L->phys.tau_f;
            a.tau_g = 0.5 + 3.0 * L->phys.nu1;
            a.kf = L->phys.kf;
            a.kf6 = L->phys.kf6;
            a.kg = L->phys.kg;
            a.T = L->phys.T;
            a.a = L->phys.a;
            a.b = L->phys.b;
            a.rtau_f = 1.0 / a.tau_f;
            a.rtau_g = 1.0 / a.tau_g;
            a.rtau_f2 = 1.0 / (a.tau_f * a.tau_f);
            a.om_f = 1.0 - 1.0 / a.tau_f;
            a.om_g = 1.0 - 1.0 / a.tau_g;
            a.wall_axis = c.wall_axis;
            a.wall_active = (c.wall_axis != DRP_WALL_NONE && nstep >= c.wall_start) ? 1 : 0;
            a.first = (rel == 0 && T->start_eq) ? 1 : 0;
            HIP_TRYA(launch_adj_collide_fp64(a, st));
            if (rel > 0) HIP_TRYA(launch_adj_stencil_fp64(c.nx, c.ny, c.nz, lapA, gradA, Pa[1 - cur], st));
            cur = 1 - cur;
        }
    }
    std::vector<double> acc((size_t)T->nblk * 3);
    HIP_TRYA(hipMemcpyAsync(acc.data(), T->acc, acc.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRYA(hipStreamSynchronize(st));
    double g3[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < T->nblk; b++)
        for (int k = 0; k < 3; k++) g3[k] += acc[(size_t)b * 3 + k];
    grad[DRP_GRAD_MOBILITY] = g3[0];
    grad[DRP_GRAD_KAPPA_G] = g3[1];
    grad[DRP_GRAD_WALL_U] = g3[2];
    return DRP_OK;
}

extern "C" int64_t drp_tape_bytes(const drp_tape* T) { return T ? T->bytes : -1; }

extern "C" void drp_tape_destroy(drp_tape* T) {
    if (!T) return;
    if (T->L) (void)hipStreamSynchronize(T->L->stream);
    for (double* p : T->ckpt) (void)hipFree(p);
    void* bufs[] = {T->seg, T->adj, T->phin, T->work, T->acc, T->plane, T->range};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    delete T;
}
