// fp32_shifted.cpp — host side of the FP32 shifted-storage mode
// (DRP_FP32_SHIFTED): initialisation through the FP64 equilibria, per-step
// launches of step_fp32, and conversions back to FP64 for the output and
// analysis hooks.
#include <hip/hip_runtime.h>

#include <cmath>

#include "d3q15.h"
#include "lattice_impl.h"

using namespace drp;

#define HIP_TRY32(expr)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return drp_fail(DRP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

static const double kF[15] = {-7.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0,
                              1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0,
                              1.0 / 24.0};

// reference phi populations of the continuous phase at rest: H phi1 + F p0(phi1)
static ShiftRef shift_ref(const drp_config& c) {
    ShiftRef r;
    const double p0 = c.phi1 * c.T / (1.0 - c.b * c.phi1) - c.a * c.phi1 * c.phi1;
    for (int q = 0; q < 15; q++) r.fr[q] = (q == 0 ? c.phi1 : 0.0) + kF[q] * p0;
    r.phi1 = c.phi1;
    return r;
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

int fp32_init_from_host_fields(drp_lattice* L, const double* const* host) {
    const int64_t n = L->n;
    DevBuf mac64, pop64;
    HIP_TRY32(hipMalloc(&mac64.p, 5 * n * sizeof(double)));
    HIP_TRY32(hipMalloc(&pop64.p, 30 * n * sizeof(double)));
    double* m = (double*)mac64.p;
    double* f = (double*)pop64.p;
    for (int k = 0; k < 5; k++)
        HIP_TRY32(hipMemcpyAsync(m + k * n, host[k], n * sizeof(double), hipMemcpyHostToDevice, L->stream));
    HIP_TRY32(launch_init_equilibrium_fp64(L->cfg.nx, L->cfg.ny, L->cfg.nz, L->phys, m, f, f + 15 * n, L->stream));
    HIP_TRY32(launch_to_shifted(n, shift_ref(L->cfg), f, m, (float*)L->pop[L->cur], (float*)L->mac[L->cur],
                                L->stream));
    int rc = reset_range(L);
    if (rc) return rc;
    HIP_TRY32(hipStreamSynchronize(L->stream));
    L->step = 0;
    L->initialized = true;
    return DRP_OK;
}

StepArgs32 make_step_args32(drp_lattice* L, int64_t nstep) {
    const drp_config& c = L->cfg;
    const Phys& P = L->phys;
    const int64_t n = L->n;
    StepArgs32 a{};
    a.nx = c.nx;
    a.ny = c.ny;
    a.nz = c.nz;
    a.n = n;
    float* fi = (float*)L->pop[L->cur];
    float* fo = (float*)L->pop[1 - L->cur];
    a.f_in = fi;
    a.g_in = fi + 15 * n;
    a.f_out = fo;
    a.g_out = fo + 15 * n;
    a.mac_in = (const float*)L->mac[L->cur];
    a.mac_out = (float*)L->mac[1 - L->cur];
    a.tau_field = P.les ? (const float*)L->tau : nullptr;
    a.tg_amp = P.tg_amp;
    a.phi1 = c.phi1;
    a.phi2 = c.phi2;
    a.phi1f = (float)c.phi1;
    a.Tf = (float)c.T;
    a.af = (float)c.a;
    a.bf = (float)c.b;
    a.one_m_bphi1 = (float)(1.0 - c.b * c.phi1);
    a.kff = (float)c.kappa_f;
    a.kf6f = (float)(c.kappa_f / 6.0);
    a.kgf = (float)c.kappa_g;
    a.rtau_f = (float)(1.0 / c.tau_f);
    a.rtau_const = (float)(1.0 / (0.5 + 3.0 * c.nu1));
    a.nu1f = (float)c.nu1;
    a.dnuf = (float)(c.nu2 - c.nu1);
    a.wall_axis = c.wall_axis;
    a.wall_u_now = (float)((nstep >= c.wall_start) ? c.wall_u : 0.0);
    a.force_kind = FORCE_NONE;
    if (c.force_kind == DRP_FORCE_TAYLOR_GREEN && nstep >= c.force_start) {
        a.force_kind = FORCE_TAYLOR_GREEN;
        a.sinx = L->tg;
        a.cosx = L->tg + c.nx;
        a.sinz = L->tg + 2 * c.nx;
        a.cosz = L->tg + 2 * c.nx + c.nz;
    } else if ((c.force_kind == DRP_FORCE_FIELD || c.force_kind == DRP_FORCE_SPECTRAL) && nstep >= c.force_start) {
        a.force_kind = FORCE_FIELD;
        a.fx = L->force;
        a.fy = L->force + n;
        a.fz = L->force + 2 * n;
    }
    a.range = L->range;
    a.slot_read = (int)((nstep - 1) % 3);
    a.slot_write = (int)(nstep % 3);
    a.slot_reset = (int)((nstep + 1) % 3);
    return a;
}

int fp32_step(drp_lattice* L, int nsteps) {
    for (int k = 0; k < nsteps; k++) {
        const int64_t nstep = L->step + 1;
        int rc = spectral_update(L, nstep);
        if (rc) return rc;
        StepArgs32 a = make_step_args32(L, nstep);
        HIP_TRY32(launch_step_fp32(a, L->shape, L->stream));
        L->cur = 1 - L->cur;
        L->step = nstep;
        rc = les_update_tau(L, false);
        if (rc) return rc;
    }
    return DRP_OK;
}

int fp32_macro_fp64_view(drp_lattice* L, const double** mac) {
    if (!L->mac64) HIP_TRY32(hipMalloc((void**)&L->mac64, 5 * L->n * sizeof(double)));
    HIP_TRY32(launch_macro_from_shifted(L->n, L->cfg.phi1, (const float*)L->mac[L->cur], L->mac64, L->stream));
    *mac = L->mac64;
    return DRP_OK;
}

int fp32_get_macro(drp_lattice* L, int field, double* host) {
    const double* m = nullptr;
    int rc = fp32_macro_fp64_view(L, &m);
    if (rc) return rc;
    HIP_TRY32(hipMemcpyAsync(host, m + field * L->n, L->n * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIP_TRY32(hipStreamSynchronize(L->stream));
    return DRP_OK;
}

int fp32_get_populations(drp_lattice* L, int set, double* host) {
    DevBuf tmp;
    HIP_TRY32(hipMalloc(&tmp.p, 15 * L->n * sizeof(double)));
    HIP_TRY32(launch_pops_from_shifted(L->n, shift_ref(L->cfg), (const float*)L->pop[L->cur], (double*)tmp.p, set,
                                       L->stream));
    HIP_TRY32(hipMemcpyAsync(host, tmp.p, 15 * L->n * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIP_TRY32(hipStreamSynchronize(L->stream));
    return DRP_OK;
}

int fp32_set_populations(drp_lattice* L, int set, const double* host, int which) {
    DevBuf tmp;
    HIP_TRY32(hipMalloc(&tmp.p, 15 * L->n * sizeof(double)));
    HIP_TRY32(hipMemcpyAsync(tmp.p, host, 15 * L->n * sizeof(double), hipMemcpyHostToDevice, L->stream));
    float* pop = (float*)L->pop[L->cur];
    HIP_TRY32(launch_pops_to_shifted(L->n, shift_ref(L->cfg), (const double*)tmp.p, pop, set, L->stream));
    HIP_TRY32(launch_moments_fp32(L->n, pop, pop + 15 * L->n, (float*)L->mac[L->cur], which, L->stream));
    HIP_TRY32(hipStreamSynchronize(L->stream));
    return DRP_OK;
}
