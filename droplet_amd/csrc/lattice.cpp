// lattice.cpp — C ABI of include/droplet.h over the HIP kernels.
//
// Owns the device state of one lattice (populations ping-pong, macroscopic
// fields ping-pong, the phi-range slots, optional LES tau / force fields) and
// replays the reference loop body (program/lbm_drop_xy.f90:787-848) one kernel
// launch per time step on the handle's stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/droplet.h"
#include "d3q15.h"
#include "kernels.h"
#include "lattice_impl.h"

using namespace drp;

static thread_local std::string g_last_error;

int drp_fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return drp_fail(DRP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

extern "C" const char* drp_last_error(void) { return g_last_error.c_str(); }

extern "C" int drp_abi_sizes(size_t* config, size_t* drop, size_t* stats) {
    if (config) *config = sizeof(drp_config);
    if (drop) *drop = sizeof(drp_drop);
    if (stats) *stats = sizeof(drp_stats);
    return DROPLET_ABI_VERSION;
}

extern "C" int drp_config_default(drp_config* c) {
    if (!c) return drp_fail(DRP_ERR_ARG, "null config");
    std::memset(c, 0, sizeof(*c));
    c->nx = c->ny = c->nz = 64;
    c->precision = DRP_FP64;
    c->tau_f = 0.7;
    c->nu1 = c->nu2 = 0.1;
    c->kappa_f = 0.01;
    c->kappa_g = 0.0;
    c->phi1 = 2.211;
    c->phi2 = 4.895;
    c->a = 9.0 / 49.0;
    c->b = 2.0 / 21.0;
    c->T = 0.55;
    c->wall_start = 1;
    c->force_start = 1;
    return DRP_OK;
}

Phys make_phys(const drp_config& c) {
    Phys p;
    p.tau_f = c.tau_f;
    p.kf = c.kappa_f;
    p.kf6 = c.kappa_f / 6.0;
    p.kg = c.kappa_g;
    p.T = c.T;
    p.a = c.a;
    p.b = c.b;
    p.nu1 = c.nu1;
    p.dnu = c.nu2 - c.nu1;
    p.phi1 = c.phi1;
    p.phi2 = c.phi2;
    p.tg_amp = c.force_amp;
    p.les_c = c.les_cs * c.les_cs * c.les_delta * c.les_delta;
    p.les = (c.les_cs != 0.0 && c.les_delta != 0.0) ? 1 : 0;
    return p;
}

static int validate(const drp_config* c) {
    if (!c) return drp_fail(DRP_ERR_ARG, "null config");
    if (c->nx < 3 || c->ny < 3 || c->nz < 3) return drp_fail(DRP_ERR_ARG, "grid must be at least 3^3");
    if ((int64_t)c->nx * c->ny * c->nz * 15 > (int64_t)1 << 40) return drp_fail(DRP_ERR_ARG, "grid too large");
    if ((int64_t)c->nx * c->ny >= (int64_t)1 << 31)
        return drp_fail(DRP_ERR_ARG, "an x-y plane of 2^31 or more nodes is not supported");
    if (c->precision != DRP_FP64 && c->precision != DRP_FP32_SHIFTED)
        return drp_fail(DRP_ERR_ARG, "unknown precision %d", c->precision);
    if (!(c->tau_f > 0.5)) return drp_fail(DRP_ERR_ARG, "tau_f must exceed 1/2");
    if (!(c->nu1 > 0.0) || !(c->nu2 > 0.0)) return drp_fail(DRP_ERR_ARG, "viscosities must be positive");
    if (c->wall_axis != DRP_WALL_NONE && c->wall_axis != DRP_WALL_Y && c->wall_axis != DRP_WALL_Z)
        return drp_fail(DRP_ERR_ARG, "wall_axis must be 0, 2 or 3");
    if (c->force_kind < DRP_FORCE_NONE || c->force_kind > DRP_FORCE_SPECTRAL)
        return drp_fail(DRP_ERR_ARG, "unknown force_kind %d", c->force_kind);
    if (c->force_kind == DRP_FORCE_SPECTRAL) {
        const double kc = c->force_k;
        if (!(kc == 2.0 || kc == 3.0 || kc == 4.0 || kc == 5.0))
            return drp_fail(DRP_ERR_UNSUPPORTED, "spectral forcing cutoff force_k must be 2, 3, 4 or 5");
        if (!(c->force_amp >= 0.0)) return drp_fail(DRP_ERR_ARG, "spectral forcing needs force_amp = epsilon >= 0");
        if (2 * ((int)kc - 1) >= c->nx || 2 * ((int)kc - 1) >= c->ny || 2 * ((int)kc - 1) >= c->nz)
            return drp_fail(DRP_ERR_ARG, "grid too small for the forced shells");
    }
    if (c->anu != 0.0)
        return drp_fail(DRP_ERR_UNSUPPORTED, "anu != 0 (velocity-gradient term of geq) is not implemented");
    if (c->model != DRP_MODEL_FREE_ENERGY && c->model != DRP_MODEL_ALLEN_CAHN)
        return drp_fail(DRP_ERR_ARG, "unknown model %d", c->model);
    if (c->les_model != DRP_LES_STRAIN && c->les_model != DRP_LES_NONEQ)
        return drp_fail(DRP_ERR_ARG, "unknown les_model %d", c->les_model);
    if (c->les_model == DRP_LES_NONEQ && c->les_cs != 0.0 && c->les_delta != 0.0 &&
        (c->precision != DRP_FP64 || c->les_ranks_x != 0 || c->model != DRP_MODEL_FREE_ENERGY))
        return drp_fail(DRP_ERR_UNSUPPORTED, "the non-equilibrium Smagorinsky model runs in FP64, free-energy "
                                             "model, without rank blocks");
    if (c->model == DRP_MODEL_ALLEN_CAHN) {
        if (!(c->ac_width > 0.0)) return drp_fail(DRP_ERR_ARG, "Allen-Cahn interface width ac_width must be positive");
        if (!(c->ac_a >= 0.0 && c->ac_a < 1.0)) return drp_fail(DRP_ERR_ARG, "Allen-Cahn ac_a must lie in [0, 1)");
        if (c->precision != DRP_FP64 || c->force_kind != DRP_FORCE_NONE || (c->les_cs != 0.0 && c->les_delta != 0.0))
            return drp_fail(DRP_ERR_UNSUPPORTED, "the Allen-Cahn model runs in FP64 without body force or LES");
    }
    if (c->advect_k_high != 0) {
        const int kh = c->advect_k_high, kl = c->advect_k_low;
        if (kh < 2 || kh > 5) return drp_fail(DRP_ERR_ARG, "advect_k_high must be 0 or 2..5 (got %d)", kh);
        if (kl < 0 || kl >= kh) return drp_fail(DRP_ERR_ARG, "advect_k_low must lie in 0 .. advect_k_high-1 (got %d)", kl);
        if (2 * (kh - 1) >= c->nx || 2 * (kh - 1) >= c->ny || 2 * (kh - 1) >= c->nz)
            return drp_fail(DRP_ERR_ARG, "grid too small for the advecting shells");
        if (c->precision != DRP_FP64 || c->model != DRP_MODEL_FREE_ENERGY || (c->les_cs != 0.0 && c->les_delta != 0.0))
            return drp_fail(DRP_ERR_UNSUPPORTED, "scale-separated advection runs in FP64, free-energy model, "
                                                 "without LES");
    } else if (c->advect_k_low != 0 || c->advect_start != 0) {
        return drp_fail(DRP_ERR_ARG, "advect_k_low and advect_start need advect_k_high");
    }
    {
        const int rx = c->les_ranks_x, ry = c->les_ranks_y, rz = c->les_ranks_z;
        const bool off = rx == 0 && ry == 0 && rz == 0;
        if (!off && (rx < 1 || ry < 1 || rz < 1))
            return drp_fail(DRP_ERR_ARG, "les_ranks must be all 0 or all positive (got %d,%d,%d)", rx, ry, rz);
        if (!off && (c->nx % rx || c->ny % ry || c->nz % rz))
            return drp_fail(DRP_ERR_ARG, "les_ranks %d,%d,%d do not divide the grid %dx%dx%d", rx, ry, rz, c->nx,
                            c->ny, c->nz);
    }
    return DRP_OK;
}

// default tile shape: the fastest at 256^3 on MI355X (drp_autotune): 32x4x2 for
// FP64 (two 512-thread workgroups per CU), 32x4x4 for FP32 (two 1024-thread
// workgroups per CU at <= 64 VGPRs); DROPLET_SHAPE overrides
int default_shape(int precision) {
    if (const char* e = std::getenv("DROPLET_SHAPE")) {
        int s = std::atoi(e);
        if (s >= 0 && s < SHAPE_COUNT) return s;
    }
    return precision == DRP_FP64 ? SHAPE_32x4x2 : SHAPE_32x4x4;
}
// per-variant defaults measured at 256^3 on MI355X (drp_autotune): the FP64 step
// with a force field (the turbulent-droplet case) runs fastest on 32x4x4 tiles
// (3.05 against 3.10 ms per step with LES and spectral forcing), the Allen-Cahn
// step on 16x4x4 tiles (1.65 against 1.76 ms)
static int pick_shape(const drp_config& c) {
    if (std::getenv("DROPLET_SHAPE")) return default_shape(c.precision);
    if (c.model == DRP_MODEL_ALLEN_CAHN) return SHAPE_16x4x4;
    if (c.precision == DRP_FP64 && (c.force_kind == DRP_FORCE_FIELD || c.force_kind == DRP_FORCE_SPECTRAL))
        return SHAPE_32x4x4;
    return default_shape(c.precision);
}

extern "C" int drp_create(const drp_config* cfg, drp_lattice** out) { return lattice_create(cfg, nullptr, out); }

int lattice_create(const drp_config* cfg, hipStream_t stream, drp_lattice** out) {
    if (!out) return drp_fail(DRP_ERR_ARG, "null output handle");
    *out = nullptr;
    int rc = validate(cfg);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return drp_fail(DRP_ERR_HIP, "no HIP device");
    drp_lattice* L = new (std::nothrow) drp_lattice();
    if (!L) return drp_fail(DRP_ERR_ALLOC, "host allocation failed");
    L->cfg = *cfg;
    L->phys = make_phys(*cfg);
    L->n = (int64_t)cfg->nx * cfg->ny * cfg->nz;
    L->shape = pick_shape(*cfg);
    L->stream = stream;
    (void)hipGetDevice(&L->dev);
    rc = lattice_alloc(L);
    if (rc) {
        lattice_free(L);
        delete L;
        return rc;
    }
    *out = L;
    return DRP_OK;
}

extern "C" void drp_destroy(drp_lattice* L) {
    if (!L) return;
    lattice_free(L);
    delete L;
}

int lattice_alloc(drp_lattice* L) {
    const int64_t n = L->n;
    auto dalloc = [&](void** p, size_t bytes) -> int {
        if (hipMalloc(p, bytes) != hipSuccess) {
            *p = nullptr;
            return drp_fail(DRP_ERR_ALLOC, "hipMalloc(%zu) failed", bytes);
        }
        return DRP_OK;
    };
    if (!L->stream) {
        HIP_TRY(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
        L->owns_stream = true;
    }
    int rc = 0;
    const size_t esz = (L->cfg.precision == DRP_FP64) ? sizeof(double) : sizeof(float);
    for (int k = 0; k < 2 && !rc; k++) {
        rc |= dalloc(&L->pop[k], 2 * 15 * (size_t)n * esz);
        rc |= dalloc((void**)&L->mac[k], 5 * (size_t)n * esz);
    }
    if (!rc) rc = dalloc((void**)&L->range, 6 * sizeof(uint64_t));
    const bool noneq = L->phys.les && L->cfg.les_model == DRP_LES_NONEQ;
    if (!rc && L->phys.les && !noneq) rc = dalloc((void**)&L->tau, (size_t)n * sizeof(double));
    for (int k = 0; k < 2 && !rc && noneq; k++) rc = dalloc((void**)&L->qh[k], (size_t)n * sizeof(double));
    if (!rc && (L->cfg.force_kind == DRP_FORCE_FIELD || L->cfg.force_kind == DRP_FORCE_SPECTRAL)) {
        rc = dalloc((void**)&L->force, 3 * (size_t)n * sizeof(double));
        if (!rc) HIP_TRY(hipMemsetAsync(L->force, 0, 3 * (size_t)n * sizeof(double), L->stream));
    }
    // truncated-DFT work space for the modes |k_a| < kc: the spectral forcing
    // (kc = force_k) and the scale-separated advection (kc = advect_k_high)
    auto spectral_alloc = [&](double** buf, SpectralWork& w, int kc) -> int {
        const int nx = L->cfg.nx, ny = L->cfg.ny, nz = L->cfg.nz;
        int r = dalloc((void**)buf, spectral_work_doubles(nx, ny, nz, kc) * sizeof(double));
        if (r) return r;
        spectral_work_layout(w, *buf, nx, ny, nz, kc);
        // e^{2 pi i j / n}, j = 0..n-1, per axis, evaluated on the host
        const int ns[3] = {nx, ny, nz};
        double2* dst[3] = {w.twx, w.twy, w.twz};
        const double pi = std::acos(-1.0);
        for (int a = 0; a < 3; a++) {
            std::vector<double2> t(ns[a]);
            for (int j = 0; j < ns[a]; j++) {
                const double th = 2.0 * pi * (double)j / (double)ns[a];
                t[j] = make_double2(std::cos(th), std::sin(th));
            }
            HIP_TRY(hipMemcpy(dst[a], t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice));
        }
        return DRP_OK;
    };
    if (!rc && L->cfg.force_kind == DRP_FORCE_SPECTRAL) rc = spectral_alloc(&L->spec, L->specw, (int)L->cfg.force_k);
    if (!rc && L->cfg.advect_k_high != 0) {
        rc = spectral_alloc(&L->spec_adv, L->specw_adv, L->cfg.advect_k_high);
        if (!rc) rc = dalloc((void**)&L->uadv, 3 * (size_t)n * sizeof(double));
    }
    if (!rc && L->cfg.force_kind == DRP_FORCE_TAYLOR_GREEN) {
        const int nx = L->cfg.nx, nz = L->cfg.nz;
        rc = dalloc((void**)&L->tg, (2 * (size_t)nx + 2 * (size_t)nz) * sizeof(double));
        if (!rc) {
            // sin/cos tables of kw*x and kw*z (externalforce :682-690), evaluated on the host
            std::vector<double> t(2 * nx + 2 * nz);
            const double k = L->cfg.force_k;
            for (int x = 0; x < nx; x++) {
                t[x] = std::sin(k * (double)x);
                t[nx + x] = std::cos(k * (double)x);
            }
            for (int z = 0; z < nz; z++) {
                t[2 * nx + z] = std::sin(k * (double)z);
                t[2 * nx + nz + z] = std::cos(k * (double)z);
            }
            HIP_TRY(hipMemcpy(L->tg, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    return rc;
}

void lattice_free(drp_lattice* L) {
    if (L->stream) (void)hipStreamSynchronize(L->stream);
    for (int k = 0; k < 2; k++) {
        if (L->pop[k]) (void)hipFree(L->pop[k]);
        if (L->mac[k]) (void)hipFree(L->mac[k]);
    }
    if (L->range) (void)hipFree(L->range);
    if (L->tau) (void)hipFree(L->tau);
    for (int k = 0; k < 2; k++)
        if (L->qh[k]) (void)hipFree(L->qh[k]);
    if (L->force) (void)hipFree(L->force);
    if (L->spec) (void)hipFree(L->spec);
    if (L->spec_adv) (void)hipFree(L->spec_adv);
    if (L->uadv) (void)hipFree(L->uadv);
    if (L->tg) (void)hipFree(L->tg);
    if (L->partial) (void)hipFree(L->partial);
    if (L->mac64) (void)hipFree(L->mac64);
    if (L->census) (void)hipFree(L->census);
    if (L->budget) (void)hipFree(L->budget);
    if (L->stream && L->owns_stream) (void)hipStreamDestroy(L->stream);
    L->stream = nullptr;
}

// -------------------------------------------------------------------- init

// (re)start the rotating phi-range slots: slot 0 holds the range used by the
// first step, (max, min) = (phi2, phi1) as ini_op uses (:278); slot 1 is in the
// neutral state (max, min) = (phi1, phi2) that physics starts its scan from (:557-558).
int reset_range(drp_lattice* L) {
    HIP_TRY(launch_range_set(L->range, 0, L->cfg.phi2, L->cfg.phi1, L->stream));
    HIP_TRY(launch_range_set(L->range, 1, L->cfg.phi1, L->cfg.phi2, L->stream));
    HIP_TRY(launch_range_set(L->range, 2, L->cfg.phi1, L->cfg.phi2, L->stream));
    return DRP_OK;
}

// LES tau_g for the next step, in the tau field the step kernels read.  After a
// step (fresh = false) from the state it left and the one it started from, with
// the neighbours outside a node's rank block one step old, as the reference
// computes it at the end of physics; for a state that replaced the lattice's
// (initialisation, set populations, restart; fresh = true) with every neighbour
// current, as ini_op does.  The phi range is the slot the next step reads.
int les_update_tau(drp_lattice* L, bool fresh) {
    if (!L->phys.les) return DRP_OK;
    const drp_config& c = L->cfg;
    if (c.les_model == DRP_LES_NONEQ) {
        // the step kernel keeps sqrt(Pi:Pi) itself; a state that replaced the
        // lattice's is an equilibrium or a restart: no non-equilibrium flux yet
        if (fresh)
            for (int k = 0; k < 2; k++) HIP_TRY(hipMemsetAsync(L->qh[k], 0, (size_t)L->n * sizeof(double), L->stream));
        return DRP_OK;
    }
    LesBlocks B{0, 0, 0};
    if (!fresh && c.les_ranks_x > 0) B = LesBlocks{c.nx / c.les_ranks_x, c.ny / c.les_ranks_y, c.nz / c.les_ranks_z};
    const int slot = (int)(L->step % 3);
    const int old = fresh ? L->cur : 1 - L->cur;
    if (c.precision == DRP_FP64)
        HIP_TRY(launch_les_tau_fp64(c.nx, c.ny, c.nz, L->phys, (const double*)L->mac[L->cur], (const double*)L->mac[old],
                                    B, L->range, slot, L->tau, L->stream));
    else
        HIP_TRY(launch_les_tau_fp32(c.nx, c.ny, c.nz, L->phys, (const float*)L->mac[L->cur], (const float*)L->mac[old],
                                    B, L->range, slot, (float*)L->tau, L->stream));
    return DRP_OK;
}

static int init_from_host_fields(drp_lattice* L, const double* const* host) {
    const int64_t n = L->n;
    if (L->cfg.precision != DRP_FP64) {
        const int rc = fp32_init_from_host_fields(L, host);
        return rc ? rc : les_update_tau(L, true);
    }
    for (int k = 0; k < 5; k++)
        HIP_TRY(hipMemcpyAsync((double*)L->mac[L->cur] + k * n, host[k], n * sizeof(double), hipMemcpyHostToDevice,
                               L->stream));
    double* f = (double*)L->pop[L->cur];
    HIP_TRY(launch_init_equilibrium_fp64(L->cfg.nx, L->cfg.ny, L->cfg.nz, L->phys, (double*)L->mac[L->cur], f,
                                         f + 15 * n, L->stream));
    int rc = reset_range(L);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(L->stream));
    L->step = 0;
    L->initialized = true;
    L->eq_init = true;
    return les_update_tau(L, true);
}

extern "C" int drp_init_droplets(drp_lattice* L, const drp_drop* drops, int ndrops, double p_init) {
    if (!L) return drp_fail(DRP_ERR_ARG, "null lattice");
    if (ndrops < 0 || (ndrops > 0 && !drops)) return drp_fail(DRP_ERR_ARG, "bad droplet list");
    const drp_config& c = L->cfg;
    const int64_t n = L->n;
    std::vector<double> phi(n), p(n, p_init), zero(n, 0.0);
    // sharp step, ini :224-235: phi2 where (x-xc)^2+(y-yc)^2+(z-zc)^2 <= (0.5 D)^2
    for (int z = 0; z < c.nz; z++)
        for (int y = 0; y < c.ny; y++)
            for (int x = 0; x < c.nx; x++) {
                double v = c.phi1;
                for (int k = 0; k < ndrops; k++) {
                    const double dx = (double)x - drops[k].xc, dy = (double)y - drops[k].yc,
                                 dz = (double)z - drops[k].zc;
                    const double dif = dx * dx + dy * dy + dz * dz;
                    const double r = 0.5 * drops[k].diameter;
                    if (dif <= r * r) v = c.phi2;
                }
                phi[x + (int64_t)c.nx * (y + (int64_t)c.ny * z)] = v;
            }
    const double* f[5] = {phi.data(), p.data(), zero.data(), zero.data(), zero.data()};
    return init_from_host_fields(L, f);
}

// Diffuse-interface droplets: phi = (phi1 + phi2)/2 + (phi2 - phi1)/2 tanh((R - r)/(sqrt2 W)),
// r the distance to the centre, the largest value over the droplets at each node.
extern "C" int drp_init_droplets_tanh(drp_lattice* L, const drp_drop* drops, int ndrops, double p_init,
                                      double width) {
    if (!L) return drp_fail(DRP_ERR_ARG, "null lattice");
    if (ndrops < 1 || !drops) return drp_fail(DRP_ERR_ARG, "bad droplet list");
    if (!(width > 0.0)) return drp_fail(DRP_ERR_ARG, "interface width must be positive");
    const drp_config& c = L->cfg;
    const int64_t n = L->n;
    const double mid = 0.5 * (c.phi1 + c.phi2), half = 0.5 * (c.phi2 - c.phi1), s = std::sqrt(2.0) * width;
    std::vector<double> phi(n), p(n, p_init), zero(n, 0.0);
    for (int z = 0; z < c.nz; z++)
        for (int y = 0; y < c.ny; y++)
            for (int x = 0; x < c.nx; x++) {
                double v = -1.0;
                for (int k = 0; k < ndrops; k++) {
                    const double dx = (double)x - drops[k].xc, dy = (double)y - drops[k].yc,
                                 dz = (double)z - drops[k].zc;
                    const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                    v = std::fmax(v, std::tanh((0.5 * drops[k].diameter - r) / s));
                }
                phi[x + (int64_t)c.nx * (y + (int64_t)c.ny * z)] = mid + half * v;
            }
    const double* f[5] = {phi.data(), p.data(), zero.data(), zero.data(), zero.data()};
    return init_from_host_fields(L, f);
}

extern "C" int drp_init_fields(drp_lattice* L, const double* phi, const double* p, const double* u1,
                               const double* u2, const double* u3) {
    if (!L || !phi || !p || !u1 || !u2 || !u3) return drp_fail(DRP_ERR_ARG, "null argument");
    const double* f[5] = {phi, p, u1, u2, u3};
    return init_from_host_fields(L, f);
}

// -------------------------------------------------------------------- step

StepArgs make_step_args(drp_lattice* L, int64_t nstep) {
    const drp_config& c = L->cfg;
    const Phys& P = L->phys;
    const int64_t n = L->n;
    StepArgs a{};
    a.nx = c.nx;
    a.ny = c.ny;
    a.nz = c.nz;
    a.n = n;
    double* fi = (double*)L->pop[L->cur];
    double* fo = (double*)L->pop[1 - L->cur];
    a.f_in = fi;
    a.g_in = fi + 15 * n;
    a.f_out = fo;
    a.g_out = fo + 15 * n;
    a.mac_in = (const double*)L->mac[L->cur];
    a.mac_out = (double*)L->mac[1 - L->cur];
    a.tau_field = P.les ? L->tau : nullptr;  // null for the non-equilibrium model (no tau field)
    if (P.les && c.les_model == DRP_LES_NONEQ) {
        const double cs = c.les_cs, dl = c.les_delta;
        a.qh_in = L->qh[L->cur];
        a.qh_out = L->qh[1 - L->cur];
        a.neq_c18 = 18.0 * cs * cs * dl * dl;
        a.neq_c6 = 6.0 * cs * cs * dl * dl;
        a.neq_cc = cs * cs * dl * dl;
    }
    a.tau_f = P.tau_f;
    a.kf = P.kf;
    a.kf6 = P.kf6;
    a.kg = P.kg;
    a.T = P.T;
    a.a = P.a;
    a.b = P.b;
    a.nu1 = P.nu1;
    a.dnu = P.dnu;
    a.tau_const = 0.5 + 3.0 * P.nu1;
    a.ekf[0] = ekappa<0>(P.kf);
    a.ekf[1] = ekappa<1>(P.kf);
    a.ekf[2] = ekappa<7>(P.kf);
    a.ekg[0] = ekappa<0>(P.kg);
    a.ekg[1] = ekappa<1>(P.kg);
    a.ekg[2] = ekappa<7>(P.kg);
    a.tg_amp = P.tg_amp;
    a.phi1 = P.phi1;
    a.phi2 = P.phi2;
    a.wall_axis = c.wall_axis;
    a.wall_u_now = (nstep >= c.wall_start) ? c.wall_u : 0.0; // bounce_back_LBM :670-674
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
    a.uadv = (c.advect_k_high != 0 && nstep >= c.advect_start) ? L->uadv : nullptr;
    a.model = c.model;
    a.ac_a = c.ac_a;
    a.ac_m = (1.0 - c.ac_a) * 1.0 / 6.0;  // M = (1 - A) ds / 6, dropac.f90:65
    a.ac_w = c.ac_width;
    a.range = L->range;
    a.slot_read = (int)((nstep - 1) % 3);
    a.slot_write = (int)(nstep % 3);
    a.slot_reset = (int)((nstep + 1) % 3);
    return a;
}

extern "C" int drp_step(drp_lattice* L, int nsteps) {
    if (!L) return drp_fail(DRP_ERR_ARG, "null lattice");
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "drp_step before drp_init_*");
    if (nsteps < 0) return drp_fail(DRP_ERR_ARG, "nsteps < 0");
    if (L->cfg.precision != DRP_FP64) return fp32_step(L, nsteps);
    for (int k = 0; k < nsteps; k++) {
        const int64_t nstep = L->step + 1;
        int rc = spectral_update(L, nstep);
        if (rc) return rc;
        rc = advect_update(L, nstep);
        if (rc) return rc;
        StepArgs a = make_step_args(L, nstep);
        const hipError_t e = launch_step_fp64(a, L->shape, L->stream);
        if (e == hipErrorInvalidConfiguration)
            return drp_fail(DRP_ERR_UNSUPPORTED, "tile shape %d: the staging of this step variant does not fit in LDS",
                            L->shape);
        HIP_TRY(e);
        L->cur = 1 - L->cur;
        L->step = nstep;
        rc = les_update_tau(L, false);
        if (rc) return rc;
    }
    return DRP_OK;
}

// externalforce at the start of loop iteration n >= step_input, from the
// velocity of the current state (program/lbm_drop_iht_xy.f90:1008-1014)
int spectral_update(drp_lattice* L, int64_t nstep) {
    const drp_config& c = L->cfg;
    if (c.force_kind != DRP_FORCE_SPECTRAL || nstep < c.force_start) return DRP_OK;
    const int64_t n = L->n;
    double *fx = L->force, *fy = L->force + n, *fz = L->force + 2 * n;
    if (c.precision == DRP_FP64) {
        const double* m = (const double*)L->mac[L->cur];
        HIP_TRY(launch_spectral_force(L->specw, c.nx, c.ny, c.nz, (int)c.force_k, c.force_amp, m + 2 * n, m + 3 * n,
                                      m + 4 * n, fx, fy, fz, L->stream));
    } else {
        const float* m = (const float*)L->mac[L->cur];
        HIP_TRY(launch_spectral_force(L->specw, c.nx, c.ny, c.nz, (int)c.force_k, c.force_amp, m + 2 * n, m + 3 * n,
                                      m + 4 * n, fx, fy, fz, L->stream));
    }
    return DRP_OK;
}

// scale-separated advection: from step advect_start on, the band-pass
// advect_k_low <= int|k| < advect_k_high of the current velocity, which the phi
// equilibria of the step carry (low_pass at the start of loop iteration n >=
// step_input, program/lbm_drop_scale_xy.f90:1022-1038)
int advect_update(drp_lattice* L, int64_t nstep) {
    const drp_config& c = L->cfg;
    if (c.advect_k_high == 0 || nstep < c.advect_start) return DRP_OK;
    const int64_t n = L->n;
    const double* m = (const double*)L->mac[L->cur];
    HIP_TRY(launch_spectral_band(L->specw_adv, c.nx, c.ny, c.nz, c.advect_k_low, c.advect_k_high, m + 2 * n,
                                 m + 3 * n, m + 4 * n, L->uadv, L->uadv + n, L->uadv + 2 * n, L->stream));
    return DRP_OK;
}

extern "C" int drp_sync(drp_lattice* L) {
    if (!L) return drp_fail(DRP_ERR_ARG, "null lattice");
    HIP_TRY(hipStreamSynchronize(L->stream));
    return DRP_OK;
}

extern "C" int64_t drp_step_count(const drp_lattice* L) { return L ? L->step : -1; }

// -------------------------------------------------------------------- access

extern "C" int drp_get_macro(drp_lattice* L, int field, double* host, size_t n) {
    if (!L || !host) return drp_fail(DRP_ERR_ARG, "null argument");
    if (field < 0 || field > 4) return drp_fail(DRP_ERR_ARG, "unknown field %d", field);
    if (n != (size_t)L->n) return drp_fail(DRP_ERR_ARG, "size %zu != %lld nodes", n, (long long)L->n);
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    if (L->cfg.precision != DRP_FP64) return fp32_get_macro(L, field, host);
    HIP_TRY(hipMemcpyAsync(host, (const double*)L->mac[L->cur] + field * L->n, n * sizeof(double),
                           hipMemcpyDeviceToHost, L->stream));
    HIP_TRY(hipStreamSynchronize(L->stream));
    return DRP_OK;
}

extern "C" int drp_get_populations(drp_lattice* L, int set, double* host, size_t n) {
    if (!L || !host) return drp_fail(DRP_ERR_ARG, "null argument");
    if (set != DRP_POP_PHI && set != DRP_POP_MOM) return drp_fail(DRP_ERR_ARG, "unknown population set %d", set);
    if (set == DRP_POP_PHI && L->cfg.model == DRP_MODEL_ALLEN_CAHN)
        return drp_fail(DRP_ERR_UNSUPPORTED, "the Allen-Cahn model has no phi populations");
    if (n != 15 * (size_t)L->n) return drp_fail(DRP_ERR_ARG, "size %zu != 15 x %lld", n, (long long)L->n);
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    if (L->cfg.precision != DRP_FP64) return fp32_get_populations(L, set, host);
    const double* src = (const double*)L->pop[L->cur] + (set == DRP_POP_MOM ? 15 * L->n : 0);
    HIP_TRY(hipMemcpyAsync(host, src, n * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIP_TRY(hipStreamSynchronize(L->stream));
    return DRP_OK;
}

// Populations from the host, moments recomputed (which: 1 phi, 2 p and u).  A
// replaced state gets a fresh LES tau; an injection (program/LES_drop.f90:919-945)
// keeps the tau the last step left.
static int set_pops(drp_lattice* L, int set, const double* host, size_t n, int which, bool replace) {
    if (!L || !host) return drp_fail(DRP_ERR_ARG, "null argument");
    if (set != DRP_POP_PHI && set != DRP_POP_MOM) return drp_fail(DRP_ERR_ARG, "unknown population set %d", set);
    if (set == DRP_POP_PHI && L->cfg.model == DRP_MODEL_ALLEN_CAHN)
        return drp_fail(DRP_ERR_UNSUPPORTED, "the Allen-Cahn model has no phi populations");
    if (n != 15 * (size_t)L->n) return drp_fail(DRP_ERR_ARG, "size %zu != 15 x %lld", n, (long long)L->n);
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    if (L->cfg.precision != DRP_FP64) {
        const int rc = fp32_set_populations(L, set, host, which);
        return rc || !replace ? rc : les_update_tau(L, true);
    }
    double* base = (double*)L->pop[L->cur];
    double* dst = base + (set == DRP_POP_MOM ? 15 * L->n : 0);
    L->eq_init = false;
    HIP_TRY(hipMemcpyAsync(dst, host, n * sizeof(double), hipMemcpyHostToDevice, L->stream));
    HIP_TRY(launch_moments_fp64(L->cfg.nx, L->cfg.ny, L->cfg.nz, base, base + 15 * L->n, (double*)L->mac[L->cur],
                                which, L->stream));
    HIP_TRY(hipStreamSynchronize(L->stream));
    return replace ? les_update_tau(L, true) : DRP_OK;
}

extern "C" int drp_set_populations(drp_lattice* L, int set, const double* host, size_t n) {
    return set_pops(L, set, host, n, set == DRP_POP_PHI ? 1 : 2, true);
}

extern "C" int drp_inject_momentum(drp_lattice* L, const double* g_host, size_t n) {
    return set_pops(L, DRP_POP_MOM, g_host, n, 2, false);
}

extern "C" int drp_set_force(drp_lattice* L, const double* fx, const double* fy, const double* fz, size_t n) {
    if (!L || !fx || !fy || !fz) return drp_fail(DRP_ERR_ARG, "null argument");
    if (L->cfg.force_kind != DRP_FORCE_FIELD) return drp_fail(DRP_ERR_STATE, "force_kind is not DRP_FORCE_FIELD");
    if (n != (size_t)L->n) return drp_fail(DRP_ERR_ARG, "size %zu != %lld nodes", n, (long long)L->n);
    const double* src[3] = {fx, fy, fz};
    for (int k = 0; k < 3; k++)
        HIP_TRY(hipMemcpyAsync(L->force + k * L->n, src[k], n * sizeof(double), hipMemcpyHostToDevice, L->stream));
    HIP_TRY(hipStreamSynchronize(L->stream));
    return DRP_OK;
}

extern "C" int drp_get_force(drp_lattice* L, double* fx, double* fy, double* fz, size_t n) {
    if (!L || !fx || !fy || !fz) return drp_fail(DRP_ERR_ARG, "null argument");
    if (!L->force) return drp_fail(DRP_ERR_STATE, "no force field (force_kind is not FIELD or SPECTRAL)");
    if (n != (size_t)L->n) return drp_fail(DRP_ERR_ARG, "size %zu != %lld nodes", n, (long long)L->n);
    double* dst[3] = {fx, fy, fz};
    for (int k = 0; k < 3; k++)
        HIP_TRY(hipMemcpyAsync(dst[k], L->force + k * L->n, n * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIP_TRY(hipStreamSynchronize(L->stream));
    return DRP_OK;
}

// -------------------------------------------------------------------- diagnostics

extern "C" int drp_reduce(drp_lattice* L, drp_stats* st) {
    if (!L || !st) return drp_fail(DRP_ERR_ARG, "null argument");
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    const double* mac = nullptr;
    if (L->cfg.precision == DRP_FP64) {
        mac = (const double*)L->mac[L->cur];
    } else {
        int rc = fp32_macro_fp64_view(L, &mac);
        if (rc) return rc;
    }
    const int64_t nb = (L->n + 255) / 256;
    if (L->partial_cap < nb) {
        if (L->partial) (void)hipFree(L->partial);
        L->partial = nullptr;
        HIP_TRY(hipMalloc((void**)&L->partial, nb * NPART * sizeof(double)));
        L->partial_cap = nb;
    }
    int nblocks = 0;
    HIP_TRY(launch_reduce_fp64(L->cfg.nx, L->cfg.ny, L->cfg.nz, L->phys, mac, L->partial, &nblocks, L->stream));
    std::vector<double> part((size_t)nblocks * NPART);
    uint64_t rg[2];
    HIP_TRY(hipMemcpyAsync(part.data(), L->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIP_TRY(hipMemcpyAsync(rg, L->range + 2 * (L->step % 3), 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, L->stream));
    HIP_TRY(hipStreamSynchronize(L->stream));
    // Neumaier-compensated sum of the workgroup partials (thousands of terms)
    double s[NPART] = {0}, comp[NPART] = {0};
    for (int b = 0; b < nblocks; b++)
        for (int k = 0; k < NPART; k++) {
            const double v = part[(size_t)b * NPART + k], t = s[k] + v;
            comp[k] += (std::fabs(s[k]) >= std::fabs(v)) ? (s[k] - t) + v : (v - t) + s[k];
            s[k] = t;
        }
    for (int k = 0; k < NPART; k++) s[k] += comp[k];
    stats_from_sums(L, s, rg, st);
    return DRP_OK;
}

// drp_stats from the compensated sums of the NPART diagnostics and the phi range slot
void stats_from_sums(const drp_lattice* L, const double* s, const uint64_t* rg, drp_stats* st) {
    const double nn = (double)L->n;
    std::memset(st, 0, sizeof(*st));
    st->phi_mass = s[0];
    st->phi_max = from_ordered_bits(rg[0]);
    st->phi_min = from_ordered_bits(rg[1]);
    st->kinetic = s[1] / nn;
    st->dissipation = s[2] / nn * 0.5 * L->cfg.nu1;
    st->interface_area = s[3];
    st->free_energy = s[4];
    st->p_mean = s[5] / nn;
    st->momentum[0] = s[6];
    st->momentum[1] = s[7];
    st->momentum[2] = s[8];
    st->interaction = s[9];
}

extern "C" int drp_deformation(drp_lattice* L, int mode, double* Lo, double* Bo, double* Do) {
    if (!L || !Lo || !Bo || !Do) return drp_fail(DRP_ERR_ARG, "null argument");
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    if (mode < 0 || mode > 2) return drp_fail(DRP_ERR_ARG, "unknown deformation mode %d", mode);
    return deformation(L, mode, Lo, Bo, Do);
}

// -------------------------------------------------------------------- measurement / tuning

extern "C" int drp_time_steps(drp_lattice* L, int nsteps, double* ms) {
    if (!L || !ms) return drp_fail(DRP_ERR_ARG, "null argument");
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, L->stream));
    int rc = drp_step(L, nsteps);
    HIP_TRY(hipEventRecord(e1, L->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t;
    return rc;
}

extern "C" int drp_tile_shape_count(void) { return SHAPE_COUNT; }

extern "C" int drp_set_tile_shape(drp_lattice* L, int shape) {
    if (!L) return drp_fail(DRP_ERR_ARG, "null lattice");
    if (shape < 0 || shape >= SHAPE_COUNT) return drp_fail(DRP_ERR_ARG, "tile shape %d out of range", shape);
    L->shape = shape;
    return DRP_OK;
}

extern "C" int drp_get_tile_shape(const drp_lattice* L) { return L ? L->shape : -1; }

extern "C" int drp_autotune(const drp_config* cfg, int nsteps, int* best, double* ms_out) {
    if (!cfg || !best || nsteps < 1) return drp_fail(DRP_ERR_ARG, "bad autotune arguments");
    drp_lattice* L = nullptr;
    int rc = drp_create(cfg, &L);
    if (rc) return rc;
    drp_drop d{0.5 * (cfg->nx - 1), 0.5 * (cfg->ny - 1), 0.5 * (cfg->nz - 1), 0.3 * cfg->nx};
    rc = drp_init_droplets(L, &d, 1, 1.0 / 3.0);
    double bestms = 1e300;
    for (int s = 0; s < SHAPE_COUNT && !rc; s++) {
        L->shape = s;
        double ms = 0.0;
        rc = drp_time_steps(L, 2, &ms);
        if (rc == DRP_ERR_UNSUPPORTED) {  // a shape this variant cannot use
            if (ms_out) ms_out[s] = -1.0;
            rc = DRP_OK;
            continue;
        }
        if (!rc) rc = drp_time_steps(L, nsteps, &ms);
        if (rc) break;
        ms /= nsteps;
        if (ms_out) ms_out[s] = ms;
        if (ms < bestms) {
            bestms = ms;
            *best = s;
        }
    }
    drp_destroy(L);
    return rc;
}
