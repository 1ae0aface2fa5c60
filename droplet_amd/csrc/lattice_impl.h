// lattice_impl.h — internal state of a drp_lattice handle (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>

#include "../../include/droplet.h"
#include "kernels.h"

struct drp_lattice {
    drp_config cfg{};
    drp::Phys phys{};
    int dev = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = false;   // false for members of a batch (shared stream)
    int64_t n = 0;
    int64_t step = 0;      // reference loop counter of the last completed step
    bool initialized = false;
    bool eq_init = false;  // populations are the equilibria of fields set by drp_init_* (no step since)
    int shape = 0;         // tile shape of the step kernel
    int cur = 0;           // which ping-pong buffer holds the current state
    void* pop[2] = {nullptr, nullptr};  // FP64: [f(15n) | g(15n)] doubles; FP32: floats
    void* mac[2] = {nullptr, nullptr};  // phi, p, u1, u2, u3 (5n)
    uint64_t* range = nullptr;          // 3 rotating (max, min) slots of the phi range
    double* tau = nullptr;              // LES relaxation-time field
    double* qh[2] = {nullptr, nullptr}; // non-equilibrium LES: sqrt(Pi:Pi) per node, with the populations' parity
    double* force = nullptr;            // FORCE_FIELD / SPECTRAL: fx | fy | fz
    double* spec = nullptr;             // SPECTRAL: transform work space and twiddle tables
    drp::SpectralWork specw{};
    double* spec_adv = nullptr;         // scale-separated advection: transform work space
    drp::SpectralWork specw_adv{};
    double* uadv = nullptr;             // scale-separated advection: band-passed u1 | u2 | u3
    double* tg = nullptr;               // Taylor-Green sin/cos tables
    double* partial = nullptr;          // diagnostics partial sums
    int64_t partial_cap = 0;
    double* mac64 = nullptr;            // FP32 mode: FP64 copy of the macroscopic fields
    int* census = nullptr;              // droplet census work space (4 n + 4 ints, census.hip)
    double* budget = nullptr;           // energy-budget work space (19 n doubles + partial sums, budget.hip)
    void* fp32_state = nullptr;         // FP32 mode: shift constants etc.
};

int drp_fail(int code, const char* fmt, ...);
drp::Phys make_phys(const drp_config& c);
int lattice_alloc(drp_lattice* L);
int lattice_create(const drp_config* cfg, hipStream_t stream, drp_lattice** out);
void lattice_free(drp_lattice* L);
drp::StepArgs make_step_args(drp_lattice* L, int64_t nstep);
int default_shape(int precision);
// SPECTRAL: recompute the force field from the current velocity (FP64 or FP32 fields)
int spectral_update(drp_lattice* L, int64_t nstep);
// scale-separated advection: the band-passed velocity the step's phi equilibria carry
int advect_update(drp_lattice* L, int64_t nstep);
int reset_range(drp_lattice* L);
// drp_stats from the compensated sums of the NPART diagnostics and the phi range slot
void stats_from_sums(const drp_lattice* L, const double* sums, const uint64_t* range_slot, drp_stats* st);
// LES tau for the next step: after a step (fresh = false, rank-block halos one
// step old) or for a replaced state (fresh = true, all neighbours current)
int les_update_tau(drp_lattice* L, bool fresh);

// deformation.cpp
int deformation(drp_lattice* L, int mode, double* Lo, double* Bo, double* Do);
// second-moment deformation (mode 2) of the current state and dD/dphi on its plane
// (n1 x n2 values, plane index third of the axis that is neither x nor wall-normal)
int deformation_inertia_grad(drp_lattice* L, double* D, double* grad, int* n1, int* n2, int* third);
int copy_phi_plane(drp_lattice* L, int third, double* host, int* n1, int* n2);

// fp32_shifted.cpp
int fp32_init_from_host_fields(drp_lattice* L, const double* const* host);
int fp32_step(drp_lattice* L, int nsteps);
drp::StepArgs32 make_step_args32(drp_lattice* L, int64_t nstep);
int fp32_get_macro(drp_lattice* L, int field, double* host);
int fp32_get_populations(drp_lattice* L, int set, double* host);
int fp32_set_populations(drp_lattice* L, int set, const double* host, int which);
int fp32_macro_fp64_view(drp_lattice* L, const double** mac);
