// kernels.h — launch interfaces between the C-ABI host code and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace drp {

enum { FORCE_NONE = 0, FORCE_TAYLOR_GREEN = 1, FORCE_FIELD = 2 };

// Map the linear workgroup id to a tile so that each XCD (workgroups are dealt
// round-robin to the 8 XCDs) gets a contiguous block of tiles: neighbouring
// tiles, which share halo lines, then share one L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    if ((nb & 7) != 0) return b;
    return (b & 7) * (nb >> 3) + (b >> 3);
}
// batched launch: per-instance step arguments and first-tile offsets in device memory
struct BatchLaunch {
    const void* args;      // StepArgs[ninst] or StepArgs32[ninst]
    const int* tile_start; // [ninst], ascending, tile_start[0] = 0
    int ninst;
    int total_tiles;
};

// tile shapes of the step kernel (nodes per workgroup in x, y, z; two threads per node)
enum { SHAPE_16x4x4 = 0, SHAPE_32x4x2 = 1, SHAPE_64x2x2 = 2, SHAPE_32x2x4 = 3, SHAPE_32x4x4 = 4, SHAPE_32x8x2 = 5,
       SHAPE_64x4x2 = 6, SHAPE_COUNT = 7 };
constexpr int SHAPE_DIMS[SHAPE_COUNT][3] = {{16, 4, 4}, {32, 4, 2}, {64, 2, 2}, {32, 2, 4},
                                            {32, 4, 4}, {32, 8, 2}, {64, 4, 2}};

// tile counts and wrap mode of one lattice for a tile shape
template <typename A> inline void set_tiles(A& a, int TX, int TY, int TZ) {
    a.tiles_x = (a.nx + TX - 1) / TX;
    a.tiles_y = (a.ny + TY - 1) / TY;
    a.tiles_z = (a.nz + TZ - 1) / TZ;
    a.wrap_mod = (a.nx <= TX || a.ny <= TY || a.nz <= TZ) ? 1 : 0;
}

// Physical constants of one lattice as the kernels need them.
struct Phys {
    double tau_f;     // phi-population relaxation time
    double kf, kf6;   // kappa_f and kappa_f/6 (the reference folds kappaf/6.0d0 at compile time)
    double kg;        // kappa_g
    double T, a, b;   // van der Waals constants
    double nu1, dnu;  // nu1 and nu2 - nu1
    double phi1, phi2;
    double tg_amp;    // Taylor-Green force amplitude
    double les_c;     // Cs*Cs*Delta*Delta (evaluated left to right)
    int les;          // Smagorinsky on
};

struct StepArgs {
    int nx, ny, nz;
    int tiles_x, tiles_y, tiles_z;
    int64_t n;
    const double* f_in;   // [15][n] phi populations (reference f)
    const double* g_in;   // [15][n] momentum populations (reference g)
    double* f_out;
    double* g_out;
    const double* mac_in; // phi, p, u1, u2, u3 (5 x n)
    double* mac_out;
    const double* tau_field;              // LES: tau_g per node, else nullptr
    const double *fx, *fy, *fz;           // FORCE_FIELD
    const double *sinx, *cosx, *sinz, *cosz;  // FORCE_TAYLOR_GREEN tables
    double tau_f, kf, kf6, kg, T, a, b, nu1, dnu, tg_amp, phi1, phi2;
    double tau_const;     // tau_g when nu1 == nu2: 0.5 + 3 nu1
    double ekf[3], ekg[3];  // E kappa_f, E kappa_g for the rest, axis and diagonal weights
    double wall_u_now;
    int wall_axis;        // 0, 2 (y) or 3 (z)
    int force_kind;
    uint64_t* range;      // 3 slots x (max, min) as order-preserving bits
    int slot_read, slot_write, slot_reset;
    int wrap_mod;         // set by the launcher: a grid extent does not exceed the tile extent
    int model;            // 0 free energy (phi populations), 1 Allen-Cahn (phi stencil update)
    double ac_a, ac_m, ac_w;  // Allen-Cahn A, mobility M = (1 - A)/6, interface width W
    // non-equilibrium Smagorinsky: sqrt(Pi:Pi) per node from the last step (in) and
    // of this step (out); 18 Cs^2 D^2, 6 Cs^2 D^2, Cs^2 D^2 (left to right)
    const double* qh_in;
    double* qh_out;
    double neq_c18, neq_c6, neq_cc;
    // scale-separated advection: the velocity the phi equilibria carry (u1 | u2 | u3),
    // or nullptr for u itself
    const double* uadv;
};

hipError_t launch_step_fp64(const StepArgs& a, int shape, hipStream_t st, const BatchLaunch* batch = nullptr);

// FP32 shifted storage (step_fp32.hip): populations and fields as FP32
// deviations from the continuous phase at rest
struct StepArgs32 {
    int nx, ny, nz;
    int tiles_x, tiles_y, tiles_z;
    int64_t n;
    const float* f_in;
    const float* g_in;
    float* f_out;
    float* g_out;
    const float* mac_in;   // phi - phi1, p - 1/3, u1, u2, u3
    float* mac_out;
    const float* tau_field;
    const double *fx, *fy, *fz;
    const double *sinx, *cosx, *sinz, *cosz;
    double tg_amp, phi1, phi2;
    float phi1f, Tf, af, bf, one_m_bphi1, kff, kf6f, kgf;
    float rtau_f, rtau_const, nu1f, dnuf;
    float wall_u_now;
    int wall_axis, force_kind;
    uint64_t* range;
    int slot_read, slot_write, slot_reset;
    int wrap_mod;
};
hipError_t launch_step_fp32(const StepArgs32& a, int shape, hipStream_t st, const BatchLaunch* batch = nullptr);

// Rank blocks of the reference's LES decomposition, in nodes per axis (all 0:
// none).  grad u for the eddy viscosity takes every neighbour outside a node's
// block, periodic wraps included, from the state the step started from.
struct LesBlocks {
    int bx, by, bz;
};

// conversions between FP64 absolute and FP32 shifted representations; fref is
// the 15 reference phi populations H_q phi1 + F_q p0(phi1) (host-computed)
struct ShiftRef {
    double fr[15];
    double phi1;
};
hipError_t launch_to_shifted(int64_t n, const ShiftRef& r, const double* pop64, const double* mac64, float* pop32,
                             float* mac32, hipStream_t st);
hipError_t launch_pops_from_shifted(int64_t n, const ShiftRef& r, const float* pop32, double* pop64, int set,
                                    hipStream_t st);
hipError_t launch_pops_to_shifted(int64_t n, const ShiftRef& r, const double* pop64, float* pop32, int set,
                                  hipStream_t st);
hipError_t launch_macro_from_shifted(int64_t n, double phi1, const float* mac32, double* mac64, hipStream_t st);
hipError_t launch_moments_fp32(int64_t n, const float* f, const float* g, float* mac, int which, hipStream_t st);
hipError_t launch_les_tau_fp32(int nx, int ny, int nz, const Phys& ph, const float* mac, const float* mac_old,
                               const LesBlocks& blk, const uint64_t* range, int slot, float* tau, hipStream_t st);

// populations := equilibria of the macroscopic fields (ini, program/lbm_drop_xy.f90:244-251)
hipError_t launch_init_equilibrium_fp64(int nx, int ny, int nz, const Phys& ph, const double* mac, double* f,
                                        double* g, hipStream_t st);
// recompute moments from populations: which & 1 -> phi from f, which & 2 -> p, u from g
hipError_t launch_moments_fp64(int nx, int ny, int nz, const double* f, const double* g, double* mac, int which,
                               hipStream_t st);
// Smagorinsky tau field for the next step (program/LES_drop.f90:598-693) from the
// state a step left (mac), the one it started from (mac_old, read across rank
// blocks) and that step's phi range slot
hipError_t launch_les_tau_fp64(int nx, int ny, int nz, const Phys& ph, const double* mac, const double* mac_old,
                               const LesBlocks& blk, const uint64_t* range, int slot, double* tau, hipStream_t st);
// derived fields of the analysis programs (budget.hip): kind = DRP_DERIVED_*
hipError_t launch_derived_fp64(int nx, int ny, int nz, const Phys& ph, const double* mac, int kind, double* out,
                               hipStream_t st);

// diagnostics partial sums: NPART doubles per workgroup
constexpr int NPART = 10;
hipError_t launch_reduce_fp64(int nx, int ny, int nz, const Phys& ph, const double* mac, double* partial,
                              int* nblocks, hipStream_t st);
// fill a range slot with (max, min) = (vmax, vmin)
hipError_t launch_range_set(uint64_t* range, int slot, double vmax, double vmin, hipStream_t st);

// spectral.hip: linear forcing of the shells 1 <= |k| < kc from the velocity field
struct SpectralWork {
    double2 *A, *B, *C, *G;   // transform stages (A also holds the inverse row stage)
    double2 *twx, *twy, *twz; // e^{2 pi i j / n} tables
    double* elow;             // energy of the forced shells (last call)
};
size_t spectral_work_doubles(int nx, int ny, int nz, int kc);
void spectral_work_layout(SpectralWork& w, double* base, int nx, int ny, int nz, int kc);
hipError_t launch_spectral_force(const SpectralWork& w, int nx, int ny, int nz, int kc, double eps, const double* u1,
                                 const double* u2, const double* u3, double* fx, double* fy, double* fz,
                                 hipStream_t st);
hipError_t launch_spectral_force(const SpectralWork& w, int nx, int ny, int nz, int kc, double eps, const float* u1,
                                 const float* u2, const float* u3, double* fx, double* fy, double* fz,
                                 hipStream_t st);
// band-pass of u keeping the shells klo <= int|k| < khi (2 <= khi <= 5; work space of kc = khi)
hipError_t launch_spectral_band(const SpectralWork& w, int nx, int ny, int nz, int klo, int khi, const double* u1,
                                const double* u2, const double* u3, double* o1, double* o2, double* o3,
                                hipStream_t st);

// adjoint.hip: one reverse step of the FP64 step (equal viscosities, no LES)
struct AdjArgs {
    int nx, ny, nz;
    int64_t n;
    const double* f;        // forward phi populations at step n (15n)
    const double* mac;      // forward macroscopic fields at step n (5n)
    const double* Fin;      // adjoint of f^{n+1} without its stencil part (15n)
    const double* phin_in;  // stencil part of the adjoint of phi^{n+1} (n), or nullptr
    const double* Gin;      // adjoint of g^{n+1} (15n)
    double* Fout;           // adjoint of f^n without its stencil part
    double* Gout;           // adjoint of g^n
    double* lapA;           // adjoint of lap phi at step n (n)
    double* gradA;          // adjoint of grad phi at step n (3n)
    double* acc;            // [workgroups][3]: d/dtau_f, d/dkappa_g, d/dU_wall, accumulated
    double tau_f, tau_g, kf, kf6, kg, T, a, b;
    double rtau_f, rtau_g, rtau_f2;  // 1 / tau_f, 1 / tau_g, 1 / (tau_f tau_f)
    double om_f, om_g;               // 1 - 1 / tau_f, 1 - 1 / tau_g
    int wall_axis;          // 0, 2 (y) or 3 (z)
    int wall_active;        // the walls move during step n -> n+1
    int first;              // step n is the initial equilibrium of constant fields
};
int adj_block_count(int64_t n);
hipError_t launch_adj_collide_fp64(const AdjArgs& a, hipStream_t st);
hipError_t launch_adj_stencil_fp64(int nx, int ny, int nz, const double* lapA, const double* gradA, double* out,
                                   hipStream_t st);
hipError_t launch_adj_seed_area_fp64(int nx, int ny, int nz, const double* phi, double w, double* lapA,
                                     double* gradA, hipStream_t st);
hipError_t launch_adj_add_plane(int nx, int ny, int n1, int n2, int wall_y, int third, const double* vals,
                                double* out, hipStream_t st);

}  // namespace drp
