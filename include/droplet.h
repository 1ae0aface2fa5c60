/*
 * droplet.h — C ABI of the MI355X-native D3Q15 two-phase lattice Boltzmann step.
 *
 * The reference (nakano-gendai/droplet) has no library boundary: its "API" is the
 * fixed call order of `program main` over `module globals`,
 *     par → ini → ini_op → DO n: [step_input injection + externalforce]
 *       → MPI_boundary → various_cal → gphi_cal → equilibrium_cal → MPI_boundary_fg
 *       → fg_cal → [bounce_back_LBM] → physics → [output | outputfg | break_up3]
 * (program/lbm_drop_xy.f90:774-868, prototype_program/two-phase/dropch.f90:1079-1912).
 * Each entry point below replaces one group of those calls; the comment on each
 * names the reference routine(s) it stands for.  The Fortran driver binds these
 * through iso_c_binding (droplet_amd/fortran/droplet_amd.f90), Python through ctypes
 * (droplet_amd/lattice.py).
 *
 * Conventions
 *  - Return 0 on success, a negative DRP_ERR_* code on failure; never abort.
 *    drp_last_error() gives a message for the calling thread's last failure.
 *  - The library owns all device memory; callers own host buffers.
 *  - One lattice handle is bound to one HIP stream; calls on one handle must be
 *    serialised by the caller.  Separate handles are independent.
 *  - Naming: "phi populations" are the order-parameter populations (reference `f`),
 *    "momentum populations" carry pressure and momentum (reference `g`).
 *  - Host field layout is x fastest, then y, then z: index = x + nx*(y + ny*z).
 *    Population arrays are [q][z][y][x] with q = 0..14 ↔ reference i = 1..15.
 *  - Lattice spacing ds = 1 (every reference file uses ds = 1.0d0).
 */
#ifndef DROPLET_AMD_H
#define DROPLET_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DROPLET_ABI_VERSION 3

#define DRP_OK 0
#define DRP_ERR_ARG (-1)         /* invalid argument / size mismatch */
#define DRP_ERR_HIP (-2)         /* HIP runtime failure              */
#define DRP_ERR_STATE (-3)       /* call out of order (e.g. step before init) */
#define DRP_ERR_UNSUPPORTED (-4) /* configuration not implemented on this path */
#define DRP_ERR_ALLOC (-5)       /* device allocation failed         */

/* storage/arithmetic precision of the populations */
#define DRP_FP64 0          /* FP64 storage and arithmetic (reference precision) */
#define DRP_FP32_SHIFTED 1  /* FP32 deviations from a rest equilibrium, FP32 collision */

/* walls: 0 = fully periodic; 2 = walls normal to y; 3 = walls normal to z (reference dropch) */
#define DRP_WALL_NONE 0
#define DRP_WALL_Y 2
#define DRP_WALL_Z 3

/* body force kinds */
#define DRP_FORCE_NONE 0
#define DRP_FORCE_TAYLOR_GREEN 1 /* externalforce, program/lbm_drop_xy.f90:668-693 */
#define DRP_FORCE_FIELD 2        /* user-supplied static field (drp_set_force)      */
#define DRP_FORCE_SPECTRAL 3     /* linear forcing of the shells 1 <= |k| < force_k with
                                    injection rate force_amp, recomputed from u every step
                                    (externalforce, program/lbm_drop_iht_xy.f90:722-841) */

/* interface models */
#define DRP_MODEL_FREE_ENERGY 0 /* Inamuro free-energy model: phi carried by populations
                                   (program/lbm_drop_xy.f90)                          */
#define DRP_MODEL_ALLEN_CAHN 1  /* conservative Allen-Cahn phase field: phi advanced by a
                                   stencil of local equilibria with mobility M = (1-A)/6,
                                   interface width W, no phi populations
                                   (prototype_program/two-phase/dropac.f90:517-535,640-667,
                                   715-728); FP64, no LES, no body force                */

/* Smagorinsky eddy-viscosity estimates */
#define DRP_LES_STRAIN 0  /* nu_e = Cs^2 Delta^2 sqrt(2 S:S), S from the D3Q15 velocity
                             gradient at the end of the step (program/LES_drop.f90:598-693) */
#define DRP_LES_NONEQ 1   /* nu_e from the non-equilibrium momentum flux of the node,
                             Pi = sum c c (g - geq), Q = Pi:Pi:
                             nu_e = Cs^2 Delta^2 (sqrt(nu^2 + 18 Cs^2 Delta^2 sqrt(Q)) - nu)
                                    / (6 Cs^2 Delta^2)
                             (program/LES_2.f90:530-551, 713-733); FP64, les_ranks 0 */

/* macroscopic fields for drp_get_macro / drp_set_macro */
#define DRP_FIELD_PHI 0
#define DRP_FIELD_P 1
#define DRP_FIELD_U1 2
#define DRP_FIELD_U2 3
#define DRP_FIELD_U3 4

/* population sets */
#define DRP_POP_PHI 0 /* reference f */
#define DRP_POP_MOM 1 /* reference g */

typedef struct drp_config {
    int32_t nx, ny, nz;   /* lattice nodes per axis (reference xmax+1, ymax+1, zmax+1); each >= 3,
                             nx * ny < 2^31 */
    int32_t precision;    /* DRP_FP64 | DRP_FP32_SHIFTED                              */
    double tau_f;         /* phi-population relaxation time (ref tauf)                 */
    double nu1, nu2;      /* kinematic viscosity of continuous / dispersed phase       */
    double kappa_f;       /* interface-width parameter (ref kappaf)                    */
    double kappa_g;       /* surface-tension parameter (ref kappag)                    */
    double phi1, phi2;    /* order parameter of continuous / dispersed phase           */
    double a, b, T;       /* van der Waals constants of p0 = phi T/(1-b phi) - a phi^2 */
    double anu;           /* ref Anu (velocity-gradient term in geq); 0 in every run   */
    int32_t wall_axis;    /* DRP_WALL_*                                                */
    int32_t wall_start;   /* walls move (+U top, -U bottom, along x) from this step (ref start) */
    double wall_u;        /* wall speed U (ref uw)                                     */
    int32_t force_kind;   /* DRP_FORCE_*                                               */
    int32_t force_start;  /* first step whose collision includes the force (ref step_input) */
    double force_amp;     /* Taylor-Green amplitude 2 pi^2 U^2/(D_vortex Re); spectral: epsilon */
    double force_k;       /* Taylor-Green wavenumber pi/D_vortex (ref kw); spectral: kc (2..5) */
    double les_cs;        /* Smagorinsky constant Cs (0 = no LES)                      */
    double les_delta;     /* filter width Delta                                        */
    /* LES: the reference computes grad u for the eddy viscosity at the end of each
       step on its MPI rank blocks, taking every neighbour outside a node's block
       (periodic wraps included) from the halo exchanged at the step's start, i.e.
       one step old.  ranks = 0,0,0: all neighbours current; otherwise each count
       >= 1 and divides its axis.  program/LES_drop.f90 runs 8 x 1 x Nz blocks,
       program/LES_drop_xy.f90 4 x Ny x 1. */
    int32_t les_ranks_x;
    int32_t les_ranks_y;
    int32_t les_ranks_z;
    int32_t model;        /* DRP_MODEL_*                                               */
    int32_t les_model;    /* DRP_LES_*: how nu_e is estimated when les_cs, les_delta != 0 */
    /* Scale-separated advection of phi (program/lbm_drop_scale_xy.f90): from step
       advect_start on, the phi-population equilibria carry the band-pass of u that
       keeps the shells advect_k_low <= int|k| < advect_k_high (low_pass :846-903,
       equilibrium_cal :507-560; the momentum equilibria keep the full u); before it,
       and with advect_k_high = 0, they carry u itself.  2 <= advect_k_high <= 5,
       0 <= advect_k_low < advect_k_high; FP64, free-energy model, no LES. */
    int32_t advect_k_high;
    int32_t advect_k_low;
    int32_t advect_start; /* ref step_input                                            */
    double ac_a;          /* Allen-Cahn: relaxation weight A of the phi update; mobility
                             M = (1 - A)/6 (ref dropac.f90 A, M)                       */
    double ac_width;      /* Allen-Cahn: interface width W (ref dropac.f90 W)          */
} drp_config;

typedef struct drp_drop {
    double xc, yc, zc;    /* centre in lattice units                                   */
    double diameter;      /* sharp step: phi2 where |x-c|^2 <= (D/2)^2 (ref ini :224-242) */
} drp_drop;

typedef struct drp_stats {
    double phi_mass;      /* sum phi                                                   */
    double phi_min, phi_max; /* range used by nu(phi) at the next step (ref physics :559-574) */
    double kinetic;       /* 0.5 <|u|^2>  (mean over nodes)                            */
    double dissipation;   /* 0.5 nu1 <(d_a u_b + d_b u_a)^2>, mean over nodes          */
    double interface_area;/* sum |grad phi| (ref drop_analysis.f90:445-474)            */
    double free_energy;   /* sum phi T ln(phi/(1-b phi)) - a phi^2 (bulk part)         */
    double p_mean;        /* <p>                                                       */
    double momentum[3];   /* sum u                                                     */
    double interaction;   /* sum over nodes of sum_ab d_a phi d_b phi S_ab, S = (grad u + grad u^T)/2:
                             a restatement of grad_u_cal / strain_cal / interaction_cal of ref
                             drop_analysis.f90:797-882, routines whose calls the reference has
                             commented out (:1271-1274); its parity is unpinned (closed forms
                             only).  The form the reference runs is the per-scale
                             kappa_g <u_k . div(grad phi grad phi)> of drp_scale_interaction. */
} drp_stats;

typedef struct drp_lattice drp_lattice; /* opaque; owns all device memory */

/* Layout check for bindings (Fortran, ctypes): sizes of the three structs as
 * this library was compiled; returns DROPLET_ABI_VERSION.  Needs no GPU. */
int drp_abi_sizes(size_t* config, size_t* drop, size_t* stats);

/* Fill a config with the reference defaults of program/lbm_drop_xy.f90:9-67
 * (phi1=2.211, phi2=4.895, a=9/49, b=2/21, T=0.55, kappa_f=0.01, tau_f=0.7). */
int drp_config_default(drp_config* cfg);

/* = par + allocation part of ini (program/lbm_drop_xy.f90:117-129, 131-222). */
int drp_create(const drp_config* cfg, drp_lattice** out);

/* = ini + ini_op with the reference sharp-step droplet(s): phi, u=0, p=p_init,
 * then f := feq, g := geq and the nu(phi) range (phi1, phi2)
 * (program/lbm_drop_xy.f90:224-284; prototype_program/two-phase/dropch.f90:186-473). */
int drp_init_droplets(drp_lattice* lat, const drp_drop* drops, int ndrops, double p_init);

/* Diffuse-interface variant of drp_init_droplets (no reference counterpart: the
 * reference starts from the sharp step): phi = (phi1+phi2)/2 + (phi2-phi1)/2 *
 * tanh((D/2 - r)/(sqrt2 * width)), r the distance to a droplet centre, the largest
 * value over the droplets; u = 0, p = p_init, populations := equilibria. */
int drp_init_droplets_tanh(drp_lattice* lat, const drp_drop* drops, int ndrops, double p_init, double width);

/* = ini from arbitrary host fields (each nx*ny*nz doubles): populations := equilibria. */
int drp_init_fields(drp_lattice* lat, const double* phi, const double* p,
                    const double* u1, const double* u2, const double* u3);

/* Advance nsteps time steps (= loop body program/lbm_drop_xy.f90:787-848 with
 * bounce_back_LBM of dropch.f90:1097 when walls are configured).  Asynchronous
 * on the handle's stream. */
int drp_step(drp_lattice* lat, int nsteps);

/* Wait for all work queued on the handle. */
int drp_sync(drp_lattice* lat);

/* Number of steps taken since init (the reference loop counter n). */
int64_t drp_step_count(const drp_lattice* lat);

/* = output (program/lbm_drop_xy.f90:695-719): copy one macroscopic field to host (FP64). */
int drp_get_macro(drp_lattice* lat, int field, double* host, size_t n);

/* = outputfg (program/lbm_drop_xy.f90:721-743): populations to host, [q][z][y][x], FP64. */
int drp_get_populations(drp_lattice* lat, int set, double* host, size_t n);

/* Restart: set one population set from host and recompute the moments it carries. */
int drp_set_populations(drp_lattice* lat, int set, const double* host, size_t n);

/* = step_input injection (program/lbm_drop_xy.f90:789-822): replace the momentum
 * populations, recompute p and u from them; phi and the phi populations are untouched. */
int drp_inject_momentum(drp_lattice* lat, const double* g_host, size_t n);

/* Static body-force field for DRP_FORCE_FIELD (three nx*ny*nz arrays). */
int drp_set_force(drp_lattice* lat, const double* fx, const double* fy, const double* fz, size_t n);

/* Diagnostics reductions (wavefront -> workgroup -> grid, FP64 accumulation). */
int drp_reduce(drp_lattice* lat, drp_stats* stats);

/* drp_reduce without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The partial sums live in one
 * file-scope device array that each GPU holds until the process ends
 * (10 MiB + 128 B); lattices of more than DRP_REDUCE_NOALLOC_MAX_NODES nodes get
 * DRP_ERR_ARG, FP32 lattices DRP_ERR_UNSUPPORTED (their FP64 view allocates).
 * Returns the same statistics as drp_reduce.  Calls that overlap on one GPU share
 * the array and race: serialise them. */
#define DRP_REDUCE_NOALLOC_MAX_NODES ((int64_t)1 << 25)
int drp_reduce_noalloc(drp_lattice* lat, drp_stats* stats);

/* Taylor deformation D = (L-B)/(L+B) in the shear plane.
 * mode 0: break_up3 (prototype_program/two-phase/dropch.f90:857-948);
 * mode 1: L = 2 max r, B = 4 area/(pi L) (prototype_program/two-phase/drop.f90:439-463);
 * mode 2: axes from the second moments of s^2, s = (phi-phi1)/(phi2-phi1), in the
 *         plane (no reference counterpart; smooth in phi, used to state FP32 error
 *         and as the differentiable deformation of drp_tape_*). */
int drp_deformation(drp_lattice* lat, int mode, double* L, double* B, double* D);

/* ---- measurement and tuning (no reference counterpart: the reference times its
 * loop with MPI_Wtime into time.d, program/lbm_drop_xy.f90:850-864) ---- */

/* Run nsteps steps bracketed by HIP events recorded on the handle's own stream;
 * *ms = elapsed device time of those steps.  Blocks until they finish. */
int drp_time_steps(drp_lattice* lat, int nsteps, double* ms);

/* Number of tile shapes of the step kernel, and selection of one for a handle. */
int drp_tile_shape_count(void);
int drp_set_tile_shape(drp_lattice* lat, int shape);
int drp_get_tile_shape(const drp_lattice* lat);

/* Time every tile shape on a temporary lattice of this configuration (a centred
 * droplet, nsteps steps each after 2 warm-up steps) and return the fastest;
 * ms_per_step (may be NULL) receives drp_tile_shape_count() timings. */
int drp_autotune(const drp_config* cfg, int nsteps, int* best_shape, double* ms_per_step);

/* ---- ensembles: many independent lattices, one launch per step ----
 * Replaces the reference's parameter ensembles run as MPI sub-communicator
 * groups (prototype_program/two-phase/multidrop2.f90:235-260,
 * program/LBM.f90:215-232) or as a sequential case loop
 * (program/lbm_drop_iht_xy_many.f90).  Members may differ in grid size and in
 * every physical parameter; they must share precision, the wall / body-force /
 * LES switches, the body-force start step and whether nu1 == nu2.  LES members
 * are not supported.  Members are initialised and read through their own
 * drp_lattice handle (drp_batch_member); they are stepped only together. */
typedef struct drp_batch drp_batch;
int drp_batch_create(const drp_config* cfgs, int count, drp_batch** out);
int drp_batch_size(const drp_batch* b);
/* Borrowed handle of member i (valid until drp_batch_destroy; do not destroy it). */
drp_lattice* drp_batch_member(drp_batch* b, int i);
/* Advance every member by nsteps: one launch per step; six-step cycles are
 * captured once into a hipGraph and replayed (drp_batch_use_graph(b, 0) disables). */
int drp_batch_step(drp_batch* b, int nsteps);
int drp_batch_use_graph(drp_batch* b, int enable);
int drp_batch_set_tile_shape(drp_batch* b, int shape);
/* Autotuner for ensembles: times every tile shape of the batched step on a
 * temporary batch of these members (each with a centred droplet), nsteps steps
 * per shape, and returns the fastest; ms_per_step (may be NULL) receives
 * drp_tile_shape_count() timings.  The caller applies it with
 * drp_batch_set_tile_shape. */
int drp_batch_autotune(const drp_config* cfgs, int count, int nsteps, int* best_shape, double* ms_per_step);
int drp_batch_sync(drp_batch* b);
int drp_batch_time_steps(drp_batch* b, int nsteps, double* ms);
void drp_batch_destroy(drp_batch* b);

/* ---- reverse-mode sensitivities of a run (hand-written discrete adjoint) ----
 * No counterpart in the reference, which has no differentiation (parameters are
 * calibrated there by re-running cases, e.g. the kappag sweep written to
 * stable5.d in prototype_program/two-phase/dropch.f90:985-1000).  A tape records
 * the FP64 run of one lattice from its current state, keeping a checkpoint of
 * the full state every K steps; drp_tape_backward recomputes each K-step segment
 * from its checkpoint and runs the exact adjoint of every step backwards.
 *
 * Loss J = w[0] * D + w[1] * A with D the second-moment Taylor deformation
 * (drp_deformation mode 2) and A = sum |grad phi| (drp_stats.interface_area).
 * Gradient entries (DRP_GRAD_*): the phi-population mobility M = tau_f - 1/2
 * (Inamuro model, lattice units: dJ/dM = dJ/dtau_f), kappa_g (surface tension,
 * sigma = 1.7039 kappa_g^0.9991) and the wall speed U (shear rate 2U/H).
 * Requirements: FP64, nu1 == nu2, no LES, fewer than 2^32 nodes.  When the tape starts right after
 * drp_init_droplets / drp_init_fields, the kappa_g dependence of the initial
 * momentum equilibria is included.  The lattice must be stepped only through
 * the tape while it is recording. */
enum { DRP_GRAD_MOBILITY = 0, DRP_GRAD_KAPPA_G = 1, DRP_GRAD_WALL_U = 2, DRP_GRAD_COUNT = 3 };
typedef struct drp_tape drp_tape;
int drp_tape_create(drp_lattice* lat, int checkpoint_every, drp_tape** out);
int drp_tape_step(drp_tape* t, int nsteps);
int drp_tape_loss(drp_tape* t, const double w[2], double* loss, double* D, double* area);
int drp_tape_backward(drp_tape* t, const double w[2], double* loss, double grad[DRP_GRAD_COUNT]);
/* Device bytes held by the tape's checkpoints and work buffers. */
int64_t drp_tape_bytes(const drp_tape* t);
void drp_tape_destroy(drp_tape* t);

/* ---- the reference's per-rank files (Fortran unformatted sequential, 4-byte
 * record markers; rank r holds x in [rx*nx/ranks_x, ...), y in [ry*ny/ranks_y, ...),
 * all z, with rx = r / ranks_y, ry = r % ranks_y as program/lbm_drop_xy.f90:145-149
 * numbers ranks; loops z, y, x (, i) in each file).
 *   DRP_FILE_FIELDS <dir>/0_<r>_<n>.bin     records (phi, u1, u2, u3, p)    = output (:695-719)
 *   DRP_FILE_FG     <dir>/0_<r>_<n>_fg.bin  records (f_i, g_i)               = outputfg (:721-743)
 *   DRP_FILE_G      <dir>/0_<r>_<n>_fg.bin  records g_i (momentum only)      = step_input input (:789-828)
 *   DRP_FILE_FG_G   <dir>/0_<r>_<n>_fg.bin  records (f_i, g_i), g_i used     = step_input input of
 *                   program/LES_drop.f90 (:919-945), which reads outputfg files and keeps g
 * drp_read_reference_files with DRP_FILE_FG is a restart: populations, moments,
 * the step counter (= n) and the phi range of the restored state.  The LES tau
 * of the restored state is formed with every neighbour current, so a run with
 * les_ranks set continues bit for bit only where no rank-block halo would have
 * been one step old (with les_ranks 0 the restart is bit-identical); with
 * DRP_FILE_G or DRP_FILE_FG_G it injects the momentum populations like
 * drp_inject_momentum (read only).
 * DRP_FILE_LAYOUT_XZ, OR-ed into kind, selects program/LES_drop.f90's
 * decomposition (:153-176): ranks_x slabs in x times ranks_y slabs in z, y whole,
 * rank r = (z slab) * ranks_x + (x slab). */
#define DRP_FILE_FIELDS 0
#define DRP_FILE_FG 1
#define DRP_FILE_G 2
#define DRP_FILE_FG_G 3
#define DRP_FILE_LAYOUT_XZ 0x100
int drp_write_reference_files(drp_lattice* lat, const char* dir, int64_t n, int ranks_x, int ranks_y, int kind);
int drp_read_reference_files(drp_lattice* lat, const char* dir, int64_t n, int ranks_x, int ranks_y, int kind);

/* Current body-force field (DRP_FORCE_FIELD / DRP_FORCE_SPECTRAL: the field the last
 * step used) to host, three arrays of nx*ny*nz doubles.  No reference counterpart
 * (forcex/forcey/forcez are module arrays there, program/lbm_drop_iht_xy.f90:722-841). */
int drp_get_force(drp_lattice* lat, double* fx, double* fy, double* fz, size_t n);

/* Achievable HBM bandwidth of a plain stream copy of `bytes` (read + write counted),
 * averaged over reps launches on a private stream. */
int drp_copy_bandwidth(size_t bytes, int reps, double* gbs);

/* Independent yardsticks for the same `bytes`: kind 0 the grid-stride copy above,
 * 1 hipMemcpyDtoDAsync, 2 a one-pass copy with four 16-byte non-temporal loads
 * in flight per lane before its stores, 3 the same with ordinary loads and
 * stores, 4 read only (bytes read), 5 write only (bytes written), 6 twenty
 * structure-of-arrays streams of bytes/20 copied at once, one double per lane,
 * non-temporal loads and stores (the step kernel's access pattern). */
int drp_copy_bandwidth_kind(size_t bytes, int kind, int reps, double* gbs);

/* The same for `streams` (10, 20 or 40) structure-of-arrays streams of
 * bytes_per_stream each, copied at once with non-temporal stores: the step
 * kernel's access pattern without its arithmetic. */
int drp_copy_bandwidth_soa(size_t bytes_per_stream, int streams, int reps, double* gbs);

/* Droplet census (program/percolation.f90:31-349, the offline Hoshen-Kopelman
   droplet count of a field file; there it reads phi, labels clusters of
   phi >= (phi1 + phi2) / 2 and writes their number).  Here on the current
   state on the GPU: clusters of nodes with phi >= threshold connected through
   face neighbours, periodic in x, y and z.  count: number of clusters;
   volumes[k] (nodes) and first[k] (smallest node index x + nx (y + ny z) in the
   cluster) for k < min(count, cap), ordered by first; labels (nullable,
   n = nodes): per node the first index of its cluster, -1 outside.  The
   reference's scan also links some diagonal (edge) neighbours and follows
   periodic wraps only partly; on clusters that neither touch diagonally nor
   cross a wrap the counts agree.  Needs fewer than 2^30 nodes. */
int drp_droplet_census(drp_lattice* lat, double threshold, int64_t* count, int64_t* volumes, int64_t* first,
                       int64_t cap, int32_t* labels, size_t n);

/* ---- FFT diagnostics of the velocity field (extents products of 2, 3 and 5, at
 * most 2048 per axis).  u_hat = FFT(u) / (nx ny nz); signed wavenumbers
 * k = i - n for i >= n/2 + 1; shell s holds the modes with int(|k|) = s. ---- */

/* Energy spectrum = enesupe_cal (program/turbulence_hit_xy.f90:588-644):
 * E[s] = sum over the half space k3 >= 0 (index 0..nz/2) of
 * 0.5 (|u1_hat|^2 + |u2_hat|^2 + |u3_hat|^2), weighted 2 off the k3 = 0 plane
 * (the Nyquist plane k3 = nz/2 included, as the reference counts it), for
 * s = 0 .. nshells-1 (the reference's enesupe(s+1), an array of nx/2 + 1;
 * modes of higher shells, which it writes past that array, are dropped here). */
int drp_energy_spectrum(drp_lattice* lat, double* E, int nshells);

/* Spectral filters of the velocity (results to host, x fastest):
 * DRP_FILTER_LOW_PASS keeps the shells int|k| < k (low_pass with k_low = k,
 * program/LES_drop_xy.f90:851-971); DRP_FILTER_SHELL keeps int|k| == k (the
 * band-pass of scale_cal2, program/drop_analysis.f90:722-797). */
#define DRP_FILTER_LOW_PASS 0
#define DRP_FILTER_SHELL 1
int drp_filter_velocity(drp_lattice* lat, int kind, int k, double* u1, double* u2, double* u3, size_t n);

/* Per-scale interface-velocity interaction (program/drop_analysis.f90:1121-1265,
 * contribution_each_scale): C[s] = kappa_g < u^(s) . w >, the mean over the nodes
 * of the shell-s band-pass of u dotted with w_b = sum_a d_a (d_b phi d_a phi)
 * (D3Q15 gradients), for s = 0 .. nshells-1 (the reference evaluates s = 1 ..
 * nx/2 + 1).  Evaluated by Parseval from one transform of u and of w. */
int drp_scale_interaction(drp_lattice* lat, double* C, int nshells);

/* The 3-D FFT these use, on a host array of nx*ny*nz complex numbers stored as
 * interleaved (re, im), x fastest; sign -1 forward, +1 inverse, unnormalised.
 * No reference counterpart (FFTW's role there); for checking the transform. */
int drp_fft3d(int nx, int ny, int nz, double* data, int sign);

/* Droplet size distribution of the current state: the census above (threshold),
 * radii r = (3 V / (4 * 3.14))^(1/3) from the cluster volumes as
 * program/percolation.f90 (:298) forms them, diameters d = 2 r / scale, those
 * with d > d_max dropped, and the histogram of program/mkdsd.f90 (:68-100):
 * counts[i] = number of droplets with int(d / bin) == i, i < nbins; *total =
 * number counted (mkdsd's frequency is counts / total). */
int drp_droplet_size_histogram(drp_lattice* lat, double threshold, double scale, double bin, double d_max,
                               int64_t* counts, int64_t nbins, int64_t* total);

/* Derived fields of the droplet analysis programs, for the current state (results
 * to host, x fastest; FP32 lattices through an FP64 copy of their fields):
 *   DRP_DERIVED_CHEMICAL_POTENTIAL  mu = T ln(phi/(1 - b phi)) + T/(1 - b phi) - 2 a phi
 *                                        - kappa_g lap phi   (chemical_potencial_cal,
 *                                        program/drop_analysis.f90:413-425; lap_cal :383-398)
 *   DRP_DERIVED_FREE_ENERGY         psi = phi T ln(phi/(1 - b phi)) - a phi^2
 *                                        (free_energy_cal, :400-411)
 *   DRP_DERIVED_ADVECTION           u . grad phi (advection_cal, :427-441; grad_cal :362-381) */
#define DRP_DERIVED_CHEMICAL_POTENTIAL 0
#define DRP_DERIVED_FREE_ENERGY 1
#define DRP_DERIVED_ADVECTION 2
int drp_get_derived(drp_lattice* lat, int kind, double* out, size_t n);

/* Cross spectrum of the phi equation (contribution, program/drop_analysis.f90:476-578):
 * with A = FFT(u . grad phi) and M = FFT(mu) (unnormalised, the chemical potential
 * above), C[s] = sum of w Re(A conj(M)) over the half space k3 = 0 .. nz/2 divided by
 * the sum of w over the same modes, for the shells int|k| == s (signed wavenumbers as
 * the reference's judge; w = 1 on k3 = 0, else 2), s = 0 .. nshells-1 with nshells <=
 * nx/2 + 1 (the reference's array; modes of higher shells, which it writes past that
 * array, are dropped).  bands[4] = the reference's energy_all, energy_large,
 * energy_medium, energy_small: minus the sums of C over s = 1 .. nx/2, 1..3, 4..9 and
 * 10 .. nx/2 (bounded by nx/2).  FP64 lattices or FP32 through an FP64 copy. */
int drp_phi_contribution(drp_lattice* lat, double* C, int nshells, double* bands);

/* One step (as drp_step with nsteps = 1), then the kinetic-energy budget of that step
 * (program/energy_budget.f90:1251-1700), summed over the nodes, in terms[7]:
 *   [0] left1  sum |p - (p0 - kappa_g phi lap phi - kappa_g/6 |grad phi|^2)|
 *   [1] left2  sum u . grad K, K = |u|^2/2
 *   [2] right1 -sum u . grad p
 *   [3] right2 sum 2 nu sum_ab d_b(u_a S_ab)
 *   [4] right3 -sum 2 nu S:S
 *   [5] right4 -(2/9) kappa_g sum_ab u_a d_b G_ab, G = 4.5 grad phi grad phi - 1.5 |grad phi|^2 I
 *   [6] right5 sum u . F  (the body force of the step)
 * with, as the reference evaluates them, p0, lap phi, grad phi, the strain S and G of
 * the fields the step started from, and p, phi, u and nu(phi) (with the new phi
 * range) of the fields it produced; D3Q15 stencils over 10 (gradients) and 5
 * (Laplacian).  Two departures from the reference: K is its commented-out definition
 * (its active code never assigns K, so its left2 is always 0), and right3 is each
 * step's own term (the reference never resets its array, so its right3 accumulates
 * over the run).  right2 is evaluated as sum 2 nu sum_b d_b (sum_a u_a S_ab), equal to
 * the reference's form but for rounding.  FP64, free-energy model; allocates 7
 * doubles per node on the first call. */
int drp_step_energy_budget(drp_lattice* lat, double* terms);

/* Message for the calling thread's last failure ("" if none). */
const char* drp_last_error(void);

void drp_destroy(drp_lattice* lat);

#ifdef __cplusplus
}
#endif
#endif /* DROPLET_AMD_H */
