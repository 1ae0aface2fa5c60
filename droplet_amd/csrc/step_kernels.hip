// step_kernels.hip — the fused D3Q15 two-phase time step for gfx950 (FP64 path).
//
// One launch = one reference time step (program/lbm_drop_xy.f90:830-840 plus
// bounce_back_LBM of prototype_program/two-phase/dropch.f90:1097):
//
//   phase 1  stage phi over the tile + 2-node halo and p, u (and, in LES mode,
//            tau) over the tile + 1-node halo in LDS (periodic wrap replaces
//            glue/MPI_boundary, :287-384)
//   phase 2  for every tile + 1-halo node: lap phi, grad phi (various_cal
//            :590-640), the F-bracket of feq and tau_g (physics :576-587)
//   phase 3  per output node, per direction q: equilibria at the upstream node
//            (equilibrium_cal :421-476, gphi_cal :642-666), pull + BGK collision
//            + body force (fg_cal :478-515), moving-wall bounce-back, moments
//            (physics :541-553), and the workgroup's part of the global phi
//            range (physics :557-574) folded into three rotating slots with
//            order-preserving 64-bit atomics.
//
// Memory traffic per node and step is the compulsory one: each population read
// once and written once, the five macroscopic fields written once and read once
// (halo re-reads hit L2).  Populations are SoA [q][z][y][x]; consecutive lanes
// own consecutive x, so every global access is coalesced.
#include "d3q15.h"
#include "kernels.h"

namespace drp {

template <int Q0, int Q1, typename Fn> __device__ __forceinline__ void static_for(Fn&& fn) {
    if constexpr (Q0 < Q1) {
        fn(std::integral_constant<int, Q0>{});
        static_for<Q0 + 1, Q1>(fn);
    }
}

// periodic wrap of a staged halo coordinate.  When every grid extent exceeds
// the tile extent the coordinate lies in [-n, 2n) and one conditional add or
// subtract suffices; otherwise (tiny grids) the launcher sets wrap_mod.
__device__ __forceinline__ int wrapi(int v, int n, bool mod) {
    if (mod) {
        v %= n;
        return v < 0 ? v + n : v;
    }
    return v < 0 ? v + n : (v >= n ? v - n : v);
}


// tau_g: constant; from nu(phi); read from a field (strain Smagorinsky, computed
// after each step); from nu(phi) and the stored sqrt(Pi:Pi) of the node's
// non-equilibrium flux (non-equilibrium Smagorinsky, program/LES_2.f90:713-733)
enum { TAU_CONST = 0, TAU_PHI = 1, TAU_FIELD = 2, TAU_NONEQ = 3 };
// interface models: the free-energy model carries phi by populations; the
// conservative Allen-Cahn model (prototype_program/two-phase/dropac.f90) has
// none and advances phi by a stencil of local equilibria
// MODEL_SCALE: the free-energy model whose phi equilibria carry a band-passed
// velocity (scale-separated advection, StepArgs::uadv)
enum { MODEL_FE = 0, MODEL_AC = 1, MODEL_SCALE = 2 };

// Resident wavefronts per SIMD the register allocation targets: two 8-wavefront
// workgroups per CU fit the LDS budget, i.e. 4 per SIMD (<= 128 VGPRs).  The
// rare wall + body-force combination would spill at 128 and gets 2.
template <bool WALL, int FORCE> constexpr int waves_per_eu() { return (WALL && FORCE != 0) ? 2 : 4; }

// Two threads per lattice node: the first NN threads of the workgroup (whole
// wavefronts) update the node's 15 phi populations, the second NN threads its
// 15 momentum populations.  Halving the per-thread state doubles the resident
// wavefronts per CU, and the shared staging / stencil phases are spread over
// twice the threads.
// workgroups of `bytes` of LDS that fit in one CU (160 KiB, 1280-byte granules)
constexpr int wg_per_cu_lds(int bytes) { return (160 * 1024) / (((bytes + 1279) / 1280) * 1280); }

template <int TX, int TY, int TZ, bool WALL, int FORCE, int TAUMODE, int MODEL = MODEL_FE>
__device__ __forceinline__ void step_tile_fp64(const StepArgs& a, int tile) {
    constexpr int NN = TX * TY * TZ, NT = 2 * NN;
    constexpr int HX2 = TX + 4, HY2 = TY + 4, HZ2 = TZ + 4, H2 = HX2 * HY2 * HZ2;
    constexpr int HX1 = TX + 2, HY1 = TY + 2, HZ1 = TZ + 2, H1 = HX1 * HY1 * HZ1;
    static_assert(NN % 64 == 0, "a role must own whole wavefronts");
    // relaxation by the shared-divisor division (d3q15.h) where it costs no
    // spilled registers: not in a role that carries a register ring (the
    // band-passed velocity of the phi role, the force field of the momentum role)
    constexpr bool FASTDIV_F = MODEL != MODEL_SCALE;
    constexpr bool FASTDIV_G = FORCE != FORCE_FIELD;

    __shared__ double s_phi[H2];
    __shared__ double s_p3[H1], s_u1[H1], s_u2[H1], s_u3[H1];  // s_p3: 3 p
    __shared__ double s_A[H1], s_g0[H1], s_g1[H1], s_g2[H1];
    // per-node 1.5 |grad phi|^2 (PRE_S) and 1.5 |u|^2 (PRE_U), formed once with the
    // same operations the per-direction equilibria would repeat (bit-identical),
    // each only where it does not lower the number of workgroups that fit in a
    // CU's 160 KiB of LDS (sizes rounded up to a 1280-byte allocation granule)
    // Allen-Cahn: the unit normal n = grad phi / (|grad phi| + eps) per halo node
    // in s_A (n0), s_n1, s_n2 (the F-bracket has no role in that model)
    __shared__ double s_n1[MODEL == MODEL_AC ? H1 : 1], s_n2[MODEL == MODEL_AC ? H1 : 1];
    constexpr int LDS_BASE = 8 * (H2 + 8 * H1 + (TAUMODE == TAU_CONST ? 1 : H1) + 2 * (NT / 64) +
                                  (MODEL == MODEL_AC ? 2 * H1 : 0));
    constexpr bool PRE_S = LDS_BASE + 8 * H1 <= 160 * 1024 && wg_per_cu_lds(LDS_BASE + 8 * H1) >= wg_per_cu_lds(LDS_BASE);
    constexpr bool PRE_U =
        PRE_S && LDS_BASE + 16 * H1 <= 160 * 1024 && wg_per_cu_lds(LDS_BASE + 16 * H1) >= wg_per_cu_lds(LDS_BASE);
    __shared__ double s_s2h[PRE_S ? H1 : 1], s_usq15[PRE_U ? H1 : 1];  // 1.5 |grad phi|^2, 1.5 |u|^2
    __shared__ double s_tau[TAUMODE == TAU_CONST ? 1 : H1];
    __shared__ double s_red[2][NT / 64];

    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t N = a.n;
    // tiles run x fastest, then z, then y: the workgroups resident at one time
    // on an XCD (a contiguous range of tiles) cover consecutive z-layers of one
    // tile row, so the z halos (the deeper ones: TZ <= TY) are read from L2
    const int x0 = (tile % a.tiles_x) * TX;
    const int z0 = ((tile / a.tiles_x) % a.tiles_z) * TZ;
    const int y0 = (tile / (a.tiles_x * a.tiles_z)) * TY;
    const int tid = threadIdx.x;
    // role: momentum populations; uniform per wavefront (NN is a multiple of 64)
    const bool mom = tid >= NN;
    const int nid = mom ? tid - NN : tid;
    __builtin_assume(nid >= 0 && nid < NN);

    const double* __restrict__ phi_in = a.mac_in;
    const double* __restrict__ p_in = a.mac_in + N;
    const double* __restrict__ u1_in = a.mac_in + 2 * N;
    const double* __restrict__ u2_in = a.mac_in + 3 * N;
    const double* __restrict__ u3_in = a.mac_in + 4 * N;

    // ---------------- phase 1a: load the macroscopic fields of the tile and its
    // halos into registers.  Thread t stages column t % HX of a halo box, in
    // rows t / HX + k (NT / HX) (row = y + HY z): its x coordinate and wrap are
    // fixed, and the rows of a tile whose halo box lies inside the grid in y and
    // z (every tile but those of the first and last layers) need no wrap
    const bool wm = a.wrap_mod != 0;
    constexpr int RP2 = NT / HX2, R2 = HY2 * HZ2, IT2 = (R2 + RP2 - 1) / RP2;
    constexpr int RP1 = NT / HX1, R1 = HY1 * HZ1, IT1 = (R1 + RP1 - 1) / RP1;
    static_assert(RP2 >= 1 && RP1 >= 1, "a workgroup spans at least one halo row");
    const int c2 = tid % HX2, r2 = tid / HX2;  // r2 < RP2: the thread stages phi rows
    const int c1 = tid % HX1, r1 = tid / HX1;  // r1 < RP1: the thread stages p, u rows
    const bool yz_in = !wm && y0 >= 2 && y0 + TY + 2 <= ny && z0 >= 2 && z0 + TZ + 2 <= nz;
    // node index of (0, y, z): z nx ny in 64 bits from 32-bit factors (an x-y
    // plane holds fewer than 2^31 nodes)
    const uint32_t nxy32 = (uint32_t)nx * (uint32_t)ny;
    auto row_offset = [&](int hy, int hz, int halo) -> int64_t {
        int gy = y0 - halo + hy, gz = z0 - halo + hz;
        if (!yz_in) {
            gy = wrapi(gy, ny, wm);
            gz = wrapi(gz, nz, wm);
        }
        return (int64_t)((uint64_t)(uint32_t)gz * nxy32 + (uint32_t)(gy * nx));
    };
    const int gx2 = wrapi(x0 - 2 + c2, nx, wm), gx1 = wrapi(x0 - 1 + c1, nx, wm);
    double vphi[IT2], vp[IT1], vu1[IT1], vu2[IT1], vu3[IT1], vtau[IT1];
    static_for<0, IT2>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        // loads are unconditional (rows clamped into the halo box) so that the
        // compiler's wait for them is a straight-line count
        const int r = r2 + k * RP2 < R2 ? r2 + k * RP2 : R2 - 1;
        vphi[k] = phi_in[gx2 + row_offset(r % HY2, r / HY2, 2)];
    });
    static_for<0, IT1>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const int r = r1 + k * RP1 < R1 ? r1 + k * RP1 : R1 - 1;
        const int64_t g = gx1 + row_offset(r % HY1, r / HY1, 1);
        vp[k] = p_in[g];
        vu1[k] = u1_in[g];
        vu2[k] = u2_in[g];
        vu3[k] = u3_in[g];
        if constexpr (TAUMODE == TAU_FIELD) vtau[k] = a.tau_field[g];
        if constexpr (TAUMODE == TAU_NONEQ) vtau[k] = a.qh_in[g];
    });

    // ---------------- phase 1b: issue the 15 upstream population loads of this
    // thread's set after the field loads, so that the wait for the fields does
    // not include them and their latency overlaps staging and stencils; they
    // are read once, non-temporally, to leave L2 to the macroscopic halos
    const int tx = nid % TX, ty = (nid / TX) % TY, tz = nid / (TX * TY);
    // LDS indices of the corner (x - 1, y - 1, z - 1) in the 1-halo arrays and of
    // (x - 2, y - 2, z - 2) in the 2-halo phi array: the upstream node of every
    // direction lies at a non-negative constant offset from them, which the LDS
    // reads take as their immediate offset
    const int hb1 = tx + HX1 * (ty + HY1 * tz);
    const int hb2 = tx + HX2 * (ty + HY2 * tz);
    const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
    const bool inside = (x < nx) && (y < ny) && (z < nz);
    // upstream coordinate x - c is entry [c + 1] of {x+1, x, x-1} (periodic wrap)
    const int xc = inside ? x : 0, yc = inside ? y : 0, zc = inside ? z : 0;
    const int xs[3] = {xc + 1 == nx ? 0 : xc + 1, xc, xc == 0 ? nx - 1 : xc - 1};
    const int zs[3] = {zc + 1 == nz ? 0 : zc + 1, zc, zc == 0 ? nz - 1 : zc - 1};
    const uint32_t yo[3] = {(uint32_t)(nx * (yc + 1 == ny ? 0 : yc + 1)), (uint32_t)(nx * yc),
                            (uint32_t)(nx * (yc == 0 ? ny - 1 : yc - 1))};
    const uint64_t zo[3] = {(uint64_t)(uint32_t)zs[0] * nxy32, (uint64_t)(uint32_t)zs[1] * nxy32,
                            (uint64_t)(uint32_t)zs[2] * nxy32};
    // node index of the upstream node x - c of direction q
    auto upstream = [&](auto qc) -> int64_t {
        using D = Dir<decltype(qc)::value>;
        return (int64_t)(zo[1 + D::cz] + (uint64_t)(yo[1 + D::cy] + (uint32_t)xs[1 + D::cx]));
    };
    // the role's population array, as a scalar (the role is uniform per wavefront)
    const double* __restrict__ pin = (__builtin_amdgcn_readfirstlane(mom ? 1 : 0) != 0) ? a.g_in : a.f_in;
    // unconditional: threads outside the grid read node (0, 0, 0)'s neighbours
    // (the Allen-Cahn model has no phi populations to load)
    double ps[15];
    auto load_pops = [&]() {
        static_for<0, 15>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            using D = Dir<q>;
            ps[q] = __builtin_nontemporal_load(&pin[q * N + upstream(qc)]);
        });
    };
    if (MODEL != MODEL_AC || mom) load_pops();

    // ---------------- phase 1c: stage the fields in LDS
    static_for<0, IT2>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const int r = r2 + k * RP2;
        if (r2 < RP2 && r < R2) s_phi[c2 + HX2 * r] = vphi[k];
    });
    static_for<0, IT1>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const int r = r1 + k * RP1;
        if (r1 < RP1 && r < R1) {
            const int h = c1 + HX1 * r;
            s_p3[h] = 3.0 * vp[k];
            s_u1[h] = vu1[k];
            s_u2[h] = vu2[k];
            s_u3[h] = vu3[k];
            if constexpr (TAUMODE == TAU_FIELD || TAUMODE == TAU_NONEQ) s_tau[h] = vtau[k];
        }
    });
    __syncthreads();

    // ---------------- phase 2: stencils, F-bracket and tau at tile + 1 halo
    const double rmax = from_ordered_bits(a.range[2 * a.slot_read + 0]);
    const double rmin = from_ordered_bits(a.range[2 * a.slot_read + 1]);
    const double range = rmax - rmin;
    static_for<0, IT1>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const int r = r1 + k * RP1;
        if (r1 >= RP1 || r >= R1) return;
        const int h = c1 + HX1 * r;
        // phi-box index of the node's (-1, -1, -1) neighbour: the stencil reads lie
        // at non-negative constant offsets from it
        const int cb = c1 + HX2 * (r % HY1 + HY2 * (r / HY1));
        const double ph = s_phi[cb + (1 + HX2 * (1 + HY2))];
        double lap = -14.0 * ph;
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        static_for<1, 15>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            using D = Dir<q>;
            const double v = s_phi[cb + ((1 + D::cx) + HX2 * ((1 + D::cy) + HY2 * (1 + D::cz)))];
            lap = lap + v;
            g0 = acc<D::cx>(g0, v);
            g1 = acc<D::cy>(g1, v);
            g2 = acc<D::cz>(g2, v);
        });
        g0 = div_const(g0, 10.0, 0.1);
        g1 = div_const(g1, 10.0, 0.1);
        g2 = div_const(g2, 10.0, 0.1);
        if constexpr (MODEL == MODEL_AC) {
            // n_vector_cal, dropac.f90:715-728 (eps = 1e-12, :67)
            const double nrm = sqrt(g0 * g0 + g1 * g1 + g2 * g2) + 1.0e-12;
            s_A[h] = g0 / nrm;
            s_n1[h] = g1 / nrm;
            s_n2[h] = g2 / nrm;
        } else {
            lap = div_const(lap, 5.0, 0.2);
            const double p0 = ph * a.T / (1.0 - a.b * ph) - a.a * (ph * ph);
            s_A[h] = p0 - a.kf * ph * lap - a.kf6 * (g0 * g0 + g1 * g1 + g2 * g2);
        }
        s_g0[h] = g0;
        if constexpr (PRE_S) s_s2h[h] = 1.5 * (g0 * g0 + g1 * g1 + g2 * g2);
        if constexpr (PRE_U) s_usq15[h] = 1.5 * (s_u1[h] * s_u1[h] + s_u2[h] * s_u2[h] + s_u3[h] * s_u3[h]);
        s_g1[h] = g1;
        s_g2[h] = g2;
        if constexpr (TAUMODE == TAU_PHI) {
            double frac = (ph - rmin) / range;
            if (range == 0.0) frac = 0.0;
            const double nu = frac * a.dnu + a.nu1;
            s_tau[h] = 0.5 + 3.0 * nu;
        }
        if constexpr (TAUMODE == TAU_NONEQ) {
            // LES_2.f90:722-730 with sqrt(Q) of the node stored by the last step:
            // S = (sqrt(nu^2 + 18 Cs^2 D^2 sqrt(Q)) - nu) / (6 Cs^2 D^2), nu_e = Cs^2 D^2 S
            double frac = (ph - rmin) / range;
            if (range == 0.0) frac = 0.0;
            const double nu = frac * a.dnu + a.nu1;
            double S = sqrt(nu * nu + a.neq_c18 * s_tau[h]) - nu;
            S = S / a.neq_c6;
            const double nue = a.neq_cc * S;
            s_tau[h] = 0.5 + 3.0 * (nu + nue);
        }
    });
    __syncthreads();

    // ---------------- phase 3: collide-stream of this thread's population set
    double vmax = -__builtin_huge_val(), vmin = __builtin_huge_val();
    if (inside) {
        const int64_t s = x + (int64_t)nx * (y + (int64_t)ny * z);
        if (!mom && MODEL == MODEL_AC) {
            // physics, dropac.f90:640-667: phi(x) <- sum_i [feq_i(x - c_i) + A E_i (phi(x) - phi(x - c_i))]
            // with feq_i = E_i phi (1 + 3 c.u) + 3 E_i M (1 - 4 phi^2) (c.n) / W (feq_cal :517-535),
            // all at the step's start, in the reference's order of operations
            const double phx = s_phi[(tx + 2) + HX2 * ((ty + 2) + HY2 * (tz + 2))];
            double ph = 0.0;
            static_for<0, 15>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                using D = Dir<q>;
                constexpr double E = weightE<q>();
                const int hu = hb1 + ((1 - D::cx) + HX1 * ((1 - D::cy) + HY1 * (1 - D::cz)));
                const int h2 = hb2 + ((2 - D::cx) + HX2 * ((2 - D::cy) + HY2 * (2 - D::cz)));
                const double phu = s_phi[h2];
                double fe = E * phu;
                if constexpr (q != 0) {
                    const double cu = cdot<q>(s_u1[hu], s_u2[hu], s_u3[hu]);
                    const double cn = cdot<q>(s_A[hu], s_n1[hu], s_n2[hu]);
                    fe = fe * (1.0 + 3.0 * cu) + (3.0 * E) * a.ac_m * (1.0 - 4.0 * (phu * phu)) * cn / a.ac_w;
                }
                ph = ph + fe + (a.ac_a * E) * (phx - phu);
                __builtin_amdgcn_sched_barrier(0);
            });
            __builtin_nontemporal_store(ph, &a.mac_out[s]);
            vmax = ph;
            vmin = ph;
        } else if (!mom) {
            // relaxation by 1/tau_f with the divisor's steps done once (SharedDivisor);
            // a node one of whose f - feq lies outside the range where that is exact
            // runs its collision again with full divisions
            const SharedDivisor dtau = shared_divisor(a.tau_f);
            bool redo = !dtau.ok;
            double ph = 0.0;
            // MODEL_SCALE: the band-passed velocity of the upstream nodes (equilibrium_cal,
            // program/lbm_drop_scale_xy.f90:537-540), loaded from L2 UPF directions ahead
            // of its use into a ring of UPF slots
            constexpr int UPF = 3;
            double uring[UPF][3] = {};
            auto load_uadv = [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                using D = Dir<q>;
                const int64_t src = upstream(qc);
                double* u = uring[q % UPF];
                if constexpr (D::cx != 0) u[0] = a.uadv[src];
                if constexpr (D::cy != 0) u[1] = a.uadv[N + src];
                if constexpr (D::cz != 0) u[2] = a.uadv[2 * N + src];
            };
            auto collide = [&](auto full_c) {
                constexpr bool FULL = decltype(full_c)::value;
                if constexpr (MODEL == MODEL_SCALE)
                    static_for<1, 1 + UPF>([&](auto qc) { load_uadv(qc); });
                static_for<0, 15>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    using D = Dir<q>;
                    const int hu = hb1 + ((1 - D::cx) + HX1 * ((1 - D::cy) + HY1 * (1 - D::cz)));
                    const int h2 = hb2 + ((2 - D::cx) + HX2 * ((2 - D::cy) + HY2 * (2 - D::cz)));
                    NodeEq<double> nd;
                    nd.phi = s_phi[h2];
                    nd.A = s_A[hu];
                    double gt = 0.0;
                    if constexpr (q != 0) {
                        if constexpr (MODEL == MODEL_SCALE) {
                            // only the components with c_a != 0 enter c . u
                            const double* u = uring[q % UPF];
                            nd.u1 = u[0];
                            nd.u2 = u[1];
                            nd.u3 = u[2];
                            if constexpr (q + UPF < 15) load_uadv(std::integral_constant<int, q + UPF>{});
                        } else {
                            nd.u1 = s_u1[hu];
                            nd.u2 = s_u2[hu];
                            nd.u3 = s_u3[hu];
                        }
                        nd.g0 = s_g0[hu];
                        nd.g1 = s_g1[hu];
                        nd.g2 = s_g2[hu];
                        const double s2h = PRE_S ? s_s2h[hu] : 1.5 * (nd.g0 * nd.g0 + nd.g1 * nd.g1 + nd.g2 * nd.g2);
                        gt = korteweg<q>(nd.g0, nd.g1, nd.g2, s2h);
                    }
                    const double fe = feq<q>(nd, gt, a.ekf[wclass<q>()]);
                    const double dx = ps[q] - fe;
                    double fv;
                    if constexpr (FULL) {
                        fv = ps[q] - dx / a.tau_f;
                    } else {
                        fv = ps[q] - div_shared_fast(dx, dtau);
                        redo |= !div_shared_exact(dx);
                    }
                    ph = (q == 0) ? fv : ph + fv;
                    __builtin_nontemporal_store(fv, &a.f_out[q * N + s]);
                    __builtin_amdgcn_sched_barrier(0);  // keep each direction's LDS reads local
                });
            };
            if constexpr (FASTDIV_F) {
                collide(std::false_type{});
                if (redo) collide(std::true_type{});
            } else {
                collide(std::true_type{});
            }
            __builtin_nontemporal_store(ph, &a.mac_out[s]);
            vmax = ph;
            vmin = ph;
        } else {
            // force-field components of the upstream nodes, loaded FPF directions
            // ahead of their use into a ring of FPF slots, so that the L2 latency
            // of these loads overlaps the collisions in between
            constexpr int FPF = 3;
            double fring[FPF][3] = {};
            auto load_force = [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                using D = Dir<q>;
                const int64_t src = upstream(qc);
                double* f = fring[q % FPF];
                if constexpr (D::cx != 0) f[0] = a.fx[src];
                if constexpr (D::cy != 0) f[1] = a.fy[src];
                if constexpr (D::cz != 0) f[2] = a.fz[src];
            };
            // with a constant tau_g, relaxation as in the phi role (SharedDivisor, and
            // the collision again with full divisions for a node that needs them)
            const SharedDivisor dtau = shared_divisor(a.tau_const);
            bool redo = TAUMODE == TAU_CONST && !dtau.ok;
            auto collide = [&](auto full_c) {
                constexpr bool FULL = decltype(full_c)::value;
                if constexpr (FORCE == FORCE_FIELD) static_for<1, 1 + FPF>(load_force);
                static_for<0, 15>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    using D = Dir<q>;
                    const int hu = hb1 + ((1 - D::cx) + HX1 * ((1 - D::cy) + HY1 * (1 - D::cz)));
                    const double u1 = s_u1[hu], u2 = s_u2[hu], u3 = s_u3[hu];
                    const double usq15 = PRE_U ? s_usq15[hu] : 1.5 * (u1 * u1 + u2 * u2 + u3 * u3);
                    double gt = 0.0;
                    if constexpr (q != 0) {
                        const double g0 = s_g0[hu], g1 = s_g1[hu], g2 = s_g2[hu];
                        const double s2h = PRE_S ? s_s2h[hu] : 1.5 * (g0 * g0 + g1 * g1 + g2 * g2);
                        gt = korteweg<q>(g0, g1, g2, s2h);
                    }
                    const double ge = geq_pre<q>(s_p3[hu], u1, u2, u3, usq15, gt, a.ekg[wclass<q>()]);
                    const double dx = ps[q] - ge;
                    double gv;
                    if constexpr (TAUMODE != TAU_CONST) {
                        gv = ps[q] - dx / s_tau[hu];
                    } else if constexpr (FULL) {
                        gv = ps[q] - dx / a.tau_const;
                    } else {
                        gv = ps[q] - div_shared_fast(dx, dtau);
                        redo |= !div_shared_exact(dx);
                    }
                    if constexpr (FORCE != FORCE_NONE && q != 0) {
                        // c . F with the zero-velocity terms dropped, as for c . u: the
                        // reference's (cx Fx + cy Fy) + cz Fz rounds identically, and
                        // only the components that enter are loaded
                        double Fx = 0.0, Fy = 0.0, Fz = 0.0;
                        if constexpr (FORCE == FORCE_TAYLOR_GREEN) {
                            const int xu = xs[1 + D::cx], zu = zs[1 + D::cz];
                            if constexpr (D::cx != 0) Fx = a.tg_amp * a.sinx[xu] * a.cosz[zu];
                            if constexpr (D::cz != 0) Fz = -(a.tg_amp * a.cosx[xu] * a.sinz[zu]);
                        } else {  // field at the upstream node, loaded FPF directions ahead
                            const double* f = fring[q % FPF];
                            if constexpr (D::cx != 0) Fx = f[0];
                            if constexpr (D::cy != 0) Fy = f[1];
                            if constexpr (D::cz != 0) Fz = f[2];
                            if constexpr (q + FPF < 15) load_force(std::integral_constant<int, q + FPF>{});
                        }
                        const double dot = cdot<q>(Fx, Fy, Fz);
                        gv = gv + 3.0 * weightE<q>() * dot;
                    }
                    ps[q] = gv;
                    __builtin_amdgcn_sched_barrier(0);  // keep each direction's LDS reads local
                });
            };
            collide(std::integral_constant<bool, !FASTDIV_G>{});
            if constexpr (FASTDIV_G && TAUMODE == TAU_CONST) {
                if (redo) {
                    // the populations this step started from, again, addressed from
                    // the node's coordinates (so that phase 1's upstream offsets
                    // need not stay in registers until here)
                    static_for<0, 15>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        using D = Dir<q>;
                        const int ux = wrapi(x - D::cx, nx, false), uy = wrapi(y - D::cy, ny, false),
                                  uz = wrapi(z - D::cz, nz, false);
                        ps[q] = __builtin_nontemporal_load(&a.g_in[q * N + ux + (int64_t)nx * (uy + (int64_t)ny * uz)]);
                    });
                    collide(std::true_type{});
                }
            }

            if constexpr (WALL) {
                // bounce_back_LBM, dropch.f90:666-698: on the first/last fluid layer
                // the populations entering the fluid are the reflected ones minus
                // 6 E (U_wall . c_opp); bottom wall moves with -U, top with +U.
                const int coord = (a.wall_axis == 2) ? y : z;
                const int nax = (a.wall_axis == 2) ? ny : nz;
                const bool bottom = coord == 0, top = coord == nax - 1;
                if (bottom || top) {
                    const double uw = bottom ? -a.wall_u_now : a.wall_u_now;
                    const int want = bottom ? 1 : -1;
                    static_for<1, 15>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        using D = Dir<q>;
                        constexpr int o = D::opp;
                        const int ca = (a.wall_axis == 2) ? D::cy : D::cz;
                        if (ca == want) {
                            const double dot = uw * (double)Dir<o>::cx;
                            ps[q] = ps[o] - 6.0 * weightE<o>() * dot;
                        }
                    });
                }
            }

            if constexpr (TAUMODE == TAU_NONEQ) {
                // tensor_cal (LES_2.f90:530-551): Pi_ab = sum_i c_ia c_ib (g_i - geq_i) with
                // this step's g and the equilibria of its start at this node; Q = Pi:Pi
                // (beta outer, alpha inner); sqrt(Q) is kept for the next step's tau
                const int hs = (tx + 1) + HX1 * ((ty + 1) + HY1 * (tz + 1));
                const double p3 = s_p3[hs], u1 = s_u1[hs], u2 = s_u2[hs], u3 = s_u3[hs];
                const double usq15 = PRE_U ? s_usq15[hs] : 1.5 * (u1 * u1 + u2 * u2 + u3 * u3);
                const double g0 = s_g0[hs], g1 = s_g1[hs], g2 = s_g2[hs];
                const double s2h = PRE_S ? s_s2h[hs] : 1.5 * (g0 * g0 + g1 * g1 + g2 * g2);
                double t[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
                static_for<0, 15>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    using D = Dir<q>;
                    constexpr int c[3] = {D::cx, D::cy, D::cz};
                    const double d = ps[q] - geq_pre<q>(p3, u1, u2, u3, usq15, korteweg<q>(g0, g1, g2, s2h), a.ekg[wclass<q>()]);
#pragma unroll
                    for (int b = 0; b < 3; b++)
#pragma unroll
                        for (int al = 0; al < 3; al++) {
                            const int cc = c[al] * c[b];
                            if (cc == 1) t[al][b] = t[al][b] + d;
                            else if (cc == -1) t[al][b] = t[al][b] - d;
                        }
                });
                double Q = 0.0;
#pragma unroll
                for (int b = 0; b < 3; b++)
#pragma unroll
                    for (int al = 0; al < 3; al++) Q = Q + t[al][b] * t[al][b];
                a.qh_out[s] = sqrt(Q);
            }

            // moments p = sum g/3, u = sum c g (physics :546-552), order i = 1..15
            double pr = div_const(ps[0], 3.0, 1.0 / 3.0);
            double m1 = 0.0, m2 = 0.0, m3 = 0.0;
            static_for<1, 15>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                using D = Dir<q>;
                pr = pr + div_const(ps[q], 3.0, 1.0 / 3.0);
                m1 = acc<D::cx>(m1, ps[q]);
                m2 = acc<D::cy>(m2, ps[q]);
                m3 = acc<D::cz>(m3, ps[q]);
            });
            static_for<0, 15>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                __builtin_nontemporal_store(ps[q], &a.g_out[q * N + s]);
            });
            __builtin_nontemporal_store(pr, &a.mac_out[N + s]);
            __builtin_nontemporal_store(m1, &a.mac_out[2 * N + s]);
            __builtin_nontemporal_store(m2, &a.mac_out[3 * N + s]);
            __builtin_nontemporal_store(m3, &a.mac_out[4 * N + s]);
        }
    }

    // ---------------- global phi range for the next step's nu(phi)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        vmax = fmax(vmax, __shfl_xor(vmax, off, 64));
        vmin = fmin(vmin, __shfl_xor(vmin, off, 64));
    }
    const int lane = tid & 63, wave = tid >> 6;
    if (lane == 0) {
        s_red[0][wave] = vmax;
        s_red[1][wave] = vmin;
    }
    __syncthreads();
    if (tid == 0) {
        double mx = s_red[0][0], mn = s_red[1][0];
        for (int w = 1; w < NN / 64; w++) {
            mx = fmax(mx, s_red[0][w]);
            mn = fmin(mn, s_red[1][w]);
        }
        atomicMax(reinterpret_cast<unsigned long long*>(&a.range[2 * a.slot_write + 0]),
                  (unsigned long long)ordered_bits(mx));
        atomicMin(reinterpret_cast<unsigned long long*>(&a.range[2 * a.slot_write + 1]),
                  (unsigned long long)ordered_bits(mn));
        if (tile == 0) {
            a.range[2 * a.slot_reset + 0] = ordered_bits(a.phi1);
            a.range[2 * a.slot_reset + 1] = ordered_bits(a.phi2);
        }
    }
}

// one lattice: workgroup id -> tile (XCD-contiguous)
template <int TX, int TY, int TZ, bool WALL, int FORCE, int TAUMODE, int MODEL>
__global__ __launch_bounds__(2 * TX* TY* TZ) __attribute__((amdgpu_waves_per_eu(waves_per_eu<WALL, FORCE>()))) void step_fp64(StepArgs a) {
    step_tile_fp64<TX, TY, TZ, WALL, FORCE, TAUMODE, MODEL>(a, xcd_tile(blockIdx.x, gridDim.x));
}

// many independent lattices in one launch: tile_start[i] is the first global
// tile of instance i (ascending, tile_start[0] = 0); the instance of a
// workgroup is found by binary search, its arguments read from global memory
template <int TX, int TY, int TZ, bool WALL, int FORCE, int TAUMODE>
__global__ __launch_bounds__(2 * TX* TY* TZ) __attribute__((amdgpu_waves_per_eu(waves_per_eu<WALL, FORCE>()))) void step_batch_fp64(const StepArgs* __restrict__ args, const int* __restrict__ tile_start, int ninst) {
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    step_tile_fp64<TX, TY, TZ, WALL, FORCE, TAUMODE>(args[lo], b - tile_start[lo]);
}

// ------------------------------------------------------------------ launcher
template <int TX, int TY, int TZ, bool WALL, int FORCE, int TAUM, int MODEL = MODEL_FE>
static hipError_t launch_shape(const StepArgs& a, hipStream_t st, const BatchLaunch* batch) {
    if constexpr (MODEL != MODEL_FE || TAUM == TAU_NONEQ) {
        // ensembles run the free-energy model without LES only
        if (batch) return hipErrorInvalidValue;
    } else {
        if (batch) {
            hipLaunchKernelGGL((step_batch_fp64<TX, TY, TZ, WALL, FORCE, TAUM>), dim3(batch->total_tiles),
                               dim3(2 * TX * TY * TZ), 0, st, (const StepArgs*)batch->args, batch->tile_start,
                               batch->ninst);
            return hipGetLastError();
        }
    }
    StepArgs b = a;
    set_tiles(b, TX, TY, TZ);
    const int nb = b.tiles_x * b.tiles_y * b.tiles_z;
    hipLaunchKernelGGL((step_fp64<TX, TY, TZ, WALL, FORCE, TAUM, MODEL>), dim3(nb), dim3(2 * TX * TY * TZ), 0, st, b);
    return hipGetLastError();
}

template <int TX, int TY, int TZ, bool WALL, int FORCE>
static hipError_t launch_tau(const StepArgs& a, hipStream_t st, const BatchLaunch* batch) {
    if (a.tau_field) return launch_shape<TX, TY, TZ, WALL, FORCE, TAU_FIELD>(a, st, batch);
    if (a.qh_in) return launch_shape<TX, TY, TZ, WALL, FORCE, TAU_NONEQ>(a, st, batch);
    if (a.dnu == 0.0) return launch_shape<TX, TY, TZ, WALL, FORCE, TAU_CONST>(a, st, batch);
    return launch_shape<TX, TY, TZ, WALL, FORCE, TAU_PHI>(a, st, batch);
}

// LDS bytes of the Allen-Cahn variant without the optional per-node tables:
// phi over the 2-halo box, p, u, n (3), grad phi (3) over the 1-halo box, tau if variable
constexpr int ac_lds_bytes(int TX, int TY, int TZ, bool var_tau) {
    return 8 * ((TX + 4) * (TY + 4) * (TZ + 4) + (10 + (var_tau ? 1 : 0)) * (TX + 2) * (TY + 2) * (TZ + 2) +
                (var_tau ? 0 : 1) + 2 * (2 * TX * TY * TZ / 64));
}

// scale-separated advection: free-energy model without LES, any wall and body force
template <int TX, int TY, int TZ, bool WALL, int FORCE>
static hipError_t launch_scale(const StepArgs& a, hipStream_t st, const BatchLaunch* batch) {
    if (a.tau_field || a.qh_in) return hipErrorInvalidValue;
    if (a.dnu == 0.0) return launch_shape<TX, TY, TZ, WALL, FORCE, TAU_CONST, MODEL_SCALE>(a, st, batch);
    return launch_shape<TX, TY, TZ, WALL, FORCE, TAU_PHI, MODEL_SCALE>(a, st, batch);
}

template <int TX, int TY, int TZ>
static hipError_t launch_variant(const StepArgs& a, hipStream_t st, const BatchLaunch* batch) {
    const bool wall = a.wall_axis != 0;
    if (a.uadv) {
        switch (a.force_kind) {
            case FORCE_NONE:
                return wall ? launch_scale<TX, TY, TZ, true, FORCE_NONE>(a, st, batch)
                            : launch_scale<TX, TY, TZ, false, FORCE_NONE>(a, st, batch);
            case FORCE_TAYLOR_GREEN:
                return wall ? launch_scale<TX, TY, TZ, true, FORCE_TAYLOR_GREEN>(a, st, batch)
                            : launch_scale<TX, TY, TZ, false, FORCE_TAYLOR_GREEN>(a, st, batch);
            case FORCE_FIELD:
                return wall ? launch_scale<TX, TY, TZ, true, FORCE_FIELD>(a, st, batch)
                            : launch_scale<TX, TY, TZ, false, FORCE_FIELD>(a, st, batch);
            default:
                return hipErrorInvalidValue;
        }
    }
    if (a.model == MODEL_AC) {  // no body force, no LES; tile shapes whose staging fits in LDS
        if (a.force_kind != FORCE_NONE || a.tau_field) return hipErrorInvalidValue;
        if (a.dnu == 0.0) {
            if constexpr (ac_lds_bytes(TX, TY, TZ, false) <= 160 * 1024)
                return wall ? launch_shape<TX, TY, TZ, true, FORCE_NONE, TAU_CONST, MODEL_AC>(a, st, batch)
                            : launch_shape<TX, TY, TZ, false, FORCE_NONE, TAU_CONST, MODEL_AC>(a, st, batch);
            return hipErrorInvalidConfiguration;
        }
        if constexpr (ac_lds_bytes(TX, TY, TZ, true) <= 160 * 1024)
            return wall ? launch_shape<TX, TY, TZ, true, FORCE_NONE, TAU_PHI, MODEL_AC>(a, st, batch)
                        : launch_shape<TX, TY, TZ, false, FORCE_NONE, TAU_PHI, MODEL_AC>(a, st, batch);
        return hipErrorInvalidConfiguration;
    }
    switch (a.force_kind) {
        case FORCE_NONE:
            return wall ? launch_tau<TX, TY, TZ, true, FORCE_NONE>(a, st, batch) : launch_tau<TX, TY, TZ, false, FORCE_NONE>(a, st, batch);
        case FORCE_TAYLOR_GREEN:
            return wall ? launch_tau<TX, TY, TZ, true, FORCE_TAYLOR_GREEN>(a, st, batch)
                        : launch_tau<TX, TY, TZ, false, FORCE_TAYLOR_GREEN>(a, st, batch);
        case FORCE_FIELD:
            return wall ? launch_tau<TX, TY, TZ, true, FORCE_FIELD>(a, st, batch) : launch_tau<TX, TY, TZ, false, FORCE_FIELD>(a, st, batch);
        default:
            return hipErrorInvalidValue;
    }
}

// The build compiles this file once per tile shape (DRP_STEP_SHAPE = k, which
// instantiates that shape's variants as launch_step_fp64_shape<k>) and once
// without it for the dispatcher, so the shapes compile in parallel.
template <int S> hipError_t launch_step_fp64_shape(const StepArgs& a, hipStream_t st, const BatchLaunch* batch);

#if defined(DRP_STEP_SHAPE)
template <>
hipError_t launch_step_fp64_shape<DRP_STEP_SHAPE>(const StepArgs& a, hipStream_t st, const BatchLaunch* batch) {
    constexpr int S = DRP_STEP_SHAPE;
    return launch_variant<SHAPE_DIMS[S][0], SHAPE_DIMS[S][1], SHAPE_DIMS[S][2]>(a, st, batch);
}
#else
template <> hipError_t launch_step_fp64_shape<0>(const StepArgs& a, hipStream_t st, const BatchLaunch* batch);
template <> hipError_t launch_step_fp64_shape<1>(const StepArgs& a, hipStream_t st, const BatchLaunch* batch);
template <> hipError_t launch_step_fp64_shape<2>(const StepArgs& a, hipStream_t st, const BatchLaunch* batch);
template <> hipError_t launch_step_fp64_shape<3>(const StepArgs& a, hipStream_t st, const BatchLaunch* batch);
template <> hipError_t launch_step_fp64_shape<4>(const StepArgs& a, hipStream_t st, const BatchLaunch* batch);
template <> hipError_t launch_step_fp64_shape<5>(const StepArgs& a, hipStream_t st, const BatchLaunch* batch);
template <> hipError_t launch_step_fp64_shape<6>(const StepArgs& a, hipStream_t st, const BatchLaunch* batch);
hipError_t launch_step_fp64(const StepArgs& a, int shape, hipStream_t st, const BatchLaunch* batch) {
    switch (shape) {
        case 0: return launch_step_fp64_shape<0>(a, st, batch);
        case 1: return launch_step_fp64_shape<1>(a, st, batch);
        case 2: return launch_step_fp64_shape<2>(a, st, batch);
        case 3: return launch_step_fp64_shape<3>(a, st, batch);
        case 4: return launch_step_fp64_shape<4>(a, st, batch);
        case 5: return launch_step_fp64_shape<5>(a, st, batch);
        case 6: return launch_step_fp64_shape<6>(a, st, batch);
        default: return hipErrorInvalidValue;
    }
}
#endif
static_assert(SHAPE_COUNT == 7, "one dispatcher case per tile shape");

}  // namespace drp
