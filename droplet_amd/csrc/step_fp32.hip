// step_fp32.hip — the D3Q15 two-phase step with FP32 shifted storage.
//
// Storage (all FP32):
//   phi populations      ft_q = f_q - fr_q,   fr_q = H_q phi1 + F_q p0(phi1)
//   momentum populations gt_q = g_q - E_q     (E_q = g_q at rest with p = 1/3)
//   fields               phit = phi - phi1, pt = p - 1/3, u1, u2, u3
// The reference states (the continuous phase at rest) are where the bulk of the
// domain sits, so the stored deviations are small and FP32 keeps their
// relative precision.  Sum_q fr_q = phi1 and Sum_q E_q (1, c) = (1, 0), so the
// moments of the deviations are the deviations of the moments.
//
// Arithmetic: equilibria and collision in FP32 with fused multiply-adds,
// written directly in deviation form (no difference of two large numbers):
//   p0(phi) - p0(phi1) = phit [T / ((1 - b phi)(1 - b phi1)) - a (phi + phi1)]
//   ft_eq = H phit + F At + 3 E phi (c.u) + E kf gt
//   gt_eq = E (3 pt + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2) + E kg gt
// with At = [p0(phi) - p0(phi1)] - kf phi lap - kf/6 |grad phi|^2.  The
// Laplacian and gradient of phi (the chemical-potential terms at the interface)
// are accumulated in FP64, as is the phi moment; diagnostics reductions run in
// FP64 on an FP64 copy of the fields.
//
// Step structure and reference mapping are those of step_kernels.hip
// (program/lbm_drop_xy.f90:421-588, prototype_program/two-phase/dropch.f90:666-698).
#include "d3q15.h"
#include "kernels.h"

namespace drp {

template <int Q0, int Q1, typename Fn> __device__ __forceinline__ void sfor32(Fn&& fn) {
    if constexpr (Q0 < Q1) {
        fn(std::integral_constant<int, Q0>{});
        sfor32<Q0 + 1, Q1>(fn);
    }
}

__device__ __forceinline__ int wrap32(int v, int n, bool mod) {
    if (mod) {
        v %= n;
        return v < 0 ? v + n : v;
    }
    return v < 0 ? v + n : (v >= n ? v - n : v);
}

template <int TX, int TY, int TZ, bool WALL, int FORCE, int TAUMODE>
__device__ __forceinline__ void step_tile_fp32(const StepArgs32& a, int tile) {
    constexpr int NN = TX * TY * TZ, NT = 2 * NN;
    constexpr int HX2 = TX + 4, HY2 = TY + 4, HZ2 = TZ + 4, H2 = HX2 * HY2 * HZ2;
    constexpr int HX1 = TX + 2, HY1 = TY + 2, HZ1 = TZ + 2, H1 = HX1 * HY1 * HZ1;
    static_assert(NN % 64 == 0, "a role must own whole wavefronts");

    // Per halo node the stencil phase leaves four float4 records, so that the
    // collision phase needs two 16-byte LDS reads and about ten FMAs per direction:
    //   fa = (At, 3 phi u1, 3 phi u2, 3 phi u3)       phi-population equilibrium
    //   ga = (3 pt - 1.5 |u|^2, u1, u2, u3)           momentum equilibrium
    //   k[0..2] = K_k = 4.5 g_k^2 - 1.5 |g|^2                       Korteweg, axis c
    //   k[3..5] = 9 g0 g1, 9 g0 g2, 9 g1 g2                          Korteweg, diagonal c
    // (six float arrays, so that lanes reading consecutive nodes hit distinct
    // banks).  56 bytes per halo node: two 1024-thread workgroups fit one CU.
    // For a diagonal c (c_a^2 = 1) the contraction 4.5 (c.g)^2 - 1.5 |g|^2 |c|^2
    // reduces to 9 (cx cy g0 g1 + cx cz g0 g2 + cy cz g1 g2).
    __shared__ float s_phi[H2];  // phit; with a tau field its first H1 entries hold 1 / tau_g in phase 3
    __shared__ float4 s_fa[H1], s_ga[H1];
    __shared__ float s_k[6][H1];
    // 1 / tau_g per halo node reuses s_phi once the stencils are done, which keeps
    // the variable-tau variants at two workgroups per CU (an own array would add
    // 4 B per halo node and exceed half of the LDS)
    float* const s_rtau = s_phi;
    // staging: thread t owns column t % HX of a halo box, rows t / HX + k (NT / HX)
    // (row = y + HY z), so that its x coordinate and wrap are fixed
    constexpr int RP2 = NT / HX2, R2 = HY2 * HZ2, IT2 = (R2 + RP2 - 1) / RP2;
    constexpr int RP1 = NT / HX1, R1 = HY1 * HZ1, IT1 = (R1 + RP1 - 1) / RP1;
    static_assert(RP2 >= 1 && RP1 >= 1, "a workgroup spans at least one halo row");
    float rt_reg[TAUMODE == 0 ? 1 : IT1];
    __shared__ double s_red[2][NT / 64];

    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t N = a.n;
    const int x0 = (tile % a.tiles_x) * TX;
    const int y0 = ((tile / a.tiles_x) % a.tiles_y) * TY;
    const int z0 = (tile / (a.tiles_x * a.tiles_y)) * TZ;
    const int tid = threadIdx.x;
    const bool mom = tid >= NN;
    const int nid = mom ? tid - NN : tid;
    __builtin_assume(nid >= 0 && nid < NN);

    const float* __restrict__ phi_in = a.mac_in;
    const float* __restrict__ p_in = a.mac_in + N;
    const float* __restrict__ u1_in = a.mac_in + 2 * N;
    const float* __restrict__ u2_in = a.mac_in + 3 * N;
    const float* __restrict__ u3_in = a.mac_in + 4 * N;

    // phase 0: this thread's 15 upstream populations
    const int tx = nid % TX, ty = (nid / TX) % TY, tz = nid / (TX * TY);
    // 1-halo index of (x - 1, y - 1, z - 1): the upstream records of every
    // direction lie at non-negative constant offsets from it
    const int hb1 = tx + HX1 * (ty + HY1 * tz);
    const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
    const bool inside = (x < nx) && (y < ny) && (z < nz);
    const int xc = inside ? x : 0, yc = inside ? y : 0, zc = inside ? z : 0;
    const int xs[3] = {xc + 1 == nx ? 0 : xc + 1, xc, xc == 0 ? nx - 1 : xc - 1};
    const int zs[3] = {zc + 1 == nz ? 0 : zc + 1, zc, zc == 0 ? nz - 1 : zc - 1};
    // node indices as 32-bit factors (an x-y plane holds fewer than 2^31 nodes)
    const uint32_t nxy32 = (uint32_t)nx * (uint32_t)ny;
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
    const float* __restrict__ pin = (__builtin_amdgcn_readfirstlane(mom ? 1 : 0) != 0) ? a.g_in : a.f_in;
    float ps[15];
    if (inside) {
        sfor32<0, 15>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            ps[q] = pin[q * N + upstream(qc)];
        });
    }

    // phase 1: stage fields (p and u into ga; p is replaced in phase 2)
    // the rows of a tile whose halo box lies inside the grid in y and z need no wrap
    const bool wm = a.wrap_mod != 0;
    const bool yz_in = !wm && y0 >= 2 && y0 + TY + 2 <= ny && z0 >= 2 && z0 + TZ + 2 <= nz;
    auto row_offset = [&](int hy, int hz, int halo) -> int64_t {
        int gy = y0 - halo + hy, gz = z0 - halo + hz;
        if (!yz_in) {
            gy = wrap32(gy, ny, wm);
            gz = wrap32(gz, nz, wm);
        }
        return (int64_t)((uint64_t)(uint32_t)gz * nxy32 + (uint32_t)(gy * nx));
    };
    const int c2 = tid % HX2, r2 = tid / HX2, c1 = tid % HX1, r1 = tid / HX1;
    const int gx2 = wrap32(x0 - 2 + c2, nx, wm), gx1 = wrap32(x0 - 1 + c1, nx, wm);
    if (r2 < RP2) {
        sfor32<0, IT2>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const int r = r2 + k * RP2;
            if (r < R2) s_phi[c2 + HX2 * r] = phi_in[gx2 + row_offset(r % HY2, r / HY2, 2)];
        });
    }
    sfor32<0, IT1>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const int r = r1 + k * RP1;
        if (r1 < RP1 && r < R1) {
            const int64_t g = gx1 + row_offset(r % HY1, r / HY1, 1);
            s_ga[c1 + HX1 * r] = make_float4(p_in[g], u1_in[g], u2_in[g], u3_in[g]);
            if constexpr (TAUMODE == 2) rt_reg[k] = 1.0f / a.tau_field[g];
        }
    });
    __syncthreads();

    // phase 2: FP64-accumulated Laplacian and gradient, deviation F-bracket, the
    // equilibrium records, 1/tau
    const double rmax = from_ordered_bits(a.range[2 * a.slot_read + 0]);
    const double rmin = from_ordered_bits(a.range[2 * a.slot_read + 1]);
    const float range = (float)(rmax - rmin), rminf = (float)(rmin - a.phi1);
    sfor32<0, IT1>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const int r = r1 + k * RP1;
        if (r1 >= RP1 || r >= R1) return;
        const int h = c1 + HX1 * r;
        // phi-box index of the node's (-1, -1, -1) neighbour: the stencil reads lie
        // at non-negative constant offsets from it
        const int cb = c1 + HX2 * (r % HY1 + HY2 * (r / HY1));
        const float pt = s_phi[cb + (1 + HX2 * (1 + HY2))];
        double lap = -14.0 * (double)pt, g0 = 0.0, g1 = 0.0, g2 = 0.0;
        sfor32<1, 15>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            using D = Dir<q>;
            const double v = (double)s_phi[cb + ((1 + D::cx) + HX2 * ((1 + D::cy) + HY2 * (1 + D::cz)))];
            lap += v;
            g0 = acc<D::cx>(g0, v);
            g1 = acc<D::cy>(g1, v);
            g2 = acc<D::cz>(g2, v);
        });
        const float lapf = (float)(lap * 0.2), gf0 = (float)(g0 * 0.1), gf1 = (float)(g1 * 0.1),
                    gf2 = (float)(g2 * 0.1);
        const float ph = a.phi1f + pt;
        const float gsq = __builtin_fmaf(gf0, gf0, __builtin_fmaf(gf1, gf1, gf2 * gf2));
        const float dp0 = pt * (a.Tf / ((1.0f - a.bf * ph) * a.one_m_bphi1) - a.af * (ph + a.phi1f));
        const float At = dp0 - a.kff * ph * lapf - a.kf6f * gsq;
        const float4 pu = s_ga[h];
        const float usq = __builtin_fmaf(pu.y, pu.y, __builtin_fmaf(pu.z, pu.z, pu.w * pu.w));
        const float ph3 = 3.0f * ph;
        s_fa[h] = make_float4(At, ph3 * pu.y, ph3 * pu.z, ph3 * pu.w);
        s_ga[h] = make_float4(__builtin_fmaf(-1.5f, usq, 3.0f * pu.x), pu.y, pu.z, pu.w);
        const float s2h = 1.5f * gsq;
        s_k[0][h] = __builtin_fmaf(4.5f * gf0, gf0, -s2h);
        s_k[1][h] = __builtin_fmaf(4.5f * gf1, gf1, -s2h);
        s_k[2][h] = __builtin_fmaf(4.5f * gf2, gf2, -s2h);
        s_k[3][h] = 9.0f * gf0 * gf1;
        s_k[4][h] = 9.0f * gf0 * gf2;
        s_k[5][h] = 9.0f * gf1 * gf2;
        if constexpr (TAUMODE == 1) {
            float frac = (range == 0.0f) ? 0.0f : (pt - rminf) / range;
            rt_reg[k] = 1.0f / (0.5f + 3.0f * (frac * a.dnuf + a.nu1f));
        }
    });
    // with a tau field, the own node's phit (the f-role's q = 0 term) is read
    // before s_phi is reused
    const int h2self = (tx + 2) + HX2 * ((ty + 2) + HY2 * (tz + 2));
    float phi_self = 0.0f;
    if constexpr (TAUMODE != 0) phi_self = s_phi[h2self];
    __syncthreads();
    if constexpr (TAUMODE != 0) {
        sfor32<0, IT1>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const int r = r1 + k * RP1;
            if (r1 < RP1 && r < R1) s_rtau[c1 + HX1 * r] = rt_reg[k];
        });
        __syncthreads();
    }

    // Korteweg contraction gt_q and c_q . v from the records of the upstream node
    auto gt_of = [&](auto qc, int hu) -> float {
        constexpr int q = decltype(qc)::value;
        using D = Dir<q>;
        if constexpr (q == 0) {
            return 0.0f;
        } else if constexpr (q <= 6) {
            return s_k[D::cx != 0 ? 0 : (D::cy != 0 ? 1 : 2)][hu];
        } else {
            const float p01 = s_k[3][hu], p02 = s_k[4][hu], p12 = s_k[5][hu];
            float t = (D::cx * D::cy > 0) ? p01 : -p01;
            t = (D::cx * D::cz > 0) ? t + p02 : t - p02;
            return (D::cy * D::cz > 0) ? t + p12 : t - p12;
        }
    };

    // phase 3: collide-stream
    double vmax = -__builtin_huge_val(), vmin = __builtin_huge_val();
    if (inside) {
        const int64_t s = x + (int64_t)nx * (y + (int64_t)ny * z);
        const int hself = (tx + 1) + HX1 * ((ty + 1) + HY1 * (tz + 1));
        if (!mom) {
            double phs = 0.0;
            const float ekf7 = (float)weightE<7>() * a.kff, ekf1 = (float)weightE<1>() * a.kff;
            sfor32<0, 15>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                using D = Dir<q>;
                constexpr float E = (float)weightE<q>(), Fw = (float)weightF<q>();
                const int hu = hb1 + ((1 - D::cx) + HX1 * ((1 - D::cy) + HY1 * (1 - D::cz)));
                float fe;
                if constexpr (q == 0) {
                    fe = __builtin_fmaf(Fw, s_fa[hself].x, TAUMODE != 0 ? phi_self : s_phi[h2self]);
                } else {
                    const float4 fa = s_fa[hu];
                    const float cv = cdot<q>(fa.y, fa.z, fa.w);
                    const float gt = gt_of(qc, hu);
                    fe = __builtin_fmaf(Fw, fa.x, __builtin_fmaf(E, cv, (q <= 6 ? ekf1 : ekf7) * gt));
                }
                const float fv = __builtin_fmaf(fe - ps[q], a.rtau_f, ps[q]);
                phs += (double)fv;
                __builtin_nontemporal_store(fv, &a.f_out[q * N + s]);
                __builtin_amdgcn_sched_barrier(0);
            });
            __builtin_nontemporal_store((float)phs, &a.mac_out[s]);
            vmax = vmin = phs + a.phi1;
        } else {
            const float ekg7 = (float)weightE<7>() * a.kgf, ekg1 = (float)weightE<1>() * a.kgf;
            sfor32<0, 15>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                using D = Dir<q>;
                constexpr float E = (float)weightE<q>();
                const int hu = hb1 + ((1 - D::cx) + HX1 * ((1 - D::cy) + HY1 * (1 - D::cz)));
                float ge;
                if constexpr (q == 0) {
                    ge = E * s_ga[hself].x;
                } else {
                    const float4 ga = s_ga[hu];
                    const float cu = cdot<q>(ga.y, ga.z, ga.w);
                    const float t = __builtin_fmaf(cu, __builtin_fmaf(4.5f, cu, 3.0f), ga.x);
                    ge = __builtin_fmaf(E, t, (q <= 6 ? ekg1 : ekg7) * gt_of(qc, hu));
                }
                const float rt = (TAUMODE == 0) ? a.rtau_const : s_rtau[hu];
                float gv = __builtin_fmaf(ge - ps[q], rt, ps[q]);
                if constexpr (FORCE != FORCE_NONE && q != 0) {
                    float Fx = 0.0f, Fy = 0.0f, Fz = 0.0f;   // only the components that enter c . F
                    if constexpr (FORCE == FORCE_TAYLOR_GREEN) {
                        const int xu = xs[1 + D::cx], zu = zs[1 + D::cz];
                        if constexpr (D::cx != 0) Fx = (float)(a.tg_amp * a.sinx[xu] * a.cosz[zu]);
                        if constexpr (D::cz != 0) Fz = (float)(-(a.tg_amp * a.cosx[xu] * a.sinz[zu]));
                    } else {
                        const int64_t src = upstream(qc);
                        if constexpr (D::cx != 0) Fx = (float)a.fx[src];
                        if constexpr (D::cy != 0) Fy = (float)a.fy[src];
                        if constexpr (D::cz != 0) Fz = (float)a.fz[src];
                    }
                    gv = __builtin_fmaf(3.0f * E, cdot<q>(Fx, Fy, Fz), gv);
                }
                ps[q] = gv;
                __builtin_amdgcn_sched_barrier(0);
            });
            if constexpr (WALL) {
                const int coord = (a.wall_axis == 2) ? y : z;
                const int nax = (a.wall_axis == 2) ? ny : nz;
                const bool bottom = coord == 0, top = coord == nax - 1;
                if (bottom || top) {
                    const float uw = bottom ? -a.wall_u_now : a.wall_u_now;
                    const int want = bottom ? 1 : -1;
                    sfor32<1, 15>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        using D = Dir<q>;
                        constexpr int o = D::opp;
                        const int ca = (a.wall_axis == 2) ? D::cy : D::cz;
                        if (ca == want) ps[q] = ps[o] - 6.0f * (float)weightE<o>() * (uw * (float)Dir<o>::cx);
                    });
                }
            }
            float pr = 0.0f, m1 = 0.0f, m2 = 0.0f, m3 = 0.0f;
            sfor32<0, 15>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                using D = Dir<q>;
                pr += ps[q];
                m1 = acc<D::cx>(m1, ps[q]);
                m2 = acc<D::cy>(m2, ps[q]);
                m3 = acc<D::cz>(m3, ps[q]);
                __builtin_nontemporal_store(ps[q], &a.g_out[q * N + s]);
            });
            __builtin_nontemporal_store(pr * (1.0f / 3.0f), &a.mac_out[N + s]);
            __builtin_nontemporal_store(m1, &a.mac_out[2 * N + s]);
            __builtin_nontemporal_store(m2, &a.mac_out[3 * N + s]);
            __builtin_nontemporal_store(m3, &a.mac_out[4 * N + s]);
        }
    }

    // phi range for nu(phi) of the next step
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

// Resident wavefronts per SIMD the register allocation targets: two workgroups
// per CU for the 1024-thread tiles (8 per SIMD, <= 64 VGPRs), three for the
// 512-thread tiles (6 per SIMD), as far as the LDS records allow.
template <int NT> constexpr int waves_per_eu32() { return NT >= 1024 ? 8 : 6; }

// one lattice: workgroup id -> tile (XCD-contiguous)
template <int TX, int TY, int TZ, bool WALL, int FORCE, int TAUMODE>
__global__ __launch_bounds__(2 * TX* TY* TZ) __attribute__((amdgpu_waves_per_eu(waves_per_eu32<2 * TX * TY * TZ>()))) void step_fp32(StepArgs32 a) {
    step_tile_fp32<TX, TY, TZ, WALL, FORCE, TAUMODE>(a, xcd_tile(blockIdx.x, gridDim.x));
}

// many independent lattices in one launch: tile_start[i] is the first global
// tile of instance i (ascending, tile_start[0] = 0); the instance of a
// workgroup is found by binary search, its arguments read from global memory
template <int TX, int TY, int TZ, bool WALL, int FORCE, int TAUMODE>
__global__ __launch_bounds__(2 * TX* TY* TZ) __attribute__((amdgpu_waves_per_eu(waves_per_eu32<2 * TX * TY * TZ>()))) void step_batch_fp32(const StepArgs32* __restrict__ args, const int* __restrict__ tile_start, int ninst) {
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    step_tile_fp32<TX, TY, TZ, WALL, FORCE, TAUMODE>(args[lo], b - tile_start[lo]);
}

template <int TX, int TY, int TZ, bool WALL, int FORCE, int TAUM>
static hipError_t launch32_shape(const StepArgs32& a, hipStream_t st, const BatchLaunch* batch) {
    if (batch) {
        hipLaunchKernelGGL((step_batch_fp32<TX, TY, TZ, WALL, FORCE, TAUM>), dim3(batch->total_tiles),
                           dim3(2 * TX * TY * TZ), 0, st, (const StepArgs32*)batch->args, batch->tile_start,
                           batch->ninst);
        return hipGetLastError();
    }
    StepArgs32 b = a;
    set_tiles(b, TX, TY, TZ);
    const int nb = b.tiles_x * b.tiles_y * b.tiles_z;
    hipLaunchKernelGGL((step_fp32<TX, TY, TZ, WALL, FORCE, TAUM>), dim3(nb), dim3(2 * TX * TY * TZ), 0, st, b);
    return hipGetLastError();
}

template <int TX, int TY, int TZ, bool WALL, int FORCE>
static hipError_t launch32_tau(const StepArgs32& a, hipStream_t st, const BatchLaunch* batch) {
    if (a.tau_field) return launch32_shape<TX, TY, TZ, WALL, FORCE, 2>(a, st, batch);
    if (a.dnuf == 0.0f) return launch32_shape<TX, TY, TZ, WALL, FORCE, 0>(a, st, batch);
    return launch32_shape<TX, TY, TZ, WALL, FORCE, 1>(a, st, batch);
}

template <int TX, int TY, int TZ>
static hipError_t launch32_variant(const StepArgs32& a, hipStream_t st, const BatchLaunch* batch) {
    const bool wall = a.wall_axis != 0;
    switch (a.force_kind) {
        case FORCE_NONE:
            return wall ? launch32_tau<TX, TY, TZ, true, FORCE_NONE>(a, st, batch)
                        : launch32_tau<TX, TY, TZ, false, FORCE_NONE>(a, st, batch);
        case FORCE_TAYLOR_GREEN:
            return wall ? launch32_tau<TX, TY, TZ, true, FORCE_TAYLOR_GREEN>(a, st, batch)
                        : launch32_tau<TX, TY, TZ, false, FORCE_TAYLOR_GREEN>(a, st, batch);
        case FORCE_FIELD:
            return wall ? launch32_tau<TX, TY, TZ, true, FORCE_FIELD>(a, st, batch)
                        : launch32_tau<TX, TY, TZ, false, FORCE_FIELD>(a, st, batch);
        default:
            return hipErrorInvalidValue;
    }
}

// Compiled once per tile shape (DRP_STEP_SHAPE = k) and once for the dispatcher,
// as step_kernels.hip.
template <int S> hipError_t launch_step_fp32_shape(const StepArgs32& a, hipStream_t st, const BatchLaunch* batch);

#if defined(DRP_STEP_SHAPE)
template <>
hipError_t launch_step_fp32_shape<DRP_STEP_SHAPE>(const StepArgs32& a, hipStream_t st, const BatchLaunch* batch) {
    constexpr int S = DRP_STEP_SHAPE;
    return launch32_variant<SHAPE_DIMS[S][0], SHAPE_DIMS[S][1], SHAPE_DIMS[S][2]>(a, st, batch);
}
#else
template <> hipError_t launch_step_fp32_shape<0>(const StepArgs32& a, hipStream_t st, const BatchLaunch* batch);
template <> hipError_t launch_step_fp32_shape<1>(const StepArgs32& a, hipStream_t st, const BatchLaunch* batch);
template <> hipError_t launch_step_fp32_shape<2>(const StepArgs32& a, hipStream_t st, const BatchLaunch* batch);
template <> hipError_t launch_step_fp32_shape<3>(const StepArgs32& a, hipStream_t st, const BatchLaunch* batch);
template <> hipError_t launch_step_fp32_shape<4>(const StepArgs32& a, hipStream_t st, const BatchLaunch* batch);
template <> hipError_t launch_step_fp32_shape<5>(const StepArgs32& a, hipStream_t st, const BatchLaunch* batch);
template <> hipError_t launch_step_fp32_shape<6>(const StepArgs32& a, hipStream_t st, const BatchLaunch* batch);
hipError_t launch_step_fp32(const StepArgs32& a, int shape, hipStream_t st, const BatchLaunch* batch) {
    switch (shape) {
        case 0: return launch_step_fp32_shape<0>(a, st, batch);
        case 1: return launch_step_fp32_shape<1>(a, st, batch);
        case 2: return launch_step_fp32_shape<2>(a, st, batch);
        case 3: return launch_step_fp32_shape<3>(a, st, batch);
        case 4: return launch_step_fp32_shape<4>(a, st, batch);
        case 5: return launch_step_fp32_shape<5>(a, st, batch);
        case 6: return launch_step_fp32_shape<6>(a, st, batch);
        default: return hipErrorInvalidValue;
    }
}
#endif

}  // namespace drp
