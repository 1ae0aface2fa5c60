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

__device__ __forceinline__ int wrapi(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

// Map the linear workgroup id to a tile so that each XCD (workgroups are dealt
// round-robin to the 8 XCDs) gets a contiguous block of tiles: neighbouring
// tiles, which share halo lines, then share one L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    if ((nb & 7) != 0) return b;
    return (b & 7) * (nb >> 3) + (b >> 3);
}

template <int TX, int TY, int TZ, bool WALL, int FORCE, bool TAUFIELD>
__global__ __launch_bounds__(TX* TY* TZ) void step_fp64(StepArgs a) {
    constexpr int NT = TX * TY * TZ;
    constexpr int HX2 = TX + 4, HY2 = TY + 4, HZ2 = TZ + 4, H2 = HX2 * HY2 * HZ2;
    constexpr int HX1 = TX + 2, HY1 = TY + 2, HZ1 = TZ + 2, H1 = HX1 * HY1 * HZ1;
    constexpr bool FFIELD = (FORCE == FORCE_FIELD);

    __shared__ double s_phi[H2];
    __shared__ double s_p[H1], s_u1[H1], s_u2[H1], s_u3[H1];
    __shared__ double s_A[H1], s_g0[H1], s_g1[H1], s_g2[H1], s_tau[H1];
    __shared__ double s_fx[FFIELD ? H1 : 1], s_fy[FFIELD ? H1 : 1], s_fz[FFIELD ? H1 : 1];
    __shared__ double s_red[2][NT / 64];

    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t N = a.n;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int x0 = (tile % a.tiles_x) * TX;
    const int y0 = ((tile / a.tiles_x) % a.tiles_y) * TY;
    const int z0 = (tile / (a.tiles_x * a.tiles_y)) * TZ;
    const int tid = threadIdx.x;

    const double* __restrict__ phi_in = a.mac_in;
    const double* __restrict__ p_in = a.mac_in + N;
    const double* __restrict__ u1_in = a.mac_in + 2 * N;
    const double* __restrict__ u2_in = a.mac_in + 3 * N;
    const double* __restrict__ u3_in = a.mac_in + 4 * N;

    // ---------------- phase 1: stage macroscopic fields
    for (int h = tid; h < H2; h += NT) {
        const int hx = h % HX2, hy = (h / HX2) % HY2, hz = h / (HX2 * HY2);
        const int gx = wrapi(x0 + hx - 2, nx), gy = wrapi(y0 + hy - 2, ny), gz = wrapi(z0 + hz - 2, nz);
        s_phi[h] = phi_in[gx + (int64_t)nx * (gy + (int64_t)ny * gz)];
    }
    for (int h = tid; h < H1; h += NT) {
        const int hx = h % HX1, hy = (h / HX1) % HY1, hz = h / (HX1 * HY1);
        const int gx = wrapi(x0 + hx - 1, nx), gy = wrapi(y0 + hy - 1, ny), gz = wrapi(z0 + hz - 1, nz);
        const int64_t g = gx + (int64_t)nx * (gy + (int64_t)ny * gz);
        s_p[h] = p_in[g];
        s_u1[h] = u1_in[g];
        s_u2[h] = u2_in[g];
        s_u3[h] = u3_in[g];
        if constexpr (TAUFIELD) s_tau[h] = a.tau_field[g];
        if constexpr (FFIELD) {
            s_fx[h] = a.fx[g];
            s_fy[h] = a.fy[g];
            s_fz[h] = a.fz[g];
        }
    }
    __syncthreads();

    // ---------------- phase 2: stencils, F-bracket and tau at tile + 1 halo
    const double rmax = from_ordered_bits(a.range[2 * a.slot_read + 0]);
    const double rmin = from_ordered_bits(a.range[2 * a.slot_read + 1]);
    const double range = rmax - rmin;
    for (int h = tid; h < H1; h += NT) {
        const int hx = h % HX1, hy = (h / HX1) % HY1, hz = h / (HX1 * HY1);
        const int c = (hx + 1) + HX2 * ((hy + 1) + HY2 * (hz + 1));
        const double ph = s_phi[c];
        double lap = -14.0 * ph;
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        static_for<1, 15>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            using D = Dir<q>;
            const double v = s_phi[c + D::cx + HX2 * (D::cy + HY2 * D::cz)];
            lap = lap + v;
            g0 = acc<D::cx>(g0, v);
            g1 = acc<D::cy>(g1, v);
            g2 = acc<D::cz>(g2, v);
        });
        lap = lap / 5.0;
        g0 = g0 / 10.0;
        g1 = g1 / 10.0;
        g2 = g2 / 10.0;
        const double p0 = ph * a.T / (1.0 - a.b * ph) - a.a * (ph * ph);
        s_A[h] = p0 - a.kf * ph * lap - a.kf6 * (g0 * g0 + g1 * g1 + g2 * g2);
        s_g0[h] = g0;
        s_g1[h] = g1;
        s_g2[h] = g2;
        if constexpr (!TAUFIELD) {
            double frac = (ph - rmin) / range;
            if (range == 0.0) frac = 0.0;
            const double nu = frac * a.dnu + a.nu1;
            s_tau[h] = 0.5 + 3.0 * nu;
        }
    }
    __syncthreads();

    // ---------------- phase 3: collide-stream at each output node
    const int tx = tid % TX, ty = (tid / TX) % TY, tz = tid / (TX * TY);
    const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
    const bool inside = (x < nx) && (y < ny) && (z < nz);
    double phin = 0.0;
    if (inside) {
        // upstream coordinate x - c is entry [c + 1] of {x+1, x, x-1} (periodic wrap)
        const int64_t xo[3] = {x + 1 == nx ? 0 : x + 1, x, x == 0 ? nx - 1 : x - 1};
        const int64_t yo[3] = {(int64_t)nx * (y + 1 == ny ? 0 : y + 1), (int64_t)nx * y,
                               (int64_t)nx * (y == 0 ? ny - 1 : y - 1)};
        const int64_t nxy = (int64_t)nx * ny;
        const int64_t zo[3] = {nxy * (z + 1 == nz ? 0 : z + 1), nxy * z, nxy * (z == 0 ? nz - 1 : z - 1)};
        const int xs[3] = {(int)xo[0], (int)xo[1], (int)xo[2]};
        const int zs[3] = {z + 1 == nz ? 0 : z + 1, z, z == 0 ? nz - 1 : z - 1};

        double fq[15], gq[15];
        static_for<0, 15>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            using D = Dir<q>;
            const int hu = (tx + 1 - D::cx) + HX1 * ((ty + 1 - D::cy) + HY1 * (tz + 1 - D::cz));
            const int h2 = (tx + 2 - D::cx) + HX2 * ((ty + 2 - D::cy) + HY2 * (tz + 2 - D::cz));
            NodeEq<double> nd;
            nd.phi = s_phi[h2];
            nd.p = s_p[hu];
            nd.u1 = s_u1[hu];
            nd.u2 = s_u2[hu];
            nd.u3 = s_u3[hu];
            nd.A = s_A[hu];
            nd.g0 = s_g0[hu];
            nd.g1 = s_g1[hu];
            nd.g2 = s_g2[hu];
            nd.tau = s_tau[hu];
            const double usq = nd.u1 * nd.u1 + nd.u2 * nd.u2 + nd.u3 * nd.u3;
            const double s2h = 1.5 * (nd.g0 * nd.g0 + nd.g1 * nd.g1 + nd.g2 * nd.g2);
            const double gt = korteweg<q>(nd.g0, nd.g1, nd.g2, s2h);
            const double fe = feq<q>(nd, gt, a.kf);
            const double ge = geq<q>(nd, usq, gt, a.kg);
            const int64_t src = xo[1 + D::cx] + yo[1 + D::cy] + zo[1 + D::cz];
            const double fs = a.f_in[q * N + src];
            const double gs = a.g_in[q * N + src];
            fq[q] = fs - (fs - fe) / a.tau_f;
            double gv = gs - (gs - ge) / nd.tau;
            if constexpr (FORCE != FORCE_NONE && q != 0) {
                double Fx, Fy, Fz;
                if constexpr (FORCE == FORCE_TAYLOR_GREEN) {
                    const int xu = xs[1 + D::cx], zu = zs[1 + D::cz];
                    Fx = a.tg_amp * a.sinx[xu] * a.cosz[zu];
                    Fy = 0.0;
                    Fz = -(a.tg_amp * a.cosx[xu] * a.sinz[zu]);
                } else {
                    Fx = s_fx[hu];
                    Fy = s_fy[hu];
                    Fz = s_fz[hu];
                }
                double dot = (double)D::cx * Fx + (double)D::cy * Fy;
                dot = dot + (double)D::cz * Fz;
                gv = gv + 3.0 * weightE<q>() * dot;
            }
            gq[q] = gv;
        });

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
                        gq[q] = gq[o] - 6.0 * weightE<o>() * dot;
                    }
                });
            }
        }

        // moments (physics :546-552), reference order i = 1..15
        double ph = fq[0], pr = gq[0] / 3.0;
        double m1 = 0.0, m2 = 0.0, m3 = 0.0;
        static_for<1, 15>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            using D = Dir<q>;
            ph = ph + fq[q];
            pr = pr + gq[q] / 3.0;
            m1 = acc<D::cx>(m1, gq[q]);
            m2 = acc<D::cy>(m2, gq[q]);
            m3 = acc<D::cz>(m3, gq[q]);
        });
        const int64_t s = x + (int64_t)nx * (y + (int64_t)ny * z);
        static_for<0, 15>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            a.f_out[q * N + s] = fq[q];
            a.g_out[q * N + s] = gq[q];
        });
        a.mac_out[s] = ph;
        a.mac_out[N + s] = pr;
        a.mac_out[2 * N + s] = m1;
        a.mac_out[3 * N + s] = m2;
        a.mac_out[4 * N + s] = m3;
        phin = ph;
    }

    // ---------------- global phi range for the next step's nu(phi)
    double vmax = inside ? phin : -__builtin_huge_val();
    double vmin = inside ? phin : __builtin_huge_val();
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
        for (int w = 1; w < NT / 64; w++) {
            mx = fmax(mx, s_red[0][w]);
            mn = fmin(mn, s_red[1][w]);
        }
        atomicMax(reinterpret_cast<unsigned long long*>(&a.range[2 * a.slot_write + 0]),
                  (unsigned long long)ordered_bits(mx));
        atomicMin(reinterpret_cast<unsigned long long*>(&a.range[2 * a.slot_write + 1]),
                  (unsigned long long)ordered_bits(mn));
        if (blockIdx.x == 0) {
            a.range[2 * a.slot_reset + 0] = ordered_bits(a.phi1);
            a.range[2 * a.slot_reset + 1] = ordered_bits(a.phi2);
        }
    }
}

// ------------------------------------------------------------------ launcher
template <int TX, int TY, int TZ, bool WALL, int FORCE, bool TAUF>
static hipError_t launch_shape(const StepArgs& a, hipStream_t st) {
    StepArgs b = a;
    b.tiles_x = (a.nx + TX - 1) / TX;
    b.tiles_y = (a.ny + TY - 1) / TY;
    b.tiles_z = (a.nz + TZ - 1) / TZ;
    const int nb = b.tiles_x * b.tiles_y * b.tiles_z;
    hipLaunchKernelGGL((step_fp64<TX, TY, TZ, WALL, FORCE, TAUF>), dim3(nb), dim3(TX * TY * TZ), 0, st, b);
    return hipGetLastError();
}

template <int TX, int TY, int TZ>
static hipError_t launch_variant(const StepArgs& a, hipStream_t st) {
    const bool wall = a.wall_axis != 0;
    const bool tauf = a.tau_field != nullptr;
#define DRP_CASE(W, F, T)                                                             \
    if (wall == W && a.force_kind == F && tauf == T) return launch_shape<TX, TY, TZ, W, F, T>(a, st);
    DRP_CASE(false, FORCE_NONE, false)
    DRP_CASE(true, FORCE_NONE, false)
    DRP_CASE(false, FORCE_TAYLOR_GREEN, false)
    DRP_CASE(true, FORCE_TAYLOR_GREEN, false)
    DRP_CASE(false, FORCE_FIELD, false)
    DRP_CASE(true, FORCE_FIELD, false)
    DRP_CASE(false, FORCE_NONE, true)
    DRP_CASE(true, FORCE_NONE, true)
    DRP_CASE(false, FORCE_TAYLOR_GREEN, true)
    DRP_CASE(true, FORCE_TAYLOR_GREEN, true)
    DRP_CASE(false, FORCE_FIELD, true)
    DRP_CASE(true, FORCE_FIELD, true)
#undef DRP_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_step_fp64(const StepArgs& a, int shape, hipStream_t st) {
    switch (shape) {
        case SHAPE_16x4x4: return launch_variant<16, 4, 4>(a, st);
        case SHAPE_32x4x2: return launch_variant<32, 4, 2>(a, st);
        case SHAPE_64x2x2: return launch_variant<64, 2, 2>(a, st);
        case SHAPE_32x2x4: return launch_variant<32, 2, 4>(a, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace drp
