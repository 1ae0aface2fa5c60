// aux_kernels.hip — initialisation, moment recomputation, LES relaxation-time
// field and diagnostics reductions (FP64).  None of these is on the per-step hot
// path except the LES tau field, the eddy-viscosity part of the reference
// `physics` (program/LES_drop.f90:598-693), computed after each step.
#include "d3q15.h"
#include "kernels.h"

namespace drp {

template <int Q0, int Q1, typename Fn> __device__ __forceinline__ void sfor(Fn&& fn) {
    if constexpr (Q0 < Q1) {
        fn(std::integral_constant<int, Q0>{});
        sfor<Q0 + 1, Q1>(fn);
    }
}

struct Grid {
    int nx, ny, nz;
    __device__ __forceinline__ int64_t at(int x, int y, int z) const {
        x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
        y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
        z = z < 0 ? z + nz : (z >= nz ? z - nz : z);
        return x + (int64_t)nx * (y + (int64_t)ny * z);
    }
};

// lap phi and grad phi at (x,y,z) from global memory (various_cal :608-633)
__device__ void phi_stencil(const Grid& G, const double* phi, int x, int y, int z, double& lap, double& g0,
                            double& g1, double& g2) {
    lap = -14.0 * phi[G.at(x, y, z)];
    g0 = g1 = g2 = 0.0;
    sfor<1, 15>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        using D = Dir<q>;
        const double v = phi[G.at(x + D::cx, y + D::cy, z + D::cz)];
        lap = lap + v;
        g0 = acc<D::cx>(g0, v);
        g1 = acc<D::cy>(g1, v);
        g2 = acc<D::cz>(g2, v);
    });
    lap = lap / 5.0;
    g0 = g0 / 10.0;
    g1 = g1 / 10.0;
    g2 = g2 / 10.0;
}

// grad u (gu[k][a] = d_a u_k) at (x,y,z) (various_cal :616-633)
__device__ void u_stencil(const Grid& G, const double* u1, const double* u2, const double* u3, int x, int y, int z,
                          double gu[3][3]) {
    for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++) gu[k][a] = 0.0;
    sfor<1, 15>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        using D = Dir<q>;
        const int64_t t = G.at(x + D::cx, y + D::cy, z + D::cz);
        const double v[3] = {u1[t], u2[t], u3[t]};
        for (int k = 0; k < 3; k++) {
            gu[k][0] = acc<D::cx>(gu[k][0], v[k]);
            gu[k][1] = acc<D::cy>(gu[k][1], v[k]);
            gu[k][2] = acc<D::cz>(gu[k][2], v[k]);
        }
    });
    for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++) gu[k][a] = gu[k][a] / 10.0;
}

// True when the neighbour (xx, yy, zz) (not wrapped) of node (x, y, z) lies
// outside the node's rank block: program/LES_drop.f90 (:153-157) and
// LES_drop_xy.f90 (:150-153) take every such neighbour from the halo exchanged
// at the start of the step.  B.bx = 0: no decomposition, every neighbour current.
__device__ __forceinline__ bool les_halo(const LesBlocks& B, int x, int y, int z, int xx, int yy, int zz) {
    if (B.bx == 0) return false;
    const int x0 = (x / B.bx) * B.bx, y0 = (y / B.by) * B.by, z0 = (z / B.bz) * B.bz;
    return xx < x0 || xx >= x0 + B.bx || yy < y0 || yy >= y0 + B.by || zz < z0 || zz >= z0 + B.bz;
}

__global__ void init_equilibrium_fp64(Grid G, Phys P, const double* __restrict__ mac, double* __restrict__ f,
                                      double* __restrict__ g) {
    const int64_t N = (int64_t)G.nx * G.ny * G.nz;
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= N) return;
    const int x = s % G.nx, y = (s / G.nx) % G.ny, z = s / ((int64_t)G.nx * G.ny);
    NodeEq<double> nd;
    nd.phi = mac[s];
    nd.p = mac[N + s];
    nd.u1 = mac[2 * N + s];
    nd.u2 = mac[3 * N + s];
    nd.u3 = mac[4 * N + s];
    double lap;
    phi_stencil(G, mac, x, y, z, lap, nd.g0, nd.g1, nd.g2);
    const double ph = nd.phi;
    const double p0 = ph * P.T / (1.0 - P.b * ph) - P.a * (ph * ph);
    nd.A = p0 - P.kf * ph * lap - P.kf6 * (nd.g0 * nd.g0 + nd.g1 * nd.g1 + nd.g2 * nd.g2);
    const double usq = nd.u1 * nd.u1 + nd.u2 * nd.u2 + nd.u3 * nd.u3;
    const double s2h = 1.5 * (nd.g0 * nd.g0 + nd.g1 * nd.g1 + nd.g2 * nd.g2);
    sfor<0, 15>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const double gt = korteweg<q>(nd.g0, nd.g1, nd.g2, s2h);
        f[q * N + s] = feq<q>(nd, gt, ekappa<q>(P.kf));
        g[q * N + s] = geq<q>(nd, usq, gt, ekappa<q>(P.kg));
    });
}

__global__ void moments_fp64(int64_t N, const double* __restrict__ f, const double* __restrict__ g,
                             double* __restrict__ mac, int which) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= N) return;
    double ph = f[s], pr = g[s] / 3.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
    sfor<1, 15>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        using D = Dir<q>;
        const double gv = g[q * N + s];
        ph = ph + f[q * N + s];
        pr = pr + gv / 3.0;
        m1 = acc<D::cx>(m1, gv);
        m2 = acc<D::cy>(m2, gv);
        m3 = acc<D::cz>(m3, gv);
    });
    if (which & 1) mac[s] = ph;
    if (which & 2) {
        mac[N + s] = pr;
        mac[2 * N + s] = m1;
        mac[3 * N + s] = m2;
        mac[4 * N + s] = m3;
    }
}

// LES tau tiles of LES_TX x LES_TY x LES_TZ nodes, one node per thread
constexpr int LES_TX = 32, LES_TY = 4, LES_TZ = 2;
// periodic index of v in [-1, n + tile): one conditional add or subtract, or a
// full modulo where a tile is longer than the grid (small)
__device__ __forceinline__ int wrap_any(int v, int n, bool small) {
    if (small) return ((v % n) + n) % n;
    return v < 0 ? v + n : (v >= n ? v - n : v);
}

// tau_g for the next step from the state a step left (mac: phi, p, u of this
// step; mac_old: the state the step started from) and that step's phi range:
// grad u as grad_u_cal takes it at the end of physics (:447-477, :660), with
// the neighbours outside a node's rank block from mac_old (les_halo), then
// nu_e = Cs^2 Delta^2 sqrt(2 S:S) with S_ab = 0.5 (d_b u_a + d_a u_b) (strain_cal
// :479-499, :661-677; amdflang lowers **0.5d0 to sqrt) and tau = 0.5 + 3 (nu + nu_e)
// (:679-690).  mac_old == mac gives the tau of a state with current halos (ini_op).
// u (and, with rank blocks, the old u) of a tile and its 1-node halo are staged
// in LDS once per workgroup.
template <bool STALE>
__global__ __launch_bounds__(LES_TX* LES_TY* LES_TZ) void les_tau_fp64(Grid G, Phys P, const double* __restrict__ mac,
                                                                       const double* __restrict__ mac_old, LesBlocks B,
                                                                       const uint64_t* __restrict__ range, int slot,
                                                                       double* __restrict__ tau, int tiles_x,
                                                                       int tiles_y) {
    constexpr int HX = LES_TX + 2, HY = LES_TY + 2, HZ = LES_TZ + 2, H = HX * HY * HZ, NT = LES_TX * LES_TY * LES_TZ;
    __shared__ double su[3][H];
    __shared__ double so[STALE ? 3 : 1][STALE ? H : 1];
    const int64_t N = (int64_t)G.nx * G.ny * G.nz;
    const int tile = blockIdx.x;
    const int x0 = (tile % tiles_x) * LES_TX, y0 = ((tile / tiles_x) % tiles_y) * LES_TY,
              z0 = (tile / (tiles_x * tiles_y)) * LES_TZ;
    const bool small = G.nx <= LES_TX || G.ny <= LES_TY || G.nz <= LES_TZ;
    for (int h = threadIdx.x; h < H; h += NT) {
        const int hx = h % HX, hy = (h / HX) % HY, hz = h / (HX * HY);
        const int64_t g = wrap_any(x0 + hx - 1, G.nx, small) +
                          (int64_t)G.nx * (wrap_any(y0 + hy - 1, G.ny, small) +
                                           (int64_t)G.ny * wrap_any(z0 + hz - 1, G.nz, small));
        for (int k = 0; k < 3; k++) {
            su[k][h] = mac[(2 + k) * N + g];
            if constexpr (STALE) so[k][h] = mac_old[(2 + k) * N + g];
        }
    }
    __syncthreads();
    const int tx = threadIdx.x % LES_TX, ty = (threadIdx.x / LES_TX) % LES_TY, tz = threadIdx.x / (LES_TX * LES_TY);
    const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
    if (x >= G.nx || y >= G.ny || z >= G.nz) return;
    const int64_t s = x + (int64_t)G.nx * (y + (int64_t)G.ny * z);
    double gu[3][3];
    for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++) gu[k][a] = 0.0;
    sfor<1, 15>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        using D = Dir<q>;
        const int h = (tx + 1 + D::cx) + HX * ((ty + 1 + D::cy) + HY * (tz + 1 + D::cz));
        const bool old = STALE && les_halo(B, x, y, z, x + D::cx, y + D::cy, z + D::cz);
        for (int k = 0; k < 3; k++) {
            const double v = old ? so[STALE ? k : 0][STALE ? h : 0] : su[k][h];
            gu[k][0] = acc<D::cx>(gu[k][0], v);
            gu[k][1] = acc<D::cy>(gu[k][1], v);
            gu[k][2] = acc<D::cz>(gu[k][2], v);
        }
    });
    for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++) gu[k][a] = gu[k][a] / 10.0;
    // S_ab = 0.5 (d_b u_a + d_a u_b), nu_e = Cs^2 Delta^2 sqrt(2 S:S)  (:671-675, strain :481-501)
    double sum = 0.0;
    for (int b = 0; b < 3; b++)
        for (int a = 0; a < 3; a++) {
            const double sab = 0.5 * (gu[a][b] + gu[b][a]);
            sum = sum + sab * sab;
        }
    const double nue = P.les_c * sqrt(2.0 * sum);
    const double rmax = from_ordered_bits(range[2 * slot]), rmin = from_ordered_bits(range[2 * slot + 1]);
    const double rg = rmax - rmin;
    double frac = (mac[s] - rmin) / rg;
    if (rg == 0.0) frac = 0.0;
    const double nu = frac * P.dnu + P.nu1;
    tau[s] = 0.5 + 3.0 * (nu + nue);
}

// The same tau with every neighbour current (no rank blocks), marching in z: a
// workgroup owns a ZM_TX x ZM_TY column of nodes over ZM_ZC planes and keeps the
// planes z-1, z, z+1 of u (with a one-node x-y halo) in a ring of three LDS
// planes.  While plane z is computed, the loads of plane z+2 are already in
// flight into registers (stored into the ring slot of plane z-1 afterwards), so
// their latency overlaps the stencil work.  u is read from L2 / HBM about
// (66 x 10)/(64 x 8) (ZM_ZC + 2)/ZM_ZC times.  Sums are formed in the order of
// les_tau_fp64 (the divisions by 10 correctly rounded), so both give identical tau.
constexpr int ZM_TX = 64, ZM_TY = 8, ZM_ZC = 16;
constexpr int ZM_HX = ZM_TX + 2, ZM_HP = ZM_HX * (ZM_TY + 2), ZM_NT = ZM_TX * ZM_TY;
constexpr int ZM_PER = (ZM_HP + ZM_NT - 1) / ZM_NT;  // plane nodes per thread

template <typename T> struct ZmPlane {
    T v[ZM_PER][3];
};

template <typename T>
__device__ __forceinline__ ZmPlane<T> zm_fetch(const Grid& G, const T* __restrict__ mac, int64_t N, int x0, int y0,
                                               int zg, bool small) {
    ZmPlane<T> r;
    const int64_t zoff = (int64_t)G.nx * G.ny * wrap_any(zg, G.nz, small);
#pragma unroll
    for (int i = 0; i < ZM_PER; i++) {
        // unconditional loads (index clamped into the plane) keep the wait a straight count
        const int h0 = threadIdx.x + i * ZM_NT, h = h0 < ZM_HP ? h0 : ZM_HP - 1;
        const int64_t g = wrap_any(x0 + h % ZM_HX - 1, G.nx, small) +
                          (int64_t)G.nx * wrap_any(y0 + h / ZM_HX - 1, G.ny, small) + zoff;
#pragma unroll
        for (int k = 0; k < 3; k++) r.v[i][k] = mac[(2 + k) * N + g];
    }
    return r;
}

template <typename T> __device__ __forceinline__ void zm_store(const ZmPlane<T>& r, T (*dst)[ZM_HP]) {
#pragma unroll
    for (int i = 0; i < ZM_PER; i++) {
        const int h = threadIdx.x + i * ZM_NT;
        if (h < ZM_HP)
#pragma unroll
            for (int k = 0; k < 3; k++) dst[k][h] = r.v[i][k];
    }
}

__global__ __launch_bounds__(ZM_NT) void les_tau_fp64_zm(Grid G, Phys P, const double* __restrict__ mac,
                                                          const uint64_t* __restrict__ range, int slot,
                                                          double* __restrict__ tau, int tiles_x, int tiles_y) {
    __shared__ double su[3][3][ZM_HP];  // [ring slot][component][plane node]
    const int64_t N = (int64_t)G.nx * G.ny * G.nz;
    const int x0 = (blockIdx.x % tiles_x) * ZM_TX, y0 = ((blockIdx.x / tiles_x) % tiles_y) * ZM_TY;
    const int zs = (blockIdx.x / (tiles_x * tiles_y)) * ZM_ZC, ze = min(zs + ZM_ZC, G.nz);
    const bool small = G.nx <= ZM_TX || G.ny <= ZM_TY || G.nz <= 2;
    const int tx = threadIdx.x % ZM_TX, ty = threadIdx.x / ZM_TX;
    const int x = x0 + tx, y = y0 + ty;
    const bool in = x < G.nx && y < G.ny;
    const double rmax = from_ordered_bits(range[2 * slot]), rmin = from_ordered_bits(range[2 * slot + 1]);
    const double rg = rmax - rmin;
    // ring slot of plane z' is (z' - zs + 1) mod 3
    zm_store(zm_fetch<double>(G, mac, N, x0, y0, zs - 1, small), su[0]);
    zm_store(zm_fetch<double>(G, mac, N, x0, y0, zs, small), su[1]);
    zm_store(zm_fetch<double>(G, mac, N, x0, y0, zs + 1, small), su[2]);
    const int64_t col = (int64_t)(in ? x : 0) + (int64_t)G.nx * (in ? y : 0);
    double phi_next = mac[col + (int64_t)G.nx * G.ny * zs];
    for (int z = zs; z < ze; z++) {
        const bool more = z + 1 < ze;  // plane z+2 is needed by plane z+1
        ZmPlane<double> next;
        if (more) next = zm_fetch<double>(G, mac, N, x0, y0, z + 2, small);
        const double phi = phi_next;
        if (more) phi_next = mac[col + (int64_t)G.nx * G.ny * (z + 1)];
        __syncthreads();
        if (in) {
            const int rz[3] = {(z - zs) % 3, (z - zs + 1) % 3, (z - zs + 2) % 3};  // planes z-1, z, z+1
            double gu[3][3];
            for (int k = 0; k < 3; k++)
                for (int a = 0; a < 3; a++) gu[k][a] = 0.0;
            sfor<1, 15>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                using D = Dir<q>;
                const int h = (tx + 1 + D::cx) + ZM_HX * (ty + 1 + D::cy);
                for (int k = 0; k < 3; k++) {
                    const double v = su[rz[1 + D::cz]][k][h];
                    gu[k][0] = acc<D::cx>(gu[k][0], v);
                    gu[k][1] = acc<D::cy>(gu[k][1], v);
                    gu[k][2] = acc<D::cz>(gu[k][2], v);
                }
            });
            for (int k = 0; k < 3; k++)
                for (int a = 0; a < 3; a++) gu[k][a] = div_const(gu[k][a], 10.0, 0.1);
            double sum = 0.0;
            for (int b = 0; b < 3; b++)
                for (int a = 0; a < 3; a++) {
                    const double sab = 0.5 * (gu[a][b] + gu[b][a]);
                    sum = sum + sab * sab;
                }
            const double nue = P.les_c * sqrt(2.0 * sum);
            double frac = (phi - rmin) / rg;
            if (rg == 0.0) frac = 0.0;
            const double nu = frac * P.dnu + P.nu1;
            tau[col + (int64_t)G.nx * G.ny * z] = 0.5 + 3.0 * (nu + nue);
        }
        __syncthreads();
        if (more) zm_store(next, su[(z - zs) % 3]);  // the slot of plane z-1
    }
}

// Diagnostics: per-workgroup partial sums, reduced in fixed order on the host.
// [0] sum phi  [1] sum 0.5|u|^2  [2] sum (d_a u_b + d_b u_a)^2  [3] sum |grad phi|
// [4] sum psi(phi)  [5] sum p  [6..8] sum u  [9] sum_ab d_a phi d_b phi S_ab
__global__ void reduce_fp64(Grid G, Phys P, const double* __restrict__ mac, double* __restrict__ partial) {
    __shared__ double sh[NPART][4];
    const int64_t N = (int64_t)G.nx * G.ny * G.nz;
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double v[NPART] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (s < N) {
        const int x = s % G.nx, y = (s / G.nx) % G.ny, z = s / ((int64_t)G.nx * G.ny);
        const double ph = mac[s], pr = mac[N + s];
        const double u1 = mac[2 * N + s], u2 = mac[3 * N + s], u3 = mac[4 * N + s];
        double lap, g0, g1, g2, gu[3][3];
        phi_stencil(G, mac, x, y, z, lap, g0, g1, g2);
        u_stencil(G, mac + 2 * N, mac + 3 * N, mac + 4 * N, x, y, z, gu);
        double e = 0.0;
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++) {
                const double sab = gu[a][b] + gu[b][a];
                e += sab * sab;
            }
        v[0] = ph;
        v[1] = 0.5 * (u1 * u1 + u2 * u2 + u3 * u3);
        v[2] = e;
        v[3] = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
        v[4] = ph * P.T * log(ph / (1.0 - P.b * ph)) - P.a * ph * ph;
        v[5] = pr;
        v[6] = u1;
        v[7] = u2;
        v[8] = u3;
        // interface-strain interaction, ref drop_analysis.f90 korteweg_cal /
        // strain_cal / interaction_cal (:830-882): beta outer, alpha inner
        const double gp[3] = {g0, g1, g2};
        double it = 0.0;
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++) it = it + (gp[a] * gp[b]) * (0.5 * (gu[a][b] + gu[b][a]));
        v[9] = it;
    }
    for (int k = 0; k < NPART; k++) {
        double t = v[k];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        v[k] = t;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < NPART; k++) sh[k][wave] = v[k];
    __syncthreads();
    if (threadIdx.x < NPART) {
        const int k = threadIdx.x;
        partial[blockIdx.x * NPART + k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
    }
}

__global__ void range_set(uint64_t* range, int slot, double vmax, double vmin) {
    range[2 * slot] = ordered_bits(vmax);
    range[2 * slot + 1] = ordered_bits(vmin);
}

// ------------------------------------------------------------------ launchers
static inline int64_t nblk(int64_t n, int bs) { return (n + bs - 1) / bs; }

hipError_t launch_init_equilibrium_fp64(int nx, int ny, int nz, const Phys& ph, const double* mac, double* f,
                                        double* g, hipStream_t st) {
    Grid G{nx, ny, nz};
    const int64_t N = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(init_equilibrium_fp64, dim3(nblk(N, 256)), dim3(256), 0, st, G, ph, mac, f, g);
    return hipGetLastError();
}

hipError_t launch_moments_fp64(int nx, int ny, int nz, const double* f, const double* g, double* mac, int which,
                               hipStream_t st) {
    const int64_t N = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(moments_fp64, dim3(nblk(N, 256)), dim3(256), 0, st, N, f, g, mac, which);
    return hipGetLastError();
}

hipError_t launch_les_tau_fp64(int nx, int ny, int nz, const Phys& ph, const double* mac, const double* mac_old,
                               const LesBlocks& blk, const uint64_t* range, int slot, double* tau, hipStream_t st) {
    Grid G{nx, ny, nz};
    const int tx = (nx + LES_TX - 1) / LES_TX, ty = (ny + LES_TY - 1) / LES_TY, tz = (nz + LES_TZ - 1) / LES_TZ;
    const dim3 grid((unsigned)(tx * ty * tz)), block(LES_TX * LES_TY * LES_TZ);
    if (blk.bx > 0 && mac_old != mac) {
        hipLaunchKernelGGL(les_tau_fp64<true>, grid, block, 0, st, G, ph, mac, mac_old, blk, range, slot, tau, tx, ty);
    } else {
        const int mx = (nx + ZM_TX - 1) / ZM_TX, my = (ny + ZM_TY - 1) / ZM_TY, mz = (nz + ZM_ZC - 1) / ZM_ZC;
        hipLaunchKernelGGL(les_tau_fp64_zm, dim3((unsigned)(mx * my * mz)), dim3(ZM_TX * ZM_TY), 0, st, G, ph, mac, range,
                           slot, tau, mx, my);
    }
    return hipGetLastError();
}

hipError_t launch_reduce_fp64(int nx, int ny, int nz, const Phys& ph, const double* mac, double* partial,
                              int* nblocks, hipStream_t st) {
    Grid G{nx, ny, nz};
    const int64_t N = (int64_t)nx * ny * nz;
    *nblocks = (int)nblk(N, 256);
    hipLaunchKernelGGL(reduce_fp64, dim3(*nblocks), dim3(256), 0, st, G, ph, mac, partial);
    return hipGetLastError();
}

hipError_t launch_range_set(uint64_t* range, int slot, double vmax, double vmin, hipStream_t st) {
    hipLaunchKernelGGL(range_set, dim3(1), dim3(1), 0, st, range, slot, vmax, vmin);
    return hipGetLastError();
}

}  // namespace drp
