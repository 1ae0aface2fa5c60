// budget.hip — derived fields of the droplet analysis programs and the per-step
// kinetic-energy budget of program/energy_budget.f90, on the GPU.
//
//   drp_get_derived          chemical potential (program/drop_analysis.f90:413-425),
//                            free-energy density (:400-411), advection u . grad phi
//                            (:427-441), with the D3Q15 lap_cal / grad_cal stencils
//   drp_step_energy_budget   one step, then the seven budget terms of
//                            program/energy_budget.f90:1251-1700 from the fields the
//                            step started from (the other ping-pong buffer) and the
//                            fields it produced
//
// Neither is on the step's hot path (energy_budget.f90 evaluates the budget after
// every step).  The budget runs in two passes over the nodes: the first forms the
// per-node terms that need no stencil of a derived quantity (left1, right3,
// right5) and stores K = |u|^2/2, V_Y = u_X S_XY and grad phi (7 doubles per
// node, allocated on the first call); the second takes the D3Q15 gradients of
// those, of p and of the Korteweg tensor G_XY formed at each stencil node (left2,
// right1, right2, right4).  Each workgroup leaves its partial sums; the host adds
// them with Neumaier compensation in workgroup order.  The per-node expressions
// keep the reference's evaluation order (FMA contraction is off for this file),
// except that right2 sums d_Y V_Y rather than d_Y (u_X S_XY) term by term.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/droplet.h"
#include "d3q15.h"
#include "kernels.h"
#include "lattice_impl.h"

namespace drp {
namespace {

#define HIP_TRYB(expr)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return drp_fail(DRP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

template <int Q0, int Q1, typename Fn> __device__ __forceinline__ void sfor(Fn&& fn) {
    if constexpr (Q0 < Q1) {
        fn(std::integral_constant<int, Q0>{});
        sfor<Q0 + 1, Q1>(fn);
    }
}

struct BGrid {
    int nx, ny, nz;
    __device__ __forceinline__ int64_t at(int x, int y, int z) const {
        x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
        y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
        z = z < 0 ? z + nz : (z >= nz ? z - nz : z);
        return x + (int64_t)nx * (y + (int64_t)ny * z);
    }
};

// component A of direction Q
template <int Q, int A> constexpr int cdir() { return A == 0 ? Dir<Q>::cx : (A == 1 ? Dir<Q>::cy : Dir<Q>::cz); }

// lap_cal and grad_cal of phi, grad of u (gu[k][a] = d_a u_k) at one node
__device__ void stencils(const BGrid& G, const double* phi, const double* u1, const double* u2, const double* u3,
                         int x, int y, int z, double& lap, double gp[3], double gu[3][3]) {
    lap = -14.0 * phi[G.at(x, y, z)];
    for (int a = 0; a < 3; a++) {
        gp[a] = 0.0;
        gu[0][a] = gu[1][a] = gu[2][a] = 0.0;
    }
    sfor<1, 15>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        using D = Dir<q>;
        const int64_t t = G.at(x + D::cx, y + D::cy, z + D::cz);
        const double v = phi[t];
        lap = lap + v;
        gp[0] = acc<D::cx>(gp[0], v);
        gp[1] = acc<D::cy>(gp[1], v);
        gp[2] = acc<D::cz>(gp[2], v);
        if (u1) {
            const double w[3] = {u1[t], u2[t], u3[t]};
            for (int k = 0; k < 3; k++) {
                gu[k][0] = acc<D::cx>(gu[k][0], w[k]);
                gu[k][1] = acc<D::cy>(gu[k][1], w[k]);
                gu[k][2] = acc<D::cz>(gu[k][2], w[k]);
            }
        }
    });
    lap = lap / 5.0;
    for (int a = 0; a < 3; a++) {
        gp[a] = gp[a] / 10.0;
        for (int k = 0; k < 3; k++) gu[k][a] = gu[k][a] / 10.0;
    }
}

__global__ __launch_bounds__(256) void derived_fp64(BGrid G, Phys P, const double* __restrict__ mac, int kind,
                                                    double* __restrict__ out) {
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int x = (int)(s % G.nx), y = (int)((s / G.nx) % G.ny), z = (int)(s / ((int64_t)G.nx * G.ny));
    const double ph = mac[s];
    if (kind == DRP_DERIVED_FREE_ENERGY) {
        out[s] = ph * P.T * log(ph / (1.0 - P.b * ph)) - P.a * ph * ph;
        return;
    }
    double lap, gp[3], gu[3][3];
    stencils(G, mac, nullptr, nullptr, nullptr, x, y, z, lap, gp, gu);
    if (kind == DRP_DERIVED_CHEMICAL_POTENTIAL) {
        out[s] = P.T * log(ph / (1.0 - P.b * ph)) + P.T / (1.0 - P.b * ph) - 2.0 * P.a * ph - P.kg * lap;
    } else {
        out[s] = mac[2 * n + s] * gp[0] + mac[3 * n + s] * gp[1] + mac[4 * n + s] * gp[2];
    }
}

constexpr int BPART = 8;  // partial sums per workgroup: the seven terms and a pad
// Work space, doubles per node: K = |u|^2/2 of the new u, V_Y = sum_X u_X S_XY
// (new u, strain of the old fields) and grad phi of the old fields.  right2 is
// taken as sum_Y d_Y V_Y, the reference's sum_Y sum_X d_Y (u_X S_XY) with the
// sum over X moved inside the (linear) stencil: equal but for rounding, and 3
// arrays instead of 9.  G_XY is formed at each stencil node from grad phi there
// with the expression pass 1 would use.
constexpr int BWORK = 7;
enum { W_K = 0, W_V = 1, W_GP = 4 };

// Workgroup brick of the budget passes: 64 x 4 columns (a wavefront per x row)
// that march over BZC planes, keeping the planes their stencils need in an LDS
// ring; bricks are dealt to the XCDs in contiguous runs (xcd_tile), so bricks
// adjacent in y read their shared halo rows through one L2.
constexpr int BTX = 64, BTY = 4, BZC = 16;
constexpr int BT = BTX * BTY, NW = BT / 64;  // threads and wavefronts per workgroup

struct BudgetArgs {
    BGrid G;
    Phys P;
    const double* old_mac;
    const double* new_mac;
    const uint64_t* range;  // the slot of the new phi range
    const double *fx, *fy, *fz;                // body force field of the step, or null
    const double *sinx, *cosx, *sinz, *cosz;   // Taylor-Green tables of the step, or null
    double tg_amp;
    double* work;     // [BWORK][ws]
    int64_t ws;       // stride of the work arrays (n)
    double* partial;  // [blocks][BPART]
};

__device__ __forceinline__ double nu_new(const BudgetArgs& a, double phn) {
    const double rmax = from_ordered_bits(a.range[0]), rmin = from_ordered_bits(a.range[1]);
    const double rg = rmax - rmin;
    const double frac = rg == 0.0 ? 0.0 : (phn - rmin) / rg;
    return frac * a.P.dnu + a.P.nu1;
}

// workgroup partial sums of v[k] for the columns in [k0, k1)
template <int K0, int K1> __device__ __forceinline__ void block_partials(double (&v)[BPART], double* partial) {
    __shared__ double sh[BPART][NW];
    for (int k = K0; k < K1; k++) {
        double t = v[k];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        v[k] = t;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = K0; k < K1; k++) sh[k][wave] = v[k];
    __syncthreads();
    if (threadIdx.x >= K0 && threadIdx.x < K1) {
        const int k = threadIdx.x;
        double t = sh[k][0];
        for (int w = 1; w < NW; w++) t += sh[k][w];
        partial[blockIdx.x * BPART + k] = t;
    }
}


__host__ __device__ inline int64_t budget_bricks(int nx, int ny, int nz) {
    return (int64_t)((nx + BTX - 1) / BTX) * ((ny + BTY - 1) / BTY) * ((nz + BZC - 1) / BZC);
}

struct Brick {
    int x, y, z0, z1;
    bool live;  // x < nx and y < ny
};
__device__ __forceinline__ Brick brick(const BGrid& G) {
    const int bx = (G.nx + BTX - 1) / BTX, by = (G.ny + BTY - 1) / BTY;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = t % bx, ty = (t / bx) % by, tz = t / (bx * by);
    Brick b;
    b.x = tx * BTX + (int)(threadIdx.x % BTX);
    b.y = ty * BTY + (int)(threadIdx.x / BTX);
    b.z0 = tz * BZC;
    b.z1 = min(b.z0 + BZC, G.nz);
    b.live = b.x < G.nx && b.y < G.ny;
    return b;
}

// Three z-planes of NA arrays over a brick and its one-node halo in x and y, in
// LDS: a ring that the brick's z-march refills one plane at a time, so that each
// value is read from memory once per brick rather than once per stencil use.
constexpr int RX = BTX + 2, RY = BTY + 2, RN = RX * RY;
template <int NA> constexpr int ring_per_thread() { return (NA * RN + BT - 1) / BT; }

__device__ __forceinline__ int wrap_any(int v, int n) {  // v in [-1, n + BTX]
    v = v < 0 ? v + n : v;
    return v >= n ? v % n : v;
}

// plane z of the NA arrays src(k) over the brick at (x0, y0) and its halo, into
// registers (ring_per_thread<NA>() per thread)
template <int NA, typename Src>
__device__ __forceinline__ void ring_fetch(const BGrid& G, int x0, int y0, int z, Src&& src,
                                           double (&r)[ring_per_thread<NA>()]) {
    const int64_t zo = (int64_t)G.nx * G.ny * wrap_any(z, G.nz);
#pragma unroll
    for (int j = 0; j < ring_per_thread<NA>(); j++) {
        const int i = (int)threadIdx.x + BT * j;
        if (i < NA * RN) {
            const int k = i / RN, c = i - k * RN, ly = c / RX, lx = c - ly * RX;
            const int gx = wrap_any(x0 - 1 + lx, G.nx), gy = wrap_any(y0 - 1 + ly, G.ny);
            r[j] = src(k)[zo + gx + (int64_t)G.nx * gy];
        }
    }
}
template <int NA> __device__ __forceinline__ void ring_store(double* slot, const double (&r)[ring_per_thread<NA>()]) {
#pragma unroll
    for (int j = 0; j < ring_per_thread<NA>(); j++) {
        const int i = (int)threadIdx.x + BT * j;
        if (i < NA * RN) slot[i] = r[j];
    }
}

// lap_cal and grad_cal of phi and grad of u (gu[k][a] = d_a u_k) from a ring of
// the old phi, u1, u2, u3 (the order and arithmetic of stencils())
__device__ __forceinline__ void stencils_ring(const double* const (&sl)[3], double& lap, double gp[3],
                                              double gu[3][3]) {
    lap = -14.0 * sl[1][0];
    for (int a = 0; a < 3; a++) {
        gp[a] = 0.0;
        gu[0][a] = gu[1][a] = gu[2][a] = 0.0;
    }
    sfor<1, 15>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        using D = Dir<q>;
        const double* t = sl[1 + D::cz] + (D::cy * RX + D::cx);
        const double v = t[0];
        lap = lap + v;
        gp[0] = acc<D::cx>(gp[0], v);
        gp[1] = acc<D::cy>(gp[1], v);
        gp[2] = acc<D::cz>(gp[2], v);
        const double w[3] = {t[RN], t[2 * RN], t[3 * RN]};
        for (int k = 0; k < 3; k++) {
            gu[k][0] = acc<D::cx>(gu[k][0], w[k]);
            gu[k][1] = acc<D::cy>(gu[k][1], w[k]);
            gu[k][2] = acc<D::cz>(gu[k][2], w[k]);
        }
        // pin the sums after each neighbour (otherwise every neighbour's reads
        // are issued first and held)
        asm volatile("" : "+v"(lap), "+v"(gp[0]), "+v"(gp[1]), "+v"(gp[2]));
#pragma unroll
        for (int k = 0; k < 3; k++) asm volatile("" : "+v"(gu[k][0]), "+v"(gu[k][1]), "+v"(gu[k][2]));
    });
    lap = lap / 5.0;
    for (int a = 0; a < 3; a++) {
        gp[a] = gp[a] / 10.0;
        for (int k = 0; k < 3; k++) gu[k][a] = gu[k][a] / 10.0;
    }
}

__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(4))) void budget_pass1(BudgetArgs a) {
    const BGrid& G = a.G;
    const Phys& P = a.P;
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const Brick b = brick(G);
    // ring of the old phi, u1, u2, u3
    constexpr int NA1 = 4;
    __shared__ double ring[3][NA1 * RN];
    const int lx = (int)threadIdx.x % BTX, ly = (int)threadIdx.x / BTX;
    const int x0 = b.x - lx, y0 = b.y - ly;
    const auto src = [&](int k) -> const double* { return a.old_mac + (k == 0 ? 0 : (k + 1) * n); };
    double r[ring_per_thread<NA1>()];
    ring_fetch<NA1>(G, x0, y0, b.z0 - 1, src, r);
    ring_store<NA1>(ring[0], r);
    ring_fetch<NA1>(G, x0, y0, b.z0, src, r);
    ring_store<NA1>(ring[1], r);
    ring_fetch<NA1>(G, x0, y0, b.z0 + 1, src, r);
    const int me = (ly + 1) * RX + (lx + 1);
    double v[BPART] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int z = b.z0; z < b.z1; z++) {
        // planes z - 1, z, z + 1 sit in slots j, j + 1, j + 2 (mod 3)
        const int j = z - b.z0;
        ring_store<NA1>(ring[(j + 2) % 3], r);
        __syncthreads();
        if (z + 2 <= b.z1) ring_fetch<NA1>(G, x0, y0, z + 2, src, r);
        // every thread reaches each barrier at the same point: a barrier inside the
        // branch would run twice in a wavefront whose lanes take both paths
        if (b.live) {
            const int x = b.x, y = b.y;
            const int64_t s = x + (int64_t)G.nx * (y + (int64_t)G.ny * z);
            const double* nm = a.new_mac;
            const double* sl[3] = {ring[j % 3] + me, ring[(j + 1) % 3] + me, ring[(j + 2) % 3] + me};
            double lap, gp[3], gu[3][3];
            stencils_ring(sl, lap, gp, gu);
            const double ph = sl[1][0];
            const double p0 = ph * P.T / (1.0 - P.b * ph) - P.a * (ph * ph);
            const double gsq = gp[0] * gp[0] + gp[1] * gp[1] + gp[2] * gp[2];
            const double phn = nm[s];
            const double un[3] = {nm[2 * n + s], nm[3 * n + s], nm[4 * n + s]};
            // left1 (kinetic_former with phi_procs already the new phi)
            const double former = p0 - P.kg * phn * lap - 1.0 / 6.0 * P.kg * gsq;
            v[0] += fabs(nm[n + s] - former);
            double S[3][3];
            for (int bb = 0; bb < 3; bb++)
                for (int c = 0; c < 3; c++) S[c][bb] = 0.5 * (gu[c][bb] + gu[bb][c]);
            const double nu = nu_new(a, phn);
            double r3 = 0.0;
            for (int bb = 0; bb < 3; bb++)
                for (int c = 0; c < 3; c++) r3 = r3 + S[c][bb] * S[c][bb];
            v[4] += r3 * (-2.0) * nu;
            double F[3] = {0.0, 0.0, 0.0};
            if (a.fx) {
                F[0] = a.fx[s];
                F[1] = a.fy[s];
                F[2] = a.fz[s];
            } else if (a.sinx) {
                F[0] = a.tg_amp * a.sinx[x] * a.cosz[z];
                F[2] = -(a.tg_amp * a.cosx[x] * a.sinz[z]);
            }
            v[6] += un[0] * F[0] + un[1] * F[1] + un[2] * F[2];
            double* w = a.work;
            const int64_t ws = a.ws;
            w[W_K * ws + s] = 0.5 * (un[0] * un[0] + un[1] * un[1] + un[2] * un[2]);
            for (int Y = 0; Y < 3; Y++) w[(W_V + Y) * ws + s] = un[0] * S[0][Y] + un[1] * S[1][Y] + un[2] * S[2][Y];
            for (int X = 0; X < 3; X++) w[(W_GP + X) * ws + s] = gp[X];
        }
        __syncthreads();
    }
    block_partials<0, BPART>(v, a.partial);
}

// ring arrays of pass 2: K, V_0..2, grad phi 0..2 (work space), p (new fields)
constexpr int NA2 = 8;
enum { R_K = 0, R_V = 1, R_GP = 4, R_P = 7 };

__global__ __launch_bounds__(BT) void budget_pass2(BudgetArgs a) {
    const BGrid& G = a.G;
    const Phys& P = a.P;
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const Brick b = brick(G);
    __shared__ double ring[3][NA2 * RN];
    const int lx = (int)threadIdx.x % BTX, ly = (int)threadIdx.x / BTX;
    const int x0 = b.x - lx, y0 = b.y - ly;
    const auto src = [&](int k) -> const double* { return k < R_P ? a.work + k * a.ws : a.new_mac + n; };
    double r[ring_per_thread<NA2>()];
    ring_fetch<NA2>(G, x0, y0, b.z0 - 1, src, r);
    ring_store<NA2>(ring[0], r);
    ring_fetch<NA2>(G, x0, y0, b.z0, src, r);
    ring_store<NA2>(ring[1], r);
    ring_fetch<NA2>(G, x0, y0, b.z0 + 1, src, r);
    double v[BPART] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int me = (ly + 1) * RX + (lx + 1);
    for (int z = b.z0; z < b.z1; z++) {
        // planes z - 1, z, z + 1 sit in slots j, j + 1, j + 2 (mod 3)
        const int j = z - b.z0;
        ring_store<NA2>(ring[(j + 2) % 3], r);
        __syncthreads();
        if (z + 2 <= b.z1) ring_fetch<NA2>(G, x0, y0, z + 2, src, r);
        if (b.live) {
            const double* sl[3] = {ring[j % 3] + me, ring[(j + 1) % 3] + me, ring[(j + 2) % 3] + me};
            // grad_cal sums (sum over the directions of c_A v(x + c), in direction
            // order) of K, p and G_XA along every axis A and of V_A along A, in one
            // pass over the 14 neighbours; G_XA is formed at the neighbour
            double aK[3] = {0.0, 0.0, 0.0}, aP[3] = {0.0, 0.0, 0.0}, aV[3] = {0.0, 0.0, 0.0};
            double aG[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
            sfor<1, 15>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                using D = Dir<q>;
                const double* t = sl[1 + D::cz] + (D::cy * RX + D::cx);
                const double Kt = t[R_K * RN], pt = t[R_P * RN];
                const double gt[3] = {t[R_GP * RN], t[(R_GP + 1) * RN], t[(R_GP + 2) * RN]};
                const double gsq = gt[0] * gt[0] + gt[1] * gt[1] + gt[2] * gt[2];
                // G_XA = 4.5 g_X g_A - 1.5 |g|^2 delta_XA; off the diagonal the
                // reference subtracts (1.5 |g|^2) * 0 = +0, which changes no finite value
                const double h[3] = {4.5 * gt[0], 4.5 * gt[1], 4.5 * gt[2]};
                const double gsq15 = 1.5 * gsq;
                sfor<0, 3>([&](auto Ac) {
                    constexpr int A = decltype(Ac)::value;
                    constexpr int c = cdir<q, A>();
                    if constexpr (c != 0) {
                        aK[A] = acc<c>(aK[A], Kt);
                        aP[A] = acc<c>(aP[A], pt);
                        aV[A] = acc<c>(aV[A], t[(R_V + A) * RN]);
                        sfor<0, 3>([&](auto Xc) {
                            constexpr int X = decltype(Xc)::value;
                            const double GXA = X == A ? h[X] * gt[A] - gsq15 : h[X] * gt[A];
                            aG[X][A] = acc<c>(aG[X][A], GXA);
                        });
                    }
                });
                // pin the sums after each neighbour: left alone, the compiler reads
                // all 14 neighbours first and holds them (over 300 VGPRs)
#pragma unroll
                for (int A = 0; A < 3; A++) {
                    asm volatile("" : "+v"(aK[A]), "+v"(aP[A]), "+v"(aV[A]));
                    asm volatile("" : "+v"(aG[0][A]), "+v"(aG[1][A]), "+v"(aG[2][A]));
                }
            });
            const int64_t s = b.x + (int64_t)G.nx * (b.y + (int64_t)G.ny * z);
            const double* nm = a.new_mac;
            const double u[3] = {nm[2 * n + s], nm[3 * n + s], nm[4 * n + s]};
            v[1] += u[0] * (aK[0] / 10.0) + u[1] * (aK[1] / 10.0) + u[2] * (aK[2] / 10.0);
            const double r1 = u[0] * (aP[0] / 10.0) + u[1] * (aP[1] / 10.0) + u[2] * (aP[2] / 10.0);
            v[2] += -1.0 * r1;
            double r2 = 0.0, r4 = 0.0;
            for (int Y = 0; Y < 3; Y++) {
                r2 = r2 + aV[Y] / 10.0;
                for (int X = 0; X < 3; X++) r4 = r4 + u[X] * (aG[X][Y] / 10.0);
            }
            v[3] += r2 * 2.0 * nu_new(a, nm[s]);
            v[5] += r4 * (-1.0) * (2.0 / 9.0) * P.kg;
        }
        __syncthreads();
    }
    // pass 1 wrote columns 0, 4, 6 (and zeros in the others); add this pass's columns
    __shared__ double sh[BPART][NW];
    for (int k = 1; k < 6; k++) {
        double t = v[k];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        v[k] = t;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 1; k < 6; k++) sh[k][wave] = v[k];
    __syncthreads();
    if (threadIdx.x == 1 || threadIdx.x == 2 || threadIdx.x == 3 || threadIdx.x == 5) {
        const int k = threadIdx.x;
        double t = sh[k][0];
        for (int w = 1; w < NW; w++) t += sh[k][w];
        a.partial[blockIdx.x * BPART + k] = t;
    }
}

}  // namespace

hipError_t launch_derived_fp64(int nx, int ny, int nz, const Phys& ph, const double* mac, int kind, double* out,
                               hipStream_t st) {
    const int64_t n = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(derived_fp64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, BGrid{nx, ny, nz}, ph, mac,
                       kind, out);
    return hipGetLastError();
}

}  // namespace drp

using namespace drp;

extern "C" int drp_get_derived(drp_lattice* L, int kind, double* host, size_t n_) {
    if (!L || !host) return drp_fail(DRP_ERR_ARG, "null argument");
    if (kind != DRP_DERIVED_CHEMICAL_POTENTIAL && kind != DRP_DERIVED_FREE_ENERGY && kind != DRP_DERIVED_ADVECTION)
        return drp_fail(DRP_ERR_ARG, "unknown derived field %d", kind);
    if (n_ != (size_t)L->n) return drp_fail(DRP_ERR_ARG, "size %zu != %lld nodes", n_, (long long)L->n);
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    const double* mac = nullptr;
    if (L->cfg.precision == DRP_FP64) {
        mac = (const double*)L->mac[L->cur];
    } else {
        const int rc = fp32_macro_fp64_view(L, &mac);
        if (rc) return rc;
    }
    double* out = nullptr;
    HIP_TRYB(hipMalloc((void**)&out, (size_t)L->n * sizeof(double)));
    hipError_t e = launch_derived_fp64(L->cfg.nx, L->cfg.ny, L->cfg.nz, L->phys, mac, kind, out, L->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(host, out, (size_t)L->n * sizeof(double), hipMemcpyDeviceToHost, L->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
    const hipError_t f = hipFree(out);
    if (e == hipSuccess) e = f;
    if (e != hipSuccess) return drp_fail(DRP_ERR_HIP, "drp_get_derived: %s", hipGetErrorString(e));
    return DRP_OK;
}

extern "C" int drp_step_energy_budget(drp_lattice* L, double* terms) {
    if (!L || !terms) return drp_fail(DRP_ERR_ARG, "null argument");
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    const drp_config& c = L->cfg;
    if (c.precision != DRP_FP64) return drp_fail(DRP_ERR_UNSUPPORTED, "energy budget: FP64 lattices only");
    if (c.model != DRP_MODEL_FREE_ENERGY) return drp_fail(DRP_ERR_UNSUPPORTED, "energy budget: free-energy model only");
    const int64_t n = L->n;
    const int64_t nb = budget_bricks(c.nx, c.ny, c.nz);
    if (!L->budget) {
        HIP_TRYB(hipMalloc((void**)&L->budget, ((size_t)BWORK * n + (size_t)nb * BPART) * sizeof(double)));
    }
    // the body force the step will use (its pointers do not depend on the ping-pong parity)
    const StepArgs sa = make_step_args(L, L->step + 1);
    int rc = drp_step(L, 1);
    if (rc) return rc;
    BudgetArgs a{};
    a.G = BGrid{c.nx, c.ny, c.nz};
    a.P = L->phys;
    a.old_mac = (const double*)L->mac[1 - L->cur];
    a.new_mac = (const double*)L->mac[L->cur];
    a.range = L->range + 2 * (L->step % 3);
    if (sa.force_kind == FORCE_FIELD) {
        a.fx = sa.fx;
        a.fy = sa.fy;
        a.fz = sa.fz;
    } else if (sa.force_kind == FORCE_TAYLOR_GREEN) {
        a.sinx = sa.sinx;
        a.cosx = sa.cosx;
        a.sinz = sa.sinz;
        a.cosz = sa.cosz;
        a.tg_amp = sa.tg_amp;
    }
    a.work = L->budget;
    a.ws = n;
    a.partial = L->budget + (size_t)BWORK * a.ws;
    hipLaunchKernelGGL(budget_pass1, dim3((unsigned)nb), dim3(BT), 0, L->stream, a);
    HIP_TRYB(hipGetLastError());
    hipLaunchKernelGGL(budget_pass2, dim3((unsigned)nb), dim3(BT), 0, L->stream, a);
    HIP_TRYB(hipGetLastError());
    std::vector<double> part((size_t)nb * BPART);
    HIP_TRYB(hipMemcpyAsync(part.data(), a.partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIP_TRYB(hipStreamSynchronize(L->stream));
    double sum[7] = {0}, comp[7] = {0};
    for (int64_t b = 0; b < nb; b++)
        for (int k = 0; k < 7; k++) {
            const double v = part[(size_t)b * BPART + k], t = sum[k] + v;
            comp[k] += (std::fabs(sum[k]) >= std::fabs(v)) ? (sum[k] - t) + v : (v - t) + sum[k];
            sum[k] = t;
        }
    for (int k = 0; k < 7; k++) terms[k] = sum[k] + comp[k];
    return DRP_OK;
}
