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
// Neither is on the per-step hot path.  The budget runs in two passes over the
// nodes: the first forms the per-node terms that need no stencil of a derived
// quantity (left1, right3, right5) and stores K = |u|^2/2, u_X S_XY and the
// Korteweg tensor G_XY (19 doubles per node, allocated on the first call); the
// second takes the D3Q15 gradients of those and of p (left2, right1, right2,
// right4).  Each workgroup leaves its partial sums; the host adds them with
// Neumaier compensation in workgroup order.  Every per-node expression keeps
// the reference's evaluation order (FMA contraction is off for this file).
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

// grad_cal along axis A: sum over the directions of c_A v(x + c), over 10
template <int A> __device__ __forceinline__ double grad_along(const BGrid& G, const double* v, int x, int y, int z) {
    double s = 0.0;
    sfor<1, 15>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        using D = Dir<q>;
        constexpr int c = A == 0 ? D::cx : (A == 1 ? D::cy : D::cz);
        if constexpr (c != 0) s = acc<c>(s, v[G.at(x + D::cx, y + D::cy, z + D::cz)]);
    });
    return s / 10.0;
}

__device__ __forceinline__ double grad_axis(const BGrid& G, const double* v, int x, int y, int z, int a) {
    return a == 0 ? grad_along<0>(G, v, x, y, z) : (a == 1 ? grad_along<1>(G, v, x, y, z) : grad_along<2>(G, v, x, y, z));
}

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

constexpr int BPART = 8;   // partial sums per workgroup: the seven terms and a pad
constexpr int BWORK = 19;  // doubles per node: K, u_X S_XY (9), G_XY (9)

struct BudgetArgs {
    BGrid G;
    Phys P;
    const double* old_mac;
    const double* new_mac;
    const uint64_t* range;  // the slot of the new phi range
    const double *fx, *fy, *fz;                // body force field of the step, or null
    const double *sinx, *cosx, *sinz, *cosz;   // Taylor-Green tables of the step, or null
    double tg_amp;
    double* work;     // [BWORK][n]
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
    __shared__ double sh[BPART][4];
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
        partial[blockIdx.x * BPART + k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
    }
}

__global__ __launch_bounds__(256) void budget_pass1(BudgetArgs a) {
    const BGrid& G = a.G;
    const Phys& P = a.P;
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double v[BPART] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (s < n) {
        const int x = (int)(s % G.nx), y = (int)((s / G.nx) % G.ny), z = (int)(s / ((int64_t)G.nx * G.ny));
        const double* om = a.old_mac;
        const double* nm = a.new_mac;
        double lap, gp[3], gu[3][3];
        stencils(G, om, om + 2 * n, om + 3 * n, om + 4 * n, x, y, z, lap, gp, gu);
        const double ph = om[s];
        const double p0 = ph * P.T / (1.0 - P.b * ph) - P.a * (ph * ph);
        const double gsq = gp[0] * gp[0] + gp[1] * gp[1] + gp[2] * gp[2];
        const double phn = nm[s];
        const double un[3] = {nm[2 * n + s], nm[3 * n + s], nm[4 * n + s]};
        // left1 (kinetic_former with phi_procs already the new phi)
        const double former = p0 - P.kg * phn * lap - 1.0 / 6.0 * P.kg * gsq;
        v[0] = fabs(nm[n + s] - former);
        double S[3][3];
        for (int b = 0; b < 3; b++)
            for (int c = 0; c < 3; c++) S[c][b] = 0.5 * (gu[c][b] + gu[b][c]);
        const double nu = nu_new(a, phn);
        double r3 = 0.0;
        for (int b = 0; b < 3; b++)
            for (int c = 0; c < 3; c++) r3 = r3 + S[c][b] * S[c][b];
        v[4] = r3 * (-2.0) * nu;
        double F[3] = {0.0, 0.0, 0.0};
        if (a.fx) {
            F[0] = a.fx[s];
            F[1] = a.fy[s];
            F[2] = a.fz[s];
        } else if (a.sinx) {
            F[0] = a.tg_amp * a.sinx[x] * a.cosz[z];
            F[2] = -(a.tg_amp * a.cosx[x] * a.sinz[z]);
        }
        v[6] = un[0] * F[0] + un[1] * F[1] + un[2] * F[2];
        double* w = a.work;
        w[s] = 0.5 * (un[0] * un[0] + un[1] * un[1] + un[2] * un[2]);
        for (int X = 0; X < 3; X++)
            for (int Y = 0; Y < 3; Y++) {
                w[(1 + X * 3 + Y) * n + s] = un[X] * S[X][Y];
                w[(10 + X * 3 + Y) * n + s] = 4.5 * gp[X] * gp[Y] - 1.5 * gsq * (X == Y ? 1.0 : 0.0);
            }
    }
    block_partials<0, BPART>(v, a.partial);
}

__global__ __launch_bounds__(256) void budget_pass2(BudgetArgs a) {
    const BGrid& G = a.G;
    const Phys& P = a.P;
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double v[BPART] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (s < n) {
        const int x = (int)(s % G.nx), y = (int)((s / G.nx) % G.ny), z = (int)(s / ((int64_t)G.nx * G.ny));
        const double* nm = a.new_mac;
        const double* w = a.work;
        const double u[3] = {nm[2 * n + s], nm[3 * n + s], nm[4 * n + s]};
        v[1] = u[0] * grad_along<0>(G, w, x, y, z) + u[1] * grad_along<1>(G, w, x, y, z) +
               u[2] * grad_along<2>(G, w, x, y, z);
        const double* p = nm + n;
        const double r1 = u[0] * grad_along<0>(G, p, x, y, z) + u[1] * grad_along<1>(G, p, x, y, z) +
                          u[2] * grad_along<2>(G, p, x, y, z);
        v[2] = -1.0 * r1;
        double r2 = 0.0, r4 = 0.0;
        for (int Y = 0; Y < 3; Y++)
            for (int X = 0; X < 3; X++) {
                r2 = r2 + grad_axis(G, w + (1 + X * 3 + Y) * n, x, y, z, Y);
                r4 = r4 + u[X] * grad_axis(G, w + (10 + X * 3 + Y) * n, x, y, z, Y);
            }
        v[3] = r2 * 2.0 * nu_new(a, nm[s]);
        v[5] = r4 * (-1.0) * (2.0 / 9.0) * P.kg;
    }
    // pass 1 wrote columns 0, 4, 6 (and zeros in the others); add this pass's columns
    __shared__ double sh[BPART][4];
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
        a.partial[blockIdx.x * BPART + k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
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
    const int64_t nb = (n + 255) / 256;
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
    a.partial = L->budget + (size_t)BWORK * n;
    hipLaunchKernelGGL(budget_pass1, dim3((unsigned)nb), dim3(256), 0, L->stream, a);
    HIP_TRYB(hipGetLastError());
    hipLaunchKernelGGL(budget_pass2, dim3((unsigned)nb), dim3(256), 0, L->stream, a);
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
