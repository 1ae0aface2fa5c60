// aux_fp32.hip — conversions between the FP64 absolute representation and the
// FP32 shifted storage of step_fp32.hip, moment recomputation and the LES
// relaxation-time field for the FP32 path.
#include "d3q15.h"
#include "kernels.h"

namespace drp {

template <int Q0, int Q1, typename Fn> __device__ __forceinline__ void sforx(Fn&& fn) {
    if constexpr (Q0 < Q1) {
        fn(std::integral_constant<int, Q0>{});
        sforx<Q0 + 1, Q1>(fn);
    }
}

static inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

__global__ void to_shifted(int64_t n, ShiftRef r, const double* __restrict__ pop64, const double* __restrict__ mac64,
                           float* __restrict__ pop32, float* __restrict__ mac32) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    sforx<0, 15>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        pop32[q * n + s] = (float)(pop64[q * n + s] - r.fr[q]);
        pop32[(15 + q) * n + s] = (float)(pop64[(15 + q) * n + s] - weightE<q>());
    });
    mac32[s] = (float)(mac64[s] - r.phi1);
    mac32[n + s] = (float)(mac64[n + s] - 1.0 / 3.0);
    mac32[2 * n + s] = (float)mac64[2 * n + s];
    mac32[3 * n + s] = (float)mac64[3 * n + s];
    mac32[4 * n + s] = (float)mac64[4 * n + s];
}

__global__ void pops_from_shifted(int64_t n, ShiftRef r, const float* __restrict__ pop32, double* __restrict__ out,
                                  int set) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    sforx<0, 15>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const double ref = set == 0 ? r.fr[q] : weightE<q>();
        out[q * n + s] = ref + (double)pop32[(set * 15 + q) * n + s];
    });
}

__global__ void pops_to_shifted(int64_t n, ShiftRef r, const double* __restrict__ in, float* __restrict__ pop32,
                                int set) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    sforx<0, 15>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const double ref = set == 0 ? r.fr[q] : weightE<q>();
        pop32[(set * 15 + q) * n + s] = (float)(in[q * n + s] - ref);
    });
}

__global__ void macro_from_shifted(int64_t n, double phi1, const float* __restrict__ m32, double* __restrict__ m64) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    m64[s] = phi1 + (double)m32[s];
    m64[n + s] = 1.0 / 3.0 + (double)m32[n + s];
    m64[2 * n + s] = (double)m32[2 * n + s];
    m64[3 * n + s] = (double)m32[3 * n + s];
    m64[4 * n + s] = (double)m32[4 * n + s];
}

__global__ void moments_fp32(int64_t n, const float* __restrict__ f, const float* __restrict__ g,
                             float* __restrict__ mac, int which) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    double ph = 0.0;
    float pr = 0.0f, m1 = 0.0f, m2 = 0.0f, m3 = 0.0f;
    sforx<0, 15>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        using D = Dir<q>;
        const float gv = g[q * n + s];
        ph += (double)f[q * n + s];
        pr += gv;
        m1 = acc<D::cx>(m1, gv);
        m2 = acc<D::cy>(m2, gv);
        m3 = acc<D::cz>(m3, gv);
    });
    if (which & 1) mac[s] = (float)ph;
    if (which & 2) {
        mac[n + s] = pr * (1.0f / 3.0f);
        mac[2 * n + s] = m1;
        mac[3 * n + s] = m2;
        mac[4 * n + s] = m3;
    }
}

// The FP32 form of les_tau_fp64 (aux_kernels.hip): tau_g for the next step from
// the state a step left (mac) and the one it started from (mac_old), every
// neighbour outside the node's rank block (B, periodic wraps included) from
// mac_old (B.bx = 0: all from mac).
__global__ void les_tau_fp32(int nx, int ny, int nz, Phys P, const float* __restrict__ mac,
                             const float* __restrict__ mac_old, LesBlocks B, const uint64_t* __restrict__ range,
                             int slot, float* __restrict__ tau) {
    const int64_t N = (int64_t)nx * ny * nz;
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= N) return;
    const int x = s % nx, y = (s / nx) % ny, z = s / ((int64_t)nx * ny);
    const int x0 = B.bx > 0 ? (x / B.bx) * B.bx : 0, y0 = B.bx > 0 ? (y / B.by) * B.by : 0,
              z0 = B.bx > 0 ? (z / B.bz) * B.bz : 0;
    float gu[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    sforx<1, 15>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        using D = Dir<q>;
        int xx = x + D::cx, yy = y + D::cy, zz = z + D::cz;
        const bool halo = B.bx > 0 && (xx < x0 || xx >= x0 + B.bx || yy < y0 || yy >= y0 + B.by || zz < z0 ||
                                       zz >= z0 + B.bz);
        const float* u = halo ? mac_old : mac;
        xx = xx < 0 ? xx + nx : (xx >= nx ? xx - nx : xx);
        yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
        zz = zz < 0 ? zz + nz : (zz >= nz ? zz - nz : zz);
        const int64_t t = xx + (int64_t)nx * (yy + (int64_t)ny * zz);
        for (int k = 0; k < 3; k++) {
            const float v = u[(2 + k) * N + t];
            gu[k][0] = acc<D::cx>(gu[k][0], v);
            gu[k][1] = acc<D::cy>(gu[k][1], v);
            gu[k][2] = acc<D::cz>(gu[k][2], v);
        }
    });
    float sum = 0.0f;
    for (int b = 0; b < 3; b++)
        for (int a = 0; a < 3; a++) {
            const float sab = 0.05f * (gu[a][b] + gu[b][a]);  // 0.5 (d_b u_a + d_a u_b), gradients / 10
            sum += sab * sab;
        }
    const double rmax = from_ordered_bits(range[2 * slot]), rmin = from_ordered_bits(range[2 * slot + 1]);
    const double rg = rmax - rmin;
    const double phi = P.phi1 + (double)mac[s];
    const double frac = rg == 0.0 ? 0.0 : (phi - rmin) / rg;
    const double nu = frac * P.dnu + P.nu1;
    tau[s] = (float)(0.5 + 3.0 * (nu + P.les_c * sqrt(2.0 * (double)sum)));
}

// The same tau with every neighbour current, marching in z over a 64 x 8 column
// of nodes with the planes z-1, z, z+1 of u (one-node x-y halo) in a ring of
// three LDS planes; the loads of plane z+2 are in flight into registers while
// plane z is computed.  The arithmetic and its order are those of les_tau_fp32.
constexpr int ZM32_TX = 64, ZM32_TY = 8, ZM32_ZC = 16;
constexpr int ZM32_HX = ZM32_TX + 2, ZM32_HP = ZM32_HX * (ZM32_TY + 2), ZM32_NT = ZM32_TX * ZM32_TY;
constexpr int ZM32_PER = (ZM32_HP + ZM32_NT - 1) / ZM32_NT;
__device__ __forceinline__ int wrap_zm32(int v, int n, bool small) {
    if (small) return ((v % n) + n) % n;
    return v < 0 ? v + n : (v >= n ? v - n : v);
}
__global__ __launch_bounds__(ZM32_NT) void les_tau_fp32_zm(int nx, int ny, int nz, Phys P, const float* __restrict__ mac,
                                                          const uint64_t* __restrict__ range, int slot,
                                                          float* __restrict__ tau, int tiles_x, int tiles_y) {
    __shared__ float su[3][3][ZM32_HP];  // [ring slot][component][plane node]
    const int64_t N = (int64_t)nx * ny * nz;
    const int x0 = (blockIdx.x % tiles_x) * ZM32_TX, y0 = ((blockIdx.x / tiles_x) % tiles_y) * ZM32_TY;
    const int zs = (blockIdx.x / (tiles_x * tiles_y)) * ZM32_ZC, ze = min(zs + ZM32_ZC, nz);
    const bool small = nx <= ZM32_TX || ny <= ZM32_TY || nz <= 2;
    const int tx = threadIdx.x % ZM32_TX, ty = threadIdx.x / ZM32_TX;
    const int x = x0 + tx, y = y0 + ty;
    const bool in = x < nx && y < ny;
    const double rmax = from_ordered_bits(range[2 * slot]), rmin = from_ordered_bits(range[2 * slot + 1]);
    const double rg = rmax - rmin;
    auto fetch = [&](int zg, float (&r)[ZM32_PER][3]) {
        const int64_t zoff = (int64_t)nx * ny * wrap_zm32(zg, nz, small);
#pragma unroll
        for (int i = 0; i < ZM32_PER; i++) {
            const int h0 = threadIdx.x + i * ZM32_NT, h = h0 < ZM32_HP ? h0 : ZM32_HP - 1;
            const int64_t g = wrap_zm32(x0 + h % ZM32_HX - 1, nx, small) +
                              (int64_t)nx * wrap_zm32(y0 + h / ZM32_HX - 1, ny, small) + zoff;
#pragma unroll
            for (int k = 0; k < 3; k++) r[i][k] = mac[(2 + k) * N + g];
        }
    };
    auto store = [&](const float (&r)[ZM32_PER][3], float (*dst)[ZM32_HP]) {
#pragma unroll
        for (int i = 0; i < ZM32_PER; i++) {
            const int h = threadIdx.x + i * ZM32_NT;
            if (h < ZM32_HP)
#pragma unroll
                for (int k = 0; k < 3; k++) dst[k][h] = r[i][k];
        }
    };
    float next[ZM32_PER][3];
    // ring slot of plane z' is (z' - zs + 1) mod 3
    for (int p = 0; p < 3; p++) {
        fetch(zs - 1 + p, next);
        store(next, su[p]);
    }
    const int64_t col = (int64_t)(in ? x : 0) + (int64_t)nx * (in ? y : 0);
    float phi_next = mac[col + (int64_t)nx * ny * zs];
    for (int z = zs; z < ze; z++) {
        const bool more = z + 1 < ze;  // plane z+2 is needed by plane z+1
        if (more) fetch(z + 2, next);
        const float phit = phi_next;
        if (more) phi_next = mac[col + (int64_t)nx * ny * (z + 1)];
        __syncthreads();
        if (in) {
            const int rz[3] = {(z - zs) % 3, (z - zs + 1) % 3, (z - zs + 2) % 3};  // planes z-1, z, z+1
            float gu[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            sforx<1, 15>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                using D = Dir<q>;
                const int h = (tx + 1 + D::cx) + ZM32_HX * (ty + 1 + D::cy);
                for (int k = 0; k < 3; k++) {
                    const float v = su[rz[1 + D::cz]][k][h];
                    gu[k][0] = acc<D::cx>(gu[k][0], v);
                    gu[k][1] = acc<D::cy>(gu[k][1], v);
                    gu[k][2] = acc<D::cz>(gu[k][2], v);
                }
            });
            float sum = 0.0f;
            for (int b = 0; b < 3; b++)
                for (int a = 0; a < 3; a++) {
                    const float sab = 0.05f * (gu[a][b] + gu[b][a]);
                    sum += sab * sab;
                }
            const double phi = P.phi1 + (double)phit;
            const double frac = rg == 0.0 ? 0.0 : (phi - rmin) / rg;
            const double nu = frac * P.dnu + P.nu1;
            tau[col + (int64_t)nx * ny * z] = (float)(0.5 + 3.0 * (nu + P.les_c * sqrt(2.0 * (double)sum)));
        }
        __syncthreads();
        if (more) store(next, su[(z - zs) % 3]);  // the slot of plane z-1
    }
}

hipError_t launch_to_shifted(int64_t n, const ShiftRef& r, const double* pop64, const double* mac64, float* pop32,
                             float* mac32, hipStream_t st) {
    hipLaunchKernelGGL(to_shifted, grid_for(n), dim3(256), 0, st, n, r, pop64, mac64, pop32, mac32);
    return hipGetLastError();
}
hipError_t launch_pops_from_shifted(int64_t n, const ShiftRef& r, const float* pop32, double* pop64, int set,
                                    hipStream_t st) {
    hipLaunchKernelGGL(pops_from_shifted, grid_for(n), dim3(256), 0, st, n, r, pop32, pop64, set);
    return hipGetLastError();
}
hipError_t launch_pops_to_shifted(int64_t n, const ShiftRef& r, const double* pop64, float* pop32, int set,
                                  hipStream_t st) {
    hipLaunchKernelGGL(pops_to_shifted, grid_for(n), dim3(256), 0, st, n, r, pop64, pop32, set);
    return hipGetLastError();
}
hipError_t launch_macro_from_shifted(int64_t n, double phi1, const float* mac32, double* mac64, hipStream_t st) {
    hipLaunchKernelGGL(macro_from_shifted, grid_for(n), dim3(256), 0, st, n, phi1, mac32, mac64);
    return hipGetLastError();
}
hipError_t launch_moments_fp32(int64_t n, const float* f, const float* g, float* mac, int which, hipStream_t st) {
    hipLaunchKernelGGL(moments_fp32, grid_for(n), dim3(256), 0, st, n, f, g, mac, which);
    return hipGetLastError();
}
hipError_t launch_les_tau_fp32(int nx, int ny, int nz, const Phys& ph, const float* mac, const float* mac_old,
                               const LesBlocks& blk, const uint64_t* range, int slot, float* tau, hipStream_t st) {
    const int64_t n = (int64_t)nx * ny * nz;
    if (blk.bx > 0 && mac_old != mac) {
        hipLaunchKernelGGL(les_tau_fp32, grid_for(n), dim3(256), 0, st, nx, ny, nz, ph, mac, mac_old, blk, range, slot,
                           tau);
    } else {
        const int mx = (nx + ZM32_TX - 1) / ZM32_TX, my = (ny + ZM32_TY - 1) / ZM32_TY, mz = (nz + ZM32_ZC - 1) / ZM32_ZC;
        hipLaunchKernelGGL(les_tau_fp32_zm, dim3((unsigned)(mx * my * mz)), dim3(ZM32_NT), 0, st, nx, ny, nz, ph, mac,
                           range, slot, tau, mx, my);
    }
    return hipGetLastError();
}

}  // namespace drp
