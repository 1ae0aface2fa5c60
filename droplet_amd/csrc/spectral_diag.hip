// spectral_diag.hip — FFT diagnostics of the velocity field on the GPU:
//
//   drp_energy_spectrum     enesupe_cal, program/turbulence_hit_xy.f90:588-644
//   drp_filter_velocity     low_pass, program/LES_drop_xy.f90:851-971 (keep the
//                           shells int|k| < k) and the band-pass of scale_cal2,
//                           program/drop_analysis.f90:722-797 (keep int|k| == k)
//   drp_scale_interaction   the per-scale interface-velocity interaction of
//                           program/drop_analysis.f90:1121-1265:
//                           C(k) = kappa_g < u^(k) . div(grad phi grad phi) >
//   drp_phi_contribution    the cross spectrum of u . grad phi and the chemical
//                           potential, contribution of drop_analysis.f90:476-578
//
// The reference transforms with 2decomp/FFTW (r2c with the halved axis z,
// u_hat = FFT(u) / (nx ny nz), signed wavenumber k = i - n for i >= n/2 + 1).
// Here a complex 3-D FFT is done by one mixed-radix Stockham pass per axis
// (radices 4, 2, 3, 5; every extent a product of those, at most 2048): each
// workgroup loads G whole lines into LDS, runs the log_R(n) stages there with
// twiddles from a table e^{-2 pi i t / n} evaluated on the host, and writes the
// lines back in place.  Lines along y and z are taken G at a time with
// consecutive x, so their loads and stores are coalesced.
//
// The per-scale interaction uses Parseval instead of one inverse transform per
// shell: < u^(k) . w > = sum over the modes of shell k of Re(u_hat w_hat*), with
// both transforms normalised by the node count; it equals the reference's
// physical-space mean of the band-passed velocity times w up to rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/droplet.h"
#include "d3q15.h"
#include "kernels.h"
#include "lattice_impl.h"

using namespace drp;

namespace {

#define HIP_TRYS(expr)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return drp_fail(DRP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

constexpr int FFT_MAX_STAGES = 12;
constexpr int FFT_MAX_LINE = 2048;   // longest line (one line per workgroup: two 32 KiB buffers)
constexpr int FFT_BLOCK_LINES = 1024;  // G * L complex numbers per buffer for shorter lines (2 x 16 KiB):
                                       // four 256-point lines per workgroup ran 1.2x faster than eight

struct FFTAxis {
    int L;                     // line length
    int64_t S;                 // element stride along the line
    int64_t nlines;            // lines of this axis
    int F;                     // lines per row of the fastest other coordinate
    int64_t sf, so;            // base = (l % F) * sf + (l / F) * so
    int G;                     // lines per workgroup
    int nst;                   // stages
    int rad[FFT_MAX_STAGES];   // radix of each stage
    const double2* tw;         // e^{-2 pi i t / L}, t = 0..L-1
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// R-point DFT o[q] = sum_r v[r] W^(r q), W = e^{s 2 pi i / R} (s = -1 forward,
// +1 inverse), with the roots of unity as constants
template <int R> __device__ __forceinline__ void small_dft(const double2 (&v)[R], double2 (&o)[R], int sign) {
    const double s = (double)sign;
    if constexpr (R == 2) {
        o[0] = make_double2(v[0].x + v[1].x, v[0].y + v[1].y);
        o[1] = make_double2(v[0].x - v[1].x, v[0].y - v[1].y);
    } else if constexpr (R == 4) {
        // W = s i: W z = (-s z.y, s z.x)
        const double2 a = make_double2(v[0].x + v[2].x, v[0].y + v[2].y);
        const double2 b = make_double2(v[0].x - v[2].x, v[0].y - v[2].y);
        const double2 c = make_double2(v[1].x + v[3].x, v[1].y + v[3].y);
        const double2 d = make_double2(v[1].x - v[3].x, v[1].y - v[3].y);
        const double2 wd = make_double2(-s * d.y, s * d.x);
        o[0] = make_double2(a.x + c.x, a.y + c.y);
        o[1] = make_double2(b.x + wd.x, b.y + wd.y);
        o[2] = make_double2(a.x - c.x, a.y - c.y);
        o[3] = make_double2(b.x - wd.x, b.y - wd.y);
    } else {
        // R = 3, 5: cos and sin of 2 pi m / R
        constexpr double C3[3] = {1.0, -0.5, -0.5};
        constexpr double S3[3] = {0.0, 0.86602540378443864676, -0.86602540378443864676};
        constexpr double C5[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410, -0.80901699437494742410,
                                  0.30901699437494742410};
        constexpr double S5[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917, -0.58778525229247312917,
                                  -0.95105651629515357212};
#pragma unroll
        for (int q = 0; q < R; q++) {
            double2 acc = v[0];
#pragma unroll
            for (int r = 1; r < R; r++) {
                const int m = (r * q) % R;
                const double wc = R == 3 ? C3[m] : C5[m], ws = s * (R == 3 ? S3[m] : S5[m]);
                acc.x += v[r].x * wc - v[r].y * ws;
                acc.y += v[r].x * ws + v[r].y * wc;
            }
            o[q] = acc;
        }
    }
}

// one Stockham stage of radix R over the ng lines of length L held in LDS
template <int R>
__device__ __forceinline__ void fft_stage(const double2* __restrict__ src, double2* __restrict__ dst, int L, int ng,
                                          int Ns, const double2* __restrict__ tw, int sign) {
    const int M = L / R, step = L / (Ns * R);
    for (int b = threadIdx.x; b < ng * M; b += blockDim.x) {
        const int g = b / M, j = b - g * M;
        const int k = j % Ns;
        const double2* sl = src + g * L;
        double2 v[R], o[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            double2 x = sl[j + r * M];
            if (r > 0 && k > 0) {
                double2 w = tw[k * r * step];  // k r step < L
                if (sign > 0) w.y = -w.y;
                x = cmul(x, w);
            }
            v[r] = x;
        }
        small_dft<R>(v, o, sign);
        double2* dl = dst + g * L + (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) dl[r * Ns] = o[r];
    }
}

// one workgroup: G lines of one axis, all stages in LDS; sign -1 forward, +1 inverse.
// Strided axes: a thread keeps one line (consecutive threads take consecutive
// lines, whose elements are adjacent in memory) and walks along it; along x the
// threads walk the line itself.
__global__ __launch_bounds__(256) void fft_lines(double2* __restrict__ data, FFTAxis A, int sign) {
    extern __shared__ double2 lds[];
    const int L = A.L, G = A.G;
    double2* buf0 = lds;
    double2* buf1 = lds + (size_t)G * L;
    const int64_t l0 = (int64_t)blockIdx.x * G;
    const int ng = (int)min<int64_t>(G, A.nlines - l0);
    const int tid = threadIdx.x, nth = blockDim.x;
    if (A.S == 1) {
        for (int g = 0; g < ng; g++) {
            const int64_t l = l0 + g;
            const double2* line = data + (l % A.F) * A.sf + (l / A.F) * A.so;
            for (int j = tid; j < L; j += nth) buf0[g * L + j] = line[j];
        }
    } else {
        const int g = tid % G, j0 = tid / G, js = nth / G;
        if (g < ng && j0 < js) {
            const int64_t l = l0 + g;
            const double2* line = data + (l % A.F) * A.sf + (l / A.F) * A.so;
            for (int j = j0; j < L; j += js) buf0[g * L + j] = line[(int64_t)j * A.S];
        }
    }
    __syncthreads();
    int Ns = 1;
    for (int st = 0; st < A.nst; st++) {
        const int R = A.rad[st];
        if (R == 4) fft_stage<4>(buf0, buf1, L, ng, Ns, A.tw, sign);
        else if (R == 2) fft_stage<2>(buf0, buf1, L, ng, Ns, A.tw, sign);
        else if (R == 3) fft_stage<3>(buf0, buf1, L, ng, Ns, A.tw, sign);
        else fft_stage<5>(buf0, buf1, L, ng, Ns, A.tw, sign);
        __syncthreads();
        double2* t = buf0;
        buf0 = buf1;
        buf1 = t;
        Ns *= R;
    }
    if (A.S == 1) {
        for (int g = 0; g < ng; g++) {
            const int64_t l = l0 + g;
            double2* line = data + (l % A.F) * A.sf + (l / A.F) * A.so;
            for (int j = tid; j < L; j += nth) line[j] = buf0[g * L + j];
        }
    } else {
        const int g = tid % G, j0 = tid / G, js = nth / G;
        if (g < ng && j0 < js) {
            const int64_t l = l0 + g;
            double2* line = data + (l % A.F) * A.sf + (l / A.F) * A.so;
            for (int j = j0; j < L; j += js) line[(int64_t)j * A.S] = buf0[g * L + j];
        }
    }
}

__global__ void to_complex(int64_t n, const double* __restrict__ a, double2* __restrict__ c) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) c[i] = make_double2(a[i], 0.0);
}

__global__ void real_part_scaled(int64_t n, const double2* __restrict__ c, double s, double* __restrict__ a) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) a[i] = c[i].x * s;
}

__device__ __forceinline__ int signed_k(int i, int n) { return i >= n / 2 + 1 ? i - n : i; }  // judge

__device__ __forceinline__ int shell_of(int x, int y, int z, int nx, int ny, int nz) {
    const int k1 = signed_k(x, nx), k2 = signed_k(y, ny), k3 = signed_k(z, nz);
    return (int)sqrt((double)(k1 * k1 + k2 * k2 + k3 * k3));
}

// spectrum over the reference's half space k3 = 0 .. nz/2 (index), weight 2 off the
// k3 = 0 plane (the Nyquist plane included, as the reference counts it); shells
// accumulated per workgroup in LDS, then one atomic per workgroup and shell
__global__ __launch_bounds__(256) void shell_energy(int nx, int ny, int nz, const double2* __restrict__ u1,
                                                    const double2* __restrict__ u2, const double2* __restrict__ u3,
                                                    double inv, int nshells, double* __restrict__ E) {
    extern __shared__ double bins[];
    for (int s = threadIdx.x; s < nshells; s += blockDim.x) bins[s] = 0.0;
    __syncthreads();
    const int64_t half = (int64_t)nx * ny * (nz / 2 + 1);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < half; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
        const int s = shell_of(x, y, z, nx, ny, nz);
        if (s >= nshells) continue;
        const double2 a = u1[i], b = u2[i], c = u3[i];
        const double e = ((a.x * a.x + a.y * a.y) + (b.x * b.x + b.y * b.y) + (c.x * c.x + c.y * c.y)) * inv * inv;
        atomicAdd(&bins[s], 0.5 * e * (z == 0 ? 1.0 : 2.0));
    }
    __syncthreads();
    for (int s = threadIdx.x; s < nshells; s += blockDim.x)
        if (bins[s] != 0.0) atomicAdd(&E[s], bins[s]);
}

// per-shell sum over the full spectrum of Re(u_hat . w_hat*), normalised
__global__ __launch_bounds__(256) void shell_cross(int nx, int ny, int nz, const double2* __restrict__ u,
                                                   const double2* __restrict__ w, double inv, int nshells,
                                                   double* __restrict__ C) {
    extern __shared__ double bins[];
    for (int s = threadIdx.x; s < nshells; s += blockDim.x) bins[s] = 0.0;
    __syncthreads();
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
        const int s = shell_of(x, y, z, nx, ny, nz);
        if (s >= nshells) continue;
        for (int c = 0; c < 3; c++) {
            const double2 a = u[c * n + i], b = w[c * n + i];
            atomicAdd(&bins[s], (a.x * b.x + a.y * b.y) * inv * inv);
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < nshells; s += blockDim.x)
        if (bins[s] != 0.0) atomicAdd(&C[s], bins[s]);
}

// half-space cross spectrum of contribution (drop_analysis.f90:476-578): per shell the
// sums of w Re(a conj(m)) and of w, w = 1 on the k3 = 0 plane and 2 elsewhere
__global__ __launch_bounds__(256) void shell_cross_half(int nx, int ny, int nz, const double2* __restrict__ a,
                                                        const double2* __restrict__ m, int nshells,
                                                        double* __restrict__ C, double* __restrict__ W) {
    extern __shared__ double bins[];  // [2][nshells]
    for (int s = threadIdx.x; s < 2 * nshells; s += blockDim.x) bins[s] = 0.0;
    __syncthreads();
    const int64_t half = (int64_t)nx * ny * (nz / 2 + 1);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < half; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
        const int s = shell_of(x, y, z, nx, ny, nz);
        if (s >= nshells) continue;
        const double w = z == 0 ? 1.0 : 2.0;
        const double2 av = a[i], mv = m[i];
        atomicAdd(&bins[s], w * (av.x * mv.x + av.y * mv.y));
        atomicAdd(&bins[nshells + s], w);
    }
    __syncthreads();
    for (int s = threadIdx.x; s < nshells; s += blockDim.x) {
        if (bins[s] != 0.0) atomicAdd(&C[s], bins[s]);
        if (bins[nshells + s] != 0.0) atomicAdd(&W[s], bins[nshells + s]);
    }
}

// keep the modes with lo <= int|k| < hi, zero the rest
__global__ void shell_mask(int nx, int ny, int nz, double2* __restrict__ a, int lo, int hi) {
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
    const int s = shell_of(x, y, z, nx, ny, nz);
    if (s < lo || s >= hi) a[i] = make_double2(0.0, 0.0);
}

// grad phi by the D3Q15 stencil (grad_cal, drop_analysis.f90; sum of c phi(x + c) / 10)
__global__ void grad_phi(int nx, int ny, int nz, const double* __restrict__ phi, double* __restrict__ g) {
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int x = (int)(s % nx), y = (int)((s / nx) % ny), z = (int)(s / ((int64_t)nx * ny));
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int q = 1; q < 15; q++) {
        const int cx[15] = {0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1};
        const int cy[15] = {0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1};
        const int cz[15] = {0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1};
        const int xx = (x + cx[q] + nx) % nx, yy = (y + cy[q] + ny) % ny, zz = (z + cz[q] + nz) % nz;
        const double v = phi[xx + (int64_t)nx * (yy + (int64_t)ny * zz)];
        g0 += cx[q] * v;
        g1 += cy[q] * v;
        g2 += cz[q] * v;
    }
    g[s] = g0 / 10.0;
    g[n + s] = g1 / 10.0;
    g[2 * n + s] = g2 / 10.0;
}

// w_b = sum_a d_a (d_b phi d_a phi), the stencil applied to the products at the
// neighbours (grad_tensor, drop_analysis.f90:1121-1217), written as complex numbers
__global__ void div_korteweg(int nx, int ny, int nz, const double* __restrict__ g, double2* __restrict__ w) {
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int x = (int)(s % nx), y = (int)((s / nx) % ny), z = (int)(s / ((int64_t)nx * ny));
    double d[3] = {0.0, 0.0, 0.0};
    for (int q = 1; q < 15; q++) {
        const int cx[15] = {0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1};
        const int cy[15] = {0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1};
        const int cz[15] = {0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1};
        const int xx = (x + cx[q] + nx) % nx, yy = (y + cy[q] + ny) % ny, zz = (z + cz[q] + nz) % nz;
        const int64_t t = xx + (int64_t)nx * (yy + (int64_t)ny * zz);
        const double ga[3] = {g[t], g[n + t], g[2 * n + t]};
        const double c[3] = {(double)cx[q], (double)cy[q], (double)cz[q]};
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++) d[b] += c[a] * (ga[b] * ga[a]);
    }
    for (int b = 0; b < 3; b++) w[b * n + s] = make_double2(d[b] / 10.0, 0.0);
}

// ------------------------------------------------------------------ host side

struct FFTPlan {
    int n[3] = {0, 0, 0};
    FFTAxis ax[3];
    double2* tw = nullptr;  // the three twiddle tables, back to back
    ~FFTPlan() {
        if (tw) (void)hipFree(tw);
    }
};

int factor(int L, int* rad, int* nst) {
    *nst = 0;
    int m = L;
    for (const int R : {4, 2, 3, 5}) {
        while (m % R == 0) {
            if (*nst >= FFT_MAX_STAGES) return -1;
            rad[(*nst)++] = R;
            m /= R;
        }
    }
    return m == 1 ? 0 : -1;
}

int make_plan(int nx, int ny, int nz, FFTPlan& P) {
    const int ns[3] = {nx, ny, nz};
    for (int a = 0; a < 3; a++) {
        if (ns[a] > FFT_MAX_LINE || factor(ns[a], P.ax[a].rad, &P.ax[a].nst))
            return drp_fail(DRP_ERR_UNSUPPORTED, "FFT diagnostics need extents that are products of 2, 3 and 5, "
                                                 "at most %d (got %d)", FFT_MAX_LINE, ns[a]);
        P.n[a] = ns[a];
    }
    HIP_TRYS(hipMalloc((void**)&P.tw, (size_t)(nx + ny + nz) * sizeof(double2)));
    std::vector<double2> t;
    for (int a = 0; a < 3; a++)
        for (int j = 0; j < ns[a]; j++) {
            const long double th = -2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)ns[a];
            t.push_back(make_double2((double)cosl(th), (double)sinl(th)));
        }
    HIP_TRYS(hipMemcpy(P.tw, t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice));
    const int64_t nxy = (int64_t)nx * ny;
    // x: lines contiguous; y: lines (x, z), stride nx; z: lines (x, y), stride nx ny
    P.ax[0].L = nx; P.ax[0].S = 1;   P.ax[0].nlines = nxy / nx * nz; P.ax[0].F = 1; P.ax[0].sf = 0; P.ax[0].so = nx;
    P.ax[1].L = ny; P.ax[1].S = nx;  P.ax[1].nlines = (int64_t)nx * nz; P.ax[1].F = nx; P.ax[1].sf = 1; P.ax[1].so = nxy;
    P.ax[2].L = nz; P.ax[2].S = nxy; P.ax[2].nlines = nxy; P.ax[2].F = nx; P.ax[2].sf = 1; P.ax[2].so = nx;
    P.ax[0].tw = P.tw;
    P.ax[1].tw = P.tw + nx;
    P.ax[2].tw = P.tw + nx + ny;
    for (int a = 0; a < 3; a++) P.ax[a].G = std::max(1, std::min(16, FFT_BLOCK_LINES / P.ax[a].L));
    return DRP_OK;
}

hipError_t fft3d(const FFTPlan& P, double2* data, int sign, hipStream_t st) {
    for (int a = 0; a < 3; a++) {
        const FFTAxis& A = P.ax[a];
        const int64_t blocks = (A.nlines + A.G - 1) / A.G;
        const size_t lds = 2 * (size_t)A.G * A.L * sizeof(double2);
        hipLaunchKernelGGL(fft_lines, dim3((unsigned)blocks), dim3(256), lds, st, data, A, sign);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// device buffer released on scope exit
struct DevMem {
    void* p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
};

int velocity_view(drp_lattice* L, const double** mac) {
    if (L->cfg.precision == DRP_FP64) {
        *mac = (const double*)L->mac[L->cur];
        return DRP_OK;
    }
    return fp32_macro_fp64_view(L, mac);
}

// forward transforms of u1, u2, u3 into c[3][n] (unnormalised)
int transform_velocity(drp_lattice* L, const FFTPlan& P, double2* c) {
    const double* mac = nullptr;
    int rc = velocity_view(L, &mac);
    if (rc) return rc;
    const int64_t n = L->n;
    const unsigned nb = (unsigned)((n + 255) / 256);
    for (int k = 0; k < 3; k++) {
        hipLaunchKernelGGL(to_complex, dim3(nb), dim3(256), 0, L->stream, n, mac + (2 + k) * n, c + k * n);
        HIP_TRYS(hipGetLastError());
        HIP_TRYS(fft3d(P, c + k * n, -1, L->stream));
    }
    return DRP_OK;
}

unsigned shell_blocks(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 2048); }

}  // namespace

// FFT round trip on a caller's complex field (for tests of the transform itself):
// data[n] complex as interleaved (re, im), sign -1 forward, +1 inverse (unnormalised).
extern "C" int drp_fft3d(int nx, int ny, int nz, double* data, int sign) {
    if (!data || nx < 1 || ny < 1 || nz < 1 || (sign != -1 && sign != 1)) return drp_fail(DRP_ERR_ARG, "bad FFT arguments");
    FFTPlan P;
    int rc = make_plan(nx, ny, nz, P);
    if (rc) return rc;
    const int64_t n = (int64_t)nx * ny * nz;
    DevMem d;
    HIP_TRYS(hipMalloc(&d.p, (size_t)n * sizeof(double2)));
    HIP_TRYS(hipMemcpy(d.p, data, (size_t)n * sizeof(double2), hipMemcpyHostToDevice));
    HIP_TRYS(fft3d(P, (double2*)d.p, sign, nullptr));
    HIP_TRYS(hipMemcpy(data, d.p, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost));
    return DRP_OK;
}

constexpr int MAX_SHELLS = 8192;  // shell bins of one workgroup in LDS

extern "C" int drp_energy_spectrum(drp_lattice* L, double* E, int nshells) {
    if (!L || !E || nshells < 1 || nshells > MAX_SHELLS) return drp_fail(DRP_ERR_ARG, "bad spectrum arguments");
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    const drp_config& c = L->cfg;
    FFTPlan P;
    int rc = make_plan(c.nx, c.ny, c.nz, P);
    if (rc) return rc;
    const int64_t n = L->n;
    DevMem cu, dE;
    HIP_TRYS(hipMalloc(&cu.p, 3 * (size_t)n * sizeof(double2)));
    HIP_TRYS(hipMalloc(&dE.p, (size_t)nshells * sizeof(double)));
    rc = transform_velocity(L, P, (double2*)cu.p);
    if (rc) return rc;
    HIP_TRYS(hipMemsetAsync(dE.p, 0, (size_t)nshells * sizeof(double), L->stream));
    const double2* u = (const double2*)cu.p;
    hipLaunchKernelGGL(shell_energy, dim3(shell_blocks(n)), dim3(256), (size_t)nshells * sizeof(double), L->stream,
                       c.nx, c.ny, c.nz, u, u + n, u + 2 * n, 1.0 / (double)n, nshells, (double*)dE.p);
    HIP_TRYS(hipGetLastError());
    HIP_TRYS(hipMemcpyAsync(E, dE.p, (size_t)nshells * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIP_TRYS(hipStreamSynchronize(L->stream));
    return DRP_OK;
}

extern "C" int drp_filter_velocity(drp_lattice* L, int kind, int k, double* u1, double* u2, double* u3, size_t n_) {
    if (!L || !u1 || !u2 || !u3) return drp_fail(DRP_ERR_ARG, "null argument");
    if (kind != DRP_FILTER_LOW_PASS && kind != DRP_FILTER_SHELL) return drp_fail(DRP_ERR_ARG, "unknown filter %d", kind);
    if (n_ != (size_t)L->n) return drp_fail(DRP_ERR_ARG, "size %zu != %lld nodes", n_, (long long)L->n);
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    const drp_config& c = L->cfg;
    FFTPlan P;
    int rc = make_plan(c.nx, c.ny, c.nz, P);
    if (rc) return rc;
    const int64_t n = L->n;
    DevMem cu, out;
    HIP_TRYS(hipMalloc(&cu.p, 3 * (size_t)n * sizeof(double2)));
    HIP_TRYS(hipMalloc(&out.p, (size_t)n * sizeof(double)));
    rc = transform_velocity(L, P, (double2*)cu.p);
    if (rc) return rc;
    // low_pass keeps int|k| + 1 <= k_low, i.e. int|k| < k; the band-pass keeps int|k| == k
    const int lo = kind == DRP_FILTER_LOW_PASS ? 0 : k, hi = kind == DRP_FILTER_LOW_PASS ? k : k + 1;
    const unsigned nb = (unsigned)((n + 255) / 256);
    double* dst[3] = {u1, u2, u3};
    for (int a = 0; a < 3; a++) {
        double2* ca = (double2*)cu.p + a * n;
        hipLaunchKernelGGL(shell_mask, dim3(nb), dim3(256), 0, L->stream, c.nx, c.ny, c.nz, ca, lo, hi);
        HIP_TRYS(hipGetLastError());
        HIP_TRYS(fft3d(P, ca, +1, L->stream));
        hipLaunchKernelGGL(real_part_scaled, dim3(nb), dim3(256), 0, L->stream, n, (const double2*)ca, 1.0 / (double)n,
                           (double*)out.p);
        HIP_TRYS(hipGetLastError());
        HIP_TRYS(hipMemcpyAsync(dst[a], out.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, L->stream));
        HIP_TRYS(hipStreamSynchronize(L->stream));
    }
    return DRP_OK;
}

extern "C" int drp_scale_interaction(drp_lattice* L, double* C, int nshells) {
    if (!L || !C || nshells < 1 || nshells > MAX_SHELLS) return drp_fail(DRP_ERR_ARG, "bad interaction arguments");
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    const drp_config& c = L->cfg;
    FFTPlan P;
    int rc = make_plan(c.nx, c.ny, c.nz, P);
    if (rc) return rc;
    const int64_t n = L->n;
    DevMem cu, cw, gp, dC;
    HIP_TRYS(hipMalloc(&cu.p, 3 * (size_t)n * sizeof(double2)));
    HIP_TRYS(hipMalloc(&cw.p, 3 * (size_t)n * sizeof(double2)));
    HIP_TRYS(hipMalloc(&gp.p, 3 * (size_t)n * sizeof(double)));
    HIP_TRYS(hipMalloc(&dC.p, (size_t)nshells * sizeof(double)));
    rc = transform_velocity(L, P, (double2*)cu.p);
    if (rc) return rc;
    const double* mac = nullptr;
    rc = velocity_view(L, &mac);
    if (rc) return rc;
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(grad_phi, dim3(nb), dim3(256), 0, L->stream, c.nx, c.ny, c.nz, mac, (double*)gp.p);
    HIP_TRYS(hipGetLastError());
    hipLaunchKernelGGL(div_korteweg, dim3(nb), dim3(256), 0, L->stream, c.nx, c.ny, c.nz, (const double*)gp.p,
                       (double2*)cw.p);
    HIP_TRYS(hipGetLastError());
    for (int a = 0; a < 3; a++) HIP_TRYS(fft3d(P, (double2*)cw.p + a * n, -1, L->stream));
    HIP_TRYS(hipMemsetAsync(dC.p, 0, (size_t)nshells * sizeof(double), L->stream));
    hipLaunchKernelGGL(shell_cross, dim3(shell_blocks(n)), dim3(256), (size_t)nshells * sizeof(double), L->stream,
                       c.nx, c.ny, c.nz, (const double2*)cu.p, (const double2*)cw.p, 1.0 / (double)n, nshells,
                       (double*)dC.p);
    HIP_TRYS(hipGetLastError());
    std::vector<double> h(nshells);
    HIP_TRYS(hipMemcpyAsync(h.data(), dC.p, (size_t)nshells * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIP_TRYS(hipStreamSynchronize(L->stream));
    for (int s = 0; s < nshells; s++) C[s] = c.kappa_g * h[s];
    return DRP_OK;
}

extern "C" int drp_phi_contribution(drp_lattice* L, double* C, int nshells, double* bands) {
    if (!L || !C || !bands) return drp_fail(DRP_ERR_ARG, "null argument");
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    const drp_config& c = L->cfg;
    const int nb = c.nx / 2 + 1;  // the reference's (xmax+1)/2 + 1 shells
    if (nshells < 1 || nshells > nb) return drp_fail(DRP_ERR_ARG, "nshells %d not in 1..%d", nshells, nb);
    FFTPlan P;
    int rc = make_plan(c.nx, c.ny, c.nz, P);
    if (rc) return rc;
    const double* mac = nullptr;
    rc = velocity_view(L, &mac);
    if (rc) return rc;
    const int64_t n = L->n;
    DevMem ca, tmp, dC;
    HIP_TRYS(hipMalloc(&ca.p, 2 * (size_t)n * sizeof(double2)));
    HIP_TRYS(hipMalloc(&tmp.p, (size_t)n * sizeof(double)));
    HIP_TRYS(hipMalloc(&dC.p, 2 * (size_t)nb * sizeof(double)));
    double2* adv = (double2*)ca.p;
    double2* mu = adv + n;
    const unsigned nblk = (unsigned)((n + 255) / 256);
    const int kinds[2] = {DRP_DERIVED_ADVECTION, DRP_DERIVED_CHEMICAL_POTENTIAL};
    double2* dst[2] = {adv, mu};
    for (int k = 0; k < 2; k++) {
        HIP_TRYS(launch_derived_fp64(c.nx, c.ny, c.nz, L->phys, mac, kinds[k], (double*)tmp.p, L->stream));
        hipLaunchKernelGGL(to_complex, dim3(nblk), dim3(256), 0, L->stream, n, (const double*)tmp.p, dst[k]);
        HIP_TRYS(hipGetLastError());
        HIP_TRYS(fft3d(P, dst[k], -1, L->stream));
    }
    HIP_TRYS(hipMemsetAsync(dC.p, 0, 2 * (size_t)nb * sizeof(double), L->stream));
    double* dsum = (double*)dC.p;
    hipLaunchKernelGGL(shell_cross_half, dim3(shell_blocks(n)), dim3(256), 2 * (size_t)nb * sizeof(double), L->stream,
                       c.nx, c.ny, c.nz, (const double2*)adv, (const double2*)mu, nb, dsum, dsum + nb);
    HIP_TRYS(hipGetLastError());
    std::vector<double> h(2 * (size_t)nb);
    HIP_TRYS(hipMemcpyAsync(h.data(), dsum, h.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
    HIP_TRYS(hipStreamSynchronize(L->stream));
    std::vector<double> res(nb);
    for (int s = 0; s < nb; s++) res[s] = h[nb + s] > 0.0 ? h[s] / h[nb + s] : 0.0;
    for (int s = 0; s < nshells; s++) C[s] = res[s];
    // energy_large: i = 2..4, energy_medium: i = 5..10, energy_small: i = 11 .. nb (1-based)
    double all = 0.0, large = 0.0, medium = 0.0, small = 0.0;
    for (int s = 1; s <= 3 && s < nb; s++) {
        large = large + res[s];
        all = all + res[s];
    }
    for (int s = 4; s <= 9 && s < nb; s++) {
        medium = medium + res[s];
        all = all + res[s];
    }
    for (int s = 10; s < nb; s++) {
        small = small + res[s];
        all = all + res[s];
    }
    bands[0] = all * (-1.0);
    bands[1] = large * (-1.0);
    bands[2] = medium * (-1.0);
    bands[3] = small * (-1.0);
    return DRP_OK;
}
