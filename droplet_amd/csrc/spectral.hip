// spectral.hip — linear forcing of the low Fourier shells for forced isotropic
// turbulence (program/lbm_drop_iht_xy.f90:722-841, externalforce).
//
// The reference transforms u with a full 3-D r2c FFT (2decomp + FFTW), keeps the
// modes with 1 <= |k| < kc (integer mode numbers, kc = 3), computes their energy
//   E_low = 1/2 sum_{k in S} |u_hat_k|^2,   u_hat = FFT(u) / N^3,
// and transforms F_hat_k = (eps / (2 E_low)) u_hat_k (zero off S) back, so that
// the injected power <F . u> is eps by construction.
//
// Only the few modes of S (|k_a| <= K = kc - 1 per axis) enter, so here the
// transform is a truncated, separable DFT instead of three full FFTs:
//   rows_fwd  A(kx, y, z)   = sum_x u e^{-2 pi i kx x/nx}        kx = 0..K (u real:
//                                                               -kx is conj(kx))
//   cols_fwd  B(kx, ky, z)  = sum_y A e^{-2 pi i ky y/ny}        ky = -K..K
//   zsum      u_hat(k)      = sum_z B e^{-2 pi i kz z/nz} / N^3
//   coef      E_low and the coefficients C(k) = eps / (2 E_low) u_hat(k) on S
//   cols_inv  G(kx, ky, z)  = sum_kz C e^{+2 pi i kz z/nz}
//   rows_inv  H(kx, y, z)   = sum_ky G e^{+2 pi i ky y/ny}
//   force     F(x, y, z)    = Re H(0) + 2 sum_{kx>0} Re(H(kx) e^{+2 pi i kx x/nx})
// The first and last passes touch every node once (read u, write F); the others
// work on N^2 rows or less, one wavefront per output with the lanes over the
// summed axis.  E_low = 0 gives F = 0 (the reference would divide
// by zero).
//
// The same passes with unit coefficients on the shells klo <= int|k| < khi give
// the band-pass of u that advects phi in the scale-separated model
// (program/lbm_drop_scale_xy.f90:846-903, low_pass; u_hat = FFT(u)/N^3 there too).
#include "kernels.h"

namespace drp {

namespace {

constexpr int SF_BLOCK = 256;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// e^{s 2 pi i m / n} from the table tw[j] = (cos 2 pi j/n, sin 2 pi j/n), s = +-1.
// Every caller has |m| <= K (n - 1) with K <= 4, so m is reduced into [0, n) by a
// few compares instead of a (64-bit) integer remainder.
__device__ __forceinline__ double2 twiddle(const double2* tw, int m, int n, int s) {
    int j = m;
    while (j >= n) j -= n;
    while (j < 0) j += n;
    const double2 t = tw[j];
    return make_double2(t.x, s * t.y);
}

// Sum of NV per-lane values over the 64 lanes of a wavefront with NV - 1 + 1
// shuffles instead of 6 NV: at each offset o = 32 .. 2 a lane keeps one half of
// its values (the upper one when lane & o) and adds the partner's copy of that
// half; after the last halving lane l holds value l >> 1 (summed over its pair's
// 32 lanes) and one exchange with lane l ^ 1 completes it.  NV <= 32.
template <int O, int HALF> __device__ __forceinline__ void halve(double (&v)[32], int lane) {
    const bool up = (lane & O) != 0;
#pragma unroll
    for (int i = 0; i < HALF; i++) {
        const double keep = up ? v[i + HALF] : v[i];
        const double give = up ? v[i] : v[i + HALF];
        v[i] = keep + __shfl_xor(give, O, 64);
    }
    if constexpr (O > 2) halve<O / 2, HALF / 2>(v, lane);
}

template <int NV> __device__ __forceinline__ double wave_sums(double (&v)[32], int lane) {
    static_assert(NV <= 32, "at most 32 values");
#pragma unroll
    for (int i = NV; i < 32; i++) v[i] = 0.0;
    halve<32, 16>(v, lane);
    return v[0] + __shfl_xor(v[0], 1, 64);
}

// pass 1: one wavefront per (y, z) row; the loads of four x-iterations are issued
// together before their products are accumulated
template <int K, typename T>
__global__ __launch_bounds__(SF_BLOCK) void sf_rows_fwd(int nx, int ny, int nz, const T* __restrict__ u1,
                                                        const T* __restrict__ u2, const T* __restrict__ u3,
                                                        const double2* __restrict__ twx, double2* __restrict__ A) {
    constexpr int NV = 6 * (K + 1), CH = 4;
    const int lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * (SF_BLOCK / 64) + (threadIdx.x >> 6);
    if (row >= (long)ny * nz) return;  // uniform per wavefront
    const T* u[3] = {u1, u2, u3};
    double acc[32];  // [c][k][re, im]
#pragma unroll
    for (int i = 0; i < NV; i++) acc[i] = 0.0;
    for (int x0 = 0; x0 < nx; x0 += 64 * CH) {
        double v[CH][3];
        double2 e[CH][K + 1];
#pragma unroll
        for (int j = 0; j < CH; j++) {
            const int x = x0 + 64 * j + lane;
            const bool ok = x < nx;
            const long s = row * nx + (ok ? x : 0);
#pragma unroll
            for (int c = 0; c < 3; c++) v[j][c] = ok ? (double)u[c][s] : 0.0;
#pragma unroll
            for (int k = 0; k <= K; k++) e[j][k] = twiddle(twx, k * (ok ? x : 0), nx, -1);
        }
#pragma unroll
        for (int j = 0; j < CH; j++)
#pragma unroll
            for (int k = 0; k <= K; k++)
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    acc[(c * (K + 1) + k) * 2 + 0] += v[j][c] * e[j][k].x;
                    acc[(c * (K + 1) + k) * 2 + 1] += v[j][c] * e[j][k].y;
                }
    }
    const double t = wave_sums<NV>(acc, lane);
    const int idx = lane >> 1;
    if ((lane & 1) == 0 && idx < NV) {
        // idx = (c (K+1) + k) 2 + part
        double* out = reinterpret_cast<double*>(A + (row * 3) * (K + 1));
        out[idx] = t;
    }
}

// pass 2: one wavefront per (z, c, kx, ky), lanes over y
template <int K>
__global__ __launch_bounds__(SF_BLOCK) void sf_cols_fwd(int ny, int nz, const double2* __restrict__ A,
                                                        const double2* __restrict__ twy, double2* __restrict__ B) {
    constexpr int NK = 2 * K + 1, NB = 3 * (K + 1) * NK;
    const int lane = threadIdx.x & 63;
    const long w = (long)blockIdx.x * (SF_BLOCK / 64) + (threadIdx.x >> 6);
    if (w >= (long)nz * NB) return;  // uniform per wavefront
    const int b = (int)(w % NB), z = (int)(w / NB);
    const int ky = b % NK - K, kx = (b / NK) % (K + 1), c = b / (NK * (K + 1));
    double sr = 0.0, si = 0.0;
    for (int y = lane; y < ny; y += 64) {
        const double2 p = cmul(A[(((long)z * ny + y) * 3 + c) * (K + 1) + kx], twiddle(twy, ky * y, ny, -1));
        sr += p.x;
        si += p.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
    }
    if (lane == 0) B[w] = make_double2(sr, si);  // [z][c][kx][ky]
}

// pass 3a: one wavefront per mode (c, kx, ky, kz), lanes over z: u_hat into U
template <int K>
__global__ __launch_bounds__(SF_BLOCK) void sf_zsum(int nx, int ny, int nz, const double2* __restrict__ B,
                                                    const double2* __restrict__ twz, double2* __restrict__ U) {
    constexpr int NK = 2 * K + 1, NB = 3 * (K + 1) * NK, NC = NB * NK;
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * (SF_BLOCK / 64) + (threadIdx.x >> 6);
    if (t >= NC) return;  // uniform per wavefront
    const int kz = t % NK - K, b = t / NK;
    double sr = 0.0, si = 0.0;
    for (int z = lane; z < nz; z += 64) {
        const double2 p = cmul(B[(long)z * NB + b], twiddle(twz, kz * z, nz, -1));
        sr += p.x;
        si += p.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
    }
    const double inv = 1.0 / ((double)nx * ny * nz);
    if (lane == 0) U[t] = make_double2(sr * inv, si * inv);
}

// pass 3b: one workgroup; the coefficients of the kept modes k2lo <= |k|^2 < k2hi,
// in place, the others zeroed.  Forcing (band = 0): scaled by eps / (2 E_low) with
// E_low the energy of the kept modes; band-pass (band = 1): unscaled.
template <int K>
__global__ __launch_bounds__(SF_BLOCK) void sf_coef(int k2lo, int k2hi, int band, double eps,
                                                    double2* __restrict__ C, double* __restrict__ elow_out) {
    constexpr int NK = 2 * K + 1, NB = 3 * (K + 1) * NK, NC = NB * NK;
    __shared__ double2 s_u[NC];
    __shared__ double s_e[SF_BLOCK / 64];
    double e = 0.0;
    for (int t = threadIdx.x; t < NC; t += SF_BLOCK) {
        const double2 s = C[t];
        s_u[t] = s;
        const int kz = t % NK - K, b = t / NK;
        const int ky = b % NK - K, kx = (b / NK) % (K + 1);
        const int k2 = kx * kx + ky * ky + kz * kz;
        if (k2 >= k2lo && k2 < k2hi) e += (kx == 0 ? 0.5 : 1.0) * (s.x * s.x + s.y * s.y);
    }
    for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
    if ((threadIdx.x & 63) == 0) s_e[threadIdx.x >> 6] = e;
    __syncthreads();
    double elow = 0.0;
    for (int w = 0; w < SF_BLOCK / 64; w++) elow += s_e[w];
    const double alpha = band ? 1.0 : (elow > 0.0 ? 0.5 * eps / elow : 0.0);
    for (int t = threadIdx.x; t < NC; t += SF_BLOCK) {
        const int kz = t % NK - K, b = t / NK;
        const int ky = b % NK - K, kx = (b / NK) % (K + 1);
        const int k2 = kx * kx + ky * ky + kz * kz;
        const bool in = k2 >= k2lo && k2 < k2hi;
        C[t] = in ? make_double2(alpha * s_u[t].x, alpha * s_u[t].y) : make_double2(0.0, 0.0);
    }
    if (threadIdx.x == 0) *elow_out = elow;
}

// pass 4: G(c, kx, ky, z); one thread per output
template <int K>
__global__ __launch_bounds__(SF_BLOCK) void sf_cols_inv(int nz, const double2* __restrict__ C,
                                                        const double2* __restrict__ twz, double2* __restrict__ G) {
    constexpr int NK = 2 * K + 1, NB = 3 * (K + 1) * NK;
    const long t = blockIdx.x * (long)SF_BLOCK + threadIdx.x;
    if (t >= (long)nz * NB) return;
    const int b = (int)(t % NB), z = (int)(t / NB);
    double2 s = make_double2(0.0, 0.0);
    for (int kz = -K; kz <= K; kz++) {
        const double2 p = cmul(C[b * NK + (kz + K)], twiddle(twz, kz * z, nz, +1));
        s.x += p.x;
        s.y += p.y;
    }
    G[t] = s;  // [z][c][kx][ky]
}

// pass 5: H(c, kx, y, z); one thread per output
template <int K>
__global__ __launch_bounds__(SF_BLOCK) void sf_rows_inv(int ny, int nz, const double2* __restrict__ G,
                                                        const double2* __restrict__ twy, double2* __restrict__ H) {
    constexpr int NK = 2 * K + 1, NB = 3 * (K + 1) * NK, NH = 3 * (K + 1);
    const long t = blockIdx.x * (long)SF_BLOCK + threadIdx.x;
    if (t >= (long)ny * nz * NH) return;
    const int ck = (int)(t % NH);                  // c * (K+1) + kx
    const long row = t / NH;
    const int y = (int)(row % ny), z = (int)(row / ny);
    double2 s = make_double2(0.0, 0.0);
    for (int ky = -K; ky <= K; ky++) {
        const double2 p = cmul(G[(long)z * NB + ck * NK + (ky + K)], twiddle(twy, ky * y, ny, +1));
        s.x += p.x;
        s.y += p.y;
    }
    H[t] = s;  // [row][c][kx]
}

// pass 6: the force field, one wavefront per (y, z) row: the row's coefficients are
// the same for every lane (read once into scalar registers), the lanes run over x
template <int K>
__global__ __launch_bounds__(SF_BLOCK) void sf_force(int nx, int ny, int nz, const double2* __restrict__ H,
                                                     const double2* __restrict__ twx, double* __restrict__ fx,
                                                     double* __restrict__ fy, double* __restrict__ fz) {
    const int lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * (SF_BLOCK / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (row >= (long)ny * nz) return;  // uniform per wavefront
    double2 h[3][K + 1];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int k = 0; k <= K; k++) h[c][k] = H[(row * 3 + c) * (K + 1) + k];
    double* const f[3] = {fx, fy, fz};
    for (int x = lane; x < nx; x += 64) {
        double2 e[K + 1];
#pragma unroll
        for (int k = 1; k <= K; k++) e[k] = twiddle(twx, k * x, nx, +1);
        const long s = row * nx + x;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double v = h[c][0].x;
#pragma unroll
            for (int k = 1; k <= K; k++) v += 2.0 * (h[c][k].x * e[k].x - h[c][k].y * e[k].y);
            f[c][s] = v;
        }
    }
}

template <int K, typename T>
hipError_t spectral_k(const SpectralWork& w, int nx, int ny, int nz, int k2lo, int k2hi, int band, double eps,
                      const T* u1, const T* u2, const T* u3, double* fx, double* fy, double* fz, hipStream_t st) {
    constexpr int NK = 2 * K + 1, NB = 3 * (K + 1) * NK;
    const long rows = (long)ny * nz, n = rows * nx;
    hipLaunchKernelGGL((sf_rows_fwd<K, T>), dim3((rows + 3) / 4), dim3(SF_BLOCK), 0, st, nx, ny, nz, u1, u2, u3,
                       w.twx, w.A);
    constexpr int NC = NB * NK, WPB = SF_BLOCK / 64;
    hipLaunchKernelGGL((sf_cols_fwd<K>), dim3((nz * NB + WPB - 1) / WPB), dim3(SF_BLOCK), 0, st, ny, nz, w.A, w.twy,
                       w.B);
    hipLaunchKernelGGL((sf_zsum<K>), dim3((NC + WPB - 1) / WPB), dim3(SF_BLOCK), 0, st, nx, ny, nz, w.B, w.twz, w.C);
    hipLaunchKernelGGL((sf_coef<K>), dim3(1), dim3(SF_BLOCK), 0, st, k2lo, k2hi, band, eps, w.C, w.elow);
    hipLaunchKernelGGL((sf_cols_inv<K>), dim3((nz * NB + SF_BLOCK - 1) / SF_BLOCK), dim3(SF_BLOCK), 0, st, nz, w.C,
                       w.twz, w.G);
    hipLaunchKernelGGL((sf_rows_inv<K>), dim3((rows * 3 * (K + 1) + SF_BLOCK - 1) / SF_BLOCK), dim3(SF_BLOCK), 0, st,
                       ny, nz, w.G, w.twy, w.A);  // H reuses A
    hipLaunchKernelGGL((sf_force<K>), dim3((rows + 3) / 4), dim3(SF_BLOCK), 0, st, nx, ny, nz, w.A,
                       w.twx, fx, fy, fz);
    return hipGetLastError();
}

// modes |k_a| <= kc - 1 transformed; kept: k2lo <= |k|^2 < k2hi
template <typename T>
hipError_t spectral_dispatch(const SpectralWork& w, int nx, int ny, int nz, int kc, int k2lo, int k2hi, int band,
                             double eps, const T* u1, const T* u2, const T* u3, double* fx, double* fy, double* fz,
                             hipStream_t st) {
    switch (kc - 1) {
        case 1: return spectral_k<1, T>(w, nx, ny, nz, k2lo, k2hi, band, eps, u1, u2, u3, fx, fy, fz, st);
        case 2: return spectral_k<2, T>(w, nx, ny, nz, k2lo, k2hi, band, eps, u1, u2, u3, fx, fy, fz, st);
        case 3: return spectral_k<3, T>(w, nx, ny, nz, k2lo, k2hi, band, eps, u1, u2, u3, fx, fy, fz, st);
        case 4: return spectral_k<4, T>(w, nx, ny, nz, k2lo, k2hi, band, eps, u1, u2, u3, fx, fy, fz, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

size_t spectral_work_doubles(int nx, int ny, int nz, int kc) {
    const size_t K = (size_t)(kc - 1), NK = 2 * K + 1, NB = 3 * (K + 1) * NK;
    const size_t rows = (size_t)ny * nz;
    const size_t cplx = rows * 3 * (K + 1)      // A / H
                        + (size_t)nz * NB        // B
                        + NB * NK                // C
                        + (size_t)nz * NB        // G
                        + (size_t)nx + ny + nz;  // twiddle tables
    return 2 * cplx + 1;                         // + E_low
}

void spectral_work_layout(SpectralWork& w, double* base, int nx, int ny, int nz, int kc) {
    const size_t K = (size_t)(kc - 1), NK = 2 * K + 1, NB = 3 * (K + 1) * NK;
    double2* p = reinterpret_cast<double2*>(base);
    w.A = p;
    p += (size_t)ny * nz * 3 * (K + 1);
    w.B = p;
    p += (size_t)nz * NB;
    w.C = p;
    p += NB * NK;
    w.G = p;
    p += (size_t)nz * NB;
    w.twx = p;
    p += nx;
    w.twy = p;
    p += ny;
    w.twz = p;
    p += nz;
    w.elow = reinterpret_cast<double*>(p);
}

hipError_t launch_spectral_force(const SpectralWork& w, int nx, int ny, int nz, int kc, double eps, const double* u1,
                                 const double* u2, const double* u3, double* fx, double* fy, double* fz,
                                 hipStream_t st) {
    return spectral_dispatch<double>(w, nx, ny, nz, kc, 1, kc * kc, 0, eps, u1, u2, u3, fx, fy, fz, st);
}

hipError_t launch_spectral_force(const SpectralWork& w, int nx, int ny, int nz, int kc, double eps, const float* u1,
                                 const float* u2, const float* u3, double* fx, double* fy, double* fz,
                                 hipStream_t st) {
    return spectral_dispatch<float>(w, nx, ny, nz, kc, 1, kc * kc, 0, eps, u1, u2, u3, fx, fy, fz, st);
}

hipError_t launch_spectral_band(const SpectralWork& w, int nx, int ny, int nz, int klo, int khi, const double* u1,
                                const double* u2, const double* u3, double* o1, double* o2, double* o3,
                                hipStream_t st) {
    return spectral_dispatch<double>(w, nx, ny, nz, khi, klo * klo, khi * khi, 1, 0.0, u1, u2, u3, o1, o2, o3, st);
}

}  // namespace drp
