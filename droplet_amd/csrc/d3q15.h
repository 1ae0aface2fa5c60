// d3q15.h — D3Q15 lattice constants and per-direction equilibrium / collision
// arithmetic shared by the HIP kernels.
//
// Every expression keeps the evaluation order of the reference Fortran
// (program/lbm_drop_xy.f90:421-515, :541-553; prototype_program/two-phase/
// dropch.f90:666-698), with the terms that are identically zero for a given
// direction left out (adding +-0 to a non-zero partial sum leaves it unchanged),
// so that with FMA contraction off the FP64 kernels round exactly like the
// reference.  Directions are compile-time template arguments: the zero terms
// vanish at compile time instead of costing multiplies.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace drp {

// q = 0..14  <->  reference i = 1..15   (program/lbm_drop_xy.f90:70-84)
template <int Q> struct Dir;
#define DRP_DIR(Q, X, Y, Z, O)                                             \
    template <> struct Dir<Q> {                                            \
        static constexpr int cx = X, cy = Y, cz = Z, opp = O;              \
    };
DRP_DIR(0, 0, 0, 0, 0)
DRP_DIR(1, 1, 0, 0, 4)
DRP_DIR(2, 0, 1, 0, 5)
DRP_DIR(3, 0, 0, 1, 6)
DRP_DIR(4, -1, 0, 0, 1)
DRP_DIR(5, 0, -1, 0, 2)
DRP_DIR(6, 0, 0, -1, 3)
DRP_DIR(7, 1, 1, 1, 11)
DRP_DIR(8, -1, 1, 1, 12)
DRP_DIR(9, 1, -1, 1, 13)
DRP_DIR(10, 1, 1, -1, 14)
DRP_DIR(11, -1, -1, -1, 7)
DRP_DIR(12, 1, -1, -1, 8)
DRP_DIR(13, -1, 1, -1, 9)
DRP_DIR(14, -1, -1, 1, 10)
#undef DRP_DIR

template <int Q> __host__ __device__ constexpr double weightE() {
    return Q == 0 ? 2.0 / 9.0 : (Q <= 6 ? 1.0 / 9.0 : 1.0 / 72.0);
}
template <int Q> __host__ __device__ constexpr double weightF() {
    return Q == 0 ? -7.0 / 3.0 : (Q <= 6 ? 1.0 / 3.0 : 1.0 / 24.0);
}

// Correctly rounded x / c for c in {3, 5, 10} with r = RN(1/c): q = RN(x r) is
// faithful, so its remainder e = x - q c is exact in one FMA, and q + e r rounds
// to RN(x / c) because x / c never lies closer than ulp/20 to a rounding
// midpoint for these c (a midpoint would need c times a 54-bit odd integer to
// fit in 53 bits) while the correction is off by at most ulp * 2^-52.
__device__ __forceinline__ double div_const(double x, double c, double r) {
    const double q = x * r;
    const double e = __builtin_fma(-q, c, x);
    return __builtin_fma(e, r, q);
}

// x / d for a divisor d that many divisions share (a constant relaxation time).
// gfx950 divides doubles by the sequence: y = rcp(d) refined by two Newton
// steps, q = x y, r = fma(-d, q, x), q + r y, then a sign / special-value fixup,
// with range scalings of x and d that apply only when |x| < 2^-968, the exponents
// of x and d differ by 768 or more, or x / d or 1 / d is subnormal.  The steps on
// d alone are done once (SharedDivisor).  For 1/4 <= d <= 2^32 and x = +0 or
// 2^-967 <= |x| < 2^766 none of the scalings applies and the fixup returns q
// unchanged, so the three remaining steps (div_shared_fast) give exactly x / d;
// div_shared_exact tells whether x is such a value.  Callers test every x and
// redo the divisions of a node with the full division when one is not.
struct SharedDivisor {
    double d, y;
    bool ok;  // 1/4 <= d <= 2^32
};
__device__ __forceinline__ SharedDivisor shared_divisor(double d) {
    SharedDivisor s;
    s.d = d;
    double y = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-d, y, 1.0);
    s.y = __builtin_fma(y, e, y);
    s.ok = d >= 0.25 && d <= 0x1p32;
    return s;
}
__device__ __forceinline__ double div_shared_fast(double x, const SharedDivisor& s) {
    const double q = x * s.y;
    const double r = __builtin_fma(-s.d, q, x);
    return __builtin_fma(r, s.y, q);
}
__device__ __forceinline__ bool div_shared_exact(double x) {
    const double ax = __builtin_fabs(x);
    return (ax >= 0x1p-967 && ax < 0x1p766) || __builtin_amdgcn_class(x, 0x040 /* +0 */);
}

// signed accumulation helper: s (+|-) v, or s unchanged when c == 0
template <int C, typename T> __device__ __forceinline__ T acc(T s, T v) {
    if constexpr (C == 1) return s + v;
    else if constexpr (C == -1) return s - v;
    else return s;
}
// first term of a signed sum started from zero: 0 + c v
template <int C, typename T> __device__ __forceinline__ T first(T v) {
    if constexpr (C == 1) return v;
    else if constexpr (C == -1) return -v;
    else return T(0);
}

// c . v with the reference's order ((cx v1 + cy v2) + cz v3), zero terms dropped
template <int Q, typename T> __device__ __forceinline__ T cdot(T v1, T v2, T v3) {
    using D = Dir<Q>;
    if constexpr (D::cx != 0) {
        T s = first<D::cx>(v1);
        s = acc<D::cy>(s, v2);
        return acc<D::cz>(s, v3);
    } else if constexpr (D::cy != 0) {
        T s = first<D::cy>(v2);
        return acc<D::cz>(s, v3);
    } else if constexpr (D::cz != 0) {
        return first<D::cz>(v3);
    } else {
        return T(0);
    }
}

// Korteweg contraction gt = sum_b sum_a G_ab c_a c_b (beta outer, alpha inner,
// equilibrium_cal :439-449) with G_ab = (4.5 g_a) g_b - (1.5 s2) delta_ab
// (gphi_cal :642-666) and s2 = (g0^2 + g1^2) + g2^2.
template <int Q, typename T>
__device__ __forceinline__ T korteweg(T g0, T g1, T g2, T s2h /* 1.5*s2 */) {
    using D = Dir<Q>;
    constexpr int c[3] = {D::cx, D::cy, D::cz};
    const T g[3] = {g0, g1, g2};
    if constexpr (Q == 0) {
        return T(0);
    } else if constexpr (Q <= 6) {
        constexpr int k = D::cx != 0 ? 0 : (D::cy != 0 ? 1 : 2);
        return T(4.5) * g[k] * g[k] - s2h;
    } else {
        T s = T(4.5) * g[0] * g[0] - s2h;                      // (0,0)
#pragma unroll
        for (int t = 1; t < 9; t++) {                           // (a,b) = (t%3, t/3)
            const int a = t % 3, b = t / 3;
            T G = (a == b) ? (T(4.5) * g[a] * g[b] - s2h) : (T(4.5) * g[a] * g[b]);
            s = (c[a] * c[b] > 0) ? s + G : s - G;
        }
        return s;
    }
}

// Per-node quantities an upstream equilibrium needs.
template <typename T> struct NodeEq {
    T phi, p, u1, u2, u3;
    T A;          // F-bracket p0 - kf phi lap - kf/6 |grad phi|^2 (:453-455)
    T g0, g1, g2; // grad phi
    T tau;        // momentum relaxation time tau_g
};

// weight class of direction q (0 rest, 1 axis, 2 diagonal) and E_q kappa, the
// factor of gt in the equilibria: the reference forms E(i) * kappa first, so
// callers may pass that product computed once per class
template <int Q> __host__ __device__ constexpr int wclass() { return Q == 0 ? 0 : (Q <= 6 ? 1 : 2); }
template <int Q, typename T> __host__ __device__ __forceinline__ T ekappa(T k) { return T(weightE<Q>()) * k; }

// feq_q (:452-459): ((H phi + F A) + ((E phi) 3) cu) + (E kf) gt, ekf = E kf
template <int Q, typename T> __device__ __forceinline__ T feq(const NodeEq<T>& n, T gt, T ekf) {
    constexpr T E = T(weightE<Q>()), Fw = T(weightF<Q>());
    T s = (Q == 0) ? (n.phi + Fw * n.A) : (Fw * n.A);
    if constexpr (Q != 0) {
        T cu = cdot<Q>(n.u1, n.u2, n.u3);
        s = s + E * n.phi * T(3.0) * cu;
        s = s + ekf * gt;
    }
    return s;
}

// geq_q (:461-469): E (((3p + 3cu) + 4.5 cu^2) - 1.5 |u|^2) + (E kg) gt   (Anu = 0), ekg = E kg,
// from the node's 3p and 1.5 |u|^2 (per-node products a caller may form once)
template <int Q, typename T>
__device__ __forceinline__ T geq_pre(T p3, T u1, T u2, T u3, T usq15, T gt, T ekg) {
    constexpr T E = T(weightE<Q>());
    if constexpr (Q == 0) {
        return E * (p3 - usq15);
    } else {
        T cu = cdot<Q>(u1, u2, u3);
        T s = p3 + T(3.0) * cu;
        s = s + T(4.5) * (cu * cu);
        s = s - usq15;
        return E * s + ekg * gt;
    }
}
template <int Q, typename T> __device__ __forceinline__ T geq(const NodeEq<T>& n, T usq, T gt, T ekg) {
    return geq_pre<Q>(T(3.0) * n.p, n.u1, n.u2, n.u3, T(1.5) * usq, gt, ekg);
}

// order-preserving map of doubles to uint64 (for atomic min/max of the phi range)
__host__ __device__ inline uint64_t ordered_bits(double v) {
    uint64_t u = *reinterpret_cast<uint64_t*>(&v);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline double from_ordered_bits(uint64_t u) {
    uint64_t r = (u & 0x8000000000000000ull) ? (u & 0x7fffffffffffffffull) : ~u;
    return *reinterpret_cast<double*>(&r);
}

}  // namespace drp
