// deformation.cpp — Taylor deformation D = (L-B)/(L+B) of a sheared droplet.
//
// The reference evaluates it on rank 0 after gathering phi
// (prototype_program/two-phase/dropch.f90:857-948, break_up3); here the one or
// two phi planes it reads are copied off the device and the same serial
// procedure runs on the host (it touches O(n^2) values once per call).
//
// Plane convention: the reference's walls are normal to z and the plane is
// (x, z) at y = ymax/2 (centroid) and y = (ymax+1)/2 (search).  With walls
// normal to y the roles of y and z swap.  Off-plane indices wrap periodically
// (the reference reads its stale 1-cell halo there).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "lattice_impl.h"

namespace {

struct Plane {
    std::vector<double> v;
    int n1 = 0, n2 = 0;
    double at(int i, int j) const {
        i %= n1;
        j %= n2;
        if (i < 0) i += n1;
        if (j < 0) j += n2;
        return v[(size_t)j * n1 + i];
    }
    // bilinear interpolation on the unit cell [int(x), int(x)+1] x [int(z), int(z)+1] (:908-919)
    double bilinear(double xj, double zj) const {
        const int x1 = (int)xj, z1 = (int)zj, x2 = x1 + 1, z2 = z1 + 1;
        const double w11 = ((double)x2 - xj) * ((double)z2 - zj), w21 = (xj - (double)x1) * ((double)z2 - zj);
        const double w12 = ((double)x2 - xj) * (zj - (double)z1), w22 = (xj - (double)x1) * (zj - (double)z1);
        return (w11 * at(x1, z1) + w21 * at(x2, z1) + w12 * at(x1, z2) + w22 * at(x2, z2)) /
               (((double)x2 - (double)x1) * ((double)z2 - (double)z1));
    }
};

}  // namespace

int copy_phi_plane(drp_lattice* L, int third, double* host, int* n1, int* n2) {
    const drp_config& c = L->cfg;
    const bool wall_y = c.wall_axis == DRP_WALL_Y;
    *n1 = c.nx;
    *n2 = wall_y ? c.ny : c.nz;
    const double* phi = nullptr;
    if (c.precision == DRP_FP64) {
        phi = (const double*)L->mac[L->cur];
    } else {
        int rc = fp32_macro_fp64_view(L, &phi);
        if (rc) return rc;
    }
    hipError_t e;
    if (wall_y) {  // plane (x, y) at z = third: contiguous
        e = hipMemcpyAsync(host, phi + (size_t)third * c.nx * c.ny, sizeof(double) * c.nx * c.ny,
                           hipMemcpyDeviceToHost, L->stream);
    } else {       // plane (x, z) at y = third: one row of nx per z
        e = hipMemcpy2DAsync(host, sizeof(double) * c.nx, phi + (size_t)third * c.nx, sizeof(double) * c.nx * c.ny,
                             sizeof(double) * c.nx, c.nz, hipMemcpyDeviceToHost, L->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
    if (e != hipSuccess) return drp_fail(DRP_ERR_HIP, "plane copy failed: %s", hipGetErrorString(e));
    return DRP_OK;
}

static int load(drp_lattice* L, int third, Plane& P) {
    const drp_config& c = L->cfg;
    P.v.resize((size_t)c.nx * (c.wall_axis == DRP_WALL_Y ? c.ny : c.nz));
    return copy_phi_plane(L, third, P.v.data(), &P.n1, &P.n2);
}

// break_up3 (dropch.f90:857-948)
static int breakup3(drp_lattice* L, double* Lo, double* Bo, double* Do) {
    const drp_config& c = L->cfg;
    const int n3 = (c.wall_axis == DRP_WALL_Y) ? c.nz : c.ny;
    const double pi = std::acos(-1.0), mid = 0.5 * (c.phi1 + c.phi2);
    Plane C, S;
    int rc = load(L, (n3 - 1) / 2, C);
    if (!rc) rc = load(L, n3 / 2, S);
    if (rc) return rc;
    // centroid weighted by phi - phi1 (:862-877)
    double xg = 0.0, zg = 0.0, w = 0.0;
    for (int j = 0; j < C.n2; j++)
        for (int i = 0; i < C.n1; i++) {
            const double d = C.at(i, j) - c.phi1;
            xg = xg + (double)i * d;
            zg = zg + (double)j * d;
            w = w + d;
        }
    xg = xg / w;
    zg = zg / w;
    // farthest interior point sets the orientation (:880-904)
    double dist = 0.0;
    int it = 0, jt = 0;
    for (int j = 0; j < S.n2; j++)
        for (int i = 0; i < S.n1; i++) {
            if (S.at(i, j) < mid) continue;
            const double r = std::sqrt(((double)i - xg) * ((double)i - xg) + ((double)j - zg) * ((double)j - zg));
            if (dist <= r) {
                dist = r;
                it = i;
                jt = j;
            }
        }
    if (dist == 0.0) return drp_fail(DRP_ERR_STATE, "no interior point in the deformation plane");
    double th = std::acos(std::fabs(((double)it - xg) / dist));
    const double dx = (double)it - xg, dz = (double)jt - zg;
    if (dx < 0.0 && dz > 0.0) th = pi - th;
    else if (dx < 0.0 && dz < 0.0) th = pi + th;
    else if (dx > 0.0 && dz < 0.0) th = 2.0 * pi - th;
    // ray-march the minor and major half-axes in steps of 0.1 (:906-946)
    const double den = std::tan(th) * std::sin(th) + std::cos(th);
    double B1 = 0.0, L1 = 0.0;
    for (int k = 1; k < 1000000; k++) {
        double zj = (0.1 * (double)k) / den;
        double xj = -zj * std::tan(th);
        xj = xj + xg;
        zj = zj + zg;
        const double v = S.bilinear(xj, zj);
        if (v < mid) break;
        B1 = std::sqrt((xj - xg) * (xj - xg) + (zj - zg) * (zj - zg)) * 2.0;
    }
    for (int k = 1; k < 1000000; k++) {
        double xj = (0.1 * (double)k) / den;
        double zj = xj * std::tan(th);
        xj = xj + xg;
        zj = zj + zg;
        const double v = S.bilinear(xj, zj);
        if (v < mid) break;
        L1 = std::sqrt((xj - xg) * (xj - xg) + (zj - zg) * (zj - zg)) * 2.0;
    }
    *Lo = L1;
    *Bo = B1;
    *Do = (L1 - B1) / (L1 + B1);
    return DRP_OK;
}

// L = 2 max r, B = 4 A / (pi L) (prototype_program/two-phase/drop.f90:439-463)
static int area_variant(drp_lattice* L, double* Lo, double* Bo, double* Do) {
    const drp_config& c = L->cfg;
    const int n3 = (c.wall_axis == DRP_WALL_Y) ? c.nz : c.ny;
    const double pi = std::acos(-1.0), mid = 0.5 * (c.phi1 + c.phi2);
    Plane S;
    int rc = load(L, n3 / 2, S);
    if (rc) return rc;
    double xg = 0.0, zg = 0.0, cnt = 0.0;
    for (int j = 0; j < S.n2; j++)
        for (int i = 0; i < S.n1; i++)
            if (S.at(i, j) >= mid) {
                xg += i;
                zg += j;
                cnt += 1.0;
            }
    if (cnt == 0.0) return drp_fail(DRP_ERR_STATE, "no interior point in the deformation plane");
    xg /= cnt;
    zg /= cnt;
    double rmax = 0.0;
    for (int j = 0; j < S.n2; j++)
        for (int i = 0; i < S.n1; i++)
            if (S.at(i, j) >= mid) rmax = std::fmax(rmax, std::sqrt((i - xg) * (i - xg) + (j - zg) * (j - zg)));
    const double Lx = 2.0 * rmax, Bx = 4.0 * cnt / (pi * Lx);
    *Lo = Lx;
    *Bo = Bx;
    *Do = (Lx - Bx) / (Lx + Bx);
    return DRP_OK;
}

// second moments of the search plane, L, B, D and optionally dD/dphi on the plane
static int inertia_core(drp_lattice* L, double* Lo, double* Bo, double* Do, std::vector<double>* grad, int* n1,
                       int* n2, int* third) {
    const drp_config& c = L->cfg;
    const int n3 = (c.wall_axis == DRP_WALL_Y) ? c.nz : c.ny;
    Plane S;
    int rc = load(L, n3 / 2, S);
    if (rc) return rc;
    const double inv = 1.0 / (c.phi2 - c.phi1);
    double W = 0.0, sx = 0.0, sz = 0.0;
    for (int j = 0; j < S.n2; j++)
        for (int i = 0; i < S.n1; i++) {
            const double sn = (S.at(i, j) - c.phi1) * inv, w = sn * sn;
            W += w;
            sx += w * i;
            sz += w * j;
        }
    if (W <= 0.0) return drp_fail(DRP_ERR_STATE, "no dispersed phase in the deformation plane");
    const double xc = sx / W, zc = sz / W;
    double ixx = 0.0, izz = 0.0, ixz = 0.0;
    for (int j = 0; j < S.n2; j++)
        for (int i = 0; i < S.n1; i++) {
            const double sn = (S.at(i, j) - c.phi1) * inv, w = sn * sn;
            const double dx = i - xc, dz = j - zc;
            ixx += w * dx * dx;
            izz += w * dz * dz;
            ixz += w * dx * dz;
        }
    const double m = 0.5 * (ixx + izz), r = std::sqrt(0.25 * (ixx - izz) * (ixx - izz) + ixz * ixz);
    const double lp = m + r, lm = std::fmax(m - r, 0.0);
    *Lo = 4.0 * std::sqrt(lp / W);
    *Bo = 4.0 * std::sqrt(lm / W);
    *Do = (*Lo - *Bo) / (*Lo + *Bo);
    if (!grad) return DRP_OK;
    // D = (a - b) / (a + b), a = sqrt(lp), b = sqrt(lm) (W cancels).  The centroid
    // terms drop out of dI/dw because sum w (i - xc) = 0, so
    // dD/dw(i, j) = Dxx (i - xc)^2 + Dzz (j - zc)^2 + Dxz (i - xc)(j - zc),
    // and dw/dphi = 2 s / (phi2 - phi1).
    const double a = std::sqrt(lp), b = std::sqrt(lm), s2 = (a + b) * (a + b);
    const double dDa = a > 0.0 ? 2.0 * b / s2 / (2.0 * a) : 0.0;   // dD/dlp
    const double dDb = b > 0.0 ? -2.0 * a / s2 / (2.0 * b) : 0.0;  // dD/dlm
    const double hd = r > 0.0 ? 0.25 * (ixx - izz) / r : 0.0, hx = r > 0.0 ? ixz / r : 0.0;
    const double Dxx = dDa * (0.5 + hd) + dDb * (0.5 - hd);
    const double Dzz = dDa * (0.5 - hd) + dDb * (0.5 + hd);
    const double Dxz = dDa * hx - dDb * hx;
    grad->assign((size_t)S.n1 * S.n2, 0.0);
    for (int j = 0; j < S.n2; j++)
        for (int i = 0; i < S.n1; i++) {
            const double sn = (S.at(i, j) - c.phi1) * inv;
            const double dx = i - xc, dz = j - zc;
            (*grad)[(size_t)j * S.n1 + i] = (Dxx * dx * dx + Dzz * dz * dz + Dxz * dx * dz) * 2.0 * sn * inv;
        }
    *n1 = S.n1;
    *n2 = S.n2;
    *third = n3 / 2;
    return DRP_OK;
}

// second-moment variant: weights w = s^2, s = (phi - phi1) / (phi2 - phi1), on the
// search plane (0 / 1 outside / inside a sharp drop, smooth in phi), centroid-relative second moments I; for a uniform ellipse with
// semi-axes a >= b the eigenvalues satisfy l+ / l- = (a / b)^2 and l+ = W a^2 / 4,
// so L = 4 sqrt(l+ / W), B = 4 sqrt(l- / W).  Continuous in phi, unlike the
// farthest-point and threshold-crossing steps of break_up3.
static int inertia_variant(drp_lattice* L, double* Lo, double* Bo, double* Do) {
    return inertia_core(L, Lo, Bo, Do, nullptr, nullptr, nullptr, nullptr);
}

int deformation_inertia_grad(drp_lattice* L, double* D, double* grad, int* n1, int* n2, int* third) {
    double Lo, Bo;
    std::vector<double> g;
    int rc = inertia_core(L, &Lo, &Bo, D, &g, n1, n2, third);
    if (rc) return rc;
    std::copy(g.begin(), g.end(), grad);
    return DRP_OK;
}

int deformation(drp_lattice* L, int mode, double* Lo, double* Bo, double* Do) {
    if (mode == 0) return breakup3(L, Lo, Bo, Do);
    if (mode == 1) return area_variant(L, Lo, Bo, Do);
    return inertia_variant(L, Lo, Bo, Do);
}
