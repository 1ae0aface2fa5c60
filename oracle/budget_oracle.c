/* budget_oracle.c — CPU restatements of the derived fields of the droplet
   analysis programs and of the per-step kinetic-energy budget.

   TEST INFRASTRUCTURE ONLY: the tests use these as the checker of the GPU entry
   points drp_get_derived and drp_step_energy_budget; the product library never
   links or calls them.

   Fields are [z][y][x] (x fastest); stencils wrap periodically (the reference's
   glue halos); ds = 1.  Every expression keeps the reference's evaluation order
   (Fortran evaluates a chain of + or * left to right, ** binds first).

   ora_derived(kind):
     0 chemical potential, chemical_potencial_cal, program/drop_analysis.f90:413-425:
         mu = (T ln(phi/(1 - b phi)) + T/(1 - b phi) - 2 a phi) - kappa_g lap phi,
         lap by lap_cal (:383-398)
     1 free-energy density, free_energy_cal (:400-411): phi T ln(phi/(1 - b phi)) - a phi^2
     2 advection, advection_cal (:427-441): u . grad phi, grad by grad_cal (:362-381)

   ora_budget_terms: the seven terms of program/energy_budget.f90 (:1251-1700) for
   one step, from the fields at the start of the step ("old": phi and u, whose
   stencils various_cal (:777-827) and gphi_cal (:829-853) take) and at its end
   ("new": p and u after physics, nu of the new phi with the new range, :760-772),
   and the body force of the step:
     [0] left1  = sum |p - (p0 - kappa_g phi lap phi - kappa_g/6 |grad phi|^2)|   (:1253-1268)
                  with p0, lap phi, grad phi of the old phi and p, phi new (the reference
                  reads phi_procs after physics has updated it)
     [1] left2  = sum u . grad K, K = |u|^2/2 of the new state                    (:1270-1300)
     [2] right1 = -sum u . grad p                                                 (:1325-1372)
     [3] right2 = sum 2 nu sum_ab d_b (u_a S_ab), S = (grad u + grad u^T)/2 (old) (:1374-1486)
     [4] right3 = -sum 2 nu S:S                                                    (:1488-1500)
     [5] right4 = -(2/9) kappa_g sum_ab u_a d_b G_ab,
                  G = 4.5 grad phi grad phi - 1.5 |grad phi|^2 I (old)             (:1502-1650)
     [6] right5 = sum u . F                                                        (:1667-1675)
   Two departures, also stated with the entry point: the reference never assigns
   kinetic_procs in its active code (the assignment is commented out, :1279-1288), so
   its left2 is identically zero; here K is that commented definition.  Its
   right3_procs array is never reset between steps, so its term accumulates over
   the run; here each step's term stands alone.

   ORACLE_PLANT 6 drops the minus sign of right1, 7 drops T/(1 - b phi) from the
   chemical potential (planted copies are built by the tests only). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "../include/droplet.h"

#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

/* D3Q15 directions, program/lbm_drop_xy.f90:70-84 (i = 1..15 -> 0..14) */
static const int DX[15] = {0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1};
static const int DY[15] = {0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1};
static const int DZ[15] = {0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1};

typedef struct {
    int nx, ny, nz;
} grid_t;

static int wrapi(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

static int64_t at(const grid_t* g, int x, int y, int z) {
    return (int64_t)wrapi(x, g->nx) + (int64_t)g->nx * ((int64_t)wrapi(y, g->ny) + (int64_t)g->ny * wrapi(z, g->nz));
}

/* grad_cal: sum over i = 2..15 of cr(a, i) v(x + c_i), over 10 ds */
static double grad_at(const grid_t* g, const double* v, int x, int y, int z, int a) {
    const int* c = a == 0 ? DX : (a == 1 ? DY : DZ);
    double s = 0.0;
    for (int i = 1; i < 15; i++) s = s + (double)c[i] * v[at(g, x + DX[i], y + DY[i], z + DZ[i])];
    return s / 10.0;
}

/* lap_cal: (-14 v(x) + sum over i = 2..15 of v(x + c_i)) / (5 ds^2) */
static double lap_at(const grid_t* g, const double* v, int x, int y, int z) {
    double s = -14.0 * v[at(g, x, y, z)];
    for (int i = 1; i < 15; i++) s = s + v[at(g, x + DX[i], y + DY[i], z + DZ[i])];
    return s / 5.0;
}

int ora_derived(const drp_config* c, const double* mac, int kind, double* out) {
    if (!c || !mac || !out || kind < 0 || kind > 2) return DRP_ERR_ARG;
    const grid_t g = {c->nx, c->ny, c->nz};
    const int64_t n = (int64_t)c->nx * c->ny * c->nz;
    const double *phi = mac, *u1 = mac + 2 * n, *u2 = mac + 3 * n, *u3 = mac + 4 * n;
    for (int z = 0; z < c->nz; z++)
        for (int y = 0; y < c->ny; y++)
            for (int x = 0; x < c->nx; x++) {
                const int64_t s = at(&g, x, y, z);
                const double ph = phi[s];
                if (kind == 0) {
                    const double lap = lap_at(&g, phi, x, y, z);
                    const double bulk = ORACLE_PLANT == 7
                                            ? c->T * log(ph / (1.0 - c->b * ph)) - 2.0 * c->a * ph
                                            : c->T * log(ph / (1.0 - c->b * ph)) + c->T / (1.0 - c->b * ph) - 2.0 * c->a * ph;
                    out[s] = bulk - c->kappa_g * lap;
                } else if (kind == 1) {
                    out[s] = ph * c->T * log(ph / (1.0 - c->b * ph)) - c->a * ph * ph;
                } else {
                    out[s] = u1[s] * grad_at(&g, phi, x, y, z, 0) + u2[s] * grad_at(&g, phi, x, y, z, 1) +
                             u3[s] * grad_at(&g, phi, x, y, z, 2);
                }
            }
    return 0;
}

int ora_budget_terms(const drp_config* c, const double* old_mac, const double* new_mac, const double* force,
                     double phi_min, double phi_max, double* terms) {
    if (!c || !old_mac || !new_mac || !terms) return DRP_ERR_ARG;
    const grid_t g = {c->nx, c->ny, c->nz};
    const int64_t n = (int64_t)c->nx * c->ny * c->nz;
    const double* phio = old_mac;
    const double* uo[3] = {old_mac + 2 * n, old_mac + 3 * n, old_mac + 4 * n};
    const double *phin = new_mac, *pn = new_mac + n;
    const double* un[3] = {new_mac + 2 * n, new_mac + 3 * n, new_mac + 4 * n};
    double* K = (double*)malloc(sizeof(double) * (size_t)n);
    double* nu = (double*)malloc(sizeof(double) * (size_t)n);
    double* T = (double*)malloc(sizeof(double) * (size_t)(9 * n)); /* T[X*3+Y] = u_X S_XY */
    double* G = (double*)malloc(sizeof(double) * (size_t)(9 * n)); /* G[X*3+Y] */
    if (!K || !nu || !T || !G) {
        free(K); free(nu); free(T); free(G);
        return DRP_ERR_ALLOC;
    }
    double left1 = 0.0, left2 = 0.0, right1 = 0.0, right2 = 0.0, right3 = 0.0, right4 = 0.0, right5 = 0.0;
    const double rg = phi_max - phi_min;
    for (int z = 0; z < c->nz; z++)
        for (int y = 0; y < c->ny; y++)
            for (int x = 0; x < c->nx; x++) {
                const int64_t s = at(&g, x, y, z);
                /* various_cal on the old fields */
                const double ph = phio[s];
                const double p0 = ph * c->T / (1.0 - c->b * ph) - c->a * (ph * ph);
                const double lap = lap_at(&g, phio, x, y, z);
                double gp[3], gu[3][3];
                for (int a = 0; a < 3; a++) {
                    gp[a] = grad_at(&g, phio, x, y, z, a);
                    for (int k = 0; k < 3; k++) gu[k][a] = grad_at(&g, uo[k], x, y, z, a);
                }
                const double gsq = gp[0] * gp[0] + gp[1] * gp[1] + gp[2] * gp[2];
                /* left1: kinetic_former (:1253-1259); p0, lap phi and grad phi are those of
                   various_cal (old phi), phi_procs has been updated by physics (new phi) */
                const double former = p0 - c->kappa_g * phin[s] * lap - 1.0 / 6.0 * c->kappa_g * gsq;
                left1 = left1 + fabs(pn[s] - former);
                /* strain_cal (:976-996) */
                double S[3][3];
                for (int b = 0; b < 3; b++)
                    for (int a = 0; a < 3; a++) S[a][b] = 0.5 * (gu[a][b] + gu[b][a]);
                /* nu of physics (:766) with the new range; 0/0 guarded */
                const double frac = rg == 0.0 ? 0.0 : (phin[s] - phi_min) / rg;
                nu[s] = frac * (c->nu2 - c->nu1) + c->nu1;
                /* right3 (:1488-1500) */
                double r3 = 0.0;
                for (int b = 0; b < 3; b++)
                    for (int a = 0; a < 3; a++) r3 = r3 + S[a][b] * S[a][b];
                right3 = right3 + r3 * (-2.0) * nu[s];
                /* right5 (:1667-1675) */
                if (force)
                    right5 = right5 + (un[0][s] * force[s] + un[1][s] * force[n + s] + un[2][s] * force[2 * n + s]);
                /* K, u_X S_XY (:1380-1408), G = gphi (:829-853) */
                K[s] = 0.5 * (un[0][s] * un[0][s] + un[1][s] * un[1][s] + un[2][s] * un[2][s]);
                for (int X = 0; X < 3; X++)
                    for (int Y = 0; Y < 3; Y++) {
                        T[(X * 3 + Y) * n + s] = un[X][s] * S[X][Y];
                        G[(X * 3 + Y) * n + s] = 4.5 * gp[X] * gp[Y] - 1.5 * gsq * (X == Y ? 1.0 : 0.0);
                    }
            }
    for (int z = 0; z < c->nz; z++)
        for (int y = 0; y < c->ny; y++)
            for (int x = 0; x < c->nx; x++) {
                const int64_t s = at(&g, x, y, z);
                const double u[3] = {un[0][s], un[1][s], un[2][s]};
                /* left2 (:1290-1300) */
                left2 = left2 + (u[0] * grad_at(&g, K, x, y, z, 0) + u[1] * grad_at(&g, K, x, y, z, 1) +
                                 u[2] * grad_at(&g, K, x, y, z, 2));
                /* right1 (:1325-1372): A = p */
                const double r1 = u[0] * grad_at(&g, pn, x, y, z, 0) + u[1] * grad_at(&g, pn, x, y, z, 1) +
                                  u[2] * grad_at(&g, pn, x, y, z, 2);
                right1 = right1 + (ORACLE_PLANT == 6 ? r1 : -1.0 * r1);
                /* right2 (:1420-1486): Y outer, X inner */
                double r2 = 0.0, r4 = 0.0;
                for (int Y = 0; Y < 3; Y++)
                    for (int X = 0; X < 3; X++) {
                        r2 = r2 + grad_at(&g, T + (X * 3 + Y) * n, x, y, z, Y);
                        r4 = r4 + u[X] * grad_at(&g, G + (X * 3 + Y) * n, x, y, z, Y);
                    }
                right2 = right2 + r2 * 2.0 * nu[s];
                /* right4 (:1572-1650) */
                right4 = right4 + r4 * (-1.0) * (2.0 / 9.0) * c->kappa_g;
            }
    terms[0] = left1;
    terms[1] = left2;
    terms[2] = right1;
    terms[3] = right2;
    terms[4] = right3;
    terms[5] = right4;
    terms[6] = right5;
    free(K); free(nu); free(T); free(G);
    return 0;
}
