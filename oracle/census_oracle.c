/* census_oracle.c — CPU restatements of the droplet census (test
   infrastructure only: the tests use them as the checker of the GPU census,
   never as a product path).

   ora_census: clusters of nodes with phi >= thr connected through face
   neighbours, periodic in x, y, z — the definition the GPU census
   implements — by a sequential union-find; clusters reported by their
   smallest node index x + nx (y + ny z), with their volumes.

   ora_percolation_count: program/percolation.f90's own procedure, step by
   step (:103-258 labelling scan, :265-279 relabelling, :288-300 count): a
   y-outer, z, x-inner scan that looks at 7 already visited neighbours,
   provisional labels 1..1000 with the equivalence array klass, then the
   number of labels 1..1000 that still own a node.  Used to pin what the
   compiled reference writes (drop_num_1.d). */
#include <stdint.h>
#include <stdlib.h>

static int64_t root_of(int64_t* par, int64_t i) {
    while (par[i] != i) {
        par[i] = par[par[i]];
        i = par[i];
    }
    return i;
}

static void join(int64_t* par, int64_t a, int64_t b) {
    a = root_of(par, a);
    b = root_of(par, b);
    if (a == b) return;
    if (a < b) par[b] = a;
    else par[a] = b;
}

/* count clusters; first[k], volumes[k] for k < min(count, cap) ascending in
   first; labels (nullable): smallest index of the node's cluster or -1 */
int ora_census(const double* phi, int nx, int ny, int nz, double thr, int64_t* count, int64_t* volumes,
               int64_t* first, int64_t cap, int32_t* labels) {
    const int64_t n = (int64_t)nx * ny * nz;
    int64_t* par = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t* vol = (int64_t*)calloc((size_t)n, sizeof(int64_t));
    if (!par || !vol) {
        free(par);
        free(vol);
        return -1;
    }
    for (int64_t i = 0; i < n; i++) par[i] = phi[i] >= thr ? i : -1;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const int64_t i = x + (int64_t)nx * (y + (int64_t)ny * z);
                if (par[i] < 0) continue;
                const int64_t j[3] = {(x + 1) % nx + (int64_t)nx * (y + (int64_t)ny * z),
                                      x + (int64_t)nx * ((y + 1) % ny + (int64_t)ny * z),
                                      x + (int64_t)nx * (y + (int64_t)ny * ((z + 1) % nz))};
                for (int d = 0; d < 3; d++)
                    if (par[j[d]] >= 0) join(par, i, j[d]);
            }
    int64_t nc = 0;
    for (int64_t i = 0; i < n; i++) {
        if (par[i] < 0) {
            if (labels) labels[i] = -1;
            continue;
        }
        const int64_t r = root_of(par, i);
        vol[r]++;
        if (labels) labels[i] = (int32_t)r;
        if (r == i) nc++;
    }
    int64_t k = 0;
    for (int64_t i = 0; i < n && k < cap; i++)
        if (par[i] == i) {
            first[k] = i;
            volumes[k] = vol[i];
            k++;
        }
    *count = nc;
    free(par);
    free(vol);
    return 0;
}

#define KLASS_MAX 1000

/* chase klass from a to its fixed point (the `go to` loops of :146-257) */
static int chase(const int* klass, int a) {
    int b;
    do {
        b = a;
        a = klass[a];
    } while (a != b);
    return a;
}

/* The reference's count for one field (phi x-fastest [z][y][x]); returns -2
   if the scan needs more than 1000 provisional labels (the reference would
   write past klass(1:1000)). */
int ora_percolation_count(const double* phi, int nx, int ny, int nz, double thr, int64_t* dropnum) {
    const int xmax = nx - 1, ymax = ny - 1, zmax = nz - 1;
    int* label = (int*)calloc((size_t)nx * ny * nz, sizeof(int));
    int klass[KLASS_MAX + 1] = {0};
    if (!label) return -1;
#define LAB(x, y, z) label[(x) + (int64_t)nx * ((y) + (int64_t)ny * (z))]
#define PHI(x, y, z) phi[(x) + (int64_t)nx * ((y) + (int64_t)ny * (z))]
    int k = 1;
    for (int yi = 0; yi <= ymax; yi++)
        for (int zi = 0; zi <= zmax; zi++)
            for (int xi = 0; xi <= xmax; xi++) {
                if (!(PHI(xi, yi, zi) >= thr)) continue;
                int x_up = xi + 1, x_down = xi - 1, y_up = yi + 1, y_down = yi - 1, z_up = zi + 1, z_down = zi - 1;
                if (xi == 0) x_down = xmax;
                else if (xi == xmax) x_up = 0;
                if (yi == 0) y_down = ymax;
                else if (yi == ymax) y_up = 0;
                if (zi == 0) z_down = zmax;
                else if (zi == zmax) z_up = 0;
                (void)z_up;
                /* the 7 neighbours in the reference's order (:129-135) */
                const int nbx[7] = {x_down, xi, xi, xi, x_up, xi, x_down};
                const int nby[7] = {yi, y_down, yi, y_up, yi, y_down, yi};
                const int nbz[7] = {zi, zi, z_down, z_down, z_down, z_down, z_down};
                int tmp[7], sum = 0;
                for (int i = 0; i < 7; i++) {
                    tmp[i] = LAB(nbx[i], nby[i], nbz[i]);
                    sum += tmp[i];
                }
                int labeltmp = xmax * ymax * zmax;
                for (int i = 0; i < 7; i++)
                    if (tmp[i] > 0 && labeltmp > tmp[i]) labeltmp = tmp[i];
                if (sum == 0) {
                    if (k > KLASS_MAX) {
                        free(label);
                        return -2;
                    }
                    LAB(xi, yi, zi) = k;
                    klass[k] = k;
                    k++;
                } else {
                    LAB(xi, yi, zi) = labeltmp;
                    for (int i = 0; i < 7; i++) {
                        int* ln = &LAB(nbx[i], nby[i], nbz[i]);
                        if (*ln > 0) {
                            klass[*ln] = LAB(xi, yi, zi);
                            *ln = chase(klass, *ln);
                        }
                    }
                }
            }
    for (int zi = 0; zi <= zmax; zi++)
        for (int yi = 0; yi <= ymax; yi++)
            for (int xi = 0; xi <= xmax; xi++)
                if (PHI(xi, yi, zi) >= thr) LAB(xi, yi, zi) = chase(klass, LAB(xi, yi, zi));
    int64_t used[KLASS_MAX + 1] = {0};
    for (int64_t s = 0; s < (int64_t)nx * ny * nz; s++)
        if (label[s] >= 1 && label[s] <= KLASS_MAX) used[label[s]]++;
    int64_t c = 0;
    for (int i = 1; i <= KLASS_MAX; i++) c += used[i] > 0;
    *dropnum = c;
#undef LAB
#undef PHI
    free(label);
    return 0;
}
