// ref_files.cpp — the reference's on-disk formats: per-rank Fortran unformatted
// sequential files (4-byte record markers around every record).
//
//   output    <dir>/0_<rank>_<n>.bin      one record (phi, u1, u2, u3, p) per node
//             (program/lbm_drop_xy.f90:695-719)
//   outputfg  <dir>/0_<rank>_<n>_fg.bin   one record (f_i, g_i) per direction and node
//             (:721-743); read back for restarts (program/LBM.f90:371-562)
//   g input   <dir>/0_<rank>_<n>_fg.bin   one record g_i per direction and node, the
//             momentum populations injected at step_input (:789-828)
//   (f,g) input of program/LES_drop.f90 (:919-945): outputfg records, g kept
// Within a file the loops run z outermost, then y, then x (then i); rank r holds
// the block x in [rx xp, (rx+1) xp), y in [ry yp, (ry+1) yp), all z, with
// rx = r / ranks_y, ry = r % ranks_y, xp = nx / ranks_x, yp = ny / ranks_y
// (comm_rank_x = comm_rank / Ny, :145-149).  With DRP_FILE_LAYOUT_XZ the split is
// program/LES_drop.f90's (:153-176): x in [ix xp, ...), z in [iz zp, ...), all y,
// ix = r % ranks_x, iz = r / ranks_x, zp = nz / ranks_y (the second count is the
// number of z slabs there).
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/droplet.h"
#include "lattice_impl.h"

namespace {

std::string file_name(const char* dir, int rank, int64_t n, bool fg) {
    std::string d(dir);
    if (!d.empty() && d.back() != '/') d += '/';
    return d + "0_" + std::to_string(rank) + "_" + std::to_string((long long)n) + (fg ? "_fg.bin" : ".bin");
}

struct Block {
    int x0, y0, z0, xp, yp, zp;
};

int check_split(const drp_config& c, int rx, int ry, bool xz) {
    if (rx < 1 || ry < 1) return drp_fail(DRP_ERR_ARG, "rank counts must be positive");
    const int second = xz ? c.nz : c.ny;
    if (c.nx % rx || second % ry)
        return drp_fail(DRP_ERR_ARG, "%dx%d ranks do not divide %dx%d (x, %s)", rx, ry, c.nx, second, xz ? "z" : "y");
    return DRP_OK;
}

Block block_of(const drp_config& c, int rank, int rx, int ry, bool xz) {
    if (xz) {
        const int xp = c.nx / rx, zp = c.nz / ry;
        return Block{(rank % rx) * xp, 0, (rank / rx) * zp, xp, c.ny, zp};
    }
    const int xp = c.nx / rx, yp = c.ny / ry;
    return Block{(rank / ry) * xp, (rank % ry) * yp, 0, xp, yp, c.nz};
}

// one file of equal records of nval doubles, written / read in loop order
struct RecordFile {
    FILE* f = nullptr;
    int nval = 0;
    ~RecordFile() {
        if (f) std::fclose(f);
    }
    bool write(const double* v) {
        const int32_t m = 8 * nval;
        return std::fwrite(&m, 4, 1, f) == 1 && std::fwrite(v, 8, nval, f) == (size_t)nval &&
               std::fwrite(&m, 4, 1, f) == 1;
    }
    bool read(double* v) {
        int32_t a = 0, b = 0;
        if (std::fread(&a, 4, 1, f) != 1 || a != 8 * nval) return false;
        if (std::fread(v, 8, nval, f) != (size_t)nval) return false;
        return std::fread(&b, 4, 1, f) == 1 && b == a;
    }
};

}  // namespace

extern "C" int drp_write_reference_files(drp_lattice* L, const char* dir, int64_t n, int ranks_x, int ranks_y,
                                         int kind) {
    if (!L || !dir) return drp_fail(DRP_ERR_ARG, "null argument");
    const bool xz = (kind & DRP_FILE_LAYOUT_XZ) != 0;
    kind &= ~DRP_FILE_LAYOUT_XZ;
    if (kind != DRP_FILE_FIELDS && kind != DRP_FILE_FG) return drp_fail(DRP_ERR_ARG, "unknown file kind %d", kind);
    const drp_config& c = L->cfg;
    int rc = check_split(c, ranks_x, ranks_y, xz);
    if (rc) return rc;
    const size_t N = (size_t)L->n;
    std::vector<double> a, b;
    if (kind == DRP_FILE_FIELDS) {
        a.resize(5 * N);
        for (int k = 0; k < 5 && !rc; k++) rc = drp_get_macro(L, k, a.data() + k * N, N);
    } else {
        a.resize(15 * N);
        b.resize(15 * N);
        rc = drp_get_populations(L, DRP_POP_PHI, a.data(), 15 * N);
        if (!rc) rc = drp_get_populations(L, DRP_POP_MOM, b.data(), 15 * N);
    }
    if (rc) return rc;
    for (int r = 0; r < ranks_x * ranks_y; r++) {
        const Block B = block_of(c, r, ranks_x, ranks_y, xz);
        const std::string name = file_name(dir, r, n, kind == DRP_FILE_FG);
        RecordFile F;
        F.nval = kind == DRP_FILE_FIELDS ? 5 : 2;
        F.f = std::fopen(name.c_str(), "wb");
        if (!F.f) return drp_fail(DRP_ERR_ARG, "cannot write %s: %s", name.c_str(), std::strerror(errno));
        for (int z = B.z0; z < B.z0 + B.zp; z++)
            for (int y = B.y0; y < B.y0 + B.yp; y++)
                for (int x = B.x0; x < B.x0 + B.xp; x++) {
                    const size_t s = (size_t)x + (size_t)c.nx * ((size_t)y + (size_t)c.ny * z);
                    bool ok = true;
                    if (kind == DRP_FILE_FIELDS) {
                        // record order (phi, u1, u2, u3, p); fields are stored phi, p, u1, u2, u3
                        const double v[5] = {a[s], a[2 * N + s], a[3 * N + s], a[4 * N + s], a[N + s]};
                        ok = F.write(v);
                    } else {
                        for (int q = 0; q < 15 && ok; q++) {
                            const double v[2] = {a[q * N + s], b[q * N + s]};
                            ok = F.write(v);
                        }
                    }
                    if (!ok) return drp_fail(DRP_ERR_ARG, "write to %s failed", name.c_str());
                }
    }
    return DRP_OK;
}

extern "C" int drp_read_reference_files(drp_lattice* L, const char* dir, int64_t n, int ranks_x, int ranks_y,
                                        int kind) {
    if (!L || !dir) return drp_fail(DRP_ERR_ARG, "null argument");
    const bool xz = (kind & DRP_FILE_LAYOUT_XZ) != 0;
    kind &= ~DRP_FILE_LAYOUT_XZ;
    if (kind != DRP_FILE_FG && kind != DRP_FILE_G && kind != DRP_FILE_FG_G)
        return drp_fail(DRP_ERR_ARG, "unknown file kind %d", kind);
    if (!L->initialized) return drp_fail(DRP_ERR_STATE, "lattice not initialised");
    if (n < 0) return drp_fail(DRP_ERR_ARG, "file number %lld < 0", (long long)n);
    const drp_config& c = L->cfg;
    int rc = check_split(c, ranks_x, ranks_y, xz);
    if (rc) return rc;
    const size_t N = (size_t)L->n;
    std::vector<double> f(kind == DRP_FILE_FG ? 15 * N : 0), g(15 * N);
    for (int r = 0; r < ranks_x * ranks_y; r++) {
        const Block B = block_of(c, r, ranks_x, ranks_y, xz);
        const std::string name = file_name(dir, r, n, true);
        RecordFile F;
        F.nval = kind == DRP_FILE_G ? 1 : 2;
        F.f = std::fopen(name.c_str(), "rb");
        if (!F.f) return drp_fail(DRP_ERR_ARG, "cannot read %s: %s", name.c_str(), std::strerror(errno));
        for (int z = B.z0; z < B.z0 + B.zp; z++)
            for (int y = B.y0; y < B.y0 + B.yp; y++)
                for (int x = B.x0; x < B.x0 + B.xp; x++) {
                    const size_t s = (size_t)x + (size_t)c.nx * ((size_t)y + (size_t)c.ny * z);
                    for (int q = 0; q < 15; q++) {
                        double v[2];
                        if (!F.read(v)) return drp_fail(DRP_ERR_ARG, "%s: short file or bad record marker", name.c_str());
                        if (kind == DRP_FILE_FG) {
                            f[q * N + s] = v[0];
                            g[q * N + s] = v[1];
                        } else {
                            g[q * N + s] = kind == DRP_FILE_FG_G ? v[1] : v[0];   // FG_G: f_i is read and dropped
                        }
                    }
                }
        if (std::fgetc(F.f) != EOF) return drp_fail(DRP_ERR_ARG, "%s: trailing data (wrong grid or rank split)", name.c_str());
    }
    if (kind == DRP_FILE_FG) {
        // a restart: the state of step n, which the next step continues from
        rc = drp_set_populations(L, DRP_POP_PHI, f.data(), 15 * N);
        if (!rc) rc = drp_set_populations(L, DRP_POP_MOM, g.data(), 15 * N);
        if (rc) return rc;
        std::vector<double> phi(N);
        rc = drp_get_macro(L, DRP_FIELD_PHI, phi.data(), N);
        if (rc) return rc;
        double mx = c.phi1, mn = c.phi2;   // physics' scan starts from (phi1, phi2) (:557-574)
        for (double v : phi) {
            mx = v > mx ? v : mx;
            mn = v < mn ? v : mn;
        }
        L->step = n;
        const int s0 = (int)(n % 3);
        if (drp::launch_range_set(L->range, s0, mx, mn, L->stream) != hipSuccess ||
            drp::launch_range_set(L->range, (s0 + 1) % 3, c.phi1, c.phi2, L->stream) != hipSuccess ||
            drp::launch_range_set(L->range, (s0 + 2) % 3, c.phi1, c.phi2, L->stream) != hipSuccess ||
            hipStreamSynchronize(L->stream) != hipSuccess)
            return drp_fail(DRP_ERR_HIP, "phi range reset failed");
        return les_update_tau(L, true);
    }
    return drp_inject_momentum(L, g.data(), 15 * N);
}
