// hc_output.cpp -- ASCII frames in the Clawpack 4.6 fort.qXXXX / fort.tXXXX
// format that ForestClaw writes (fclaw2d_clawpatch_output_ascii,
// src/patches/clawpatch/fclaw2d_clawpatch_output_ascii.c, with the Fortran
// writers of src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_output_ascii.f),
// so that VisClaw / Matlab readers and frame diffs work on this solver's
// output.  Host code only: the context's solution is downloaded once per frame.
//
// Record layouts (Fortran edit descriptors of the reference):
//   fort.tXXXX  time E30.20, then meqn, ngrids, num_aux, num_dim (I5) with labels
//   fort.qXXXX  per patch: grid_number, AMR_level, block_number, mpi_rank, mx, my
//               (I5), xlow, ylow, dx, dy (E24.16), an empty record, then for
//               j = 1..my, i = 1..mx one record of the meqn values (E26.16)
//               and after every j a list-directed ' ' record.
// Values with |q| < 1e-99 are written as zero, as the reference does (it
// also zeroes them in the patch itself; here the device solution is left
// unchanged).  The list-directed blank record is written as two blanks, the
// form of the Fortran runtime of this toolchain (amdflang).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hipclaw.h"
#include "hc_internal.hpp"

using hc::set_error;

namespace {

// Fortran Ew.d output of a double: [-]0.<d digits>E[+-]ee (|exponent| <= 99)
// or [-]0.<d digits>[+-]eee, right-justified in w columns.
void put_e(std::string& out, double x, int w, int d)
{
    std::string s;
    if (x == 0.0) {
        s = std::signbit(x) ? "-0." : "0.";
        s.append(d, '0');
        s += "E+00";
    } else {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.*e", d - 1, x);  // [-]D.DDDe[+-]XX, d digits
        const char* p = buf;
        const bool neg = *p == '-';
        if (neg) p++;
        const char* dot = std::strchr(p, '.');
        const char* ep = std::strchr(p, 'e');
        s = neg ? "-0." : "0.";
        s += p[0];
        if (dot && dot < ep) s.append(dot + 1, ep);
        const int ex = std::atoi(ep + 1) + 1;
        char eb[8];
        if (std::abs(ex) <= 99)
            std::snprintf(eb, sizeof eb, "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
        else
            std::snprintf(eb, sizeof eb, "%c%03d", ex < 0 ? '-' : '+', std::abs(ex));
        s += eb;
    }
    if ((int)s.size() < w) out.append(w - s.size(), ' ');
    out += s;
}

void put_i5(std::string& out, int v, const char* label)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%5d%s\n", v, label);
    out += buf;
}

std::string frame_name(const char* dir, char kind, int iframe)
{
    char buf[32];
    std::snprintf(buf, sizeof buf, "fort.%c%04d", kind, iframe);
    std::string path = dir && *dir ? std::string(dir) : std::string(".");
    return path + "/" + buf;
}

void patch_text(std::string& out, int mx, int my, int meqn, int mbc, double xlower, double ylower,
                double dx, double dy, const double* q, int patch_num, int level, int blockno, int mpirank)
{
    put_i5(out, patch_num, "                 grid_number");
    put_i5(out, level, "                 AMR_level");
    put_i5(out, blockno, "                 block_number");
    put_i5(out, mpirank, "                 mpi_rank");
    put_i5(out, mx, "                 mx");
    put_i5(out, my, "                 my");
    put_e(out, xlower, 24, 16);
    out += "    xlow\n";
    put_e(out, ylower, 24, 16);
    out += "    ylow\n";
    put_e(out, dx, 24, 16);
    out += "    dx\n";
    put_e(out, dy, 24, 16);
    out += "    dy\n\n";
    const int nx = mx + 2 * mbc, ny = my + 2 * mbc;
    for (int j = 0; j < my; j++) {
        for (int i = 0; i < mx; i++) {
            for (int m = 0; m < meqn; m++) {
                double v = q[((size_t)m * ny + (j + mbc)) * nx + (i + mbc)];
                if (std::fabs(v) < 1e-99) v = 0.0;
                put_e(out, v, 26, 16);
            }
            out += "\n";
        }
        out += "  \n";
    }
}

int write_file(const std::string& path, const std::string& text, const char* mode)
{
    FILE* f = std::fopen(path.c_str(), mode);
    if (!f) return set_error(HIPCLAW_ERR_ARG, ("cannot open " + path).c_str());
    const size_t n = std::fwrite(text.data(), 1, text.size(), f);
    const int rc = std::fclose(f);
    if (n != text.size() || rc != 0) return set_error(HIPCLAW_ERR_ARG, ("cannot write " + path).c_str());
    return HIPCLAW_OK;
}

uint64_t interleave(uint32_t x, uint32_t y)
{
    uint64_t k = 0;
    for (int b = 0; b < 32; b++) k |= (uint64_t)((x >> b) & 1) << (2 * b) | (uint64_t)((y >> b) & 1) << (2 * b + 1);
    return k;
}

}  // namespace

extern "C" {

int hipclaw_ascii_header(const char* dir, int iframe, double time, int meqn, int maux, int ngrids)
{
    if (iframe < 0 || iframe > 9999) return set_error(HIPCLAW_ERR_ARG, "frame number 0..9999");
    std::string t;
    put_e(t, time, 30, 20);
    t += "    time\n";
    put_i5(t, meqn, "                 meqn");
    put_i5(t, ngrids, "                 ngrids");
    put_i5(t, maux, "                 num_aux");
    put_i5(t, 2, "                 num_dim");
    int rc = write_file(frame_name(dir, 't', iframe), t, "w");
    if (!rc) rc = write_file(frame_name(dir, 'q', iframe), std::string(), "w");
    return rc;
}

int hipclaw_ascii_patch(const char* dir, int iframe, int mx, int my, int meqn, int mbc, double xlower,
                        double ylower, double dx, double dy, const double* q, int patch_num, int level,
                        int blockno, int mpirank)
{
    if (iframe < 0 || iframe > 9999) return set_error(HIPCLAW_ERR_ARG, "frame number 0..9999");
    if (!q || mx < 1 || my < 1 || mbc < 0) return set_error(HIPCLAW_ERR_ARG, "bad patch");
    if (meqn < 1 || meqn > 20) return set_error(HIPCLAW_ERR_ARG, "meqn must be 1..20 (format 20E26.16)");
    std::string out;
    patch_text(out, mx, my, meqn, mbc, xlower, ylower, dx, dy, q, patch_num, level, blockno, mpirank);
    return write_file(frame_name(dir, 'q', iframe), out, "a");
}

}  // extern "C"

namespace hc {

int write_frame(const char* dir, int iframe, double time, int mi, int mx, int my, int meqn, int maux, int n,
                const int* lev, const int* gx, const int* gy, const double* xl, const double* yl,
                const double* dx, const double* dy, const double* q)
{
    const int mbc = 2;
    if (meqn < 1 || meqn > 20) return set_error(HIPCLAW_ERR_ARG, "meqn must be 1..20 (format 20E26.16)");
    const size_t per = (size_t)(mx + 2 * mbc) * (my + 2 * mbc) * meqn;
    // the reference's patch order: blocks in order, leaves of a block in
    // Morton order (p4est's quadrant order), numbered from 0
    int maxl = 0;
    for (int k = 0; k < n; k++) maxl = std::max(maxl, lev[k]);
    std::vector<int> order(n), block(n);
    std::vector<uint64_t> key(n);
    for (int k = 0; k < n; k++) {
        const int L = lev[k];
        const int bx = gx[k] >> L, by = gy[k] >> L;
        block[k] = by * mi + bx;
        const uint32_t lx = (uint32_t)(gx[k] - (bx << L)) << (maxl - L);
        const uint32_t ly = (uint32_t)(gy[k] - (by << L)) << (maxl - L);
        key[k] = interleave(lx, ly);
        order[k] = k;
    }
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        return block[a] != block[b] ? block[a] < block[b] : key[a] < key[b];
    });
    int rc = hipclaw_ascii_header(dir, iframe, time, meqn, maux, n);
    if (rc) return rc;
    std::string out;
    out.reserve((size_t)n * (size_t)mx * my * (26 * meqn + 1));
    for (int r = 0; r < n; r++) {
        const int k = order[r];
        patch_text(out, mx, my, meqn, mbc, xl[k], yl[k], dx[k], dy[k], q + (size_t)k * per, r, lev[k],
                   block[k], 0);
    }
    return write_file(frame_name(dir, 'q', iframe), out, "a");
}

}  // namespace hc
