// hc_output.cpp -- ASCII frames in the Clawpack 4.6 fort.qXXXX / fort.tXXXX
// format that ForestClaw writes (fclaw2d_clawpatch_output_ascii,
// src/patches/clawpatch/fclaw2d_clawpatch_output_ascii.c, with the Fortran
// writers of src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_output_ascii.f),
// so that VisClaw / Matlab readers and frame diffs work on this solver's
// output, and the frame diff of two runs (hipclaw_compare_frames).  Host code
// only: the context's solution is downloaded once per frame.
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

// ---------------------------------------------------------------------------
// Frame diff (scripts/compare_files/compare_files.f90): two runs' frames of
// the same mesh, cell by cell.
// ---------------------------------------------------------------------------
namespace {

struct Lines {
    std::vector<std::string> v;
    size_t k = 0;
    bool load(const std::string& path)
    {
        FILE* f = std::fopen(path.c_str(), "r");
        if (!f) return false;
        std::string cur;
        int c;
        while ((c = std::fgetc(f)) != EOF) {
            if (c == '\n') {
                v.push_back(cur);
                cur.clear();
            } else {
                cur += (char)c;
            }
        }
        if (!cur.empty()) v.push_back(cur);
        std::fclose(f);
        return true;
    }
    // the next record with a value (list-directed input skips blank records)
    const char* next()
    {
        while (k < v.size()) {
            const std::string& l = v[k++];
            if (l.find_first_not_of(" \t\r") != std::string::npos) return l.c_str();
        }
        return nullptr;
    }
};

// one value of an E-edit field: [-]0.dddE[+-]ee or [-]0.ddd[+-]eee
bool read_real(const char*& p, double& x)
{
    while (*p == ' ' || *p == '\t' || *p == ',') p++;
    if (!*p) return false;
    char* e = nullptr;
    x = std::strtod(p, &e);
    if (e == p) return false;
    if ((*e == '+' || *e == '-') && e[1] >= '0' && e[1] <= '9') {  // three-digit exponent, no E
        char* e2 = nullptr;
        const long ex = std::strtol(e, &e2, 10);
        x *= std::pow(10.0, (double)ex);
        e = e2;
    }
    p = e;
    return true;
}

bool first_int(Lines& L, int& v)
{
    const char* l = L.next();
    if (!l) return false;
    char* e = nullptr;
    v = (int)std::strtol(l, &e, 10);
    return e != l;
}

bool first_real(Lines& L, double& x)
{
    const char* l = L.next();
    return l && read_real(l, x);
}

}  // namespace

extern "C" {

int hipclaw_compare_frames(const char* dir1, const char* dir2, const char* dir3, int iframe, int nmax,
                           double* global_max, int* grid_max)
{
    if (iframe < 0 || iframe > 9999) return set_error(HIPCLAW_ERR_ARG, "frame number 0..9999");
    Lines t1, t2, q1, q2;
    const std::string ft1 = frame_name(dir1, 't', iframe), ft2 = frame_name(dir2, 't', iframe);
    const std::string fq1 = frame_name(dir1, 'q', iframe), fq2 = frame_name(dir2, 'q', iframe);
    if (!t1.load(ft1)) return set_error(HIPCLAW_ERR_ARG, ("could not open file " + ft1).c_str());
    if (!t2.load(ft2)) return set_error(HIPCLAW_ERR_ARG, ("could not open file " + ft2).c_str());
    double tm1 = 0, tm2 = 0;
    int meqn1 = 0, meqn2 = 0, ng1 = 0, ng2 = 0;
    if (!first_real(t1, tm1) || !first_int(t1, meqn1) || !first_int(t1, ng1) || !first_real(t2, tm2) ||
        !first_int(t2, meqn2) || !first_int(t2, ng2))
        return set_error(HIPCLAW_ERR_ARG, "unreadable fort.t file");
    if (meqn1 != meqn2) return set_error(HIPCLAW_ERR_ARG, "meqn1 .ne. meqn2");
    if (std::fabs(tm1 - tm2) > 1e-5) return set_error(HIPCLAW_ERR_ARG, "t1 and t2 are not close enough to compare files");
    if (ng1 != ng2) return set_error(HIPCLAW_ERR_ARG, "ngrids1 .ne. ngrids2");
    const int meqn = meqn1, ngrids = ng1;
    if (meqn < 1 || meqn > 10) return set_error(HIPCLAW_ERR_ARG, "meqn must be 1..10 (format 10E26.16)");
    if (!q1.load(fq1)) return set_error(HIPCLAW_ERR_ARG, ("could not open file " + fq1).c_str());
    if (!q2.load(fq2)) return set_error(HIPCLAW_ERR_ARG, ("could not open file " + fq2).c_str());

    // fort.tXXXX of the difference (format e18.8 / i5 records, two empty records)
    std::string tt;
    put_e(tt, tm1, 18, 8);
    tt += "    time\n";
    put_i5(tt, meqn, "                 meqn");
    put_i5(tt, ngrids, "                 ngrids");
    put_i5(tt, 0, "                 maux");
    tt += "\n\n";
    std::vector<double> gmax(meqn, 0.0), gridmax((size_t)std::max(ngrids, 0) * meqn, 0.0);
    std::vector<int> gwhere(meqn, 0);
    std::string qt;
    for (int n = 0; n < ngrids; n++) {
        int hd1[6], hd2[6];
        double geo1[4], geo2[4];
        for (int k = 0; k < 6; k++)
            if (!first_int(q1, hd1[k]) || !first_int(q2, hd2[k]))
                return set_error(HIPCLAW_ERR_ARG, "unreadable grid header");
        for (int k = 0; k < 4; k++)
            if (!first_real(q1, geo1[k]) || !first_real(q2, geo2[k]))
                return set_error(HIPCLAW_ERR_ARG, "unreadable grid header");
        if (hd1[0] != hd2[0]) return set_error(HIPCLAW_ERR_ARG, "ngrid1 .ne. ngrid2");
        if (hd1[4] != hd2[4]) return set_error(HIPCLAW_ERR_ARG, "mx1 .ne. mx2");
        if (geo1[0] != geo2[0]) return set_error(HIPCLAW_ERR_ARG, "xlow1 .ne. xlow2");
        if (geo1[1] != geo2[1]) return set_error(HIPCLAW_ERR_ARG, "ylow1 .ne. ylow2");
        const int ngrid = hd1[0], level = hd1[1], blockno = hd1[2], rank = std::max(hd1[3], hd2[3]);
        const int mx = hd1[4], my = hd1[5];
        if (mx < 1 || my < 1 || ngrid < 0 || ngrid >= ngrids) return set_error(HIPCLAW_ERR_ARG, "bad grid header");
        std::vector<double> qc((size_t)meqn * mx * my);
        for (int m = 0; m < meqn; m++) gridmax[(size_t)ngrid * meqn + m] = 0.0;
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                double a[10], b[10];
                for (int pass = 0; pass < 2; pass++) {
                    Lines& L = pass ? q2 : q1;
                    double* dst = pass ? b : a;
                    int got = 0;
                    while (got < meqn) {
                        const char* l = L.next();
                        if (!l) return set_error(HIPCLAW_ERR_ARG, "fort.q file ends early");
                        while (got < meqn && read_real(l, dst[got])) got++;
                    }
                }
                for (int m = 0; m < meqn; m++) {
                    const double d = a[m] - b[m];
                    qc[((size_t)m * my + j) * mx + i] = d;
                    if (std::fabs(d) > gmax[m]) {
                        gmax[m] = std::fabs(d);
                        gwhere[m] = ngrid;
                    }
                    double& gm = gridmax[(size_t)ngrid * meqn + m];
                    gm = std::max(gm, std::fabs(d));
                }
            }
        patch_text(qt, mx, my, meqn, 0, geo1[0], geo1[1], geo1[2], geo1[3], qc.data(), ngrid, level, blockno, rank);
    }
    // compare_summary.XXXX
    std::string sm;
    for (int m = 0; m < meqn; m++) {
        put_e(sm, gmax[m], 30, 16);
        char buf[64];
        std::snprintf(buf, sizeof buf, "     occurs in field%5d on grid %5d\n", m + 1, gwhere[m]);
        sm += buf;
    }
    sm += "  \n";
    for (int n = 0; n < ngrids; n++) {
        char buf[16];
        std::snprintf(buf, sizeof buf, "%5d", n);
        sm += buf;
        for (int m = 0; m < meqn; m++) put_e(sm, gridmax[(size_t)n * meqn + m], 16, 4);
        sm += "\n";
    }
    char sname[32];
    std::snprintf(sname, sizeof sname, "compare_summary.%04d", iframe);
    const std::string d3 = dir3 && *dir3 ? std::string(dir3) : std::string(".");
    int rc = write_file(frame_name(dir3, 't', iframe), tt, "w");
    if (!rc) rc = write_file(frame_name(dir3, 'q', iframe), qt, "w");
    if (!rc) rc = write_file(d3 + "/" + sname, sm, "w");
    if (rc) return rc;
    for (int m = 0; m < meqn && m < nmax; m++) {
        if (global_max) global_max[m] = gmax[m];
        if (grid_max) grid_max[m] = gwhere[m];
    }
    return HIPCLAW_OK;
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
