/*
 * claw46_ghost.c -- CPU FP64 restatement of the clawpatch 4.6 ghost-cell and
 * regrid kernels, one patch pair per call.  TEST INFRASTRUCTURE ONLY.
 *
 * Follows the reference loops, validity checks and summation orders:
 *   copy face / corner        src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_copy.f:17-133
 *   average face / corner     .../fclaw2d_clawpatch46_average.f:16-272
 *   average2coarse            .../fclaw2d_clawpatch46_average.f:282-364
 *   interpolate face / corner .../fclaw2d_clawpatch46_interpolate.f:30-314
 *   interpolate2fine          .../fclaw2d_clawpatch46_interpolate.f:324-400
 *   slopes (mth = 5)          src/patches/clawpatch/fclaw2d_clawpatch_utils.f:11-66
 *   is_valid_average/interp   src/patches/clawpatch/fclaw2d_clawpatch_utils.f (is_valid_* functions)
 *   time interpolation        .../fclaw2d_clawpatch46_timeinterp.f:11-78
 *
 * The index transforms come from the caller as the reference's same-block
 * closed forms (src/forestclaw2d.c:655-663 and 798-841, based = 1):
 *   same size:  i2 = i1 + sx,                       j2 = j1 + sy
 *   half size:  i2(m) = 1 + ox + 2*(i1 - 1) + kx,   j2(m) = 1 + oy + 2*(j1 - 1) + ky,
 *               m = 2*ky + kx,  with ox = (xlower_c - xlower_f)*2^lf*mx.
 * Indices below are the reference's 1-based ones; Q() maps them to the tile.
 */
#include <math.h>
#include <string.h>

#include "claw46_oracle.h"

#define Q(a, i, j, m) a[((size_t)(m) * (my + 2 * mbc) + (size_t)((j) + mbc - 1)) * (mx + 2 * mbc) + (size_t)((i) + mbc - 1)]

static int valid_average(int i, int j, int mx, int my) { return 1 <= i && i <= mx && 1 <= j && j <= my; }
static int valid_interp(int i, int j, int mx, int my, int mbc)
{
    return 1 - mbc <= i && i <= mx + mbc && 1 - mbc <= j && j <= my + mbc;
}

static double fsign1(double x) { return copysign(1.0, x); }
static double dmin_(double a, double b) { return a < b ? a : b; }
static double dmax_(double a, double b) { return a > b ? a : b; }

double orc_compute_slopes(double sl, double sr)
{
    double sc = (sl + sr) / 2.0;
    return dmin_(dmin_(2 * fabs(sl), 2 * fabs(sr)), fabs(sc)) * dmax_(0.0, fsign1(sl * sr)) * fsign1(sc);
}

static void half(int i1, int j1, int ox, int oy, int i2[4], int j2[4])
{
    for (int ky = 0; ky < 2; ky++)
        for (int kx = 0; kx < 2; kx++) {
            i2[2 * ky + kx] = 1 + ox + 2 * (i1 - 1) + kx;
            j2[2 * ky + kx] = 1 + oy + 2 * (j1 - 1) + ky;
        }
}

void orc_copy_face(int mx, int my, int mbc, int meqn, double* qthis, const double* qnbr,
                   int iface, int sx, int sy)
{
    for (int m = 0; m < meqn; m++) {
        if (iface == 0) {
            for (int j = 1; j <= my; j++)
                for (int i = 1 - mbc; i <= 0; i++) Q(qthis, i, j, m) = Q(qnbr, i + sx, j + sy, m);
        } else if (iface == 1) {
            for (int j = 1; j <= my; j++)
                for (int i = mx + 1; i <= mx + mbc; i++) Q(qthis, i, j, m) = Q(qnbr, i + sx, j + sy, m);
        } else if (iface == 2) {
            for (int j = 1 - mbc; j <= 0; j++)
                for (int i = 1; i <= mx; i++) Q(qthis, i, j, m) = Q(qnbr, i + sx, j + sy, m);
        } else {
            for (int j = my + 1; j <= my + mbc; j++)
                for (int i = 1; i <= mx; i++) Q(qthis, i, j, m) = Q(qnbr, i + sx, j + sy, m);
        }
    }
}

static void corner_cell(int icorner, int ibc, int jbc, int mx, int my, int* i1, int* j1)
{
    switch (icorner) {
    case 0: *i1 = 1 - ibc; *j1 = 1 - jbc; break;
    case 1: *i1 = mx + ibc; *j1 = 1 - jbc; break;
    case 2: *i1 = 1 - ibc; *j1 = my + jbc; break;
    default: *i1 = mx + ibc; *j1 = my + jbc; break;
    }
}

void orc_copy_corner(int mx, int my, int mbc, int meqn, double* qthis, const double* qnbr,
                     int icorner, int sx, int sy)
{
    for (int m = 0; m < meqn; m++)
        for (int ibc = 1; ibc <= mbc; ibc++)
            for (int jbc = 1; jbc <= mbc; jbc++) {
                int i1, j1;
                corner_cell(icorner, ibc, jbc, mx, my, &i1, &j1);
                Q(qthis, i1, j1, m) = Q(qnbr, i1 + sx, j1 + sy, m);
            }
}

void orc_average_face(int mx, int my, int mbc, int meqn, double* qcoarse, const double* qfine,
                      int iface_coarse, int ox, int oy)
{
    int i2[4], j2[4];
    for (int m = 0; m < meqn; m++) {
        if (iface_coarse < 2) {
            for (int jc = 1; jc <= my; jc++)
                for (int ibc = 1; ibc <= mbc; ibc++) {
                    int ic = iface_coarse == 0 ? 1 - ibc : mx + ibc;
                    half(ic, jc, ox, oy, i2, j2);
                    int skip = 0;
                    for (int k = 0; k < 4; k++)
                        if (!valid_average(i2[k], j2[k], mx, my)) skip = 1;
                    if (skip) continue;
                    double sum = 0;
                    for (int k = 0; k < 4; k++) sum = sum + Q(qfine, i2[k], j2[k], m);
                    Q(qcoarse, ic, jc, m) = sum / (double)4;
                }
        } else {
            for (int jbc = 1; jbc <= mbc; jbc++)
                for (int ic = 1; ic <= mx; ic++) {
                    int jc = iface_coarse == 2 ? 1 - jbc : my + jbc;
                    half(ic, jc, ox, oy, i2, j2);
                    int skip = 0;
                    for (int k = 0; k < 4; k++)
                        if (!valid_average(i2[k], j2[k], mx, my)) skip = 1;
                    if (skip) continue;
                    double sum = 0;
                    for (int k = 0; k < 4; k++) sum = sum + Q(qfine, i2[k], j2[k], m);
                    Q(qcoarse, ic, jc, m) = sum / (double)4;
                }
        }
    }
}

void orc_average_corner(int mx, int my, int mbc, int meqn, double* qcoarse, const double* qfine,
                        int icorner_coarse, int ox, int oy)
{
    int i2[4], j2[4];
    for (int ibc = 1; ibc <= mbc; ibc++)
        for (int jbc = 1; jbc <= mbc; jbc++) {
            int i1, j1;
            corner_cell(icorner_coarse, ibc, jbc, mx, my, &i1, &j1);
            half(i1, j1, ox, oy, i2, j2);
            for (int m = 0; m < meqn; m++) {
                double sum = 0;
                for (int k = 0; k < 4; k++) sum = sum + Q(qfine, i2[k], j2[k], m);
                Q(qcoarse, i1, j1, m) = sum / (double)4;
            }
        }
}

static void interp_cell(int mx, int my, int mbc, int m, const double* qcoarse, double* qfine, int ic,
                        int jc, const int i2[4], const int j2[4])
{
    double qc = Q(qcoarse, ic, jc, m);
    double sl = qc - Q(qcoarse, ic - 1, jc, m);
    double sr = Q(qcoarse, ic + 1, jc, m) - qc;
    double gradx = orc_compute_slopes(sl, sr);
    sl = qc - Q(qcoarse, ic, jc - 1, m);
    sr = Q(qcoarse, ic, jc + 1, m) - qc;
    double grady = orc_compute_slopes(sl, sr);
    int k = 0;
    for (int jj = 0; jj <= 1; jj++)
        for (int ii = 0; ii <= 1; ii++) {
            /* df = (a*dc)/refratio with a = 2*identity in the same block */
            int iff = i2[0] + ii, jff = j2[0] + jj;
            double shiftx = (ii - 0.5) / 2.0, shifty = (jj - 0.5) / 2.0;
            Q(qfine, iff, jff, m) = qc + gradx * shiftx + grady * shifty;
            k++;
        }
}

void orc_interpolate_face(int mx, int my, int mbc, int meqn, const double* qcoarse, double* qfine,
                          int iface_coarse, int ox, int oy)
{
    int i2[4], j2[4];
    for (int m = 0; m < meqn; m++) {
        if (iface_coarse < 2) {
            for (int ibc = 1; ibc <= mbc / 2; ibc++) {
                int ic = iface_coarse == 0 ? ibc : mx - ibc + 1;
                for (int jc = 1; jc <= mx; jc++) {
                    half(ic, jc, ox, oy, i2, j2);
                    int skip = 0;
                    for (int k = 0; k < 4; k++)
                        if (!valid_interp(i2[k], j2[k], mx, my, mbc)) skip = 1;
                    if (!skip) interp_cell(mx, my, mbc, m, qcoarse, qfine, ic, jc, i2, j2);
                }
            }
        } else {
            for (int jbc = 1; jbc <= mbc / 2; jbc++) {
                int jc = iface_coarse == 2 ? jbc : my - jbc + 1;
                for (int ic = 1; ic <= mx; ic++) {
                    half(ic, jc, ox, oy, i2, j2);
                    int skip = 0;
                    for (int k = 0; k < 4; k++)
                        if (!valid_interp(i2[k], j2[k], mx, my, mbc)) skip = 1;
                    if (!skip) interp_cell(mx, my, mbc, m, qcoarse, qfine, ic, jc, i2, j2);
                }
            }
        }
    }
}

void orc_interpolate_corner(int mx, int my, int mbc, int meqn, const double* qcoarse, double* qfine,
                            int icorner_coarse, int ox, int oy)
{
    int i2[4], j2[4];
    for (int ibc = 1; ibc <= mbc / 2; ibc++)
        for (int jbc = 1; jbc <= mbc / 2; jbc++) {
            int ic, jc;
            switch (icorner_coarse) {
            case 0: ic = ibc; jc = jbc; break;
            case 1: ic = mx - ibc + 1; jc = jbc; break;
            case 2: ic = ibc; jc = my - jbc + 1; break;
            default: ic = mx - ibc + 1; jc = my - jbc + 1; break;
            }
            half(ic, jc, ox, oy, i2, j2);
            for (int m = 0; m < meqn; m++) interp_cell(mx, my, mbc, m, qcoarse, qfine, ic, jc, i2, j2);
        }
}

void orc_interpolate2fine(int mx, int my, int mbc, int meqn, const double* qcoarse, double* qfine,
                          int igrid)
{
    const int refratio = 2;
    int ig = igrid % refratio, jg = (igrid - ig) / refratio;
    int i1 = 1 - ig, i2 = mx / 2 + (1 - ig), ic_add = ig * mx / 2;
    int j1 = 1 - jg, j2 = my / 2 + (1 - jg), jc_add = jg * my / 2;
    for (int m = 0; m < meqn; m++)
        for (int j = j1; j <= j2; j++)
            for (int i = i1; i <= i2; i++) {
                int ic = i + ic_add, jc = j + jc_add;
                double qc = Q(qcoarse, ic, jc, m);
                double sl = qc - Q(qcoarse, ic - 1, jc, m), sr = Q(qcoarse, ic + 1, jc, m) - qc;
                double gradx = orc_compute_slopes(sl, sr);
                sl = qc - Q(qcoarse, ic, jc - 1, m);
                sr = Q(qcoarse, ic, jc + 1, m) - qc;
                double grady = orc_compute_slopes(sl, sr);
                for (int ii = 1; ii <= refratio; ii++)
                    for (int jj = 1; jj <= refratio; jj++) {
                        double shiftx = (ii - refratio / 2.0 - 0.5) / refratio;
                        double shifty = (jj - refratio / 2.0 - 0.5) / refratio;
                        Q(qfine, (i - 1) * refratio + ii, (j - 1) * refratio + jj, m) =
                            qc + shiftx * gradx + shifty * grady;
                    }
            }
}

void orc_average2coarse(int mx, int my, int mbc, int meqn, double* qcoarse, const double* qfine,
                        int igrid)
{
    const int refratio = 2;
    int ig = igrid % refratio, jg = (igrid - ig) / refratio;
    int ic_add = ig * mx / 2, jc_add = jg * mx / 2;
    for (int m = 0; m < meqn; m++)
        for (int j = 1; j <= my / 2; j++)
            for (int i = 1; i <= mx / 2; i++) {
                double sum = 0;
                for (int jj = 1; jj <= refratio; jj++)
                    for (int ii = 1; ii <= refratio; ii++)
                        sum = sum + Q(qfine, (i - 1) * refratio + ii, (j - 1) * refratio + jj, m);
                Q(qcoarse, i + ic_add, j + jc_add, m) = sum / (refratio * refratio);
            }
}

int orc_timeinterp(int mx, int my, int mbc, int meqn, const double* qcurr, const double* qlast,
                   double* qinterp, double alpha)
{
    int mint = mbc;
    for (int m = 0; m < meqn; m++) {
        for (int j = 1; j <= my - mint; j++)
            for (int i = 1; i <= mint; i++)
                Q(qinterp, i, j, m) = Q(qlast, i, j, m) + alpha * (Q(qcurr, i, j, m) - Q(qlast, i, j, m));
        for (int j = 1; j <= mint; j++)
            for (int i = mint + 1; i <= mx; i++)
                Q(qinterp, i, j, m) = Q(qlast, i, j, m) + alpha * (Q(qcurr, i, j, m) - Q(qlast, i, j, m));
        for (int j = mint + 1; j <= my; j++)
            for (int i = mx - mint + 1; i <= mx; i++)
                Q(qinterp, i, j, m) = Q(qlast, i, j, m) + alpha * (Q(qcurr, i, j, m) - Q(qlast, i, j, m));
        for (int j = my - mint + 1; j <= my; j++)
            for (int i = 1; i <= mx - mint; i++)
                Q(qinterp, i, j, m) = Q(qlast, i, j, m) + alpha * (Q(qcurr, i, j, m) - Q(qlast, i, j, m));
    }
    return 0;
}

/* Gradient tag of one patch (refine test): 1 if any interior cell of field
 * ivar has |grad q| (central differences) > threshold. */
int orc_tag_gradient(int mx, int my, int mbc, const double* q, int ivar, double dx, double dy,
                     double threshold)
{
    const double* qv = q + (size_t)ivar * (mx + 2 * mbc) * (my + 2 * mbc);
    const int m = 0;
    for (int j = 1; j <= my; j++)
        for (int i = 1; i <= mx; i++) {
            double dqx = (Q(qv, i + 1, j, m) - Q(qv, i - 1, j, m)) / (2 * dx);
            double dqy = (Q(qv, i, j + 1, m) - Q(qv, i, j - 1, m)) / (2 * dy);
            double ds = sqrt(dqx * dqx + dqy * dqy + 0.0 * 0.0);
            if (ds > threshold) return 1;
        }
    return 0;
}
#undef Q
