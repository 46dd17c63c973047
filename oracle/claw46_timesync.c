/*
 * claw46_timesync.c -- CPU FP64 restatement of ForestClaw's conservative time
 * synchronisation corrections for Cartesian patches of one orientation (a
 * brick of unit blocks).  TEST INFRASTRUCTURE ONLY: the GPU product path never
 * links or calls this code.
 *
 * Register layout per patch (orc_reg_offset): for face f = 0 (left), 1
 * (right), 2 (bottom), 3 (top) with length L = my (f < 2) or mx, four blocks
 * of meqn*L doubles, index m*L + j:
 *   block 0  P : fp (x faces) / gp (y faces)   fclaw2d_clawpatch_registers fp[], gp[]
 *   block 1  M : fm / gm                        fm[], gm[]
 *   block 2  E0: edge flux at the interior cell  edge_fluxes[f](:, :, 0)
 *   block 3  E1: edge flux at the ghost cell     edge_fluxes[f](:, :, 1)
 * (src/patches/clawpatch/fclaw2d_clawpatch_conservation.c:80-110).
 *
 * q is one tile, q(1-mbc:mx+mbc, 1-mbc:my+mbc, meqn) in C order
 * ((m*NY + j)*NX + i).
 */
#include <stddef.h>

#include "claw46_oracle.h"

#define REG(r, f, b, m, k) (r)[orc_reg_offset(mx, my, meqn, (f), (b)) + (long)(m) * ((f) < 2 ? my : mx) + (k)]

/* Fine-to-coarse correction from one fine neighbour igrid (0 lower, 1 upper
 * half of the coarse face) across coarse face iface
 * (clawpack46_fort_time_sync_f2c, clawpack46_time_sync.f:254-498, with the
 * same-orientation transform fine = 2*coarse - 1 - len*igrid and the
 * fc46_skip_this_grid validity test, :706-740).  Every coarse edge cell of the
 * face is visited; cells the fine patch does not cover receive + 0/area. */
void orc_time_sync_f2c(int mx, int my, int mbc, int meqn, int iface, int igrid, double area,
                       double* qc, const double* rc, const double* rf)
{
    const int NX = mx + 2 * mbc, NY = my + 2 * mbc;
    const int len = iface < 2 ? my : mx;
    for (int m = 0; m < meqn; m++)
        for (int c = 1; c <= len; c++) {
            const int f0 = 2 * c - 1 - len * igrid;  /* first fine index (1-based) */
            const int valid = f0 >= 1 && f0 + 1 <= len;
            double delta = 0.0;
            if (valid) {
                const int a = f0 - 1, b = f0;    /* 0-based fine pair */
                const int k = c - 1;
                if (iface == 0) {
                    double fp = REG(rf, 1, 0, m, a) + REG(rf, 1, 0, m, b);
                    double ef = REG(rf, 1, 3, m, a) + REG(rf, 1, 3, m, b);
                    delta = (ef - REG(rc, 0, 2, m, k)) - (fp - REG(rc, 0, 0, m, k));
                } else if (iface == 1) {
                    double fm = REG(rf, 0, 1, m, a) + REG(rf, 0, 1, m, b);
                    double ef = REG(rf, 0, 3, m, a) + REG(rf, 0, 3, m, b);
                    delta = (REG(rc, 1, 2, m, k) - ef) - (fm - REG(rc, 1, 1, m, k));
                } else if (iface == 2) {
                    double gp = REG(rf, 3, 0, m, a) + REG(rf, 3, 0, m, b);
                    double ef = REG(rf, 3, 3, m, a) + REG(rf, 3, 3, m, b);
                    delta = (ef - REG(rc, 2, 2, m, k)) - (gp - REG(rc, 2, 0, m, k));
                } else {
                    double gm = REG(rf, 2, 1, m, a) + REG(rf, 2, 1, m, b);
                    double ef = REG(rf, 2, 3, m, a) + REG(rf, 2, 3, m, b);
                    delta = (REG(rc, 3, 2, m, k) - ef) - (gm - REG(rc, 3, 1, m, k));
                }
            }
            int i, j;
            if (iface < 2) { i = iface == 0 ? 1 : mx; j = c; }
            else { i = c; j = iface == 2 ? 1 : my; }
            double* qq = &qc[((size_t)m * NY + (j - 1 + mbc)) * NX + (i - 1 + mbc)];
            *qq = *qq + delta / area;
        }
}

/* Same-size correction of this patch across face iface
 * (clawpack46_fort_time_sync_samesize, clawpack46_time_sync.f:528-704):
 * half of the mismatch is applied from each side. */
void orc_time_sync_samesize(int mx, int my, int mbc, int meqn, int iface, double area,
                            double* q, const double* rt, const double* rn)
{
    const int NX = mx + 2 * mbc, NY = my + 2 * mbc;
    const int len = iface < 2 ? my : mx;
    for (int m = 0; m < meqn; m++)
        for (int k = 0; k < len; k++) {
            double delta;
            if (iface == 0) {
                double fp = REG(rn, 1, 0, m, k), ef = REG(rn, 1, 3, m, k);
                delta = (ef - REG(rt, 0, 2, m, k)) - (fp - REG(rt, 0, 0, m, k));
            } else if (iface == 1) {
                double fm = REG(rn, 0, 1, m, k), ef = REG(rn, 0, 3, m, k);
                delta = (REG(rt, 1, 2, m, k) - ef) - (fm - REG(rt, 1, 1, m, k));
            } else if (iface == 2) {
                double gp = REG(rn, 3, 0, m, k), ef = REG(rn, 3, 3, m, k);
                delta = (ef - REG(rt, 2, 2, m, k)) - (gp - REG(rt, 2, 0, m, k));
            } else {
                double gm = REG(rn, 2, 1, m, k), ef = REG(rn, 2, 3, m, k);
                delta = (REG(rt, 3, 2, m, k) - ef) - (gm - REG(rt, 3, 1, m, k));
            }
            int i, j;
            if (iface < 2) { i = iface == 0 ? 1 : mx; j = k + 1; }
            else { i = k + 1; j = iface == 2 ? 1 : my; }
            double* qq = &q[((size_t)m * NY + (j - 1 + mbc)) * NX + (i - 1 + mbc)];
            *qq = *qq + 0.5 * delta / area;
        }
}

/* Zero the four register blocks of one face
 * (fclaw2d_clawpatch_time_sync_reset, fclaw2d_clawpatch_conservation.c:207-280). */
void orc_time_sync_reset_face(int mx, int my, int meqn, int iface, double* reg)
{
    const long n = 4L * meqn * (iface < 2 ? my : mx);
    double* r = reg + orc_reg_offset(mx, my, meqn, iface, 0);
    for (long t = 0; t < n; t++) r[t] = 0.0;
}
