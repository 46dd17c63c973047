/*
 * claw46_oracle.c -- CPU FP64 restatement of the Clawpack-4.6 patch step.
 *
 * TEST INFRASTRUCTURE ONLY (checker and CPU baseline); the GPU product path
 * never links or calls it.
 *
 * Follows, operation by operation and in the same evaluation order so that
 * results are bitwise equal to the reference Fortran built without FMA
 * contraction:
 *   step2 driver      src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_step2.f:52-236
 *   flux update       .../clawpack46_step2_wrap.f:104-125
 *   1-D flux kernel   .../clawpack46_flux2.f:90-224
 *   limiter           .../clawpack46_inlinelimiter.f:39-115
 *   advection RP      applications/clawpack/advection/2d/all/clawpack46_rpn2adv.f:20-58,
 *                     clawpack46_rpt2adv.f:16-25
 *   SWE Roe RP        applications/clawpack/shallow/2d/rp/clawpack46_rpn2.f:70-245,
 *                     clawpack46_rpt2.f:27-72
 *   Euler5 Roe RP     applications/clawpack/euler/2d/rp/clawpack46_rpn2_euler5.f:51-278,
 *                     clawpack46_rpt2_euler5.f:40-118
 *   swirl b4step2     applications/clawpack/advection/2d/swirl/user_4.6/b4step2.f:19-39,
 *                     swirl/psi.f:1-12
 *   physical BC       src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_bc2_default.f:14-211
 *
 * Layout: tile (i, j, m) -> q[(m*NY + j)*NX + i], NX = mx+2mbc, NY = my+2mbc.
 * A 1-D slice has N = n + 2mbc cells; face k (1 <= k <= N-1) separates slice
 * cells k-1 and k (Fortran interface i = k - mbc + 1).
 */
#include "claw46_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXW 5
#define MAXE 5
#define ROE_N 7

typedef struct {
    int N;         /* slice length incl. ghosts                 */
    int n;         /* interior length (mx or my)                */
    int mbc, meqn, maux, mwaves;
    /* work arrays, each indexed [...][N]                        */
    double *q1d, *aux1, *aux2, *aux3;
    double *wave, *s, *amdq, *apdq, *cqxx, *bmasdq, *bpasdq;
    double *faddm, *faddp, *gaddm, *gaddp;
    double *roe;
} slice_t;

#define Q1(m, k) (sl->q1d[(m) * sl->N + (k)])
#define A1(a, k) (sl->aux1[(a) * sl->N + (k)])
#define A2(a, k) (sl->aux2[(a) * sl->N + (k)])
#define A3(a, k) (sl->aux3[(a) * sl->N + (k)])
#define WV(w, m, k) (sl->wave[((w) * sl->meqn + (m)) * sl->N + (k)])
#define SP(w, k) (sl->s[(w) * sl->N + (k)])
#define AM(m, k) (sl->amdq[(m) * sl->N + (k)])
#define AP(m, k) (sl->apdq[(m) * sl->N + (k)])
#define CQ(m, k) (sl->cqxx[(m) * sl->N + (k)])
#define BM(m, k) (sl->bmasdq[(m) * sl->N + (k)])
#define BP(m, k) (sl->bpasdq[(m) * sl->N + (k)])
#define FM(m, k) (sl->faddm[(m) * sl->N + (k)])
#define FP(m, k) (sl->faddp[(m) * sl->N + (k)])
#define GM(side, m, k) (sl->gaddm[((side) * sl->meqn + (m)) * sl->N + (k)])
#define GP(side, m, k) (sl->gaddp[((side) * sl->meqn + (m)) * sl->N + (k)])
#define ROE(r, k) (sl->roe[(r) * sl->N + (k)])

static double dmin(double a, double b) { return a < b ? a : b; }
static double dmax(double a, double b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------ */
/* Normal Riemann solvers.  Left state of face k is slice cell k-1 ("qr(i-1)"),
 * right state is cell k ("ql(i)").                                          */
/* ------------------------------------------------------------------------ */

static void rpn2_adv(const orc_cfg* c, int ixy, slice_t* sl)
{
    (void)c;
    const int iface = ixy - 1; /* auxl(i, ixy) */
    for (int k = 1; k < sl->N; k++) {
        for (int m = 0; m < sl->meqn; m++) {
            double delta = Q1(m, k) - Q1(m, k - 1);
            for (int w = 0; w < sl->mwaves; w++) WV(w, m, k) = delta;
        }
        for (int w = 0; w < sl->mwaves; w++) SP(w, k) = A2(iface, k);
        for (int m = 0; m < sl->meqn; m++) {
            double am = 0.0, ap = 0.0;
            for (int w = 0; w < sl->mwaves; w++) {
                am = am + dmin(SP(w, k), 0.0) * WV(w, m, k);
                ap = ap + dmax(SP(w, k), 0.0) * WV(w, m, k);
            }
            AM(m, k) = am;
            AP(m, k) = ap;
        }
    }
}

/* Roe cache slots for SWE: 0 u, 1 v, 2 a, 3 h */
static void rpn2_swe(const orc_cfg* c, int ixy, slice_t* sl)
{
    const double g = c->grav;
    const int mu = (ixy == 1) ? 1 : 2;
    const int mv = (ixy == 1) ? 2 : 1;
    for (int k = 1; k < sl->N; k++) {
        double hl = Q1(0, k - 1), hr = Q1(0, k);
        double h = (hl + hr) / 2.0;
        double sl_ = sqrt(hl), sr_ = sqrt(hr);
        double hsq2 = sl_ + sr_;
        double u = (Q1(mu, k - 1) / sl_ + Q1(mu, k) / sr_) / hsq2;
        double v = (Q1(mv, k - 1) / sl_ + Q1(mv, k) / sr_) / hsq2;
        double a = sqrt(g * h);
        ROE(0, k) = u; ROE(1, k) = v; ROE(2, k) = a; ROE(3, k) = h;

        double d0 = Q1(0, k) - Q1(0, k - 1);
        double d1 = Q1(mu, k) - Q1(mu, k - 1);
        double d2 = Q1(mv, k) - Q1(mv, k - 1);
        double a1, a3;
        double a2 = -v * d0 + d2;
        if (c->claw5) {
            a1 = ((u + a) * d0 - d1) * (0.5 / a);
            a3 = (-(u - a) * d0 + d1) * (0.5 / a);
        } else {
            a1 = ((u + a) * d0 - d1) / (2.0 * a);
            a3 = (-(u - a) * d0 + d1) / (2.0 * a);
        }

        WV(0, 0, k) = a1; WV(0, mu, k) = a1 * (u - a); WV(0, mv, k) = a1 * v;
        SP(0, k) = u - a;
        WV(1, 0, k) = 0.0; WV(1, mu, k) = 0.0; WV(1, mv, k) = a2;
        SP(1, k) = u;
        WV(2, 0, k) = a3; WV(2, mu, k) = a3 * (u + a); WV(2, mv, k) = a3 * v;
        SP(2, k) = u + a;
    }

    if (!c->efix) {
        for (int k = 1; k < sl->N; k++)
            for (int m = 0; m < sl->meqn; m++) {
                double am = 0.0, ap = 0.0;
                for (int w = 0; w < sl->mwaves; w++) {
                    am = am + dmin(SP(w, k), 0.0) * WV(w, m, k);
                    ap = ap + dmax(SP(w, k), 0.0) * WV(w, m, k);
                }
                AM(m, k) = am;
                AP(m, k) = ap;
            }
        return;
    }

    for (int k = 1; k < sl->N; k++) {
        /* 1-wave */
        double him1 = Q1(0, k - 1);
        double s0 = Q1(mu, k - 1) / him1 - sqrt(g * him1);
        double sfract;
        if (s0 > 0.0 && SP(0, k) > 0.0) {
            for (int m = 0; m < 3; m++) AM(m, k) = 0.0;
            goto done;
        }
        {
            double h1 = Q1(0, k - 1) + WV(0, 0, k);
            double hu1 = Q1(mu, k - 1) + WV(0, mu, k);
            double s1 = hu1 / h1 - sqrt(g * h1);
            if (s0 < 0.0 && s1 > 0.0)
                sfract = s0 * ((s1 - SP(0, k)) / (s1 - s0));
            else if (SP(0, k) < 0.0)
                sfract = SP(0, k);
            else
                sfract = 0.0;
        }
        for (int m = 0; m < 3; m++) AM(m, k) = sfract * WV(0, m, k);
        /* 2-wave */
        if (SP(1, k) > 0.0) goto done;
        for (int m = 0; m < 3; m++) AM(m, k) = AM(m, k) + SP(1, k) * WV(1, m, k);
        /* 3-wave */
        {
            double hi = Q1(0, k);
            double s03 = Q1(mu, k) / hi + sqrt(g * hi);
            double h3 = Q1(0, k) - WV(2, 0, k);
            double hu3 = Q1(mu, k) - WV(2, mu, k);
            double s3 = hu3 / h3 + sqrt(g * h3);
            if (s3 < 0.0 && s03 > 0.0)
                sfract = s3 * ((s03 - SP(2, k)) / (s03 - s3));
            else if (SP(2, k) < 0.0)
                sfract = SP(2, k);
            else
                goto done;
        }
        for (int m = 0; m < 3; m++) AM(m, k) = AM(m, k) + sfract * WV(2, m, k);
    done:;
    }
    for (int m = 0; m < 3; m++)
        for (int k = 1; k < sl->N; k++) {
            double df = 0.0;
            for (int w = 0; w < sl->mwaves; w++) df = df + SP(w, k) * WV(w, m, k);
            AP(m, k) = df - AM(m, k);
        }
}

/* Roe cache slots for Euler: 0 u, 1 v, 2 enth, 3 a, 4 g1a2, 5 euv, 6 u2v2 */
static void euler_roe(const orc_cfg* c, int mu, int mv, const double* ql,
                      const double* qr, double* r)
{
    /* ql = right cell (i), qr = left cell (i-1); components 0..3 */
    const double g1 = c->gamma - 1.0;
    double rl = sqrt(qr[0]);
    double rr = sqrt(ql[0]);
    double pl = g1 * (qr[3] - 0.5 * (qr[1] * qr[1] + qr[2] * qr[2]) / qr[0]);
    double pr = g1 * (ql[3] - 0.5 * (ql[1] * ql[1] + ql[2] * ql[2]) / ql[0]);
    double rsq2 = rl + rr;
    double u = (qr[mu] / rl + ql[mu] / rr) / rsq2;
    double v = (qr[mv] / rl + ql[mv] / rr) / rsq2;
    double enth = (((qr[3] + pl) / rl + (ql[3] + pr) / rr)) / rsq2;
    double u2v2 = u * u + v * v;
    double a2 = g1 * (enth - .5 * u2v2);
    double a = sqrt(a2);
    r[0] = u; r[1] = v; r[2] = enth; r[3] = a;
    r[4] = g1 / a2; r[5] = enth - u2v2; r[6] = u2v2;
}

static void rpn2_euler5(const orc_cfg* c, int ixy, slice_t* sl)
{
    const double gam = c->gamma, g1 = c->gamma - 1.0;
    const int mu = (ixy == 1) ? 1 : 2;
    const int mv = (ixy == 1) ? 2 : 1;
    for (int k = 1; k < sl->N; k++) {
        double ql[5], qr[5], r[ROE_N];
        for (int m = 0; m < 5; m++) { ql[m] = Q1(m, k); qr[m] = Q1(m, k - 1); }
        euler_roe(c, mu, mv, ql, qr, r);
        for (int t = 0; t < ROE_N; t++) ROE(t, k) = r[t];
        double u = r[0], v = r[1], enth = r[2], a = r[3], g1a2 = r[4], euv = r[5],
               u2v2 = r[6];
        double d0 = ql[0] - qr[0], d1 = ql[mu] - qr[mu], d2 = ql[mv] - qr[mv],
               d3 = ql[3] - qr[3];
        double a3 = g1a2 * (euv * d0 + u * d1 + v * d2 - d3);
        double a2 = d2 - v * d0;
        double a4 = (d1 + (a - u) * d0 - a * a3) / (2.0 * a);
        double a1 = d0 - a3 - a4;

        WV(0, 0, k) = a1; WV(0, mu, k) = a1 * (u - a); WV(0, mv, k) = a1 * v;
        WV(0, 3, k) = a1 * (enth - u * a); WV(0, 4, k) = 0.0;
        SP(0, k) = u - a;
        WV(1, 0, k) = 0.0; WV(1, mu, k) = 0.0; WV(1, mv, k) = a2;
        WV(1, 3, k) = a2 * v; WV(1, 4, k) = 0.0;
        SP(1, k) = u;
        WV(2, 0, k) = a3; WV(2, mu, k) = a3 * u; WV(2, mv, k) = a3 * v;
        WV(2, 3, k) = a3 * 0.5 * u2v2; WV(2, 4, k) = 0.0;
        SP(2, k) = u;
        WV(3, 0, k) = a4; WV(3, mu, k) = a4 * (u + a); WV(3, mv, k) = a4 * v;
        WV(3, 3, k) = a4 * (enth + u * a); WV(3, 4, k) = 0.0;
        SP(3, k) = u + a;
        WV(4, 0, k) = 0.0; WV(4, mu, k) = 0.0; WV(4, mv, k) = 0.0;
        WV(4, 3, k) = 0.0; WV(4, 4, k) = ql[4] - qr[4];
        SP(4, k) = u;
    }

    if (!c->efix) {
        for (int m = 0; m < 5; m++)
            for (int k = 1; k < sl->N; k++) {
                double am = 0.0, ap = 0.0;
                for (int w = 0; w < 5; w++) {
                    if (SP(w, k) < 0.0)
                        am = am + SP(w, k) * WV(w, m, k);
                    else
                        ap = ap + SP(w, k) * WV(w, m, k);
                }
                AM(m, k) = am;
                AP(m, k) = ap;
            }
        return;
    }

    for (int k = 1; k < sl->N; k++) {
        double sfract;
        /* 1-wave */
        double rhoim1 = Q1(0, k - 1);
        double pim1 = g1 * (Q1(3, k - 1) -
                            0.5 * (Q1(mu, k - 1) * Q1(mu, k - 1) +
                                   Q1(mv, k - 1) * Q1(mv, k - 1)) / rhoim1);
        double cim1 = sqrt(gam * pim1 / rhoim1);
        double s0 = Q1(mu, k - 1) / rhoim1 - cim1;
        if (s0 >= 0.0 && SP(0, k) > 0.0) {
            for (int m = 0; m < 5; m++) AM(m, k) = 0.0;
            goto done;
        }
        {
            double rho1 = Q1(0, k - 1) + WV(0, 0, k);
            double rhou1 = Q1(mu, k - 1) + WV(0, mu, k);
            double rhov1 = Q1(mv, k - 1) + WV(0, mv, k);
            double en1 = Q1(3, k - 1) + WV(0, 3, k);
            double p1 = g1 * (en1 - 0.5 * (rhou1 * rhou1 + rhov1 * rhov1) / rho1);
            double c1 = sqrt(gam * p1 / rho1);
            double s1 = rhou1 / rho1 - c1;
            if (s0 < 0.0 && s1 > 0.0)
                sfract = s0 * (s1 - SP(0, k)) / (s1 - s0);
            else if (SP(0, k) < 0.0)
                sfract = SP(0, k);
            else
                sfract = 0.0;
        }
        for (int m = 0; m < 5; m++) AM(m, k) = sfract * WV(0, m, k);
        /* contact / shear / tracer */
        if (SP(1, k) >= 0.0) goto done;
        for (int m = 0; m < 5; m++) {
            AM(m, k) = AM(m, k) + SP(1, k) * WV(1, m, k);
            AM(m, k) = AM(m, k) + SP(2, k) * WV(2, m, k);
            AM(m, k) = AM(m, k) + SP(4, k) * WV(4, m, k);
        }
        /* 4-wave */
        {
            double rhoi = Q1(0, k);
            double pi_ = g1 * (Q1(3, k) -
                               0.5 * (Q1(mu, k) * Q1(mu, k) + Q1(mv, k) * Q1(mv, k)) /
                                   rhoi);
            double ci = sqrt(gam * pi_ / rhoi);
            double s3 = Q1(mu, k) / rhoi + ci;
            double rho2 = Q1(0, k) - WV(3, 0, k);
            double rhou2 = Q1(mu, k) - WV(3, mu, k);
            double rhov2 = Q1(mv, k) - WV(3, mv, k);
            double en2 = Q1(3, k) - WV(3, 3, k);
            double p2 = g1 * (en2 - 0.5 * (rhou2 * rhou2 + rhov2 * rhov2) / rho2);
            double c2 = sqrt(gam * p2 / rho2);
            double s2 = rhou2 / rho2 + c2;
            if (s2 < 0.0 && s3 > 0.0)
                sfract = s2 * (s3 - SP(3, k)) / (s3 - s2);
            else if (SP(3, k) < 0.0)
                sfract = SP(3, k);
            else
                goto done;
        }
        for (int m = 0; m < 5; m++) AM(m, k) = AM(m, k) + sfract * WV(3, m, k);
    done:;
    }
    for (int m = 0; m < 5; m++)
        for (int k = 1; k < sl->N; k++) {
            double df = 0.0;
            for (int w = 0; w < 5; w++) df = df + SP(w, k) * WV(w, m, k);
            AP(m, k) = df - AM(m, k);
        }
}

/* HLLE (Harten, Lax and van Leer 1983, with the wave-speed bounds of
 * Einfeldt 1988; LeVeque, Finite Volume Methods for Hyperbolic Problems,
 * section 15.3.7).  No counterpart in the reference (its Euler solver is
 * Roe only, applications/clawpack/euler/2d/rp/README); the interface is that
 * of clawpack46_rpn2_euler5.f.  Two waves W1 = q_m - q_a, W2 = q_b - q_m
 * with the HLL middle state q_m; W1 is formed from the jumps, so equal
 * states give exactly zero waves:
 *   s1 = min(u_a - c_a, u_roe - a_roe),  s2 = max(u_b + c_b, u_roe + a_roe)
 *   W1 = (s2 (q_b - q_a) - (f(q_b) - f(q_a))) / (s2 - s1),  W2 = (q_b - q_a) - W1
 * The tracer (field 5) is a conserved density with flux u q_5. */
static void euler_flux(double g1, int mu, int mv, const double* q, double* f, double* un, double* c,
                       double gam)
{
    const double u = q[mu] / q[0];
    const double p = g1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
    f[0] = q[mu];
    f[mu] = q[mu] * u + p;
    f[mv] = q[mv] * u;
    f[3] = u * (q[3] + p);
    f[4] = q[4] * u;
    *un = u;
    *c = sqrt(gam * p / q[0]);
}

void orc_rpn2_hlle_face(double gamma, int ixy, const double* qa, const double* qb, double* w,
                        double* s, double* amdq, double* apdq, double* r)
{
    orc_cfg c;
    memset(&c, 0, sizeof(c));
    c.gamma = gamma;
    const int mu = (ixy == 1) ? 1 : 2;
    const int mv = (ixy == 1) ? 2 : 1;
    const double g1 = gamma - 1.0;
    euler_roe(&c, mu, mv, qb, qa, r);
    double fa[5], fb[5], ua, ca, ub, cb;
    euler_flux(g1, mu, mv, qa, fa, &ua, &ca, gamma);
    euler_flux(g1, mu, mv, qb, fb, &ub, &cb, gamma);
    const double s1 = dmin(ua - ca, r[0] - r[3]);
    const double s2 = dmax(ub + cb, r[0] + r[3]);
    for (int m = 0; m < 5; m++) {
        const double dq = qb[m] - qa[m];
        const double w1 = (s2 * dq - (fb[m] - fa[m])) / (s2 - s1);
        const double w2 = dq - w1;
        w[m] = w1;
        w[5 + m] = w2;
        amdq[m] = dmin(s1, 0.0) * w1 + dmin(s2, 0.0) * w2;
        apdq[m] = dmax(s1, 0.0) * w1 + dmax(s2, 0.0) * w2;
    }
    s[0] = s1;
    s[1] = s2;
}

static void rpn2_hlle(const orc_cfg* c, int ixy, slice_t* sl)
{
    for (int k = 1; k < sl->N; k++) {
        double qa[5], qb[5], w[10], sp[2], am[5], ap[5], r[ROE_N];
        for (int m = 0; m < 5; m++) { qa[m] = Q1(m, k - 1); qb[m] = Q1(m, k); }
        orc_rpn2_hlle_face(c->gamma, ixy, qa, qb, w, sp, am, ap, r);
        for (int t = 0; t < ROE_N; t++) ROE(t, k) = r[t];
        for (int m = 0; m < 5; m++) {
            WV(0, m, k) = w[m];
            WV(1, m, k) = w[5 + m];
            AM(m, k) = am[m];
            AP(m, k) = ap[m];
        }
        SP(0, k) = sp[0];
        SP(1, k) = sp[1];
    }
}

/* ------------------------------------------------------------------------ */
/* Transverse solvers: split asdq(face k) into down/up-going parts.          */
/* ------------------------------------------------------------------------ */

static void rpt2_adv(const orc_cfg* c, int ixy, int imp, const double* asdq,
                     slice_t* sl)
{
    (void)c;
    const int kv = 2 - ixy; /* Fortran kv = 3 - ixy, 1-based */
    const int N = sl->N;
    for (int k = 1; k < N; k++) {
        int i1 = k - 2 + imp;
        for (int m = 0; m < sl->meqn; m++) {
            BM(m, k) = dmin(A2(kv, i1), 0.0) * asdq[m * N + k];
            BP(m, k) = dmax(A3(kv, i1), 0.0) * asdq[m * N + k];
        }
    }
}

static void rpt2_swe(const orc_cfg* c, int ixy, int imp, const double* asdq,
                     slice_t* sl)
{
    (void)imp;
    const int mu = (ixy == 1) ? 1 : 2;
    const int mv = (ixy == 1) ? 2 : 1;
    const int N = sl->N;
    for (int k = 1; k < N; k++) {
        double u = ROE(0, k), v = ROE(1, k), a = ROE(2, k);
        double q0 = asdq[0 * N + k], qu = asdq[mu * N + k], qv = asdq[mv * N + k];
        double a1, a3;
        double a2 = qu - u * q0;
        if (c->claw5) {
            a1 = (0.5 / a) * ((v + a) * q0 - qv);
            a3 = (0.5 / a) * (-(v - a) * q0 + qv);
        } else {
            a1 = ((v + a) * q0 - qv) / (2 * a);
            a3 = (-(v - a) * q0 + qv) / (2 * a);
        }
        double wb[3][3], sb[3];
        wb[0][0] = a1; wb[0][mu] = a1 * u; wb[0][mv] = a1 * (v - a); sb[0] = v - a;
        wb[1][0] = 0.0; wb[1][mu] = a2; wb[1][mv] = 0.0; sb[1] = v;
        wb[2][0] = a3; wb[2][mu] = a3 * u; wb[2][mv] = a3 * (v + a); sb[2] = v + a;
        for (int m = 0; m < 3; m++) {
            double bm = 0.0, bp = 0.0;
            for (int w = 0; w < 3; w++) {
                bm = bm + dmin(sb[w], 0.0) * wb[w][m];
                bp = bp + dmax(sb[w], 0.0) * wb[w][m];
            }
            BM(m, k) = bm;
            BP(m, k) = bp;
        }
    }
}

static void rpt2_euler5(const orc_cfg* c, int ixy, int imp, const double* asdq,
                        slice_t* sl)
{
    (void)imp;
    const int mu = (ixy == 1) ? 1 : 2;
    const int mv = (ixy == 1) ? 2 : 1;
    const int N = sl->N;
    for (int k = 1; k < N; k++) {
        /* rpt2_euler5 recomputes the Roe averages from the same two cells with
         * the same formulas as rpn2_euler5; the cached values are identical. */
        double u = ROE(0, k), v = ROE(1, k), enth = ROE(2, k), a = ROE(3, k),
               g1a2 = ROE(4, k), euv = ROE(5, k), u2v2 = ROE(6, k);
        if (c->claw5) g1a2 = (c->gamma - 1.0) / (a * a);
        double s0 = asdq[0 * N + k], su = asdq[mu * N + k], sv = asdq[mv * N + k],
               s3 = asdq[3 * N + k], s4 = asdq[4 * N + k];
        double a3 = g1a2 * (euv * s0 + u * su + v * sv - s3);
        double a2 = su - u * s0;
        double a4 = (sv + (a - v) * s0 - a * a3) / (2.0 * a);
        double a1 = s0 - a3 - a4;
        double wb[4][5], sb[4];
        wb[0][0] = a1; wb[0][mu] = a1 * u; wb[0][mv] = a1 * (v - a);
        wb[0][3] = a1 * (enth - v * a); wb[0][4] = 0.0; sb[0] = v - a;
        wb[1][0] = a3; wb[1][mu] = a3 * u + a2; wb[1][mv] = a3 * v;
        wb[1][3] = a3 * 0.5 * u2v2 + a2 * u; wb[1][4] = 0.0; sb[1] = v;
        wb[2][0] = a4; wb[2][mu] = a4 * u; wb[2][mv] = a4 * (v + a);
        wb[2][3] = a4 * (enth + v * a); wb[2][4] = 0.0; sb[2] = v + a;
        wb[3][0] = 0.0; wb[3][mu] = 0.0; wb[3][mv] = 0.0; wb[3][3] = 0.0;
        wb[3][4] = s4; sb[3] = v;
        for (int m = 0; m < 5; m++) {
            double bm = 0.0, bp = 0.0;
            for (int w = 0; w < 4; w++) {
                if (sb[w] < 0.0)
                    bm = bm + sb[w] * wb[w][m];
                else
                    bp = bp + sb[w] * wb[w][m];
            }
            BM(m, k) = bm;
            BP(m, k) = bp;
        }
    }
}

/* Transverse split of the HLLE pair: the HLL form of the Roe transverse
 * matrix B at the Roe average of the face.  B asdq is the Roe split summed
 * over its waves (sb_w wb_w, as rpt2_euler5 forms them); with the bounds
 * t1 = v - a, t2 = v + a (v the transverse Roe velocity):
 *   V1 = (t2 asdq - B asdq) / (t2 - t1),  V2 = asdq - V1
 *   bmasdq = min(t1, 0) V1 + min(t2, 0) V2,  bpasdq = max(t1, 0) V1 + max(t2, 0) V2
 * so bmasdq + bpasdq = B asdq.  A fluctuation that the Roe split gives no
 * transverse speed (a contact's) still spreads with the acoustic speeds, as
 * the normal HLLE solver spreads it. */
void orc_rpt2_hlle_face(int ixy, const double* r, const double* d, double* bm, double* bp)
{
    const int mu = (ixy == 1) ? 1 : 2;
    const int mv = (ixy == 1) ? 2 : 1;
    double u = r[0], v = r[1], enth = r[2], a = r[3], g1a2 = r[4], euv = r[5], u2v2 = r[6];
    double a3 = g1a2 * (euv * d[0] + u * d[mu] + v * d[mv] - d[3]);
    double a2 = d[mu] - u * d[0];
    double a4 = (d[mv] + (a - v) * d[0] - a * a3) / (2.0 * a);
    double a1 = d[0] - a3 - a4;
    double wb[4][5], sb[4];
    wb[0][0] = a1; wb[0][mu] = a1 * u; wb[0][mv] = a1 * (v - a);
    wb[0][3] = a1 * (enth - v * a); wb[0][4] = 0.0; sb[0] = v - a;
    wb[1][0] = a3; wb[1][mu] = a3 * u + a2; wb[1][mv] = a3 * v;
    wb[1][3] = a3 * 0.5 * u2v2 + a2 * u; wb[1][4] = 0.0; sb[1] = v;
    wb[2][0] = a4; wb[2][mu] = a4 * u; wb[2][mv] = a4 * (v + a);
    wb[2][3] = a4 * (enth + v * a); wb[2][4] = 0.0; sb[2] = v + a;
    wb[3][0] = 0.0; wb[3][mu] = 0.0; wb[3][mv] = 0.0; wb[3][3] = 0.0;
    wb[3][4] = d[4]; sb[3] = v;
    const double t1 = v - a, t2 = v + a;
    for (int m = 0; m < 5; m++) {
        double bq = 0.0;
        for (int w = 0; w < 4; w++) bq = bq + sb[w] * wb[w][m];
        const double v1 = (t2 * d[m] - bq) / (t2 - t1);
        const double v2 = d[m] - v1;
        bm[m] = dmin(t1, 0.0) * v1 + dmin(t2, 0.0) * v2;
        bp[m] = dmax(t1, 0.0) * v1 + dmax(t2, 0.0) * v2;
    }
}

static void rpt2_hlle(const orc_cfg* c, int ixy, int imp, const double* asdq, slice_t* sl)
{
    (void)c;
    (void)imp;
    const int N = sl->N;
    for (int k = 1; k < N; k++) {
        double r[ROE_N], d[5], bm[5], bp[5];
        for (int t = 0; t < ROE_N; t++) r[t] = ROE(t, k);
        for (int m = 0; m < 5; m++) d[m] = asdq[m * N + k];
        orc_rpt2_hlle_face(ixy, r, d, bm, bp);
        for (int m = 0; m < 5; m++) {
            BM(m, k) = bm[m];
            BP(m, k) = bp[m];
        }
    }
}

/* Conservative transport with f-waves (applications/clawpack/transport/2d/
 * all/clawpack46_rpn2cons_fw_manifold.f:23-46): edge length g of cell k,
 * normal velocities on the two sides of face k scaled by it, the f-wave
 * (flux difference) and the average speed; everything goes one way. */
static void rpn2_fw(const orc_cfg* c, int ixy, slice_t* sl)
{
    (void)c;
    const int idir = ixy - 1;
    for (int k = 1; k < sl->N; k++) {
        const double g = A2(5 + idir, k);
        const double urrot = g * A2(1 + 2 * idir, k);
        const double ulrot = g * A2(2 + 2 * idir, k - 1);
        const double qrr = Q1(0, k), qll = Q1(0, k - 1);
        const double uhat = (ulrot + urrot) / 2.0;
        if (uhat >= 0) {
            AM(0, k) = 0.0;
            AP(0, k) = urrot * qrr - ulrot * qll;
        } else {
            AM(0, k) = urrot * qrr - ulrot * qll;
            AP(0, k) = 0.0;
        }
        WV(0, 0, k) = urrot * qrr - ulrot * qll;
        SP(0, k) = uhat;
    }
}

/* clawpack46_rpt2cons_manifold.f:25-63: the split of face k's fluctuation
 * with the averages of the transverse normal velocities on the face's two
 * cells in the neighbouring slices (aux2: this slice, aux3: the next), scaled
 * by the transverse edge length of cell i1 = k-1 (imp 1) or k (imp 2). */
static void rpt2_fw(const orc_cfg* c, int ixy, int imp, const double* asdq, slice_t* sl)
{
    (void)c;
    const int idir = 2 - ixy;
    const int N = sl->N;
    for (int k = 1; k < N; k++) {
        const int i1 = k - 2 + imp;
        double g = A2(5 + idir, i1);
        double vrrot = g * A2(1 + 2 * idir, k);
        double vlrot = g * A2(2 + 2 * idir, k - 1);
        double vhat = (vrrot + vlrot) / 2.0;
        BM(0, k) = dmin(vhat, 0.0) * asdq[k];
        g = A3(5 + idir, i1);
        vrrot = g * A3(1 + 2 * idir, k);
        vlrot = g * A3(2 + 2 * idir, k - 1);
        vhat = (vrrot + vlrot) / 2.0;
        BP(0, k) = dmax(vhat, 0.0) * asdq[k];
    }
}

static void rpn2(const orc_cfg* c, int ixy, slice_t* sl)
{
    switch (c->solver) {
    case ORC_SOLVER_TRANSPORT_FW: rpn2_fw(c, ixy, sl); break;
    case ORC_SOLVER_ADV: rpn2_adv(c, ixy, sl); break;
    case ORC_SOLVER_SWE_ROE: rpn2_swe(c, ixy, sl); break;
    case ORC_SOLVER_EULER_HLLE: rpn2_hlle(c, ixy, sl); break;
    default: rpn2_euler5(c, ixy, sl); break;
    }
}

static void rpt2(const orc_cfg* c, int ixy, int imp, const double* asdq, slice_t* sl)
{
    switch (c->solver) {
    case ORC_SOLVER_TRANSPORT_FW: rpt2_fw(c, ixy, imp, asdq, sl); break;
    case ORC_SOLVER_ADV: rpt2_adv(c, ixy, imp, asdq, sl); break;
    case ORC_SOLVER_SWE_ROE: rpt2_swe(c, ixy, imp, asdq, sl); break;
    case ORC_SOLVER_EULER_HLLE: rpt2_hlle(c, ixy, imp, asdq, sl); break;
    default: rpt2_euler5(c, ixy, imp, asdq, sl); break;
    }
}

/* ------------------------------------------------------------------------ */
/* Limiter                                                                  */
/* ------------------------------------------------------------------------ */

double orc_limiter_phi(int mthlim, double r)
{
    switch (mthlim) {
    case 1: return dmax(0.0, dmin(1.0, r));                                  /* minmod   */
    case 2: return dmax(dmax(0.0, dmin(1.0, 2.0 * r)), dmin(2.0, r));       /* superbee */
    case 3: return (r + fabs(r)) / (1.0 + fabs(r));                          /* van Leer */
    case 4: {                                                                /* MC       */
        double c = (1.0 + r) / 2.0;
        return dmax(0.0, dmin(dmin(c, 2.0), 2.0 * r));
    }
    case 5: return r;                                                        /* B-W      */
    case 6: {                                                                /* gen. mm  */
        /* th = 1.3 is a default-real (single precision) literal assigned to a
         * double in the reference, so its value is (double)1.3f. */
        const double th = (double)1.3f;
        return dmax(0.0, dmin(dmin(th * r, (1 + r) / 2.0), th));
    }
    default: return 1.0;
    }
}

static void limiter(const orc_cfg* c, slice_t* sl)
{
    for (int w = 0; w < sl->mwaves; w++) {
        if (c->mthlim[w] == 0) continue;
        double dotr = 0.0;
        /* Fortran i = 0 .. mx+1 -> slice cell k = i + mbc - 1 */
        for (int k = sl->mbc - 1; k <= sl->n + sl->mbc; k++) {
            double wnorm2 = 0.0;
            double dotl = dotr;
            dotr = 0.0;
            for (int m = 0; m < sl->meqn; m++) {
                wnorm2 = wnorm2 + WV(w, m, k) * WV(w, m, k);
                dotr = dotr + WV(w, m, k) * WV(w, m, k + 1);
            }
            if (k == sl->mbc - 1) continue;
            if (wnorm2 == 0.0) continue;
            double r = (SP(w, k) > 0.0) ? dotl / wnorm2 : dotr / wnorm2;
            /* Clawpack 5: computed GO TO over 1..5, 6 falls through to minmod */
            double phi = orc_limiter_phi(c->claw5 && c->mthlim[w] == 6 ? 1 : c->mthlim[w], r);
            for (int m = 0; m < sl->meqn; m++) WV(w, m, k) = phi * WV(w, m, k);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* 1-D wave-propagation kernel on one slice (clawpack46_flux2)               */
/* ------------------------------------------------------------------------ */

static double flux1d(const orc_cfg* c, int ixy, double dtdx, slice_t* sl)
{
    const int N = sl->N, meqn = sl->meqn, mw = sl->mwaves;
    int limit = 0;
    for (int w = 0; w < mw; w++)
        if (c->mthlim[w] > 0) limit = 1;

    memset(sl->faddm, 0, sizeof(double) * meqn * N);
    memset(sl->faddp, 0, sizeof(double) * meqn * N);
    memset(sl->gaddm, 0, sizeof(double) * 2 * meqn * N);
    memset(sl->gaddp, 0, sizeof(double) * 2 * meqn * N);

    rpn2(c, ixy, sl);

    /* clawpack46_flux2fw.f works on faces i = 1..mx+1 only (slice cells
     * mbc..n+mbc), which is all the interior update reads; clawpack46_flux2.f
     * on every face */
    const int fw = c->solver == ORC_SOLVER_TRANSPORT_FW;
    const int k0 = fw ? sl->mbc : 1;
    const int kg = fw ? sl->n + sl->mbc : N - 2;  /* last face of the Godunov fluxes */
    const int kc = fw ? sl->n + sl->mbc : N - 1;  /* last face of the corrections, splits */
    for (int k = k0; k <= kg; k++)
        for (int m = 0; m < meqn; m++) {
            FP(m, k) = FP(m, k) - AP(m, k);
            FM(m, k) = FM(m, k) + AM(m, k);
        }

    double cfl = 0.0;
    for (int w = 0; w < mw; w++)
        for (int k = sl->mbc; k <= sl->n + sl->mbc; k++) {
            /* dmax1(cfl, dtdx(i)*s, -dtdx(i-1)*s) with constant dtdx */
            double a = dtdx * SP(w, k);
            double b = -dtdx * SP(w, k);
            cfl = dmax(dmax(cfl, a), b);
        }

    if (c->order2 == 2) {
        if (limit) limiter(c, sl);
        for (int k = k0; k <= kc; k++) {
            double dtdxave = 0.5 * (dtdx + dtdx);
            for (int m = 0; m < meqn; m++) {
                double cq = 0.0;
                for (int w = 0; w < mw; w++) {
                    if (fw) /* dsign(1, s) (1 - |s| dtdx) fwave, clawpack46_flux2fw.f:176-177 */
                        cq = cq + copysign(1.0, SP(w, k)) * (1.0 - fabs(SP(w, k)) * dtdxave) * WV(w, m, k);
                    else
                        cq = cq + fabs(SP(w, k)) * (1.0 - fabs(SP(w, k)) * dtdxave) * WV(w, m, k);
                }
                CQ(m, k) = cq;
                FM(m, k) = FM(m, k) + 0.5 * cq;
                FP(m, k) = FP(m, k) + 0.5 * cq;
            }
        }
    }

    if (c->transverse > 0) {
        if (c->order2 > 1 && c->transverse == 2) {
            for (int k = k0; k <= (fw ? kc : N - 2); k++)
                for (int m = 0; m < meqn; m++) {
                    AM(m, k) = AM(m, k) + CQ(m, k);
                    AP(m, k) = AP(m, k) - CQ(m, k);
                }
        }
        rpt2(c, ixy, 1, sl->amdq, sl);
        for (int m = 0; m < meqn; m++)
            for (int k = k0; k <= kc; k++) {
                double gu = 0.5 * dtdx * BM(m, k);
                GM(0, m, k - 1) = GM(0, m, k - 1) - gu;
                GP(0, m, k - 1) = GP(0, m, k - 1) - gu;
                gu = 0.5 * dtdx * BP(m, k);
                GM(1, m, k - 1) = GM(1, m, k - 1) - gu;
                GP(1, m, k - 1) = GP(1, m, k - 1) - gu;
            }
        rpt2(c, ixy, 2, sl->apdq, sl);
        for (int m = 0; m < meqn; m++)
            for (int k = k0; k <= kc; k++) {
                double gu = 0.5 * dtdx * BM(m, k);
                GM(0, m, k) = GM(0, m, k) - gu;
                GP(0, m, k) = GP(0, m, k) - gu;
                gu = 0.5 * dtdx * BP(m, k);
                GM(1, m, k) = GM(1, m, k) - gu;
                GP(1, m, k) = GP(1, m, k) - gu;
            }
    }
    return cfl;
}

/* ------------------------------------------------------------------------ */
/* Patch step                                                               */
/* ------------------------------------------------------------------------ */

static int check_cfg(const orc_cfg* c)
{
    if (c->mx < 1 || c->my < 1 || c->mbc < 2) return 1;
    if (c->mwaves < 1 || c->mwaves > MAXW || c->meqn < 1 || c->meqn > MAXE) return 1;
    if (c->order2 < 1 || c->order2 > 2 || c->transverse < 0 || c->transverse > 2) return 1;
    for (int w = 0; w < c->mwaves; w++)
        if (c->mthlim[w] < 0 || c->mthlim[w] > 6) return 1;
    switch (c->solver) {
    case ORC_SOLVER_ADV: if (c->maux < 2) return 1; break;
    case ORC_SOLVER_TRANSPORT_FW: if (c->maux < 7 || c->meqn != 1 || c->mwaves != 1) return 1; break;
    case ORC_SOLVER_SWE_ROE: if (c->meqn != 3 || c->mwaves != 3) return 1; break;
    case ORC_SOLVER_EULER5_ROE: if (c->meqn != 5 || c->mwaves != 5) return 1; break;
    case ORC_SOLVER_EULER_HLLE: if (c->meqn != 5 || c->mwaves != 2 || c->claw5) return 1; break;
    default: return 1;
    }
    return 0;
}

/* Conservative flux f(q) of the shallow-water equations in direction idir
 * (0 = x, 1 = y), rpn2_cons_update of the radial-dam application
 * (applications/clawpack/shallow/2d/radialdam/rpn2cons_update.f:7-51). */
void orc_swe_flux(int idir, double grav, const double* q, double* flux)
{
    const int mu = idir == 0 ? 1 : 2, mv = idir == 0 ? 2 : 1;
    const double hun = q[mu], hut = q[mv], h = q[0];
    const double un = hun / h;
    flux[0] = hun;
    flux[mu] = hun * hun / h + 0.5 * grav * (h * h);
    flux[mv] = un * hut;
}

long orc_reg_offset(int mx, int my, int meqn, int face, int block)
{
    long off = 0;
    for (int f = 0; f < face; f++) off += 4L * meqn * (f < 2 ? my : mx);
    return off + (long)block * meqn * (face < 2 ? my : mx);
}

long orc_reg_size(int mx, int my, int meqn) { return 8L * meqn * (mx + my); }

/* Edge-flux registers before the update
 * (clawpack46_time_sync_store_flux, clawpack46_time_sync.f:37-156):
 * E_k += dt*el*f(q) at the first interior (k = 0) and first ghost (k = 1)
 * cell of every edge cell; el = dy on x faces, dx on y faces (Cartesian
 * clawpatch_time_sync_setup, fclaw2d_clawpatch_utils.f:296-312). */
static void store_flux(const orc_cfg* c, const double* q, double dx, double dy, double dt,
                       double* reg)
{
    const int mx = c->mx, my = c->my, mbc = c->mbc, meqn = c->meqn;
    const int NX = mx + 2 * mbc, NY = my + 2 * mbc;
    double qv[8], fl[8];
#define QV(i, j) for (int m = 0; m < meqn; m++) qv[m] = q[((size_t)m * NY + (j)) * NX + (i)]
    for (int j = 0; j < my; j++)
        for (int k = 0; k < 2; k++) {
            double* e0 = reg + orc_reg_offset(mx, my, meqn, 0, 2 + k);
            double* e1 = reg + orc_reg_offset(mx, my, meqn, 1, 2 + k);
            QV(mbc - k, mbc + j);
            orc_swe_flux(0, c->grav, qv, fl);
            for (int m = 0; m < meqn; m++) e0[m * my + j] = e0[m * my + j] + dt * dy * fl[m];
            QV(mbc + mx - 1 + k, mbc + j);
            orc_swe_flux(0, c->grav, qv, fl);
            for (int m = 0; m < meqn; m++) e1[m * my + j] = e1[m * my + j] + dt * dy * fl[m];
        }
    for (int i = 0; i < mx; i++)
        for (int k = 0; k < 2; k++) {
            double* e2 = reg + orc_reg_offset(mx, my, meqn, 2, 2 + k);
            double* e3 = reg + orc_reg_offset(mx, my, meqn, 3, 2 + k);
            QV(mbc + i, mbc - k);
            orc_swe_flux(1, c->grav, qv, fl);
            for (int m = 0; m < meqn; m++) e2[m * mx + i] = e2[m * mx + i] + dt * dx * fl[m];
            QV(mbc + i, mbc + my - 1 + k);
            orc_swe_flux(1, c->grav, qv, fl);
            for (int m = 0; m < meqn; m++) e3[m * mx + i] = e3[m * mx + i] + dt * dx * fl[m];
        }
#undef QV
}

/* Fluctuation registers after the update
 * (clawpack46_time_sync_accumulate_waves, clawpack46_time_sync.f:163-226). */
static void accumulate_waves(const orc_cfg* c, const double* fp, const double* fm,
                             const double* gp, const double* gm, double dx, double dy,
                             double dt, double* reg)
{
    const int mx = c->mx, my = c->my, mbc = c->mbc, meqn = c->meqn;
    const int NX = mx + 2 * mbc, NY = my + 2 * mbc;
#define F2(arr, i, j, m) arr[((size_t)(m) * NY + (j)) * NX + (i)]
    double* P0 = reg + orc_reg_offset(mx, my, meqn, 0, 0);
    double* M0 = reg + orc_reg_offset(mx, my, meqn, 0, 1);
    double* P1 = reg + orc_reg_offset(mx, my, meqn, 1, 0);
    double* M1 = reg + orc_reg_offset(mx, my, meqn, 1, 1);
    for (int j = 0; j < my; j++)
        for (int m = 0; m < meqn; m++) {
            const int r = m * my + j;
            P0[r] = P0[r] - dt * dy * F2(fp, mbc, mbc + j, m);
            M0[r] = M0[r] + dt * dy * F2(fm, mbc, mbc + j, m);
            P1[r] = P1[r] - dt * dy * F2(fp, mbc + mx, mbc + j, m);
            M1[r] = M1[r] + dt * dy * F2(fm, mbc + mx, mbc + j, m);
        }
    double* P2 = reg + orc_reg_offset(mx, my, meqn, 2, 0);
    double* M2 = reg + orc_reg_offset(mx, my, meqn, 2, 1);
    double* P3 = reg + orc_reg_offset(mx, my, meqn, 3, 0);
    double* M3 = reg + orc_reg_offset(mx, my, meqn, 3, 1);
    for (int i = 0; i < mx; i++)
        for (int m = 0; m < meqn; m++) {
            const int r = m * mx + i;
            P2[r] = P2[r] - dt * dx * F2(gp, mbc + i, mbc, m);
            M2[r] = M2[r] + dt * dx * F2(gm, mbc + i, mbc, m);
            P3[r] = P3[r] - dt * dx * F2(gp, mbc + i, mbc + my, m);
            M3[r] = M3[r] + dt * dx * F2(gm, mbc + i, mbc + my, m);
        }
#undef F2
}

static int step2_impl(const orc_cfg* c, double* q, const double* aux, double dx, double dy,
                      double dt, double* cflout, double* reg);

int orc_step2(const orc_cfg* c, double* q, const double* aux, double dx, double dy,
              double dt, double* cflout)
{
    return step2_impl(c, q, aux, dx, dy, dt, cflout, NULL);
}

int orc_step2_ts(const orc_cfg* c, double* q, const double* aux, double dx, double dy,
                 double dt, double* cflout, double* reg)
{
    if (c->solver != ORC_SOLVER_SWE_ROE) return 1;  /* only SWE has a conservative flux */
    return step2_impl(c, q, aux, dx, dy, dt, cflout, reg);
}

static int step2_impl(const orc_cfg* c, double* q, const double* aux, double dx, double dy,
                      double dt, double* cflout, double* reg)
{
    if (check_cfg(c)) return 1;
    const int mx = c->mx, my = c->my, mbc = c->mbc, meqn = c->meqn;
    const int maux = c->maux, mw = c->mwaves;
    const int NX = mx + 2 * mbc, NY = my + 2 * mbc;
    const int NM = (NX > NY ? NX : NY);
    const size_t T = (size_t)NX * NY;

    double* fm = calloc(T * meqn, sizeof(double));
    double* fp = calloc(T * meqn, sizeof(double));
    double* gm = calloc(T * meqn, sizeof(double));
    double* gp = calloc(T * meqn, sizeof(double));

    slice_t S, *sl = &S;
    sl->mbc = mbc; sl->meqn = meqn; sl->maux = maux; sl->mwaves = mw;
    size_t nw = (size_t)NM;
    double* pool = calloc(nw * (size_t)(meqn * (8 + 2 * 2) + 3 * (maux > 0 ? maux : 1) +
                                        mw * meqn + mw + ROE_N),
                          sizeof(double));
    double* p = pool;
    sl->q1d = p; p += nw * meqn;
    sl->aux1 = p; p += nw * (maux > 0 ? maux : 1);
    sl->aux2 = p; p += nw * (maux > 0 ? maux : 1);
    sl->aux3 = p; p += nw * (maux > 0 ? maux : 1);
    sl->wave = p; p += nw * mw * meqn;
    sl->s = p; p += nw * mw;
    sl->amdq = p; p += nw * meqn;
    sl->apdq = p; p += nw * meqn;
    sl->cqxx = p; p += nw * meqn;
    sl->bmasdq = p; p += nw * meqn;
    sl->bpasdq = p; p += nw * meqn;
    sl->faddm = p; p += nw * meqn;
    sl->faddp = p; p += nw * meqn;
    sl->gaddm = p; p += nw * 2 * meqn;
    sl->gaddp = p; p += nw * 2 * meqn;
    sl->roe = p; p += nw * ROE_N;

    const double dtdx = dt / dx, dtdy = dt / dy;
    double cflgrid = 0.0;

#define QT(i, j, m) q[((size_t)(m) * NY + (j)) * NX + (i)]
#define AT(i, j, a) aux[((size_t)(a) * NY + (j)) * NX + (i)]
#define F2(arr, i, j, m) arr[((size_t)(m) * NY + (j)) * NX + (i)]

    /* x-sweeps: rows j = 2-mbc .. my+mbc-1 (Fortran) -> 1 .. NY-2 */
    sl->n = mx; sl->N = NX;
    for (int j = 1; j <= NY - 2; j++) {
        for (int m = 0; m < meqn; m++)
            for (int i = 0; i < NX; i++) sl->q1d[m * NX + i] = QT(i, j, m);
        for (int a = 0; a < maux; a++)
            for (int i = 0; i < NX; i++) {
                sl->aux1[a * NX + i] = AT(i, j - 1, a);
                sl->aux2[a * NX + i] = AT(i, j, a);
                sl->aux3[a * NX + i] = AT(i, j + 1, a);
            }
        double cfl1d = flux1d(c, 1, dtdx, sl);
        cflgrid = dmax(cflgrid, cfl1d);
        for (int m = 0; m < meqn; m++)
            for (int i = 1; i <= NX - 1; i++) {
                F2(fm, i, j, m) = F2(fm, i, j, m) + FM(m, i);
                F2(fp, i, j, m) = F2(fp, i, j, m) + FP(m, i);
                F2(gm, i, j, m) = F2(gm, i, j, m) + GM(0, m, i);
                F2(gp, i, j, m) = F2(gp, i, j, m) + GP(0, m, i);
                F2(gm, i, j + 1, m) = F2(gm, i, j + 1, m) + GM(1, m, i);
                F2(gp, i, j + 1, m) = F2(gp, i, j + 1, m) + GP(1, m, i);
            }
    }

    /* y-sweeps: columns i = 2-mbc .. mx+mbc-1 -> 1 .. NX-2 */
    sl->n = my; sl->N = NY;
    for (int i = 1; i <= NX - 2; i++) {
        for (int m = 0; m < meqn; m++)
            for (int j = 0; j < NY; j++) sl->q1d[m * NY + j] = QT(i, j, m);
        for (int a = 0; a < maux; a++)
            for (int j = 0; j < NY; j++) {
                sl->aux1[a * NY + j] = AT(i - 1, j, a);
                sl->aux2[a * NY + j] = AT(i, j, a);
                sl->aux3[a * NY + j] = AT(i + 1, j, a);
            }
        double cfl1d = flux1d(c, 2, dtdy, sl);
        cflgrid = dmax(cflgrid, cfl1d);
        for (int m = 0; m < meqn; m++)
            for (int j = 1; j <= NY - 1; j++) {
                F2(gm, i, j, m) = F2(gm, i, j, m) + FM(m, j);
                F2(gp, i, j, m) = F2(gp, i, j, m) + FP(m, j);
                F2(fm, i, j, m) = F2(fm, i, j, m) + GM(0, m, j);
                F2(fp, i, j, m) = F2(fp, i, j, m) + GP(0, m, j);
                F2(fm, i + 1, j, m) = F2(fm, i + 1, j, m) + GM(1, m, j);
                F2(fp, i + 1, j, m) = F2(fp, i + 1, j, m) + GP(1, m, j);
            }
    }

    if (reg) {
        store_flux(c, q, dx, dy, dt, reg);  /* q still holds the data before the update */
        accumulate_waves(c, fp, fm, gp, gm, dx, dy, dt, reg);
    }

    /* flux-difference update over [1-mbc, mx+mbc-1] x [1-mbc, my+mbc-1] */
    for (int m = 0; m < meqn; m++)
        for (int i = 0; i <= NX - 2; i++)
            for (int j = 0; j <= NY - 2; j++)
                QT(i, j, m) = QT(i, j, m) - dtdx * (F2(fm, i + 1, j, m) - F2(fp, i, j, m)) -
                              dtdy * (F2(gm, i, j + 1, m) - F2(gp, i, j, m));
#undef QT
#undef AT
#undef F2

    free(pool);
    free(fm); free(fp); free(gm); free(gp);
    if (cflout) *cflout = cflgrid;
    return 0;
}

int orc_step2_batch(const orc_cfg* c, int npatch, double* q, const double* aux,
                    const double* dx, const double* dy, double dt, double* cfl)
{
    if (check_cfg(c)) return 1;
    const size_t T = (size_t)(c->mx + 2 * c->mbc) * (c->my + 2 * c->mbc);
    int err = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(| : err)
    for (int p = 0; p < npatch; p++) {
        double cp = 0.0;
        err |= orc_step2(c, q + (size_t)p * T * c->meqn,
                         aux ? aux + (size_t)p * T * c->maux : NULL, dx[p], dy[p], dt,
                         &cp);
        if (cfl) cfl[p] = cp;
    }
    return err;
}

/* ------------------------------------------------------------------------ */
/* Swirl velocity field (b4step2 + psi)                                     */
/* ------------------------------------------------------------------------ */

static double swirl_psi(double x, double y, double pi)
{
    double sx = sin(pi * x), sy = sin(pi * y);
    return ((sx * sx) * (sy * sy)) / pi;
}

void orc_swirl_aux(int mx, int my, int mbc, double xlower, double ylower, double dx,
                   double dy, double t, double dt, double tperiod, double* aux)
{
    const int NX = mx + 2 * mbc, NY = my + 2 * mbc;
    const double pi = 4.0 * atan(1.0);
    const double pi2 = 2.0 * pi;
    if (tperiod == 0.0) return;
    double vt = cos(pi2 * (t + dt / 2.0) / tperiod);
    double vty = vt / dy, vtx = vt / dx;
    for (int j = 0; j < NY; j++) {
        int jf = j - mbc + 1;
        double yll = ylower + (double)(jf - 1) * dy;
        for (int i = 0; i < NX; i++) {
            int ifo = i - mbc + 1;
            double xll = xlower + (double)(ifo - 1) * dx;
            double pij = swirl_psi(xll, yll, pi);
            aux[(size_t)j * NX + i] = vty * (swirl_psi(xll, yll + dy, pi) - pij);
            aux[((size_t)NY + j) * NX + i] = -vtx * (swirl_psi(xll + dx, yll, pi) - pij);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Physical boundary conditions                                             */
/* ------------------------------------------------------------------------ */

int orc_bc2(int mx, int my, int mbc, int meqn, const int mthbc[4], double* q)
{
    const int NX = mx + 2 * mbc, NY = my + 2 * mbc;
#define QB(i, j, m) q[((size_t)(m) * NY + (j)) * NX + (i)]
    /* Fortran i_f -> C i = i_f + mbc - 1 */
    const int i1 = mbc, imx = mx + mbc - 1, j1 = mbc, jmy = my + mbc - 1;
    for (int f = 0; f < 4; f++)
        if (mthbc[f] == 0 || mthbc[f] == 2) return 1;
    /* left */
    if (mthbc[0] == 1) {
        for (int m = 0; m < meqn; m++)
            for (int ibc = 1; ibc <= mbc; ibc++)
                for (int j = 0; j < NY; j++) QB(i1 - ibc, j, m) = QB(i1, j, m);
    } else if (mthbc[0] == 3) {
        for (int m = 0; m < meqn; m++)
            for (int ibc = 1; ibc <= mbc; ibc++)
                for (int j = 0; j < NY; j++) QB(i1 - ibc, j, m) = QB(i1 + ibc - 1, j, m);
        for (int ibc = 1; ibc <= mbc; ibc++)
            for (int j = 0; j < NY; j++) QB(i1 - ibc, j, 1) = -QB(i1 + ibc - 1, j, 1);
    }
    /* right */
    if (mthbc[1] == 1) {
        for (int m = 0; m < meqn; m++)
            for (int ibc = 1; ibc <= mbc; ibc++)
                for (int j = 0; j < NY; j++) QB(imx + ibc, j, m) = QB(imx, j, m);
    } else if (mthbc[1] == 3) {
        for (int m = 0; m < meqn; m++)
            for (int ibc = 1; ibc <= mbc; ibc++)
                for (int j = 0; j < NY; j++) QB(imx + ibc, j, m) = QB(imx + 1 - ibc, j, m);
        for (int ibc = 1; ibc <= mbc; ibc++)
            for (int j = 0; j < NY; j++) QB(imx + ibc, j, 1) = -QB(imx + 1 - ibc, j, 1);
    }
    /* bottom */
    if (mthbc[2] == 1) {
        for (int m = 0; m < meqn; m++)
            for (int jbc = 1; jbc <= mbc; jbc++)
                for (int i = 0; i < NX; i++) QB(i, j1 - jbc, m) = QB(i, j1, m);
    } else if (mthbc[2] == 3) {
        for (int m = 0; m < meqn; m++)
            for (int jbc = 1; jbc <= mbc; jbc++)
                for (int i = 0; i < NX; i++) QB(i, j1 - jbc, m) = QB(i, j1 + jbc - 1, m);
        for (int jbc = 1; jbc <= mbc; jbc++)
            for (int i = 0; i < NX; i++) QB(i, j1 - jbc, 2) = -QB(i, j1 + jbc - 1, 2);
    }
    /* top */
    if (mthbc[3] == 1) {
        for (int m = 0; m < meqn; m++)
            for (int jbc = 1; jbc <= mbc; jbc++)
                for (int i = 0; i < NX; i++) QB(i, jmy + jbc, m) = QB(i, jmy, m);
    } else if (mthbc[3] == 3) {
        for (int m = 0; m < meqn; m++)
            for (int jbc = 1; jbc <= mbc; jbc++)
                for (int i = 0; i < NX; i++) QB(i, jmy + jbc, m) = QB(i, jmy + 1 - jbc, m);
        for (int jbc = 1; jbc <= mbc; jbc++)
            for (int i = 0; i < NX; i++) QB(i, jmy + jbc, 2) = -QB(i, jmy + 1 - jbc, 2);
    }
#undef QB
    return 0;
}
