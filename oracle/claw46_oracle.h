/*
 * claw46_oracle.h -- CPU FP64 restatement of the Clawpack-4.6 wave-propagation
 * patch step used by ForestClaw.  TEST INFRASTRUCTURE ONLY: the GPU product
 * path never links or calls this code; tests, smoke() and the bench's
 * cpu_baseline leg use it as the checker and as the CPU baseline.
 *
 * Storage follows the clawpatch FArrayBox layout
 * q(1-mbc:mx+mbc, 1-mbc:my+mbc, meqn), i fastest, field-major
 * (src/patches/clawpatch/fclaw2d_clawpatch.cpp:305-335).  In C the tile index
 * is (i, j) with i in [0, mx+2mbc), j in [0, my+2mbc); Fortran index
 * i_f = i - mbc + 1.
 */
#ifndef CLAW46_ORACLE_H
#define CLAW46_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_SOLVER_ADV = 0,        /* clawpack46_rpn2adv / rpt2adv            */
    ORC_SOLVER_SWE_ROE = 1,    /* shallow/2d/rp clawpack46_rpn2 / rpt2    */
    ORC_SOLVER_EULER5_ROE = 2, /* euler/2d/rp clawpack46_rpn2/rpt2_euler5 */
    ORC_SOLVER_EULER_HLLE = 3, /* Euler + tracer, HLLE normal solver (2 waves) and
                                  the HLL form of the Roe transverse split        */
    ORC_SOLVER_TRANSPORT_FW = 4 /* conservative transport, f-waves: the transport
                                   applications' clawpack46_rpn2cons_fw_manifold /
                                   clawpack46_rpt2cons_manifold with the f-wave
                                   flux function clawpack46_flux2fw; aux per cell
                                   (1-based) 1 capacity, 2-5 normal velocities on
                                   the left, right, bottom and top edges, 6-7 the
                                   x- and y-edge lengths over dy and dx           */
};

typedef struct {
    int solver;
    int mx, my, mbc;
    int meqn, maux, mwaves;
    int order2;        /* method(2): 1 = Godunov only, 2 = + limited corrections */
    int transverse;    /* method(3): 0, 1, 2                                    */
    int mthlim[8];     /* limiter per wave family, 0..6                         */
    int efix;          /* entropy fix (reference: always on, data efix /.true./)*/
    double grav;       /* /cparam/ grav  (SWE)                                  */
    double gamma;      /* /cparam/ gamma (Euler), gamma1 = gamma - 1            */
    int claw5;         /* 1: the Clawpack 5 solvers' operation order (SWE:
                          clawpack5_rpn2.f90:76,78, clawpack5_rpt2.f90:59,61
                          scale by 0.5/a; Euler5: clawpack5_rpt2_euler5.f90:67
                          takes gamma1/c**2 for the transverse split)           */
} orc_cfg;

/* HLLE normal solver at one face for the Euler equations with a conserved
 * tracer (meqn 5: rho, rho u, rho v, E, rho phi), gamma-law gas.  qa is the
 * state on the low side, qb on the high side, ixy 1 = x face, 2 = y face.
 * Two waves w[2][5] with speeds s[2] (Einfeldt's bounds), fluctuations
 * amdq / apdq.  Returns the Roe averages r[7] (u, v, enth, a, g1/a^2,
 * enth - u^2 - v^2, u^2 + v^2) that the transverse split uses. */
void orc_rpn2_hlle_face(double gamma, int ixy, const double* qa, const double* qb, double* w,
                        double* s, double* amdq, double* apdq, double* r);

/* Transverse split of the HLLE pair at one face: d (5) the fluctuation, r
 * the Roe averages orc_rpn2_hlle_face returned, ixy the normal direction of
 * the face (the split is along the other one).  bm + bp = B d with B the Roe
 * transverse matrix; the split is its HLL form with the bounds v -+ a. */
void orc_rpt2_hlle_face(int ixy, const double* r, const double* d, double* bm, double* bp);

/* One patch step in place (clawpack46_step2_wrap semantics, mcapa = 0).
 * Returns 0 on success, nonzero on bad configuration.  *cfl receives the
 * patch Courant number (cflgrid). */
int orc_step2(const orc_cfg* cfg, double* q, const double* aux,
              double dx, double dy, double dt, double* cfl);

/* npatch independent patches, tiles packed back to back; OpenMP over
 * patches when compiled with -fopenmp.  cfl[p] per patch. */
int orc_step2_batch(const orc_cfg* cfg, int npatch, double* q,
                    const double* aux, const double* dx, const double* dy,
                    double dt, double* cfl);

/* Swirl b4step2: edge velocities from the stream function (2 aux fields). */
void orc_swirl_aux(int mx, int my, int mbc, double xlower, double ylower,
                   double dx, double dy, double t, double dt, double tperiod,
                   double* aux);

/* clawpack46_bc2_default: mthbc[4] (left,right,bottom,top), -1 = not a
 * physical face, 1 = extrapolation, 3 = solid wall.  Returns nonzero for
 * 0 (user) or 2 (periodic), which the reference rejects. */
int orc_bc2(int mx, int my, int mbc, int meqn, const int mthbc[4], double* q);

/* Limiter function phi(r) of clawpack46_inlinelimiter for one family. */
double orc_limiter_phi(int mthlim, double r);

/* Conservative time synchronisation (SWE only; claw46_oracle.c,
 * claw46_timesync.c).  Registers: orc_reg_size doubles per patch. */
void orc_swe_flux(int idir, double grav, const double* q, double* flux);
long orc_reg_offset(int mx, int my, int meqn, int face, int block);
long orc_reg_size(int mx, int my, int meqn);
/* orc_step2 plus store_flux (before) and accumulate_waves (after) into reg. */
int orc_step2_ts(const orc_cfg* cfg, double* q, const double* aux, double dx, double dy,
                 double dt, double* cfl, double* reg);
void orc_time_sync_f2c(int mx, int my, int mbc, int meqn, int iface, int igrid, double area,
                       double* qc, const double* rc, const double* rf);
void orc_time_sync_samesize(int mx, int my, int mbc, int meqn, int iface, double area,
                            double* q, const double* rt, const double* rn);
void orc_time_sync_reset_face(int mx, int my, int meqn, int iface, double* reg);

/* Ghost-cell and regrid kernels for one patch pair (claw46_ghost.c).
 * (sx, sy): same-size index shift; (ox, oy): half-size offset
 * (xlower_coarse - xlower_fine) * 2^level_fine * mx (resp. my). */
double orc_compute_slopes(double sl, double sr);
void orc_copy_face(int mx, int my, int mbc, int meqn, double* qthis, const double* qnbr,
                   int iface, int sx, int sy);
void orc_copy_corner(int mx, int my, int mbc, int meqn, double* qthis, const double* qnbr,
                     int icorner, int sx, int sy);
void orc_average_face(int mx, int my, int mbc, int meqn, double* qcoarse, const double* qfine,
                      int iface_coarse, int ox, int oy);
void orc_average_corner(int mx, int my, int mbc, int meqn, double* qcoarse, const double* qfine,
                        int icorner_coarse, int ox, int oy);
void orc_interpolate_face(int mx, int my, int mbc, int meqn, const double* qcoarse, double* qfine,
                          int iface_coarse, int ox, int oy);
void orc_interpolate_corner(int mx, int my, int mbc, int meqn, const double* qcoarse, double* qfine,
                            int icorner_coarse, int ox, int oy);
void orc_interpolate2fine(int mx, int my, int mbc, int meqn, const double* qcoarse, double* qfine,
                          int igrid);
void orc_average2coarse(int mx, int my, int mbc, int meqn, double* qcoarse, const double* qfine,
                        int igrid);
int orc_timeinterp(int mx, int my, int mbc, int meqn, const double* qcurr, const double* qlast,
                   double* qinterp, double alpha);
int orc_tag_gradient(int mx, int my, int mbc, const double* q, int ivar, double dx, double dy,
                     double threshold);

#ifdef __cplusplus
}
#endif
#endif
