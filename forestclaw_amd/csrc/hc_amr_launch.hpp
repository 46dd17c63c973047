// hc_amr_launch.hpp -- host launchers of the AMR kernels (hc_amr_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "hc_amr.hpp"

namespace hc {

enum QinitKind {
    QINIT_CONST = 0,        // q = p[0] (field 0), 0 elsewhere
    QINIT_SWIRL = 1,        // swirl/user_4.6/qinit.f: q = 1 for x < p[0]
    QINIT_RADIAL_DAM = 2,   // radialdam: h = p[3] inside circle (p[0],p[1]) radius p[2], else p[4]
    QINIT_HUMPS = 3,        // h = p[0] + sum of p[1] Gaussians (cx, cy, A, w) from p[2]
    QINIT_SHOCK_BUBBLE = 4, // gamma, xshock, rho_inf, u_inf, p_inf, x0, y0, r0, rho_bubble
    QINIT_SHOCKBUBBLE_APP = 5  // the shockbubble application: gamma, x0, y0, r0, rhoin, pinf
};

struct QinitParams {
    double p[56];
    int nsub;  // sub-samples per direction for cell averages of discontinuous data
};

template <class R> void launch_ghost(const GhostOp*, int, const Pools<R>&, int, int, int, hipStream_t);
template <class R> void launch_bc(const BcOp*, int, const Pools<R>&, int, int, int, hipStream_t);
// the same over several lists in one launch
template <class R> void launch_ghost(const Segs<GhostOp>&, const Pools<R>&, int, int, int, hipStream_t);
// ghost operations and, beside them in the same launch, time interpolations
// (list k of level ti_levels[k], weight alpha[k]) of the ring cells that ghost
// fills read; a copy from the interpolated pool of such a level interpolates
// its source cell itself
template <class R>
void launch_ghost_ti(const Segs<GhostOp>&, const Segs<TiOp>&, const double* alpha, const int* ti_levels,
                     const Pools<R>&, int, int, int, hipStream_t);
template <class R> void launch_bc(const Segs<BcOp>&, const Pools<R>&, int, int, int, hipStream_t);
// the same with the coarse step's Courant-number reduction as an extra workgroup
template <class R>
void launch_bc_cfl(const Segs<BcOp>&, const Pools<R>&, int, int, int, const CflFin&, hipStream_t);
template <class R>
void launch_timeinterp(const TiOp*, int, const Pools<R>&, int, int, int, double, hipStream_t);
template <class R> void launch_xfer(const XferOp*, int, const Pools<R>&, R*, int, int, int, hipStream_t);
template <class R>
void launch_tag(const int*, const int*, int, const Pools<R>&, const double*, int, int, int, int,
                double, int, int*, hipStream_t);
template <class R>
void launch_tag_stats(const int*, const int*, int, const Pools<R>&, const double*, int, int, int, int, double*,
                      hipStream_t);
template <class R>
void launch_qinit(int, const QinitParams&, const int*, const int*, int, const Pools<R>&,
                  const double*, const double*, int, int, int, hipStream_t);
template <class R>
void launch_mass(const int*, const int*, int, const Pools<R>&, const double*, int, int, int, double*,
                 hipStream_t);
template <class R> void launch_copy_tiles(const int*, int, const R*, R*, long long, hipStream_t);
template <class R>
void launch_ts_correct(const TsOp*, int, R*, const double*, long long, const double*, int, int, int,
                       hipStream_t);
void launch_ts_reset(const int*, int, double*, long long, int, int, int, hipStream_t);
// several levels in one launch: list s corrects patches whose q is qs[s]
template <class R>
void launch_ts_correct(const Segs<TsOp>&, R* const* qs, const double*, long long, const double*, int, int,
                       int, hipStream_t);
void launch_ts_reset(const ResetSegs&, double*, long long, int, int, int, hipStream_t);
// n doubles set to zero by a kernel (safe under stream capture)
void launch_zero(double* p, long long n, hipStream_t s);
// physical BCs, then time-sync corrections, one patch per workgroup
template <class R>
void launch_bc_ts(const Segs<BcTsOp>&, const Pools<R>&, const double* reg, long long regstride,
                  const double* dxdy, int, int, int, const CflFin&, hipStream_t);
// Network input tiles of n patches (bf16 [n][36][36][8]); returns HIPCLAW_*.
template <class R>
int launch_make_tiles(const int*, const int*, int, const Pools<R>&, const double*, int, int, int, double,
                      void*, hipStream_t);

}  // namespace hc
