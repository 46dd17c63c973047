// hc_amr_kernels.hip -- device kernels of the AMR driver (gfx950):
// ghost fill, physical BCs on lists, time interpolation, regrid transfers,
// tagging, initial conditions and conservation sums.  Built with
// -ffp-contract=off so that the FP64 instantiations round like the
// reference's FP64 Fortran (copies and averages are then bitwise equal).
//
// Tile layout: q[(m*NY + j)*NX + i], NX = mx+4, NY = my+4 (mbc = 2); the
// 0-based tile index i corresponds to the reference's i_f = i - 1.
#include "hc_amr.hpp"
#include "hc_amr_launch.hpp"
#include "hc_bc.hpp"
#include "hc_common.hpp"

namespace hc {

constexpr int MBC = 2;

__device__ inline int floordiv2(int a) { return a >= 0 ? a / 2 : -((1 - a) / 2); }

// Fortran sign(1, x): +1 or -1 with the sign bit of x
__device__ inline double sign1(double x) { return __builtin_copysign(1.0, x); }
__device__ inline float sign1(float x) { return __builtin_copysignf(1.0f, x); }

template <class R> __device__ inline R* pool_ptr(const Pools<R>& P, int level, bool ti)
{
    return ti ? P.ti : P.q[level];
}

// AMRClaw slope limiter used for ghost interpolation (mth = 5,
// src/patches/clawpatch/fclaw2d_clawpatch_utils.f:52-55):
// min(2|sl|, 2|sr|, |sc|) * max(0, sign(1, sl*sr)) * sign(1, sc)
template <class R> __device__ inline R amr_slope(R sl, R sr)
{
    R sc = (sl + sr) / R(2);
    R a = R(2) * M<R>::abs_(sl), b = R(2) * M<R>::abs_(sr), c = M<R>::abs_(sc);
    R mn = dmin(dmin(a, b), c);
    R f = dmax(R(0), sign1(sl * sr));
    return (mn * f) * sign1(sc);
}

// ---------------------------------------------------------------------------
// Ghost fill.  One thread per (operation, item); an item is one ghost cell
// (copy / average) or one coarse cell whose four children are written
// (interpolation).  All meqn fields are handled by the thread.
// ---------------------------------------------------------------------------
__host__ __device__ inline int ti_ring_size(int mx, int my)
{
    return (mx < 2 * MBC || my < 2 * MBC) ? mx * my : 2 * MBC * mx + (my - 2 * MBC) * 2 * MBC;
}

// Time interpolation of one cell of the ring of interior cells within MBC of
// the patch edge (the cells that ghost fills read), item k of the ring:
// rows jj < MBC and jj >= my - MBC first, then the side columns.
template <class R>
__device__ inline void ti_ring_cell(const TiOp& op, int k, const Pools<R>& P, int mx, int my, int meqn, R alpha)
{
    const int NX = mx + 2 * MBC, NY = my + 2 * MBC;
    const int nrow = 2 * MBC * mx;
    int ii, jj;
    if (mx < 2 * MBC || my < 2 * MBC) {  // every cell is within MBC of an edge
        jj = k / mx;
        ii = k - jj * mx;
    } else if (k < nrow) {
        const int r = k / mx;
        ii = k - r * mx;
        jj = r < MBC ? r : my - 2 * MBC + r;
    } else {
        const int k2 = k - nrow;
        const int r = k2 / (2 * MBC), cc = k2 - r * (2 * MBC);
        jj = MBC + r;
        ii = cc < MBC ? cc : mx - 2 * MBC + cc;
    }
    const long long base = (long long)op.slot * NX * NY * meqn;
    const R* qc = P.q[op.level] + base;
    const R* ql = P.qlast[op.level] + base;
    R* qi = P.ti + base;
    for (int m = 0; m < meqn; m++) {
        const long long idx = ((long long)m * NY + jj + MBC) * NX + ii + MBC;
        qi[idx] = ql[idx] + alpha * (qc[idx] - ql[idx]);
    }
}

// A copy from the time-interpolated pool of a level whose interpolation runs
// in the same launch: the source cell interpolated here, with the operations
// of ti_ring_cell (so the same bits), instead of read from the pool that the
// launch is still writing.
template <class R>
__device__ inline void copy_ti_direct(R* __restrict__ D, const Pools<R>& P, const GhostOp& op, int i, int j,
                                      long long T, int NX, int NY, int meqn, R alpha)
{
    const long long base = (long long)op.src * T * meqn;
    const R* qc = P.q[op.slev] + base;
    const R* ql = P.qlast[op.slev] + base;
    for (int m = 0; m < meqn; m++) {
        const long long s = ((long long)m * NY + j + op.oy) * NX + i + op.ox;
        D[((long long)m * NY + j) * NX + i] = ql[s] + alpha * (qc[s] - ql[s]);
    }
}

template <class R>
__global__ void k_ghost(const Segs<GhostOp> SG, const Segs<TiOp> ST, const TiWeights<R> TW, Pools<R> P, int mx,
                        int my, int meqn)
{
    const int NX = mx + 2 * MBC, NY = my + 2 * MBC;
    const long long T = (long long)NX * NY;
    const int items = MBC * (mx > my ? mx : my);
    const long long gid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long ng = (long long)SG.total() * items;
    if (gid >= ng) {
        // time interpolations that run beside the ghost operations
        const int ring = ti_ring_size(mx, my);
        const long long t = gid - ng;
        if (t >= (long long)ST.total() * ring) return;
        const int o = (int)(t / ring);
        const int s = ST.seg(o);
        ti_ring_cell<R>(ST.ops[s][o - ST.start[s]], (int)(t - (long long)o * ring), P, mx, my, meqn, TW.a[s]);
        return;
    }
    const int o = (int)(gid / items);
    const int k = (int)(gid - (long long)o * items);
    const GhostOp op = SG.at(o);
    R* __restrict__ D = pool_ptr(P, op.dlev, op.flags & 1) + (long long)op.dst * T * meqn;
    const R* __restrict__ S = pool_ptr(P, op.slev, op.flags & 2) + (long long)op.src * T * meqn;
#define AT(B, i, j, m) B[((long long)(m) * NY + (j)) * NX + (i)]

    if (op.kind == G_COPY_FACE || op.kind == G_AVG_FACE) {
        const int f = op.sub;
        const int len = f < 2 ? my : mx;
        if (k >= MBC * len) return;
        int i, j;
        const int layer = k % MBC, along = MBC + k / MBC;
        switch (f) {
        case 0: i = layer; j = along; break;
        case 1: i = MBC + mx + layer; j = along; break;
        case 2: j = layer; i = along; break;
        default: j = MBC + my + layer; i = along; break;
        }
        if (op.kind == G_COPY_FACE) {
            if ((op.flags & 2) && ((TW.pend >> op.slev) & 1u)) {
                copy_ti_direct<R>(D, P, op, i, j, T, NX, NY, meqn, TW.al[op.slev]);
                return;
            }
            for (int m = 0; m < meqn; m++) AT(D, i, j, m) = AT(S, i + op.ox, j + op.oy, m);
        } else {
            const int fi = 2 * i - MBC + op.ox, fj = 2 * j - MBC + op.oy;
            if (fi < MBC || fi + 1 > MBC + mx - 1 || fj < MBC || fj + 1 > MBC + my - 1) return;
            for (int m = 0; m < meqn; m++) {
                R s = R(0);
                s = s + AT(S, fi, fj, m);
                s = s + AT(S, fi + 1, fj, m);
                s = s + AT(S, fi, fj + 1, m);
                s = s + AT(S, fi + 1, fj + 1, m);
                AT(D, i, j, m) = s / R(4);
            }
        }
        return;
    }
    if (op.kind == G_COPY_CORNER || op.kind == G_AVG_CORNER) {
        if (k >= MBC * MBC) return;
        const int ibc = k % MBC + 1, jbc = k / MBC + 1;
        const int c = op.sub;
        const int i = (c & 1) ? MBC + mx - 1 + ibc : MBC - ibc;
        const int j = (c & 2) ? MBC + my - 1 + jbc : MBC - jbc;
        if (op.kind == G_COPY_CORNER) {
            if ((op.flags & 2) && ((TW.pend >> op.slev) & 1u)) {
                copy_ti_direct<R>(D, P, op, i, j, T, NX, NY, meqn, TW.al[op.slev]);
                return;
            }
            for (int m = 0; m < meqn; m++) AT(D, i, j, m) = AT(S, i + op.ox, j + op.oy, m);
        } else {
            const int fi = 2 * i - MBC + op.ox, fj = 2 * j - MBC + op.oy;
            for (int m = 0; m < meqn; m++) {
                R s = R(0);
                s = s + AT(S, fi, fj, m);
                s = s + AT(S, fi + 1, fj, m);
                s = s + AT(S, fi, fj + 1, m);
                s = s + AT(S, fi + 1, fj + 1, m);
                AT(D, i, j, m) = s / R(4);
            }
        }
        return;
    }
    // interpolation: coarse source cell (ci, cj), fine destination children
    int ci, cj;
    bool check;
    if (op.kind == G_INTERP_FACE) {
        const int f = op.sub;
        const int len = f < 2 ? my : mx;
        if (k >= len) return;
        switch (f) {
        case 0: ci = MBC; cj = MBC + k; break;
        case 1: ci = MBC + mx - 1; cj = MBC + k; break;
        case 2: cj = MBC; ci = MBC + k; break;
        default: cj = MBC + my - 1; ci = MBC + k; break;
        }
        check = true;
    } else {
        if (k >= 1) return;
        const int c = op.sub;
        ci = (c & 1) ? MBC + mx - 1 : MBC;
        cj = (c & 2) ? MBC + my - 1 : MBC;
        check = false;
    }
    const int fi = 2 * ci - MBC + op.ox, fj = 2 * cj - MBC + op.oy;
    if (check && (fi < 0 || fi + 1 > NX - 1 || fj < 0 || fj + 1 > NY - 1)) return;
    for (int m = 0; m < meqn; m++) {
        const R qc = AT(S, ci, cj, m);
        R gx = amr_slope(qc - AT(S, ci - 1, cj, m), AT(S, ci + 1, cj, m) - qc);
        R gy = amr_slope(qc - AT(S, ci, cj - 1, m), AT(S, ci, cj + 1, m) - qc);
        const R lo = R(-0.25), hi = R(0.25);
        AT(D, fi, fj, m) = (qc + gx * lo) + gy * lo;
        AT(D, fi + 1, fj, m) = (qc + gx * hi) + gy * lo;
        AT(D, fi, fj + 1, m) = (qc + gx * lo) + gy * hi;
        AT(D, fi + 1, fj + 1, m) = (qc + gx * hi) + gy * hi;
    }
#undef AT
}

// ---------------------------------------------------------------------------
// clawpack46_bc2_default on a list of patches (left/right over the full
// height, then bottom/top over the full width).
// ---------------------------------------------------------------------------
template <class R>
__global__ void k_bc_list(const Segs<BcOp> S, Pools<R> P, int mx, int my, int meqn, CflFin cf)
{
    const int o = blockIdx.x;
    if (o >= S.total()) {
        if (o == S.total() && cf.acc) cfl_final_block(cf);
        return;
    }
    const BcOp op = S.at(o);
    const long long T = (long long)(mx + 2 * MBC) * (my + 2 * MBC);
    bc_patch<R>(op, pool_ptr(P, op.level, op.flags & 1) + (long long)op.slot * T * meqn, mx, my, meqn);
}

// ---------------------------------------------------------------------------
// Time interpolation on the interior band of width mbc
// (src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_timeinterp.f:33-69).
// ---------------------------------------------------------------------------
template <class R>
__global__ void k_timeinterp(const TiOp* __restrict__ ops, int nops, Pools<R> P, int mx, int my,
                             int meqn, R alpha)
{
    const int NX = mx + 2 * MBC, NY = my + 2 * MBC;
    const long long cells = (long long)mx * my;
    const long long gid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (gid >= nops * cells) return;
    const int o = (int)(gid / cells);
    const int c = (int)(gid - (long long)o * cells);
    const int ii = c % mx, jj = c / mx;  // 0-based interior indices
    if (!(ii < MBC || ii >= mx - MBC || jj < MBC || jj >= my - MBC)) return;
    const TiOp op = ops[o];
    const long long base = (long long)op.slot * NX * NY * meqn;
    const R* qc = P.q[op.level] + base;
    const R* ql = P.qlast[op.level] + base;
    R* qi = P.ti + base;
    for (int m = 0; m < meqn; m++) {
        const long long idx = ((long long)m * NY + jj + MBC) * NX + ii + MBC;
        qi[idx] = ql[idx] + alpha * (qc[idx] - ql[idx]);
    }
}

// ---------------------------------------------------------------------------
// Regrid transfers into new pools: keep (copy whole tile), interpolate2fine
// (fclaw2d_clawpatch46_interpolate.f:324-400) and average2coarse
// (fclaw2d_clawpatch46_average.f:282-364).
// ---------------------------------------------------------------------------
template <class R>
__global__ void k_xfer(const XferOp* __restrict__ ops, int nops, Pools<R> Old, R* __restrict__ qnew,
                       int mx, int my, int meqn)
{
    const int NX = mx + 2 * MBC, NY = my + 2 * MBC;
    const long long T = (long long)NX * NY;
    const long long gid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (gid >= nops * T) return;
    const int o = (int)(gid / T);
    const int cell = (int)(gid - (long long)o * T);
    const int i = cell % NX, j = cell / NX;
    const XferOp op = ops[o];
    R* D = qnew + (long long)op.dst * T * meqn;
#define AT(B, i, j, m) B[((long long)(m) * NY + (j)) * NX + (i)]
    if (op.kind == 0) {
        const R* S = Old.q[op.srclev] + (long long)op.src[0] * T * meqn;
        for (int m = 0; m < meqn; m++) AT(D, i, j, m) = AT(S, i, j, m);
        return;
    }
    const int ig = op.igrid & 1, jg = op.igrid >> 1;
    if (op.kind == 1) {
        // reference loops: coarse-local i in [1-ig, mx/2+1-ig], child ii in
        // {1,2}, fine index (i-1)*2 + ii (1-based); invert for this fine cell
        const int fi_f = i - MBC + 1, fj_f = j - MBC + 1;
        const int icl = floordiv2(fi_f - 1) + 1, jcl = floordiv2(fj_f - 1) + 1;
        const int ii = fi_f - (icl - 1) * 2, jj = fj_f - (jcl - 1) * 2;
        if (icl < 1 - ig || icl > mx / 2 + 1 - ig || jcl < 1 - jg || jcl > my / 2 + 1 - jg) return;
        const int ci = icl + ig * (mx / 2) + MBC - 1, cj = jcl + jg * (my / 2) + MBC - 1;
        const R* S = Old.q[op.srclev] + (long long)op.src[0] * T * meqn;
        const R sx = (R(ii) - R(1.0) - R(0.5)) / R(2), sy = (R(jj) - R(1.0) - R(0.5)) / R(2);
        for (int m = 0; m < meqn; m++) {
            const R qc = AT(S, ci, cj, m);
            R gx = amr_slope(qc - AT(S, ci - 1, cj, m), AT(S, ci + 1, cj, m) - qc);
            R gy = amr_slope(qc - AT(S, ci, cj - 1, m), AT(S, ci, cj + 1, m) - qc);
            AT(D, i, j, m) = (qc + sx * gx) + sy * gy;
        }
        return;
    }
    // average2coarse: coarse interior (i_f + ic_add, j_f + jc_add), i_f in [1, mx/2]
    const int ci_f = i - MBC + 1, cj_f = j - MBC + 1;
    if (ci_f < 1 || ci_f > mx || cj_f < 1 || cj_f > my) return;
    const int g = ((ci_f > mx / 2) ? 1 : 0) + 2 * ((cj_f > my / 2) ? 1 : 0);
    const int il = ci_f - (g & 1) * (mx / 2), jl = cj_f - (g >> 1) * (mx / 2);
    const R* S = Old.q[op.srclev] + (long long)op.src[g] * T * meqn;
    for (int m = 0; m < meqn; m++) {
        R s = R(0);
        const int f0 = (il - 1) * 2 + 1 + MBC - 1, g0 = (jl - 1) * 2 + 1 + MBC - 1;
        s = s + AT(S, f0, g0, m);
        s = s + AT(S, f0 + 1, g0, m);
        s = s + AT(S, f0, g0 + 1, m);
        s = s + AT(S, f0 + 1, g0 + 1, m);
        AT(D, i, j, m) = s / R(4);
    }
#undef AT
}

// ---------------------------------------------------------------------------
// Gradient tagging (fclaw2d_clawpatch_gradient_exceeds_th.f90:47-109 on the
// interior cells, fclaw2d_clawpatch46_tag4refinement.f:33-74).  One block
// per patch; tag = 1 if any interior cell exceeds the threshold.
// mode 0: refine test -> tags[p] = exceeds;  mode 1: coarsen test for the
// patch -> tags[p] = !exceeds (a family coarsens only if all four agree).
// ---------------------------------------------------------------------------
template <class R>
__global__ void k_tag_gradient(const int* __restrict__ slots, const int* __restrict__ levels,
                               int n, Pools<R> P, const double* __restrict__ dxdy, int mx, int my,
                               int meqn, int ivar, double threshold, int mode, int* __restrict__ tags)
{
    const int p = blockIdx.x;
    if (p >= n) return;
    const int NX = mx + 2 * MBC, NY = my + 2 * MBC;
    const int slot = slots[p];
    const R* q = P.q[levels[p]] + ((long long)slot * meqn + ivar) * NX * NY;
    const double dx2 = 2 * dxdy[2 * slot], dy2 = 2 * dxdy[2 * slot + 1];
    int hit = 0;
    for (int c = threadIdx.x; c < mx * my; c += blockDim.x) {
        const int i = c % mx + MBC, j = c / mx + MBC;
        const double dqx = ((double)q[j * NX + i + 1] - (double)q[j * NX + i - 1]) / dx2;
        const double dqy = ((double)q[(j + 1) * NX + i] - (double)q[(j - 1) * NX + i]) / dy2;
        const double ds = sqrt(dqx * dqx + dqy * dqy + 0.0 * 0.0);
        if (ds > threshold) hit = 1;
    }
    hit = __syncthreads_or(hit);
    if (threadIdx.x == 0) tags[p] = mode == 0 ? hit : !hit;
}

// ---------------------------------------------------------------------------
// Per-patch statistics of the tagging variable, from which every criterion of
// the clawpatch tag pass is decided (src/patches/clawpatch/fort_4.6/
// fclaw2d_clawpatch46_tag4refinement.f:40-75 / _tag4coarsening.f:30-118 with
// fclaw2d_clawpatch_{value,minmax,difference,gradient}_exceeds_th.f90):
//   st[0], st[1]  min and max of q over the whole tile, ghost cells included
//                 (the running qmin/qmax of the reference loops; "value" tests
//                 every cell, so its decision is max > threshold)
//   st[2]         max over interior cells of max(|dq_x|, |dq_y|), central
//                 differences ("difference")
//   st[3]         max over interior cells of |grad q| with dq/(2 dx) ("gradient")
// A criterion exceeds its threshold somewhere in the patch iff the statistic
// exceeds it, so the host decides refinement (per patch) and coarsening (per
// family: the minmax range over the four siblings together) exactly as the
// reference's early-exit loops do.
// ---------------------------------------------------------------------------
template <class R>
__global__ void __launch_bounds__(256)
k_tag_stats(const int* __restrict__ slots, const int* __restrict__ levels, int n, Pools<R> P,
            const double* __restrict__ dxdy, int mx, int my, int meqn, int ivar, double* __restrict__ st)
{
    __shared__ double red[4][256];
    const int p = blockIdx.x;
    if (p >= n) return;
    const int NX = mx + 2 * MBC, NY = my + 2 * MBC;
    const int slot = slots[p];
    const R* q = P.q[levels[p]] + ((long long)slot * meqn + ivar) * NX * NY;
    const double dx2 = 2 * dxdy[2 * slot], dy2 = 2 * dxdy[2 * slot + 1];
    double qmin = (double)q[MBC * NX + MBC], qmax = qmin, dmax = 0.0, gmax = 0.0;
    for (int c = threadIdx.x; c < NX * NY; c += blockDim.x) {
        const double v = (double)q[c];
        qmin = v < qmin ? v : qmin;
        qmax = v > qmax ? v : qmax;
    }
    for (int c = threadIdx.x; c < mx * my; c += blockDim.x) {
        const int i = c % mx + MBC, j = c / mx + MBC;
        const double ex = (double)q[j * NX + i + 1] - (double)q[j * NX + i - 1];
        const double ey = (double)q[(j + 1) * NX + i] - (double)q[(j - 1) * NX + i];
        const double d = fmax(fabs(ex), fabs(ey));
        const double dqx = ex / dx2, dqy = ey / dy2;
        const double g = sqrt(dqx * dqx + dqy * dqy + 0.0 * 0.0);
        dmax = d > dmax ? d : dmax;
        gmax = g > gmax ? g : gmax;
    }
    red[0][threadIdx.x] = qmin;
    red[1][threadIdx.x] = qmax;
    red[2][threadIdx.x] = dmax;
    red[3][threadIdx.x] = gmax;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][threadIdx.x + h]);
            red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + h]);
            red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + h]);
            red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + h]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) st[4 * (long long)p + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// Initial conditions (the applications' qinit hooks) on cell centres.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Cell average of the disc indicator by the reference's polygon rule
// (cellave2, src/fortran_source2d/cellave2.f:1-95, with the radialdam
// fdisc, applications/clawpack/shallow/2d/radialdam/fdisc.f): the cell
// corners inside the disc and the crossings of the cell edges with the
// circle, located by the zero finder zeroin (Brent's method in the form of
// Forsythe, Malcolm and Moler, "Computer Methods for Mathematical
// Computations", 1977, as the reference's copy has it, including its step
// "b - xm" when xm < 0), bound a polygon whose area fraction is the weight.
// FP64 with the reference's operation order (this file is built without FMA
// contraction), so the values are the reference's.
// ---------------------------------------------------------------------------
struct Disc {
    double x0, y0, r0;
    __device__ double f(double x, double y) const
    {
        const double r2 = (x - x0) * (x - x0) + (y - y0) * (y - y0);
        return r2 - r0 * r0;
    }
};

__device__ double zeroin_edge(const Disc& dc, double xa, double ya, double xb, double yb, double tol)
{
    auto fs = [&](double s) { return dc.f(xa + s * (xb - xa), ya + s * (yb - ya)); };
    double eps = 1.0, tol1;
    do {
        eps = eps / 2.0;
        tol1 = 1.0 + eps;
    } while (tol1 > 1.0);
    double a = 0.0, b = 1.0, fa = fs(a), fb = fs(b), c, fc, d, e, xm, s, p, q, r;
    bool restart = true;
    for (int it = 0; it < 400; it++) {
        if (restart) {
            c = a;
            fc = fa;
            d = b - a;
            e = d;
        }
        if (fabs(fc) < fabs(fb)) {
            a = b;
            b = c;
            c = a;
            fa = fb;
            fb = fc;
            fc = fa;
        }
        tol1 = 2.0 * eps * fabs(b) + 0.5 * tol;
        xm = 0.5 * (c - b);
        if (fabs(xm) <= tol1 || fb == 0.0) return b;
        bool bisect = fabs(e) < tol1 || fabs(fa) <= fabs(fb);
        if (!bisect) {
            if (a == c) {
                s = fb / fa;
                p = 2.0 * xm * s;
                q = 1.0 - s;
            } else {
                q = fa / fc;
                r = fb / fc;
                s = fb / fa;
                p = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0));
                q = (q - 1.0) * (r - 1.0) * (s - 1.0);
            }
            if (p > 0.0) q = -q;
            p = fabs(p);
            if ((2.0 * p) >= (3.0 * xm * q - fabs(tol1 * q)) || p >= fabs(0.5 * e * q)) {
                bisect = true;
            } else {
                e = d;
                d = p / q;
            }
        }
        if (bisect) {
            d = xm;
            e = d;
        }
        a = b;
        fa = fb;
        if (fabs(d) > tol1) b = b + d;
        if (fabs(d) <= tol1) {
            if (xm >= 0) b = b + tol1;
            else b = b - xm;
        }
        fb = fs(b);
        restart = (fb * (fc / fabs(fc))) > 0.0;
    }
    return b;
}

__device__ double cellave_disc(const Disc& dc, double xlow, double ylow, double dx, double dy)
{
    const double xx[5] = {xlow, xlow, xlow + dx, xlow + dx, xlow};
    const double yy[5] = {ylow, ylow + dy, ylow + dy, ylow, ylow};
    bool fl[5];
    bool alll = true, allr = true;
    for (int i = 0; i < 4; i++) {
        fl[i] = dc.f(xx[i], yy[i]) < 0.0;
        alll = alll && fl[i];
        allr = allr && !fl[i];
    }
    fl[4] = fl[0];
    if (alll) return 1.0;
    if (allr) return 0.0;
    double x[10], y[10];
    int iv = 0;
    for (int i = 0; i < 4; i++) {
        if (fl[i]) {
            x[iv] = xx[i];
            y[iv] = yy[i];
            iv++;
        }
        if (fl[i] != fl[i + 1]) {
            const double ss = zeroin_edge(dc, xx[i], yy[i], xx[i + 1], yy[i + 1], 1e-14);
            x[iv] = xx[i] + ss * (xx[i + 1] - xx[i]);
            y[iv] = yy[i] + ss * (yy[i + 1] - yy[i]);
            iv++;
        }
    }
    if (iv == 0) return 0.0;
    x[iv] = x[0];
    y[iv] = y[0];
    double area = 0.0;
    for (int i = 0; i < iv; i++) area = area + 0.5 * (y[i] + y[i + 1]) * (x[i + 1] - x[i]);
    return area / (dx * dy);
}

template <class R>
__global__ void k_qinit(int kind, QinitParams prm, const int* __restrict__ slots,
                        const int* __restrict__ levels, int n, Pools<R> P,
                        const double* __restrict__ xylower, const double* __restrict__ dxdy,
                        int mx, int my, int meqn)
{
    const int NX = mx + 2 * MBC, NY = my + 2 * MBC;
    const long long T = (long long)NX * NY;
    const long long gid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (gid >= n * T) return;
    const int p = (int)(gid / T);
    const int cell = (int)(gid - (long long)p * T);
    const int i = cell % NX, j = cell / NX;
    const int slot = slots[p];
    R* q = P.q[levels[p]] + (long long)slot * T * meqn;
    const double dx = dxdy[2 * slot], dy = dxdy[2 * slot + 1];
    const double xlo = xylower[2 * slot] + (i - MBC) * dx, ylo = xylower[2 * slot + 1] + (j - MBC) * dy;
    const double xc = xlo + 0.5 * dx, yc = ylo + 0.5 * dy;
    double v[5] = {0, 0, 0, 0, 0};
    const int S = prm.nsub > 0 ? prm.nsub : 1;
    switch (kind) {
    case QINIT_SWIRL:  // q = 1 for x < 0.5 (swirl/user_4.6/qinit.f:13-25)
        v[0] = xc < prm.p[0] ? 1.0 : 0.0;
        break;
    case QINIT_RADIAL_DAM: {  // h = hin inside r0, hout outside; u = v = 0
        if (prm.nsub == 0) {  // the reference's polygon cell average
            const double w = cellave_disc(Disc{prm.p[0], prm.p[1], prm.p[2]}, xlo, ylo, dx, dy);
            v[0] = prm.p[3] * w + prm.p[4] * (1.0 - w);
            break;
        }
        int inside = 0;
        for (int b = 0; b < S; b++)
            for (int a = 0; a < S; a++) {
                const double x = xlo + (a + 0.5) * dx / S - prm.p[0];
                const double y = ylo + (b + 0.5) * dy / S - prm.p[1];
                inside += (x * x + y * y < prm.p[2] * prm.p[2]);
            }
        const double fr = (double)inside / (S * S);
        v[0] = prm.p[3] * fr + prm.p[4] * (1.0 - fr);
        break;
    }
    case QINIT_HUMPS: {  // h = h0 + sum A exp(-|x-c|^2/w^2), at rest
        double h = prm.p[0];
        const int nh = (int)prm.p[1];
        for (int k = 0; k < nh && k < 12; k++) {
            const double cx = prm.p[2 + 4 * k], cy = prm.p[3 + 4 * k];
            const double A = prm.p[4 + 4 * k], w = prm.p[5 + 4 * k];
            const double r2 = (xc - cx) * (xc - cx) + (yc - cy) * (yc - cy);
            h += A * exp(-r2 / (w * w));
        }
        v[0] = h;
        break;
    }
    case QINIT_SHOCK_BUBBLE: {
        // p: gamma, xshock, rinf, vinf, pinf, x0, y0, r0, rhob
        const double g1 = prm.p[0] - 1.0;
        double acc[5] = {0, 0, 0, 0, 0};
        for (int b = 0; b < S; b++)
            for (int a = 0; a < S; a++) {
                const double x = xlo + (a + 0.5) * dx / S, y = ylo + (b + 0.5) * dy / S;
                double rho = 1.0, u = 0.0, pr = 1.0, tr = 0.0;
                if (x < prm.p[1]) {
                    rho = prm.p[2];
                    u = prm.p[3];
                    pr = prm.p[4];
                } else {
                    const double rx = x - prm.p[5], ry = y - prm.p[6];
                    if (rx * rx + ry * ry < prm.p[7] * prm.p[7]) {
                        rho = prm.p[8];
                        tr = 1.0;
                    }
                }
                acc[0] += rho;
                acc[1] += rho * u;
                acc[3] += pr / g1 + 0.5 * rho * u * u;
                acc[4] += rho * tr;
            }
        for (int m = 0; m < 5; m++) v[m] = acc[m] / (S * S);
        break;
    }
    case QINIT_SHOCKBUBBLE_APP: {
        // applications/clawpack/euler/2d/shockbubble: setprob.f:33-58 (states),
        // user_4.6/qinit.f:23-64 (bubble by cellave2 with the idisc = 2
        // circle of fdisc.f:36-38, post-shock state left of x = 0.2, tracer
        // doubled in the interior rows left of x = 0.5)
        const double gamma = prm.p[0], gamma1 = gamma - 1.0;
        const double rhoin = prm.p[4], pinf = prm.p[5];
        const double qin[5] = {rhoin, 0.0, 0.0, 1.0 / gamma1, 1.0};
        const double qout[5] = {1.0, 0.0, 0.0, 1.0 / gamma1, 0.0};
        const double win = cellave_disc(Disc{prm.p[1], prm.p[2], prm.p[3]}, xlo, ylo, dx, dy);
        for (int m = 0; m < 5; m++) v[m] = win * qin[m] + (1.0 - win) * qout[m];
        if (xlo < 0.2) {
            const double rinf = (gamma1 + (gamma + 1) * pinf) / ((gamma + 1) + gamma1 * pinf);
            const double vinf = (1.0 / sqrt(gamma)) * (pinf - 1.0) /
                                sqrt(0.5 * ((gamma + 1) / gamma) * pinf + 0.5 * gamma1 / gamma);
            const double einf = 0.5 * rinf * vinf * vinf + pinf / gamma1;
            v[0] = rinf;
            v[1] = rinf * vinf;
            v[2] = 0.0;
            v[3] = einf;
            v[4] = 1.0;
        }
        if (xlo < 0.5 && j >= MBC && j < MBC + my) v[4] = 2.0 * v[4];
        break;
    }
    default:
        v[0] = prm.p[0];
        break;
    }
    for (int m = 0; m < meqn; m++) q[((long long)m * NY + j) * NX + i] = R(v[m]);
}

// ---------------------------------------------------------------------------
// FP64 conservation sums per patch: sum over interior of q*dx*dy with
// compensated accumulation per thread
// (fclaw2d_clawpatch46_diagnostics.f:13-57), then a fixed-order tree.
// ---------------------------------------------------------------------------
template <class R>
__global__ void k_mass(const int* __restrict__ slots, const int* __restrict__ levels, int n,
                       Pools<R> P, const double* __restrict__ dxdy, int mx, int my, int meqn,
                       double* __restrict__ out)
{
    __shared__ double red[256];
    const int p = blockIdx.x;
    if (p >= n) return;
    const int NX = mx + 2 * MBC, NY = my + 2 * MBC;
    const int slot = slots[p];
    const double area = dxdy[2 * slot] * dxdy[2 * slot + 1];
    for (int m = 0; m < meqn; m++) {
        const R* q = P.q[levels[p]] + ((long long)slot * meqn + m) * NX * NY;
        double s = 0.0, c = 0.0;
        for (int k = threadIdx.x; k < mx * my; k += blockDim.x) {
            const int i = k % mx + MBC, j = k / mx + MBC;
            const double y = (double)q[j * NX + i] * area - c;
            const double t = s + y;
            c = (t - s) - y;
            s = t;
        }
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(long long)p * meqn + m] = red[0];
        __syncthreads();
    }
}

template <class R>
__global__ void k_copy_tiles(const int* __restrict__ slots, int n, const R* __restrict__ src,
                             R* __restrict__ dst, long long tile)
{
    const long long gid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (gid >= n * tile) return;
    const int p = (int)(gid / tile);
    const long long off = (long long)slots[p] * tile + (gid - (long long)p * tile);
    dst[off] = src[off];
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline unsigned nblocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <class R>
void launch_ghost_ti(const Segs<GhostOp>& S, const Segs<TiOp>& T, const double* alpha, const int* ti_levels,
                     const Pools<R>& P, int mx, int my, int meqn, hipStream_t s)
{
    if (S.total() <= 0 && T.total() <= 0) return;
    TiWeights<R> w{};
    for (int k = 0; k < T.n; k++) {
        w.a[k] = R(alpha[k]);
        if (ti_levels && ti_levels[k] >= 0 && ti_levels[k] < MAXL) {
            w.al[ti_levels[k]] = w.a[k];
            w.pend |= 1u << ti_levels[k];
        }
    }
    const long long n =
        (long long)S.total() * MBC * (mx > my ? mx : my) + (long long)T.total() * ti_ring_size(mx, my);
    hipLaunchKernelGGL(k_ghost<R>, dim3(nblocks(n, 256)), dim3(256), 0, s, S, T, w, P, mx, my, meqn);
}

template <class R>
void launch_ghost(const Segs<GhostOp>& S, const Pools<R>& P, int mx, int my, int meqn, hipStream_t s)
{
    launch_ghost_ti<R>(S, Segs<TiOp>(), nullptr, nullptr, P, mx, my, meqn, s);
}

template <class R>
void launch_ghost(const GhostOp* ops, int nops, const Pools<R>& P, int mx, int my, int meqn,
                  hipStream_t s)
{
    Segs<GhostOp> S;
    S.add(ops, nops);
    launch_ghost<R>(S, P, mx, my, meqn, s);
}

template <class R>
void launch_bc_cfl(const Segs<BcOp>& S, const Pools<R>& P, int mx, int my, int meqn, const CflFin& cf,
                   hipStream_t s)
{
    const int nb = S.total() + (cf.acc ? 1 : 0);
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_bc_list<R>, dim3(nb), dim3(64), 0, s, S, P, mx, my, meqn, cf);
}

template <class R>
void launch_bc(const Segs<BcOp>& S, const Pools<R>& P, int mx, int my, int meqn, hipStream_t s)
{
    launch_bc_cfl<R>(S, P, mx, my, meqn, CflFin{nullptr, 0, nullptr}, s);
}

template <class R>
void launch_bc(const BcOp* ops, int nops, const Pools<R>& P, int mx, int my, int meqn, hipStream_t s)
{
    Segs<BcOp> S;
    S.add(ops, nops);
    launch_bc<R>(S, P, mx, my, meqn, s);
}

template <class R>
void launch_timeinterp(const TiOp* ops, int nops, const Pools<R>& P, int mx, int my, int meqn,
                       double alpha, hipStream_t s)
{
    if (nops <= 0) return;
    const long long n = (long long)nops * mx * my;
    hipLaunchKernelGGL(k_timeinterp<R>, dim3(nblocks(n, 256)), dim3(256), 0, s, ops, nops, P, mx, my,
                       meqn, R(alpha));
}

template <class R>
void launch_xfer(const XferOp* ops, int nops, const Pools<R>& Old, R* qnew, int mx, int my, int meqn,
                 hipStream_t s)
{
    if (nops <= 0) return;
    const long long n = (long long)nops * (mx + 4) * (my + 4);
    hipLaunchKernelGGL(k_xfer<R>, dim3(nblocks(n, 256)), dim3(256), 0, s, ops, nops, Old, qnew, mx, my,
                       meqn);
}

template <class R>
void launch_tag(const int* slots, const int* levels, int n, const Pools<R>& P, const double* dxdy,
                int mx, int my, int meqn, int ivar, double thr, int mode, int* tags, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_tag_gradient<R>, dim3(n), dim3(256), 0, s, slots, levels, n, P, dxdy, mx, my,
                       meqn, ivar, thr, mode, tags);
}

template <class R>
void launch_tag_stats(const int* slots, const int* levels, int n, const Pools<R>& P, const double* dxdy,
                      int mx, int my, int meqn, int ivar, double* st, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_tag_stats<R>, dim3(n), dim3(256), 0, s, slots, levels, n, P, dxdy, mx, my, meqn, ivar,
                       st);
}

template <class R>
void launch_qinit(int kind, const QinitParams& prm, const int* slots, const int* levels, int n,
                  const Pools<R>& P, const double* xylower, const double* dxdy, int mx, int my,
                  int meqn, hipStream_t s)
{
    if (n <= 0) return;
    const long long tot = (long long)n * (mx + 4) * (my + 4);
    hipLaunchKernelGGL(k_qinit<R>, dim3(nblocks(tot, 256)), dim3(256), 0, s, kind, prm, slots, levels,
                       n, P, xylower, dxdy, mx, my, meqn);
}

template <class R>
void launch_mass(const int* slots, const int* levels, int n, const Pools<R>& P, const double* dxdy,
                 int mx, int my, int meqn, double* out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mass<R>, dim3(n), dim3(256), 0, s, slots, levels, n, P, dxdy, mx, my, meqn, out);
}

template <class R>
void launch_copy_tiles(const int* slots, int n, const R* src, R* dst, long long tile, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy_tiles<R>, dim3(nblocks((long long)n * tile, 256)), dim3(256), 0, s, slots,
                       n, src, dst, tile);
}

#define HC_INST(R)                                                                                \
    template void launch_ghost<R>(const GhostOp*, int, const Pools<R>&, int, int, int, hipStream_t); \
    template void launch_bc<R>(const BcOp*, int, const Pools<R>&, int, int, int, hipStream_t);       \
    template void launch_ghost<R>(const Segs<GhostOp>&, const Pools<R>&, int, int, int, hipStream_t); \
    template void launch_ghost_ti<R>(const Segs<GhostOp>&, const Segs<TiOp>&, const double*, const int*, \
                                     const Pools<R>&, int, int, int, hipStream_t);                     \
    template void launch_bc<R>(const Segs<BcOp>&, const Pools<R>&, int, int, int, hipStream_t);       \
    template void launch_bc_cfl<R>(const Segs<BcOp>&, const Pools<R>&, int, int, int, const CflFin&,     \
                                   hipStream_t);                                                      \
    template void launch_timeinterp<R>(const TiOp*, int, const Pools<R>&, int, int, int, double,     \
                                       hipStream_t);                                              \
    template void launch_xfer<R>(const XferOp*, int, const Pools<R>&, R*, int, int, int, hipStream_t); \
    template void launch_tag<R>(const int*, const int*, int, const Pools<R>&, const double*, int, int, \
                                int, int, double, int, int*, hipStream_t);                        \
    template void launch_tag_stats<R>(const int*, const int*, int, const Pools<R>&, const double*, int, \
                                      int, int, int, double*, hipStream_t);                      \
    template void launch_qinit<R>(int, const QinitParams&, const int*, const int*, int,             \
                                  const Pools<R>&, const double*, const double*, int, int, int,   \
                                  hipStream_t);                                                   \
    template void launch_mass<R>(const int*, const int*, int, const Pools<R>&, const double*, int,  \
                                 int, int, double*, hipStream_t);                                 \
    template void launch_copy_tiles<R>(const int*, int, const R*, R*, long long, hipStream_t);

HC_INST(double)
HC_INST(float)

}  // namespace hc
