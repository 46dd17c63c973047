// hc_timesync.hip -- conservative time synchronisation on the device.
//
// Registers (FP64, one block of 8*meqn*(mx+my) doubles per slot) are filled
// by the patch-step kernel (k_step2 with SYNC).  After a ghost update that
// synchronises levels [minlevel, maxlevel], every patch of those levels with a
// same-size or finer face neighbour is corrected here, then the registers the
// reference clears are zeroed:
//   corrections  clawpack46_fort_time_sync_samesize / _f2c
//                (src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_time_sync.f:254-704)
//   driver order src/fclaw2d_time_sync.c:118-202
//   resets       src/patches/clawpatch/fclaw2d_clawpatch_conservation.c:207-280
//
// A patch's corrections read only registers, which are reset only after all
// corrections of the sync, so all levels are corrected by independent
// workgroups.  Within a patch the order of the reference is kept per edge
// cell: same-size faces 0..3, then finer faces 0..3 (igrid 0 then 1), every
// addend applied in FP64 (a corner cell receives two faces' addends).
//
// In the merged-level driver the corrections run in the ghost fill's final
// launch (k_bc_ts: each workgroup applies its patch's physical BCs first,
// then the corrections), and the clears are left out: the step kernel writes
// the first sums after a clear onto zero instead of reading the register
// (StepArgs::reg_fresh).
#include "hc_amr_launch.hpp"
#include "hc_bc.hpp"
#include "hc_common.hpp"
#include "hc_step2.hpp"

namespace hc {

__device__ inline double ts_delta_same(const double* rt, const double* rn, int mx, int my,
                                       int meqn, int f, int m, int k)
{
#pragma clang fp contract(off)
    const int len = f < 2 ? my : mx;
    auto R = [&](const double* r, int face, int blk, int idx) {
        return r[reg_off(mx, my, meqn, face, blk) + (long long)m * len + idx];
    };
    if (f == 0) return (R(rn, 1, 3, k) - R(rt, 0, 2, k)) - (R(rn, 1, 0, k) - R(rt, 0, 0, k));
    if (f == 1) return (R(rt, 1, 2, k) - R(rn, 0, 3, k)) - (R(rn, 0, 1, k) - R(rt, 1, 1, k));
    if (f == 2) return (R(rn, 3, 3, k) - R(rt, 2, 2, k)) - (R(rn, 3, 0, k) - R(rt, 2, 0, k));
    return (R(rt, 3, 2, k) - R(rn, 2, 3, k)) - (R(rn, 2, 1, k) - R(rt, 3, 1, k));
}

// Coarse edge cell k (0-based along face f) against fine neighbour g: the fine
// pair (a, a+1) = (2k - len*g, 2k - len*g + 1) when it lies in the fine patch.
__device__ inline double ts_delta_f2c(const double* rc, const double* rf, int mx, int my,
                                      int meqn, int f, int g, int m, int k)
{
#pragma clang fp contract(off)
    const int len = f < 2 ? my : mx;
    const int a = 2 * k - len * g;
    if (a < 0 || a + 1 > len - 1) return 0.0;
    auto R = [&](const double* r, int face, int blk, int idx) {
        return r[reg_off(mx, my, meqn, face, blk) + (long long)m * len + idx];
    };
    auto P = [&](int face, int blk) { return R(rf, face, blk, a) + R(rf, face, blk, a + 1); };
    if (f == 0) return (P(1, 3) - R(rc, 0, 2, k)) - (P(1, 0) - R(rc, 0, 0, k));
    if (f == 1) return (R(rc, 1, 2, k) - P(0, 3)) - (P(0, 1) - R(rc, 1, 1, k));
    if (f == 2) return (P(3, 3) - R(rc, 2, 2, k)) - (P(3, 0) - R(rc, 2, 0, k));
    return (R(rc, 3, 2, k) - P(2, 3)) - (P(2, 1) - R(rc, 3, 1, k));
}

// q pool of each list of a launch (one per level)
template <class R> struct QList {
    R* q[SEG_MAX];
};

// Corrections of one patch by one workgroup (the addends of all four faces in
// one pass, then the edge cells).  add: LDS add[f][g][m][k], f < 4, g < 2.
template <class R>
__device__ inline void ts_correct_patch(const TsOp& o, R* __restrict__ q, const double* __restrict__ reg,
                                        long long regstride, const double* __restrict__ dxdy, int mx, int my,
                                        int meqn, double* add)
{
#pragma clang fp contract(off)
    const int maxlen = mx > my ? mx : my;
    const int NX = mx + 4, NY = my + 4;
    const double area = dxdy[2 * o.slot] * dxdy[2 * o.slot + 1];
    const double* rt = reg + (long long)o.slot * regstride;
    auto A = [&](int f, int g, int m, int k) -> double& {
        return add[(((long long)f * 2 + g) * meqn + m) * maxlen + k];
    };
    // phase 1: addends of every face, items (f, m, k) of the four faces in turn
    const int nx_items = meqn * my, ntot = 2 * meqn * (mx + my);
    for (int t = threadIdx.x; t < ntot; t += blockDim.x) {
        int f, u;
        if (t < 2 * nx_items) { f = t < nx_items ? 0 : 1; u = t - f * nx_items; }
        else { const int v = t - 2 * nx_items; f = v < meqn * mx ? 2 : 3; u = v - (f - 2) * meqn * mx; }
        // the face's entries by selects (f differs across the lanes)
        const int kind = f == 0 ? o.kind[0] : f == 1 ? o.kind[1] : f == 2 ? o.kind[2] : o.kind[3];
        if (kind == TS_NONE) continue;
        const int n0 = f == 0 ? o.nbr[0][0] : f == 1 ? o.nbr[1][0] : f == 2 ? o.nbr[2][0] : o.nbr[3][0];
        const int n1 = f == 0 ? o.nbr[0][1] : f == 1 ? o.nbr[1][1] : f == 2 ? o.nbr[2][1] : o.nbr[3][1];
        const int len = f < 2 ? my : mx;
        const int m = u / len, k = u - m * len;
        if (kind == TS_SAME) {
            const double* rn = reg + (long long)n0 * regstride;
            A(f, 0, m, k) = 0.5 * ts_delta_same(rt, rn, mx, my, meqn, f, m, k) / area;
        } else {
            A(f, 0, m, k) = ts_delta_f2c(rt, reg + (long long)n0 * regstride, mx, my, meqn, f, 0, m, k) / area;
            A(f, 1, m, k) = ts_delta_f2c(rt, reg + (long long)n1 * regstride, mx, my, meqn, f, 1, m, k) / area;
        }
    }
    __syncthreads();
    // phase 2: every edge cell applies its addends in the reference order
    const int ncell = 2 * my + 2 * (mx - 2);
    for (int t = threadIdx.x; t < meqn * ncell; t += blockDim.x) {
        const int m = t / ncell, e = t - m * ncell;
        int i, j;  // 0-based interior indices
        if (e < my) { i = 0; j = e; }
        else if (e < 2 * my) { i = mx - 1; j = e - my; }
        else if (e < 2 * my + mx - 2) { i = e - 2 * my + 1; j = 0; }
        else { i = e - 2 * my - (mx - 2) + 1; j = my - 1; }
        const bool on[4] = {i == 0, i == mx - 1, j == 0, j == my - 1};
        const int idx[4] = {j, j, i, i};
        R* qp = q + (long long)o.slot * NX * NY * meqn + ((long long)m * NY + j + 2) * NX + i + 2;
        double v = (double)*qp;
        bool touched = false;
        for (int f = 0; f < 4; f++)
            if (on[f] && o.kind[f] == TS_SAME) { v = v + A(f, 0, m, idx[f]); touched = true; }
        for (int f = 0; f < 4; f++)
            if (on[f] && o.kind[f] == TS_FINE) {
                v = v + A(f, 0, m, idx[f]);
                v = v + A(f, 1, m, idx[f]);
                touched = true;
            }
        if (touched) *qp = (R)v;
    }
}

// One workgroup of TS_THREADS per patch.
constexpr int TS_THREADS = 256;
template <class R>
__global__ void __launch_bounds__(TS_THREADS)
k_ts_correct(const Segs<TsOp> S, const QList<R> qs, const double* __restrict__ reg, long long regstride,
             const double* __restrict__ dxdy, int mx, int my, int meqn)
{
    extern __shared__ double add[];
    const int op = blockIdx.x;
    if (op >= S.total()) return;
    ts_correct_patch<R>(S.at(op), qs.q[S.seg(op)], reg, regstride, dxdy, mx, my, meqn, add);
}

// Physical BCs, then the corrections, of one patch per workgroup.
template <class R>
__global__ void __launch_bounds__(TS_THREADS)
k_bc_ts(const Segs<BcTsOp> S, Pools<R> P, const double* __restrict__ reg, long long regstride,
        const double* __restrict__ dxdy, int mx, int my, int meqn, CflFin cf)
{
    extern __shared__ double add[];
    const int op = blockIdx.x;
    if (op >= S.total()) {
        if (op == S.total() && cf.acc) cfl_final_block(cf);
        return;
    }
    const BcTsOp o = S.at(op);
    const long long T = (long long)(mx + 4) * (my + 4) * meqn;
    if (o.bc.slot >= 0) {
        bc_patch<R>(o.bc, ((o.bc.flags & 1) ? P.ti : P.q[o.level]) + (long long)o.bc.slot * T, mx, my, meqn);
        __syncthreads();
    }
    if (o.ts.slot >= 0) ts_correct_patch<R>(o.ts, P.q[o.level], reg, regstride, dxdy, mx, my, meqn, add);
}

// Zero the four register blocks of listed faces (entry = slot*4 + face).
__global__ void k_ts_reset(const ResetSegs S, double* __restrict__ reg, long long regstride, int mx, int my,
                           int meqn)
{
    const int maxblk = 4 * meqn * (mx > my ? mx : my);
    const long long tot = (long long)S.total() * maxblk;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < tot;
         t += (long long)gridDim.x * blockDim.x) {
        const int e = (int)(t / maxblk), r = (int)(t - (long long)e * maxblk);
        const int fe = S.at(e);
        const int slot = fe >> 2, f = fe & 3;
        const int len = f < 2 ? my : mx;
        if (r < 4 * meqn * len) reg[(long long)slot * regstride + reg_off(mx, my, meqn, f, 0) + r] = 0.0;
    }
}

__global__ void k_zero(double* __restrict__ p, long long n)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        p[t] = 0.0;
}

void launch_zero(double* p, long long n, hipStream_t s)
{
    if (n <= 0) return;
    const long long nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)(nb < 65536 ? nb : 65536)), dim3(256), 0, s, p, n);
}

template <class R>
void launch_ts_correct(const Segs<TsOp>& S, R* const* qs, const double* reg, long long regstride,
                       const double* dxdy, int mx, int my, int meqn, hipStream_t s)
{
    if (S.total() <= 0) return;
    QList<R> ql{};
    for (int k = 0; k < S.n; k++) ql.q[k] = qs[k];
    const size_t lds = sizeof(double) * 8 * meqn * (mx > my ? mx : my);
    hipLaunchKernelGGL(k_ts_correct<R>, dim3(S.total()), dim3(TS_THREADS), lds, s, S, ql, reg, regstride, dxdy,
                       mx, my, meqn);
}

template <class R>
void launch_bc_ts(const Segs<BcTsOp>& S, const Pools<R>& P, const double* reg, long long regstride,
                  const double* dxdy, int mx, int my, int meqn, const CflFin& cf, hipStream_t s)
{
    const int nb = S.total() + (cf.acc ? 1 : 0);
    if (nb <= 0) return;
    const size_t lds = sizeof(double) * 8 * meqn * (mx > my ? mx : my);
    hipLaunchKernelGGL(k_bc_ts<R>, dim3(nb), dim3(TS_THREADS), lds, s, S, P, reg, regstride, dxdy, mx, my,
                       meqn, cf);
}

template <class R>
void launch_ts_correct(const TsOp* ops, int nops, R* q, const double* reg, long long regstride,
                       const double* dxdy, int mx, int my, int meqn, hipStream_t s)
{
    Segs<TsOp> S;
    S.add(ops, nops);
    launch_ts_correct<R>(S, &q, reg, regstride, dxdy, mx, my, meqn, s);
}

void launch_ts_reset(const ResetSegs& S, double* reg, long long regstride, int mx, int my, int meqn,
                     hipStream_t s)
{
    if (S.total() <= 0) return;
    const long long tot = (long long)S.total() * 4 * meqn * (mx > my ? mx : my);
    const long long nb = (tot + 255) / 256;
    hipLaunchKernelGGL(k_ts_reset, dim3((unsigned)(nb < 65536 ? nb : 65536)), dim3(256), 0, s, S, reg,
                       regstride, mx, my, meqn);
}

void launch_ts_reset(const int* faces, int n, double* reg, long long regstride, int mx, int my,
                     int meqn, hipStream_t s)
{
    ResetSegs S;
    S.add(faces, n);
    launch_ts_reset(S, reg, regstride, mx, my, meqn, s);
}

template void launch_bc_ts<double>(const Segs<BcTsOp>&, const Pools<double>&, const double*, long long,
                                   const double*, int, int, int, const CflFin&, hipStream_t);
template void launch_bc_ts<float>(const Segs<BcTsOp>&, const Pools<float>&, const double*, long long,
                                  const double*, int, int, int, const CflFin&, hipStream_t);
template void launch_ts_correct<double>(const Segs<TsOp>&, double* const*, const double*, long long,
                                        const double*, int, int, int, hipStream_t);
template void launch_ts_correct<float>(const Segs<TsOp>&, float* const*, const double*, long long,
                                       const double*, int, int, int, hipStream_t);
template void launch_ts_correct<double>(const TsOp*, int, double*, const double*, long long,
                                        const double*, int, int, int, hipStream_t);
template void launch_ts_correct<float>(const TsOp*, int, float*, const double*, long long,
                                       const double*, int, int, int, hipStream_t);

}  // namespace hc
