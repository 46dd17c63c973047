// hc_step2.hpp -- fused wave-propagation patch step for gfx950.
//
// Computes, for every patch of a batch, what clawpack46_step2_wrap computes
// for the patch interior (src/solvers/fc2d_clawpack4.6/fortran_source/
// clawpack46_step2_wrap.f:104-125, clawpack46_step2.f:113-236,
// clawpack46_flux2.f:116-224, clawpack46_inlinelimiter.f:39-115), plus the
// patch Courant number (clawpack46_flux2.f:129-137, clawpack46_step2.f:153,222).
//
// Mapping.  A patch tile of NX = mx+4 columns is owned by LPP = ceil(NX/CPL)
// consecutive lanes of one wave; lane lp owns columns lp*CPL .. lp*CPL+CPL-1
// and walks the tile rows J = 0 .. NY-1 once.  At row J the lane
//   * solves the x-face to the left of each of its columns on row J
//     (neighbour columns arrive by cross-lane shuffles),
//   * solves the y-face below (c, J) of its own columns,
//   * finalises y-face J-1 (its limiter needed face J), and
//   * writes the updated interior cell (c, J-2).
// The whole x/y sweep therefore reads each tile row from HBM once, keeps the
// sliding windows in registers and uses no LDS.  Every flux sum is assembled
// in one lane in the order the reference accumulates fm/fp/gm/gp, so the FP64
// instantiation (no FMA contraction) is bitwise equal to the reference on the
// interior.  Only interior cells are written; the ghost band of q_out is left
// untouched (the reference's post-step ghost values are overwritten by the
// next ghost fill in any case).
//
// mbc must be 2 (ForestClaw's default); the host rejects other values.
#pragma once
#include <type_traits>

#include "hc_riemann.hpp"

namespace hc {

constexpr int STEP_MAXL = 22;  // levels 0..21, as the AMR driver

struct StepArgs {
    int npatch;
    int mx, my;
    int order2, transverse;
    int mthlim[5];
    int lpp, ppw;          // lanes per patch, patches per wave
    PhysParams pp;
    double dt;
    const double* dxdy;    // per patch (dx, dy)
    const int* patch_ids;  // optional indirection (nullptr = identity)
    int ntiles;            // tiles addressable in q / aux / dxdy / cflp
    long long qstride;     // elements per patch in q (NX*NY*meqn)
    long long astride;     // elements per patch in aux (NX*NY*maux)
    double* reg;           // time-sync registers (SYNC kernels only), FP64
    long long regstride;   // doubles per patch in reg = 8*meqn*(mx+my)
    // Multi-level launch (levels != nullptr): tile p of the launch is slot
    // slot0 + p of the slot-indexed arrays (dxdy, aux, cfl, reg and the q
    // pools), and its level levels[slot] selects the q pool pair and dt.
    const int* levels;
    int slot0;
    double dtl[STEP_MAXL];
    const void* qin_l[STEP_MAXL];
    void* qout_l[STEP_MAXL];
    // PUSH kernels: per slot, the same-level neighbours across faces 0..3 and
    // corners 0..3 (-1: none); updated edge cells are also written into their
    // ghost cells (the same-level copy of the next ghost update).
    const int* push;
    // Advection with the swirl velocity field evaluated in the kernel
    // (swirl != 0) instead of read from aux: the velocities b4step2 would
    // store, from the patch corner (xylower, per slot / tile like dxdy) and
    // vt = cos(2 pi (t + dt/2) / T) of the step (vtl[level] in multi-level
    // launches).
    int swirl;
    const double* xylower;
    double vt;
    double vtl[STEP_MAXL];
    // Multi-level launches replayed from a hipGraph: dt and vt of level L
    // read from the device at par[L] and par[STEP_MAXL + L] (written before
    // each replay) instead of dtl / vtl, so one graph serves every dt.
    const double* par;
    // optional launch maximum of the patch Courant numbers, kept as the bits
    // of a non-negative double (integer order = double order) by atomic max
    // of every positive value (one per wave, or per patch with row bands);
    // the caller clears it beforehand
    unsigned long long* cflmax;
    // SYNC kernels of a merged subcycled step (reg_fresh != 0): the registers
    // are zero on entry, without having been cleared, except those of the
    // faces of a level-reg_lo patch that have a coarser neighbour (bit f of
    // reg_cfaces[slot]), which hold the patch's earlier substeps; only those
    // are read, the others are written as sums onto zero.  In such a step
    // every level's own faces are cleared by the sync that ends each of its
    // steps, and the coarse-fine faces of level L by the sync of level L-1,
    // which precedes L's first substep; reg_lo is the coarsest level of the
    // launch, whose coarser neighbour level is between steps.
    int reg_fresh, reg_lo;
    const int* reg_cfaces;
    // Row bands (nband > 1, q_out != q_in): a workgroup of
    // nband waves takes the same ppw patches, wave b the interior rows
    // [2 + b*band_rows, 2 + (b+1)*band_rows) (clipped to my), sweeping from
    // at least two rows below to two rows above them; the dependence of an
    // interior row on the data reaches two rows each way, as the ghost width
    // does, so every band computes its rows exactly as a whole-patch sweep
    // would.  The bands of a patch stage its edge cells for the ghost push
    // in one shared area and push them together.
    int nband, band_rows;
};

// Time-sync register blocks of one face: P (fp/gp), M (fm/gm), E0/E1 (edge
// flux at the interior / ghost cell), each meqn x len with len fastest, faces
// 0 left, 1 right, 2 bottom, 3 top (the clawpatch registers fp[], fm[], gp[],
// gm[], edge_fluxes[] of src/patches/clawpatch/fclaw2d_clawpatch_conservation.c:80-110).
__device__ inline long long reg_off(int mx, int my, int meqn, int face, int block)
{
    long long off = 0;
    for (int f = 0; f < face; f++) off += 4LL * meqn * (f < 2 ? my : mx);
    return off + (long long)block * meqn * (face < 2 ? my : mx);
}

// r + s * v and r - s * v in FP64 without contraction, as
// clawpack46_time_sync.f:37-156 and :163-226 accumulate.
__device__ inline double reg_plus(double r, double s, double v)
{
#pragma clang fp contract(off)
    return r + s * v;
}
__device__ inline double reg_minus(double r, double s, double v)
{
#pragma clang fp contract(off)
    return r - s * v;
}

// 0 + x and 0 - x: the reference accumulates fluxes into zeroed arrays, and
// FP64 keeps those additions (they decide the sign of a zero result); FP32
// takes x and -x, which drops an instruction each.
template <class R> __device__ inline R zadd(R x)
{
    if constexpr (sizeof(R) == 4) return x;
    else return R(0) + x;
}
template <class R> __device__ inline R zsub(R x)
{
    if constexpr (sizeof(R) == 4) return -x;
    else return R(0) - x;
}

// Block b of n, dealt to XCD b mod 8, renumbered so that XCD x takes the
// contiguous range of the blocks it runs: a bijection of [0, n).
#ifndef HC_STEP_XCD_ORDER
#define HC_STEP_XCD_ORDER 1
#endif
constexpr int HC_NXCD = 8;
__device__ inline int xcd_block(int b, int n)
{
    const int q = n / HC_NXCD, r = n % HC_NXCD;
    const int x = b % HC_NXCD, i = b / HC_NXCD;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// One-component limiter ratio (FP32): limited when w^2 would be normal, and
// the reciprocal of the wave.  The FP64 forms are never instantiated.
template <class R> __device__ inline bool lim_ok1(R w) { return M<R>::abs_(w) >= R(1.0842022e-19f); }
template <class R> __device__ inline R rcp1(R w)
{
    if constexpr (sizeof(R) == 4) return __builtin_amdgcn_rcpf(w);
    else return R(1) / w;
}

template <int CPL, class T> __device__ inline T from_left(const T (&v)[CPL], int k)
{
    T s = shl_up(v[CPL - 1]);
    return k > 0 ? v[k > 0 ? k - 1 : 0] : s;
}
template <int CPL, class T> __device__ inline T from_right(const T (&v)[CPL], int k)
{
    T s = shl_dn(v[0]);
    return k < CPL - 1 ? v[k < CPL - 1 ? k + 1 : 0] : s;
}

// MC22: compile-time specialisation for the configuration every shipped
// application uses (order 2, transverse 2, MC limiter on every wave family);
// the generic instantiation reads these from the arguments.
// SYNC: also accumulate the conservative time-sync registers (solvers with a
// conservative flux function only): store_flux on the data before the update
// and accumulate_waves on the final face fluctuations
// (src/solvers/fc2d_clawpack4.6/fc2d_clawpack46.cpp:317-341, 376-390).
// PUSH: also write every updated cell within two cells of a patch edge into
// the ghost cells of the same-level neighbour across that face or corner
// (what fclaw2d_clawpatch46_fort_copy_face / _copy_corner would copy from it,
// src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_copy.f:17-133).
// Row-loop shape.  UNR rows per loop iteration (4: the unrolled, renamed
// windows; 1: rotate after every row); the time-sync and the FP64
// two-column kernels keep 1, whose unrolled bodies would not fit their
// register budget.  WPE is the occupancy floor (waves per SIMD) the kernel is
// register-allocated for: the FP32 kernels and the FP64 one-column kernels
// keep two waves (at most 256 VGPRs).  HC_STEP_UNROLL / HC_STEP_MIN_WAVES
// override the defaults.
#ifndef HC_STEP_UNROLL
#define HC_STEP_UNROLL 4
#endif
#ifndef HC_STEP_MIN_WAVES
#define HC_STEP_MIN_WAVES 2
#endif
template <class R, int CPL, bool SYNC>
constexpr int STEP_UNR = ((SYNC && sizeof(R) == 8) || (sizeof(R) == 8 && CPL == 2)) ? 1 : HC_STEP_UNROLL;
// HC_STEP_PREFETCH: the unrolled kernels load tile row J+1 while row J is
// computed (one row of q in flight per lane) instead of at the start of row J.
#ifndef HC_STEP_PREFETCH
#define HC_STEP_PREFETCH 1
#endif
template <class R, int CPL, bool SYNC>
constexpr int STEP_WPE = (!SYNC && (sizeof(R) == 4 || CPL == 1)) ? HC_STEP_MIN_WAVES : 1;

template <class R, template <class> class SolverT, int CPL, bool MC22, bool SYNC, bool PUSH,
          bool BAND = false, int UNR = STEP_UNR<R, CPL, SYNC>>
__global__ void __launch_bounds__(BAND ? 512 : 256) __attribute__((amdgpu_waves_per_eu(STEP_WPE<R, CPL, SYNC>)))
k_step2(const StepArgs A, const R* __restrict__ qin, R* __restrict__ qout,
        const R* __restrict__ aux, double* __restrict__ cflp)
{
    using S = SolverT<R>;
    constexpr int ME = S::ME, MW = S::MW;
    constexpr bool PF = UNR == 4 && HC_STEP_PREFETCH;
    using Face = typename S::Face;
    using Trans = typename S::Trans;
    // FP32 sums over a wave's components skip those that are zero by
    // construction (FP64 keeps every term of the reference's sums)
    constexpr bool SKIPZ = sizeof(R) == 4;
    // FP32 one-equation solvers: the limiter ratio <w_up, w> / <w, w> of a
    // one-component wave is w_up / w, taken as such (the neighbouring faces'
    // waves are carried instead of the dot products); a wave is limited when
    // |w| >= sqrt(FLT_MIN), i.e. when w^2 is normal, as in the general form
    constexpr bool RAT = sizeof(R) == 4 && ME == 1 && MW == 1;

    const int lane = threadIdx.x & (HC_WAVE - 1);
    // XCD-contiguous block order: the hardware deals workgroups to the 8 XCDs
    // in turn, so block b runs on XCD b mod 8; renumbering the blocks so that
    // each XCD takes one contiguous range of patches (slots are in Morton
    // order) keeps most face neighbours in one L2, where a patch's interior
    // stores and its neighbours' ghost pushes into the same lines combine
    const int blk = HC_STEP_XCD_ORDER ? xcd_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;
    const int gwave = (int)(((unsigned)blk * (unsigned)blockDim.x + threadIdx.x) >> 6);
    const int slot = lane / A.lpp;
    const int lp = lane - slot * A.lpp;
    // wave-uniform (scalar) band index: the row loop and its conditions stay
    // scalar when the band bounds derive from it
    const int band = BAND ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
    const int pidx = (BAND ? blk : gwave) * A.ppw + slot;
    const bool pin = (slot < A.ppw) && (pidx < A.npatch);
    const int praw = pin ? (A.patch_ids ? A.patch_ids[pidx] : pidx) : 0;
    // a list entry outside the tile range is skipped, never dereferenced
    const bool pact = pin && praw >= 0 && praw < A.ntiles;
    const int patch = pact ? praw : 0;
    const int NX = A.mx + 4, NY = A.my + 4;
    // rows of this wave: interior rows [ra, rb) are written, the sweep runs
    // over rows J0 .. Jlast (the whole tile without bands)
    const int ra = BAND ? 2 + band * A.band_rows : 2;
    const int rb = BAND ? min(2 + (band + 1) * A.band_rows, NY - 2) : NY - 2;
    const bool first_band = !BAND || band == 0;
    const bool last_band = !BAND || band == A.nband - 1;
    // a band's sweep starts on a multiple of the unroll factor, so every row
    // runs through the unrolled copy (and its FP32 contractions) that a
    // whole-patch sweep runs it through
    const int J0 = first_band ? 0 : ((ra - 2) & ~(UNR - 1));
    const int Jlast = last_band ? NY - 1 : rb + 1;

    // tile index into the tile-indexed arrays; multi-level launches address
    // slots and take q pools and dt from the slot's level
    const int tix = A.levels ? A.slot0 + patch : patch;
    const int lev = (A.levels && pact) ? A.levels[tix] : 0;
    const R* qbase = A.levels ? static_cast<const R*>(A.qin_l[lev]) : qin;
    R* obase = A.levels ? static_cast<R*>(A.qout_l[lev]) : qout;
    const double dtp = A.levels ? (A.par ? A.par[lev] : A.dtl[lev]) : A.dt;
    const R* __restrict__ qp = qbase + (long long)tix * A.qstride;
    R* __restrict__ qo = obase + (long long)tix * A.qstride;
    const R* __restrict__ ap = (S::MAUX > 0) ? aux + (long long)tix * A.astride : nullptr;
    // SYNC: faces whose registers are read (bit f), the others being zero
    int reg_rd = 15;
    if constexpr (SYNC) {
        if (A.reg_fresh) reg_rd = (pact && lev == A.reg_lo) ? A.reg_cfaces[tix] : 0;
    }

    // PUSH: this patch's edge cells are staged in LDS during the sweep, per
    // patch 4*meqn*(mx+my) values: x edges [side][col][m][row], then y edges
    // [side][row][m][col] (sides: low, high; two cells deep); the row bands
    // of a patch share its area
    extern __shared__ char hc_smem[];
    const int pstage = 4 * ME * (A.mx + A.my);
    const int swave = BAND ? 0 : (int)(threadIdx.x >> 6);  // staging wave
    const int nstage = BAND ? 1 : (int)(blockDim.x >> 6);   // staging waves per group
    R* stage = reinterpret_cast<R*>(hc_smem) + (swave * A.ppw + (slot < A.ppw ? slot : 0)) * pstage;
    const double dx = pact ? A.dxdy[2 * tix] : 1.0;
    const double dy = pact ? A.dxdy[2 * tix + 1] : 1.0;
    const R dtdx = R(dtp / dx);
    const R dtdy = R(dtp / dy);
    const R hdtdx = R(0.5) * dtdx;
    const R hdtdy = R(0.5) * dtdy;
    const PhysParams pp = A.pp;
    const bool second = MC22 || A.order2 == 2;
    const bool trans = MC22 || A.transverse > 0;
    const bool cqsplit = MC22 || ((A.order2 > 1) && (A.transverse == 2));
    int mth[MW];
#pragma unroll
    for (int w = 0; w < MW; w++) mth[w] = MC22 ? 4 : A.mthlim[w];

    int col[CPL];
    bool cv[CPL];
#pragma unroll
    for (int k = 0; k < CPL; k++) {
        col[k] = lp * CPL + k;
        cv[k] = pact && col[k] < NX;
    }

    // ---- carried state -------------------------------------------------
    // The sliding windows are rings indexed by the row number counted from
    // the first row J0 of the sweep: a value made at row J lives in slot
    // (J - J0) mod 4 (windows three rows deep: q, u, the x-fluxes and G) or
    // (J - J0) mod 2 (two rows deep: the y-face, its transverse
    // splits and limiter dot products, d2).  With UNR = 4 the row loop is
    // unrolled four times and every slot index is a compile-time constant,
    // so the windows never move between registers; UNR = 1 keeps one row per
    // iteration and rotates the rings after it.
    R qr[4][CPL][ME], ur[4][CPL];
    // advection: v of row J+1 (and of the left neighbour column), made by
    // row J and taken by row J+1 as its own (slot (J - J0) mod 2)
    R vnr[2][CPL], vlr[2][CPL];
    // FP32 swirl: the LDS row factors a row takes (dB of its own row for u,
    // B of the next row for v), read by the row before it (same ring slots),
    // so that no row waits for an LDS read it has just issued
    float ydb[2] = {0.0f, 0.0f}, yb1[2] = {0.0f, 0.0f};
    R fmr[4][CPL][ME], fpr[4][CPL][ME], Gr[4][CPL][ME];
    Face fyr[2][CPL] = {};
    Trans tyr[2][CPL] = {};
    R dotr2[2][CPL][MW];
    R hmpr[2][CPL][ME], hppr[2][CPL][ME], fdpr[2][CPL][ME], d2r[2][CPL][ME];
    R cfl[CPL];
#pragma unroll
    for (int k = 0; k < CPL; k++) {
        cfl[k] = R(0);
#pragma unroll
        for (int s = 0; s < 4; s++) {
            ur[s][k] = R(0);
            if (s < 2) vnr[s][k] = vlr[s][k] = R(0);
#pragma unroll
            for (int m = 0; m < ME; m++) qr[s][k][m] = fmr[s][k][m] = fpr[s][k][m] = Gr[s][k][m] = R(0);
        }
#pragma unroll
        for (int s = 0; s < 2; s++) {
#pragma unroll
            for (int m = 0; m < ME; m++) hmpr[s][k][m] = hppr[s][k][m] = fdpr[s][k][m] = d2r[s][k][m] = R(0);
#pragma unroll
            for (int w = 0; w < MW; w++) dotr2[s][k][w] = R(0);
        }
    }

    // ---- advection: the swirl velocity field in the kernel ---------------
    // psi(x, y) = sin^2(pi x) sin^2(pi y) / pi at the corners (xll, yll),
    // (xll, yll+dy), (xll+dx, yll) of every cell, with the operations of the
    // velocity-field kernel (k_swirl_aux), so the velocities are the ones it
    // would store: the column sines are kept per lane, the row sines of each
    // patch are evaluated once into LDS (after the push staging area).
    double swx0[CPL], swx1[CPL];
    const double* swy = nullptr;
    double swvtx = 0.0, swvty = 0.0;
    const double swpi = 4.0 * atan(1.0);
    // FP32: psi = A(x) B(y) / pi is separable.  Per column A = sin^2(pi xll)
    // and dA = A(xll + dx) - A, per row (in LDS) B and dB likewise, and the
    // velocities are u = vt/(pi dy) A dB, v = -vt/(pi dx) B dA: two
    // multiplications each instead of the FP64 differences of psi, which
    // FP64 keeps so that its velocities equal the stored field bit for bit.
    float swa0[CPL], swda[CPL];
    const float* swyf = nullptr;
    float swcu = 0.0f, swcv = 0.0f;
    if constexpr (S::MAUX > 0 && !S::FW) {
        if (A.swirl) {
            const double x0 = pact ? A.xylower[2 * tix] : 0.0;
            const double y0 = pact ? A.xylower[2 * tix + 1] : 0.0;
            if constexpr (sizeof(R) == 8) {
#pragma unroll
                for (int k = 0; k < CPL; k++) {
                    const double xll = x0 + (double)(col[k] - 2) * dx;
                    swx0[k] = sin(swpi * xll);
                    swx1[k] = sin(swpi * (xll + dx));
                }
            }
            const size_t stage_bytes = PUSH ? sizeof(R) * (size_t)nstage * A.ppw * pstage : 0;
            double* ys = reinterpret_cast<double*>(hc_smem + ((stage_bytes + 7) & ~(size_t)7)) +
                         ((threadIdx.x >> 6) * A.ppw + (slot < A.ppw ? slot : 0)) * 2 * NY;
            const double vt = A.levels ? (A.par ? A.par[STEP_MAXL + lev] : A.vtl[lev]) : A.vt;
            if constexpr (sizeof(R) == 4) {
                // sin(pi z) by the hardware sine, which takes revolutions
                // (v_sin_f32(z/2)); the corner coordinates are dyadic, so
                // their halves are exact in FP32
                auto sinpi = [](double z) { return __builtin_amdgcn_sinf((float)(0.5 * z)); };
                float* yf = reinterpret_cast<float*>(ys);
                if (slot < A.ppw) {
                    for (int j = lp; j < NY; j += A.lpp) {
                        const double yll = y0 + (double)(j - 2) * dy;
                        const float s0 = sinpi(yll), s1 = sinpi(yll + dy);
                        yf[j] = s0 * s0;
                        yf[NY + j] = s1 * s1 - s0 * s0;
                    }
                }
#pragma unroll
                for (int k = 0; k < CPL; k++) {
                    const double xll = x0 + (double)(col[k] - 2) * dx;
                    const float s0 = sinpi(xll), s1 = sinpi(xll + dx);
                    swa0[k] = s0 * s0;
                    swda[k] = s1 * s1 - s0 * s0;
                }
                swyf = yf;
                swcu = (float)(vt / (swpi * dy));
                swcv = (float)(-vt / (swpi * dx));
            } else {
                if (slot < A.ppw) {
                    for (int t = lp; t < 2 * NY; t += A.lpp) {
                        const int j = t < NY ? t : t - NY;
                        const double yll = y0 + (double)(j - 2) * dy;
                        ys[t] = sin(swpi * (t < NY ? yll : yll + dy));
                    }
                }
                swy = ys;
                swvty = vt / dy;
                swvtx = vt / dx;
            }
            __syncthreads();
        }
    }

    // advection: v at the bottom edge of (col[k], r), 0 for lanes without a
    // cell; as the velocity-field kernel evaluates it (or as aux holds it)
    auto vel_v = [&](int k, int r) __attribute__((always_inline)) -> R {
        if constexpr (S::MAUX > 0 && !S::FW) {
            if (A.swirl && sizeof(R) == 4) {
                return cv[k] ? R(swcv * (swyf[r] * swda[k])) : R(0);
            } else if (A.swirl) {
                const double a0 = swx0[k] * swx0[k], a1 = swx1[k] * swx1[k];
                const double b0 = swy[r] * swy[r];
                const double pij = (a0 * b0) / swpi;
                const double v = -swvtx * ((a1 * b0) / swpi - pij);
                return cv[k] ? R(v) : R(0);
            } else {
                return cv[k] ? ap[((long long)NY + r) * NX + col[k]] : R(0);
            }
        }
        return R(0);
    };

    auto row = [&](auto phase, const int J) __attribute__((always_inline)) {
        constexpr int P = decltype(phase)::value;
        constexpr int S0 = P & 3, S1 = (P + 3) & 3, S2 = (P + 2) & 3;  // rows J, J-1, J-2
        constexpr int T0 = P & 1, T1 = (P + 1) & 1;                     // rows J, J-1
        R(&qc)[CPL][ME] = qr[S0];
        R(&qm1)[CPL][ME] = qr[S1];
        R(&qm2)[CPL][ME] = qr[S2];
        R(&ucur)[CPL] = ur[S0];
        R(&um1)[CPL] = ur[S1];  // advection: u(c, J-1), u(c, J-2)
        R(&um2)[CPL] = ur[S2];
        R(&fmxn_w)[CPL][ME] = fmr[S0];
        R(&fmx1)[CPL][ME] = fmr[S1];
        R(&fmx2)[CPL][ME] = fmr[S2];
        R(&fpxn_w)[CPL][ME] = fpr[S0];
        R(&fpx1)[CPL][ME] = fpr[S1];
        R(&fpx2)[CPL][ME] = fpr[S2];
        R(&Gn_w)[CPL][ME] = Gr[S0];
        R(&G1)[CPL][ME] = Gr[S1];
        R(&G2)[CPL][ME] = Gr[S2];
        R(&d2n_w)[CPL][ME] = d2r[T0];
        R(&d2p)[CPL][ME] = d2r[T1];
        R(&hmp_new_w)[CPL][ME] = hmpr[T0];
        R(&hmp_prev)[CPL][ME] = hmpr[T1];
        R(&hpp_new_w)[CPL][ME] = hppr[T0];
        R(&hpp_prev)[CPL][ME] = hppr[T1];
        R(&fdp_w)[CPL][ME] = fdpr[T0];
        R(&fdp_prev)[CPL][ME] = fdpr[T1];
        R(&dotry_w)[CPL][MW] = dotr2[T0];
        R(&dotly)[CPL][MW] = dotr2[T1];
        Face(&fyn_w)[CPL] = fyr[T0];
        Face(&fyp)[CPL] = fyr[T1];
        Trans(&tyn_w)[CPL] = tyr[T0];
        Trans(&typ)[CPL] = tyr[T1];

        // ---- load row J (rows past the tile, run by the unrolled loop when
        // NY is not a multiple of 4, load zeros and store nothing) ----------
        const bool rowin = J <= Jlast;
        // row within the sweep: a band starts as a whole-patch sweep starts
        // (no face below its first row, its first row not x-swept), so the
        // rows it writes see the operations of a whole-patch sweep
        const int Jr = J - J0;
        // row offsets through an opaque copy of J: addresses are formed per
        // row from the per-lane tile base instead of one pointer induction
        // variable per unrolled row and field kept live across the loop
        const int Jo = __builtin_amdgcn_readfirstlane(J);
        R vcur[CPL], vnext[CPL];
#pragma unroll
        for (int k = 0; k < CPL; k++) {
            if constexpr (!PF) {
#pragma unroll
                for (int m = 0; m < ME; m++)
                    qc[k][m] = (cv[k] && rowin) ? qp[((long long)m * NY + Jo) * NX + col[k]] : R(0);
            }
            if constexpr (S::MAUX > 0 && !S::FW) {
                // u of (c, J) and v of (c, J+1); v of (c, J) is what row J-1
                // made as its v of the next row (or the prologue, for J0)
                if (A.swirl && sizeof(R) == 4) {
                    const float db = rowin ? ydb[T1] : 0.0f;
                    ucur[k] = (cv[k] && rowin) ? R(swcu * (swa0[k] * db)) : R(0);
                } else if (A.swirl) {
                    ucur[k] = R(0);
                    if (rowin) {
                        const double a0 = swx0[k] * swx0[k];
                        const double b0 = swy[Jo] * swy[Jo], b1 = swy[NY + Jo] * swy[NY + Jo];
                        const double pij = (a0 * b0) / swpi;
                        const double u = swvty * ((a0 * b1) / swpi - pij);
                        ucur[k] = cv[k] ? R(u) : R(0);
                    }
                } else {
                    ucur[k] = (cv[k] && rowin) ? ap[(long long)Jo * NX + col[k]] : R(0);
                }
                vcur[k] = rowin ? vnr[T1][k] : R(0);
                if (A.swirl && sizeof(R) == 4)
                    vnext[k] = (J + 1 < NY && cv[k]) ? R(swcv * (yb1[T1] * swda[k])) : R(0);
                else
                    vnext[k] = J + 1 < NY ? vel_v(k, Jo + 1) : R(0);
            } else {
                ucur[k] = vcur[k] = vnext[k] = R(0);
            }
        }
        R vl_next[CPL];
#pragma unroll
        for (int k = 0; k < CPL; k++) vl_next[k] = (S::MAUX > 0 && !S::FW) ? from_left<CPL>(vnext, k) : R(0);

        // f-wave transport: edge data of row J (left/right/bottom/top edge
        // velocities, x/y edge lengths, aux components 1..6), of row J+1 (the
        // x-faces' transverse splits: bottom/top velocities, y length) and of
        // row J-1 (the y-face below row J: top and right velocities, x length)
        R fuL[CPL], fuR[CPL], fvB[CPL], fvT[CPL], fex[CPL], fey[CPL];
        R fvBn[CPL], fvTn[CPL], feyn[CPL], fvTm[CPL], fuRm[CPL], fexm[CPL];
        if constexpr (S::FW) {
#pragma unroll
            for (int k = 0; k < CPL; k++) {
                auto ld = [&](bool in, int comp, int r) {
                    return in ? ap[((long long)comp * NY + r) * NX + col[k]] : R(0);
                };
                const bool i0 = cv[k] && rowin, ip = i0 && J + 1 < NY, im = i0 && J >= 1;
                fuL[k] = ld(i0, 1, Jo);
                fuR[k] = ld(i0, 2, Jo);
                fvB[k] = ld(i0, 3, Jo);
                fvT[k] = ld(i0, 4, Jo);
                fex[k] = ld(i0, 5, Jo);
                fey[k] = ld(i0, 6, Jo);
                fvBn[k] = ld(ip, 3, Jo + 1);
                fvTn[k] = ld(ip, 4, Jo + 1);
                feyn[k] = ld(ip, 6, Jo + 1);
                fvTm[k] = ld(im, 4, Jo - 1);
                fuRm[k] = ld(im, 2, Jo - 1);
                fexm[k] = ld(im, 5, Jo - 1);
            }
        }

        if constexpr (PF) {
            // row J was loaded during row J-1 (or before the loop); row J+1
            // goes into the slot of row J-3, which no row from J on reads, so
            // its load latency is hidden behind this row's arithmetic
            // (loaded unconditionally and zeroed by a select: without
            // branches around the loads the wait for row J counts exactly the
            // loads issued after it).  Issued after the row's velocity and
            // edge-data loads, so the waits for those leave it in flight.  Every address is valid: a lane without
            // a patch reads tile 0 of the launch (patch 0: slot slot0 of a
            // merged launch, whose q pool pointer step2_launch sets for every
            // level, or tile 0 of a batch, which it requires to exist); rows
            // past the sweep read row Jlast (<= NY-1) of the tile; lanes past
            // the tile's columns read column 0.
            constexpr int SN = (P + 1) & 3;
            const bool nin = J + 1 <= Jlast;
            const int Jn = nin ? Jo + 1 : Jlast;
#pragma unroll
            for (int k = 0; k < CPL; k++) {
                const int cc = cv[k] ? col[k] : 0;
#pragma unroll
                for (int m = 0; m < ME; m++) {
                    const R v = qp[((long long)m * NY + Jn) * NX + cc];
                    qr[SN][k][m] = (cv[k] && nin) ? v : R(0);
                }
            }
        }
        // ---- per-cell solver data of row J (shared by its four faces) -------
        constexpr int NCD = S::NC > 0 ? S::NC : 1;
        R cdc[CPL][NCD];
#pragma unroll
        for (int k = 0; k < CPL; k++) {
            if constexpr (S::NC > 0) S::cell(qc[k], pp, cdc[k]);
            else cdc[k][0] = R(0);
        }

        // ---- time-sync registers of this row ------------------------------
        // Every register element takes one addend per step (FP64, old + s*v or
        // old - s*v, uncontracted, as clawpack46_time_sync.f accumulates).  The
        // elements this row updates are loaded here, ahead of the row's
        // arithmetic, and written where their addends are made:
        //   0  x-face edge flux of row J           lanes 1, 2, NX-3, NX-2
        //   1  x-face fluctuation of row J-2       the same lanes
        //   2  y-face edge flux of rows 1, 2, NY-3, NY-2
        //   3  y-face fluctuation of rows 2, NY-3  (row J-2)
        //   4  y-face fluctuation of ghost rows 1, NY-2 (rows J = 3, NY-1)
        // Slots 0 and 2 take their addends from the row's data (the edge
        // fluxes before the update) and are written at the end of the row.
        int tso[CPL][5];
        double tsold[CPL][5][ME];
        double tsv0[CPL][ME], tsv2[CPL][ME];
        double* rg = nullptr;
        double dtel_x = 0.0, dtel_y = 0.0;
        if constexpr (SYNC) {
            rg = A.reg + (long long)tix * A.regstride;
            dtel_y = dtp * dy;
            dtel_x = dtp * dx;
            const int Rw2 = J - 2;
#pragma unroll
            for (int k = 0; k < CPL; k++) {
                const int c = col[k];
                const bool xe = cv[k] && (c == 1 || c == 2 || c == NX - 3 || c == NX - 2);
                const bool yi = cv[k] && c >= 2 && c <= NX - 3;
                const int xf = c <= 2 ? 0 : 1;
                tso[k][0] = (xe && J >= ra && J < rb)
                                ? (int)reg_off(A.mx, A.my, ME, xf, (c == 2 || c == NX - 3) ? 2 : 3) + (J - 2)
                                : -1;
                tso[k][1] = (xe && Rw2 >= ra && Rw2 < rb)
                                ? (int)reg_off(A.mx, A.my, ME, xf, (c == 1 || c == NX - 3) ? 1 : 0) + (Rw2 - 2)
                                : -1;
                tso[k][2] = (yi && (((J == 1 || J == 2) && first_band) || ((J == NY - 3 || J == NY - 2) && last_band)))
                                ? (int)reg_off(A.mx, A.my, ME, J <= 2 ? 2 : 3, (J == 2 || J == NY - 3) ? 2 : 3) + (c - 2)
                                : -1;
                tso[k][3] = (yi && Rw2 >= ra && Rw2 < rb && (Rw2 == 2 || Rw2 == NY - 3))
                                ? (int)reg_off(A.mx, A.my, ME, Rw2 == 2 ? 2 : 3, Rw2 == 2 ? 0 : 1) + (c - 2)
                                : -1;
                tso[k][4] = (yi && ((J == 3 && first_band) || (J == NY - 1 && last_band)))
                                ? (int)reg_off(A.mx, A.my, ME, J == 3 ? 2 : 3, J == 3 ? 1 : 0) + (c - 2)
                                : -1;
                const int fsl[5] = {xf, xf, J <= 2 ? 2 : 3, Rw2 == 2 ? 2 : 3, J == 3 ? 2 : 3};
#pragma unroll
                for (int sl = 0; sl < 5; sl++) {
                    const int st = sl < 2 ? A.my : A.mx;  // element stride of the field m
                    const bool rd = tso[k][sl] >= 0 && ((reg_rd >> fsl[sl]) & 1);
#pragma unroll
                    for (int m = 0; m < ME; m++) tsold[k][sl][m] = rd ? rg[tso[k][sl] + m * st] : 0.0;
                }
                double qd[ME];
#pragma unroll
                for (int m = 0; m < ME; m++) qd[m] = (double)qc[k][m];
                if (tso[k][0] >= 0) S::template cons_flux<0>(qd, pp.grav, tsv0[k]);
                if (tso[k][2] >= 0) S::template cons_flux<1>(qd, pp.grav, tsv2[k]);
            }
        }

        // ---- x-sweep on row J (reference rows 2-mbc .. my+mbc-1) -----------
        R fmxn[CPL][ME], fpxn[CPL][ME], Gn[CPL][ME], d2n[CPL][ME];
        if (Jr >= 1 && J <= NY - 2) {
            Face fx[CPL];
            Trans tx[CPL];
            R qa[CPL][ME];
#pragma unroll
            for (int m = 0; m < ME; m++) {
                R t[CPL];
#pragma unroll
                for (int k = 0; k < CPL; k++) t[k] = qc[k][m];
#pragma unroll
                for (int k = 0; k < CPL; k++) qa[k][m] = from_left<CPL>(t, k);
            }
            R cda[CPL][NCD];
#pragma unroll
            for (int n = 0; n < NCD; n++) {
                R t[CPL];
#pragma unroll
                for (int k = 0; k < CPL; k++) t[k] = cdc[k][n];
#pragma unroll
                for (int k = 0; k < CPL; k++) cda[k][n] = S::NC > 0 ? from_left<CPL>(t, k) : R(0);
            }
            R vl_cur[CPL];
#pragma unroll
            for (int k = 0; k < CPL; k++) vl_cur[k] = rowin ? vlr[T1][k] : R(0);
            // f-wave: the low-side cell's right-edge velocity and its top-edge
            // velocity and y length in rows J and J+1
            R fw_uRl[CPL], fw_vTl[CPL], fw_eyl[CPL], fw_vTnl[CPL], fw_eynl[CPL];
            if constexpr (S::FW) {
#pragma unroll
                for (int k = 0; k < CPL; k++) {
                    fw_uRl[k] = from_left<CPL>(fuR, k);
                    fw_vTl[k] = from_left<CPL>(fvT, k);
                    fw_eyl[k] = from_left<CPL>(fey, k);
                    fw_vTnl[k] = from_left<CPL>(fvTn, k);
                    fw_eynl[k] = from_left<CPL>(feyn, k);
                }
            }
#pragma unroll
            for (int k = 0; k < CPL; k++) {
                if constexpr (S::FW) {
                    S::rpn2f(qa[k][0], qc[k][0], fex[k], fuL[k], fw_uRl[k], fx[k]);
                    tx[k] = Trans{S::mean(fw_eyl[k], fvB[k], fw_vTl[k]), S::mean(fw_eynl[k], fvBn[k], fw_vTnl[k]),
                                  S::mean(fey[k], fvB[k], fw_vTl[k]), S::mean(feyn[k], fvBn[k], fw_vTnl[k])};
                } else {
                    S::template rpn2<0>(qa[k], qc[k], cda[k], cdc[k], ucur[k], pp, fx[k], tx[k]);
                }
                if constexpr (S::MAUX > 0 && !S::FW) {
                    tx[k] = Trans{vl_cur[k], vl_next[k], vcur[k], vnext[k]};
                }
                // Courant number on faces i = 1 .. mx+1 of every swept row
                // (FP32: max(dt/dx s, -dt/dx s) as dt/dx |s|, dt/dx > 0)
                if (cv[k] && col[k] >= 2 && col[k] <= NX - 2 && J <= Jlast) {
                    if constexpr (S::SMAX) cfl[k] = pmax(cfl[k], dtdx * S::smax(fx[k], tx[k]));
                    else
#pragma unroll
                    for (int w = 0; w < MW; w++) {
                        if constexpr (sizeof(R) == 4) {
                            cfl[k] = pmax(cfl[k], dtdx * M<R>::abs_(fx[k].s[w]));
                        } else {
                            R a = dtdx * fx[k].s[w];
                            R b = -dtdx * fx[k].s[w];
                            cfl[k] = dmax(dmax(cfl[k], a), b);
                        }
                    }
                }
            }
            R cq[CPL][ME];
#pragma unroll
            for (int k = 0; k < CPL; k++)
#pragma unroll
                for (int m = 0; m < ME; m++) cq[k][m] = R(0);
            if (second) {
                // limiter: needs the unlimited waves of the faces on each side
                R dotr[CPL][MW];
#pragma unroll
                for (int w = 0; w < MW; w++) {
                    R wr[CPL][ME];
#pragma unroll
                    for (int m = 0; m < ME; m++) {
                        R t[CPL];
#pragma unroll
                        for (int k = 0; k < CPL; k++) t[k] = fx[k].w[w][m];
#pragma unroll
                        for (int k = 0; k < CPL; k++) wr[k][m] = from_right<CPL>(t, k);
                    }
#pragma unroll
                    for (int k = 0; k < CPL; k++) {
                        R d = R(0);
#pragma unroll
                        for (int m = 0; m < ME; m++)
                            if (!(SKIPZ && S::template wz<0>(w, m))) d = d + fx[k].w[w][m] * wr[k][m];
                        dotr[k][w] = RAT ? wr[k][0] : d;
                    }
                }
#pragma unroll
                for (int w = 0; w < MW; w++) {
                    R t[CPL];
#pragma unroll
                    for (int k = 0; k < CPL; k++) t[k] = RAT ? fx[k].w[w][0] : dotr[k][w];
#pragma unroll
                    for (int k = 0; k < CPL; k++) {
                        R dotl = from_left<CPL>(t, k);
                        if (mth[w] != 0 && col[k] >= 2 && col[k] <= NX - 2) {
                            R wn = R(0);
                            if constexpr (!RAT) {
#pragma unroll
                                for (int m = 0; m < ME; m++)
                                    if (!(SKIPZ && S::template wz<0>(w, m))) wn = wn + fx[k].w[w][m] * fx[k].w[w][m];
                            }
                            if (RAT ? lim_ok1(fx[k].w[w][0]) : M<R>::lim_ok(wn)) {
                                R r = RAT ? (fx[k].s[w] > R(0) ? dotl : dotr[k][w]) * rcp1(fx[k].w[w][0])
                                          : M<R>::divl(fx[k].s[w] > R(0) ? dotl : dotr[k][w], wn);
                                R phi = limiter_phi<R>(mth[w], r);
#pragma unroll
                                for (int m = 0; m < ME; m++)
                                    if (!(SKIPZ && S::template wz<0>(w, m))) fx[k].w[w][m] = phi * fx[k].w[w][m];
                            }
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < CPL; k++)
#pragma unroll
                    for (int m = 0; m < ME; m++) {
                        R c = R(0);
#pragma unroll
                        for (int w = 0; w < MW; w++) {
                            if (SKIPZ && S::template wz<0>(w, m)) continue;
                            R as = M<R>::abs_(fx[k].s[w]);
                            if constexpr (S::FW) c = c + S::fw_weight(fx[k].s[w]) * (R(1) - as * dtdx) * fx[k].w[w][m];
                            else c = c + as * (R(1) - as * dtdx) * fx[k].w[w][m];
                        }
                        cq[k][m] = c;
                    }
            }
            R fam[CPL][ME], fap[CPL][ME];
#pragma unroll
            for (int k = 0; k < CPL; k++)
#pragma unroll
                for (int m = 0; m < ME; m++) {
                    R a = zadd(fx[k].am[m]);
                    R b = zsub(fx[k].ap[m]);
                    if (second) {
                        a = a + R(0.5) * cq[k][m];
                        b = b + R(0.5) * cq[k][m];
                    }
                    fam[k][m] = a;
                    fap[k][m] = b;
                }
            R hmm[CPL][ME], hpm[CPL][ME], hmp[CPL][ME], hpp[CPL][ME];
#pragma unroll
            for (int k = 0; k < CPL; k++) {
                if (trans) {
                    R am2[ME], ap2[ME], bm[ME], bp[ME];
#pragma unroll
                    for (int m = 0; m < ME; m++) {
                        am2[m] = cqsplit ? fx[k].am[m] + cq[k][m] : fx[k].am[m];
                        ap2[m] = cqsplit ? fx[k].ap[m] - cq[k][m] : fx[k].ap[m];
                    }
                    if constexpr (S::RPT2_SCALED) {
                        S::template rpt2_scaled<0>(am2, tx[k], hdtdx, hmm[k], hpm[k]);
                        S::template rpt2_scaled<0>(ap2, tx[k], hdtdx, hmp[k], hpp[k]);
                    } else {
                        S::template rpt2<0>(1, am2, tx[k], pp, bm, bp);
#pragma unroll
                        for (int m = 0; m < ME; m++) {
                            hmm[k][m] = hdtdx * bm[m];
                            hpm[k][m] = hdtdx * bp[m];
                        }
                        S::template rpt2<0>(2, ap2, tx[k], pp, bm, bp);
#pragma unroll
                        for (int m = 0; m < ME; m++) {
                            hmp[k][m] = hdtdx * bm[m];
                            hpp[k][m] = hdtdx * bp[m];
                        }
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < ME; m++) hmm[k][m] = hpm[k][m] = hmp[k][m] = hpp[k][m] = R(0);
                }
            }
#pragma unroll
            for (int m = 0; m < ME; m++) {
                R t1[CPL], t2[CPL], t3[CPL];
#pragma unroll
                for (int k = 0; k < CPL; k++) {
                    t1[k] = hmm[k][m];
                    t2[k] = hpm[k][m];
                    t3[k] = fam[k][m];
                }
#pragma unroll
                for (int k = 0; k < CPL; k++) {
                    R d1 = trans ? zsub(from_right<CPL>(t1, k)) - hmp[k][m] : R(0);
                    R d2 = trans ? zsub(from_right<CPL>(t2, k)) - hpp[k][m] : R(0);
                    fmxn[k][m] = zadd(from_right<CPL>(t3, k));
                    fpxn[k][m] = zadd(fap[k][m]);
                    Gn[k][m] = zadd(d2p[k][m]) + d1;
                    d2n[k][m] = d2;
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < CPL; k++)
#pragma unroll
                for (int m = 0; m < ME; m++) fmxn[k][m] = fpxn[k][m] = Gn[k][m] = d2n[k][m] = R(0);
        }

        // ---- y-sweep: solve face J of each column, finalise face J-1 --------
        R e1[CPL][ME], e2[CPL][ME], fdm[CPL][ME], fdp[CPL][ME];
        R hmp_new[CPL][ME], hpp_new[CPL][ME];
        Face fyn[CPL];
        Trans tyn[CPL];
        R dotry[CPL][MW];
        // These are made only on the rows whose faces exist (Jr >= 2, 3) and
        // left unassigned on the first rows of a sweep, since nothing reads
        // them there: e1, e2, fdm and fdp are read by the update and the
        // time-sync slots, which start at Jr >= 4; through the rings, fdp is
        // read by the next row's update, dotry by the limiter of the next
        // face finalised (Jr >= 3), and hmp_new / hpp_new by the next row's
        // e1, e2, which again only the update reads.  Zeroing them cost one
        // move per value and row, placed before the row's branches.
        if (Jr == 0) {
#pragma unroll
            for (int k = 0; k < CPL; k++) {
                fyn[k] = Face{};
                tyn[k] = Trans{};
            }
        } else {
            // f-wave: the transverse splits of face J are made here and
            // carried with the face (low-side split: x length of row J-1,
            // high-side: of row J; "up" is the column to the right)
            R fw_dn[CPL], fw_up[CPL], fw_dnr[CPL], fw_upr[CPL];
            if constexpr (S::FW) {
#pragma unroll
                for (int k = 0; k < CPL; k++) {
                    fw_dn[k] = S::mean(fexm[k], fuL[k], fuRm[k]);
                    fw_up[k] = S::mean(fex[k], fuL[k], fuRm[k]);
                }
#pragma unroll
                for (int k = 0; k < CPL; k++) {
                    fw_dnr[k] = from_right<CPL>(fw_dn, k);
                    fw_upr[k] = from_right<CPL>(fw_up, k);
                }
            }
#pragma unroll
            for (int k = 0; k < CPL; k++) {
                if constexpr (S::FW) {
                    S::rpn2f(qm1[k][0], qc[k][0], fey[k], fvB[k], fvTm[k], fyn[k]);
                    tyn[k] = Trans{fw_dn[k], fw_dnr[k], fw_up[k], fw_upr[k]};
                } else {
                R cdm[NCD];  // row J-1, recomputed rather than carried
                if constexpr (S::NC > 0) S::cell(qm1[k], pp, cdm);
                else cdm[0] = R(0);
                S::template rpn2<1>(qm1[k], qc[k], cdm, cdc[k], vcur[k], pp, fyn[k], tyn[k]);
                }
                if (cv[k] && col[k] >= 1 && col[k] <= NX - 2 && J >= 2 && J <= NY - 2 && J > J0 && J <= Jlast) {
                    if constexpr (S::SMAX) cfl[k] = pmax(cfl[k], dtdy * S::smax(fyn[k], tyn[k]));
                    else
#pragma unroll
                    for (int w = 0; w < MW; w++) {
                        if constexpr (sizeof(R) == 4) {
                            cfl[k] = pmax(cfl[k], dtdy * M<R>::abs_(fyn[k].s[w]));
                        } else {
                            R a = dtdy * fyn[k].s[w];
                            R b = -dtdy * fyn[k].s[w];
                            cfl[k] = dmax(dmax(cfl[k], a), b);
                        }
                    }
                }
            }
        }
        if (Jr >= 2) {
            // dot product of face J-1 with face J (both unlimited)
#pragma unroll
            for (int k = 0; k < CPL; k++)
#pragma unroll
                for (int w = 0; w < MW; w++) {
                    R d = R(0);
#pragma unroll
                    for (int m = 0; m < ME; m++)
                        if (!(SKIPZ && S::template wz<1>(w, m))) d = d + fyp[k].w[w][m] * fyn[k].w[w][m];
                    // one-equation FP32: the unlimited wave of face J-1 itself
                    dotry[k][w] = RAT ? fyp[k].w[w][0] : d;
                }
        }
        if (Jr >= 3) {
            // finalise y-face F = J-1 (2 <= F <= NY-2)
            R urt_m1[CPL], urt_m2[CPL];
#pragma unroll
            for (int k = 0; k < CPL; k++) {
                urt_m1[k] = from_right<CPL>(um1, k);
                urt_m2[k] = from_right<CPL>(um2, k);
            }
#pragma unroll
            for (int k = 0; k < CPL; k++) {
                Face& f = fyp[k];
                R cq[ME];
#pragma unroll
                for (int m = 0; m < ME; m++) cq[m] = R(0);
                if (second) {
#pragma unroll
                    for (int w = 0; w < MW; w++) {
                        if (mth[w] != 0) {
                            R wn = R(0);
                            if constexpr (!RAT) {
#pragma unroll
                                for (int m = 0; m < ME; m++)
                                    if (!(SKIPZ && S::template wz<1>(w, m))) wn = wn + f.w[w][m] * f.w[w][m];
                            }
                            // RAT: dotly is the wave of face J-2, fyn the (unlimited) face J
                            if (RAT ? lim_ok1(f.w[w][0]) : M<R>::lim_ok(wn)) {
                                R r = RAT ? (f.s[w] > R(0) ? dotly[k][w] : fyn[k].w[w][0]) * rcp1(f.w[w][0])
                                          : M<R>::divl(f.s[w] > R(0) ? dotly[k][w] : dotry[k][w], wn);
                                R phi = limiter_phi<R>(mth[w], r);
#pragma unroll
                                for (int m = 0; m < ME; m++)
                                    if (!(SKIPZ && S::template wz<1>(w, m))) f.w[w][m] = phi * f.w[w][m];
                            }
                        }
                    }
#pragma unroll
                    for (int m = 0; m < ME; m++) {
                        R c = R(0);
#pragma unroll
                        for (int w = 0; w < MW; w++) {
                            if (SKIPZ && S::template wz<1>(w, m)) continue;
                            R as = M<R>::abs_(f.s[w]);
                            if constexpr (S::FW) c = c + S::fw_weight(f.s[w]) * (R(1) - as * dtdy) * f.w[w][m];
                            else c = c + as * (R(1) - as * dtdy) * f.w[w][m];
                        }
                        cq[m] = c;
                    }
                }
#pragma unroll
                for (int m = 0; m < ME; m++) {
                    R a = zadd(f.am[m]);
                    R b = zsub(f.ap[m]);
                    if (second) {
                        a = a + R(0.5) * cq[m];
                        b = b + R(0.5) * cq[m];
                    }
                    fdm[k][m] = a;
                    fdp[k][m] = b;
                }
                if (trans) {
                    Trans t = typ[k];
                    if constexpr (S::MAUX > 0 && !S::FW) {
                        // amdq split uses cell F-1 = J-2, apdq split cell F = J-1;
                        // "down" is this column, "up" the column to the right
                        t = Trans{um2[k], urt_m2[k], um1[k], urt_m1[k]};
                    }
                    R am2[ME], ap2[ME], bm[ME], bp[ME];
#pragma unroll
                    for (int m = 0; m < ME; m++) {
                        am2[m] = cqsplit ? f.am[m] + cq[m] : f.am[m];
                        ap2[m] = cqsplit ? f.ap[m] - cq[m] : f.ap[m];
                    }
                    R hmm_y[ME], hpm_y[ME];
                    if constexpr (S::RPT2_SCALED) {
                        S::template rpt2_scaled<1>(am2, t, hdtdy, hmm_y, hpm_y);
                        S::template rpt2_scaled<1>(ap2, t, hdtdy, hmp_new[k], hpp_new[k]);
                    } else {
                        S::template rpt2<1>(1, am2, t, pp, bm, bp);
#pragma unroll
                        for (int m = 0; m < ME; m++) {
                            hmm_y[m] = hdtdy * bm[m];
                            hpm_y[m] = hdtdy * bp[m];
                        }
                        S::template rpt2<1>(2, ap2, t, pp, bm, bp);
#pragma unroll
                        for (int m = 0; m < ME; m++) {
                            hmp_new[k][m] = hdtdy * bm[m];
                            hpp_new[k][m] = hdtdy * bp[m];
                        }
                    }
#pragma unroll
                    for (int m = 0; m < ME; m++) {
                        e1[k][m] = zsub(hmm_y[m]) - hmp_prev[k][m];
                        e2[k][m] = zsub(hpm_y[m]) - hpp_prev[k][m];
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < ME; m++) e1[k][m] = e2[k][m] = R(0);
                }
            }
        }

        // ---- interior cell row R = J-2 ------------------------------------
        const int Rw = J - 2;
        if (Rw >= ra && Rw < rb) {
            // neighbour terms by DPP from every lane first, then one branch
            // per column for the update and its stores
            R e1rv[CPL][ME], e2lv[CPL][ME];
#pragma unroll
            for (int m = 0; m < ME; m++) {
                R t1[CPL], t2[CPL];
#pragma unroll
                for (int k = 0; k < CPL; k++) {
                    t1[k] = e1[k][m];
                    t2[k] = e2[k][m];
                }
#pragma unroll
                for (int k = 0; k < CPL; k++) {
                    e1rv[k][m] = from_right<CPL>(t1, k);
                    e2lv[k][m] = from_left<CPL>(t2, k);
                }
            }
#pragma unroll
            for (int k = 0; k < CPL; k++) {
                if (cv[k] && col[k] >= 2 && col[k] <= NX - 3) {
                    // push staging slots of this cell: x-edge column (two
                    // deep on each side) and y-edge row, -1 where none
                    const int c = col[k];
                    const int sxc = c <= 3 ? c - 2 : (c >= NX - 4 ? 2 + c - (NX - 4) : -1);
                    const int syr = Rw <= 3 ? Jo - 4 : (Rw >= NY - 4 ? Jo - NY + 4 : -1);
#pragma unroll
                    for (int m = 0; m < ME; m++) {
                        R fmr = (fmx2[k][m] + e2[k][m]) + e1rv[k][m];
                        R fpc = (fpx2[k][m] + e2lv[k][m]) + e1[k][m];
                        R gmu = G1[k][m] + fdm[k][m];
                        R gpc = G2[k][m] + fdp_prev[k][m];
                        R qn = (qm2[k][m] - dtdx * (fmr - fpc)) - dtdy * (gmu - gpc);
                        qo[((long long)m * NY + (Jo - 2)) * NX + col[k]] = qn;
                        if constexpr (PUSH) {
                            if (sxc >= 0) stage[(sxc * ME + m) * A.my + (Jo - 4)] = qn;
                            if (syr >= 0) stage[4 * ME * A.my + (syr * ME + m) * A.mx + (c - 2)] = qn;
                        }
                    }
                }
            }
#pragma unroll
            for (int m = 0; m < ME; m++) {
#pragma unroll
                for (int k = 0; k < CPL; k++) {
                    const R e1r = e1rv[k][m], e2l = e2lv[k][m];
                    (void)e1r;
                    (void)e2l;
                    if constexpr (SYNC) {
                        // fluctuations on the patch edges (accumulate_waves)
                        const int c = col[k];
                        if (tso[k][1] >= 0) {
                            double* r = rg + tso[k][1] + m * A.my;
                            if (c == 1 || c == NX - 3)
                                *r = reg_plus(tsold[k][1][m], dtel_y, (double)((fmx2[k][m] + e2[k][m]) + e1r));
                            else
                                *r = reg_minus(tsold[k][1][m], dtel_y, (double)((fpx2[k][m] + e2l) + e1[k][m]));
                        }
                        if (tso[k][3] >= 0) {
                            double* r = rg + tso[k][3] + m * A.mx;
                            if (Rw == 2)
                                *r = reg_minus(tsold[k][3][m], dtel_x, (double)(G2[k][m] + fdp_prev[k][m]));
                            else
                                *r = reg_plus(tsold[k][3][m], dtel_x, (double)(G1[k][m] + fdm[k][m]));
                        }
                    }
                }
            }
        }

        // ---- time-sync fluctuations on the two ghost-row edge faces, and
        // the edge fluxes of the row's data (slots 0 and 2) ------------------
        if constexpr (SYNC) {
#pragma unroll
            for (int k = 0; k < CPL; k++) {
#pragma unroll
                for (int m = 0; m < ME; m++) {
                    if (tso[k][4] >= 0) {
                        // gm on the bottom edge face (top face of ghost row 1),
                        // gp on the top edge face (bottom face of ghost row NY-2)
                        double* r = rg + tso[k][4] + m * A.mx;
                        if (J == 3)
                            *r = reg_plus(tsold[k][4][m], dtel_x, (double)(G1[k][m] + fdm[k][m]));
                        else
                            *r = reg_minus(tsold[k][4][m], dtel_x, (double)(G1[k][m] + fdp[k][m]));
                    }
                    if (tso[k][0] >= 0)
                        rg[tso[k][0] + m * A.my] = reg_plus(tsold[k][0][m], dtel_y, tsv0[k][m]);
                    if (tso[k][2] >= 0)
                        rg[tso[k][2] + m * A.mx] = reg_plus(tsold[k][2][m], dtel_x, tsv2[k][m]);
                }
            }
        }

        // ---- this row's values enter the rings (renames when unrolled) ----
#pragma unroll
        for (int k = 0; k < CPL; k++) {
#pragma unroll
            for (int m = 0; m < ME; m++) {
                fmxn_w[k][m] = fmxn[k][m];
                fpxn_w[k][m] = fpxn[k][m];
                Gn_w[k][m] = Gn[k][m];
                d2n_w[k][m] = d2n[k][m];
                hmp_new_w[k][m] = hmp_new[k][m];
                hpp_new_w[k][m] = hpp_new[k][m];
                fdp_w[k][m] = fdp[k][m];
            }
#pragma unroll
            for (int w = 0; w < MW; w++) dotry_w[k][w] = dotry[k][w];
            fyn_w[k] = fyn[k];
            if constexpr (S::MAUX > 0 && !S::FW) {
                vnr[T0][k] = vnext[k];
                vlr[T0][k] = vl_next[k];
            }
            tyn_w[k] = tyn[k];
        }
        if constexpr (S::MAUX > 0 && !S::FW && sizeof(R) == 4) {
            // factors of row J+1, issued after this row's stores
            if (A.swirl) {
                const int Jn = __builtin_amdgcn_readfirstlane(min(J + 1, NY - 1));
                ydb[T0] = swyf[NY + Jn];
                yb1[T0] = swyf[min(Jn + 1, NY - 1)];
            }
        }
    };

    // the advection velocity v of the sweep's first row, as its predecessor
    // row would have made it
    if constexpr (S::MAUX > 0 && !S::FW) {
#pragma unroll
        for (int k = 0; k < CPL; k++) vnr[1][k] = vel_v(k, J0);
#pragma unroll
        for (int k = 0; k < CPL; k++) vlr[1][k] = from_left<CPL>(vnr[1], k);
        if (A.swirl && sizeof(R) == 4) {
            ydb[1] = swyf[NY + J0];
            yb1[1] = swyf[min(J0 + 1, NY - 1)];
        }
    }
    if constexpr (UNR == 4) {
        if constexpr (PF) {
#pragma unroll
            for (int k = 0; k < CPL; k++) {
#pragma unroll
                for (int m = 0; m < ME; m++)
                    qr[0][k][m] = cv[k] ? qp[((long long)m * NY + J0) * NX + col[k]] : R(0);
            }
        }
        for (int J = J0; J <= Jlast; J += 4) {
            row(std::integral_constant<int, 0>{}, J);
            row(std::integral_constant<int, 1>{}, J + 1);
            row(std::integral_constant<int, 2>{}, J + 2);
            row(std::integral_constant<int, 3>{}, J + 3);
        }
    } else {
        for (int J = J0; J <= Jlast; J++) {
            row(std::integral_constant<int, 0>{}, J);
            // rotate the rings: slot 0 holds row J, slots 3 / 2 / 1 are read
            // as rows J-1 / J-2 / J-1 by the next row
#pragma unroll
            for (int k = 0; k < CPL; k++) {
                ur[2][k] = ur[3][k];
                ur[3][k] = ur[0][k];
                vnr[1][k] = vnr[0][k];
                vlr[1][k] = vlr[0][k];
                if (k == 0) {
                    ydb[1] = ydb[0];
                    yb1[1] = yb1[0];
                }
#pragma unroll
                for (int m = 0; m < ME; m++) {
                    qr[2][k][m] = qr[3][k][m];
                    qr[3][k][m] = qr[0][k][m];
                    fmr[2][k][m] = fmr[3][k][m];
                    fmr[3][k][m] = fmr[0][k][m];
                    fpr[2][k][m] = fpr[3][k][m];
                    fpr[3][k][m] = fpr[0][k][m];
                    Gr[2][k][m] = Gr[3][k][m];
                    Gr[3][k][m] = Gr[0][k][m];
                    d2r[1][k][m] = d2r[0][k][m];
                    hmpr[1][k][m] = hmpr[0][k][m];
                    hppr[1][k][m] = hppr[0][k][m];
                    fdpr[1][k][m] = fdpr[0][k][m];
                }
#pragma unroll
                for (int w = 0; w < MW; w++) dotr2[1][k][w] = dotr2[0][k][w];
                fyr[1][k] = fyr[0][k];
                tyr[1][k] = tyr[0][k];
            }
        }
    }

    // ---- same-level ghost push --------------------------------------------
    // After the sweep the lanes of the wave (of the workgroup, with row
    // bands) store the staged edge cells of the wave's patches into the
    // ghost cells of their same-level neighbours across faces and corners,
    // one staged value per lane and iteration over the patch's 4*meqn*(mx+my)
    // values.  x-edge values go with the edge column fastest (lane pairs
    // store 8 contiguous bytes), then the row; y-edge values with the column
    // fastest (contiguous stores).
    if constexpr (PUSH) {
        __syncthreads();  // every thread of the block reaches this point
        const int tl = BAND ? (int)threadIdx.x : lane;    // pushing thread
        const int tn = BAND ? (int)blockDim.x : HC_WAVE;  // pushing threads
        const int nxi = 4 * ME * A.my;                    // x-edge values per patch
        const float rmy = 1.0f / (float)A.my, rmx = 1.0f / (float)A.mx;
        // a / b for 0 <= a < 2^20 from a float reciprocal (one correction)
        auto qdiv = [](int a, int b, float rb) {
            int q = (int)((float)a * rb);
            if (q * b > a) q--;
            else if ((q + 1) * b <= a) q++;
            return q;
        };
        for (int sp = 0; sp < A.ppw; sp++) {
            const int pi = (BAND ? blk : gwave) * A.ppw + sp;
            if (pi >= A.npatch || pi >= A.ntiles) break;  // uniform across the wave
            const int st = A.slot0 + pi;
            // the patch's neighbour table, loaded once (uniform across the wave)
            const int* tb = A.push + 8 * st;
            int nt[8];
#pragma unroll
            for (int k = 0; k < 8; k++) nt[k] = tb[k];
            R* ob = static_cast<R*>(A.qout_l[A.levels[st]]);
            const R* sg = reinterpret_cast<const R*>(hc_smem) + (swave * A.ppw + sp) * pstage;
            auto put = [&](int nb, int cc, int rr, int m, R v) {
                ob[(long long)nb * A.qstride + ((long long)m * NY + rr) * NX + cc] = v;
            };
            for (int e = tl; e < pstage; e += tn) {
                if (e < nxi) {
                    // x edges: e = ((side*ME + m)*my + r)*2 + cb
                    const int cb = e & 1, h = e >> 1;
                    const int g = qdiv(h, A.my, rmy);
                    const int r = h - g * A.my;
                    const int side = g / ME, m = g - side * ME;
                    const R v = sg[((side * 2 + cb) * ME + m) * A.my + r];
                    const int nb = side ? nt[1] : nt[0];
                    const int nbl = side ? nt[5] : nt[4], nbh = side ? nt[7] : nt[6];
                    const int c = (side ? NX - 4 - A.mx : 2 + A.mx) + cb;
                    const int Rw = r + 2;
                    if (nb >= 0) put(nb, c, Rw, m, v);
                    if (Rw <= 3 && nbl >= 0) put(nbl, c, Rw + A.my, m, v);
                    if (Rw >= NY - 4 && nbh >= 0) put(nbh, c, Rw - A.my, m, v);
                } else {
                    // y edges: e - nxi = ((side*2 + rr)*ME + m)*mx + i
                    const int h = e - nxi;
                    const int g = qdiv(h, A.mx, rmx);
                    const int i = h - g * A.mx;
                    const int sr = g / ME, m = g - sr * ME;
                    const int side = sr >> 1, rr = sr & 1;
                    const int nb = side ? nt[3] : nt[2];
                    const int Rt = side ? NY - 4 + rr - A.my : 2 + rr + A.my;
                    if (nb >= 0) put(nb, i + 2, Rt, m, sg[h + nxi]);
                }
            }
        }
    }

    // ---- per-patch Courant number: segmented max over the patch's lanes ----
    R c = R(0);
#pragma unroll
    for (int k = 0; k < CPL; k++) c = dmax(c, cfl[k]);
    double cd = (double)c;
#pragma unroll
    for (int d = 1; d < HC_WAVE; d <<= 1) {
        double t = __shfl_up(cd, d, HC_WAVE);
        if (lp - d >= 0) cd = dmax(cd, t);
    }
    if constexpr (BAND) {
        // the bands of a patch are the waves of one workgroup: max over them
        __shared__ double cfl_band[8][HC_WAVE];
        if (lp == A.lpp - 1 && slot < A.ppw) cfl_band[band][slot] = cd;
        __syncthreads();
        if (band == 0 && pact && lp == A.lpp - 1 && cflp) {
            double m = cfl_band[0][slot];
            for (int b = 1; b < A.nband; b++) m = dmax(m, cfl_band[b][slot]);
            cflp[tix] = m;
            if (A.cflmax && m > 0.0) atomicMax(A.cflmax, (unsigned long long)__double_as_longlong(m));
        }
    } else {
        if (pact && lp == A.lpp - 1 && cflp) cflp[tix] = cd;
        if (A.cflmax) {
            // one atomic per wave, of the maximum over its patches
            double wm = (pact && lp == A.lpp - 1) ? cd : 0.0;
#pragma unroll
            for (int d = 1; d < HC_WAVE; d <<= 1) wm = dmax(wm, __shfl_xor(wm, d, HC_WAVE));
            if (lane == 0 && wm > 0.0) atomicMax(A.cflmax, (unsigned long long)__double_as_longlong(wm));
        }
    }
}

}  // namespace hc
