// hc_api.hip -- C ABI entry points of libhipclaw (declared in include/hipclaw.h)
// and the small batched kernels that sit around the patch step: swirl
// velocities, physical boundary conditions and the level CFL reduction.
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <mutex>
#include <string>

#include "hc_internal.hpp"

namespace hc {

// Per-tile Courant numbers of hipclaw_step2_batch_noalloc: a file-scope
// device array (each GPU holds its own copy until the process ends).  Host
// code never passes the array's name to a kernel: its device address on the
// current GPU comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) double scratch[HIPCLAW_NOALLOC_MAX_TILES];
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Device address of `scratch` on the current device (looked up on the first
// call per device).
static int scratch_ptr(double** out)
{
    int dev = -1;
    int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= kMaxDevices) return set_error(HIPCLAW_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (rc) return rc;
        if (!p) return set_error(HIPCLAW_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return HIPCLAW_OK;
}

__global__ void k_fill_zero(double* __restrict__ x, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int check_hip(hipError_t e, const char* what)
{
    if (e == hipSuccess) return HIPCLAW_OK;
    return set_error(HIPCLAW_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int validate_physics(const hipclaw_physics* p)
{
    if (!p) return set_error(HIPCLAW_ERR_ARG, "physics is NULL");
    if (p->mbc != 2) return set_error(HIPCLAW_ERR_UNSUPPORTED, "mbc must be 2");
    if (p->mx < 4 || p->my < 4) return set_error(HIPCLAW_ERR_ARG, "mx, my must be >= 4");
    if (p->precision != HIPCLAW_FP64 && p->precision != HIPCLAW_FP32)
        return set_error(HIPCLAW_ERR_ARG, "precision must be HIPCLAW_FP64 or HIPCLAW_FP32");
    if (p->order[0] < 1 || p->order[0] > 2 || p->order[1] < 0 || p->order[1] > 2)
        return set_error(HIPCLAW_ERR_ARG, "order must be (1|2, 0|1|2)");
    if ((p->solver & ~(HIPCLAW_CLAW5 | 0xff)) != 0) return set_error(HIPCLAW_ERR_ARG, "unknown solver flags");
    switch (solver_family(p->solver)) {
    case HIPCLAW_SOLVER_ADV:
        if (p->meqn != 1 || p->mwaves != 1 || p->maux != 2)
            return set_error(HIPCLAW_ERR_ARG, "advection needs meqn=mwaves=1, maux=2");
        break;
    case HIPCLAW_SOLVER_SWE_ROE:
        if (p->meqn != 3 || p->mwaves != 3 || p->maux != 0)
            return set_error(HIPCLAW_ERR_ARG, "shallow water needs meqn=mwaves=3, maux=0");
        if (!(p->grav > 0)) return set_error(HIPCLAW_ERR_ARG, "grav must be > 0");
        break;
    case HIPCLAW_SOLVER_EULER5_ROE:
        // maux = 1 carries the radius of the axisymmetric source (the
        // shock-bubble setaux); the step does not read it
        if (p->meqn != 5 || p->mwaves != 5 || p->maux < 0 || p->maux > 1)
            return set_error(HIPCLAW_ERR_ARG, "euler5 needs meqn=mwaves=5, maux=0|1");
        if (!(p->gamma > 1)) return set_error(HIPCLAW_ERR_ARG, "gamma must be > 1");
        break;
    case HIPCLAW_SOLVER_EULER_HLLE:
        if (p->meqn != 5 || p->mwaves != 2 || p->maux < 0 || p->maux > 1)
            return set_error(HIPCLAW_ERR_ARG, "euler HLLE needs meqn=5, mwaves=2, maux=0|1");
        if (!(p->gamma > 1)) return set_error(HIPCLAW_ERR_ARG, "gamma must be > 1");
        if (is_claw5(p->solver)) return set_error(HIPCLAW_ERR_UNSUPPORTED, "HLLE has no Clawpack 5 variant");
        break;
    case HIPCLAW_SOLVER_TRANSPORT_FW:
        if (p->meqn != 1 || p->mwaves != 1 || p->maux != 7)
            return set_error(HIPCLAW_ERR_ARG, "f-wave transport needs meqn=mwaves=1, maux=7");
        if (is_claw5(p->solver))
            return set_error(HIPCLAW_ERR_UNSUPPORTED, "f-wave transport has no Clawpack 5 variant");
        break;
    default:
        return set_error(HIPCLAW_ERR_ARG, "unknown solver");
    }
    for (int w = 0; w < p->mwaves; w++)
        if (p->mthlim[w] < 0 || p->mthlim[w] > 6)
            return set_error(HIPCLAW_ERR_ARG, "mthlim must be in 0..6");
    return HIPCLAW_OK;
}

// ---------------------------------------------------------------------------
// Swirl edge velocities from the stream function psi = sin^2(pi x) sin^2(pi y)/pi
// (applications/clawpack/advection/2d/swirl/user_4.6/b4step2.f:25-39).
// ---------------------------------------------------------------------------
template <class R>
// psi(x, y) = sin^2(pi x) sin^2(pi y) / pi is a product of one factor per
// coordinate, and the three corners b4step2 evaluates per cell, (xll, yll),
// (xll, yll+dy), (xll+dx, yll), use the coordinates xll_i, xll_i + dx, yll_j,
// yll_j + dy as the reference forms them.  One workgroup per patch evaluates
// the 2 (NX + NY) sines of those coordinates once into LDS; every cell then
// combines them with the same operations as psi, so the values are those of
// evaluating psi per cell (6 sines per cell before).
__global__ void __launch_bounds__(256) k_swirl_aux(int npatch, int mx, int my, const double* xylower,
                                                   const double* dxdy, double vt, R* aux)
{
    extern __shared__ double sn[];  // sin(pi*xll), sin(pi*(xll+dx)), sin(pi*yll), sin(pi*(yll+dy))
    const int NX = mx + 4, NY = my + 4;
    const long long T = (long long)NX * NY;
    const int p = blockIdx.x;
    if (p >= npatch) return;
    const double pi = 4.0 * atan(1.0);
    const double dx = dxdy[2 * p], dy = dxdy[2 * p + 1];
    const double x0 = xylower[2 * p], y0 = xylower[2 * p + 1];
    double* sx0 = sn;
    double* sx1 = sn + NX;
    double* sy0 = sn + 2 * NX;
    double* sy1 = sn + 2 * NX + NY;
    for (int t = threadIdx.x; t < 2 * (NX + NY); t += blockDim.x) {
        double arg;
        if (t < 2 * NX) {
            const int i = t < NX ? t : t - NX;
            const double xll = x0 + (double)(i - 2) * dx;
            arg = t < NX ? xll : xll + dx;
        } else {
            const int u = t - 2 * NX;
            const int j = u < NY ? u : u - NY;
            const double yll = y0 + (double)(j - 2) * dy;
            arg = u < NY ? yll : yll + dy;
        }
        sn[t] = sin(pi * arg);
    }
    __syncthreads();
    const double vty = vt / dy, vtx = vt / dx;
    R* au = aux + (long long)p * 2 * T;
    for (int cell = threadIdx.x; cell < (int)T; cell += blockDim.x) {
        const int j = cell / NX, i = cell - j * NX;
        const double a0 = sx0[i] * sx0[i], a1 = sx1[i] * sx1[i];
        const double b0 = sy0[j] * sy0[j], b1 = sy1[j] * sy1[j];
        const double pij = (a0 * b0) / pi;
        const double u = vty * ((a0 * b1) / pi - pij);
        const double v = -vtx * ((a1 * b0) / pi - pij);
        au[cell] = R(u);
        au[T + cell] = R(v);
    }
}

// ---------------------------------------------------------------------------
// clawpack46_bc2_default: left/right strips over the full height first, then
// bottom/top over the full width (so corners come from the second pass).
// One 64-lane group per patch.
// ---------------------------------------------------------------------------
template <class R>
__global__ void k_bc2(int npatch, int mx, int my, int meqn, R* q, const int* ids, int ntiles,
                      const int* mthbc)
{
    const int p = blockIdx.x;
    if (p >= npatch) return;
    const int tile = ids ? ids[p] : p;
    if (tile < 0 || tile >= ntiles) return;
    const int NX = mx + 4, NY = my + 4;
    R* qp = q + (long long)tile * NX * NY * meqn;
    const int* bc = mthbc + 4 * p;
    const int mbc = 2;
    const int i1 = mbc, imx = mx + mbc - 1, j1 = mbc, jmy = my + mbc - 1;
#define QB(i, j, m) qp[((long long)(m) * NY + (j)) * NX + (i)]
    for (int t = threadIdx.x; t < NY * meqn; t += blockDim.x) {
        const int m = t / NY, j = t - m * NY;
        for (int ibc = 1; ibc <= mbc; ibc++) {
            if (bc[0] == 1) QB(i1 - ibc, j, m) = QB(i1, j, m);
            else if (bc[0] == 3) QB(i1 - ibc, j, m) = (m == 1) ? -QB(i1 + ibc - 1, j, m) : QB(i1 + ibc - 1, j, m);
            if (bc[1] == 1) QB(imx + ibc, j, m) = QB(imx, j, m);
            else if (bc[1] == 3) QB(imx + ibc, j, m) = (m == 1) ? -QB(imx + 1 - ibc, j, m) : QB(imx + 1 - ibc, j, m);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < NX * meqn; t += blockDim.x) {
        const int m = t / NX, i = t - m * NX;
        for (int jbc = 1; jbc <= mbc; jbc++) {
            if (bc[2] == 1) QB(i, j1 - jbc, m) = QB(i, j1, m);
            else if (bc[2] == 3) QB(i, j1 - jbc, m) = (m == 2) ? -QB(i, j1 + jbc - 1, m) : QB(i, j1 + jbc - 1, m);
            if (bc[3] == 1) QB(i, jmy + jbc, m) = QB(i, jmy, m);
            else if (bc[3] == 3) QB(i, jmy + jbc, m) = (m == 2) ? -QB(i, jmy + 1 - jbc, m) : QB(i, jmy + 1 - jbc, m);
        }
    }
#undef QB
}

// Maximum of n non-negative doubles (one workgroup; n is a patch count).
// The result is +0 or positive (NaN and -0 entries never win the comparisons);
// as bits (OUT = unsigned long long) it orders as the doubles do.
template <class OUT> __device__ inline void max_reduce_body(const double* x, int n, OUT* out)
{
    __shared__ double s[1024];
    // eight independent loads in flight per thread: a level has up to ~10^5
    // patches, and one dependent load per iteration left the single
    // workgroup waiting on memory (max is exact, so the grouping is free)
    constexpr int U = 8;
    const int B = blockDim.x;
    double vv[U];
#pragma unroll
    for (int u = 0; u < U; u++) vv[u] = 0.0;
    int i = threadIdx.x;
    for (; i + (U - 1) * B < n; i += U * B) {
        double t[U];
#pragma unroll
        for (int u = 0; u < U; u++) t[u] = x[i + u * B];
#pragma unroll
        for (int u = 0; u < U; u++) vv[u] = t[u] > vv[u] ? t[u] : vv[u];
    }
    for (; i < n; i += B) vv[0] = x[i] > vv[0] ? x[i] : vv[0];
    double v = vv[0];
#pragma unroll
    for (int u = 1; u < U; u++) v = vv[u] > v ? vv[u] : v;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x + w] > s[threadIdx.x] ? s[threadIdx.x + w] : s[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if constexpr (sizeof(OUT) == 8 && !std::is_floating_point<OUT>::value)
            *out = (OUT)__double_as_longlong(s[0]);
        else
            *out = s[0];
    }
}
__global__ void __launch_bounds__(1024) k_max_reduce(const double* x, int n, double* out)
{
    max_reduce_body(x, n, out);
}
__global__ void __launch_bounds__(1024) k_max_reduce_bits(const double* x, int n, unsigned long long* out)
{
    max_reduce_body(x, n, out);
}

// Maximum of the n slots of a coarse step (bits of non-negative doubles)
// into out as a double; every slot is cleared after it is read.
__global__ void __launch_bounds__(1024) k_cfl_final(unsigned long long* acc, int n, double* out)
{
    __shared__ unsigned long long s[1024];
    unsigned long long v = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const unsigned long long a = acc[i];
        v = a > v ? a : v;
        acc[i] = 0;
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && s[threadIdx.x + w] > s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = __longlong_as_double((long long)s[0]);
}

int cfl_slot_reduce(const double* x, int n, unsigned long long* slot, hipStream_t s)
{
    hipLaunchKernelGGL(k_max_reduce_bits, dim3(1), dim3(1024), 0, s, x, n, slot);
    return check_hip(hipGetLastError(), "cfl slot reduce");
}

int cfl_final(unsigned long long* acc, int n, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(1024), 0, s, acc, n, out);
    return check_hip(hipGetLastError(), "cfl final");
}

}  // namespace hc

using namespace hc;

extern "C" {

int hipclaw_abi_version(void) { return HIPCLAW_ABI_VERSION; }

const char* hipclaw_last_error(void) { return g_last_error.c_str(); }

int hipclaw_physics_default(int solver, int mx, int my, hipclaw_physics* out)
{
    if (!out) return set_error(HIPCLAW_ERR_ARG, "out is NULL");
    std::memset(out, 0, sizeof(*out));
    out->solver = solver;
    out->mx = mx;
    out->my = my;
    out->mbc = 2;
    out->order[0] = 2;
    out->order[1] = 2;
    out->efix = 1;
    out->precision = HIPCLAW_FP64;
    out->grav = 1.0;
    out->gamma = 1.4;
    for (int f = 0; f < 4; f++) out->mthbc[f] = 1;
    switch (solver_family(solver)) {
    case HIPCLAW_SOLVER_ADV: out->meqn = out->mwaves = 1; out->maux = 2; break;
    case HIPCLAW_SOLVER_SWE_ROE: out->meqn = out->mwaves = 3; break;
    case HIPCLAW_SOLVER_EULER5_ROE: out->meqn = out->mwaves = 5; break;
    case HIPCLAW_SOLVER_EULER_HLLE: out->meqn = 5; out->mwaves = 2; break;
    case HIPCLAW_SOLVER_TRANSPORT_FW: out->meqn = out->mwaves = 1; out->maux = 7; break;
    default: return set_error(HIPCLAW_ERR_ARG, "unknown solver");
    }
    for (int w = 0; w < 5; w++) out->mthlim[w] = w < out->mwaves ? 4 : 0;
    return HIPCLAW_OK;
}

static int step2_batch_impl(const hipclaw_physics* phys, int npatch, const void* q_in,
                            void* q_out, const void* aux, const double* dxdy,
                            const int* patch_ids, int ntiles, double dt, double* cfl_patch,
                            double* reg, const hipclaw_launch* launch, void* stream);

int hipclaw_step2_batch(const hipclaw_physics* phys, int npatch, const void* q_in, void* q_out,
                        const void* aux, const double* dxdy, const int* patch_ids, int ntiles,
                        double dt, double* cfl_patch, const hipclaw_launch* launch, void* stream)
{
    return step2_batch_impl(phys, npatch, q_in, q_out, aux, dxdy, patch_ids, ntiles, dt,
                            cfl_patch, nullptr, launch, stream);
}

int hipclaw_step2_batch_noalloc(const hipclaw_physics* phys, int npatch, const void* q_in, void* q_out,
                                const void* aux, const double* dxdy, const int* patch_ids, int ntiles,
                                double dt, double* maxcfl, const hipclaw_launch* launch, void* stream)
{
    int rc = validate_physics(phys);
    if (rc) return rc;
    if (!maxcfl) return set_error(HIPCLAW_ERR_ARG, "maxcfl is NULL");
    if (npatch < 0 || ntiles < 0) return set_error(HIPCLAW_ERR_ARG, "npatch, ntiles must be >= 0");
    if (!patch_ids && ntiles < npatch) return set_error(HIPCLAW_ERR_ARG, "ntiles < npatch without a patch list");
    if (ntiles > HIPCLAW_NOALLOC_MAX_TILES || npatch > HIPCLAW_NOALLOC_MAX_TILES)
        return set_error(HIPCLAW_ERR_ARG, "step2_batch_noalloc: at most " +
                                              std::to_string(HIPCLAW_NOALLOC_MAX_TILES) + " tiles per call");
    double* cfl = nullptr;
    rc = scratch_ptr(&cfl);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the array keeps the previous call's values: clear the tiles this call
    // reduces over (a fill kernel, so the call can be captured in a graph)
    const int nclear = ntiles > 0 ? ntiles : 1;
    hipLaunchKernelGGL(k_fill_zero, dim3((nclear + 255) / 256), dim3(256), 0, s, cfl, nclear);
    rc = check_hip(hipGetLastError(), "k_fill_zero launch");
    if (rc) return rc;
    rc = step2_batch_impl(phys, npatch, q_in, q_out, aux, dxdy, patch_ids, ntiles, dt, cfl, nullptr,
                          launch, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(1024), 0, s, cfl, ntiles, maxcfl);
    return check_hip(hipGetLastError(), "k_max_reduce launch");
}

int hipclaw_step2_batch_sync(const hipclaw_physics* phys, int npatch, const void* q_in,
                             void* q_out, const void* aux, const double* dxdy,
                             const int* patch_ids, int ntiles, double dt, double* cfl_patch,
                             double* registers, const hipclaw_launch* launch, void* stream)
{
    if (!registers) return set_error(HIPCLAW_ERR_ARG, "registers is NULL");
    if (phys && solver_family(phys->solver) != HIPCLAW_SOLVER_SWE_ROE)
        return set_error(HIPCLAW_ERR_UNSUPPORTED, "time-sync registers need the SWE solver");
    return step2_batch_impl(phys, npatch, q_in, q_out, aux, dxdy, patch_ids, ntiles, dt,
                            cfl_patch, registers, launch, stream);
}

static int step2_batch_impl(const hipclaw_physics* phys, int npatch, const void* q_in,
                            void* q_out, const void* aux, const double* dxdy,
                            const int* patch_ids, int ntiles, double dt, double* cfl_patch,
                            double* reg, const hipclaw_launch* launch, void* stream)
{
    return hc::step2_launch(phys, npatch, q_in, q_out, aux, dxdy, patch_ids, ntiles, dt, cfl_patch,
                            reg, launch, stream, nullptr);
}

}  // extern "C"

namespace hc {

// Launch shape of one patch-step launch.  The library default: two columns
// per lane on FP32 launches of 2,048 patches or more with wide patches (three
// patches per wave pay there), else one; one wave per workgroup below 1,024
// patches (a small launch spreads its waves one per workgroup); and below
// 2,048 patches row bands (8 up to 256 patches, 4 up to 512, else 2): a
// launch of fewer patches than the chip has wave slots is bound by the
// latency of one sweep, which bands shorten.  A caller's shape with
// row_bands = -1 is used for launches of 2,048 patches or more and the
// default below that; row bands apply to one column per lane only (two
// columns per lane sweep whole patches with the caller's waves per group).
int choose_launch(const hipclaw_physics* phys, int npatch, const hipclaw_launch* launch, int* cpl,
                  int* wpg, int* nband)
{
    if (launch && (launch->cols_per_lane < 1 || launch->cols_per_lane > 2 || launch->waves_per_group < 1 ||
                   launch->waves_per_group > 4 || launch->row_bands < -1 || launch->row_bands > 8))
        return set_error(HIPCLAW_ERR_ARG,
                         "launch: cols_per_lane 1..2, waves_per_group 1..4, row_bands -1..8");
    const bool small = npatch < 2048;
    if (launch && !(launch->row_bands < 0 && small)) {
        *cpl = launch->cols_per_lane;
        *wpg = launch->waves_per_group;
        *nband = launch->row_bands < 0 ? 1 : launch->row_bands;
    } else {
        *cpl = (phys->precision == HIPCLAW_FP32 && phys->mx + 4 > 32 && !small) ? 2 : 1;
        *wpg = npatch < 1024 ? 1 : 4;
        *nband = npatch <= 256 ? 8 : (npatch <= 512 ? 4 : (small ? 2 : 1));
    }
    if (*cpl != 1) *nband = 1;
    return HIPCLAW_OK;
}

int step2_launch(const hipclaw_physics* phys, int npatch, const void* q_in, void* q_out,
                 const void* aux, const double* dxdy, const int* patch_ids, int ntiles, double dt,
                 double* cfl_patch, double* reg, const hipclaw_launch* launch, void* stream,
                 const MultiLevel* ml)
{
    int rc = validate_physics(phys);
    if (rc) return rc;
    if (npatch < 0) return set_error(HIPCLAW_ERR_ARG, "npatch < 0");
    if (!patch_ids && ntiles < npatch)
        return set_error(HIPCLAW_ERR_ARG, "ntiles < npatch without a patch list");
    if (npatch == 0) return HIPCLAW_OK;
    // lanes of the kernels that own no patch read tile 0 (their row loads are
    // unconditional), so it must exist
    if (ntiles < 1) return set_error(HIPCLAW_ERR_ARG, "ntiles must be >= 1");
    if (!q_in || !q_out || !dxdy) return set_error(HIPCLAW_ERR_ARG, "NULL buffer");
    if (phys->maux > 0 && !aux) return set_error(HIPCLAW_ERR_ARG, "aux is NULL but maux > 0");
    if (!(dt >= 0)) return set_error(HIPCLAW_ERR_ARG, "dt must be >= 0");
    int cpl, wpg, nband;
    rc = choose_launch(phys, npatch, launch, &cpl, &wpg, &nband);
    if (rc) return rc;
    StepArgs a{};
    a.npatch = npatch;
    a.mx = phys->mx;
    a.my = phys->my;
    a.order2 = phys->order[0];
    a.transverse = phys->order[1];
    for (int w = 0; w < 5; w++) a.mthlim[w] = w < phys->mwaves ? phys->mthlim[w] : 0;
    // the Clawpack 5 inline limiter selects with a computed GO TO over five
    // labels (clawpack5_inlinelimiter.f:59); 6 falls through to minmod
    if (is_claw5(phys->solver))
        for (int w = 0; w < 5; w++)
            if (a.mthlim[w] == 6) a.mthlim[w] = 1;
    a.pp.grav = phys->grav;
    a.pp.sgrav = sqrt(phys->grav);
    a.pp.gamma = phys->gamma;
    a.pp.efix = phys->efix;
    a.dt = dt;
    a.dxdy = dxdy;
    a.patch_ids = patch_ids;
    a.ntiles = ntiles;
    const long long T = (long long)(phys->mx + 4) * (phys->my + 4);
    a.qstride = T * phys->meqn;
    a.astride = T * phys->maux;
    a.reg = reg;
    a.regstride = 8LL * phys->meqn * (phys->mx + phys->my);
    if (reg && phys->mbc != 2) return set_error(HIPCLAW_ERR_UNSUPPORTED, "time sync needs mbc = 2");
    if (ml) {
        if (patch_ids) return set_error(HIPCLAW_ERR_ARG, "multi-level launches take no patch list");
        a.levels = ml->levels;
        a.slot0 = ml->slot0;
        a.push = ml->push;
        a.swirl = ml->swirl;
        a.xylower = ml->xylower;
        a.par = ml->par;
        a.cflmax = ml->cflmax;
        a.reg_fresh = ml->reg_fresh;
        a.reg_lo = ml->reg_lo;
        a.reg_cfaces = ml->reg_cfaces;
        for (int L = 0; L < STEP_MAXL; L++) a.vtl[L] = ml->vtl[L];
        // levels the launch does not cover get q_in, the pool every slot of
        // the launch lives in: the kernels' lanes without a patch read from
        // the tile of slot slot0 through the pointer of level 0
        for (int L = 0; L < STEP_MAXL; L++) {
            a.dtl[L] = ml->dtl[L];
            a.qin_l[L] = ml->qin_l[L] ? ml->qin_l[L] : q_in;
            a.qout_l[L] = ml->qout_l[L];
        }
    }
    // bands of at least two rows each, none empty; an update in place must
    // not overwrite rows that another band of the patch still reads, so it
    // sweeps unbanded (choose_launch gives bands to one column per lane only)
    if (nband < 2 || q_in == q_out) {
        a.nband = 1;
        a.band_rows = a.my;
    } else {
        a.band_rows = (a.my + nband - 1) / nband;
        if (a.band_rows < 2) a.band_rows = 2;
        a.nband = (a.my + a.band_rows - 1) / a.band_rows;
        if (a.nband < 2) a.band_rows = a.my;
        else wpg = a.nband;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (phys->precision == HIPCLAW_FP64)
        return launch_step2_fp64(a, phys->solver, cpl, wpg, q_in, q_out, aux, cfl_patch, s);
    return launch_step2_fp32(a, phys->solver, cpl, wpg, q_in, q_out, aux, cfl_patch, s);
}

}  // namespace hc

extern "C" {

int hipclaw_launch_shape(const hipclaw_physics* phys, int npatch, const hipclaw_launch* launch,
                         hipclaw_launch* out)
{
    int rc = validate_physics(phys);
    if (rc) return rc;
    if (!out || npatch < 0) return set_error(HIPCLAW_ERR_ARG, "bad launch_shape args");
    int cpl, wpg, nband;
    rc = choose_launch(phys, npatch, launch, &cpl, &wpg, &nband);
    if (rc) return rc;
    // the band count the launch really uses: bands of at least two rows
    int rows = nband > 1 ? (phys->my + nband - 1) / nband : phys->my;
    if (rows < 2) rows = 2;
    nband = nband > 1 ? (phys->my + rows - 1) / rows : 1;
    if (nband < 2) nband = 1;
    out->cols_per_lane = cpl;
    out->waves_per_group = nband > 1 ? nband : wpg;
    out->row_bands = nband;
    return HIPCLAW_OK;
}

int hipclaw_max_reduce(const double* x_dev, int n, double* out_dev, void* stream)
{
    if (!x_dev || !out_dev || n < 0) return set_error(HIPCLAW_ERR_ARG, "bad max_reduce args");
    hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(1024), 0, static_cast<hipStream_t>(stream),
                       x_dev, n, out_dev);
    return check_hip(hipGetLastError(), "k_max_reduce launch");
}

int hipclaw_swirl_aux_batch(const hipclaw_physics* phys, int npatch, const double* xylower,
                            const double* dxdy, double t, double dt, double tperiod, void* aux,
                            void* stream)
{
    int rc = validate_physics(phys);
    if (rc) return rc;
    if (solver_family(phys->solver) != HIPCLAW_SOLVER_ADV)
        return set_error(HIPCLAW_ERR_ARG, "swirl aux needs the advection solver");
    if (npatch <= 0) return HIPCLAW_OK;
    if (tperiod == 0.0) return HIPCLAW_OK;  // velocities fixed by setaux
    const double pi2 = 8.0 * atan(1.0);
    const double vt = cos(pi2 * (t + dt / 2.0) / tperiod);
    const size_t lds = sizeof(double) * 2 * (size_t)(phys->mx + 4 + phys->my + 4);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (phys->precision == HIPCLAW_FP64)
        hipLaunchKernelGGL(k_swirl_aux<double>, dim3(npatch), dim3(256), lds, s, npatch, phys->mx, phys->my,
                           xylower, dxdy, vt, static_cast<double*>(aux));
    else
        hipLaunchKernelGGL(k_swirl_aux<float>, dim3(npatch), dim3(256), lds, s, npatch, phys->mx, phys->my,
                           xylower, dxdy, vt, static_cast<float*>(aux));
    return check_hip(hipGetLastError(), "k_swirl_aux launch");
}

int hipclaw_bc2_batch(const hipclaw_physics* phys, int npatch, void* q, const int* patch_ids,
                      int ntiles, const int* mthbc, void* stream)
{
    int rc = validate_physics(phys);
    if (rc) return rc;
    if (npatch <= 0) return HIPCLAW_OK;
    if (!patch_ids && ntiles < npatch)
        return set_error(HIPCLAW_ERR_ARG, "ntiles < npatch without a patch list");
    if (!q || !mthbc) return set_error(HIPCLAW_ERR_ARG, "NULL buffer");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (phys->precision == HIPCLAW_FP64)
        hipLaunchKernelGGL(k_bc2<double>, dim3(npatch), dim3(64), 0, s, npatch, phys->mx, phys->my,
                           phys->meqn, static_cast<double*>(q), patch_ids, ntiles, mthbc);
    else
        hipLaunchKernelGGL(k_bc2<float>, dim3(npatch), dim3(64), 0, s, npatch, phys->mx, phys->my,
                           phys->meqn, static_cast<float*>(q), patch_ids, ntiles, mthbc);
    return check_hip(hipGetLastError(), "k_bc2 launch");
}

}  // extern "C"
