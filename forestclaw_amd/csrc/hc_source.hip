// hc_source.hip -- source terms applied after the patch step, as
// clawpack46_update calls claw46_vt->src2 after step2 with the same t and dt
// (reference src/solvers/fc2d_clawpack4.6/fc2d_clawpack46.cpp:402-444, the
// src2 wrapper :186-215).
//
// HIPCLAW_SRC2_AXISYMMETRIC is the shock-bubble application's source
// (applications/clawpack/euler/2d/shockbubble/user_4.6/src2.f:1-77): the
// geometric source of the axisymmetric Euler equations about y = 0, with
// y playing the radius (setaux.f stores rad = ylower + (j-1/2) dy), integrated
// by a two-stage Runge-Kutta step over the interior cells, fields 1-4 (the
// tracer has no source).  Pointwise FP64 in the reference's operation order;
// this file is built without FMA contraction, so the values are bitwise the
// reference's.
//
// One workgroup per patch, threads over the interior cells; the patch tile is
// read and written once (HBM-bound, 2 x 4 fields x w bytes per cell).
#include <hip/hip_runtime.h>

#include "../../include/hipclaw.h"
#include "hc_internal.hpp"

namespace hc {

namespace {

constexpr int MBC = 2;

template <class R>
__global__ void k_src2_axisym(Src2Args a, const R* __restrict__ aux)
{
    const int k = blockIdx.x;
    const int slot = a.slot0 + k;
    const int L = a.levels ? a.levels[slot] : 0;
    const double dt = a.par ? a.par[L] : a.dtl[L];
    R* q = static_cast<R*>(a.q_l[L]);
    const int NX = a.mx + 2 * MBC, NY = a.my + 2 * MBC;
    const long long T = (long long)NX * NY;
    q += (long long)slot * T * a.meqn;
    const double ylower = a.xylower ? a.xylower[2 * slot + 1] : 0.0;
    const double dy = a.dxdy ? a.dxdy[2 * slot + 1] : 0.0;
    const double gamma1 = a.gamma1;
    const double dt2 = dt / 2.0;
    const int ndim = 2;
    for (int c = threadIdx.x; c < a.mx * a.my; c += blockDim.x) {
        const int ii = c % a.mx, jj = c / a.mx;
        const long long o = (long long)(jj + MBC) * NX + (ii + MBC);
        double rad;
        if (aux)
            rad = (double)aux[(long long)slot * T * a.maux + o];
        else
            rad = ylower + ((double)(jj + 1) - 0.5) * dy;
        const double q1 = (double)q[o], q2 = (double)q[T + o], q3 = (double)q[2 * T + o],
                     q4 = (double)q[3 * T + o];
        double rho = q1;
        double u = q2 / q1;
        double v = q3 / q1;
        double press = gamma1 * (q4 - 0.5 * rho * (u * u + v * v));
        const double f = dt2 * (ndim - 1);
        const double s1 = q1 - f / rad * q3;
        const double s2 = q2 - f / rad * (rho * u * v);
        const double s3 = q3 - f / rad * (rho * v * v);
        const double s4 = q4 - f / rad * v * (q4 + press);
        rho = s1;
        u = s2 / s1;
        v = s3 / s1;
        press = gamma1 * (s4 - 0.5 * rho * (u * u + v * v));
        const double g = dt * (ndim - 1);
        q[o] = R(q1 - g / rad * s3);
        q[T + o] = R(q2 - g / rad * (rho * u * v));
        q[2 * T + o] = R(q3 - g / rad * (rho * v * v));
        q[3 * T + o] = R(q4 - g / rad * v * (s4 + press));
    }
}

}  // namespace

int src2_launch(const Src2Args& a, int kind, int precision, const void* aux, hipStream_t s)
{
    if (kind == HIPCLAW_SRC2_NONE || a.n <= 0) return HIPCLAW_OK;
    if (kind != HIPCLAW_SRC2_AXISYMMETRIC) return set_error(HIPCLAW_ERR_ARG, "unknown source term");
    if (a.meqn < 4) return set_error(HIPCLAW_ERR_ARG, "axisymmetric source needs meqn >= 4");
    if (!aux && (!a.xylower || !a.dxdy)) return set_error(HIPCLAW_ERR_ARG, "source needs aux or patch geometry");
    if (aux && a.maux < 1) return set_error(HIPCLAW_ERR_ARG, "source reads the radius from aux field 1");
    if (precision == HIPCLAW_FP64)
        hipLaunchKernelGGL(k_src2_axisym<double>, dim3(a.n), dim3(256), 0, s, a,
                           static_cast<const double*>(aux));
    else
        hipLaunchKernelGGL(k_src2_axisym<float>, dim3(a.n), dim3(256), 0, s, a,
                           static_cast<const float*>(aux));
    return check_hip(hipGetLastError(), "k_src2_axisym launch");
}

}  // namespace hc

extern "C" int hipclaw_src2_batch(const hipclaw_physics* phys, int kind, int npatch, void* q, const void* aux,
                                  const double* xylower, const double* dxdy, double t, double dt,
                                  void* stream)
{
    int rc = hc::validate_physics(phys);
    if (rc) return rc;
    (void)t;
    if (npatch < 0 || (npatch > 0 && !q)) return hc::set_error(HIPCLAW_ERR_ARG, "bad source batch");
    hc::Src2Args a{};
    a.n = npatch;
    a.slot0 = 0;
    a.levels = nullptr;
    a.q_l[0] = q;
    a.dtl[0] = dt;
    a.xylower = xylower;
    a.dxdy = dxdy;
    a.mx = phys->mx;
    a.my = phys->my;
    a.meqn = phys->meqn;
    a.maux = phys->maux;
    a.gamma1 = phys->gamma - 1.0;
    return hc::src2_launch(a, kind, phys->precision, aux, static_cast<hipStream_t>(stream));
}
