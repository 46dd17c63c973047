// hc_step2_inst.hpp -- instantiates the patch-step kernel for one precision
// (included by hc_step2_fp64.hip / hc_step2_fp32.hip, which differ only in
// the element type and in their floating-point contraction flags).
#pragma once
#include "hc_internal.hpp"

namespace hc {

template <class R, template <class> class SolverT, int CPL, bool MC22, bool SYNC, bool PUSH>
static void launch_k(const StepArgs& a, long long blocks, int wpg, const void* qin, void* qout,
                     const void* aux, double* cfl, hipStream_t s)
{
    // push staging per wave's patches (shared by the row bands of a patch),
    // then the swirl row sines per wave
    const size_t nstage = a.nband > 1 ? 1 : (size_t)wpg;
    size_t lds = PUSH ? sizeof(R) * nstage * a.ppw * 4 * SolverT<R>::ME * (a.mx + a.my) : 0;
    if (a.swirl) lds = ((lds + 7) & ~(size_t)7) + sizeof(double) * (size_t)wpg * a.ppw * 2 * (a.my + 4);
    // row bands are instantiated for one column per lane (the small-launch shape)
    if constexpr (CPL == 1) {
        if (a.nband > 1) {
            hipLaunchKernelGGL((k_step2<R, SolverT, CPL, MC22, SYNC, PUSH, true>), dim3((unsigned)blocks),
                               dim3(HC_WAVE * wpg), lds, s, a, static_cast<const R*>(qin),
                               static_cast<R*>(qout), static_cast<const R*>(aux), cfl);
            return;
        }
    }
    hipLaunchKernelGGL((k_step2<R, SolverT, CPL, MC22, SYNC, PUSH, false>), dim3((unsigned)blocks),
                       dim3(HC_WAVE * wpg), lds, s, a, static_cast<const R*>(qin),
                       static_cast<R*>(qout), static_cast<const R*>(aux), cfl);
}

template <class R, template <class> class SolverT, int CPL>
static int launch_one(const StepArgs& a0, int wpg, const void* qin, void* qout,
                      const void* aux, double* cfl, hipStream_t s)
{
    constexpr int MW = SolverT<R>::MW;
    StepArgs a = a0;
    const int NX = a.mx + 4;
    if (CPL != 1 && a.nband > 1)  // choose_launch gives bands to one column per lane only
        return set_error(HIPCLAW_ERR_ARG, "row bands need one column per lane");
    a.lpp = (NX + CPL - 1) / CPL;
    if (a.lpp > HC_WAVE) return set_error(HIPCLAW_ERR_UNSUPPORTED, "patch too wide for one wave");
    a.ppw = HC_WAVE / a.lpp;
    if (a.push) {
        // the push stage must fit a workgroup's LDS: fewer waves per group
        // (fewer row bands) if needed
        const size_t stg = sizeof(R) * (size_t)a.ppw * 4 * SolverT<R>::ME * (a.mx + a.my);
        const size_t sw = a.swirl ? sizeof(double) * (size_t)a.ppw * 2 * (a.my + 4) + 8 : 0;
        if (a.nband > 1) {
            int nb = a.nband;
            while (nb > 1 && stg + sw * nb > 65536) nb--;
            if (nb < a.nband) {
                a.band_rows = nb > 1 ? (a.my + nb - 1) / nb : a.my;
                a.nband = nb > 1 ? (a.my + a.band_rows - 1) / a.band_rows : 1;
                wpg = a.nband;
            }
        }
        while (a.nband < 2 && wpg > 1 && (stg + sw) * wpg > 65536) wpg--;
        if (stg + sw > 65536) return set_error(HIPCLAW_ERR_UNSUPPORTED, "patches too large for the ghost push");
    }
    if (a.swirl && (!(SolverT<R>::MAUX > 0) || SolverT<R>::FW)) return set_error(HIPCLAW_ERR_ARG, "the swirl field needs the advection solver");
    if (a.swirl && !a.xylower) return set_error(HIPCLAW_ERR_ARG, "the swirl field needs the patch corners");
    const long long waves = (a.npatch + a.ppw - 1) / a.ppw;
    // banded: one workgroup of nband waves per wave's worth of patches
    if (a.nband > 1 && (wpg != a.nband || a.nband > 8 || a.band_rows < 2))
        return set_error(HIPCLAW_ERR_ARG, "row bands: 2..8 waves of one workgroup, 2 rows or more each");
    const long long blocks = a.nband > 1 ? waves : (waves + wpg - 1) / wpg;
    if (blocks == 0) return HIPCLAW_OK;
    bool mc22 = a.order2 == 2 && a.transverse == 2;
    for (int w = 0; w < MW; w++) mc22 = mc22 && a.mthlim[w] == 4;
    if (a.reg && a.push) return set_error(HIPCLAW_ERR_ARG, "ghost push and time-sync registers exclude each other");
    if (a.push && !a.levels) return set_error(HIPCLAW_ERR_ARG, "ghost push needs slot addressing");
    if (a.reg) {
        if constexpr (SolverT<R>::HAS_CONS) {
            if (mc22)
                launch_k<R, SolverT, CPL, true, true, false>(a, blocks, wpg, qin, qout, aux, cfl, s);
            else
                launch_k<R, SolverT, CPL, false, true, false>(a, blocks, wpg, qin, qout, aux, cfl, s);
        } else {
            return set_error(HIPCLAW_ERR_UNSUPPORTED,
                             "time-sync registers need a conservative flux (SWE only)");
        }
    } else if (a.push) {
        if (mc22)
            launch_k<R, SolverT, CPL, true, false, true>(a, blocks, wpg, qin, qout, aux, cfl, s);
        else
            launch_k<R, SolverT, CPL, false, false, true>(a, blocks, wpg, qin, qout, aux, cfl, s);
    } else if (mc22) {
        launch_k<R, SolverT, CPL, true, false, false>(a, blocks, wpg, qin, qout, aux, cfl, s);
    } else {
        launch_k<R, SolverT, CPL, false, false, false>(a, blocks, wpg, qin, qout, aux, cfl, s);
    }
    return check_hip(hipGetLastError(), "k_step2 launch");
}

template <class R>
static int launch_step2_t(const StepArgs& a, int solver_id, int cpl, int wpg, const void* qin,
                          void* qout, const void* aux, double* cfl, hipStream_t s)
{
    const int solver = solver_family(solver_id);
    // the Clawpack 5 solvers differ from the 4.6 ones only in FP64 rounding
    // (operation order); the FP32 kernels serve both
    if constexpr (sizeof(R) == 8) {
        if (is_claw5(solver_id)) {
            if (cpl == 1) {
                switch (solver) {
                case SOLVER_ADV: return launch_one<R, Adv, 1>(a, wpg, qin, qout, aux, cfl, s);
                case SOLVER_SWE: return launch_one<R, Swe5, 1>(a, wpg, qin, qout, aux, cfl, s);
                case SOLVER_EULER5: return launch_one<R, Euler5C5, 1>(a, wpg, qin, qout, aux, cfl, s);
                }
            } else if (cpl == 2) {
                switch (solver) {
                case SOLVER_ADV: return launch_one<R, Adv, 2>(a, wpg, qin, qout, aux, cfl, s);
                case SOLVER_SWE: return launch_one<R, Swe5, 2>(a, wpg, qin, qout, aux, cfl, s);
                case SOLVER_EULER5: return launch_one<R, Euler5C5, 2>(a, wpg, qin, qout, aux, cfl, s);
                }
            }
            return set_error(HIPCLAW_ERR_ARG, "unsupported solver / cols_per_lane");
        }
    }
    if (cpl == 1) {
        switch (solver) {
        case SOLVER_ADV: return launch_one<R, Adv, 1>(a, wpg, qin, qout, aux, cfl, s);
        case SOLVER_SWE: return launch_one<R, Swe, 1>(a, wpg, qin, qout, aux, cfl, s);
        case SOLVER_EULER5: return launch_one<R, Euler5, 1>(a, wpg, qin, qout, aux, cfl, s);
        case SOLVER_EULER_HLLE: return launch_one<R, EulerHlle, 1>(a, wpg, qin, qout, aux, cfl, s);
        case SOLVER_TRANSPORT_FW: return launch_one<R, FwCons, 1>(a, wpg, qin, qout, aux, cfl, s);
        }
    } else if (cpl == 2) {
        switch (solver) {
        case SOLVER_ADV: return launch_one<R, Adv, 2>(a, wpg, qin, qout, aux, cfl, s);
        case SOLVER_SWE: return launch_one<R, Swe, 2>(a, wpg, qin, qout, aux, cfl, s);
        case SOLVER_EULER5: return launch_one<R, Euler5, 2>(a, wpg, qin, qout, aux, cfl, s);
        case SOLVER_EULER_HLLE: return launch_one<R, EulerHlle, 2>(a, wpg, qin, qout, aux, cfl, s);
        case SOLVER_TRANSPORT_FW: return launch_one<R, FwCons, 2>(a, wpg, qin, qout, aux, cfl, s);
        }
    }
    return set_error(HIPCLAW_ERR_ARG, "unsupported solver / cols_per_lane");
}

}  // namespace hc
