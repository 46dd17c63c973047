// FP64 instantiation of the patch-step kernel.  Built with -ffp-contract=off
// so every operation rounds as in the reference's FP64 Fortran.
#include "hc_step2_inst.hpp"

namespace hc {
int launch_step2_fp64(const StepArgs& a, int solver, int cpl, int wpg, const void* qin,
                      void* qout, const void* aux, double* cfl, hipStream_t s)
{
    return launch_step2_t<double>(a, solver, cpl, wpg, qin, qout, aux, cfl, s);
}
}  // namespace hc
