// FP32 instantiation of the patch-step kernel (bulk Riemann / limiter / flux
// arithmetic in single precision; FMA contraction allowed).
#include "hc_step2_inst.hpp"

namespace hc {
int launch_step2_fp32(const StepArgs& a, int solver, int cpl, int wpg, const void* qin,
                      void* qout, const void* aux, double* cfl, hipStream_t s)
{
    return launch_step2_t<float>(a, solver, cpl, wpg, qin, qout, aux, cfl, s);
}
}  // namespace hc
