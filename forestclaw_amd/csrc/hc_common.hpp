// hc_common.hpp -- device helpers shared by the hipclaw kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HC_WAVE 64

namespace hc {

// Arithmetic helpers.  The FP64 path uses IEEE division and square root and
// is compiled without FMA contraction so that it reproduces the reference's
// FP64 arithmetic operation for operation.
//   div   general division, IEEE in both precisions (divisors may be tiny or
//         denormal, e.g. the limiter's squared wave norm ahead of a front)
//   divp  division by a positive normal number (sqrt of a depth or density,
//         2*sound speed, 1+|r|): the FP32 path uses the hardware reciprocal
//         (about 1 ulp), far inside its stated error budget
//   sqrt_ FP32: hardware square root (about 1 ulp) of depths / pressures
template <class R> struct M;
template <> struct M<double> {
    __device__ static inline double sqrt_(double x) { return __builtin_sqrt(x); }
    __device__ static inline double div(double a, double b) { return a / b; }
    __device__ static inline double divp(double a, double b) { return a / b; }
    __device__ static inline double abs_(double x) { return __builtin_fabs(x); }
};
template <> struct M<float> {
    __device__ static inline float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
    __device__ static inline float div(float a, float b) { return a / b; }
    __device__ static inline float divp(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static inline float abs_(float x) { return __builtin_fabsf(x); }
};

template <class R> __device__ inline R dmin(R a, R b) { return a < b ? a : b; }
template <class R> __device__ inline R dmax(R a, R b) { return a > b ? a : b; }

// Whole-wave neighbour exchange (lane l receives lane l-1 / l+1).  Lanes at
// the wave edges get their own value back; callers never consume those.
__device__ inline float shl_up(float v) { return __shfl_up(v, 1, HC_WAVE); }
__device__ inline float shl_dn(float v) { return __shfl_down(v, 1, HC_WAVE); }
__device__ inline double shl_up(double v) { return __shfl_up(v, 1, HC_WAVE); }
__device__ inline double shl_dn(double v) { return __shfl_down(v, 1, HC_WAVE); }

// Non-negative doubles order like their bit patterns.
__device__ inline unsigned long long dbits(double x) { return __double_as_longlong(x); }

}  // namespace hc
