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
//   lim_ok / divl  the limiter ratio r = dot / |wave|^2.  FP64 limits every
//         wave with a nonzero norm, as the reference does.  FP32 limits waves
//         whose squared norm is a normal number and divides with the hardware
//         reciprocal; a wave with a denormal squared norm (|wave| < 1.1e-19)
//         is left unlimited, which moves q by less than dt/dx*|s|*1.1e-19.
template <class R> struct M;
template <> struct M<double> {
    __device__ static inline double sqrt_(double x) { return __builtin_sqrt(x); }
    __device__ static inline double div(double a, double b) { return a / b; }
    __device__ static inline double divp(double a, double b) { return a / b; }
    __device__ static inline double abs_(double x) { return __builtin_fabs(x); }
    __device__ static inline bool lim_ok(double wn) { return wn != 0.0; }
    __device__ static inline double divl(double a, double wn) { return a / wn; }
};
template <> struct M<float> {
    __device__ static inline float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
    __device__ static inline float div(float a, float b) { return a / b; }
    __device__ static inline float divp(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static inline float abs_(float x) { return __builtin_fabsf(x); }
    __device__ static inline bool lim_ok(float wn) { return wn >= 1.17549435e-38f; }
    __device__ static inline float divl(float a, float wn) { return a * __builtin_amdgcn_rcpf(wn); }
};

template <class R> __device__ inline R dmin(R a, R b) { return a < b ? a : b; }
template <class R> __device__ inline R dmax(R a, R b) { return a > b ? a : b; }
// FP32: the hardware v_min_f32 / v_max_f32 (one instruction instead of a
// compare and a select; equal for the finite data of a step up to the sign
// of a zero).  FP64 keeps the reference's comparisons.
template <> __device__ inline float dmin<float>(float a, float b) { return __builtin_fminf(a, b); }
template <> __device__ inline float dmax<float>(float a, float b) { return __builtin_fmaxf(a, b); }
// min(x, 0), max(x, 0) and the maximum of two non-negative values.  FP32: the
// integer min / max of the bit patterns (a negative float, -0 included, is a
// negative integer; non-negative floats order like their bits), one
// instruction each with no operand quieting: IEEE-mode v_min_f32 / v_max_f32
// need operands known to be canonical, and values made in another basic
// block get an extra v_max_f32 x, x first.  FP64: the comparisons above.
template <class R> __device__ inline R zmin(R x) { return dmin(x, R(0)); }
template <class R> __device__ inline R zmax(R x) { return dmax(x, R(0)); }
template <class R> __device__ inline R pmax(R a, R b) { return dmax(a, b); }
// max(0, x) with the zero first, as the limiters write it (FP64 keeps the
// comparison order, which decides the sign of a zero result)
template <class R> __device__ inline R zmax0(R x) { return dmax(R(0), x); }
template <> __device__ inline float zmin<float>(float x) { return __int_as_float(min(__float_as_int(x), 0)); }
template <> __device__ inline float zmax<float>(float x) { return __int_as_float(max(__float_as_int(x), 0)); }
template <> __device__ inline float zmax0<float>(float x) { return zmax<float>(x); }
template <> __device__ inline float pmax<float>(float a, float b)
{
    return __int_as_float(max(__float_as_int(a), __float_as_int(b)));
}

// Whole-wave neighbour exchange (lane l receives lane l-1 / l+1) with the
// GFX9 DPP wave_shr:1 / wave_shl:1 row controls: one VALU move, no LDS round
// trip.  Lanes at the wave edges get zero (bound_ctrl on, so the move needs
// no copy of an old value first); callers never consume those.  Called only
// in wave-uniform control flow, so every source lane is active.
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;
template <int CTRL> __device__ inline int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, true);
}
template <int CTRL> __device__ inline float dpp_f(float v)
{
    return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL> __device__ inline double dpp_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<CTRL>((int)(b & 0xffffffffll));
    const int hi = dpp_i<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline float shl_up(float v) { return dpp_f<DPP_WAVE_SHR1>(v); }
__device__ inline float shl_dn(float v) { return dpp_f<DPP_WAVE_SHL1>(v); }
__device__ inline double shl_up(double v) { return dpp_d<DPP_WAVE_SHR1>(v); }
__device__ inline double shl_dn(double v) { return dpp_d<DPP_WAVE_SHL1>(v); }

// Non-negative doubles order like their bit patterns.
__device__ inline unsigned long long dbits(double x) { return __double_as_longlong(x); }

}  // namespace hc
