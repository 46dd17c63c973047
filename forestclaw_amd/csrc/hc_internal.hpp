// hc_internal.hpp -- host-side declarations shared by the hipclaw translation
// units (not part of the public C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/hipclaw.h"
#include "hc_step2.hpp"

namespace hc {

// Solver id = family (HIPCLAW_SOLVER_*) | optional HIPCLAW_CLAW5 flag.
inline int solver_family(int solver) { return solver & ~HIPCLAW_CLAW5; }
inline bool is_claw5(int solver) { return (solver & HIPCLAW_CLAW5) != 0; }

// Host patch arrays.  The Clawpack 4.6 layout q(i, j, m) is the library's
// tile layout; the Clawpack 5 layout q(m, i, j) (fclaw2d_clawpatch5) has the
// fields fastest.  nf fields over NX*NY cells.
template <class T>
inline void patch_from_host(const double* src, T* dst, int nf, long long cells, bool claw5)
{
    if (!claw5) {
        for (long long i = 0; i < cells * nf; i++) dst[i] = (T)src[i];
        return;
    }
    for (long long c = 0; c < cells; c++)
        for (int m = 0; m < nf; m++) dst[m * cells + c] = (T)src[c * nf + m];
}
template <class T>
inline void patch_to_host(const T* src, double* dst, int nf, long long cells, bool claw5)
{
    if (!claw5) {
        for (long long i = 0; i < cells * nf; i++) dst[i] = (double)src[i];
        return;
    }
    for (long long c = 0; c < cells; c++)
        for (int m = 0; m < nf; m++) dst[c * nf + m] = (double)src[m * cells + c];
}

// Records a message for hipclaw_last_error() and returns code.
int set_error(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

// Kernel launchers, one per precision translation unit.  Return HIPCLAW_*.
int launch_step2_fp64(const StepArgs& a, int solver, int cpl, int wpg, const void* qin,
                      void* qout, const void* aux, double* cfl, hipStream_t s);
int launch_step2_fp32(const StepArgs& a, int solver, int cpl, int wpg, const void* qin,
                      void* qout, const void* aux, double* cfl, hipStream_t s);

// Validates a physics description; returns HIPCLAW_OK or an error code with
// the message set.
int validate_physics(const hipclaw_physics* p);

// Effective launch shape (columns per lane, waves per workgroup, row bands)
// of a patch-step launch of npatch patches for a caller's shape (NULL: the
// library default).  Host-only.
int choose_launch(const hipclaw_physics* phys, int npatch, const hipclaw_launch* launch, int* cpl,
                  int* wpg, int* nband);

// Patch-step launch behind hipclaw_step2_batch(_sync); with ml, one launch
// over the slots [ml->slot0, ml->slot0 + npatch) of several levels (q_in /
// q_out unused, dxdy / aux / cfl / reg addressed by slot).
struct MultiLevel {
    const int* levels;
    int slot0;
    double dtl[STEP_MAXL];
    const void* qin_l[STEP_MAXL];
    void* qout_l[STEP_MAXL];
    const int* push;  // optional same-level ghost push table (8 ints per slot)
    // advection: the swirl velocity field evaluated in the step kernel
    // (xylower by slot, vt = cos(2 pi (t + dt/2) / T) per level) instead of aux
    int swirl;
    const double* xylower;
    double vtl[STEP_MAXL];
    // optional device parameter block: dt of level L at par[L], vt at
    // par[STEP_MAXL + L] (graph replays; see StepArgs::par)
    const double* par;
    // optional atomic launch maximum of the patch Courant numbers (StepArgs::cflmax)
    unsigned long long* cflmax;
    // time-sync registers known to be zero (StepArgs::reg_fresh)
    int reg_fresh, reg_lo;
    const int* reg_cfaces;
};
// Courant-number slots of a coarse step (hc_api.hip): the maximum of n
// patch values stored into one slot as double bits; the maximum of n slots
// into out, clearing the slots.
int cfl_slot_reduce(const double* x, int n, unsigned long long* slot, hipStream_t s);
int cfl_final(unsigned long long* acc, int n, double* out, hipStream_t s);
int step2_launch(const hipclaw_physics* phys, int npatch, const void* q_in, void* q_out,
                 const void* aux, const double* dxdy, const int* patch_ids, int ntiles, double dt,
                 double* cfl_patch, double* reg, const hipclaw_launch* launch, void* stream,
                 const MultiLevel* ml);

// Source term after the patch step (hc_source.hip) over the slots
// [slot0, slot0 + n): q_l / dtl indexed by levels[slot] (levels NULL: index 0);
// the radius from aux field 1 when aux is given, else from xylower / dxdy.
struct Src2Args {
    int n, slot0;
    const int* levels;
    void* q_l[STEP_MAXL];
    double dtl[STEP_MAXL];
    const double* xylower;
    const double* dxdy;
    int mx, my, meqn, maux;
    double gamma1;
    const double* par;  // optional device parameter block: dt of level L at par[L]
};
int src2_launch(const Src2Args& a, int kind, int precision, const void* aux, hipStream_t s);

// fort.tXXXX / fort.qXXXX frame of n patches (host arrays, q in the tile
// layout T*meqn per patch, mbc = 2); hc_output.cpp
int write_frame(const char* dir, int iframe, double time, int mi, int mx, int my, int meqn, int maux, int n,
                const int* lev, const int* gx, const int* gy, const double* xl, const double* yl,
                const double* dx, const double* dy, const double* q);

}  // namespace hc
