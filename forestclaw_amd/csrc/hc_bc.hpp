// hc_bc.hpp -- clawpack46_bc2_default on one patch, run by one workgroup
// (src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_bc2_default.f):
// left / right over the full height, then bottom / top over the full width,
// whose corner cells read the left / right ghost cells of the first pass.
#pragma once
#include <hip/hip_runtime.h>

#include "hc_amr.hpp"

namespace hc {

// qp: the patch's tile (q[(m*NY + j)*NX + i], two ghost layers).  Every
// thread of the workgroup must call it (one barrier between the passes).
// mthbc 1 extrapolates, 3 reflects (normal momentum negated: field 1 across
// x faces, 2 across y faces); other values leave the face alone.
template <class R>
__device__ inline void bc_patch(const BcOp& op, R* __restrict__ qp, int mx, int my, int meqn)
{
    constexpr int mb = 2;
    const int NX = mx + 2 * mb, NY = my + 2 * mb;
    const int i1 = mb, imx = mx + mb - 1, j1 = mb, jmy = my + mb - 1;
#define QB(i, j, m) qp[((long long)(m) * NY + (j)) * NX + (i)]
    for (int t = threadIdx.x; t < NY * meqn; t += blockDim.x) {
        const int m = t / NY, j = t - m * NY;
        for (int ibc = 1; ibc <= mb; ibc++) {
            if (op.mthbc[0] == 1) QB(i1 - ibc, j, m) = QB(i1, j, m);
            else if (op.mthbc[0] == 3) QB(i1 - ibc, j, m) = (m == 1) ? -QB(i1 + ibc - 1, j, m) : QB(i1 + ibc - 1, j, m);
            if (op.mthbc[1] == 1) QB(imx + ibc, j, m) = QB(imx, j, m);
            else if (op.mthbc[1] == 3) QB(imx + ibc, j, m) = (m == 1) ? -QB(imx + 1 - ibc, j, m) : QB(imx + 1 - ibc, j, m);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < NX * meqn; t += blockDim.x) {
        const int m = t / NX, i = t - m * NX;
        for (int jbc = 1; jbc <= mb; jbc++) {
            if (op.mthbc[2] == 1) QB(i, j1 - jbc, m) = QB(i, j1, m);
            else if (op.mthbc[2] == 3) QB(i, j1 - jbc, m) = (m == 2) ? -QB(i, j1 + jbc - 1, m) : QB(i, j1 + jbc - 1, m);
            if (op.mthbc[3] == 1) QB(i, jmy + jbc, m) = QB(i, jmy, m);
            else if (op.mthbc[3] == 3) QB(i, jmy + jbc, m) = (m == 2) ? -QB(i, jmy + 1 - jbc, m) : QB(i, jmy + 1 - jbc, m);
        }
    }
#undef QB
}

// CflFin's reduction by one workgroup (blockDim a power of two, at most 256):
// the slots hold the bits of non-negative doubles, which order like the
// doubles, so the integer maximum is the maximum Courant number.
__device__ inline void cfl_final_block(const CflFin& c)
{
    __shared__ unsigned long long s[256];
    unsigned long long v = 0;
    for (int i = threadIdx.x; i < c.n; i += blockDim.x) {
        const unsigned long long a = c.acc[i];
        v = a > v ? a : v;
        c.acc[i] = 0;
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && s[threadIdx.x + w] > s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *c.out = __longlong_as_double((long long)s[0]);
}

}  // namespace hc
