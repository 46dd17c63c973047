// hc_cnn.hip -- learned refinement indicator: a 4-layer 3x3 CNN on patch
// tiles, trained and run with hand-written BF16 MFMA kernels (FP32
// accumulate) for gfx950.
//
// Network (fixed): input 36x36 tile (32x32 interior + 2 ghost layers), NHWC,
// 8 channels (normalised h, hu, hv, zero padding) -> conv 8->16 -> ReLU ->
// conv 16->32 -> ReLU -> conv 32->32 -> ReLU -> conv 32->16 -> ReLU -> global
// average pool -> linear 16->2 (logits: keep, refine).  Convolutions are 3x3,
// zero 'same' padding, with bias.  It replaces the per-patch tag pass of
// fclaw2d_regrid (src/fclaw2d_regrid.c:45-72); the plain gradient tag stays
// as fallback and judge.
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16.  Lane layout
// of that instruction (A 16x32, B 32x16, D 16x16 per wave): lane l holds
// A[l%16][8*(l/16) + i], B[8*(l/16) + i][l%16], D[4*(l/16) + i][l%16].
// With channels-last activations and k = tap*Cin + ci, one A or B fragment
// is 8 consecutive channels: a single 16-byte load.
//   forward      M = pixels, N = Cout, K = 9*Cin   (bias + ReLU fused)
//   data grad    M = pixels, N = Cin,  K = 9*Cout  (ReLU mask fused, W^T)
//   weight grad  M = Cout,   N = 9*Cin, K = pixels (per-wave partial sums
//                in registers, bias fused; the pixel-major fragments come
//                from 16-byte channel loads transposed through LDS; the
//                partials are summed in a fixed order, so gradients are
//                deterministic)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "../../include/hipclaw_cnn.h"
#include "hc_amr.hpp"
#include "hc_internal.hpp"

using hc::check_hip;
using hc::set_error;

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TW = 36;             // tile width = height
constexpr int NPIX = TW * TW;      // 1296 = 81 blocks of 16 pixels
constexpr int PBLK = NPIX / 16;    // 81
constexpr int C0 = 8;              // input channels (3 used)

// layer shapes
constexpr int CIN[4] = {8, 16, 32, 32};
constexpr int COUT[4] = {16, 32, 32, 16};

struct ParamLayout {
    long long w[4], b[4], fcw, fcb, total;
    ParamLayout()
    {
        long long o = 0;
        for (int l = 0; l < 4; l++) {
            w[l] = o;
            o += (long long)COUT[l] * 9 * CIN[l];
            b[l] = o;
            o += COUT[l];
        }
        fcw = o;
        o += 2 * 16;
        fcb = o;
        o += 2;
        total = o;
    }
};

__device__ inline bf16x8 ld8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ inline bf16x8 zero8()
{
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; i++) z[i] = (bf16)0.0f;
    return z;
}

// ---------------------------------------------------------------------------
// Forward convolution (+ bias + ReLU).  One wave = 16 pixels x COUT.
// in: [.., NPIX, CIN] (tile t of the batch reads tile idx[t] when idx given)
// W: bf16 [COUT][9*CIN]; out: [n, NPIX, COUT]
// ---------------------------------------------------------------------------
template <int CI, int CO>
__global__ void __launch_bounds__(256)
k_conv_fwd(const bf16* __restrict__ in, const int* __restrict__ idx, const bf16* __restrict__ W,
           const float* __restrict__ bias, bf16* __restrict__ out, int n)
{
    constexpr int K = 9 * CI, NCH = (K + 31) / 32, NB = CO / 16;
    const int lane = threadIdx.x & 63;
    const long long blk = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (blk >= (long long)n * PBLK) return;  // whole wave leaves together
    const int t = (int)(blk / PBLK), pb = (int)(blk - (long long)t * PBLK);
    const bf16* src = in + (long long)(idx ? idx[t] : t) * NPIX * CI;
    const int m = lane & 15, kq = lane >> 4;
    const int p = pb * 16 + m, y = p / TW, x = p - (p / TW) * TW;
    f32x4 acc[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ch = 0; ch < NCH; ch++) {
        const int k0 = ch * 32 + kq * 8;
        bf16x8 a = zero8();
        if (k0 < K) {
            const int tap = k0 / CI, ci = k0 - tap * CI;
            const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
            if (yy >= 0 && yy < TW && xx >= 0 && xx < TW) a = ld8(src + (yy * TW + xx) * CI + ci);
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const bf16x8 w = k0 < K ? ld8(W + (long long)(b * 16 + m) * K + k0) : zero8();
            acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, acc[b], 0, 0, 0);
        }
    }
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int co = b * 16 + m;
        const float bb = bias[co];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int pp = pb * 16 + kq * 4 + i;
            const float v = acc[b][i] + bb;
            out[((long long)t * NPIX + pp) * CO + co] = (bf16)(v > 0.f ? v : 0.f);
        }
    }
}

// ---------------------------------------------------------------------------
// Data gradient of a convolution: din = conv^T(dout * (out > 0), W).
// WT: bf16 [CIN][9][COUT]; din: [n, NPIX, CIN]
// ---------------------------------------------------------------------------
template <int CI, int CO>
__global__ void __launch_bounds__(256)
k_conv_bwd_data(const bf16* __restrict__ dout, const bf16* __restrict__ out,
                const bf16* __restrict__ WT, bf16* __restrict__ din, int n)
{
    constexpr int K = 9 * CO, NCH = (K + 31) / 32, NB = CI / 16;
    const int lane = threadIdx.x & 63;
    const long long blk = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (blk >= (long long)n * PBLK) return;
    const int t = (int)(blk / PBLK), pb = (int)(blk - (long long)t * PBLK);
    const bf16* g = dout + (long long)t * NPIX * CO;
    const bf16* o = out + (long long)t * NPIX * CO;
    const int m = lane & 15, kq = lane >> 4;
    const int p = pb * 16 + m, y = p / TW, x = p - (p / TW) * TW;
    f32x4 acc[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) acc[b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ch = 0; ch < NCH; ch++) {
        const int k0 = ch * 32 + kq * 8;
        const int tap = k0 / CO, co = k0 - tap * CO;
        const int yy = y - (tap / 3 - 1), xx = x - (tap % 3 - 1);
        bf16x8 a = zero8();
        if (k0 < K && yy >= 0 && yy < TW && xx >= 0 && xx < TW) {
            const bf16x8 gv = ld8(g + (yy * TW + xx) * CO + co);
            const bf16x8 ov = ld8(o + (yy * TW + xx) * CO + co);
#pragma unroll
            for (int i = 0; i < 8; i++) a[i] = (float)ov[i] > 0.f ? gv[i] : (bf16)0.0f;
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const bf16x8 w = k0 < K ? ld8(WT + (long long)(b * 16 + m) * K + k0) : zero8();
            acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, acc[b], 0, 0, 0);
        }
    }
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int pp = pb * 16 + kq * 4 + i;
            din[((long long)t * NPIX + pp) * CI + b * 16 + m] = (bf16)acc[b][i];
        }
}

// Weight and bias gradient, deterministic: every wave of the grid owns the
// 32-pixel chunks w, w + W, ... (chunk = tile * WG_NCHUNK + c, pixels c*32 ..
// c*32+31 of the tile) and accumulates the whole dW[CO][9*CIN] plus the bias
// sums in registers; it writes them to its row of part[W][CO*K + CO], which
// k_wgrad_reduce sums in a fixed order.  The MFMA reduction runs over pixels,
// so an operand fragment is 8 pixels of one channel: each lane loads 8
// consecutive channels of one pixel (one 16-byte load from the channels-last
// activations), and a transpose through the wave's own LDS hands every lane
// its 8 pixels of one channel.
constexpr int WG_NCHUNK = (NPIX + 31) / 32;  // 41 (the last chunk is half empty)
constexpr int WG_WAVES = 1024;               // waves (partial rows) per launch

template <int CI, int CO>
__global__ void __launch_bounds__(256)
k_conv_wgrad(const bf16* __restrict__ in, const int* __restrict__ idx, const bf16* __restrict__ dout,
             const bf16* __restrict__ out, float* __restrict__ part, int n)
{
    constexpr int K = 9 * CI, MB = CO / 16, NBT = (K + 15) / 16;
    static_assert(CI % 8 == 0 && CO % 16 == 0, "channel counts");
    __shared__ __align__(16) bf16 sA[4][MB][32][16];
    __shared__ __align__(16) bf16 sB[4][32][16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int gw = blockIdx.x * 4 + wv;
    const int m = lane & 15, kq = lane >> 4;
    const int li = m >> 1, lh = m & 1;  // loading role: pixel kq*8 + li, channel half lh
    f32x4 acc[MB][NBT];
    float bsum[MB];
#pragma unroll
    for (int b = 0; b < MB; b++) {
        bsum[b] = 0.f;
#pragma unroll
        for (int j = 0; j < NBT; j++) acc[b][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    const long long nchunks = (long long)n * WG_NCHUNK;
    for (long long ch = gw; ch < nchunks; ch += WG_WAVES) {
        const int t = (int)(ch / WG_NCHUNK), c = (int)(ch - (long long)t * WG_NCHUNK);
        const int p = c * 32 + kq * 8 + li;  // the pixel this lane loads
        const bool pin = p < NPIX;
        const int y = p / TW, x = p - (p / TW) * TW;
        // A: masked output gradient, 8 pixels of channel mb*16 + m per lane
        bf16x8 a[MB];
#pragma unroll
        for (int b = 0; b < MB; b++) {
            bf16x8 v = zero8();
            if (pin) {
                const long long e = ((long long)t * NPIX + p) * CO + b * 16 + lh * 8;
                const bf16x8 g = ld8(dout + e), o = ld8(out + e);
#pragma unroll
                for (int i = 0; i < 8; i++) v[i] = (float)o[i] > 0.f ? g[i] : (bf16)0.0f;
            }
            *reinterpret_cast<bf16x8*>(&sA[wv][b][kq * 8 + li][lh * 8]) = v;
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int b = 0; b < MB; b++) {
#pragma unroll
            for (int i = 0; i < 8; i++) {
                a[b][i] = sA[wv][b][kq * 8 + i][m];
                bsum[b] += (float)a[b][i];
            }
        }
        const bf16* src = in + (long long)(idx ? idx[t] : t) * NPIX * CI;
#pragma unroll
        for (int j = 0; j < NBT; j++) {
            // B: the input at pixel + tap, 8 pixels of column j*16 + m per lane
            const int col0 = j * 16 + lh * 8;
            bf16x8 v = zero8();
            if (pin && col0 < K) {
                const int tap = col0 / CI, ci0 = col0 - tap * CI;
                const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
                if (yy >= 0 && yy < TW && xx >= 0 && xx < TW) v = ld8(src + (yy * TW + xx) * CI + ci0);
            }
            __builtin_amdgcn_wave_barrier();
            *reinterpret_cast<bf16x8*>(&sB[wv][kq * 8 + li][lh * 8]) = v;
            __builtin_amdgcn_wave_barrier();
            bf16x8 bb;
#pragma unroll
            for (int i = 0; i < 8; i++) bb[i] = sB[wv][kq * 8 + i][m];
#pragma unroll
            for (int b = 0; b < MB; b++) acc[b][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[b], bb, acc[b][j], 0, 0, 0);
        }
        __builtin_amdgcn_wave_barrier();
    }
    float* row = part + (long long)gw * (CO * K + CO);
#pragma unroll
    for (int b = 0; b < MB; b++) {
#pragma unroll
        for (int j = 0; j < NBT; j++) {
            const int col = j * 16 + m;
            if (col < K)
#pragma unroll
                for (int i = 0; i < 4; i++) row[(b * 16 + kq * 4 + i) * K + col] = acc[b][j][i];
        }
        float s = bsum[b];
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        if (kq == 0) row[CO * K + b * 16 + m] = s;
    }
}

// G[e] = sum over the WG_WAVES partial rows of part[.][e], e < len, in a fixed
// order: a block takes 64 entries; its four waves sum a quarter of the rows
// each, and the quarters are added in wave order.
__global__ void __launch_bounds__(256)
k_wgrad_reduce(const float* __restrict__ part, int len, float* __restrict__ G)
{
    __shared__ float red[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int e = blockIdx.x * 64 + lane;
    float s = 0.f;
    if (e < len)
        for (int w = wv * (WG_WAVES / 4); w < (wv + 1) * (WG_WAVES / 4); w++) s += part[(long long)w * len + e];
    red[wv][lane] = s;
    __syncthreads();
    if (wv == 0 && e < len) G[e] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}


// ---------------------------------------------------------------------------
// Head: global average pool of conv4 (16 ch) -> logits = Wfc f + bfc.  With
// labels: softmax cross-entropy (mean over the batch), its gradient into
// dWfc, dbfc and the pooled broadcast into dout4.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_head(const bf16* __restrict__ act4, const float* __restrict__ P, long long fcw, long long fcb,
       const int* __restrict__ labels, const int* __restrict__ idx, int n, float* __restrict__ logits,
       float* __restrict__ grad, bf16* __restrict__ dout4, float* __restrict__ loss)
{
    __shared__ float red[256];
    __shared__ float feat[16], dfeat[16];
    const int t = blockIdx.x;
    if (t >= n) return;
    const int c = threadIdx.x & 15, grp = threadIdx.x >> 4;
    float s = 0.f;
    for (int p = grp; p < NPIX; p += 16) s += (float)act4[((long long)t * NPIX + p) * 16 + c];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < 16) {
        float tot = 0.f;
        for (int gI = 0; gI < 16; gI++) tot += red[gI * 16 + threadIdx.x];
        feat[threadIdx.x] = tot / (float)NPIX;
    }
    __syncthreads();
    __shared__ float lg[2];
    if (threadIdx.x < 2) {
        float v = P[fcb + threadIdx.x];
        for (int k = 0; k < 16; k++) v += P[fcw + threadIdx.x * 16 + k] * feat[k];
        lg[threadIdx.x] = v;
        if (logits) logits[2 * t + threadIdx.x] = v;
    }
    __syncthreads();
    if (!labels) return;
    const int lab = labels[idx ? idx[t] : t];
    const float mx = fmaxf(lg[0], lg[1]);
    const float e0 = expf(lg[0] - mx), e1 = expf(lg[1] - mx);
    const float p0 = e0 / (e0 + e1), p1 = e1 / (e0 + e1);
    const float inv_n = 1.f / (float)n;
    const float d0 = (p0 - (lab == 0 ? 1.f : 0.f)) * inv_n, d1 = (p1 - (lab == 1 ? 1.f : 0.f)) * inv_n;
    if (threadIdx.x == 0) atomicAdd(loss, -logf(lab == 1 ? p1 : p0) * inv_n);
    if (threadIdx.x < 16) {
        const int k = threadIdx.x;
        atomicAdd(&grad[fcw + k], d0 * feat[k]);
        atomicAdd(&grad[fcw + 16 + k], d1 * feat[k]);
        dfeat[k] = P[fcw + k] * d0 + P[fcw + 16 + k] * d1;
    }
    if (threadIdx.x == 16) atomicAdd(&grad[fcb], d0);
    if (threadIdx.x == 17) atomicAdd(&grad[fcb + 1], d1);
    __syncthreads();
    for (int e = threadIdx.x; e < NPIX * 16; e += blockDim.x)
        dout4[(long long)t * NPIX * 16 + e] = (bf16)(dfeat[e & 15] / (float)NPIX);
}

// bf16 weight copies: W[co][k] and WT[ci][tap][co]
__global__ void k_pack(const float* __restrict__ P, long long off, int ci_n, int co_n, bf16* __restrict__ W,
                       bf16* __restrict__ WT)
{
    const int K = 9 * ci_n;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < co_n * K; e += gridDim.x * blockDim.x) {
        const int co = e / K, k = e - co * K, tap = k / ci_n, ci = k - tap * ci_n;
        const bf16 v = (bf16)P[off + e];
        W[e] = v;
        if (WT) WT[((long long)ci * 9 + tap) * co_n + co] = v;
    }
}

__global__ void k_adam(float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, long long n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const float gg = g[e];
        const float a = b1 * m1[e] + (1.f - b1) * gg;
        const float v = b2 * m2[e] + (1.f - b2) * gg * gg;
        m1[e] = a;
        m2[e] = v;
        P[e] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
    }
}

__global__ void k_fill0(float* __restrict__ x, long long n)
{
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x)
        x[e] = 0.f;
}

// Network input of one patch per workgroup: x_c = clamp((q_c - mean_c) /
// (2 dx thr), -16, 16) / 4 over the whole 36x36 tile, mean over the interior.
template <class R>
__global__ void __launch_bounds__(256)
k_make_tiles(const int* __restrict__ slots, const int* __restrict__ levels, int n, hc::Pools<R> P,
             const double* __restrict__ dxdy, int meqn, double thr, bf16* __restrict__ tiles)
{
    __shared__ float red[256];
    __shared__ float mean[3];
    const int p = blockIdx.x;
    if (p >= n) return;
    const int slot = slots[p];
    const R* q = P.q[levels[p]] + (long long)slot * NPIX * meqn;
    const int nc = meqn < 3 ? meqn : 3;
    for (int c = 0; c < nc; c++) {
        float s = 0.f;
        for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x)
            s += (float)q[(long long)c * NPIX + (e / 32 + 2) * TW + e % 32 + 2];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) mean[c] = red[0] / 1024.f;
        __syncthreads();
    }
    const float scale = (float)(1.0 / (2.0 * dxdy[2 * slot] * thr));
    for (int e = threadIdx.x; e < NPIX * C0; e += blockDim.x) {
        const int pix = e / C0, c = e - pix * C0;
        float v = 0.f;
        if (c < nc) {
            v = ((float)q[(long long)c * NPIX + pix] - mean[c]) * scale;
            v = fminf(fmaxf(v, -16.f), 16.f) * 0.25f;
        }
        tiles[(long long)p * NPIX * C0 + e] = (bf16)v;
    }
}

}  // namespace

namespace hc {
template <class R>
int launch_make_tiles(const int* slots, const int* levels, int n, const Pools<R>& P, const double* dxdy,
                      int mx, int my, int meqn, double thr, void* tiles, hipStream_t s)
{
    if (mx + 4 != TW || my + 4 != TW) return set_error(HIPCLAW_ERR_UNSUPPORTED, "CNN tiles need 32x32 patches");
    if (!(thr > 0)) return set_error(HIPCLAW_ERR_ARG, "CNN tiles need a positive refine threshold");
    if (n <= 0) return HIPCLAW_OK;
    hipLaunchKernelGGL(k_make_tiles<R>, dim3(n), dim3(256), 0, s, slots, levels, n, P, dxdy, meqn, thr,
                       static_cast<bf16*>(tiles));
    return check_hip(hipGetLastError(), "make tiles");
}
template int launch_make_tiles<double>(const int*, const int*, int, const Pools<double>&, const double*, int,
                                       int, int, double, void*, hipStream_t);
template int launch_make_tiles<float>(const int*, const int*, int, const Pools<float>&, const double*, int,
                                      int, int, double, void*, hipStream_t);
}  // namespace hc

struct hipclaw_cnn {
    hipStream_t stream = nullptr;
    ParamLayout L;
    float* P = nullptr;        // FP32 master parameters
    float* G = nullptr;        // gradients
    float* M1 = nullptr;       // Adam moments
    float* M2 = nullptr;
    float* loss = nullptr;     // device scalar
    bf16* Wb[4] = {};          // bf16 [COUT][9*CIN]
    bf16* WTb[4] = {};         // bf16 [CIN][9][COUT] (layers 1..3)
    int cap = 0;               // batch capacity of the activation buffers
    bf16* act[4] = {};         // conv outputs [cap, NPIX, COUT]
    bf16* dact[4] = {};        // gradients w.r.t. conv outputs
    float* logits = nullptr;   // [cap, 2]
    float* part = nullptr;     // weight-gradient partials [WG_WAVES][CO*K + CO] (largest layer)
    long long step = 0;
};

namespace {

int ensure_cap(hipclaw_cnn* c, int n)
{
    if (n <= c->cap) return HIPCLAW_OK;
    for (int l = 0; l < 4; l++) {
        if (c->act[l]) (void)hipFree(c->act[l]);
        if (c->dact[l]) (void)hipFree(c->dact[l]);
        c->act[l] = c->dact[l] = nullptr;
    }
    if (c->logits) (void)hipFree(c->logits);
    c->logits = nullptr;
    c->cap = 0;
    for (int l = 0; l < 4; l++) {
        const size_t b = sizeof(bf16) * (size_t)n * NPIX * COUT[l];
        int rc = check_hip(hipMalloc((void**)&c->act[l], b), "cnn activations");
        if (!rc) rc = check_hip(hipMalloc((void**)&c->dact[l], b), "cnn gradients");
        if (rc) return rc;
    }
    int rc = check_hip(hipMalloc((void**)&c->logits, sizeof(float) * 2 * (size_t)n), "cnn logits");
    if (rc) return rc;
    c->cap = n;
    return HIPCLAW_OK;
}

int pack(hipclaw_cnn* c)
{
    for (int l = 0; l < 4; l++)
        hipLaunchKernelGGL(k_pack, dim3(64), dim3(256), 0, c->stream, c->P, c->L.w[l], CIN[l], COUT[l],
                           c->Wb[l], l > 0 ? c->WTb[l] : nullptr);
    return check_hip(hipGetLastError(), "cnn pack");
}

unsigned blocks_for(long long n_waves) { return (unsigned)((n_waves + 3) / 4); }

int forward(hipclaw_cnn* c, const void* tiles, const int* idx, int n)
{
    const bf16* x = static_cast<const bf16*>(tiles);
    const unsigned nb = blocks_for((long long)n * PBLK);
    hipStream_t s = c->stream;
    hipLaunchKernelGGL((k_conv_fwd<8, 16>), dim3(nb), dim3(256), 0, s, x, idx, c->Wb[0], c->P + c->L.b[0],
                       c->act[0], n);
    hipLaunchKernelGGL((k_conv_fwd<16, 32>), dim3(nb), dim3(256), 0, s, c->act[0], nullptr, c->Wb[1],
                       c->P + c->L.b[1], c->act[1], n);
    hipLaunchKernelGGL((k_conv_fwd<32, 32>), dim3(nb), dim3(256), 0, s, c->act[1], nullptr, c->Wb[2],
                       c->P + c->L.b[2], c->act[2], n);
    hipLaunchKernelGGL((k_conv_fwd<32, 16>), dim3(nb), dim3(256), 0, s, c->act[2], nullptr, c->Wb[3],
                       c->P + c->L.b[3], c->act[3], n);
    return check_hip(hipGetLastError(), "cnn forward");
}

// forward + loss + all gradients into c->G (zeroed first)
int backward(hipclaw_cnn* c, const void* tiles, const int* labels, const int* idx, int n)
{
    int rc = forward(c, tiles, idx, n);
    if (rc) return rc;
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_fill0, dim3(64), dim3(256), 0, s, c->G, c->L.total);
    hipLaunchKernelGGL(k_fill0, dim3(1), dim3(64), 0, s, c->loss, 1LL);
    hipLaunchKernelGGL(k_head, dim3(n), dim3(256), 0, s, c->act[3], c->P, c->L.fcw, c->L.fcb, labels, idx, n,
                       c->logits, c->G, c->dact[3], c->loss);
    const bf16* x = static_cast<const bf16*>(tiles);
    const unsigned nb = blocks_for((long long)n * PBLK);
    // weight and bias gradients of layer l: per-wave partials, then a fixed-order sum
    auto wgrad = [&](int l, auto launch) {
        launch();
        const int len = COUT[l] * 9 * CIN[l] + COUT[l];
        hipLaunchKernelGGL(k_wgrad_reduce, dim3((unsigned)((len + 63) / 64)), dim3(256), 0, s, c->part, len,
                           c->G + c->L.w[l]);
    };
    const dim3 gwg(WG_WAVES / 4);
    // layer 4 (32 -> 16)
    wgrad(3, [&] {
        hipLaunchKernelGGL((k_conv_wgrad<32, 16>), gwg, dim3(256), 0, s, c->act[2], nullptr, c->dact[3],
                           c->act[3], c->part, n);
    });
    hipLaunchKernelGGL((k_conv_bwd_data<32, 16>), dim3(nb), dim3(256), 0, s, c->dact[3], c->act[3],
                       c->WTb[3], c->dact[2], n);
    // layer 3 (32 -> 32)
    wgrad(2, [&] {
        hipLaunchKernelGGL((k_conv_wgrad<32, 32>), gwg, dim3(256), 0, s, c->act[1], nullptr, c->dact[2],
                           c->act[2], c->part, n);
    });
    hipLaunchKernelGGL((k_conv_bwd_data<32, 32>), dim3(nb), dim3(256), 0, s, c->dact[2], c->act[2],
                       c->WTb[2], c->dact[1], n);
    // layer 2 (16 -> 32)
    wgrad(1, [&] {
        hipLaunchKernelGGL((k_conv_wgrad<16, 32>), gwg, dim3(256), 0, s, c->act[0], nullptr, c->dact[1],
                           c->act[1], c->part, n);
    });
    hipLaunchKernelGGL((k_conv_bwd_data<16, 32>), dim3(nb), dim3(256), 0, s, c->dact[1], c->act[1],
                       c->WTb[1], c->dact[0], n);
    // layer 1 (8 -> 16): weights only
    wgrad(0, [&] {
        hipLaunchKernelGGL((k_conv_wgrad<8, 16>), gwg, dim3(256), 0, s, x, idx, c->dact[0], c->act[0],
                           c->part, n);
    });
    return check_hip(hipGetLastError(), "cnn backward");
}

int check_batch(hipclaw_cnn* c, const void* tiles, int n)
{
    if (!c) return set_error(HIPCLAW_ERR_ARG, "cnn is NULL");
    if (!tiles || n < 1) return set_error(HIPCLAW_ERR_ARG, "no tiles");
    return ensure_cap(c, n);
}

}  // namespace

extern "C" {

int hipclaw_cnn_create(void* stream, hipclaw_cnn** out)
{
    if (!out) return set_error(HIPCLAW_ERR_ARG, "out is NULL");
    auto* c = new (std::nothrow) hipclaw_cnn;
    if (!c) return set_error(HIPCLAW_ERR_NOMEM, "host allocation failed");
    c->stream = static_cast<hipStream_t>(stream);
    const size_t pb = sizeof(float) * (size_t)c->L.total;
    int rc = check_hip(hipMalloc((void**)&c->P, pb), "cnn params");
    if (!rc) rc = check_hip(hipMalloc((void**)&c->G, pb), "cnn grads");
    if (!rc) rc = check_hip(hipMalloc((void**)&c->M1, pb), "cnn adam");
    if (!rc) rc = check_hip(hipMalloc((void**)&c->M2, pb), "cnn adam");
    if (!rc) rc = check_hip(hipMalloc((void**)&c->loss, sizeof(float)), "cnn loss");
    size_t prow = 0;  // the largest layer's weight + bias count
    for (int l = 0; l < 4; l++) prow = std::max(prow, (size_t)COUT[l] * 9 * CIN[l] + COUT[l]);
    if (!rc) rc = check_hip(hipMalloc((void**)&c->part, sizeof(float) * WG_WAVES * prow), "cnn partials");
    for (int l = 0; l < 4 && !rc; l++) {
        const size_t wb = sizeof(bf16) * (size_t)COUT[l] * 9 * CIN[l];
        rc = check_hip(hipMalloc((void**)&c->Wb[l], wb), "cnn bf16 weights");
        if (!rc) rc = check_hip(hipMalloc((void**)&c->WTb[l], wb), "cnn bf16 weights");
    }
    if (rc) {
        hipclaw_cnn_destroy(c);
        return rc;
    }
    rc = hipclaw_cnn_init_params(c, 1234);
    if (rc) {
        hipclaw_cnn_destroy(c);
        return rc;
    }
    *out = c;
    return HIPCLAW_OK;
}

int hipclaw_cnn_destroy(hipclaw_cnn* c)
{
    if (!c) return HIPCLAW_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void* bufs[] = {c->P, c->G, c->M1, c->M2, c->loss, c->logits, c->part};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (int l = 0; l < 4; l++) {
        void* more[] = {c->Wb[l], c->WTb[l], c->act[l], c->dact[l]};
        for (void* b : more)
            if (b) (void)hipFree(b);
    }
    delete c;
    return HIPCLAW_OK;
}

int hipclaw_cnn_num_params(hipclaw_cnn* c, long long* n)
{
    if (!c || !n) return set_error(HIPCLAW_ERR_ARG, "NULL argument");
    *n = c->L.total;
    return HIPCLAW_OK;
}

int hipclaw_cnn_init_params(hipclaw_cnn* c, unsigned long long seed)
{
    if (!c) return set_error(HIPCLAW_ERR_ARG, "cnn is NULL");
    std::vector<float> h((size_t)c->L.total, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < 4; l++) {
        // He initialisation over the real fan-in (3 input channels for layer 1)
        const int fan = 9 * (l == 0 ? 3 : CIN[l]);
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / fan));
        for (int co = 0; co < COUT[l]; co++)
            for (int k = 0; k < 9 * CIN[l]; k++) {
                const int ci = k % CIN[l];
                h[(size_t)(c->L.w[l] + (long long)co * 9 * CIN[l] + k)] = (l == 0 && ci >= 3) ? 0.f : nd(rng);
            }
        for (int co = 0; co < COUT[l]; co++) h[(size_t)(c->L.b[l] + co)] = 0.01f;
    }
    std::normal_distribution<float> nf(0.f, std::sqrt(1.f / 16));
    for (int k = 0; k < 32; k++) h[(size_t)(c->L.fcw + k)] = nf(rng);
    return hipclaw_cnn_set_params(c, h.data());
}

int hipclaw_cnn_set_params(hipclaw_cnn* c, const float* host)
{
    if (!c || !host) return set_error(HIPCLAW_ERR_ARG, "NULL argument");
    const size_t pb = sizeof(float) * (size_t)c->L.total;
    int rc = check_hip(hipMemcpy(c->P, host, pb, hipMemcpyHostToDevice), "cnn params");
    if (rc) return rc;
    hipLaunchKernelGGL(k_fill0, dim3(64), dim3(256), 0, c->stream, c->M1, c->L.total);
    hipLaunchKernelGGL(k_fill0, dim3(64), dim3(256), 0, c->stream, c->M2, c->L.total);
    c->step = 0;
    rc = pack(c);
    if (rc) return rc;
    return check_hip(hipStreamSynchronize(c->stream), "cnn set params");
}

int hipclaw_cnn_get_params(hipclaw_cnn* c, float* host)
{
    if (!c || !host) return set_error(HIPCLAW_ERR_ARG, "NULL argument");
    int rc = check_hip(hipStreamSynchronize(c->stream), "sync");
    if (rc) return rc;
    return check_hip(hipMemcpy(host, c->P, sizeof(float) * (size_t)c->L.total, hipMemcpyDeviceToHost),
                     "cnn params");
}

int hipclaw_cnn_forward(hipclaw_cnn* c, const void* tiles, const int* idx, int n, float* logits)
{
    int rc = check_batch(c, tiles, n);
    if (rc) return rc;
    rc = forward(c, tiles, idx, n);
    if (rc) return rc;
    hipLaunchKernelGGL(k_head, dim3(n), dim3(256), 0, c->stream, c->act[3], c->P, c->L.fcw, c->L.fcb,
                       nullptr, nullptr, n, logits, nullptr, nullptr, nullptr);
    return check_hip(hipGetLastError(), "cnn head");
}

int hipclaw_cnn_tag_patches(hipclaw_cnn* c, int n, const double* const* host_q, const double* dx,
                            int meqn, double threshold, int* tags)
{
    if (!c) return set_error(HIPCLAW_ERR_ARG, "cnn is NULL");
    if (n < 0 || (n > 0 && (!host_q || !dx || !tags))) return set_error(HIPCLAW_ERR_ARG, "NULL argument");
    if (meqn < 1 || meqn > 16) return set_error(HIPCLAW_ERR_ARG, "meqn must be 1..16");
    if (!(threshold > 0)) return set_error(HIPCLAW_ERR_ARG, "threshold must be > 0");
    if (n == 0) return HIPCLAW_OK;
    int rc = ensure_cap(c, n);
    if (rc) return rc;
    // stage the host tiles (32x32 patches, mbc 2, q(i, j, m)) into one device
    // buffer in slot order, then the same tile-making kernel and network
    // forward pass as the device-resident regrid
    const size_t tile = (size_t)NPIX * meqn;
    std::vector<double> hq(tile * (size_t)n), hdx(2 * (size_t)n);
    std::vector<int> hslot(n), hlev(n, 0);
    for (int p = 0; p < n; p++) {
        if (!host_q[p]) return set_error(HIPCLAW_ERR_ARG, "NULL patch pointer");
        std::memcpy(hq.data() + tile * p, host_q[p], sizeof(double) * tile);
        hdx[2 * p] = hdx[2 * p + 1] = dx[p];
        hslot[p] = p;
    }
    double *dq = nullptr, *ddx = nullptr;
    int *dslot = nullptr, *dlev = nullptr;
    float* dlog = nullptr;
    void* dt = nullptr;
    rc = check_hip(hipMalloc((void**)&dq, sizeof(double) * hq.size()), "tag q");
    if (!rc) rc = check_hip(hipMalloc((void**)&ddx, sizeof(double) * hdx.size()), "tag dx");
    if (!rc) rc = check_hip(hipMalloc((void**)&dslot, sizeof(int) * 2 * (size_t)n), "tag slots");
    if (!rc) rc = check_hip(hipMalloc((void**)&dlog, sizeof(float) * 2 * (size_t)n), "tag logits");
    if (!rc) rc = check_hip(hipMalloc(&dt, sizeof(bf16) * (size_t)NPIX * C0 * n), "tag tiles");
    if (!rc) {
        dlev = dslot + n;
        rc = check_hip(hipMemcpyAsync(dq, hq.data(), sizeof(double) * hq.size(), hipMemcpyHostToDevice, c->stream),
                       "tag q");
        if (!rc) rc = check_hip(hipMemcpyAsync(ddx, hdx.data(), sizeof(double) * hdx.size(), hipMemcpyHostToDevice,
                                               c->stream), "tag dx");
        if (!rc) rc = check_hip(hipMemcpyAsync(dslot, hslot.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream),
                                "tag slots");
        if (!rc) rc = check_hip(hipMemcpyAsync(dlev, hlev.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream),
                                "tag levels");
    }
    if (!rc) {
        hc::Pools<double> P{};
        P.q[0] = dq;
        rc = hc::launch_make_tiles<double>(dslot, dlev, n, P, ddx, 32, 32, meqn, threshold, dt, c->stream);
    }
    if (!rc) rc = hipclaw_cnn_forward(c, dt, nullptr, n, dlog);
    std::vector<float> lg(2 * (size_t)n);
    if (!rc) rc = check_hip(hipMemcpyAsync(lg.data(), dlog, sizeof(float) * lg.size(), hipMemcpyDeviceToHost,
                                           c->stream), "tag logits");
    if (!rc) rc = check_hip(hipStreamSynchronize(c->stream), "tag sync");
    if (!rc)
        for (int p = 0; p < n; p++) tags[p] = (double)lg[2 * p + 1] - (double)lg[2 * p] >= 0.0 ? 1 : 0;
    void* bufs[] = {dq, ddx, dslot, dlog, dt};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    return rc;
}

int hipclaw_cnn_gradients(hipclaw_cnn* c, const void* tiles, const int* labels, const int* idx, int n,
                          float* grad_host, float* loss_host)
{
    int rc = check_batch(c, tiles, n);
    if (!rc && !labels) rc = set_error(HIPCLAW_ERR_ARG, "labels is NULL");
    if (!rc) rc = backward(c, tiles, labels, idx, n);
    if (!rc) rc = check_hip(hipStreamSynchronize(c->stream), "sync");
    if (!rc && grad_host)
        rc = check_hip(hipMemcpy(grad_host, c->G, sizeof(float) * (size_t)c->L.total, hipMemcpyDeviceToHost),
                       "cnn grads");
    if (!rc && loss_host) rc = check_hip(hipMemcpy(loss_host, c->loss, sizeof(float), hipMemcpyDeviceToHost), "loss");
    return rc;
}

int hipclaw_cnn_train_step(hipclaw_cnn* c, const void* tiles, const int* labels, const int* idx, int n,
                           float lr, float* loss_dev)
{
    int rc = check_batch(c, tiles, n);
    if (!rc && !labels) rc = set_error(HIPCLAW_ERR_ARG, "labels is NULL");
    if (!rc) rc = backward(c, tiles, labels, idx, n);
    if (rc) return rc;
    c->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)c->step), c2 = 1.f - std::pow(b2, (float)c->step);
    hipLaunchKernelGGL(k_adam, dim3(80), dim3(256), 0, c->stream, c->P, c->G, c->M1, c->M2, c->L.total, lr, b1,
                       b2, 1e-8f, c1, c2);
    rc = pack(c);
    if (!rc && loss_dev)
        rc = check_hip(hipMemcpyAsync(loss_dev, c->loss, sizeof(float), hipMemcpyDeviceToDevice, c->stream),
                       "loss");
    return rc;
}

}  // extern "C"
