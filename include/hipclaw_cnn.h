/*
 * hipclaw_cnn.h -- C ABI of the learned refinement indicator: a 4-layer 3x3
 * CNN on patch tiles with hand-written BF16 MFMA (FP32 accumulate) training
 * and inference kernels for gfx950.  It replaces the per-patch refinement tag
 * pass of fclaw2d_regrid (reference src/fclaw2d_regrid.c:45-72,
 * src/patches/clawpatch/fclaw2d_clawpatch.cpp:799-850); the plain gradient
 * tag (fclaw2d_clawpatch46_fort_tag4refinement) stays as fallback and judge.
 *
 * Tiles: bf16, [n][36][36][8] (row-major pixels, channels last; channels
 * 0..2 = normalised h, hu, hv, 3..7 = 0), made from a context's patches by
 * hipclaw_make_tiles.  Logits: float [n][2] = (keep, refine).
 * All device pointers; the network owns its parameters and work buffers.
 */
#ifndef HIPCLAW_CNN_H
#define HIPCLAW_CNN_H

#include "hipclaw.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hipclaw_cnn hipclaw_cnn;

#define HIPCLAW_TILE_W 36
#define HIPCLAW_TILE_C 8

int hipclaw_cnn_create(void* stream, hipclaw_cnn** out);
int hipclaw_cnn_destroy(hipclaw_cnn* net);
/* Number of FP32 parameters: conv l weights [Cout][3][3][Cin] + bias for
 * (Cin, Cout) = (8,16), (16,32), (32,32), (32,16); then fc [2][16] + bias. */
int hipclaw_cnn_num_params(hipclaw_cnn* net, long long* n);
int hipclaw_cnn_init_params(hipclaw_cnn* net, unsigned long long seed);
int hipclaw_cnn_set_params(hipclaw_cnn* net, const float* host);
int hipclaw_cnn_get_params(hipclaw_cnn* net, float* host);
/* Logits of tiles idx[0..n) (idx = NULL: tiles 0..n-1). */
int hipclaw_cnn_forward(hipclaw_cnn* net, const void* tiles, const int* idx, int n, float* logits);
/* Mean softmax cross-entropy over the batch and its gradient w.r.t. every
 * parameter, copied to host arrays (either may be NULL).  labels[tile] in {0, 1}. */
int hipclaw_cnn_gradients(hipclaw_cnn* net, const void* tiles, const int* labels, const int* idx,
                          int n, float* grad_host, float* loss_host);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the batch; *loss_dev (device,
 * optional) receives the batch loss. */
int hipclaw_cnn_train_step(hipclaw_cnn* net, const void* tiles, const int* labels, const int* idx,
                           int n, float lr, float* loss_dev);

/* Tiles of every patch of a context, in slot order, normalised for the
 * network: x_c = clamp((q_c - mean_c) / (2 dx thr), -16, 16) / 4 with mean_c the
 * patch-interior mean and thr the context's refine threshold, so that one
 * unit of a cell-to-cell jump is the plain criterion's threshold on every
 * level.  tags_dev (optional, int per patch) receives the plain refinement tag. */
int hipclaw_make_tiles(hipclaw_ctx* ctx, void* tiles_dev, int* tags_dev);
/* Refinement tags of regrid from the network: refine when p(refine) >= 0.5,
 * except that patches with |p - 0.5| < margin take the plain tag.  net = NULL
 * restores plain tagging.  Coarsening keeps the plain test. */
int hipclaw_set_tagging_cnn(hipclaw_ctx* ctx, hipclaw_cnn* net, float margin);
/* Refinement tags of n host-resident patches from the network, for a
 * ForestClaw that keeps its patches on the host (patch_vt->tag4refinement,
 * src/patches/clawpatch/fclaw2d_clawpatch.cpp:1373, called per patch from
 * cb_fclaw2d_regrid_tag4refinement, src/fclaw2d_regrid.c:45-72): the caller
 * gathers the patches of a regrid pass once and answers each per-patch call
 * from tags[].  host_q[p]: the patch's q(1-mbc:mx+mbc, 1-mbc:my+mbc, meqn)
 * array, FP64, 32x32 patches, mbc 2; dx[p] its cell width; threshold the
 * plain criterion's refine threshold (the tile normalisation of
 * hipclaw_make_tiles).  tags[p] = 1: refine (p(refine) >= 0.5).  Staging
 * buffers are allocated per call. */
int hipclaw_cnn_tag_patches(hipclaw_cnn* net, int n, const double* const* host_q, const double* dx, int meqn,
                            double threshold, int* tags);
/* Statistics of the last regrid's tag pass: patches, refine tags of the
 * network, of the plain test, misses (plain refine, network not), false
 * refines, fallbacks. */
int hipclaw_tag_stats(hipclaw_ctx* ctx, int* stats6);

#ifdef __cplusplus
}
#endif
#endif
