/*
 * hipclaw.h -- C ABI of the MI355X-native (gfx950) Clawpack-4.6 wave-propagation
 * patch solver for ForestClaw.
 *
 * Plain C types only (no HIP or torch types): device buffers are passed as
 * void* device addresses, HIP streams as void* (NULL = the default stream).
 * Every entry point returns an int status (HIPCLAW_OK = 0) and never exits;
 * hipclaw_last_error() returns a message for the calling thread's last
 * failure.
 *
 * Storage layout of one patch is the clawpatch FArrayBox layout
 * q(1-mbc:mx+mbc, 1-mbc:my+mbc, meqn), i fastest, field-major
 * (reference src/patches/clawpatch/fclaw2d_clawpatch.cpp:305-335); a batch of
 * patches is packed back to back, T*meqn elements per patch with
 * T = (mx+2mbc)(my+2mbc).  Element type is double (HIPCLAW_FP64) or float
 * (HIPCLAW_FP32) as selected in hipclaw_physics.precision.
 */
#ifndef HIPCLAW_H
#define HIPCLAW_H

#ifdef __cplusplus
extern "C" {
#endif

#define HIPCLAW_ABI_VERSION 1

#define HIPCLAW_OK 0
#define HIPCLAW_ERR_ARG 1         /* invalid argument or configuration      */
#define HIPCLAW_ERR_HIP 2         /* HIP runtime failure                    */
#define HIPCLAW_ERR_UNSUPPORTED 3 /* valid in the reference, not built here */
#define HIPCLAW_ERR_NOMEM 4       /* device allocation failed               */
#define HIPCLAW_ERR_STATE 5       /* call out of sequence                   */

/* Riemann solver families (replace the fc2d_clawpack46_vtable fort_rpn2 /
 * fort_rpt2 function pointers, which cannot run on the device). */
#define HIPCLAW_SOLVER_ADV 0        /* clawpack46_rpn2adv / rpt2adv (meqn 1, maux 2) */
#define HIPCLAW_SOLVER_SWE_ROE 1    /* shallow-water Roe + entropy fix (meqn 3)      */
#define HIPCLAW_SOLVER_EULER5_ROE 2 /* Euler + tracer Roe + entropy fix (meqn 5; maux
                                       0, or 1 = the radius for the axisymmetric
                                       source, not read by the step)             */
#define HIPCLAW_SOLVER_EULER_HLLE 3 /* Euler + conserved tracer, HLLE normal solver with
                                       Einfeldt's speeds (2 waves; meqn 5, mwaves 2)
                                       and the HLL form of rpt2_euler5's Roe split.
                                       No counterpart in the reference (its Euler
                                       solver is Roe only); same interface as
                                       clawpack46_rpn2_euler5.f:1-280 /
                                       clawpack46_rpt2_euler5.f:1-122             */
#define HIPCLAW_SOLVER_TRANSPORT_FW 4 /* conservative transport in f-wave form: the
                                       clawpack46_flux2fw.f step (step2_wrap with
                                       use_fwaves) with the transport solver pair
                                       clawpack46_rpn2cons_fw_manifold.f /
                                       clawpack46_rpt2cons_manifold.f (meqn 1,
                                       mwaves 1, maux 7: capacity, left / right /
                                       bottom / top edge velocities, x / y edge
                                       lengths).  hipclaw_step2_batch only: the
                                       AMR context does not fill these edge data */

/* Flag OR-ed into the solver id: the Clawpack 5 variant of the solver family
 * (the fc2d_clawpack5 solver, src/solvers/fc2d_clawpack5/, with the
 * applications' clawpack5_rpn2 / rpt2).  Its FP64 arithmetic follows the
 * Clawpack 5 solvers' operation order, and host patch arrays exchanged with
 * the library (hipclaw_single_step_update, level upload / download,
 * set_q / get_q, set_aux) use the Clawpack 5 layout q(meqn, i, j)
 * (fclaw2d_clawpatch5 storage) instead of q(i, j, meqn).  Device buffers keep
 * the library's tile layout either way. */
#define HIPCLAW_CLAW5 0x100

#define HIPCLAW_FP64 0
#define HIPCLAW_FP32 1

/* Solver configuration.  Replaces the Fortran common blocks /cparam/
 * (grav | gamma, gamma1), the [clawpatch] and [clawpack46] option groups
 * (reference src/solvers/fc2d_clawpack4.6/fc2d_clawpack46_options.c:93-99:
 * method(2) = order[0], method(3) = order[1]) and the entropy-fix flag. */
typedef struct hipclaw_physics {
    int solver;
    int mx, my, mbc;  /* mbc must be 2 */
    int meqn, maux, mwaves;
    int order[2];
    int mthlim[5];    /* limiter per wave family: 0 none, 1 minmod, 2 superbee,
                         3 van Leer, 4 MC, 5 Beam-Warming, 6 generalized minmod */
    int mthbc[4];     /* left, right, bottom, top: 1 extrapolation, 3 wall     */
    int efix;         /* 1 = entropy fix (the reference always uses it)        */
    int precision;    /* HIPCLAW_FP64 or HIPCLAW_FP32                          */
    double grav;      /* shallow water                                         */
    double gamma;     /* Euler                                                 */
} hipclaw_physics;

/* Launch shape of the patch-step kernel (library default when NULL):
 * cols_per_lane 1 or 2; waves_per_group 1..4; row_bands 0 or 1 sweeps every
 * patch with one wave, 2..8 splits the interior rows of each patch into that
 * many bands swept by the waves of one workgroup (waves_per_group is then
 * ignored; launches in place or with cols_per_lane 2 sweep whole patches
 * with the given waves_per_group).  row_bands = -1: the given columns per
 * lane and waves per group for launches of 2,048 patches or more, the library
 * default below that.  Launch shapes change no bits in FP64; FP32 results of
 * one and two columns per lane differ by contraction rounding.  The library
 * default: two columns per lane for FP32 launches of 2,048 patches or more
 * (mx > 28), else one; one wave per workgroup below 1,024 patches; 8 bands
 * for launches of at most 256 patches, 4 for at most 512 and 2 for fewer
 * than 2,048. */
typedef struct hipclaw_launch {
    int cols_per_lane;
    int waves_per_group;
    int row_bands;
} hipclaw_launch;

int hipclaw_abi_version(void);
const char* hipclaw_last_error(void);

/* Fill in the defaults of the reference options for a solver family. */
int hipclaw_physics_default(int solver, int mx, int my, hipclaw_physics* out);

/* The launch shape a patch-step launch of npatch patches uses for the
 * caller's shape launch (NULL: the library default), with row_bands the
 * band count it really sweeps (1 = whole patches) and waves_per_group the
 * workgroup size in waves.  Host-only: no device call.  No reference
 * counterpart (the reference has one fixed loop nest per patch). */
int hipclaw_launch_shape(const hipclaw_physics* phys, int npatch, const hipclaw_launch* launch,
                         hipclaw_launch* out);

/* ------------------------------------------------------------------------
 * Batched single-step update.
 *
 * Replaces patch_vt->single_step_update = clawpack46_update
 * (reference src/solvers/fc2d_clawpack4.6/fc2d_clawpack46.cpp:402-444 ->
 * clawpack46_step2 :258-400 -> CLAWPACK46_STEP2_WRAP, fortran_source/
 * clawpack46_step2_wrap.f:1-127) for npatch patches at once, as the deferred
 * batching of fclaw2d_single_step_buffer_data allows
 * (src/fclaw2d_update_single_step.h:47-55).
 *
 * q_in   device, npatch tiles (ghost cells filled)
 * q_out  device, npatch tiles; interior cells receive q after the step, the
 *        ghost band is not written.  q_out == q_in updates in place; a
 *        distinct q_out leaves q_in as the reference's save_current_step copy
 *        (fc2d_clawpack46.cpp:295).
 * aux    device, npatch aux tiles (maux fields) or NULL when maux == 0
 * dxdy   device, 2 doubles per patch (dx, dy)
 * patch_ids device, optional list of npatch tile indices (NULL = 0..npatch-1)
 * ntiles number of tiles addressable in q_in / q_out / aux / dxdy / cfl_patch;
 *        list entries outside [0, ntiles) are skipped (dxdy and cfl_patch
 *        are indexed by tile)
 * cfl_patch device, one double per tile: the per-patch Courant number
 *        returned by clawpack46_update (may be NULL)
 * ------------------------------------------------------------------------ */
int hipclaw_step2_batch(const hipclaw_physics* phys, int npatch, const void* q_in,
                        void* q_out, const void* aux, const double* dxdy,
                        const int* patch_ids, int ntiles, double dt, double* cfl_patch,
                        const hipclaw_launch* launch, void* stream);

/* The same update that also accumulates the conservative time-sync
 * registers of every patch (SWE only), as clawpack46_step2 does when
 * fclaw_opt->time_sync is set: clawpack46_time_sync_store_flux on the data
 * before the update and clawpack46_time_sync_accumulate_waves on the final
 * edge fluctuations (reference src/solvers/fc2d_clawpack4.6/fc2d_clawpack46.cpp:317-341,
 * 376-390; fortran_source/clawpack46_time_sync.f:37-226).  registers: device
 * FP64, 8*meqn*(mx+my) doubles per tile addressed like q; per face f = 0 left,
 * 1 right, 2 bottom, 3 top (length len = my or mx) four blocks of meqn*len
 * (len fastest): fp|gp, fm|gm, edge flux at the interior cell, edge flux at
 * the ghost cell (the fclaw2d_clawpatch_registers arrays,
 * src/patches/clawpatch/fclaw2d_clawpatch_conservation.h:56-77). */
int hipclaw_step2_batch_sync(const hipclaw_physics* phys, int npatch, const void* q_in,
                             void* q_out, const void* aux, const double* dxdy,
                             const int* patch_ids, int ntiles, double dt, double* cfl_patch,
                             double* registers, const hipclaw_launch* launch, void* stream);

/* hipclaw_step2_batch that allocates nothing and returns the batch's maximum
 * Courant number, the value fclaw2d_update_single_step takes over a level,
 * in *maxcfl (one device double), instead of per-tile values.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The per-tile Courant numbers, which
 * hipclaw_step2_batch writes to the caller's cfl_patch workspace, live in a
 * static device array of HIPCLAW_NOALLOC_MAX_TILES doubles (2 MiB per GPU,
 * held until the process ends): ntiles and npatch are limited to that many
 * (HIPCLAW_ERR_ARG beyond it).  Every call on one GPU shares the array, so
 * calls that overlap on one GPU (other streams, other threads) race; calls
 * ordered on one stream do not.  Other arguments as hipclaw_step2_batch. */
#define HIPCLAW_NOALLOC_MAX_TILES 262144
int hipclaw_step2_batch_noalloc(const hipclaw_physics* phys, int npatch, const void* q_in,
                                void* q_out, const void* aux, const double* dxdy,
                                const int* patch_ids, int ntiles, double dt, double* maxcfl,
                                const hipclaw_launch* launch, void* stream);

/* ------------------------------------------------------------------------
 * Host-resident patches: deferred single-step update.
 *
 * For a host that keeps its clawpatch arrays in host memory (an unmodified
 * ForestClaw): register a patch_vt->single_step_update callback that calls
 * hipclaw_single_step_update with the patch's q / aux arrays and the
 * fclaw2d_single_step_buffer_data_t counters it receives (reference
 * src/fclaw2d_patch.h:937-943, src/fclaw2d_update_single_step.c:45-103).
 * Calls stage patches into pinned memory; the call that fills the batch
 * (capacity patches) or carries iter == total_count-1 runs one batched launch
 * and copies the updated interior cells back (ghost cells left as given).
 * *maxcfl is the batch Courant number on a flushing call, 0 otherwise.
 * All calls of one level must share t and dt.  No time-sync registers on
 * this path (use the device-resident context for that).
 * ------------------------------------------------------------------------ */
typedef struct hipclaw_batch hipclaw_batch;
int hipclaw_batch_create(const hipclaw_physics* phys, int capacity, void* stream,
                         hipclaw_batch** out);
int hipclaw_batch_destroy(hipclaw_batch* b);
int hipclaw_single_step_update(hipclaw_batch* b, double* q, const double* aux, double dx,
                               double dy, double t, double dt, int iter, int total_count,
                               double* maxcfl);
/* The source term clawpack46_update applies after step2 (HIPCLAW_SRC2_*);
 * the axisymmetric source reads the radius from aux field 1, so it needs
 * maux >= 1 with the application's setaux values. */
int hipclaw_batch_set_src2(hipclaw_batch* b, int kind);

/* Maximum of n device doubles into *out_dev (device), the level CFL of
 * fclaw2d_update_single_step (src/fclaw2d_update_single_step.c:68). */
int hipclaw_max_reduce(const double* x_dev, int n, double* out_dev, void* stream);

/* Swirl velocity field (clawpack46_b4step2 of the swirl application,
 * applications/clawpack/advection/2d/swirl/user_4.6/b4step2.f:1-41 + psi.f).
 * xylower: 2 doubles per patch; writes maux = 2 fields per patch. */
int hipclaw_swirl_aux_batch(const hipclaw_physics* phys, int npatch, const double* xylower,
                            const double* dxdy, double t, double dt, double tperiod,
                            void* aux, void* stream);

/* Source terms (claw46_vt->fort_src2, called by clawpack46_update after
 * step2 with the same t and dt when src_term > 0; reference
 * src/solvers/fc2d_clawpack4.6/fc2d_clawpack46.cpp:186-215, 434-441).
 * HIPCLAW_SRC2_AXISYMMETRIC: the shockbubble application's src2
 * (applications/clawpack/euler/2d/shockbubble/user_4.6/src2.f): the
 * axisymmetric Euler source about y = 0, radius = aux field 1 (its setaux:
 * ylower + (j-1/2) dy) when aux is given, else computed from xylower / dxdy.
 * Updates the interior cells of npatch tiles of q in place. */
#define HIPCLAW_SRC2_NONE 0
#define HIPCLAW_SRC2_AXISYMMETRIC 1
int hipclaw_src2_batch(const hipclaw_physics* phys, int kind, int npatch, void* q, const void* aux,
                       const double* xylower, const double* dxdy, double t, double dt, void* stream);

/* Physical boundary conditions, clawpack46_bc2_default
 * (src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_bc2_default.f:1-213),
 * per patch mthbc[4] (-1 where the face is not on the physical boundary,
 * as clawpack46_bc2 builds it, fc2d_clawpack46.cpp:120-128), indexed by list
 * entry; tiles outside [0, ntiles) are skipped. */
int hipclaw_bc2_batch(const hipclaw_physics* phys, int npatch, void* q, const int* patch_ids,
                      int ntiles, const int* mthbc, void* stream);


/* ------------------------------------------------------------------------
 * Device-resident AMR solver context.
 *
 * Holds the forest of leaf patches (a quadtree per block of an mi x mj brick,
 * replacing what the reference obtains from p4est through src/forestclaw2d.c),
 * all patch data packed per slot in device pools, the precomputed ghost-fill
 * plans and the subcycling state.  Leaves are ordered by level, then Morton
 * order within each block; host arrays exchanged with the context follow that
 * order (slot k = leaf k).
 * ------------------------------------------------------------------------ */
typedef struct hipclaw_domain {
    int mi, mj;            /* brick of mi x mj unit blocks (1 x 1 = single block) */
    double ax, bx, ay, by; /* physical extent of the brick                        */
    int minlevel, maxlevel;
    int subcycle;          /* [Options] subcycle: 1 = level time stepping (default) */
} hipclaw_domain;

typedef struct hipclaw_ctx hipclaw_ctx;

/* Initial-condition families (the applications' clawpack46_qinit hooks). */
#define HIPCLAW_QINIT_CONST 0
#define HIPCLAW_QINIT_SWIRL 1        /* swirl/user_4.6/qinit.f: q = 1 for x < p[0]      */
#define HIPCLAW_QINIT_RADIAL_DAM 2   /* p = x0, y0, r0, h_in, h_out (cell averages)      */
#define HIPCLAW_QINIT_HUMPS 3        /* p = h0, n, then n x (cx, cy, A, w)               */
#define HIPCLAW_QINIT_SHOCK_BUBBLE 4 /* p = gamma, xs, rho, u, p post-shock, x0, y0, r, rho_b */
/* The shockbubble application's qinit (euler/2d/shockbubble/user_4.6/qinit.f
 * with setprob.f and the idisc = 2 fdisc): p = gamma, x0, y0, r0, rhoin, pinf;
 * meqn = 5; the bubble is averaged by the polygon rule (nsub ignored). */
#define HIPCLAW_QINIT_SHOCKBUBBLE_APP 5

int hipclaw_ctx_create(const hipclaw_physics* phys, const hipclaw_domain* dom, void* stream,
                       hipclaw_ctx** out);
int hipclaw_ctx_destroy(hipclaw_ctx* ctx);

/* Uniform forest at one level (fclaw2d_domain_new_brick / _unitsquare). */
int hipclaw_init_uniform(hipclaw_ctx* ctx, int level);
/* Forest from an explicit leaf list (level, gx, gy), e.g. to restore a saved
 * mesh (the reference's restart path, src/fclaw2d_forestclaw.c) or to run two
 * precisions on one mesh.  The list must cover the brick exactly with levels in
 * [minlevel, maxlevel] and be 2:1 balanced across faces and corners; patches are
 * renumbered into level-then-Morton order (read them back with patch_info). */
int hipclaw_init_leaves(hipclaw_ctx* ctx, int nleaves, const int* level, const int* gx,
                        const int* gy);
int hipclaw_num_patches(hipclaw_ctx* ctx, int* n);
/* Conservative time synchronisation (fclaw_opt->time_sync; SWE only): the
 * context keeps FP64 registers per patch, accumulated by every level update,
 * and hipclaw_advance applies fclaw2d_time_sync after each synchronising ghost
 * update (reference src/fclaw2d_advance.c:280-341). */
int hipclaw_set_time_sync(hipclaw_ctx* ctx, int on);
/* Source term applied after every patch step ([clawpack46] src_term with the
 * application's fort_src2; HIPCLAW_SRC2_*, see hipclaw_src2_batch).  The
 * radius of the axisymmetric source comes from the patch geometry.  Turns
 * the same-level ghost push off (pushed ghosts would predate the source). */
int hipclaw_set_src2(hipclaw_ctx* ctx, int kind);
/* fclaw2d_time_sync(glob, minlevel, maxlevel) (src/fclaw2d_time_sync.c:157-202):
 * same-size and fine-to-coarse corrections, then the register resets. */
int hipclaw_time_sync(hipclaw_ctx* ctx, int minlevel, int maxlevel);
/* Copy of all registers (num_patches * 8*meqn*(mx+my) doubles, slot order). */
int hipclaw_get_registers(hipclaw_ctx* ctx, double* host);
/* Execution features of hipclaw_advance (results are bitwise the same
 * either way):
 *   MERGE_LEVELS  (default on) one patch-step launch for all levels due at a
 *                 subcycled time, and one launch per ghost-fill phase and per
 *                 time-sync phase for all levels
 *   GHOST_PUSH    the step writes updated edge cells into the ghost cells of
 *                 same-level neighbours, replacing the same-level copies of the
 *                 next ghost fill (not used while time sync is on or with a
 *                 source term).  Default: on for patches of at least 32x32
 *                 cells, off below (the staging costs more than the copy it
 *                 saves); setting the feature applies it to every patch size
 *   GRAPH         each coarse step is captured once as a hipGraph and
 *                 replayed while the pool parities, the mesh and the features
 *                 are unchanged, for every t and dt: the step launches read dt
 *                 and the swirl time factor from a device parameter block
 *                 uploaded before each replay.  Not used while profiling,
 *                 with MERGE_LEVELS off, or with the swirl field stored in
 *                 aux (FUSE_AUX off).  Default when the context's stream is
 *                 not the null stream: graphs for coarse steps of more than
 *                 one fine step (subcycled or multi-rate schedules), eager
 *                 launches otherwise; setting the feature on or off applies
 *                 it to every coarse step. */
#define HIPCLAW_FEATURE_MERGE_LEVELS 1
#define HIPCLAW_FEATURE_GHOST_PUSH 2
#define HIPCLAW_FEATURE_GRAPH 3
/*   INIT_GHOSTCELL (off by default, as the reference's init_ghostcell option,
 *                 src/fclaw_options.c:186) ghost updates between the initial
 *                 refinement passes of hipclaw_initialize and after them; off,
 *                 the tags and the first step see the qinit ghost values */
#define HIPCLAW_FEATURE_INIT_GHOSTCELL 4
/* Swirl advection: evaluate the velocity field inside the step kernel instead
 * of storing it in aux before each step (default on; the same values). */
#define HIPCLAW_FEATURE_FUSE_AUX 5
int hipclaw_set_feature(hipclaw_ctx* ctx, int feature, int on);
/* Coarse steps captured as graphs and replays of captured graphs so far. */
int hipclaw_graph_stats(hipclaw_ctx* ctx, long long* captures, long long* replays);
/* Patch-step launch shape used by this context (NULL = library default);
 * _get returns it (cols_per_lane 0 = the library default). */
int hipclaw_set_launch(hipclaw_ctx* ctx, const hipclaw_launch* launch);
int hipclaw_get_launch(hipclaw_ctx* ctx, hipclaw_launch* launch);
/* Autotuner: times every launch shape (cols_per_lane 1..2 x waves_per_group
 * 1, 2, 4; reps timed launches each, after one untimed) on the current data of
 * one level (level < 0: the level with the most patches), writing into a
 * temporary buffer, and keeps the fastest for this context with row_bands =
 * -1 (launches of fewer than 2,048 patches keep the library default).  The kernel
 * variant timed is the one the level updates launch (with the same-level
 * ghost push when that is on).  ms (optional) receives the 6 average times in
 * that order.  No reference counterpart: the reference has one fixed loop
 * nest per patch. */
int hipclaw_autotune(hipclaw_ctx* ctx, int level, int reps, hipclaw_launch* best, float* ms);
/* Per-patch host arrays of length num_patches (any pointer may be NULL). */
int hipclaw_patch_info(hipclaw_ctx* ctx, int* level, int* gx, int* gy, double* xlower,
                       double* ylower, double* dx, double* dy);
/* Per-patch host arrays of one level (n = patches of the level, in slot
 * order; each array T*meqn doubles, aux T*maux or aux = NULL) to / from the
 * context: the clawpatch soln / aux arrays (fclaw2d_clawpatch_soln_data,
 * src/patches/clawpatch/fclaw2d_clawpatch.cpp:166-189).  Uploading aux
 * switches off the swirl velocity field. */
int hipclaw_level_upload(hipclaw_ctx* ctx, int level, int n, const double* const* q,
                         const double* const* aux);
int hipclaw_level_download(hipclaw_ctx* ctx, int level, int n, double* const* q);
/* Whole solution in leaf order, FP64 on the host, T*meqn per patch. */
int hipclaw_set_q(hipclaw_ctx* ctx, const double* host_q);
int hipclaw_get_q(hipclaw_ctx* ctx, double* host_q);
/* Swirl velocity field recomputed before every patch step (b4step2), period
 * tperiod; tperiod = 0 keeps the aux arrays as set.  With
 * HIPCLAW_FEATURE_FUSE_AUX on (the default) the step kernel evaluates the
 * field itself and the aux arrays are not written, so after tperiod = 0 they
 * hold what hipclaw_set_aux / hipclaw_level_upload or an unfused step left. */
int hipclaw_set_swirl(hipclaw_ctx* ctx, double tperiod);
int hipclaw_set_aux(hipclaw_ctx* ctx, const double* host_aux);

/* patch_vt->initialize on every leaf.  Cell averages use nsub x nsub
 * sub-cell samples; for HIPCLAW_QINIT_RADIAL_DAM, nsub = 0 selects the
 * reference's exact polygon average (cellave2 with the radialdam fdisc,
 * src/fortran_source2d/cellave2.f:1-95, radialdam/user_4.6/qinit.f:25-33). */
int hipclaw_qinit(hipclaw_ctx* ctx, int kind, const double* params, int nparams, int nsub);

/* fclaw2d_ghost_update (src/fclaw2d_ghost_fill.c:838-856): average, copy,
 * physical BC, interpolate, physical BC over levels minlevel..maxlevel, with
 * level minlevel-1 supplying time-interpolated data when time_interp != 0. */
int hipclaw_ghost_update(hipclaw_ctx* ctx, int minlevel, int maxlevel, double sync_time,
                         int time_interp);

/* fclaw2d_update_single_step for one level (src/fclaw2d_update_single_step.c:73-103):
 * every patch of the level in one batched launch; *cfl = level maximum. */
int hipclaw_update_level(hipclaw_ctx* ctx, int level, double t, double dt, double* cfl);

/* save_time_step / restore_time_step of the run loop (src/fclaw2d_run.c:155-158,
 * 220-236; patch_vt->save_step / restore_step = clawpatch_save_step /
 * clawpatch_restore_step, src/patches/clawpatch/fclaw2d_clawpatch.cpp:445-471):
 * _save copies every patch's current solution (ghost cells included) into a
 * device save pool, _restore copies it back, so a step whose Courant number
 * exceeds max_cfl is retaken without leaving the device.  Both are
 * stream-ordered copies.  The save pool is allocated on the first save after
 * each mesh change; restore fails (HIPCLAW_ERR_STATE) when nothing was saved
 * on the current mesh. */
int hipclaw_save_step(hipclaw_ctx* ctx);
int hipclaw_restore_step(hipclaw_ctx* ctx);

/* fclaw2d_timeinterp (src/fclaw2d_timeinterp.c:51-57). */
int hipclaw_timeinterp(hipclaw_ctx* ctx, int level, double alpha);

/* fclaw2d_advance_all_levels (src/fclaw2d_advance.c:247-357): one coarse step
 * of size dt with subcycling, intermediate and final ghost updates; returns
 * the maximum Courant number over all level updates. */
int hipclaw_advance(hipclaw_ctx* ctx, double t, double dt, double* maxcfl);
/* nsteps (1..256) coarse steps of one dt, as the fixed-dt run loop takes them
 * between outputs and regrids (src/fclaw2d_run.c:150-245 with use_fixed_dt:
 * no save, no retake): step k starts at t + dt + ... + dt (k additions, as
 * the loop accumulates t_curr) and maxcfl[k] (NULL: not returned) is its
 * maximum Courant number, all read back once at the end.  Bitwise the same
 * as nsteps calls of hipclaw_advance; the steps are one hipGraph launch when
 * graphs are on. */
int hipclaw_advance_steps(hipclaw_ctx* ctx, double t, double dt, int nsteps, double* maxcfl);
/* Captures the hipGraph that hipclaw_advance_steps(nsteps) would replay from
 * the current state (mesh, pool parities), without running anything and
 * leaving the solution and counters unchanged; no effect when graphs are off.
 * No counterpart in the reference: it moves graph capture out of the first
 * call (e.g. ahead of a latency-bound serving loop). */
int hipclaw_prepare_steps(hipclaw_ctx* ctx, int nsteps);

/* Tagging variable (0-based field ivar, the reference's threshold-variable
 * minus 1) and thresholds of the regrid tag pass. */
int hipclaw_set_tagging(hipclaw_ctx* ctx, int ivar, double refine_threshold,
                        double coarsen_threshold);
/* Refinement criterion (clawpatch option refinement-criteria,
 * src/patches/clawpatch/fclaw2d_clawpatch_options.c:107-110, values of
 * fclaw2d_clawpatch_options.h:48-54; the library default is GRADIENT):
 *   VALUE       q > threshold at some cell, ghost cells included
 *   DIFFERENCE  max(|dq_x|, |dq_y|) > threshold at some interior cell
 *   MINMAX      max q - min q > threshold over the patch, ghost cells included
 *               (coarsening: over the four siblings together)
 *   GRADIENT    |grad q| > threshold at some interior cell
 * (fort_user/fclaw2d_clawpatch_*_exceeds_th.f90 with the tag loops of
 * fort_4.6/fclaw2d_clawpatch46_tag4refinement.f / _tag4coarsening.f). */
#define HIPCLAW_TAG_VALUE 0
#define HIPCLAW_TAG_DIFFERENCE 1
#define HIPCLAW_TAG_MINMAX 2
#define HIPCLAW_TAG_GRADIENT 3
int hipclaw_set_tag_criterion(hipclaw_ctx* ctx, int criterion);
/* fclaw2d_domain_set_refinement (src/forestclaw2d.c:1386-1397): refinement
 * smoothing (the smooth-refine / smooth-level options; every patch takes the
 * largest target level of itself and its face and corner neighbours of
 * level >= smooth_level, src/fclaw2d_convenience.c:506-643).  coarsen_delay
 * must be 0 (delayed coarsening is not built). */
int hipclaw_set_refinement(hipclaw_ctx* ctx, int smooth_refine, int smooth_level, int coarsen_delay);
/* fclaw2d_regrid (src/fclaw2d_regrid.c:234-335): tag, adapt with 2:1 balance,
 * repopulate (interpolate2fine / average2coarse), rebuild plans, ghost update.
 * init_flag = 1 disables coarsening as in the initial refinement. */
int hipclaw_regrid(hipclaw_ctx* ctx, int init_flag, int* changed);
/* fclaw2d_initialize: qinit on minlevel, then refine / re-initialise until the
 * forest stops changing (at most maxlevel-minlevel passes); ghost cells filled. */
int hipclaw_initialize(hipclaw_ctx* ctx, int qinit_kind, const double* params, int nparams,
                       int nsub);

/* ASCII frames in the Clawpack 4.6 format of fclaw2d_clawpatch_output_ascii
 * (src/patches/clawpatch/fclaw2d_clawpatch_output_ascii.c:32-118 and the
 * writers of src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_output_ascii.f):
 * <dir>/fort.tXXXX (time, meqn, ngrids, num_aux, num_dim) and
 * <dir>/fort.qXXXX (every patch in the reference's order: blocks, then Morton
 * order within a block, numbered from 0).  dir NULL or "" = current directory.
 * The values are the context's solution (FP64; FP32 contexts widen). */
int hipclaw_output_ascii(hipclaw_ctx* ctx, int iframe, double time, const char* dir);
/* The two Fortran writers themselves, for host-resident patches:
 * fclaw2d_clawpatch46_fort_header_ascii (creates fort.tXXXX, empties
 * fort.qXXXX) and fclaw2d_clawpatch46_fort_output_ascii (appends one patch;
 * q is the (mx+2mbc) x (my+2mbc) x meqn tile, meqn <= 20). */
int hipclaw_ascii_header(const char* dir, int iframe, double time, int meqn, int maux, int ngrids);
int hipclaw_ascii_patch(const char* dir, int iframe, int mx, int my, int meqn, int mbc, double xlower,
                        double ylower, double dx, double dy, const double* q, int patch_num, int level,
                        int blockno, int mpirank);
/* Frame diff of two runs on the same mesh (scripts/compare_files/
 * compare_files.f90): reads <dir1>/fort.{t,q}XXXX and <dir2>/fort.{t,q}XXXX,
 * fails like the reference tool when meqn, the number of grids, a grid's
 * number, mx, xlow or ylow differ or the times differ by more than 1e-5, and
 * writes into dir3 the difference frame q1 - q2 (fort.qXXXX in the writers'
 * format, values below 1e-99 as zero, and fort.tXXXX) and
 * compare_summary.XXXX (per field the largest |q1 - q2| and the grid where it
 * first occurs; per grid the largest |q1 - q2| of every field).
 * global_max / grid_max (nmax entries each, either may be NULL) receive the
 * per-field maxima and grids.  meqn <= 10, as the reference's format. */
int hipclaw_compare_frames(const char* dir1, const char* dir2, const char* dir3, int iframe, int nmax,
                           double* global_max, int* grid_max);
/* FP64 conservation sums sum(q dx dy) per field (meqn doubles). */
int hipclaw_mass(hipclaw_ctx* ctx, double* sums);
/* Cell updates performed so far (count_single_step * mx * my). */
int hipclaw_cell_updates(hipclaw_ctx* ctx, long long* n);

/* HIP-event timing of every patch-step launch on the context's stream (the
 * ADVANCE_STEP2 timer of src/fclaw_timer.h on device events).  enable != 0
 * starts a fresh record; _read synchronises and returns the summed kernel
 * time, the number of launches and the interior cells they updated. */
int hipclaw_profile(hipclaw_ctx* ctx, int enable);
int hipclaw_profile_read(hipclaw_ctx* ctx, double* kernel_ms, long long* launches,
                         long long* cells);

#ifdef __cplusplus
}
#endif
#endif /* HIPCLAW_H */
