// hc_ctx.hpp -- the device-resident solver context behind the hipclaw_ctx_*
// entry points: forest, slot pools, ghost-fill plans, subcycling state.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <utility>
#include <vector>

#include "../../include/hipclaw.h"
#include "../../include/hipclaw_cnn.h"
#include "hc_amr.hpp"
#include "hc_amr_launch.hpp"
#include "hc_forest.hpp"
#include "hc_internal.hpp"

namespace hc {

template <class T> struct DevArray {
    T* p = nullptr;
    int n = 0;
    void free_()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int upload(const std::vector<T>& v);
};

struct LevelPlan {
    DevArray<GhostOp> avg, avg_ti;       // coarse = this level
    DevArray<GhostOp> copy, copy_ti;     // this level
    DevArray<GhostOp> interp, interp_ti; // coarse = this level
    DevArray<BcOp> bc, bc_ti;
    DevArray<TiOp> ti;                   // patches with finer neighbours
    DevArray<TsOp> ts;                   // time-sync corrections of this level's patches
    DevArray<int> ts_same, ts_f2c, ts_phys;  // register faces reset by a sync
    DevArray<BcTsOp> bcts, bcts_ti;      // physical BCs + corrections per patch
    int s0 = 0, s1 = 0;                  // slot range of the level
};

struct Ctx {
    hipclaw_physics phys{};
    hipclaw_domain dom{};
    Forest forest;
    hipStream_t stream = nullptr;
    int esz = 8;
    long long tile = 0, atile = 0;  // elements per slot (q, aux)
    int n = 0, cap = 0;
    void* pool[2] = {nullptr, nullptr};
    void* ti = nullptr;
    void* save = nullptr;
    void* aux = nullptr;
    double* d_xylower = nullptr;   // 2 per slot
    double* d_dxdy = nullptr;      // 2 per slot
    double* d_cfl = nullptr;       // per slot
    double* d_cflacc = nullptr;    // [4095]: the last coarse step's maximum Courant number
    // Courant-number slots of a coarse step: one per step launch, the bits of
    // a non-negative double (small launches: atomic max from the step kernel,
    // so zero before it; larger ones: stored by a reduction); cleared by the
    // reduction that reads them (cfl_final)
    unsigned long long* d_cflslot = nullptr;
    bool cflslot_dirty = false;    // slots possibly not cleared (an interrupted step)
    // time interpolations (level, weight) of a coarse step waiting for the
    // next ghost fill, whose average launch runs them (merged levels)
    std::vector<std::pair<int, double>> ti_pending;
    int clear_cfl_slots();
    int* d_levels = nullptr;       // level per slot
    int* d_iota = nullptr;         // 0..cap-1
    int* d_tags = nullptr;         // per slot
    double* d_mass = nullptr;      // per slot * meqn
    int parity[MAXL] = {};
    LevelPlan plan[MAXL];
    std::vector<int> has_fine;     // per slot
    std::vector<int> level_of;     // per slot
    double swirl_tperiod = 0.0;    // > 0: advection velocities from the swirl stream function
    int cflacc_count = 0;
    hipclaw_launch launch{0, 0};   // patch-step launch shape; 0 = library default
    hipclaw_cnn* cnn = nullptr;    // learned refinement tags (not owned)
    float cnn_margin = 0.0f;
    void* d_tiles = nullptr;       // network input of every slot (bf16)
    float* d_logits = nullptr;
    int tiles_cap = 0;
    int tag_stats[6] = {};         // patches, net refine, plain refine, misses, false refines, fallbacks
    bool tsync = false;            // conservative time synchronisation on
    int src2 = HIPCLAW_SRC2_NONE;  // source term after every patch step (src_term)
    bool fuse_aux = true;          // swirl velocities evaluated in the step kernel
    double* d_reg = nullptr;       // time-sync registers, regstride doubles per slot (FP64)
    int* d_cfaces = nullptr;       // per slot: bit f set when face f has a coarser neighbour
    // Register state.  reg_clean: every register is zero as far as the
    // drivers are concerned (after a mesh change and after every coarse
    // step, whose last sync covers all levels); reg_stale: they are not
    // zero in memory (a merged coarse step writes its first sums onto zero
    // instead of clearing after each sync, StepArgs::reg_fresh) and are
    // cleared before anything else reads them (flush_registers).
    bool reg_clean = true, reg_stale = false;
    bool reg_fresh_run = false;    // the coarse steps being run or recorded work so
    int flush_registers();
    long long regstride = 0;
    // tagging parameters: clawpatch refinement-criteria (HIPCLAW_TAG_*),
    // threshold variable and thresholds; the domain's refinement strategy
    // (fclaw2d_domain_set_refinement: smooth-refine, smooth-level)
    double refine_threshold = 0.0, coarsen_threshold = 0.0;
    int tag_ivar = 0;
    int tag_criterion = HIPCLAW_TAG_GRADIENT;
    bool smooth_refine = false;
    int smooth_level = 0;
    double* d_tagst = nullptr;     // tag statistics, 4 doubles per slot
    int tagst_cap = 0;
    // ghost update after the initial refinement passes (init_ghostcell)
    bool init_ghostcell = false;
    long long cell_updates = 0;    // sum over patch-steps of mx*my
    // HIP-event timing of the patch-step launches (on this context's stream)
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::vector<long long> ev_cells;  // interior cells updated by each launch
    void prof_clear();

    explicit Ctx(const hipclaw_physics& p, const hipclaw_domain& d);
    ~Ctx();
    void release_pools();
    int rebuild(bool keep_data, const std::vector<Transfer>* xfer);
    template <class R> Pools<R> pools() const;
    // sync: the time sync of the same levels follows (time_sync).  cfl_out:
    // the fill ends a coarse step, whose Courant-number reduction (cfl_final
    // of the step's slots into cfl_out) the merged path runs in its last
    // launch, setting cfl_done
    int ghost_update(int minlevel, int maxlevel, int time_interp, bool sync = false, double* cfl_out = nullptr);
    bool cfl_done = false;
    int update_level(int level, double t, double dt);
    Src2Args src2_args(int s0, int np) const;
    bool fused_swirl() const;
    double swirl_vt(double t, double dt) const;
    // one merged launch for the contiguous levels [lo, hi] due at the same
    // subcycled time (t[L], dt[L] per level)
    int update_levels(int lo, int hi, const double* t, const double* dt);
    bool merge_levels = true;
    // same-level ghost push from the step kernel (off with time sync, whose
    // corrections change edge cells between a step and the next ghost fill)
    bool push_ghosts = true;
    // default: push only for patches of at least 32x32 cells; below that the
    // edge staging costs the step kernel more than the copy launch it saves
    bool push_auto = true;
    bool use_push() const
    {
        return push_ghosts && !tsync && !src2 && phys.mx >= 4 && phys.my >= 4 &&
               (!push_auto || (phys.mx >= 32 && phys.my >= 32));
    }
    int* d_push = nullptr;          // 8 ints per slot
    bool pushed[MAXL] = {};         // level data unchanged since a pushing step
    void clear_pushed()
    {
        for (bool& b : pushed) b = false;
    }
    int timeinterp(int level, double alpha);
    int advance(double t, double dt, double* maxcfl);
    // nstep coarse steps of one dt (t, t + dt, ... accumulated as the run
    // loop does); maxcfl[k] = the Courant number of step k
    // prepare: capture the graph these steps would replay (graph mode only)
    // and leave the state as it was, running nothing
    int advance_n(double t, double dt, int nstep, double* maxcfl, bool prepare = false);
    static constexpr int ADVANCE_MAXSTEPS = 256;
    int advance_body(double t, double dt, double* cfl_out = nullptr);
    // hipGraph of a whole coarse step, replayed when dt, the level pool
    // parities, the ghost-push state, the launch shape and the mesh match
    bool use_graph = false;
    bool graph_auto = false;  // default: graphs only for coarse steps of several fine steps
    long long mesh_epoch = 0;
    struct GraphEntry {
        unsigned parity_in, pushed_in;
        long long epoch;
        int nstep;
        bool reg_fresh;
        hipGraphExec_t exec;
        unsigned parity_out, pushed_out;
        long long cells;
        int cflacc;
    };
    std::vector<GraphEntry> graphs;
    long long graph_captures = 0, graph_replays = 0;
    void clear_graphs();
    // Step parameters of graph replays: per step launch of a coarse step,
    // dt and the swirl vt of every level (PAR_STRIDE doubles), computed on
    // the host by a dry pass of the schedule, staged in a pinned ring slot
    // and copied to d_par ahead of each graph launch on the context stream.
    static constexpr int PAR_STRIDE = 2 * STEP_MAXL;
    static constexpr int PAR_MAXLAUNCH = 4096;
    static constexpr int PAR_RING = 4;
    double* d_par = nullptr;
    double* h_par = nullptr;          // pinned, PAR_RING slots
    hipEvent_t par_ev[PAR_RING] = {};
    bool par_ev_used[PAR_RING] = {};
    int par_slot = 0;
    double* par_host = nullptr;       // slot being filled (nullptr: no recording)
    int par_n = 0;                    // launches recorded in this coarse step
    bool par_dry = false;             // schedule only: record parameters, launch nothing
    int ensure_par();
    void free_par();
    // save_time_step / restore_time_step of the run loop (clawpatch
    // griddata_save): every level's current tiles copied on the device
    bool saved = false;
    long long saved_epoch = -1;
    bool saved_pushed[MAXL] = {};
    int save_step();
    int restore_step();
    int qinit(int kind, const QinitParams& prm);
    int tag_and_adapt(int init_flag, int* changed);
    int mass(double* sums);
    int level_range(int level, int* s0, int* s1) const;
    int autotune(int level, int reps, hipclaw_launch* best, float* ms);
    int set_time_sync(bool on);
    int make_tiles(void* tiles, int* tags);
    int time_sync(int minlevel, int maxlevel);
    bool reset_segs(int minlevel, int maxlevel, ResetSegs& rs) const;
};

}  // namespace hc
