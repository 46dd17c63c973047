// hc_plan.hpp -- host construction of the ghost-fill plans from the forest
// (host-only; shared by the context and the CPU plan checker).
#pragma once
#include <vector>

#include "hc_amr.hpp"
#include "hc_forest.hpp"

namespace hc {

struct HostPlans {
    std::vector<GhostOp> avg[MAXL], copy[MAXL], interp[MAXL];
    std::vector<BcOp> bc[MAXL];
    std::vector<TiOp> ti[MAXL];
    std::vector<int> has_fine;  // per leaf: a finer face or corner neighbour
    std::vector<int> coarse_faces;  // per leaf: bit f set when face f has a coarser neighbour
    // time sync: corrections of level-L patches, and register faces (slot*4 +
    // face) reset after a sync that includes level L
    // (fclaw2d_clawpatch_time_sync_reset, fclaw2d_clawpatch_conservation.c:207-280):
    //   ts_same: SAMESIZE faces of level-L patches
    //   ts_f2c : HALFSIZE faces of level-L patches + DOUBLESIZE faces of level L+1
    //   ts_phys: physical-boundary faces of level-L patches
    std::vector<TsOp> ts[MAXL];
    std::vector<int> ts_same[MAXL], ts_f2c[MAXL], ts_phys[MAXL];
};

// Turns the per-pair dispatch of cb_face_fill / cb_corner_fill
// (src/fclaw2d_face_neighbors.c:195-450, src/fclaw2d_corner_neighbors.c:334-535)
// into flat per-level operation lists.  Slot k is leaf k.
void build_host_plans(const Forest& forest, int mx, int my, const int mthbc[4], HostPlans& out);

// Same-level ghost push table of the step kernel: 8 ints per slot, the
// neighbours that receive its edge cells across faces 0..3 and corners 0..3
// (-1: none).  A copy op (dst k, face / corner c, src nb) becomes "nb writes
// into k across its opposite face / corner".
void build_push_table(const HostPlans& hp, int nslots, std::vector<int>& push);

}  // namespace hc
