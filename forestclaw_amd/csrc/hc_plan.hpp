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
};

// Turns the per-pair dispatch of cb_face_fill / cb_corner_fill
// (src/fclaw2d_face_neighbors.c:195-450, src/fclaw2d_corner_neighbors.c:334-535)
// into flat per-level operation lists.  Slot k is leaf k.
void build_host_plans(const Forest& forest, int mx, int my, const int mthbc[4], HostPlans& out);

}  // namespace hc
