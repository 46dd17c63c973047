// hc_forest.hpp -- minimal quadtree forest over an mi x mj brick of unit
// blocks (single block = 1 x 1), standing in for the p4est topology that the
// reference reaches through src/forestclaw2d.c.
//
// A leaf is (level, gx, gy): its integer position on the level's uniform
// grid of (mi << level) x (mj << level) patches spanning the whole brick.
// All blocks of a brick share one orientation, so every inter-block
// transform reduces to an index shift in these global coordinates (the
// same-block branch of fclaw2d_patch_transform_face/face2/corner/corner2,
// src/forestclaw2d.c:655-663, 798-841).
#pragma once
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace hc {

enum NbrKind {
    NB_SAME = 0,      // FCLAW2D_PATCH_SAMESIZE
    NB_FINE = 1,      // FCLAW2D_PATCH_HALFSIZE  (neighbour is finer)
    NB_COARSE = -1,   // FCLAW2D_PATCH_DOUBLESIZE (neighbour is coarser)
    NB_BOUNDARY = -2, // physical boundary
    NB_NONE = -3      // hanging corner: no corner neighbour
};

struct Leaf {
    int level;
    int gx, gy;
};

struct Transfer {      // how a leaf of the new forest gets its data
    enum { KEEP, REFINE, COARSEN } kind;
    int new_leaf;
    int old_leaf[4];   // KEEP: [0]; REFINE: parent in [0]; COARSEN: 4 children (igrid order)
    int igrid;         // REFINE: child position 0..3 (x fastest)
};

class Forest {
public:
    Forest(int mi, int mj, int minlevel, int maxlevel);

    int mi, mj, minlevel, maxlevel;
    std::vector<Leaf> leaves;  // sorted by level, then block-major Morton order

    int nleaves() const { return (int)leaves.size(); }
    int find(int level, int gx, int gy) const;
    bool inside(int level, int gx, int gy) const
    {
        return gx >= 0 && gy >= 0 && gx < (mi << level) && gy < (mj << level);
    }
    int global_minlevel() const;
    int global_maxlevel() const;
    int block_of(const Leaf& l) const { return (l.gy >> l.level) * mi + (l.gx >> l.level); }

    void init_uniform(int level);
    // Replace the leaves by a given list.  Returns an empty string on success,
    // else why the list is not a valid forest: every leaf inside the brick and
    // within [minlevel, maxlevel], no duplicates or nested leaves, an exact
    // cover of the brick, and 2:1 balance across faces and corners.
    std::string set_leaves(const std::vector<Leaf>& in);

    // face 0 left, 1 right, 2 bottom, 3 top.  Returns the kind; nb[0..1]
    // receive neighbour leaves (two for NB_FINE, ordered by increasing
    // coordinate along the face, i.e. igrid 0 then 1).
    int face_neighbors(int leaf, int face, int nb[2]) const;
    // corner 0 lower-left, 1 lower-right, 2 upper-left, 3 upper-right.
    int corner_neighbor(int leaf, int corner, int* nb) const;

    // Refine (+1) / coarsen (-1, needs the whole family) / keep (0), then
    // restore 2:1 balance across faces and corners.  Fills the transfer list
    // describing how each new leaf is populated from the old forest.
    void adapt(const std::vector<int>& tags, std::vector<Transfer>& out);
    // Refinement smoothing of fclaw2d_domain_adapt (src/fclaw2d_convenience.c:506-643):
    // every leaf takes the largest target level (level + tag) over itself and
    // its face and corner neighbours of level >= smooth_level; the result is
    // refine (+1) if that exceeds its level, coarsen (-1) if it is below, else 0.
    std::vector<int> smooth_marks(const std::vector<int>& tags, int smooth_level) const;

private:
    std::unordered_map<uint64_t, int> index_;
    static uint64_t key(int level, int gx, int gy)
    {
        return ((uint64_t)level << 58) | ((uint64_t)(uint32_t)gx << 29) | (uint64_t)(uint32_t)gy;
    }
    void sort_and_index();
    int covering(int level, int gx, int gy) const;  // leaf at level <= given containing the cell
};

}  // namespace hc
