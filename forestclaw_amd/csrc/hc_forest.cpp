// hc_forest.cpp -- quadtree forest: neighbour queries and 2:1-balanced adapt.
//
// Neighbour semantics follow what ForestClaw receives from p4est:
//   * face neighbours: same size, one double-size leaf, or two half-size
//     leaves ordered along the face (fclaw2d_patch_face_neighbors,
//     src/forestclaw2d.c:443-540);
//   * corner neighbours: the leaf diagonally across the corner that touches
//     only the corner; none when the corner is a hanging node, i.e. lies in
//     the middle of a coarser face neighbour's edge (p4est mesh corner
//     lists, src/forestclaw2d.c:986-1116, case 2 of
//     src/fclaw2d_corner_neighbors.c:103-121).
// Balance is enforced across faces and corners (p4est P4EST_CONNECT_FULL).
#include "hc_forest.hpp"

#include <algorithm>
#include <array>
#include <tuple>
#include <stdexcept>

namespace hc {

Forest::Forest(int mi_, int mj_, int minlevel_, int maxlevel_)
    : mi(mi_), mj(mj_), minlevel(minlevel_), maxlevel(maxlevel_)
{
    if (mi < 1 || mj < 1 || minlevel < 0 || maxlevel < minlevel || maxlevel > 20)
        throw std::invalid_argument("bad forest dimensions");
}

static uint64_t morton(uint32_t x, uint32_t y)
{
    uint64_t m = 0;
    for (int b = 0; b < 31; b++) {
        m |= (uint64_t)((x >> b) & 1u) << (2 * b);
        m |= (uint64_t)((y >> b) & 1u) << (2 * b + 1);
    }
    return m;
}

void Forest::sort_and_index()
{
    auto sortkey = [this](const Leaf& l) {
        int bx = l.gx >> l.level, by = l.gy >> l.level;
        uint32_t lx = (uint32_t)(l.gx - (bx << l.level)) << (24 - l.level);
        uint32_t ly = (uint32_t)(l.gy - (by << l.level)) << (24 - l.level);
        return std::make_tuple(l.level, by * mi + bx, morton(lx, ly));
    };
    std::sort(leaves.begin(), leaves.end(),
              [&](const Leaf& a, const Leaf& b) { return sortkey(a) < sortkey(b); });
    index_.clear();
    index_.reserve(leaves.size() * 2);
    for (int i = 0; i < (int)leaves.size(); i++)
        index_[key(leaves[i].level, leaves[i].gx, leaves[i].gy)] = i;
}

int Forest::find(int level, int gx, int gy) const
{
    if (level < 0 || !inside(level, gx, gy)) return -1;
    auto it = index_.find(key(level, gx, gy));
    return it == index_.end() ? -1 : it->second;
}

int Forest::covering(int level, int gx, int gy) const
{
    for (int l = level; l >= 0; l--) {
        int k = find(l, gx >> (level - l), gy >> (level - l));
        if (k >= 0) return k;
    }
    return -1;
}

int Forest::global_minlevel() const
{
    int m = 1 << 30;
    for (auto& l : leaves) m = std::min(m, l.level);
    return m;
}

int Forest::global_maxlevel() const
{
    int m = -1;
    for (auto& l : leaves) m = std::max(m, l.level);
    return m;
}

void Forest::init_uniform(int level)
{
    leaves.clear();
    for (int gy = 0; gy < (mj << level); gy++)
        for (int gx = 0; gx < (mi << level); gx++) leaves.push_back({level, gx, gy});
    sort_and_index();
}

std::string Forest::set_leaves(const std::vector<Leaf>& in)
{
    if (in.empty()) return "empty leaf list";
    int top = 0;
    for (const Leaf& l : in) {
        if (l.level < minlevel || l.level > maxlevel) return "leaf level outside [minlevel, maxlevel]";
        if (!inside(l.level, l.gx, l.gy)) return "leaf outside the brick";
        top = std::max(top, l.level);
    }
    // Cover: the areas (in finest-level cells) add up to the brick and no
    // leaf has an ancestor in the list; together these make an exact cover.
    unsigned long long area = 0;
    for (const Leaf& l : in) area += 1ull << (2 * (top - l.level));
    if (area != (unsigned long long)mi * mj << (2 * top)) return "leaves do not cover the brick exactly";
    std::vector<Leaf> old;
    old.swap(leaves);
    leaves = in;
    sort_and_index();
    std::string err;
    if (index_.size() != leaves.size()) err = "duplicate leaves";
    for (int k = 0; err.empty() && k < (int)leaves.size(); k++) {
        const Leaf& l = leaves[k];
        for (int a = l.level - 1; a >= 0 && err.empty(); a--)
            if (find(a, l.gx >> (l.level - a), l.gy >> (l.level - a)) >= 0) err = "nested leaves";
    }
    // 2:1 balance: the leaf covering each of the eight neighbours of a leaf
    // at its own level is at most one level coarser.  (A finer neighbour is
    // checked from its own side.)
    for (int k = 0; err.empty() && k < (int)leaves.size(); k++) {
        const Leaf& l = leaves[k];
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                if ((dx == 0 && dy == 0) || !inside(l.level, l.gx + dx, l.gy + dy)) continue;
                int c = covering(l.level, l.gx + dx, l.gy + dy);
                if (c >= 0 && leaves[c].level < l.level - 1) err = "leaves are not 2:1 balanced";
            }
    }
    if (!err.empty()) {
        leaves.swap(old);
        sort_and_index();
    }
    return err;
}

int Forest::face_neighbors(int leaf, int face, int nb[2]) const
{
    const Leaf& l = leaves[leaf];
    static const int DX[4] = {-1, 1, 0, 0}, DY[4] = {0, 0, -1, 1};
    const int nx = l.gx + DX[face], ny = l.gy + DY[face];
    nb[0] = nb[1] = -1;
    if (!inside(l.level, nx, ny)) return NB_BOUNDARY;
    int k = find(l.level, nx, ny);
    if (k >= 0) {
        nb[0] = k;
        return NB_SAME;
    }
    if (l.level > 0) {
        k = find(l.level - 1, nx >> 1, ny >> 1);
        if (k >= 0) {
            nb[0] = k;
            return NB_COARSE;
        }
    }
    int c0x, c0y, c1x, c1y;
    switch (face) {
    case 0: c0x = c1x = 2 * nx + 1; c0y = 2 * ny; c1y = 2 * ny + 1; break;
    case 1: c0x = c1x = 2 * nx; c0y = 2 * ny; c1y = 2 * ny + 1; break;
    case 2: c0y = c1y = 2 * ny + 1; c0x = 2 * nx; c1x = 2 * nx + 1; break;
    default: c0y = c1y = 2 * ny; c0x = 2 * nx; c1x = 2 * nx + 1; break;
    }
    nb[0] = find(l.level + 1, c0x, c0y);
    nb[1] = find(l.level + 1, c1x, c1y);
    if (nb[0] < 0 || nb[1] < 0) throw std::logic_error("forest is not 2:1 balanced across a face");
    return NB_FINE;
}

int Forest::corner_neighbor(int leaf, int corner, int* nb) const
{
    const Leaf& l = leaves[leaf];
    const int dx = (corner & 1) ? 1 : -1, dy = (corner & 2) ? 1 : -1;
    const int nx = l.gx + dx, ny = l.gy + dy;
    *nb = -1;
    if (!inside(l.level, nx, ny)) return NB_BOUNDARY;
    int k = find(l.level, nx, ny);
    if (k >= 0) {
        *nb = k;
        return NB_SAME;
    }
    k = find(l.level + 1, 2 * nx + (dx > 0 ? 0 : 1), 2 * ny + (dy > 0 ? 0 : 1));
    if (k >= 0) {
        *nb = k;
        return NB_FINE;
    }
    const int vx = l.gx + (dx > 0 ? 1 : 0), vy = l.gy + (dy > 0 ? 1 : 0);
    if (l.level > 0 && (vx % 2) == 0 && (vy % 2) == 0) {
        k = find(l.level - 1, nx >> 1, ny >> 1);
        if (k >= 0) {
            *nb = k;
            return NB_COARSE;
        }
    }
    return NB_NONE;
}

std::vector<int> Forest::smooth_marks(const std::vector<int>& tags, int smooth_level) const
{
    const int n = nleaves();
    if ((int)tags.size() != n) throw std::invalid_argument("tag count mismatch");
    std::vector<int> tgt(n), out(n, 0);
    for (int k = 0; k < n; k++) tgt[k] = leaves[k].level + (tags[k] > 0 ? 1 : tags[k] < 0 ? -1 : 0);
    for (int k = 0; k < n; k++) {
        const int level = leaves[k].level;
        int mt = tgt[k];
        auto take = [&](int nb) {
            if (nb >= 0 && leaves[nb].level >= smooth_level) mt = std::max(mt, tgt[nb]);
        };
        for (int f = 0; mt <= level && f < 4; f++) {
            int nb[2];
            const int rel = face_neighbors(k, f, nb);
            if (rel == NB_SAME || rel == NB_COARSE) take(nb[0]);
            else if (rel == NB_FINE) {
                take(nb[0]);
                take(nb[1]);
            }
        }
        for (int c = 0; mt <= level && c < 4; c++) {
            int nb = -1;
            corner_neighbor(k, c, &nb);
            take(nb);
        }
        out[k] = mt < level ? -1 : (mt > level ? 1 : 0);
    }
    return out;
}

void Forest::adapt(const std::vector<int>& tags, std::vector<Transfer>& out)
{
    if (tags.size() != leaves.size()) throw std::invalid_argument("tag count mismatch");
    const std::vector<Leaf> old = leaves;

    // origin of each current leaf: old leaf index (KEEP) or parent (REFINE, -2-parent)
    struct Cur { Leaf l; int origin; int igrid; int tag; };
    std::vector<Cur> cur;
    cur.reserve(old.size() * 2);
    for (int i = 0; i < (int)old.size(); i++) {
        const Leaf& l = old[i];
        if (tags[i] > 0 && l.level < maxlevel) {
            for (int c = 0; c < 4; c++)
                cur.push_back({{l.level + 1, 2 * l.gx + (c & 1), 2 * l.gy + (c >> 1)}, i, c, 0});
        } else {
            cur.push_back({l, i, -1, tags[i]});
        }
    }
    auto rebuild = [&]() {
        leaves.clear();
        for (auto& c : cur) leaves.push_back(c.l);
        index_.clear();
        for (int i = 0; i < (int)cur.size(); i++) index_[key(cur[i].l.level, cur[i].l.gx, cur[i].l.gy)] = i;
    };
    rebuild();

    // 2:1 balance by refinement (faces and corners)
    for (;;) {
        std::vector<char> mark(cur.size(), 0);
        bool any = false;
        for (int i = 0; i < (int)cur.size(); i++) {
            const Leaf l = cur[i].l;
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    if (!dx && !dy) continue;
                    if (!inside(l.level, l.gx + dx, l.gy + dy)) continue;
                    int c = covering(l.level, l.gx + dx, l.gy + dy);
                    if (c >= 0 && cur[c].l.level < l.level - 1 && !mark[c]) {
                        mark[c] = 1;
                        any = true;
                    }
                }
        }
        if (!any) break;
        std::vector<Cur> nxt;
        nxt.reserve(cur.size() + 64);
        for (int i = 0; i < (int)cur.size(); i++) {
            if (!mark[i]) {
                nxt.push_back(cur[i]);
                continue;
            }
            if (cur[i].igrid >= 0) throw std::logic_error("balance refined a new leaf twice");
            const Leaf l = cur[i].l;
            for (int c = 0; c < 4; c++)
                nxt.push_back({{l.level + 1, 2 * l.gx + (c & 1), 2 * l.gy + (c >> 1)}, cur[i].origin, c, 0});
        }
        cur.swap(nxt);
        rebuild();
    }

    // coarsen complete families of untouched old leaves tagged -1, when the
    // parent stays balanced against every leaf around it
    std::vector<char> gone(cur.size(), 0);
    std::vector<Leaf> parents;
    std::vector<std::array<int, 4>> families;  // old leaf indices, igrid order
    for (int i = 0; i < (int)cur.size(); i++) {
        const Cur& c = cur[i];
        if (gone[i] || c.igrid >= 0 || c.tag >= 0 || c.l.level <= minlevel) continue;
        if ((c.l.gx & 1) || (c.l.gy & 1)) continue;  // visit each family once, from child 0
        int fam[4];
        bool ok = true;
        for (int k = 0; k < 4 && ok; k++) {
            int j = find(c.l.level, c.l.gx + (k & 1), c.l.gy + (k >> 1));
            if (j < 0 || gone[j] || cur[j].igrid >= 0 || cur[j].tag >= 0) ok = false;
            fam[k] = j;
        }
        if (!ok) continue;
        const int L = c.l.level, px = c.l.gx >> 1, py = c.l.gy >> 1;
        for (int dy = -1; dy <= 1 && ok; dy++)
            for (int dx = -1; dx <= 1 && ok; dx++) {
                if (!dx && !dy) continue;
                if (!inside(L - 1, px + dx, py + dy)) continue;
                for (int k = 0; k < 4 && ok; k++) {
                    // only the children of the neighbour that touch the parent
                    if (dx == -1 && (k & 1) == 0) continue;
                    if (dx == 1 && (k & 1) == 1) continue;
                    if (dy == -1 && (k >> 1) == 0) continue;
                    if (dy == 1 && (k >> 1) == 1) continue;
                    int cx = 2 * (px + dx) + (k & 1), cy = 2 * (py + dy) + (k >> 1);
                    if (covering(L, cx, cy) < 0) ok = false;  // finer leaves there
                }
            }
        if (!ok) continue;
        for (int k = 0; k < 4; k++) gone[fam[k]] = 1;
        parents.push_back({L - 1, px, py});
        families.push_back({cur[fam[0]].origin, cur[fam[1]].origin, cur[fam[2]].origin,
                            cur[fam[3]].origin});
    }

    std::vector<Cur> fin;
    for (int i = 0; i < (int)cur.size(); i++)
        if (!gone[i]) fin.push_back(cur[i]);
    const int nkeep = (int)fin.size();
    for (auto& p : parents) fin.push_back({p, -1, -1, 0});

    // final order and transfer list
    leaves.clear();
    for (auto& c : fin) leaves.push_back(c.l);
    sort_and_index();
    out.clear();
    out.reserve(fin.size());
    for (int i = 0; i < (int)fin.size(); i++) {
        const Cur& c = fin[i];
        Transfer t{};
        t.new_leaf = find(c.l.level, c.l.gx, c.l.gy);
        if (i >= nkeep) {
            t.kind = Transfer::COARSEN;
            const auto& f = families[i - nkeep];
            for (int k = 0; k < 4; k++) t.old_leaf[k] = f[k];
            t.igrid = -1;
        } else if (c.igrid >= 0) {
            t.kind = Transfer::REFINE;
            t.old_leaf[0] = c.origin;
            t.old_leaf[1] = t.old_leaf[2] = t.old_leaf[3] = -1;
            t.igrid = c.igrid;
        } else {
            t.kind = Transfer::KEEP;
            t.old_leaf[0] = c.origin;
            t.old_leaf[1] = t.old_leaf[2] = t.old_leaf[3] = -1;
            t.igrid = -1;
        }
        out.push_back(t);
    }
    std::sort(out.begin(), out.end(), [](const Transfer& a, const Transfer& b) { return a.new_leaf < b.new_leaf; });
}

}  // namespace hc
