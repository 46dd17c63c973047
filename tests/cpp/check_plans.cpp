// check_plans.cpp -- CPU bounds and coverage check of the ghost-fill plans
// and regrid transfers that the GPU kernels execute.
//
// Builds forests (uniform, random 2:1-balanced AMR, 4x1 brick), constructs
// the plans with the library's own host code (hc_forest.cpp, hc_plan.cpp),
// and replays the index arithmetic of k_ghost / k_xfer
// (forestclaw_amd/csrc/hc_amr_kernels.hip) for every operation, checking
// that every read and write stays inside its tile and slot range.  It also
// reports ghost cells that no operation writes.
//
//   g++ -std=c++17 -O1 -I forestclaw_amd/csrc tests/cpp/check_plans.cpp \
//       forestclaw_amd/csrc/hc_forest.cpp forestclaw_amd/csrc/hc_plan.cpp
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <tuple>
#include <vector>

#include "hc_forest.hpp"
#include "hc_plan.hpp"

using namespace hc;

static int g_err = 0;
static long long g_checked = 0;

#define REQUIRE(c, ...)                          \
    do {                                         \
        g_checked++;                             \
        if (!(c)) {                              \
            if (g_err < 20) {                    \
                std::printf("FAIL: " __VA_ARGS__); \
                std::printf("\n");               \
            }                                    \
            g_err++;                             \
        }                                        \
    } while (0)

static const int MBC = 2;

struct Tile {
    int NX, NY;
    bool in(int i, int j) const { return i >= 0 && i < NX && j >= 0 && j < NY; }
};

// replay of k_ghost's index computations for one op; marks written ghost cells
static void replay(const GhostOp& op, int n, int mx, int my, std::vector<std::vector<char>>& written)
{
    const Tile t{mx + 2 * MBC, my + 2 * MBC};
    REQUIRE(op.dst >= 0 && op.dst < n && op.src >= 0 && op.src < n, "slot out of range (%d,%d)", op.dst, op.src);
    const int items = MBC * (mx > my ? mx : my);
    for (int k = 0; k < items; k++) {
        if (op.kind == G_COPY_FACE || op.kind == G_AVG_FACE) {
            const int f = op.sub;
            const int len = f < 2 ? my : mx;
            if (k >= MBC * len) continue;
            int i, j;
            const int layer = k % MBC, along = MBC + k / MBC;
            switch (f) {
            case 0: i = layer; j = along; break;
            case 1: i = MBC + mx + layer; j = along; break;
            case 2: j = layer; i = along; break;
            default: j = MBC + my + layer; i = along; break;
            }
            REQUIRE(t.in(i, j), "face dst (%d,%d)", i, j);
            if (op.kind == G_COPY_FACE) {
                REQUIRE(t.in(i + op.ox, j + op.oy), "copy src (%d,%d) kind %d sub %d", i + op.ox, j + op.oy, op.kind, op.sub);
                written[op.dst][j * t.NX + i] = 1;
            } else {
                const int fi = 2 * i - MBC + op.ox, fj = 2 * j - MBC + op.oy;
                if (fi < MBC || fi + 1 > MBC + mx - 1 || fj < MBC || fj + 1 > MBC + my - 1) continue;
                REQUIRE(t.in(fi, fj) && t.in(fi + 1, fj + 1), "avg src");
                written[op.dst][j * t.NX + i] = 1;
            }
        } else if (op.kind == G_COPY_CORNER || op.kind == G_AVG_CORNER) {
            if (k >= MBC * MBC) continue;
            const int ibc = k % MBC + 1, jbc = k / MBC + 1;
            const int c = op.sub;
            const int i = (c & 1) ? MBC + mx - 1 + ibc : MBC - ibc;
            const int j = (c & 2) ? MBC + my - 1 + jbc : MBC - jbc;
            REQUIRE(t.in(i, j), "corner dst");
            if (op.kind == G_COPY_CORNER) {
                REQUIRE(t.in(i + op.ox, j + op.oy), "copy corner src (%d,%d)", i + op.ox, j + op.oy);
            } else {
                const int fi = 2 * i - MBC + op.ox, fj = 2 * j - MBC + op.oy;
                REQUIRE(t.in(fi, fj) && t.in(fi + 1, fj + 1), "avg corner src (%d,%d)", fi, fj);
                // the four fine cells must be interior cells of the fine patch
                REQUIRE(fi >= MBC && fi + 1 <= MBC + mx - 1 && fj >= MBC && fj + 1 <= MBC + my - 1,
                        "avg corner reads fine ghost (%d,%d)", fi, fj);
            }
            written[op.dst][j * t.NX + i] = 1;
        } else {
            int ci, cj;
            bool check;
            if (op.kind == G_INTERP_FACE) {
                const int f = op.sub;
                const int len = f < 2 ? my : mx;
                if (k >= len) continue;
                switch (f) {
                case 0: ci = MBC; cj = MBC + k; break;
                case 1: ci = MBC + mx - 1; cj = MBC + k; break;
                case 2: cj = MBC; ci = MBC + k; break;
                default: cj = MBC + my - 1; ci = MBC + k; break;
                }
                check = true;
            } else {
                if (k >= 1) continue;
                const int c = op.sub;
                ci = (c & 1) ? MBC + mx - 1 : MBC;
                cj = (c & 2) ? MBC + my - 1 : MBC;
                check = false;
            }
            const int fi = 2 * ci - MBC + op.ox, fj = 2 * cj - MBC + op.oy;
            if (check && (fi < 0 || fi + 1 > t.NX - 1 || fj < 0 || fj + 1 > t.NY - 1)) continue;
            REQUIRE(t.in(ci - 1, cj - 1) && t.in(ci + 1, cj + 1), "interp coarse stencil");
            REQUIRE(t.in(fi, fj) && t.in(fi + 1, fj + 1), "interp fine dst (%d,%d) kind %d sub %d", fi, fj, op.kind, op.sub);
            for (int dj = 0; dj < 2; dj++)
                for (int di = 0; di < 2; di++) {
                    const int ii = fi + di, jj = fj + dj;
                    const bool ghost = ii < MBC || ii >= MBC + mx || jj < MBC || jj >= MBC + my;
                    REQUIRE(ghost, "interpolation writes an interior cell (%d,%d)", ii, jj);
                    written[op.dst][jj * t.NX + ii] = 1;
                }
        }
    }
}

static int floordiv2(int a) { return a >= 0 ? a / 2 : -((1 - a) / 2); }

static void check_xfer(const std::vector<Transfer>& xf, int nold, int nnew, int mx, int my)
{
    const Tile t{mx + 2 * MBC, my + 2 * MBC};
    for (const Transfer& x : xf) {
        REQUIRE(x.new_leaf >= 0 && x.new_leaf < nnew, "xfer dst");
        if (x.kind == Transfer::COARSEN) {
            for (int g = 0; g < 4; g++) REQUIRE(x.old_leaf[g] >= 0 && x.old_leaf[g] < nold, "xfer coarsen src");
        } else {
            REQUIRE(x.old_leaf[0] >= 0 && x.old_leaf[0] < nold, "xfer src");
        }
        if (x.kind == Transfer::REFINE) {
            const int ig = x.igrid & 1, jg = x.igrid >> 1;
            for (int j = 0; j < t.NY; j++)
                for (int i = 0; i < t.NX; i++) {
                    const int fi_f = i - MBC + 1, fj_f = j - MBC + 1;
                    const int icl = floordiv2(fi_f - 1) + 1, jcl = floordiv2(fj_f - 1) + 1;
                    const int ii = fi_f - (icl - 1) * 2, jj = fj_f - (jcl - 1) * 2;
                    REQUIRE(ii == 1 || ii == 2, "child offset %d", ii);
                    REQUIRE(jj == 1 || jj == 2, "child offset %d", jj);
                    if (icl < 1 - ig || icl > mx / 2 + 1 - ig || jcl < 1 - jg || jcl > my / 2 + 1 - jg) continue;
                    const int ci = icl + ig * (mx / 2) + MBC - 1, cj = jcl + jg * (my / 2) + MBC - 1;
                    REQUIRE(t.in(ci - 1, cj - 1) && t.in(ci + 1, cj + 1), "refine stencil (%d,%d)", ci, cj);
                }
        }
        if (x.kind == Transfer::COARSEN) {
            for (int j = 0; j < t.NY; j++)
                for (int i = 0; i < t.NX; i++) {
                    const int ci_f = i - MBC + 1, cj_f = j - MBC + 1;
                    if (ci_f < 1 || ci_f > mx || cj_f < 1 || cj_f > my) continue;
                    const int g = ((ci_f > mx / 2) ? 1 : 0) + 2 * ((cj_f > my / 2) ? 1 : 0);
                    const int il = ci_f - (g & 1) * (mx / 2), jl = cj_f - (g >> 1) * (mx / 2);
                    const int f0 = (il - 1) * 2 + 1 + MBC - 1, g0 = (jl - 1) * 2 + 1 + MBC - 1;
                    REQUIRE(f0 >= MBC && f0 + 1 <= MBC + mx - 1 && g0 >= MBC && g0 + 1 <= MBC + my - 1,
                            "coarsen reads (%d,%d)", f0, g0);
                }
        }
    }
}

// The step kernel's same-level ghost push (PUSH epilogue of k_step2,
// forestclaw_amd/csrc/hc_step2.hpp) must write exactly the cells, from exactly
// the sources, that the same-level copy ops of the plans would.
using CellMap = std::map<std::tuple<int, int, int>, std::tuple<int, int, int>>;

static void copy_sources(const HostPlans& hp, int mx, int my, CellMap& out)
{
    for (int L = 0; L < MAXL; L++)
        for (const GhostOp& o : hp.copy[L]) {
            auto rec = [&](int i, int j) {
                auto key = std::make_tuple(o.dst, i, j);
                REQUIRE(!out.count(key), "ghost (%d,%d,%d) copied twice", o.dst, i, j);
                out[key] = std::make_tuple(o.src, i + o.ox, j + o.oy);
            };
            if (o.kind == G_COPY_FACE) {
                const int len = o.sub < 2 ? my : mx;
                for (int layer = 0; layer < MBC; layer++)
                    for (int a = MBC; a < MBC + len; a++) {
                        switch (o.sub) {
                        case 0: rec(layer, a); break;
                        case 1: rec(MBC + mx + layer, a); break;
                        case 2: rec(a, layer); break;
                        default: rec(a, MBC + my + layer); break;
                        }
                    }
            } else if (o.kind == G_COPY_CORNER) {
                for (int jb = 1; jb <= MBC; jb++)
                    for (int ib = 1; ib <= MBC; ib++) {
                        const int i = (o.sub & 1) ? MBC + mx - 1 + ib : MBC - ib;
                        const int j = (o.sub & 2) ? MBC + my - 1 + jb : MBC - jb;
                        rec(i, j);
                    }
            }
        }
}

static void push_sources(const std::vector<int>& tab, int n, int mx, int my, CellMap& out)
{
    const int NX = mx + 2 * MBC, NY = my + 2 * MBC;
    auto put = [&](int nb, int cc, int rr, int src, int c, int r) {
        REQUIRE(nb >= 0 && nb < n, "push target %d out of range", nb);
        REQUIRE(cc >= 0 && cc < NX && rr >= 0 && rr < NY, "push cell (%d,%d) outside the tile", cc, rr);
        auto key = std::make_tuple(nb, cc, rr);
        REQUIRE(!out.count(key), "ghost (%d,%d,%d) pushed twice", nb, cc, rr);
        out[key] = std::make_tuple(src, c, r);
    };
    for (int s = 0; s < n; s++) {
        const int* tb = &tab[(size_t)s * 8];
        for (int c = 2; c <= NX - 3; c++) {
            const bool lx = c <= 3, hx = c >= NX - 4;
            const int nbx = lx ? tb[0] : (hx ? tb[1] : -1);
            const int nbcl = lx ? tb[4] : (hx ? tb[5] : -1);
            const int nbch = lx ? tb[6] : (hx ? tb[7] : -1);
            const int xo = lx ? mx : -mx;
            if (nbx >= 0)
                for (int rr = 2; rr <= NY - 3; rr++) put(nbx, c + xo, rr, s, c, rr);
            if (nbcl >= 0)
                for (int rr = 2; rr <= 3; rr++) put(nbcl, c + xo, rr + my, s, c, rr);
            if (nbch >= 0)
                for (int rr = NY - 4; rr <= NY - 3; rr++) put(nbch, c + xo, rr - my, s, c, rr);
            if (tb[2] >= 0)
                for (int rr = 2; rr <= 3; rr++) put(tb[2], c, rr + my, s, c, rr);
            if (tb[3] >= 0)
                for (int rr = NY - 4; rr <= NY - 3; rr++) put(tb[3], c, rr - my, s, c, rr);
        }
    }
}

static void check_push(const HostPlans& hp, int n, int mx, int my, const char* name)
{
    std::vector<int> tab;
    build_push_table(hp, n, tab);
    CellMap a, b;
    copy_sources(hp, mx, my, a);
    push_sources(tab, n, mx, my, b);
    REQUIRE(a == b, "%s: push writes differ from the same-level copies (%zu vs %zu cells)", name, b.size(),
            a.size());
}

static void check_forest(const Forest& F, int mx, int my, const char* name)
{
    const int n = F.nleaves();
    // 2:1 balance over faces and corners
    for (int k = 0; k < n; k++) {
        int nb[2];
        for (int f = 0; f < 4; f++) (void)F.face_neighbors(k, f, nb);  // throws if unbalanced
    }
    HostPlans hp;
    const int mthbc[4] = {1, 1, 1, 1};
    build_host_plans(F, mx, my, mthbc, hp);
    check_push(hp, n, mx, my, name);
    const int NX = mx + 2 * MBC, NY = my + 2 * MBC;
    std::vector<std::vector<char>> written(n, std::vector<char>(NX * NY, 0));
    long long nops = 0;
    for (int L = 0; L < MAXL; L++) {
        for (auto& o : hp.copy[L]) replay(o, n, mx, my, written), nops++;
        for (auto& o : hp.avg[L]) replay(o, n, mx, my, written), nops++;
        for (auto& o : hp.interp[L]) replay(o, n, mx, my, written), nops++;
        for (auto& b : hp.bc[L]) {
            REQUIRE(b.slot >= 0 && b.slot < n, "bc slot");
            // bc2 overwrites the whole strip of every physical face
            for (int f = 0; f < 4; f++) {
                if (b.mthbc[f] < 0) continue;
                for (int j = 0; j < NY; j++)
                    for (int i = 0; i < NX; i++) {
                        bool in = (f == 0 && i < MBC) || (f == 1 && i >= MBC + mx) ||
                                  (f == 2 && j < MBC) || (f == 3 && j >= MBC + my);
                        if (in) written[b.slot][j * NX + i] = 1;
                    }
            }
        }
    }
    long long missing = 0;
    for (int k = 0; k < n; k++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++) {
                const bool ghost = i < MBC || i >= MBC + mx || j < MBC || j >= MBC + my;
                if (ghost && !written[k][j * NX + i]) missing++;
            }
    std::printf("%s: %d leaves, levels %d..%d, %lld ops, unwritten ghost cells %lld\n", name, n,
                F.global_minlevel(), F.global_maxlevel(), nops, missing);
    REQUIRE(missing == 0, "%s: %lld ghost cells are never written", name, missing);
}

// Forest::set_leaves accepts every adapted forest (in any order, giving the
// same numbering) and rejects gaps, duplicates, nesting and imbalance.
static void check_set_leaves(const Forest& F)
{
    std::vector<Leaf> in(F.leaves.rbegin(), F.leaves.rend());
    Forest G(F.mi, F.mj, F.minlevel, F.maxlevel);
    std::string err = G.set_leaves(in);
    REQUIRE(err.empty(), "set_leaves rejected an adapted forest: %s", err.c_str());
    bool same = G.nleaves() == F.nleaves();
    for (int k = 0; same && k < F.nleaves(); k++)
        same = G.leaves[k].level == F.leaves[k].level && G.leaves[k].gx == F.leaves[k].gx &&
               G.leaves[k].gy == F.leaves[k].gy;
    REQUIRE(same, "set_leaves numbering differs from the adapted forest");
    std::vector<Leaf> gap(F.leaves.begin() + 1, F.leaves.end());
    REQUIRE(!G.set_leaves(gap).empty(), "gap accepted");
    REQUIRE(G.nleaves() == F.nleaves(), "a rejected list changed the forest");
    std::vector<Leaf> dup(F.leaves);
    dup.back() = dup.front();
    REQUIRE(!G.set_leaves(dup).empty(), "duplicate accepted");
    // split a leaf into its four children but keep the parent: nested (area off too)
    const Leaf p = F.leaves.front();
    if (p.level < F.maxlevel) {
        std::vector<Leaf> nest(F.leaves);
        for (int c = 0; c < 4; c++) nest.push_back({p.level + 1, 2 * p.gx + (c & 1), 2 * p.gy + (c >> 1)});
        REQUIRE(!G.set_leaves(nest).empty(), "nested leaves accepted");
    }
}

int main()
{
    {
        // imbalance: a level-1 leaf next to level-3 leaves
        Forest F(1, 1, 1, 3);
        std::vector<Leaf> in = {{1, 1, 0}, {1, 0, 1}, {1, 1, 1}};
        for (int y = 0; y < 2; y++)
            for (int x = 0; x < 2; x++)
                if (x == 1 && y == 1) {
                    for (int c = 0; c < 4; c++) in.push_back({3, 2 + (c & 1), 2 + (c >> 1)});
                } else {
                    in.push_back({2, x, y});
                }
        std::string err = F.set_leaves(in);
        REQUIRE(err.find("balanced") != std::string::npos, "imbalanced forest accepted (%s)", err.c_str());
        in.resize(3);
        for (int y = 0; y < 2; y++)
            for (int x = 0; x < 2; x++) in.push_back({2, x, y});
        REQUIRE(F.set_leaves(in).empty(), "balanced forest rejected");
    }
    const int mx = 8, my = 8;
    {
        Forest F(1, 1, 3, 3);
        F.init_uniform(3);
        check_forest(F, mx, my, "uniform L3");
    }
    {
        Forest F(4, 1, 1, 4);
        F.init_uniform(1);
        check_forest(F, mx, my, "brick 4x1 L1");
    }
    std::mt19937 rng(12345);
    for (int trial = 0; trial < 12; trial++) {
        const int mi = trial % 3 == 2 ? 4 : 1, mj = 1;
        Forest F(mi, mj, 1, 5);
        F.init_uniform(1);
        for (int round = 0; round < 6; round++) {
            std::vector<int> tags(F.nleaves());
            for (auto& t : tags) {
                int r = rng() % 10;
                t = r < 3 ? 1 : (r < 6 ? -1 : 0);
            }
            const int nold = F.nleaves();
            std::vector<Transfer> xf;
            F.adapt(tags, xf);
            REQUIRE((int)xf.size() == F.nleaves(), "transfer count");
            check_xfer(xf, nold, F.nleaves(), mx, my);
            char name[64];
            std::snprintf(name, sizeof(name), "random trial %d round %d", trial, round);
            check_forest(F, mx, my, name);
            check_set_leaves(F);
        }
    }
    std::printf("checked %lld conditions, %d failures\n", g_checked, g_err);
    return g_err ? 1 : 0;
}
