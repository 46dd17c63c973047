// hc_plan.cpp -- ghost-fill plan construction (see hc_plan.hpp).
//
// For every leaf P at level L:
//   face with a same-size neighbour N  -> copy op into P (src = dst + offset)
//   face with two half-size neighbours -> average op into P from each, and an
//                                         interpolation op from P into each
//   face on the physical boundary       -> mthbc entry of P's bc op
//   corner with a same-size neighbour  -> copy op; with a half-size one
//                                         -> average + interpolation ops
// Double-size neighbours get nothing here: the coarse side owns the pair,
// exactly as cb_face_fill only acts on FINER_GRID / SAMESIZE_GRID from the
// coarse patch's point of view.  Offsets follow the same-block transforms of
// src/forestclaw2d.c:655-663 (same size) and 798-841 (half size).
#include "hc_plan.hpp"

namespace hc {

void build_host_plans(const Forest& forest, int mx, int my, const int mthbc[4], HostPlans& hp)
{
    const int nn = forest.nleaves();
    for (int L = 0; L < MAXL; L++) {
        hp.avg[L].clear(); hp.copy[L].clear(); hp.interp[L].clear();
        hp.bc[L].clear(); hp.ti[L].clear();
        hp.ts[L].clear(); hp.ts_same[L].clear(); hp.ts_f2c[L].clear(); hp.ts_phys[L].clear();
    }
    hp.has_fine.assign(nn, 0);
    hp.coarse_faces.assign(nn, 0);
    for (int k = 0; k < nn; k++) {
        const Leaf& P = forest.leaves[k];
        const int L = P.level;
        BcOp b{};
        b.slot = k;
        b.level = L;
        bool phys_any = false;
        TsOp ts{};
        ts.slot = k;
        bool ts_any = false;
        for (int f = 0; f < 4; f++) {
            int nb[2];
            const int kind = forest.face_neighbors(k, f, nb);
            ts.kind[f] = TS_NONE;
            ts.nbr[f][0] = ts.nbr[f][1] = -1;
            if (kind == NB_SAME) {
                ts.kind[f] = TS_SAME;
                ts.nbr[f][0] = nb[0];
                hp.ts_same[L].push_back(4 * k + f);
                ts_any = true;
            } else if (kind == NB_FINE) {
                ts.kind[f] = TS_FINE;
                ts.nbr[f][0] = nb[0];
                ts.nbr[f][1] = nb[1];
                hp.ts_f2c[L].push_back(4 * k + f);
                ts_any = true;
            } else if (kind == NB_COARSE) {
                hp.coarse_faces[k] |= 1 << f;
                if (L >= 1) hp.ts_f2c[L - 1].push_back(4 * k + f);
            } else if (kind == NB_BOUNDARY) {
                hp.ts_phys[L].push_back(4 * k + f);
            }
            b.mthbc[f] = -1;
            if (kind == NB_BOUNDARY) {
                b.mthbc[f] = mthbc[f];
                phys_any = true;
            } else if (kind == NB_SAME) {
                const Leaf& N = forest.leaves[nb[0]];
                hp.copy[L].push_back({G_COPY_FACE, f, k, nb[0], L, L, (P.gx - N.gx) * mx,
                                      (P.gy - N.gy) * my, 0});
            } else if (kind == NB_FINE) {
                hp.has_fine[k] = 1;
                for (int g = 0; g < 2; g++) {
                    const Leaf& Fl = forest.leaves[nb[g]];
                    const int ox = (2 * P.gx - Fl.gx) * mx, oy = (2 * P.gy - Fl.gy) * my;
                    hp.avg[L].push_back({G_AVG_FACE, f, k, nb[g], L, L + 1, ox, oy, 0});
                    hp.interp[L].push_back({G_INTERP_FACE, f, nb[g], k, L + 1, L, ox, oy, 0});
                }
            }
        }
        for (int c = 0; c < 4; c++) {
            int nb;
            const int kind = forest.corner_neighbor(k, c, &nb);
            if (kind == NB_SAME) {
                const Leaf& N = forest.leaves[nb];
                hp.copy[L].push_back({G_COPY_CORNER, c, k, nb, L, L, (P.gx - N.gx) * mx,
                                      (P.gy - N.gy) * my, 0});
            } else if (kind == NB_FINE) {
                hp.has_fine[k] = 1;
                const Leaf& Fl = forest.leaves[nb];
                const int ox = (2 * P.gx - Fl.gx) * mx, oy = (2 * P.gy - Fl.gy) * my;
                hp.avg[L].push_back({G_AVG_CORNER, c, k, nb, L, L + 1, ox, oy, 0});
                hp.interp[L].push_back({G_INTERP_CORNER, c, nb, k, L + 1, L, ox, oy, 0});
            }
        }
        if (phys_any) hp.bc[L].push_back(b);
        if (ts_any) hp.ts[L].push_back(ts);
        if (hp.has_fine[k]) hp.ti[L].push_back({k, L});
    }
}

void build_push_table(const HostPlans& hp, int nslots, std::vector<int>& push)
{
    static const int opp_face[4] = {1, 0, 3, 2};
    push.assign((size_t)nslots * 8, -1);
    for (int L = 0; L < MAXL; L++)
        for (const GhostOp& o : hp.copy[L]) {
            if (o.kind == G_COPY_FACE) push[(size_t)o.src * 8 + opp_face[o.sub]] = o.dst;
            else if (o.kind == G_COPY_CORNER) push[(size_t)o.src * 8 + 4 + (3 - o.sub)] = o.dst;
        }
}

}  // namespace hc
