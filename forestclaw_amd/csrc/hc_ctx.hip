// hc_ctx.hip -- the device-resident AMR driver behind the hipclaw_ctx_* C ABI.
//
// Host-side restatement of the ForestClaw drivers that sit around the patch
// step, organised for the GPU: every phase of a ghost update, every level
// update and every regrid transfer is one batched launch over a precomputed
// device list instead of one callback per patch or patch pair.
//   ghost update   src/fclaw2d_ghost_fill.c:611-856 (phase order), with the
//                  per-pair dispatch of src/fclaw2d_face_neighbors.c:195-450 and
//                  src/fclaw2d_corner_neighbors.c:334-535 turned into plans
//   level update   src/fclaw2d_update_single_step.c:44-103
//   subcycling     src/fclaw2d_advance.c:54-357
//   time interp    src/fclaw2d_timeinterp.c:30-57
//   regrid         src/fclaw2d_regrid.c:45-335 (tagging, adapt, repopulate)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "hc_ctx.hpp"
#include "hc_plan.hpp"
#include "hc_internal.hpp"

namespace hc {

template <class T> int DevArray<T>::upload(const std::vector<T>& v)
{
    free_();
    n = (int)v.size();
    if (n == 0) return HIPCLAW_OK;
    if (hipMalloc(&p, sizeof(T) * n) != hipSuccess) {
        p = nullptr;
        n = 0;
        return set_error(HIPCLAW_ERR_NOMEM, "hipMalloc failed for a plan");
    }
    return check_hip(hipMemcpy(p, v.data(), sizeof(T) * n, hipMemcpyHostToDevice), "plan upload");
}

Ctx::Ctx(const hipclaw_physics& p, const hipclaw_domain& d)
    : phys(p), dom(d), forest(d.mi, d.mj, d.minlevel, d.maxlevel)
{
    esz = p.precision == HIPCLAW_FP64 ? 8 : 4;
    const long long T = (long long)(p.mx + 4) * (p.my + 4);
    tile = T * p.meqn;
    atile = T * p.maux;
}

Ctx::~Ctx()
{
    clear_graphs();
    free_par();
    prof_clear();
    release_pools();
    if (d_tiles) (void)hipFree(d_tiles);
    if (d_logits) (void)hipFree(d_logits);
}

void Ctx::clear_graphs()
{
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
}

int Ctx::ensure_par()
{
    if (d_par) return HIPCLAW_OK;
    const size_t n = (size_t)PAR_MAXLAUNCH * PAR_STRIDE;
    if (hipMalloc(reinterpret_cast<void**>(&d_par), sizeof(double) * n) != hipSuccess) {
        d_par = nullptr;
        return set_error(HIPCLAW_ERR_NOMEM, "hipMalloc of the step parameter block failed");
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&h_par), sizeof(double) * n * PAR_RING, hipHostMallocDefault) !=
        hipSuccess) {
        h_par = nullptr;
        free_par();
        return set_error(HIPCLAW_ERR_NOMEM, "hipHostMalloc of the step parameter ring failed");
    }
    for (int k = 0; k < PAR_RING; k++) {
        if (hipEventCreateWithFlags(&par_ev[k], hipEventDisableTiming) != hipSuccess) {
            par_ev[k] = nullptr;
            free_par();
            return set_error(HIPCLAW_ERR_HIP, "event creation failed");
        }
        par_ev_used[k] = false;
    }
    return HIPCLAW_OK;
}

void Ctx::free_par()
{
    for (int k = 0; k < PAR_RING; k++) {
        if (par_ev[k]) {
            if (par_ev_used[k]) (void)hipEventSynchronize(par_ev[k]);
            (void)hipEventDestroy(par_ev[k]);
        }
        par_ev[k] = nullptr;
        par_ev_used[k] = false;
    }
    if (h_par) (void)hipHostFree(h_par);
    if (d_par) (void)hipFree(d_par);
    h_par = nullptr;
    d_par = nullptr;
}

void Ctx::prof_clear()
{
    for (auto& e : ev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    ev.clear();
    ev_cells.clear();
}

void Ctx::release_pools()
{
    for (auto& b : pool) {
        if (b) (void)hipFree(b);
        b = nullptr;
    }
    void** bufs[] = {&ti, &save, &aux};
    for (auto b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    if (d_cflslot) (void)hipFree(d_cflslot);
    d_cflslot = nullptr;
    double** dbufs[] = {&d_xylower, &d_dxdy, &d_cfl, &d_cflacc, &d_mass, &d_reg, &d_tagst};
    for (auto b : dbufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    int** ibufs[] = {&d_levels, &d_iota, &d_tags, &d_push, &d_cfaces};
    for (auto b : ibufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    for (auto& lp : plan) {
        lp.avg.free_(); lp.avg_ti.free_(); lp.copy.free_(); lp.copy_ti.free_();
        lp.interp.free_(); lp.interp_ti.free_(); lp.bc.free_(); lp.bc_ti.free_(); lp.ti.free_();
        lp.ts.free_(); lp.ts_same.free_(); lp.ts_f2c.free_(); lp.ts_phys.free_();
    }
    n = cap = 0;
    tagst_cap = 0;
}

template <class R> Pools<R> Ctx::pools() const
{
    Pools<R> P{};
    for (int L = 0; L < MAXL; L++) {
        P.q[L] = static_cast<R*>(pool[parity[L]]);
        P.qlast[L] = static_cast<R*>(pool[parity[L] ^ 1]);
    }
    P.ti = static_cast<R*>(ti);
    return P;
}

int Ctx::level_range(int level, int* s0, int* s1) const
{
    if (level < 0 || level >= MAXL) return set_error(HIPCLAW_ERR_ARG, "level out of range");
    *s0 = plan[level].s0;
    *s1 = plan[level].s1;
    return HIPCLAW_OK;
}

#define HC_TRY(x)                   \
    do {                            \
        int rc_ = (x);              \
        if (rc_) return rc_;        \
    } while (0)

static int dmalloc(void** p, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return set_error(HIPCLAW_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    // zero on the null stream and wait: the context's stream may be a
    // non-blocking one, which the null stream does not order against
    HC_TRY(check_hip(hipMemsetAsync(*p, 0, bytes, nullptr), "hipMemset"));
    return check_hip(hipStreamSynchronize(nullptr), "hipMemset");
}

// Build pools, geometry and ghost plans for the current forest.  With a
// transfer list the previous pools are repopulated into the new ones.
int Ctx::rebuild(bool keep_data, const std::vector<Transfer>* xfer)
{
    HC_TRY(check_hip(hipStreamSynchronize(stream), "rebuild"));
    const int nn = forest.nleaves();
    const int mx = phys.mx, my = phys.my;
    void* npool[2] = {nullptr, nullptr};
    void* nti = nullptr;
    void* naux = nullptr;
    HC_TRY(dmalloc(&npool[0], (size_t)nn * tile * esz));
    HC_TRY(dmalloc(&npool[1], (size_t)nn * tile * esz));
    HC_TRY(dmalloc(&nti, (size_t)nn * tile * esz));
    if (phys.maux > 0) HC_TRY(dmalloc(&naux, (size_t)nn * atile * esz));

    if (keep_data && xfer && pool[0]) {
        std::vector<XferOp> ops;
        ops.reserve(xfer->size());
        std::vector<XferOp> by_level[MAXL];
        for (const Transfer& t : *xfer) {
            XferOp o{};
            o.dst = t.new_leaf;
            o.kind = t.kind == Transfer::KEEP ? 0 : (t.kind == Transfer::REFINE ? 1 : 2);
            for (int k = 0; k < 4; k++) o.src[k] = t.old_leaf[k];
            o.srclev = level_of[t.old_leaf[0]];
            o.igrid = t.igrid;
            ops.push_back(o);
        }
        DevArray<XferOp> dops;
        HC_TRY(dops.upload(ops));
        if (esz == 8)
            launch_xfer<double>(dops.p, dops.n, pools<double>(), static_cast<double*>(npool[0]), mx, my,
                                phys.meqn, stream);
        else
            launch_xfer<float>(dops.p, dops.n, pools<float>(), static_cast<float*>(npool[0]), mx, my,
                               phys.meqn, stream);
        HC_TRY(check_hip(hipStreamSynchronize(stream), "regrid transfer"));
        dops.free_();
    }
    // drop the previous pools and plans
    {
        const long long cu = cell_updates;
        const double tp = swirl_tperiod;
        const double rt = refine_threshold, ct = coarsen_threshold;
        const int iv = tag_ivar;
        release_pools();
        cell_updates = cu;
        swirl_tperiod = tp;
        refine_threshold = rt;
        coarsen_threshold = ct;
        tag_ivar = iv;
    }
    pool[0] = npool[0];
    pool[1] = npool[1];
    ti = nti;
    aux = naux;
    n = cap = nn;
    for (int L = 0; L < MAXL; L++) parity[L] = 0;

    // geometry
    const double wx = (dom.bx - dom.ax) / dom.mi, wy = (dom.by - dom.ay) / dom.mj;
    std::vector<double> xy(2 * (size_t)nn), dd(2 * (size_t)nn);
    std::vector<int> lev(nn), iota(nn);
    level_of.assign(nn, 0);
    for (int k = 0; k < nn; k++) {
        const Leaf& l = forest.leaves[k];
        const double pw = wx / (double)(1 << l.level), ph = wy / (double)(1 << l.level);
        xy[2 * k] = dom.ax + l.gx * pw;
        xy[2 * k + 1] = dom.ay + l.gy * ph;
        dd[2 * k] = pw / mx;
        dd[2 * k + 1] = ph / my;
        lev[k] = l.level;
        level_of[k] = l.level;
        iota[k] = k;
    }
    HC_TRY(dmalloc((void**)&d_xylower, sizeof(double) * 2 * nn));
    HC_TRY(dmalloc((void**)&d_dxdy, sizeof(double) * 2 * nn));
    HC_TRY(dmalloc((void**)&d_cfl, sizeof(double) * nn));
    HC_TRY(dmalloc((void**)&d_cflacc, sizeof(double) * 4096));
    HC_TRY(dmalloc((void**)&d_cflslot, sizeof(unsigned long long) * 4096));
    HC_TRY(check_hip(hipMemset(d_cflslot, 0, sizeof(unsigned long long) * 4096), "cfl slots"));
    cflslot_dirty = false;
    HC_TRY(dmalloc((void**)&d_levels, sizeof(int) * nn));
    HC_TRY(dmalloc((void**)&d_iota, sizeof(int) * nn));
    HC_TRY(dmalloc((void**)&d_tags, sizeof(int) * nn));
    HC_TRY(dmalloc((void**)&d_mass, sizeof(double) * nn * phys.meqn));
    HC_TRY(check_hip(hipMemcpy(d_xylower, xy.data(), sizeof(double) * 2 * nn, hipMemcpyHostToDevice), "geom"));
    HC_TRY(check_hip(hipMemcpy(d_dxdy, dd.data(), sizeof(double) * 2 * nn, hipMemcpyHostToDevice), "geom"));
    HC_TRY(check_hip(hipMemcpy(d_levels, lev.data(), sizeof(int) * nn, hipMemcpyHostToDevice), "levels"));
    HC_TRY(check_hip(hipMemcpy(d_iota, iota.data(), sizeof(int) * nn, hipMemcpyHostToDevice), "iota"));

    // level ranges (leaves are sorted by level)
    for (int L = 0; L < MAXL; L++) plan[L].s0 = plan[L].s1 = 0;
    for (int k = 0; k < nn;) {
        int L = lev[k], e = k;
        while (e < nn && lev[e] == L) e++;
        plan[L].s0 = k;
        plan[L].s1 = e;
        k = e;
    }

    // neighbour relations -> ghost plans
    HostPlans hp;
    build_host_plans(forest, mx, my, phys.mthbc, hp);
    has_fine = hp.has_fine;
    auto& avg = hp.avg;
    auto& copy = hp.copy;
    auto& interp = hp.interp;
    auto& bc = hp.bc;
    auto& tio = hp.ti;
    for (int L = 0; L < MAXL; L++) {
        LevelPlan& lp = plan[L];
        auto with = [](std::vector<GhostOp> v, int flags) {
            for (auto& o : v) o.flags = flags;
            return v;
        };
        HC_TRY(lp.avg.upload(avg[L]));
        HC_TRY(lp.avg_ti.upload(with(avg[L], 1)));
        HC_TRY(lp.copy.upload(copy[L]));
        HC_TRY(lp.copy_ti.upload(with(copy[L], 3)));
        HC_TRY(lp.interp.upload(interp[L]));
        HC_TRY(lp.interp_ti.upload(with(interp[L], 2)));
        HC_TRY(lp.bc.upload(bc[L]));
        std::vector<BcOp> bti = bc[L];
        for (auto& o : bti) o.flags = 1;
        HC_TRY(lp.bc_ti.upload(bti));
        HC_TRY(lp.ti.upload(tio[L]));
        HC_TRY(lp.ts.upload(hp.ts[L]));
        HC_TRY(lp.ts_same.upload(hp.ts_same[L]));
        HC_TRY(lp.ts_f2c.upload(hp.ts_f2c[L]));
        HC_TRY(lp.ts_phys.upload(hp.ts_phys[L]));
        // the physical BCs and the time-sync corrections of a patch as one
        // operation (a fill's last phase fused with the sync that follows)
        std::vector<BcTsOp> bt, btti;
        std::vector<int> at((size_t)nn, -1);
        auto get = [&](int slot) -> BcTsOp& {
            if (at[slot] < 0) {
                BcTsOp o{};
                o.bc.slot = -1;
                o.ts.slot = -1;
                o.level = L;
                at[slot] = (int)bt.size();
                bt.push_back(o);
            }
            return bt[at[slot]];
        };
        for (const BcOp& b : bc[L]) get(b.slot).bc = b;
        for (const TsOp& o : hp.ts[L]) get(o.slot).ts = o;
        for (const BcOp& b : bti) {
            BcTsOp o{};
            o.bc = b;
            o.ts.slot = -1;
            o.level = L;
            btti.push_back(o);
        }
        HC_TRY(lp.bcts.upload(bt));
        HC_TRY(lp.bcts_ti.upload(btti));
    }
    // same-level push table: the copy op (dst k, face / corner c, src nb)
    // becomes "nb writes into k across its opposite face / corner"
    {
        std::vector<int> push;
        build_push_table(hp, nn, push);
        HC_TRY(dmalloc((void**)&d_push, sizeof(int) * push.size()));
        HC_TRY(check_hip(hipMemcpy(d_push, push.data(), sizeof(int) * push.size(), hipMemcpyHostToDevice),
                         "push table"));
    }
    clear_pushed();
    clear_graphs();
    mesh_epoch++;
    // registers start at zero on a new mesh (a regrid happens after a full
    // coarse step, when every synchronised register has been reset)
    regstride = 8LL * phys.meqn * (phys.mx + phys.my);
    if (tsync) HC_TRY(dmalloc((void**)&d_reg, sizeof(double) * (size_t)nn * regstride));
    reg_clean = true;
    reg_stale = false;
    HC_TRY(dmalloc((void**)&d_cfaces, sizeof(int) * (size_t)nn));
    HC_TRY(check_hip(hipMemcpy(d_cfaces, hp.coarse_faces.data(), sizeof(int) * (size_t)nn, hipMemcpyHostToDevice),
                     "coarse faces"));
    return check_hip(hipDeviceSynchronize(), "rebuild");
}

// ---------------------------------------------------------------------------
// Ghost update, phase order of fclaw2d_ghost_update_async for one rank:
// average (maxcoarse..mincoarse [+ time-interpolated level]) -> copy
// (maxlevel..minlevel [+ti]) -> physical BC on coarse levels [+ti] ->
// interpolate (maxcoarse..mincoarse [+ti]) -> physical BC on all levels [+ti].
// The physical BC pass on the fine levels that precedes the final pass in
// the reference is subsumed by it (bc2 is idempotent on unchanged data).
// ---------------------------------------------------------------------------
int Ctx::ghost_update(int minlevel, int maxlevel, int time_interp, bool sync, double* cfl_out)
{
    if (par_dry) {
        ti_pending.clear();
        return HIPCLAW_OK;
    }
    if (sync && !(merge_levels && tsync)) {
        // a sync after the fill reads q: the reduction (slots only) may
        // still ride on the fill's last launch
        HC_TRY(ghost_update(minlevel, maxlevel, time_interp, false, cfl_out));
        return time_sync(minlevel, maxlevel);
    }
    const int mx = phys.mx, my = phys.my, meqn = phys.meqn;
    const int mincoarse = minlevel, maxcoarse = maxlevel - 1;
    const int tl = minlevel - 1;
    const bool use_ti = time_interp && tl >= 0;
    auto run = [&](auto tag) -> int {
        using R = decltype(tag);
        const Pools<R> P = pools<R>();
        if (merge_levels) {
            // one launch per phase: within the average, copy and BC phases
            // every operation writes ghost cells of its own patch from
            // interior cells only, so the levels of a phase are independent;
            // interpolation reads coarse ghost cells that the interpolation
            // of the next coarser level writes and keeps its level order.
            // With the time sync following (sync), the final BC phase and the
            // corrections run per patch in one launch: a patch's BCs and
            // corrections touch only that patch's cells, and its
            // corrections read registers only.
            Segs<GhostOp> avg, copy;
            Segs<BcOp> bcc, bca;
            for (int L = maxcoarse; L >= mincoarse; L--) avg.add(plan[L].avg.p, plan[L].avg.n);
            if (use_ti) avg.add(plan[tl].avg_ti.p, plan[tl].avg_ti.n);
            for (int L = maxlevel; L >= minlevel; L--)
                if (!pushed[L]) copy.add(plan[L].copy.p, plan[L].copy.n);
            if (use_ti) copy.add(plan[tl].copy_ti.p, plan[tl].copy_ti.n);
            for (int L = maxcoarse; L >= mincoarse; L--) bcc.add(plan[L].bc.p, plan[L].bc.n);
            if (use_ti) bcc.add(plan[tl].bc_ti.p, plan[tl].bc_ti.n);
            for (int L = maxlevel; L >= minlevel; L--) bca.add(plan[L].bc.p, plan[L].bc.n);
            if (use_ti) bca.add(plan[tl].bc_ti.p, plan[tl].bc_ti.n);
            // the pending time interpolations write the interior ring of the
            // interpolated pool, which no average or copy writes
            Segs<TiOp> tis;
            double alpha[SEG_MAX];
            int tilev[SEG_MAX];
            for (const auto& ti : ti_pending) {
                if (plan[ti.first].ti.n <= 0) continue;
                if (tis.n < SEG_MAX) {
                    alpha[tis.n] = ti.second;
                    tilev[tis.n] = ti.first;
                }
                tis.add(plan[ti.first].ti.p, plan[ti.first].ti.n);
            }
            ti_pending.clear();
            Segs<BcTsOp> bct;
            ResetSegs rs;
            if (sync) {
                for (int L = maxlevel; L >= minlevel; L--) bct.add(plan[L].bcts.p, plan[L].bcts.n);
                if (use_ti) bct.add(plan[tl].bcts_ti.p, plan[tl].bcts_ti.n);
                if (!bct.ok || !reset_segs(minlevel, maxlevel, rs))
                    return set_error(HIPCLAW_ERR_UNSUPPORTED, "ghost update: too many operation lists");
            }
            if (!avg.ok || !copy.ok || !bcc.ok || !bca.ok || !tis.ok)
                return set_error(HIPCLAW_ERR_UNSUPPORTED, "ghost update: too many operation lists");
            // averages and copies share a launch when their lists fit one:
            // each face and corner of a patch has one neighbour, so the two
            // phases write disjoint ghost cells, and both read interior cells
            // only.  A copy from the interpolated pool of a level whose time
            // interpolation runs in that launch interpolates its source cell
            // itself.
            Segs<GhostOp> first = avg;
            for (int s = 0; s < copy.n; s++) first.add(copy.ops[s], copy.start[s + 1] - copy.start[s]);
            if (first.ok) {
                launch_ghost_ti<R>(first, tis, alpha, tilev, P, mx, my, meqn, stream);
            } else {
                launch_ghost_ti<R>(avg, tis, alpha, tilev, P, mx, my, meqn, stream);
                launch_ghost<R>(copy, P, mx, my, meqn, stream);
            }
            launch_bc<R>(bcc, P, mx, my, meqn, stream);
            for (int L = maxcoarse; L >= mincoarse; L--)
                launch_ghost<R>(plan[L].interp.p, plan[L].interp.n, P, mx, my, meqn, stream);
            if (use_ti) launch_ghost<R>(plan[tl].interp_ti.p, plan[tl].interp_ti.n, P, mx, my, meqn, stream);
            const CflFin cf{cfl_out ? d_cflslot : nullptr, cflacc_count < 4095 ? cflacc_count : 4095, cfl_out};
            if (cfl_out) cfl_done = true;
            if (sync) {
                launch_bc_ts<R>(bct, P, d_reg, regstride, d_dxdy, mx, my, meqn, cf, stream);
                clear_pushed();
                // a merged coarse step's updates write their first sums onto
                // zero instead (StepArgs::reg_fresh)
                if (!reg_fresh_run) launch_ts_reset(rs, d_reg, regstride, mx, my, meqn, stream);
                return check_hip(hipGetLastError(), "ghost update and time sync");
            }
            launch_bc_cfl<R>(bca, P, mx, my, meqn, cf, stream);
            return check_hip(hipGetLastError(), "ghost update");
        }
        for (int L = maxcoarse; L >= mincoarse; L--)
            launch_ghost<R>(plan[L].avg.p, plan[L].avg.n, P, mx, my, meqn, stream);
        if (use_ti) launch_ghost<R>(plan[tl].avg_ti.p, plan[tl].avg_ti.n, P, mx, my, meqn, stream);
        for (int L = maxlevel; L >= minlevel; L--)
            if (!pushed[L]) launch_ghost<R>(plan[L].copy.p, plan[L].copy.n, P, mx, my, meqn, stream);
        if (use_ti) launch_ghost<R>(plan[tl].copy_ti.p, plan[tl].copy_ti.n, P, mx, my, meqn, stream);
        for (int L = maxcoarse; L >= mincoarse; L--)
            launch_bc<R>(plan[L].bc.p, plan[L].bc.n, P, mx, my, meqn, stream);
        if (use_ti) launch_bc<R>(plan[tl].bc_ti.p, plan[tl].bc_ti.n, P, mx, my, meqn, stream);
        for (int L = maxcoarse; L >= mincoarse; L--)
            launch_ghost<R>(plan[L].interp.p, plan[L].interp.n, P, mx, my, meqn, stream);
        if (use_ti) launch_ghost<R>(plan[tl].interp_ti.p, plan[tl].interp_ti.n, P, mx, my, meqn, stream);
        for (int L = maxlevel; L >= minlevel; L--)
            launch_bc<R>(plan[L].bc.p, plan[L].bc.n, P, mx, my, meqn, stream);
        if (use_ti) launch_bc<R>(plan[tl].bc_ti.p, plan[tl].bc_ti.n, P, mx, my, meqn, stream);
        return check_hip(hipGetLastError(), "ghost update");
    };
    return esz == 8 ? run(double(0)) : run(float(0));
}

Src2Args Ctx::src2_args(int s0, int np) const
{
    Src2Args a{};
    a.n = np;
    a.slot0 = s0;
    a.xylower = d_xylower;
    a.dxdy = d_dxdy;
    a.mx = phys.mx;
    a.my = phys.my;
    a.meqn = phys.meqn;
    a.maux = 0;
    a.gamma1 = phys.gamma - 1.0;
    return a;
}

// Swirl advection: the velocity field b4step2 stores before each step is
// evaluated inside the step kernel (same operations) unless switched off.
bool Ctx::fused_swirl() const
{
    return fuse_aux && swirl_tperiod > 0 && solver_family(phys.solver) == HIPCLAW_SOLVER_ADV;
}

double Ctx::swirl_vt(double t, double dt) const
{
    const double pi2 = 8.0 * atan(1.0);
    return cos(pi2 * (t + dt / 2.0) / swirl_tperiod);
}

int Ctx::update_level(int level, double t, double dt)
{
    const int s0 = plan[level].s0, s1 = plan[level].s1;
    const int np = s1 - s0;
    if (np <= 0) return HIPCLAW_OK;
    char* qin = static_cast<char*>(pool[parity[level]]) + (size_t)s0 * tile * esz;
    char* qout = static_cast<char*>(pool[parity[level] ^ 1]) + (size_t)s0 * tile * esz;
    char* a = aux ? static_cast<char*>(aux) + (size_t)s0 * atile * esz : nullptr;
    const bool fz = fused_swirl();
    if (!fz && swirl_tperiod > 0 && solver_family(phys.solver) == HIPCLAW_SOLVER_ADV)
        HC_TRY(hipclaw_swirl_aux_batch(&phys, np, d_xylower + 2 * s0, d_dxdy + 2 * s0, t, dt,
                                       swirl_tperiod, a, stream));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (prof) {
        HC_TRY(check_hip(hipEventCreate(&e0), "event"));
        HC_TRY(check_hip(hipEventCreate(&e1), "event"));
        HC_TRY(check_hip(hipEventRecord(e0, stream), "event"));
    }
    const hipclaw_launch* lc = launch.cols_per_lane > 0 ? &launch : nullptr;
    if (tsync) {
        HC_TRY(hipclaw_step2_batch_sync(&phys, np, qin, qout, a, d_dxdy + 2 * s0, nullptr, np, dt,
                                        d_cfl + s0, d_reg + (size_t)s0 * regstride, lc, stream));
    } else if (fz) {
        MultiLevel ml{};
        ml.levels = d_levels;
        ml.slot0 = s0;
        ml.dtl[level] = dt;
        ml.qin_l[level] = pool[parity[level]];
        ml.qout_l[level] = pool[parity[level] ^ 1];
        ml.swirl = 1;
        ml.xylower = d_xylower;
        ml.vtl[level] = swirl_vt(t, dt);
        HC_TRY(step2_launch(&phys, np, pool[0], pool[1], aux, d_dxdy, nullptr, np, dt, d_cfl, nullptr, lc,
                            stream, &ml));
    } else
        HC_TRY(hipclaw_step2_batch(&phys, np, qin, qout, a, d_dxdy + 2 * s0, nullptr, np, dt,
                                   d_cfl + s0, lc, stream));
    if (src2) {
        Src2Args sa = src2_args(s0, np);
        sa.q_l[0] = pool[parity[level] ^ 1];
        sa.dtl[0] = dt;
        HC_TRY(src2_launch(sa, src2, phys.precision, nullptr, stream));
    }
    if (prof) {
        HC_TRY(check_hip(hipEventRecord(e1, stream), "event"));
        ev.push_back({e0, e1});
        ev_cells.push_back((long long)np * phys.mx * phys.my);
    }
    if (cflacc_count < 4095) {
        HC_TRY(cfl_slot_reduce(d_cfl + s0, np, d_cflslot + cflacc_count, stream));
        cflacc_count++;
    }
    parity[level] ^= 1;
    pushed[level] = false;
    cell_updates += (long long)np * phys.mx * phys.my;
    return HIPCLAW_OK;
}

// Time every patch-step launch shape on one level's current data (output into
// a temporary buffer, so the solution is untouched) and keep the fastest.
int Ctx::autotune(int level, int reps, hipclaw_launch* best, float* ms)
{
    if (level < 0) {  // the level with the most patches
        int bn = -1;
        for (int L = 0; L < MAXL; L++)
            if (plan[L].s1 - plan[L].s0 > bn) bn = plan[L].s1 - plan[L].s0, level = L;
    }
    if (level >= MAXL) return set_error(HIPCLAW_ERR_ARG, "level out of range");
    const int s0 = plan[level].s0, np = plan[level].s1 - s0;
    if (np <= 0) return set_error(HIPCLAW_ERR_ARG, "level has no patches");
    if (reps < 1) reps = 1;
    const char* qin = static_cast<const char*>(pool[parity[level]]) + (size_t)s0 * tile * esz;
    const char* a = aux ? static_cast<const char*>(aux) + (size_t)s0 * atile * esz : nullptr;
    void* tmp = nullptr;
    double* cfl = nullptr;
    HC_TRY(dmalloc(&tmp, (size_t)np * tile * esz));
    int rc = dmalloc(reinterpret_cast<void**>(&cfl), sizeof(double) * np);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc) rc = check_hip(hipEventCreate(&e0), "event");
    if (!rc) rc = check_hip(hipEventCreate(&e1), "event");
    // a small step (Courant number ~0.1 for unit wave speeds) keeps the data valid
    const double dx = (dom.bx - dom.ax) / dom.mi / (double)(1 << level) / phys.mx;
    const double dt = 0.1 * dx;
    // Time the kernel variant the level updates launch: with the ghost push
    // on, the slot-addressed push kernel, whose output pool is the scratch
    // tiles (same-level neighbours are slots of the same level, so every
    // push lands in the scratch too).
    const bool push = use_push() && merge_levels;
    MultiLevel ml{};
    ml.levels = d_levels;
    ml.slot0 = s0;
    ml.dtl[level] = dt;
    ml.qin_l[level] = pool[parity[level]];
    ml.qout_l[level] = static_cast<char*>(tmp) - (ptrdiff_t)s0 * tile * esz;
    ml.push = d_push;
    if (fused_swirl()) {
        ml.swirl = 1;
        ml.xylower = d_xylower;
        ml.vtl[level] = swirl_vt(0.0, dt);
    }
    auto run = [&](const hipclaw_launch& l) {
        if (push)
            return step2_launch(&phys, np, pool[0], pool[1], aux, d_dxdy, nullptr, np, dt, cfl - s0, nullptr,
                                &l, stream, &ml);
        return hipclaw_step2_batch(&phys, np, qin, tmp, a, d_dxdy + 2 * s0, nullptr, np, dt, cfl, &l, stream);
    };
    float bestms = 1e30f;
    int k = 0;
    for (int cpl = 1; cpl <= 2 && !rc; cpl++)
        for (int wpg : {1, 2, 4}) {
            if (rc) break;
            hipclaw_launch l{cpl, wpg, 0};
            // one untimed launch, then reps timed ones
            rc = run(l);
            if (!rc) rc = check_hip(hipEventRecord(e0, stream), "event");
            for (int r = 0; r < reps && !rc; r++) rc = run(l);
            if (!rc) rc = check_hip(hipEventRecord(e1, stream), "event");
            if (!rc) rc = check_hip(hipEventSynchronize(e1), "event");
            float t = 0;
            if (!rc) rc = check_hip(hipEventElapsedTime(&t, e0, e1), "event");
            t /= reps;
            if (ms) ms[k] = t;
            k++;
            if (!rc && t < bestms) {
                bestms = t;
                // the shape for launches of this level's size; smaller
                // launches keep the library's size-based shape and bands
                launch = hipclaw_launch{cpl, wpg, -1};
            }
        }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(tmp);
    if (cfl) (void)hipFree(cfl);
    clear_graphs();
    if (rc) return rc;
    if (best) *best = launch;
    return HIPCLAW_OK;
}

int Ctx::set_time_sync(bool on)
{
    if (on && solver_family(phys.solver) != HIPCLAW_SOLVER_SWE_ROE)
        return set_error(HIPCLAW_ERR_UNSUPPORTED, "time sync needs a conservative flux (SWE only)");
    tsync = on;
    clear_graphs();
    if (on && !d_reg && n > 0) {
        HC_TRY(dmalloc((void**)&d_reg, sizeof(double) * (size_t)n * regstride));
        reg_stale = false;
    }
    return HIPCLAW_OK;
}

// Registers left by merged coarse steps: cleared before a per-level update,
// a stand-alone time sync, a read-back or a coarse step that reads them.
int Ctx::flush_registers()
{
    if (!reg_stale || !d_reg) return HIPCLAW_OK;
    launch_zero(d_reg, (long long)n * regstride, stream);
    reg_stale = false;
    return check_hip(hipGetLastError(), "register clear");
}

// fclaw2d_time_sync(minlevel, maxlevel): corrections of every synchronised
// level, then the register resets of correct_coarse_cells
// (src/fclaw2d_time_sync.c:118-154): same-size faces of levels
// [minlevel, maxlevel], coarse/fine faces between levels L and L+1 for L in
// [minlevel, maxlevel-1], physical faces of [minlevel, maxlevel].
// register faces a sync of [minlevel, maxlevel] resets
// (fclaw2d_clawpatch_conservation.c:207-280): same-size and physical faces of
// every level, coarse-fine faces between two synchronised levels
bool Ctx::reset_segs(int minlevel, int maxlevel, ResetSegs& rs) const
{
    for (int L = minlevel; L <= maxlevel; L++) {
        rs.add(plan[L].ts_same.p, plan[L].ts_same.n);
        if (L < maxlevel) rs.add(plan[L].ts_f2c.p, plan[L].ts_f2c.n);
        rs.add(plan[L].ts_phys.p, plan[L].ts_phys.n);
    }
    return rs.ok;
}

int Ctx::time_sync(int minlevel, int maxlevel)
{
    if (!tsync || par_dry) return HIPCLAW_OK;
    clear_pushed();
    if (minlevel < 0 || maxlevel >= MAXL || minlevel > maxlevel)
        return set_error(HIPCLAW_ERR_ARG, "time_sync: bad level range");
    const int mx = phys.mx, my = phys.my, meqn = phys.meqn;
    if (merge_levels) {
        // one correction launch over the levels (each operation adds to the
        // interior of its own patch and only reads registers) and one reset
        // launch after it
        Segs<TsOp> cs;
        ResetSegs rs;
        void* qs[SEG_MAX];
        for (int L = minlevel; L <= maxlevel; L++) {
            if (plan[L].ts.n <= 0) continue;
            qs[cs.n] = esz == 8 ? (void*)pools<double>().q[L] : (void*)pools<float>().q[L];
            cs.add(plan[L].ts.p, plan[L].ts.n);
        }
        if (!cs.ok || !reset_segs(minlevel, maxlevel, rs))
            return set_error(HIPCLAW_ERR_UNSUPPORTED, "time sync: too many face lists");
        if (esz == 8)
            launch_ts_correct<double>(cs, (double* const*)qs, d_reg, regstride, d_dxdy, mx, my, meqn, stream);
        else
            launch_ts_correct<float>(cs, (float* const*)qs, d_reg, regstride, d_dxdy, mx, my, meqn, stream);
        launch_ts_reset(rs, d_reg, regstride, mx, my, meqn, stream);
        return check_hip(hipGetLastError(), "time sync");
    }
    for (int L = minlevel; L <= maxlevel; L++) {
        const LevelPlan& lp = plan[L];
        if (esz == 8)
            launch_ts_correct<double>(lp.ts.p, lp.ts.n, pools<double>().q[L], d_reg, regstride, d_dxdy,
                                      mx, my, meqn, stream);
        else
            launch_ts_correct<float>(lp.ts.p, lp.ts.n, pools<float>().q[L], d_reg, regstride, d_dxdy,
                                     mx, my, meqn, stream);
    }
    for (int L = minlevel; L <= maxlevel; L++) {
        launch_ts_reset(plan[L].ts_same.p, plan[L].ts_same.n, d_reg, regstride, mx, my, meqn, stream);
        if (L < maxlevel)
            launch_ts_reset(plan[L].ts_f2c.p, plan[L].ts_f2c.n, d_reg, regstride, mx, my, meqn, stream);
        launch_ts_reset(plan[L].ts_phys.p, plan[L].ts_phys.n, d_reg, regstride, mx, my, meqn, stream);
    }
    return check_hip(hipGetLastError(), "time sync");
}

// After a coarse step that did not finish: clear every Courant-number slot
// (the final reduction of the slots, into an unused element).
int Ctx::clear_cfl_slots()
{
    HC_TRY(cfl_final(d_cflslot, 4095, d_cflacc + 4094, stream));
    cflslot_dirty = false;
    return HIPCLAW_OK;
}

int Ctx::update_levels(int lo, int hi, const double* tl, const double* dtl)
{
    // graph mode: this launch's dt and vt per level into the parameter block
    // (written in the dry pass; the capture that follows only takes the
    // launch's slot address, since the slot's upload may still be in flight)
    const double* par_dev = nullptr;
    if (par_host) {
        if (par_n >= PAR_MAXLAUNCH) return set_error(HIPCLAW_ERR_STATE, "too many step launches for a graph");
        par_dev = d_par + (size_t)par_n * PAR_STRIDE;
        if (par_dry) {
            double* hp = par_host + (size_t)par_n * PAR_STRIDE;
            const bool fz = fused_swirl();
            for (int L = lo; L <= hi; L++) {
                hp[L] = dtl[L];
                hp[STEP_MAXL + L] = fz ? swirl_vt(tl[L], dtl[L]) : 0.0;
            }
            par_n++;
            return HIPCLAW_OK;
        }
        par_n++;
    }
    if (!merge_levels) {
        for (int L = hi; L >= lo; L--) HC_TRY(update_level(L, tl[L], dtl[L]));
        return HIPCLAW_OK;
    }
    const int s0 = plan[lo].s0, s1 = plan[hi].s1;
    for (int L = lo; L <= hi; L++)
        if (plan[L].s1 <= plan[L].s0 || (L > lo && plan[L].s0 != plan[L - 1].s1))
            return set_error(HIPCLAW_ERR_STATE, "merged update needs contiguous non-empty levels");
    const int np = s1 - s0;
    MultiLevel ml{};
    ml.levels = d_levels;
    ml.slot0 = s0;
    for (int L = lo; L <= hi; L++) {
        ml.dtl[L] = dtl[L];
        ml.qin_l[L] = pool[parity[L]];
        ml.qout_l[L] = pool[parity[L] ^ 1];
        if (fused_swirl()) {
            ml.swirl = 1;
            ml.xylower = d_xylower;
            ml.vtl[L] = swirl_vt(tl[L], dtl[L]);
        } else if (swirl_tperiod > 0 && solver_family(phys.solver) == HIPCLAW_SOLVER_ADV) {
            const int a0 = plan[L].s0, na = plan[L].s1 - a0;
            char* a = static_cast<char*>(aux) + (size_t)a0 * atile * esz;
            HC_TRY(hipclaw_swirl_aux_batch(&phys, na, d_xylower + 2 * a0, d_dxdy + 2 * a0, tl[L], dtl[L],
                                           swirl_tperiod, a, stream));
        }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (prof) {
        HC_TRY(check_hip(hipEventCreate(&e0), "event"));
        HC_TRY(check_hip(hipEventCreate(&e1), "event"));
        HC_TRY(check_hip(hipEventRecord(e0, stream), "event"));
    }
    const hipclaw_launch* lc = launch.cols_per_lane > 0 ? &launch : nullptr;
    const bool push = use_push();
    ml.push = push ? d_push : nullptr;
    ml.par = par_dev;
    // the launch's Courant-number maximum is taken by atomics in the step
    // kernel (one per wave): no reduction launch per fine step
    const bool cfl_atomic = cflacc_count < 4095;
    ml.cflmax = cfl_atomic ? d_cflslot + cflacc_count : nullptr;
    ml.reg_fresh = reg_fresh_run ? 1 : 0;
    ml.reg_lo = lo;
    ml.reg_cfaces = d_cfaces;
    HC_TRY(step2_launch(&phys, np, pool[0], pool[1], aux, d_dxdy, nullptr, np, dtl[hi], d_cfl,
                        tsync ? d_reg : nullptr, lc, stream, &ml));
    if (src2) {
        Src2Args sa = src2_args(s0, np);
        sa.levels = d_levels;
        sa.par = par_dev;
        for (int L = lo; L <= hi; L++) {
            sa.q_l[L] = ml.qout_l[L];
            sa.dtl[L] = dtl[L];
        }
        HC_TRY(src2_launch(sa, src2, phys.precision, nullptr, stream));
    }
    for (int L = lo; L <= hi; L++) pushed[L] = push;
    if (prof) {
        HC_TRY(check_hip(hipEventRecord(e1, stream), "event"));
        ev.push_back({e0, e1});
        ev_cells.push_back((long long)np * phys.mx * phys.my);
    }
    if (cflacc_count < 4095) {
        if (!cfl_atomic) HC_TRY(cfl_slot_reduce(d_cfl + s0, np, d_cflslot + cflacc_count, stream));
        cflacc_count++;
    }
    for (int L = lo; L <= hi; L++) parity[L] ^= 1;
    cell_updates += (long long)np * phys.mx * phys.my;
    return HIPCLAW_OK;
}

int Ctx::timeinterp(int level, double alpha)
{
    if (par_dry) return HIPCLAW_OK;
    const LevelPlan& lp = plan[level];
    if (esz == 8)
        launch_timeinterp<double>(lp.ti.p, lp.ti.n, pools<double>(), phys.mx, phys.my, phys.meqn, alpha, stream);
    else
        launch_timeinterp<float>(lp.ti.p, lp.ti.n, pools<float>(), phys.mx, phys.my, phys.meqn, alpha, stream);
    return check_hip(hipGetLastError(), "timeinterp");
}

// Subcycled coarse step (fclaw2d_advance_all_levels).
static unsigned level_bits(const int* parity)
{
    unsigned b = 0;
    for (int L = 0; L < MAXL; L++) b |= (parity[L] & 1u) << L;
    return b;
}
static unsigned level_bits(const bool* f)
{
    unsigned b = 0;
    for (int L = 0; L < MAXL; L++) b |= (f[L] ? 1u : 0u) << L;
    return b;
}

// Coarse steps with the Courant-number reduction; eager or as a hipGraph.
// A graph records the device work of nstep coarse steps; what depends on
// (t, dt), each step launch's dt and swirl vt per level, comes from the
// parameter block d_par, which a dry pass of the schedule fills on the host
// (the same arithmetic as the eager path) and one copy uploads before each
// launch.  The schedule itself (which levels step when, the
// time-interpolation weights) depends only on the step counters, so a graph
// serves every dt; it is keyed by the pool parities at entry, the mesh and
// the number of steps.  Step k's maximum Courant number goes to
// d_cflacc[4095] (one step) or d_cflacc[k] (several), read once at the end.
int Ctx::advance(double t, double dt, double* maxcfl)
{
    return advance_n(t, dt, 1, maxcfl);
}

int Ctx::advance_n(double t, double dt, int nstep, double* maxcfl, bool prepare)
{
    if (nstep < 1 || nstep > ADVANCE_MAXSTEPS) return set_error(HIPCLAW_ERR_ARG, "advance: nstep out of range");
    // fine steps of the finest level per coarse step (the step counters of
    // advance_body); a single coarse step of one fine step is a handful of
    // launches, for which launching a graph costs more than it saves
    const int flo = forest.global_minlevel(), fhi = forest.global_maxlevel();
    const int fine_steps = dom.subcycle ? 1 << (fhi - std::min(dom.minlevel, flo))
                                        : 1 << std::max(dom.maxlevel - dom.minlevel, fhi - flo);
    const bool graph_ok = use_graph && stream != nullptr && !prof && merge_levels &&
                          (!graph_auto || fine_steps * nstep > 1) &&
                          !(swirl_tperiod > 0 && solver_family(phys.solver) == HIPCLAW_SOLVER_ADV &&
                            !fused_swirl());
    double* const cfl_dev = nstep == 1 ? d_cflacc + 4095 : d_cflacc;
    auto steps = [&]() -> int {
        double tk = t;
        for (int k = 0; k < nstep; k++) {
            cfl_done = false;
            HC_TRY(advance_body(tk, dt, cfl_dev + k));
            if (!par_dry && !cfl_done) HC_TRY(cfl_final(d_cflslot, cflacc_count, cfl_dev + k, stream));
            tk += dt;
        }
        return HIPCLAW_OK;
    };
    if (prepare && !graph_ok) return HIPCLAW_OK;  // nothing to capture
    // time-sync registers: merged coarse steps from a clean state skip the
    // register clears (StepArgs::reg_fresh); others clear what earlier
    // merged steps left first.  Either way every register is zero for the
    // drivers afterwards (the last sync of a coarse step covers all levels).
    // A run that stops on an error leaves them to be cleared.
    reg_fresh_run = tsync && merge_levels && reg_clean;
    if (tsync && !prepare) {
        if (!reg_fresh_run) HC_TRY(flush_registers());
        reg_clean = true;
        reg_stale = true;
    }
    if (cflslot_dirty) HC_TRY(clear_cfl_slots());
    cflslot_dirty = true;
    if (!graph_ok) {
        HC_TRY(steps());
    } else {
        HC_TRY(ensure_par());
        // host parameters of these coarse steps into a free ring slot
        const int slot = par_slot;
        par_slot = (par_slot + 1) % PAR_RING;
        if (par_ev_used[slot]) HC_TRY(check_hip(hipEventSynchronize(par_ev[slot]), "parameter slot"));
        par_host = h_par + (size_t)slot * PAR_MAXLAUNCH * PAR_STRIDE;
        par_n = 0;
        par_dry = true;
        const int cnt0 = cflacc_count;
        int rc = steps();
        par_dry = false;
        cflacc_count = cnt0;
        const int nlaunch = par_n;
        if (rc) {
            par_host = nullptr;
            return rc;
        }
        if (!prepare) {
            rc = check_hip(hipMemcpyAsync(d_par, par_host, sizeof(double) * (size_t)nlaunch * PAR_STRIDE,
                                          hipMemcpyHostToDevice, stream),
                           "parameter upload");
            if (!rc) rc = check_hip(hipEventRecord(par_ev[slot], stream), "parameter event");
            if (rc) {
                par_host = nullptr;
                return rc;
            }
            par_ev_used[slot] = true;
        }
        // Key: pool parities, mesh, steps.  The ghost-push flags at entry do
        // not change the recorded work: the first fine step updates every
        // level, which sets them, before any ghost fill reads them.
        const unsigned pin = level_bits(parity), qin = level_bits(pushed);
        GraphEntry* e = nullptr;
        for (auto& g : graphs)
            if (g.parity_in == pin && g.epoch == mesh_epoch && g.nstep == nstep && g.reg_fresh == reg_fresh_run)
                e = &g;
        if (!e) {
            // capture: the body runs its host-side bookkeeping as usual while the
            // device work is recorded (its kernels read d_par); restore the
            // bookkeeping if capture fails
            int par0[MAXL];
            bool push0[MAXL];
            std::memcpy(par0, parity, sizeof(par0));
            std::memcpy(push0, pushed, sizeof(push0));
            const long long cells0 = cell_updates;
            par_n = 0;
            rc = check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
            if (!rc) {
                rc = steps();
                if (!rc && par_n != nlaunch) rc = set_error(HIPCLAW_ERR_STATE, "graph schedule mismatch");
                hipGraph_t g = nullptr;
                const hipError_t ec = hipStreamEndCapture(stream, &g);
                hipGraphExec_t ex = nullptr;
                if (!rc && ec == hipSuccess && g)
                    rc = check_hip(hipGraphInstantiateWithFlags(&ex, g, 0), "instantiate");
                else if (!rc)
                    rc = check_hip(ec, "end capture");
                if (g) (void)hipGraphDestroy(g);
                if (!rc) {
                    if (graphs.size() >= 4) {
                        (void)hipGraphExecDestroy(graphs.front().exec);
                        graphs.erase(graphs.begin());
                    }
                    graphs.push_back({pin, qin, mesh_epoch, nstep, reg_fresh_run, ex, level_bits(parity),
                                      level_bits(pushed), cell_updates - cells0, cflacc_count});
                    e = &graphs.back();
                    graph_captures++;
                }
            }
            par_host = nullptr;
            if (rc || prepare) {
                // failed capture, or nothing to run: back to the state at entry
                std::memcpy(parity, par0, sizeof(par0));
                std::memcpy(pushed, push0, sizeof(push0));
                cell_updates = cells0;
                cflacc_count = cnt0;
                cflslot_dirty = false;
                return rc;
            }
        } else if (prepare) {
            par_host = nullptr;
            cflslot_dirty = false;
            return HIPCLAW_OK;  // already captured
        } else {
            par_host = nullptr;
            graph_replays++;
            for (int L = 0; L < MAXL; L++) {
                parity[L] = (e->parity_out >> L) & 1u;
                pushed[L] = ((e->pushed_out >> L) & 1u) != 0;
            }
            cell_updates += e->cells;
            cflacc_count = e->cflacc;
        }
        HC_TRY(check_hip(hipGraphLaunch(e->exec, stream), "graph launch"));
    }
    if (tsync) reg_stale = reg_fresh_run;
    cflslot_dirty = false;
    if (maxcfl) {
        HC_TRY(check_hip(hipMemcpyAsync(maxcfl, cfl_dev, sizeof(double) * nstep, hipMemcpyDeviceToHost, stream),
                         "cfl D2H"));
        HC_TRY(check_hip(hipStreamSynchronize(stream), "advance"));
    }
    return HIPCLAW_OK;
}

int Ctx::advance_body(double t, double dt, double* cfl_out)
{
    const int lo = forest.global_minlevel(), hi = forest.global_maxlevel();
    struct Cnt { int last_step, step_inc, total_steps; double t, dt_step; } c[MAXL];
    const bool sub = dom.subcycle != 0;
    // time-step counters of initialize_timestep_counters (src/fclaw2d_advance.c:54-118):
    // they count from the options' minlevel m0, whose step dt is; a coarsest
    // present level lo > m0 therefore takes 2^(lo - m0) steps per coarse step.
    // Without subcycling every level takes 2^(maxlevel - minlevel) steps of
    // the options' level range.
    const int m0 = std::min(dom.minlevel, lo);
    const int rf = 1 << std::max(dom.maxlevel - dom.minlevel, hi - lo);
    double d = dt;  // the reference's dt_level /= 2 per level from minlevel
    for (int L = m0; L <= hi; L++) {
        c[L].last_step = 0;
        c[L].t = t;
        if (sub) {
            c[L].dt_step = d;
            c[L].step_inc = 1 << (hi - L);
            c[L].total_steps = 1 << (L - m0);
            d /= 2;
        } else {
            c[L].step_inc = 1;
            c[L].dt_step = dt / (double)rf;
            c[L].total_steps = rf;
        }
    }
    cflacc_count = 0;
    ti_pending.clear();
    // recursive level advance (advance_level).  The recursion is walked on the
    // host first: the levels it updates at one fine step are independent of
    // each other (each reads its own pool and writes its other one), so they
    // form one launch; the time interpolations follow in the recursion's order.
    double tl[MAXL], dtl[MAXL];
    int ulo = hi;
    std::vector<std::pair<int, double>> tis;
    auto advance_level = [&](auto& self, int level, int curr_fine_step) -> void {
        tl[level] = c[level].t;
        dtl[level] = c[level].dt_step;
        ulo = level;
        c[level].last_step += c[level].step_inc;
        c[level].t += c[level].dt_step;
        if (level > lo) {
            const int coarser = level - 1;
            if (c[coarser].last_step == curr_fine_step) {
                self(self, coarser, c[coarser].last_step);
                if (sub) {
                    const int coarse_inc = c[level - 1].step_inc;
                    tis.push_back({coarser, (double)(c[level].last_step % coarse_inc) / coarse_inc});
                }
            }
        }
    };
    const int nfine = c[hi].total_steps;
    for (int nf = 0; nf < nfine; nf++) {
        ulo = hi;
        tis.clear();
        advance_level(advance_level, hi, nf);
        HC_TRY(update_levels(ulo, hi, tl, dtl));
        if (merge_levels)  // run by the average launch of the next ghost fill
            ti_pending.insert(ti_pending.end(), tis.begin(), tis.end());
        else
            for (const auto& ti : tis) HC_TRY(timeinterp(ti.first, ti.second));
        if (c[hi].last_step < c[hi].total_steps) {
            if (sub) {
                int tl = hi - 1;
                while (c[hi].last_step % c[tl].step_inc == 0) tl--;
                // all present levels synchronised (a time-interpolated level
                // below lo has no patches): the update of [lo, hi]
                if (tl < lo) HC_TRY(ghost_update(lo, hi, 0, true));
                else HC_TRY(ghost_update(tl + 1, hi, 1, true));
            } else {
                HC_TRY(ghost_update(lo, hi, 0, true));
            }
        }
    }
    return ghost_update(lo, hi, 0, true, cfl_out);
}

// Copies of every level's current tiles (interiors and ghost cells) into the
// save pool and back, as clawpatch_save_step / clawpatch_restore_step do per
// patch (griddata_save = griddata, src/patches/clawpatch/fclaw2d_clawpatch.cpp:445-471)
// for save_time_step / restore_time_step (src/fclaw2d_run.c:155-158, 220-236).
// Each level's slots are contiguous, so a level is one device copy.  The
// time-sync registers need no copy: the final time sync of a coarse step
// resets every register, so they are zero at every step boundary.
int Ctx::save_step()
{
    if (n <= 0) return set_error(HIPCLAW_ERR_STATE, "save_step: no mesh");
    if (!save) HC_TRY(dmalloc(&save, (size_t)n * tile * esz));
    for (int L = 0; L < MAXL; L++) {
        const int s0 = plan[L].s0, np = plan[L].s1 - s0;
        if (np <= 0) continue;
        const size_t off = (size_t)s0 * tile * esz, bytes = (size_t)np * tile * esz;
        HC_TRY(check_hip(hipMemcpyAsync(static_cast<char*>(save) + off,
                                        static_cast<const char*>(pool[parity[L]]) + off, bytes,
                                        hipMemcpyDeviceToDevice, stream),
                         "save_step copy"));
    }
    std::memcpy(saved_pushed, pushed, sizeof(pushed));
    saved = true;
    saved_epoch = mesh_epoch;
    return HIPCLAW_OK;
}

int Ctx::restore_step()
{
    if (!saved || saved_epoch != mesh_epoch || !save)
        return set_error(HIPCLAW_ERR_STATE, "restore_step: no step saved on the current mesh");
    for (int L = 0; L < MAXL; L++) {
        const int s0 = plan[L].s0, np = plan[L].s1 - s0;
        if (np <= 0) continue;
        const size_t off = (size_t)s0 * tile * esz, bytes = (size_t)np * tile * esz;
        HC_TRY(check_hip(hipMemcpyAsync(static_cast<char*>(pool[parity[L]]) + off,
                                        static_cast<const char*>(save) + off, bytes,
                                        hipMemcpyDeviceToDevice, stream),
                         "restore_step copy"));
    }
    std::memcpy(pushed, saved_pushed, sizeof(pushed));
    return HIPCLAW_OK;
}

int Ctx::qinit(int kind, const QinitParams& prm)
{
    clear_pushed();
    if (esz == 8)
        launch_qinit<double>(kind, prm, d_iota, d_levels, n, pools<double>(), d_xylower, d_dxdy, phys.mx,
                             phys.my, phys.meqn, stream);
    else
        launch_qinit<float>(kind, prm, d_iota, d_levels, n, pools<float>(), d_xylower, d_dxdy, phys.mx,
                            phys.my, phys.meqn, stream);
    return check_hip(hipGetLastError(), "qinit");
}

// Tag every leaf (refine) and every complete family (coarsen), adapt the
// forest, repopulate the new leaves and refill ghost cells.
int Ctx::make_tiles(void* tiles, int* tags)
{
    if (esz == 8)
        HC_TRY(launch_make_tiles<double>(d_iota, d_levels, n, pools<double>(), d_dxdy, phys.mx, phys.my,
                                         phys.meqn, refine_threshold, tiles, stream));
    else
        HC_TRY(launch_make_tiles<float>(d_iota, d_levels, n, pools<float>(), d_dxdy, phys.mx, phys.my,
                                        phys.meqn, refine_threshold, tiles, stream));
    if (tags) {
        if (esz == 8)
            launch_tag<double>(d_iota, d_levels, n, pools<double>(), d_dxdy, phys.mx, phys.my, phys.meqn,
                               tag_ivar, refine_threshold, 0, tags, stream);
        else
            launch_tag<float>(d_iota, d_levels, n, pools<float>(), d_dxdy, phys.mx, phys.my, phys.meqn,
                              tag_ivar, refine_threshold, 0, tags, stream);
    }
    return check_hip(hipGetLastError(), "make tiles");
}

int Ctx::tag_and_adapt(int init_flag, int* changed)
{
    const int nn = n;
    // statistics of the tagging variable of every patch (k_tag_stats)
    if (tagst_cap < nn) {
        if (d_tagst) (void)hipFree(d_tagst);
        d_tagst = nullptr;
        tagst_cap = 0;
        HC_TRY(dmalloc(reinterpret_cast<void**>(&d_tagst), sizeof(double) * 4 * (size_t)nn));
        tagst_cap = nn;
    }
    if (esz == 8)
        launch_tag_stats<double>(d_iota, d_levels, nn, pools<double>(), d_dxdy, phys.mx, phys.my, phys.meqn,
                                 tag_ivar, d_tagst, stream);
    else
        launch_tag_stats<float>(d_iota, d_levels, nn, pools<float>(), d_dxdy, phys.mx, phys.my, phys.meqn,
                                tag_ivar, d_tagst, stream);
    std::vector<double> st(4 * (size_t)nn);
    // synchronous copies run on the null stream: wait for this context's stream
    HC_TRY(check_hip(hipStreamSynchronize(stream), "tags"));
    HC_TRY(check_hip(hipMemcpy(st.data(), d_tagst, sizeof(double) * st.size(), hipMemcpyDeviceToHost), "tags"));
    // the criterion of clawpatch's refinement-criteria on (min, max, max
    // difference, max gradient) of a patch or of a family
    auto exceeds = [&](const double* s4, double thr) {
        switch (tag_criterion) {
        case HIPCLAW_TAG_VALUE: return s4[1] > thr;
        case HIPCLAW_TAG_DIFFERENCE: return s4[2] > thr;
        case HIPCLAW_TAG_MINMAX: return s4[1] - s4[0] > thr;
        default: return s4[3] > thr;
        }
    };
    std::vector<int> rtag(nn), ctag(nn, 0);
    for (int k = 0; k < nn; k++) rtag[k] = exceeds(&st[4 * (size_t)k], refine_threshold) ? 1 : 0;
    if (cnn) {
        // learned refinement tags; the plain tag is the fallback and the judge
        if (tiles_cap < nn) {
            if (d_tiles) (void)hipFree(d_tiles);
            if (d_logits) (void)hipFree(d_logits);
            d_tiles = nullptr;
            d_logits = nullptr;
            tiles_cap = 0;
            HC_TRY(check_hip(hipMalloc(&d_tiles, (size_t)nn * 36 * 36 * 8 * 2), "tiles"));
            HC_TRY(check_hip(hipMalloc((void**)&d_logits, sizeof(float) * 2 * (size_t)nn), "logits"));
            tiles_cap = nn;
        }
        HC_TRY(make_tiles(d_tiles, nullptr));
        HC_TRY(hipclaw_cnn_forward(cnn, d_tiles, nullptr, nn, d_logits));
        std::vector<float> lg(2 * (size_t)nn);
        HC_TRY(check_hip(hipStreamSynchronize(stream), "cnn"));
        HC_TRY(check_hip(hipMemcpy(lg.data(), d_logits, sizeof(float) * lg.size(), hipMemcpyDeviceToHost),
                         "logits"));
        int ts6[6] = {nn, 0, 0, 0, 0, 0};
        for (int k = 0; k < nn; k++) {
            const double d = (double)lg[2 * k + 1] - (double)lg[2 * k];
            const double pr = 1.0 / (1.0 + std::exp(-d));
            int t = pr >= 0.5 ? 1 : 0;
            ts6[2] += rtag[k] != 0;
            if (std::fabs(pr - 0.5) < cnn_margin) {
                t = rtag[k] != 0;
                ts6[5]++;
            } else {
                ts6[3] += (rtag[k] != 0 && !t);
                ts6[4] += (rtag[k] == 0 && t);
            }
            ts6[1] += t;
            rtag[k] = t;
        }
        std::memcpy(tag_stats, ts6, sizeof(ts6));
    }
    if (!init_flag) {
        // families of four sibling leaves above minlevel coarsen when the
        // criterion does not exceed the coarsen threshold on the four together
        // (tag4coarsening never coarsens at initialisation)
        for (int k = 0; k < nn; k++) {
            const Leaf& l = forest.leaves[k];
            if (l.level <= dom.minlevel || (l.gx & 1) || (l.gy & 1)) continue;
            int fam[4];
            bool ok = true;
            for (int c = 0; c < 4 && ok; c++) {
                fam[c] = forest.find(l.level, l.gx + (c & 1), l.gy + (c >> 1));
                ok = fam[c] >= 0;
            }
            if (!ok) continue;
            double f4[4] = {st[4 * (size_t)fam[0]], st[4 * (size_t)fam[0] + 1], 0.0, 0.0};
            for (int c = 0; c < 4; c++) {
                const double* s4 = &st[4 * (size_t)fam[c]];
                f4[0] = std::min(f4[0], s4[0]);
                f4[1] = std::max(f4[1], s4[1]);
                f4[2] = std::max(f4[2], s4[2]);
                f4[3] = std::max(f4[3], s4[3]);
            }
            if (!exceeds(f4, coarsen_threshold))
                for (int c = 0; c < 4; c++) ctag[fam[c]] = 1;
        }
    }
    std::vector<int> tags(nn, 0);
    for (int k = 0; k < nn; k++) {
        const Leaf& l = forest.leaves[k];
        if (rtag[k] && l.level < dom.maxlevel) tags[k] = 1;
        else if (ctag[k]) tags[k] = -1;
    }
    if (smooth_refine) tags = forest.smooth_marks(tags, smooth_level);
    std::vector<Leaf> before = forest.leaves;
    std::vector<Transfer> xfer;
    try {
        forest.adapt(tags, xfer);
    } catch (const std::exception& e) {
        return set_error(HIPCLAW_ERR_STATE, std::string("adapt: ") + e.what());
    }
    bool same = before.size() == forest.leaves.size();
    for (size_t k = 0; same && k < before.size(); k++)
        same = before[k].level == forest.leaves[k].level && before[k].gx == forest.leaves[k].gx &&
               before[k].gy == forest.leaves[k].gy;
    if (changed) *changed = !same;
    if (same) return HIPCLAW_OK;
    HC_TRY(rebuild(true, &xfer));
    return HIPCLAW_OK;
}

int Ctx::mass(double* sums)
{
    if (esz == 8)
        launch_mass<double>(d_iota, d_levels, n, pools<double>(), d_dxdy, phys.mx, phys.my, phys.meqn, d_mass, stream);
    else
        launch_mass<float>(d_iota, d_levels, n, pools<float>(), d_dxdy, phys.mx, phys.my, phys.meqn, d_mass, stream);
    std::vector<double> h((size_t)n * phys.meqn);
    HC_TRY(check_hip(hipStreamSynchronize(stream), "mass"));
    HC_TRY(check_hip(hipMemcpy(h.data(), d_mass, sizeof(double) * h.size(), hipMemcpyDeviceToHost), "mass"));
    for (int m = 0; m < phys.meqn; m++) {
        double s = 0.0, c = 0.0;  // compensated sum over patches in leaf order
        for (int k = 0; k < n; k++) {
            double y = h[(size_t)k * phys.meqn + m] - c;
            double t = s + y;
            c = (t - s) - y;
            s = t;
        }
        sums[m] = s;
    }
    return HIPCLAW_OK;
}

}  // namespace hc

// ===========================================================================
// C ABI
// ===========================================================================
using namespace hc;

struct hipclaw_ctx {
    Ctx* c;
};

#define CTX_CHECK(ctx)                                                             \
    do {                                                                           \
        if (!(ctx) || !(ctx)->c) return set_error(HIPCLAW_ERR_ARG, "ctx is NULL"); \
    } while (0)
#define CTX_READY(ctx)                                                                    \
    do {                                                                                  \
        CTX_CHECK(ctx);                                                                   \
        if ((ctx)->c->n == 0) return set_error(HIPCLAW_ERR_STATE, "no forest: call hipclaw_init_uniform"); \
    } while (0)

extern "C" {

int hipclaw_ctx_create(const hipclaw_physics* phys, const hipclaw_domain* dom, void* stream,
                       hipclaw_ctx** out)
{
    int rc = validate_physics(phys);
    if (rc) return rc;
    if (!dom || !out) return set_error(HIPCLAW_ERR_ARG, "NULL argument");
    if (dom->mi < 1 || dom->mj < 1 || dom->minlevel < 0 || dom->maxlevel < dom->minlevel ||
        dom->maxlevel >= MAXL - 1 || !(dom->bx > dom->ax) || !(dom->by > dom->ay))
        return set_error(HIPCLAW_ERR_ARG, "bad domain");
    if (phys->mx % 2 || phys->my % 2) return set_error(HIPCLAW_ERR_ARG, "mx, my must be even for AMR");
    if (solver_family(phys->solver) == HIPCLAW_SOLVER_TRANSPORT_FW)
        return set_error(HIPCLAW_ERR_UNSUPPORTED,
                         "f-wave transport steps patch batches (hipclaw_step2_batch); the AMR context has no fill of its edge data");
    for (int f = 0; f < 4; f++)
        if (phys->mthbc[f] != 1 && phys->mthbc[f] != 3)
            return set_error(HIPCLAW_ERR_UNSUPPORTED, "mthbc must be 1 (extrapolation) or 3 (wall)");
    if (phys->meqn < 3)
        for (int f = 0; f < 4; f++)
            if (phys->mthbc[f] == 3)
                return set_error(HIPCLAW_ERR_ARG, "wall boundaries need a momentum field (meqn >= 3)");
    auto* h = new (std::nothrow) hipclaw_ctx;
    if (!h) return set_error(HIPCLAW_ERR_NOMEM, "host allocation failed");
    try {
        h->c = new Ctx(*phys, *dom);
    } catch (const std::exception& e) {
        delete h;
        return set_error(HIPCLAW_ERR_ARG, e.what());
    }
    h->c->stream = static_cast<hipStream_t>(stream);
    // whole coarse steps replay as graphs where the stream can be captured
    h->c->use_graph = h->c->graph_auto = h->c->stream != nullptr;
    *out = h;
    return HIPCLAW_OK;
}

int hipclaw_ctx_destroy(hipclaw_ctx* ctx)
{
    if (!ctx) return HIPCLAW_OK;
    if (ctx->c) (void)hipStreamSynchronize(ctx->c->stream);
    delete ctx->c;
    delete ctx;
    return HIPCLAW_OK;
}

int hipclaw_init_uniform(hipclaw_ctx* ctx, int level)
{
    CTX_CHECK(ctx);
    Ctx& c = *ctx->c;
    if (level < c.dom.minlevel || level > c.dom.maxlevel) return set_error(HIPCLAW_ERR_ARG, "level outside [minlevel, maxlevel]");
    c.forest.init_uniform(level);
    return c.rebuild(false, nullptr);
}

int hipclaw_init_leaves(hipclaw_ctx* ctx, int nleaves, const int* level, const int* gx, const int* gy)
{
    CTX_CHECK(ctx);
    Ctx& c = *ctx->c;
    if (nleaves < 1 || !level || !gx || !gy) return set_error(HIPCLAW_ERR_ARG, "bad leaf list");
    std::vector<Leaf> in((size_t)nleaves);
    for (int k = 0; k < nleaves; k++) in[k] = {level[k], gx[k], gy[k]};
    std::string err = c.forest.set_leaves(in);
    if (!err.empty()) return set_error(HIPCLAW_ERR_ARG, err);
    return c.rebuild(false, nullptr);
}

int hipclaw_set_launch(hipclaw_ctx* ctx, const hipclaw_launch* launch)
{
    CTX_CHECK(ctx);
    if (!launch) {
        ctx->c->launch = hipclaw_launch{0, 0, 0};
        ctx->c->clear_graphs();
        return HIPCLAW_OK;
    }
    int cpl, wpg, nband;
    HC_TRY(choose_launch(&ctx->c->phys, 1, launch, &cpl, &wpg, &nband));  // validates the shape
    ctx->c->launch = *launch;
    ctx->c->clear_graphs();
    return HIPCLAW_OK;
}

int hipclaw_set_time_sync(hipclaw_ctx* ctx, int on)
{
    CTX_CHECK(ctx);
    return ctx->c->set_time_sync(on != 0);
}

int hipclaw_set_src2(hipclaw_ctx* ctx, int kind)
{
    CTX_CHECK(ctx);
    if (kind != HIPCLAW_SRC2_NONE && kind != HIPCLAW_SRC2_AXISYMMETRIC)
        return set_error(HIPCLAW_ERR_ARG, "unknown source term");
    if (kind == HIPCLAW_SRC2_AXISYMMETRIC && ctx->c->phys.meqn < 4)
        return set_error(HIPCLAW_ERR_ARG, "axisymmetric source needs the Euler solver");
    ctx->c->src2 = kind;
    ctx->c->clear_graphs();
    return HIPCLAW_OK;
}

int hipclaw_time_sync(hipclaw_ctx* ctx, int minlevel, int maxlevel)
{
    CTX_READY(ctx);
    HC_TRY(ctx->c->flush_registers());
    return ctx->c->time_sync(minlevel, maxlevel);
}

int hipclaw_get_registers(hipclaw_ctx* ctx, double* host)
{
    CTX_READY(ctx);
    Ctx& c = *ctx->c;
    if (!host) return set_error(HIPCLAW_ERR_ARG, "host is NULL");
    if (!c.d_reg) return set_error(HIPCLAW_ERR_ARG, "time sync is off");
    HC_TRY(c.flush_registers());
    int rc = check_hip(hipStreamSynchronize(c.stream), "sync");
    if (rc) return rc;
    return check_hip(hipMemcpy(host, c.d_reg, sizeof(double) * (size_t)c.n * c.regstride,
                               hipMemcpyDeviceToHost), "registers");
}

int hipclaw_set_feature(hipclaw_ctx* ctx, int feature, int on)
{
    CTX_CHECK(ctx);
    Ctx& c = *ctx->c;
    c.clear_graphs();
    switch (feature) {
    case HIPCLAW_FEATURE_MERGE_LEVELS: c.merge_levels = on != 0; break;
    case HIPCLAW_FEATURE_GHOST_PUSH:
        c.push_ghosts = on != 0;
        c.push_auto = false;
        c.clear_pushed();
        break;
    case HIPCLAW_FEATURE_GRAPH: c.use_graph = on != 0; c.graph_auto = false; break;
    case HIPCLAW_FEATURE_INIT_GHOSTCELL: c.init_ghostcell = on != 0; break;
    case HIPCLAW_FEATURE_FUSE_AUX: c.fuse_aux = on != 0; break;
    default: return set_error(HIPCLAW_ERR_ARG, "unknown feature");
    }
    return HIPCLAW_OK;
}

int hipclaw_set_tag_criterion(hipclaw_ctx* ctx, int criterion)
{
    CTX_CHECK(ctx);
    if (criterion < HIPCLAW_TAG_VALUE || criterion > HIPCLAW_TAG_GRADIENT)
        return set_error(HIPCLAW_ERR_ARG, "criterion: HIPCLAW_TAG_VALUE, _DIFFERENCE, _MINMAX or _GRADIENT");
    ctx->c->tag_criterion = criterion;
    return HIPCLAW_OK;
}

int hipclaw_set_refinement(hipclaw_ctx* ctx, int smooth_refine, int smooth_level, int coarsen_delay)
{
    CTX_CHECK(ctx);
    if (smooth_level < 0) return set_error(HIPCLAW_ERR_ARG, "smooth_level must be >= 0");
    if (coarsen_delay != 0) return set_error(HIPCLAW_ERR_UNSUPPORTED, "coarsen_delay > 0 is not supported");
    ctx->c->smooth_refine = smooth_refine != 0;
    ctx->c->smooth_level = smooth_level;
    return HIPCLAW_OK;
}

int hipclaw_graph_stats(hipclaw_ctx* ctx, long long* captures, long long* replays)
{
    CTX_CHECK(ctx);
    if (captures) *captures = ctx->c->graph_captures;
    if (replays) *replays = ctx->c->graph_replays;
    return HIPCLAW_OK;
}

int hipclaw_make_tiles(hipclaw_ctx* ctx, void* tiles_dev, int* tags_dev)
{
    CTX_READY(ctx);
    if (!tiles_dev) return set_error(HIPCLAW_ERR_ARG, "tiles_dev is NULL");
    return ctx->c->make_tiles(tiles_dev, tags_dev);
}

int hipclaw_set_tagging_cnn(hipclaw_ctx* ctx, hipclaw_cnn* net, float margin)
{
    CTX_CHECK(ctx);
    if (net && (ctx->c->phys.mx != 32 || ctx->c->phys.my != 32))
        return set_error(HIPCLAW_ERR_UNSUPPORTED, "the network takes 32x32 patches");
    if (!(margin >= 0.f && margin <= 0.5f)) return set_error(HIPCLAW_ERR_ARG, "margin must lie in [0, 0.5]");
    ctx->c->cnn = net;
    ctx->c->cnn_margin = margin;
    return HIPCLAW_OK;
}

int hipclaw_tag_stats(hipclaw_ctx* ctx, int* stats6)
{
    CTX_CHECK(ctx);
    if (!stats6) return set_error(HIPCLAW_ERR_ARG, "stats6 is NULL");
    std::memcpy(stats6, ctx->c->tag_stats, sizeof(ctx->c->tag_stats));
    return HIPCLAW_OK;
}

int hipclaw_save_step(hipclaw_ctx* ctx)
{
    CTX_READY(ctx);
    return ctx->c->save_step();
}

int hipclaw_restore_step(hipclaw_ctx* ctx)
{
    CTX_READY(ctx);
    return ctx->c->restore_step();
}

int hipclaw_get_launch(hipclaw_ctx* ctx, hipclaw_launch* launch)
{
    CTX_CHECK(ctx);
    if (!launch) return set_error(HIPCLAW_ERR_ARG, "launch is NULL");
    *launch = ctx->c->launch;
    return HIPCLAW_OK;
}

int hipclaw_autotune(hipclaw_ctx* ctx, int level, int reps, hipclaw_launch* best, float* ms)
{
    CTX_READY(ctx);
    return ctx->c->autotune(level, reps, best, ms);
}

int hipclaw_num_patches(hipclaw_ctx* ctx, int* n)
{
    CTX_CHECK(ctx);
    if (!n) return set_error(HIPCLAW_ERR_ARG, "n is NULL");
    *n = ctx->c->n;
    return HIPCLAW_OK;
}

int hipclaw_patch_info(hipclaw_ctx* ctx, int* level, int* gx, int* gy, double* xlower, double* ylower,
                       double* dx, double* dy)
{
    CTX_READY(ctx);
    Ctx& c = *ctx->c;
    std::vector<double> xy(2 * (size_t)c.n), dd(2 * (size_t)c.n);
    int rc = check_hip(hipMemcpy(xy.data(), c.d_xylower, sizeof(double) * xy.size(), hipMemcpyDeviceToHost), "info");
    if (rc) return rc;
    rc = check_hip(hipMemcpy(dd.data(), c.d_dxdy, sizeof(double) * dd.size(), hipMemcpyDeviceToHost), "info");
    if (rc) return rc;
    for (int k = 0; k < c.n; k++) {
        const Leaf& l = c.forest.leaves[k];
        if (level) level[k] = l.level;
        if (gx) gx[k] = l.gx;
        if (gy) gy[k] = l.gy;
        if (xlower) xlower[k] = xy[2 * k];
        if (ylower) ylower[k] = xy[2 * k + 1];
        if (dx) dx[k] = dd[2 * k];
        if (dy) dy[k] = dd[2 * k + 1];
    }
    return HIPCLAW_OK;
}

static int copy_q(Ctx& c, double* host, bool to_device, bool auxdata)
{
    if (to_device) c.clear_pushed();
    HC_TRY(check_hip(hipStreamSynchronize(c.stream), "q copy"));
    const long long per = auxdata ? c.atile : c.tile;
    const size_t tot = (size_t)c.n * per;
    if (tot == 0) return HIPCLAW_OK;
    std::vector<char> tmp;
    for (int L = 0; L < MAXL; L++) {
        const int s0 = c.plan[L].s0, s1 = c.plan[L].s1;
        if (s1 <= s0) continue;
        char* dev = static_cast<char*>(auxdata ? c.aux : c.pool[c.parity[L]]) + (size_t)s0 * per * c.esz;
        const size_t cnt = (size_t)(s1 - s0) * per;
        double* h = host + (size_t)s0 * per;
        if (c.esz == 8) {
            int rc = check_hip(hipMemcpy(to_device ? (void*)dev : (void*)h, to_device ? (void*)h : (void*)dev,
                                         cnt * 8, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost),
                               "q copy");
            if (rc) return rc;
        } else {
            std::vector<float> f(cnt);
            if (to_device) {
                for (size_t i = 0; i < cnt; i++) f[i] = (float)h[i];
                int rc = check_hip(hipMemcpy(dev, f.data(), cnt * 4, hipMemcpyHostToDevice), "q copy");
                if (rc) return rc;
            } else {
                int rc = check_hip(hipMemcpy(f.data(), dev, cnt * 4, hipMemcpyDeviceToHost), "q copy");
                if (rc) return rc;
                for (size_t i = 0; i < cnt; i++) h[i] = (double)f[i];
            }
        }
    }
    return HIPCLAW_OK;
}

// Per-patch host arrays of one level <-> the level's slots of the current pool.
static int level_copy(Ctx& c, int level, int n, double* const* q, const double* const* aux, bool up)
{
    if (up) c.clear_pushed();
    HC_TRY(check_hip(hipStreamSynchronize(c.stream), "level copy"));
    if (level < 0 || level >= MAXL) return set_error(HIPCLAW_ERR_ARG, "level out of range");
    const int s0 = c.plan[level].s0, np = c.plan[level].s1 - s0;
    if (n != np) return set_error(HIPCLAW_ERR_ARG, "n differs from the number of patches of the level");
    if (np == 0) return HIPCLAW_OK;
    if (!q) return set_error(HIPCLAW_ERR_ARG, "q is NULL");
    for (int k = 0; k < n; k++)
        if (!q[k]) return set_error(HIPCLAW_ERR_ARG, "a patch pointer is NULL");
    auto move = [&](double* const* ptrs, long long per, void* dev_base) -> int {
        char* dev = static_cast<char*>(dev_base) + (size_t)s0 * per * c.esz;
        const size_t cnt = (size_t)np * per;
        if (c.esz == 8) {
            std::vector<double> h(cnt);
            if (up) {
                for (int k = 0; k < n; k++) std::memcpy(&h[(size_t)k * per], ptrs[k], sizeof(double) * per);
                return check_hip(hipMemcpy(dev, h.data(), cnt * 8, hipMemcpyHostToDevice), "level upload");
            }
            int rc = check_hip(hipMemcpy(h.data(), dev, cnt * 8, hipMemcpyDeviceToHost), "level download");
            if (rc) return rc;
            for (int k = 0; k < n; k++) std::memcpy(ptrs[k], &h[(size_t)k * per], sizeof(double) * per);
            return HIPCLAW_OK;
        }
        std::vector<float> h(cnt);
        if (up) {
            for (int k = 0; k < n; k++)
                for (long long i = 0; i < per; i++) h[(size_t)k * per + i] = (float)ptrs[k][i];
            return check_hip(hipMemcpy(dev, h.data(), cnt * 4, hipMemcpyHostToDevice), "level upload");
        }
        int rc = check_hip(hipMemcpy(h.data(), dev, cnt * 4, hipMemcpyDeviceToHost), "level download");
        if (rc) return rc;
        for (int k = 0; k < n; k++)
            for (long long i = 0; i < per; i++) ptrs[k][i] = (double)h[(size_t)k * per + i];
        return HIPCLAW_OK;
    };
    if (!up) {
        int rc = check_hip(hipStreamSynchronize(c.stream), "sync");
        if (rc) return rc;
    }
    int rc = move(q, c.tile, c.pool[c.parity[level]]);
    if (rc) return rc;
    if (up && aux && c.phys.maux > 0) {
        for (int k = 0; k < n; k++)
            if (!aux[k]) return set_error(HIPCLAW_ERR_ARG, "an aux pointer is NULL");
        c.swirl_tperiod = 0;
        rc = move(const_cast<double* const*>(aux), c.atile, c.aux);
    }
    return rc;
}

// Host patch arrays in the Clawpack 5 layout (HIPCLAW_CLAW5 solvers) go
// through tile-layout copies.
static std::vector<double> to_tiles(const double* host, int npatch, int nf, long long cells)
{
    std::vector<double> t((size_t)npatch * nf * cells);
    for (int k = 0; k < npatch; k++)
        patch_from_host(host + (size_t)k * nf * cells, t.data() + (size_t)k * nf * cells, nf, cells, true);
    return t;
}

int hipclaw_level_upload(hipclaw_ctx* ctx, int level, int n, const double* const* q,
                         const double* const* aux)
{
    CTX_READY(ctx);
    Ctx& c = *ctx->c;
    if (!is_claw5(c.phys.solver) || n <= 0 || !q)
        return level_copy(c, level, n, const_cast<double* const*>(q), aux, true);
    const long long cells = (long long)(c.phys.mx + 4) * (c.phys.my + 4);
    std::vector<std::vector<double>> tq(n), ta;
    std::vector<double*> pq(n), pa;
    for (int k = 0; k < n; k++) {
        if (!q[k]) return set_error(HIPCLAW_ERR_ARG, "a patch pointer is NULL");
        tq[k] = to_tiles(q[k], 1, c.phys.meqn, cells);
        pq[k] = tq[k].data();
    }
    if (aux && c.phys.maux > 0) {
        ta.resize(n);
        pa.resize(n);
        for (int k = 0; k < n; k++) {
            if (!aux[k]) return set_error(HIPCLAW_ERR_ARG, "an aux pointer is NULL");
            ta[k] = to_tiles(aux[k], 1, c.phys.maux, cells);
            pa[k] = ta[k].data();
        }
    }
    return level_copy(c, level, n, pq.data(), pa.empty() ? nullptr : pa.data(), true);
}

int hipclaw_level_download(hipclaw_ctx* ctx, int level, int n, double* const* q)
{
    CTX_READY(ctx);
    Ctx& c = *ctx->c;
    if (!is_claw5(c.phys.solver) || n <= 0 || !q) return level_copy(c, level, n, q, nullptr, false);
    const long long cells = (long long)(c.phys.mx + 4) * (c.phys.my + 4);
    std::vector<std::vector<double>> tq(n);
    std::vector<double*> pq(n);
    for (int k = 0; k < n; k++) {
        if (!q[k]) return set_error(HIPCLAW_ERR_ARG, "a patch pointer is NULL");
        tq[k].resize((size_t)c.phys.meqn * cells);
        pq[k] = tq[k].data();
    }
    int rc = level_copy(c, level, n, pq.data(), nullptr, false);
    if (rc) return rc;
    for (int k = 0; k < n; k++) patch_to_host(pq[k], q[k], c.phys.meqn, cells, true);
    return HIPCLAW_OK;
}

int hipclaw_set_q(hipclaw_ctx* ctx, const double* host_q)
{
    CTX_READY(ctx);
    if (!host_q) return set_error(HIPCLAW_ERR_ARG, "host_q is NULL");
    Ctx& c = *ctx->c;
    if (!is_claw5(c.phys.solver)) return copy_q(c, const_cast<double*>(host_q), true, false);
    std::vector<double> t = to_tiles(host_q, c.n, c.phys.meqn, (long long)(c.phys.mx + 4) * (c.phys.my + 4));
    return copy_q(c, t.data(), true, false);
}

int hipclaw_get_q(hipclaw_ctx* ctx, double* host_q)
{
    CTX_READY(ctx);
    if (!host_q) return set_error(HIPCLAW_ERR_ARG, "host_q is NULL");
    Ctx& c = *ctx->c;
    int rc = check_hip(hipStreamSynchronize(c.stream), "sync");
    if (rc) return rc;
    if (!is_claw5(c.phys.solver)) return copy_q(c, host_q, false, false);
    const long long cells = (long long)(c.phys.mx + 4) * (c.phys.my + 4);
    std::vector<double> t((size_t)c.n * c.phys.meqn * cells);
    rc = copy_q(c, t.data(), false, false);
    if (rc) return rc;
    for (int k = 0; k < c.n; k++)
        patch_to_host(t.data() + (size_t)k * c.phys.meqn * cells, host_q + (size_t)k * c.phys.meqn * cells,
                      c.phys.meqn, cells, true);
    return HIPCLAW_OK;
}

int hipclaw_output_ascii(hipclaw_ctx* ctx, int iframe, double time, const char* dir)
{
    CTX_READY(ctx);
    Ctx& c = *ctx->c;
    const int n = c.n;
    std::vector<int> lev(n), gx(n), gy(n);
    std::vector<double> xl(n), yl(n), dx(n), dy(n);
    int rc = hipclaw_patch_info(ctx, lev.data(), gx.data(), gy.data(), xl.data(), yl.data(), dx.data(), dy.data());
    if (rc) return rc;
    std::vector<double> q((size_t)n * c.tile);  // tile = elements per slot, tile layout
    rc = check_hip(hipStreamSynchronize(c.stream), "sync");
    if (!rc) rc = copy_q(c, q.data(), false, false);
    if (rc) return rc;
    return write_frame(dir, iframe, time, c.dom.mi, c.phys.mx, c.phys.my, c.phys.meqn, c.phys.maux, n, lev.data(),
                       gx.data(), gy.data(), xl.data(), yl.data(), dx.data(), dy.data(), q.data());
}

int hipclaw_set_aux(hipclaw_ctx* ctx, const double* host_aux)
{
    CTX_READY(ctx);
    if (!host_aux || ctx->c->phys.maux == 0) return set_error(HIPCLAW_ERR_ARG, "no aux for this solver");
    Ctx& c = *ctx->c;
    c.swirl_tperiod = 0;
    if (!is_claw5(c.phys.solver)) return copy_q(c, const_cast<double*>(host_aux), true, true);
    std::vector<double> t = to_tiles(host_aux, c.n, c.phys.maux, (long long)(c.phys.mx + 4) * (c.phys.my + 4));
    return copy_q(c, t.data(), true, true);
}

int hipclaw_set_swirl(hipclaw_ctx* ctx, double tperiod)
{
    CTX_CHECK(ctx);
    if (solver_family(ctx->c->phys.solver) != HIPCLAW_SOLVER_ADV) return set_error(HIPCLAW_ERR_ARG, "swirl needs the advection solver");
    if (!(tperiod >= 0)) return set_error(HIPCLAW_ERR_ARG, "tperiod must be >= 0");
    ctx->c->swirl_tperiod = tperiod;
    return HIPCLAW_OK;
}

int hipclaw_qinit(hipclaw_ctx* ctx, int kind, const double* params, int nparams, int nsub)
{
    CTX_READY(ctx);
    if (kind < 0 || kind > QINIT_SHOCKBUBBLE_APP) return set_error(HIPCLAW_ERR_ARG, "unknown qinit kind");
    if (nparams < 0 || nparams > 56 || (nparams > 0 && !params)) return set_error(HIPCLAW_ERR_ARG, "bad qinit params");
    if (nsub < 0 || nsub > 64) return set_error(HIPCLAW_ERR_ARG, "nsub must be 0..64");
    QinitParams prm{};
    for (int k = 0; k < nparams; k++) prm.p[k] = params[k];
    prm.nsub = nsub;
    return ctx->c->qinit(kind, prm);
}

int hipclaw_ghost_update(hipclaw_ctx* ctx, int minlevel, int maxlevel, double sync_time, int time_interp)
{
    CTX_READY(ctx);
    (void)sync_time;
    if (minlevel < 0 || maxlevel >= MAXL - 1 || maxlevel < minlevel) return set_error(HIPCLAW_ERR_ARG, "bad level range");
    return ctx->c->ghost_update(minlevel, maxlevel, time_interp);
}

int hipclaw_update_level(hipclaw_ctx* ctx, int level, double t, double dt, double* cfl)
{
    CTX_READY(ctx);
    Ctx& c = *ctx->c;
    if (level < 0 || level >= MAXL) return set_error(HIPCLAW_ERR_ARG, "level out of range");
    if (c.cflslot_dirty) HC_TRY(c.clear_cfl_slots());
    c.cflacc_count = 0;
    c.cflslot_dirty = true;
    if (c.tsync) {
        HC_TRY(c.flush_registers());
        c.reg_clean = false;  // its sums stay until a sync clears them
    }
    int rc = c.update_level(level, t, dt);
    if (rc) return rc;
    if (c.cflacc_count > 0) HC_TRY(cfl_final(c.d_cflslot, c.cflacc_count, c.d_cflacc + 4095, c.stream));
    c.cflslot_dirty = false;
    if (cfl) {
        if (c.cflacc_count == 0) {
            *cfl = 0.0;
            return HIPCLAW_OK;
        }
        rc = check_hip(hipMemcpyAsync(cfl, c.d_cflacc + 4095, sizeof(double), hipMemcpyDeviceToHost, c.stream),
                       "cfl");
        if (rc) return rc;
        return check_hip(hipStreamSynchronize(c.stream), "update_level");
    }
    return HIPCLAW_OK;
}

int hipclaw_timeinterp(hipclaw_ctx* ctx, int level, double alpha)
{
    CTX_READY(ctx);
    if (level < 0 || level >= MAXL) return set_error(HIPCLAW_ERR_ARG, "level out of range");
    return ctx->c->timeinterp(level, alpha);
}

int hipclaw_advance(hipclaw_ctx* ctx, double t, double dt, double* maxcfl)
{
    CTX_READY(ctx);
    if (!(dt > 0)) return set_error(HIPCLAW_ERR_ARG, "dt must be > 0");
    return ctx->c->advance(t, dt, maxcfl);
}

int hipclaw_advance_steps(hipclaw_ctx* ctx, double t, double dt, int nsteps, double* maxcfl)
{
    CTX_READY(ctx);
    if (!(dt > 0)) return set_error(HIPCLAW_ERR_ARG, "dt must be > 0");
    if (nsteps < 1 || nsteps > Ctx::ADVANCE_MAXSTEPS)
        return set_error(HIPCLAW_ERR_ARG, "nsteps must be 1..256");
    return ctx->c->advance_n(t, dt, nsteps, maxcfl);
}

int hipclaw_prepare_steps(hipclaw_ctx* ctx, int nsteps)
{
    CTX_READY(ctx);
    if (nsteps < 1 || nsteps > Ctx::ADVANCE_MAXSTEPS)
        return set_error(HIPCLAW_ERR_ARG, "nsteps must be 1..256");
    // the schedule does not depend on (t, dt): any positive dt records it
    return ctx->c->advance_n(0.0, 1.0, nsteps, nullptr, true);
}

int hipclaw_set_tagging(hipclaw_ctx* ctx, int ivar, double refine_threshold, double coarsen_threshold)
{
    CTX_CHECK(ctx);
    if (ivar < 0 || ivar >= ctx->c->phys.meqn) return set_error(HIPCLAW_ERR_ARG, "ivar out of range");
    ctx->c->tag_ivar = ivar;
    ctx->c->refine_threshold = refine_threshold;
    ctx->c->coarsen_threshold = coarsen_threshold;
    return HIPCLAW_OK;
}

int hipclaw_regrid(hipclaw_ctx* ctx, int init_flag, int* changed)
{
    CTX_READY(ctx);
    Ctx& c = *ctx->c;
    int ch = 0;
    int rc = c.tag_and_adapt(init_flag, &ch);
    if (rc) return rc;
    if (changed) *changed = ch;
    if (ch) rc = c.ghost_update(c.forest.global_minlevel(), c.forest.global_maxlevel(), 0);
    return rc;
}

int hipclaw_initialize(hipclaw_ctx* ctx, int qinit_kind, const double* params, int nparams, int nsub)
{
    CTX_CHECK(ctx);
    Ctx& c = *ctx->c;
    if (c.n == 0) {
        int rc = hipclaw_init_uniform(ctx, c.dom.minlevel);
        if (rc) return rc;
    }
    int rc = hipclaw_qinit(ctx, qinit_kind, params, nparams, nsub);
    if (rc) return rc;
    // initial refinement (fclaw2d_initialize, src/fclaw2d_initialize.c:160-223):
    // one tag and adapt pass per level from minlevel to maxlevel, new patches
    // initialised by qinit, stopping at the first pass that changes nothing.
    // The tags see the qinit values in the ghost cells unless init_ghostcell
    // asks for ghost updates (the reference's option of that name).
    for (int pass = c.dom.minlevel; pass <= c.dom.maxlevel; pass++) {
        if (c.init_ghostcell) {
            rc = c.ghost_update(c.forest.global_minlevel(), c.forest.global_maxlevel(), 0);
            if (rc) return rc;
        }
        int ch = 0;
        rc = c.tag_and_adapt(1, &ch);
        if (rc) return rc;
        if (!ch) break;
        rc = hipclaw_qinit(ctx, qinit_kind, params, nparams, nsub);
        if (rc) return rc;
    }
    if (!c.init_ghostcell) return HIPCLAW_OK;
    return c.ghost_update(c.forest.global_minlevel(), c.forest.global_maxlevel(), 0);
}

int hipclaw_mass(hipclaw_ctx* ctx, double* sums)
{
    CTX_READY(ctx);
    if (!sums) return set_error(HIPCLAW_ERR_ARG, "sums is NULL");
    return ctx->c->mass(sums);
}

int hipclaw_profile(hipclaw_ctx* ctx, int enable)
{
    CTX_CHECK(ctx);
    ctx->c->prof_clear();
    ctx->c->prof = enable != 0;
    return HIPCLAW_OK;
}

int hipclaw_profile_read(hipclaw_ctx* ctx, double* kernel_ms, long long* launches, long long* cells)
{
    CTX_CHECK(ctx);
    Ctx& c = *ctx->c;
    int rc = check_hip(hipStreamSynchronize(c.stream), "profile sync");
    if (rc) return rc;
    double tot = 0.0;
    long long cl = 0;
    for (size_t k = 0; k < c.ev.size(); k++) {
        float ms = 0.f;
        rc = check_hip(hipEventElapsedTime(&ms, c.ev[k].first, c.ev[k].second), "elapsed");
        if (rc) return rc;
        tot += ms;
        cl += c.ev_cells[k];
    }
    if (kernel_ms) *kernel_ms = tot;
    if (launches) *launches = (long long)c.ev.size();
    if (cells) *cells = cl;
    return HIPCLAW_OK;
}

int hipclaw_cell_updates(hipclaw_ctx* ctx, long long* n)
{
    CTX_CHECK(ctx);
    if (!n) return set_error(HIPCLAW_ERR_ARG, "n is NULL");
    *n = ctx->c->cell_updates;
    return HIPCLAW_OK;
}

}  // extern "C"
