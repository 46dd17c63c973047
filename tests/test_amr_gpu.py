"""GPU parity of the device-resident AMR driver (hipclaw_ctx_* through
forestclaw_amd.forest) against the Python restatement of the ForestClaw
drivers on top of the C oracle (oracle/amr.py).

FP64: every patch interior after several subcycled coarse steps (ghost fill,
time interpolation, level updates) is bitwise equal to the oracle, and so is
the maximum Courant number.  FP32: relative L1 error against the FP64 oracle
is reported and bounded.
"""
import numpy as np
import pytest

from forestclaw_amd import forest as F
from forestclaw_amd import hipclaw as H
from oracle import amr as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FP32_REL_L1 = 1e-5


def make(solver, mx, minlevel, maxlevel, *, precision=H.FP64, mi=1, mj=1, box=(0, 1, 0, 1),
         mthbc=(1, 1, 1, 1), refine=0.5, coarsen=0.05, subcycle=True):
    opt = F.Options(minlevel=minlevel, maxlevel=maxlevel, mx=mx, my=mx, solver=solver,
                    precision=precision, mi=mi, mj=mj, ax=box[0], bx=box[1], ay=box[2],
                    by=box[3], mthbc=mthbc, refine_threshold=refine, coarsen_threshold=coarsen,
                    subcycle=subcycle)
    return F.Hipclaw46(opt)


def oracle_from(sim, solver, swirl=0.0):
    info = sim.patch_info()
    leaves = list(zip(info["level"], info["gx"], info["gy"]))
    geom = np.stack([info["xlower"], info["ylower"], info["dx"], info["dy"]], 1)
    cfg = O.make_cfg(solver, sim.opt.mx, sim.opt.my, 2,
                     gamma=sim.opt.gamma, grav=sim.opt.grav)
    amr = A.AmrOracle(cfg, leaves, geom, sim.opt.mthbc, brick=(sim.opt.mi, sim.opt.mj),
                      subcycle=sim.opt.subcycle, swirl_tperiod=swirl, minlevel=sim.opt.minlevel,
                      maxlevel=sim.opt.maxlevel)
    q = sim.get_q()
    for k in range(len(leaves)):
        amr.q[k][...] = q[k]
    return amr


def interiors(arrs):
    return np.stack([a[:, 2:-2, 2:-2] for a in arrs])


DAM = (0.5, 0.5, 0.23, 2.0, 1.0)
SMALL_DAM = (0.3, 0.35, 0.15, 2.0, 1.0)  # touches 4 of the 16 level-2 patches


def test_uniform_swe_bitwise(dev):
    sim = make(H.SOLVER_SWE_ROE, 16, 2, 2)
    sim.initialize("radial_dam", DAM, nsub=4)
    amr = oracle_from(sim, O.SOLVER_SWE_ROE)
    dt = 0.004
    for s in range(4):
        c_gpu = sim.advance(s * dt, dt)
        c_orc = amr.advance(s * dt, dt)
        assert c_gpu == c_orc
    got = sim.get_q()
    assert np.array_equal(interiors(got), interiors(amr.q))
    # ghost cells after the final ghost update match too
    assert np.array_equal(got, np.stack(amr.q))
    assert sim.cell_updates() == 4 * 16 * 16 * 16


@pytest.mark.parametrize("subcycle", [True, False])
def test_amr_swe_subcycled_bitwise(dev, subcycle):
    sim = make(H.SOLVER_SWE_ROE, 8, 2, 4, refine=0.5, subcycle=subcycle)
    sim.initialize("radial_dam", SMALL_DAM, nsub=4)
    lev = sim.patch_info()["level"]
    assert lev.min() == 2 and lev.max() == 4, f"test needs three levels, got {sorted(set(lev))}"
    amr = oracle_from(sim, O.SOLVER_SWE_ROE)
    dt = 0.01
    for s in range(3):
        c_gpu = sim.advance(s * dt, dt)
        c_orc = amr.advance(s * dt, dt)
        assert c_gpu == c_orc
    assert np.array_equal(interiors(sim.get_q()), interiors(amr.q))
    assert sim.cell_updates() == amr.cell_updates


@pytest.mark.parametrize("subcycle", [True, False])
def test_coarsest_level_above_minlevel_bitwise(dev, subcycle):
    """Options minlevel 1 / maxlevel 4 on a mesh of levels 2-3 only: dt is the
    step of level 1, so level 2 takes two steps per coarse step with a fully
    synchronised ghost update between them (subcycling), or every level takes
    2^(4-1) steps (no subcycling), as the reference's counters give."""
    sim = make(H.SOLVER_SWE_ROE, 8, 1, 4, refine=0.5, subcycle=subcycle)
    lv, gx, gy = [], [], []
    for y in range(4):
        for x in range(4):
            if x < 2 and y < 2:
                for c in range(4):
                    lv.append(3), gx.append(2 * x + (c & 1)), gy.append(2 * y + (c >> 1))
            else:
                lv.append(2), gx.append(x), gy.append(y)
    sim.init_leaves(lv, gx, gy)
    sim.qinit("radial_dam", SMALL_DAM, nsub=4)
    sim.ghost_update(2, 3)
    amr = oracle_from(sim, O.SOLVER_SWE_ROE)
    dt = 0.02
    for s in range(2):
        assert sim.advance(s * dt, dt) == amr.advance(s * dt, dt)
    assert np.array_equal(interiors(sim.get_q()), interiors(amr.q))
    per_coarse = (8 * 8) * ((12 * 2 + 16 * 4) if subcycle else (28 * 8))
    assert sim.cell_updates() == amr.cell_updates == 2 * per_coarse


@pytest.mark.parametrize("solver", [(H.SOLVER_EULER5_ROE, O.SOLVER_EULER5_ROE),
                                    (H.SOLVER_EULER_HLLE, O.SOLVER_EULER_HLLE)])
def test_euler_brick_bitwise(dev, solver):
    g = 1.4
    pinf = 1.5698
    rinf = ((g - 1) + (g + 1) * pinf) / ((g + 1) + (g - 1) * pinf)
    vinf = (1 / np.sqrt(g)) * (pinf - 1) / np.sqrt(0.5 * ((g + 1) / g) * pinf + 0.5 * (g - 1) / g)
    sim = make(solver[0], 16, 1, 2, mi=4, mj=1, box=(0, 2, 0, 0.5), refine=2.0)
    sim.initialize("shock_bubble", (g, 0.2, rinf, vinf, pinf, 0.5, 0.25, 0.2, 0.138), nsub=4)
    amr = oracle_from(sim, solver[1])
    dt = 0.002
    for s in range(2):
        assert sim.advance(s * dt, dt) == amr.advance(s * dt, dt)
    assert np.array_equal(interiors(sim.get_q()), interiors(amr.q))


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
@pytest.mark.parametrize("merge", [True, False])
def test_swirl_fused_velocities_bitwise(dev, precision, merge):
    """The step kernel's own swirl velocities (HIPCLAW_FEATURE_FUSE_AUX) give
    the bits of the velocity-field kernel followed by the step, on a
    subcycled two-level mesh with non-square cells (FP64).  FP32 evaluates
    the separable stream function with FP32 factors instead of rounding the
    FP64 field: there the two agree to FP32 rounding (relative L1 1e-6)."""
    runs = []
    for fuse in (False, True):
        sim = make(H.SOLVER_ADV, 16, 2, 3, precision=precision, box=(0.0, 1.0, 0.0, 0.75))
        sim.set_feature(H.FEATURE_MERGE_LEVELS, merge)
        sim.set_feature(H.FEATURE_FUSE_AUX, fuse)
        sim.set_swirl(8.0)
        sim.initialize("swirl", (0.5,))
        dt = 0.004
        cfl = [sim.advance(s * dt, dt) for s in range(4)]
        runs.append((sim.get_q(), cfl, sim.num_patches()))
        sim.close()
    assert runs[0][2] == runs[1][2] and runs[0][2] > 16
    if precision == H.FP64:
        assert runs[0][1] == runs[1][1]
        assert np.array_equal(interiors(runs[0][0]), interiors(runs[1][0]))
    else:
        np.testing.assert_allclose(runs[1][1], runs[0][1], rtol=1e-5)
        a, b = interiors(runs[1][0]), interiors(runs[0][0])
        assert np.abs(a - b).sum() <= 1e-6 * np.abs(b).sum()


def test_swirl_advection_device_velocities(dev):
    sim = make(H.SOLVER_ADV, 16, 2, 2)
    sim.set_swirl(8.0)
    sim.initialize("swirl", (0.5,))
    amr = oracle_from(sim, O.SOLVER_ADV, swirl=8.0)
    dt = 0.01
    for s in range(3):
        c_gpu = sim.advance(s * dt, dt)
        c_orc = amr.advance(s * dt, dt)
        # velocities use the device sin/cos, the oracle glibc's
        assert c_gpu == pytest.approx(c_orc, rel=1e-13)
    d = interiors(sim.get_q()) - interiors(amr.q)
    assert np.abs(d).max() <= 1e-12


def test_regrid_transfers_match_oracle_kernels(dev):
    sim = make(H.SOLVER_SWE_ROE, 8, 2, 4, refine=0.5, coarsen=0.2)
    sim.initialize("radial_dam", SMALL_DAM, nsub=4)
    for s in range(3):
        sim.advance(s * 0.01, 0.01)
    old_info = sim.patch_info()
    old_q = sim.get_q()
    changed = sim.regrid()
    assert changed
    new_info = sim.patch_info()
    new_q = sim.get_q()
    old = {(l, x, y): k for k, (l, x, y) in enumerate(zip(old_info["level"], old_info["gx"],
                                                            old_info["gy"]))}
    L = A._setup_sigs()
    mx = 8
    kinds = {"keep": 0, "refine": 0, "coarsen": 0}
    for k, (l, x, y) in enumerate(zip(new_info["level"], new_info["gx"], new_info["gy"])):
        if (l, x, y) in old:
            exp = old_q[old[(l, x, y)]]
            kinds["keep"] += 1
        elif (l - 1, x >> 1, y >> 1) in old:
            exp = np.zeros_like(new_q[k])
            par = old_q[old[(l - 1, x >> 1, y >> 1)]]
            L.orc_interpolate2fine(mx, mx, 2, 3, A._p(par), A._p(exp), (x & 1) + 2 * (y & 1))
            kinds["refine"] += 1
        else:
            exp = np.zeros_like(new_q[k])
            for ig in range(4):
                ch = old_q[old[(l + 1, 2 * x + (ig & 1), 2 * y + (ig >> 1))]]
                L.orc_average2coarse(mx, mx, 2, 3, A._p(exp), A._p(ch), ig)
            kinds["coarsen"] += 1
        assert np.array_equal(new_q[k][:, 2:-2, 2:-2], exp[:, 2:-2, 2:-2]), (k, l, x, y)
    assert kinds["keep"] > 0 and (kinds["refine"] > 0 or kinds["coarsen"] > 0)


def test_fp32_amr_error_vs_fp64_oracle(dev):
    sim = make(H.SOLVER_SWE_ROE, 8, 2, 4, precision=H.FP32, refine=0.5)
    sim.initialize("radial_dam", SMALL_DAM, nsub=4)
    amr = oracle_from(sim, O.SOLVER_SWE_ROE)  # starts from the FP32-rounded data
    dt = 0.01
    for s in range(3):
        c32 = sim.advance(s * dt, dt)
        c64 = amr.advance(s * dt, dt)
        assert c32 == pytest.approx(c64, rel=1e-5)
    a, b = interiors(sim.get_q()), interiors(amr.q)
    rel_l1 = np.abs(a - b).sum() / np.abs(b).sum()
    print(f"\nFP32 AMR rel L1 {rel_l1:.3e} rel Linf {np.abs(a - b).max() / np.abs(b).max():.3e}")
    assert rel_l1 <= FP32_REL_L1


def test_wall_boundaries_conserve_mass(dev):
    sim = make(H.SOLVER_SWE_ROE, 16, 2, 2, mthbc=(3, 3, 3, 3))
    sim.initialize("radial_dam", DAM, nsub=4)
    m0 = sim.mass()
    for s in range(10):
        sim.advance(s * 0.004, 0.004)
    m1 = sim.mass()
    assert abs(m1[0] - m0[0]) <= 1e-13 * abs(m0[0])


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_autotune_and_launch_shapes_bitwise(dev, precision):
    """Every launch shape gives the same bits (FP32 too: its step contracts
    within source expressions only, so every instantiation rounds alike); the
    autotuner keeps one of them."""
    sim = make(H.SOLVER_SWE_ROE, 16, 2, 3, refine=0.5, precision=precision)
    sim.initialize("radial_dam", SMALL_DAM, nsub=4)
    q0 = sim.get_q()
    ref = None
    for cpl in (1, 2):
        for wpg in (1, 2, 4):
            sim.set_q(q0)
            sim.set_launch(cpl, wpg, 0)
            cfl = [sim.advance(s * 0.005, 0.005) for s in range(2)]
            got = (cfl, interiors(sim.get_q()))
            if ref is None:
                ref = got
            assert got[0] == ref[0] and np.array_equal(got[1], ref[1]), (cpl, wpg)
    best, sweep = sim.autotune(-1, 2)
    assert best in sweep and len(sweep) == 6 and all(v > 0 for v in sweep.values())
    # the tuned shape applies to large launches; this mesh's small level
    # launches keep the library's banded shape
    assert sim.get_launch() == best + (-1,)
    np_small = int(sim.num_patches())
    assert H.launch_shape(sim.phys, np_small, sim.get_launch()) == H.launch_shape(sim.phys, np_small)
    assert H.launch_shape(sim.phys, np_small)[2] >= 2
    sim.set_q(q0)
    cfl = [sim.advance(s * 0.005, 0.005) for s in range(2)]
    assert cfl == ref[0] and np.array_equal(interiors(sim.get_q()), ref[1])
    sim.set_launch(None)


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
@pytest.mark.parametrize("case", ["swe", "swe_push", "swe_time_sync", "swirl"])
@pytest.mark.parametrize("mx", [8, 10, 16])
def test_row_bands_change_no_bits(dev, precision, case, mx):
    """Patches swept in 2, 3, 4 or 8 row bands by the waves of one workgroup give
    the bits of whole-patch sweeps: interiors, per-patch Courant numbers and,
    with time sync, the corrections from the flux registers; with the
    same-level ghost push (swe_push, swirl), the pushed ghost cells (the last band
    of a 10-row patch holds one row; 4 bands of an 8-row patch hold 2 each).
    Each banded run is compared with the unbanded one of the same columns per
    lane (in FP32 the two instantiations contract differently); with two
    columns per lane the library sweeps whole patches whatever row_bands says."""
    runs = []
    for cpl in (1, 2):
        for bands in (0, 2, 3, 4, 8):
            if case == "swirl":
                sim = make(H.SOLVER_ADV, mx, 2, 3, precision=precision, box=(0.0, 1.0, 0.0, 0.75))
                sim.set_swirl(8.0)
                sim.initialize("swirl", (0.5,))
                dt = 0.004
            else:
                opt = F.Options(minlevel=2, maxlevel=3, mx=mx, my=mx, solver=H.SOLVER_SWE_ROE,
                                precision=precision, refine_threshold=0.5, coarsen_threshold=0.05,
                                time_sync=case == "swe_time_sync")
                sim = F.Hipclaw46(opt)
                sim.initialize("radial_dam", SMALL_DAM, nsub=4)
                dt = 0.005
            sim.set_feature(H.FEATURE_GHOST_PUSH, case in ("swe_push", "swirl"))
            sim.set_launch(cpl, 1, bands)
            cfl = [sim.advance(s * dt, dt) for s in range(3)]
            runs.append(((cpl, bands), cfl, sim.get_q(), sim.num_patches()))
            sim.close()
    for shape, cfl, q, n in runs:
        ref = next(r for r in runs if r[0] == (shape[0], 0))
        assert ref[3] > 16
        assert n == ref[3] and cfl == ref[1], shape
        # whole tiles: interiors and every ghost cell after the final ghost
        # update, pushed ghost cells included
        assert np.array_equal(q, ref[2]), shape


@pytest.mark.parametrize("precision,launch", [(H.FP64, None), (H.FP32, None), (H.FP32, (1, 2, 0)),
                                              (H.FP64, (1, 1, 0))])
def test_merged_launch_and_ghost_push_change_no_bits(dev, precision, launch):
    """hipclaw_advance with one launch per subcycled time and the same-level
    ghost push gives the same bits as per-level launches with ghost copies."""
    res = []
    for on in (True, False):
        sim = make(H.SOLVER_SWE_ROE, 8, 2, 4, precision=precision, refine=0.5)
        sim.initialize("radial_dam", SMALL_DAM, nsub=4)
        if launch is not None:
            sim.set_launch(*launch)
        sim.set_feature(H.FEATURE_MERGE_LEVELS, on)
        sim.set_feature(H.FEATURE_GHOST_PUSH, on)
        cfl = [sim.advance(s * 0.01, 0.01) for s in range(3)]
        res.append((cfl, sim.get_q()))
    assert res[0][0] == res[1][0]
    assert np.array_equal(res[0][1], res[1][1])


@pytest.mark.parametrize("ts", [False, True])
def test_graph_replay_matches_eager(dev, ts):
    """Coarse steps captured as hipGraphs and replayed (two pool-parity
    states alternate, each replayed several times, with eager work and host
    synchronisation in between) give the eager bits after every step."""
    import torch
    stream = torch.cuda.Stream()
    sims = []
    for graph in (False, True):
        opt = F.Options(minlevel=2, maxlevel=4, mx=8, my=8, solver=H.SOLVER_SWE_ROE,
                        refine_threshold=0.5, coarsen_threshold=0.05, time_sync=ts)
        sim = F.Hipclaw46(opt, stream=stream.cuda_stream)
        sim.initialize("radial_dam", SMALL_DAM, nsub=4)
        sim.set_feature(H.FEATURE_GRAPH, graph)
        sim.set_feature(H.FEATURE_GHOST_PUSH, True)  # default only from 32x32 patches
        sims.append(sim)
    side = torch.ones(1 << 20, device=dev)
    for s in range(7):
        c = [sim.advance(s * 0.005, 0.005) for sim in sims]
        with torch.cuda.stream(stream):
            side.mul_(1.0001)  # eager work on the same stream between replays
        assert c[0] == c[1], s
        assert np.array_equal(sims[0].get_q(), sims[1].get_q()), s
    assert sims[0].cell_updates() == sims[1].cell_updates()
    caps, reps = sims[1].graph_stats()
    # one capture per pool-parity state (two alternate); the rest replays
    assert caps == 2 and reps == 5, (caps, reps)
    assert sims[0].graph_stats() == (0, 0)


@pytest.mark.parametrize("case", ["swe", "swe_time_sync", "swirl", "euler_src2"])
def test_graph_replays_serve_every_dt(dev, case):
    """One captured coarse step per pool-parity state serves every (t, dt):
    with dt changing from step to step (as under CFL control), and for the
    swirl field whose velocities depend on t and dt, graph replays give the
    eager bits after every step."""
    import torch
    stream = torch.cuda.Stream()
    g = 1.4
    pinf = 1.5698
    rinf = ((g - 1) + (g + 1) * pinf) / ((g + 1) + (g - 1) * pinf)
    vinf = (1 / np.sqrt(g)) * (pinf - 1) / np.sqrt(0.5 * ((g + 1) / g) * pinf + 0.5 * (g - 1) / g)
    sims = []
    for graph in (False, True):
        if case == "swirl":
            opt = F.Options(minlevel=2, maxlevel=4, mx=8, my=8, solver=H.SOLVER_ADV,
                            refine_threshold=0.2, coarsen_threshold=0.05,
                            refinement_criteria="minmax")
            sim = F.Hipclaw46(opt, stream=stream.cuda_stream)
            sim.set_swirl(8.0)
            sim.initialize("swirl", (0.5,))
        elif case == "euler_src2":
            opt = F.Options(minlevel=1, maxlevel=3, mx=8, my=8, mi=4, mj=1, ax=0, bx=2, ay=0, by=0.5,
                            solver=H.SOLVER_EULER5_ROE, src2=H.SRC2_AXISYMMETRIC,
                            refine_threshold=0.1, coarsen_threshold=0.02, mthbc=(1, 1, 3, 1))
            sim = F.Hipclaw46(opt, stream=stream.cuda_stream)
            sim.initialize("shock_bubble", (g, 0.2, rinf, vinf, pinf, 0.5, 0.25, 0.2, 0.138), nsub=4)
        else:
            opt = F.Options(minlevel=2, maxlevel=4, mx=8, my=8, solver=H.SOLVER_SWE_ROE,
                            refine_threshold=0.5, coarsen_threshold=0.05,
                            time_sync=case == "swe_time_sync")
            sim = F.Hipclaw46(opt, stream=stream.cuda_stream)
            sim.initialize("radial_dam", SMALL_DAM, nsub=4)
        sim.set_feature(H.FEATURE_GRAPH, graph)
        sims.append(sim)
    lev = sims[0].patch_info()["level"]
    assert len(set(lev)) >= 2, f"{case}: needs two levels"
    base = 0.004 if case != "euler_src2" else 0.002
    t = 0.0
    for s, f in enumerate((1.0, 0.8, 0.93, 0.71, 1.0, 0.84, 0.77)):
        dt = base * f
        c = [sim.advance(t, dt) for sim in sims]
        t += dt
        assert c[0] == c[1], (case, s)
        assert np.array_equal(sims[0].get_q(), sims[1].get_q()), (case, s)
    caps, reps = sims[1].graph_stats()
    assert caps <= 2 and caps + reps == 7 and reps >= 5, (caps, reps)
    assert sims[0].graph_stats() == (0, 0)


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_mass_is_the_interior_sum(dev, precision):
    """hipclaw_mass (the conservation check of
    fclaw2d_clawpatch46_diagnostics.f:13-57): per field, the sum over every
    leaf's interior cells of q dx dy, against an exactly rounded sum
    (math.fsum) of the same products from the solution as stored.  A wrong
    cell area or ghost cells in the sum are off by far more than rounding."""
    import math
    sim = make(H.SOLVER_SWE_ROE, 8, 2, 4, precision=precision, refine=0.5)
    sim.initialize("radial_dam", SMALL_DAM, nsub=4)
    dt = 0.01
    for s in range(2):
        sim.advance(s * dt, dt)
    info = sim.patch_info()
    assert len(set(info["level"])) == 3
    q = sim.get_q()
    got = sim.mass()
    area = info["dx"] * info["dy"]
    for m in range(q.shape[1]):
        prod = q[:, m, 2:-2, 2:-2] * area[:, None, None]
        exact = math.fsum(prod.ravel())
        scale = math.fsum(np.abs(prod).ravel())
        assert abs(got[m] - exact) <= 1e-14 * scale, (m, got[m], exact)
    # the ghost band would change h's sum by far more than the tolerance
    ghost = math.fsum((q[:, 0] * area[:, None, None]).ravel()) - math.fsum(
        (q[:, 0, 2:-2, 2:-2] * area[:, None, None]).ravel())
    assert abs(ghost) > 1e6 * 1e-14 * abs(got[0])


def test_save_and_restore_step_on_the_device(dev):
    """hipclaw_save_step / hipclaw_restore_step: a restore after two coarse
    steps gives back the saved solution bit for bit (ghost cells included);
    restore without a save on the current mesh is an error."""
    sim = make(H.SOLVER_SWE_ROE, 8, 2, 4, refine=0.5)
    sim.initialize("radial_dam", SMALL_DAM, nsub=4)
    with pytest.raises(H.HipclawError):
        sim.restore_step()
    q0 = sim.get_q()
    sim.save_step()
    for s in range(2):
        sim.advance(s * 0.01, 0.01)
    assert not np.array_equal(sim.get_q(), q0)
    sim.restore_step()
    assert np.array_equal(sim.get_q(), q0)
    # the restored state steps like the original one
    c1 = sim.advance(0.0, 0.01)
    q1 = sim.get_q()
    sim.restore_step()
    assert sim.advance(0.0, 0.01) == c1 and np.array_equal(sim.get_q(), q1)


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_rejected_step_is_retaken_from_the_device_copy(dev, precision):
    """fclaw2d_run with variable dt: an initial dt whose Courant number exceeds
    max_cfl is rejected, the solution restored on the device and the step
    retaken at dt * desired_cfl / cfl.  The run ends bitwise equal to one
    started at that smaller dt, whose attempts are the first run's after its
    first rejection (later rejections, if the flow or a regrid raise the
    Courant number, happen in both)."""
    def run(dt0):
        opt = F.Options(minlevel=2, maxlevel=4, mx=8, my=8, solver=H.SOLVER_SWE_ROE,
                        precision=precision, refine_threshold=0.5, coarsen_threshold=0.05,
                        use_fixed_dt=False, max_cfl=1.0, desired_cfl=0.9, regrid_interval=2)
        sim = F.Hipclaw46(opt)
        sim.initialize("radial_dam", SMALL_DAM, nsub=4)
        sim.run(0.2, nout=1, dt0=dt0, max_steps=5)
        return sim
    a = run(0.05)
    assert a.retakes >= 1 and a.dt_history[0][1] > 1.0
    b = run(a.dt_history[1][0])
    assert b.retakes == a.retakes - 1
    assert b.dt_history == a.dt_history[1:]
    assert np.array_equal(a.get_q(), b.get_q())
    assert a.t == b.t


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
@pytest.mark.parametrize("case", ["swirl_uniform", "swe_time_sync"])
def test_advance_steps_equal_single_steps(dev, precision, case):
    """hipclaw_advance_steps (the fixed-dt steps between outputs in one call,
    one graph launch when graphs are on) gives the bits and the per-step
    Courant numbers of one hipclaw_advance per step, for the uniform swirl
    mesh (one fine step per coarse step, time-dependent velocities) and a
    subcycled SWE mesh with time sync; the run loop with use_fixed_dt takes
    its steps in batches and ends where step-by-step calls end."""
    import torch
    stream = torch.cuda.Stream()
    sims = []
    for mode in ("single", "eager", "graph"):
        if case == "swirl_uniform":
            opt = F.Options(minlevel=3, maxlevel=3, mx=8, my=8, solver=H.SOLVER_ADV, precision=precision)
            sim = F.Hipclaw46(opt, stream=stream.cuda_stream)
            sim.set_swirl(8.0)
            sim.initialize("swirl", (0.5,))
        else:
            opt = F.Options(minlevel=2, maxlevel=4, mx=8, my=8, solver=H.SOLVER_SWE_ROE, precision=precision,
                            refine_threshold=0.5, coarsen_threshold=0.05, time_sync=True)
            sim = F.Hipclaw46(opt, stream=stream.cuda_stream)
            sim.initialize("radial_dam", SMALL_DAM, nsub=4)
        sim.set_feature(H.FEATURE_GRAPH, mode == "graph")
        sims.append(sim)
    if case == "swe_time_sync":
        assert len(set(sims[0].patch_info()["level"])) >= 2
    dt = 0.004 if case != "swirl_uniform" else 0.9 / 64
    ts = [0.0]  # step start times, accumulated as the run loop does
    for _ in range(15):
        ts.append(ts[-1] + dt)
    c_single = [sims[0].advance(ts[k], dt) for k in range(7)]
    for sim in sims[1:]:
        c = np.concatenate([sim.advance_steps(ts[0], dt, 4), sim.advance_steps(ts[4], dt, 3)])
        assert np.array_equal(c, np.array(c_single)), case
        assert np.array_equal(sim.get_q(), sims[0].get_q()), case
    caps, reps = sims[2].graph_stats()
    assert caps == 2 and reps == 0, (caps, reps)
    # two more calls of 4 steps: the second starts in the pool-parity state
    # of the first and replays its graph
    c_single = [sims[0].advance(ts[k], dt) for k in range(7, 15)]
    for sim in sims[1:]:
        c = np.concatenate([sim.advance_steps(ts[7], dt, 4), sim.advance_steps(ts[11], dt, 4)])
        assert np.array_equal(c, np.array(c_single)), case
        assert np.array_equal(sim.get_q(), sims[0].get_q()), case
    caps, reps = sims[2].graph_stats()
    assert caps == 3 and reps == 1, (caps, reps)
    # the fixed-dt run loop: batched and one step at a time end equal
    runs = []
    for batch in (1, 5):
        opt = F.Options(minlevel=3, maxlevel=3, mx=8, my=8, solver=H.SOLVER_ADV, precision=precision,
                        use_fixed_dt=True)
        sim = F.Hipclaw46(opt, stream=stream.cuda_stream)
        sim.set_swirl(8.0)
        sim.initialize("swirl", (0.5,))
        sim.run(13 * 0.01, nout=2, dt0=0.01, batch=batch)
        runs.append((sim.get_q(), sim.t, list(sim.dt_history), sim.n_inner))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert runs[0][1:] == runs[1][1:]


def test_prepare_steps_captures_without_running(dev):
    """hipclaw_prepare_steps records the graph of the next advance_steps call
    and leaves solution, counters and pool state as they were; the call that
    follows replays it and ends bitwise where an eager run ends."""
    import torch
    stream = torch.cuda.Stream()
    sims = []
    for graph in (False, True):
        opt = F.Options(minlevel=2, maxlevel=4, mx=8, my=8, solver=H.SOLVER_SWE_ROE,
                        refine_threshold=0.5, coarsen_threshold=0.05, time_sync=True)
        sim = F.Hipclaw46(opt, stream=stream.cuda_stream)
        sim.initialize("radial_dam", SMALL_DAM, nsub=4)
        sim.set_feature(H.FEATURE_GRAPH, graph)
        sims.append(sim)
    q0 = sims[1].get_q()
    sims[1].prepare_steps(3)
    sims[1].prepare_steps(3)  # already captured: nothing new
    assert sims[1].graph_stats() == (1, 0)
    assert sims[1].cell_updates() == 0
    assert np.array_equal(sims[1].get_q(), q0)
    sims[0].prepare_steps(3)  # graphs off: no effect
    assert sims[0].graph_stats() == (0, 0)
    dt = 0.004
    c = [sim.advance_steps(0.0, dt, 3) for sim in sims]
    assert np.array_equal(c[0], c[1])
    assert np.array_equal(sims[0].get_q(), sims[1].get_q())
    assert sims[1].graph_stats() == (1, 1)
    assert sims[0].cell_updates() == sims[1].cell_updates() > 0
