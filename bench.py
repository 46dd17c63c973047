"""Benchmark: cell-updates/s of the wave-propagation patch solver on one MI355X.

Workload (shallow water, large uniform-plus-AMR case): unit square, one block,
32x32 patches (mbc 2), base level 7 uniform (16,384 patches = 16.8 M cells)
refined to level 9 around 8 seeded Gaussian humps of amplitude U(0.1, 0.5)
and width U(0.02, 0.1) on h = 1 at rest; Roe solver with entropy fix, MC
limiter, transverse corrections, extrapolation boundaries, subcycling.

A step is one coarse step (fclaw2d_advance_all_levels): every level update
at its subcycled times, time interpolation and all ghost updates.  Regrid is
not part of a step.  The headline value is the batched FP32 path (one step
launch per subcycled time for all due levels, same-level ghost cells pushed
by the step kernel, the fixed-dt coarse steps in one hipclaw_advance_steps
call replayed from a hipGraph).  The FP64 path runs the same steps from the
same data twice: batched like the FP32 run, and plain (one call and one
launch per level step, ghost copies by the ghost kernel, eager launches).
Both are bitwise equal to the CPU oracle in the parity tests, so the FP64
run is also the full-size oracle that gives the FP32 error.  The CPU
baseline times the reference's own Fortran patch step
(built into oracle/_ref) on a bounded sample of the same patches, one
process per host core, with the oracle's C restatement under OpenMP beside it.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s on 1 MI355X, FP32-batched vs FP64 plain, at <=1e-5 rel L1"
HBM_PEAK = 8.0e12

# HBM bytes per step launch (average over the 4 launches of a coarse step,
# ghost push on) of the default workload, whose mesh is fixed by the seed:
# 2 x FETCH_SIZE + 1 x WRITE_SIZE from two rocprofv3 --pmc runs, the factors
# being what the same counters read for a 1 GiB device copy
# (profiles/k_step2_swe_humps_l7-9_hbm_bytes.md).  VALU_ACTIVE is
# SQ_ACTIVE_INST_VALU x 4 cycles over the SIMD-cycles of the step launches of
# a coarse step (profiles/k_step2_swe_humps_l7-9_sq_counters.md).
MEASURED_MESH = {7: 10481, 8: 2349, 9: 85052}
TRAFFIC_PER_LAUNCH = 3.586e9
VALU_ACTIVE = 0.87
# FP32 VALU work of the step launches of a coarse step per cell update:
# wave-level VALU instructions, and FP32 operations issued (64 lanes x (ADD +
# MUL + 2 FMA + TRANS) wave instructions, SQ_INSTS_VALU_*_F32;
# profiles/k_step2_swe_humps_l7-9_valu_mix.md).
VALU_INSTS_PER_CELL = 11.86
FLOP_PER_CELL = 748.2
# MI355X FP32 vector peak without packed math: 256 CUs x 4 SIMDs x 16 lanes
# x 2 (FMA) x 2.4 GHz.  The step kernels are built without packed FP32.
FP32_VALU_PEAK = 78.6e12
# VALU issue peak: one wave64 instruction per SIMD every 4 cycles.
VALU_ISSUE_PEAK = 1024 * 2.4e9 / 4


_STREAM = []
_T0 = time.perf_counter()


def log(msg):
    """Progress line on stderr (the JSON result goes to stdout)."""
    print(f"[bench {time.perf_counter() - _T0:6.1f}s] {msg}", file=sys.stderr, flush=True)


def make_sim(precision, args, leaves=None, q=None):
    """Simulation on a dedicated (default-priority) stream: graph capture
    needs a stream other than the null stream."""
    import torch
    from forestclaw_amd import configs as CF
    from forestclaw_amd import forest as F
    if not _STREAM:
        _STREAM.append(torch.cuda.Stream())
    opt = CF.c5_options(precision, args.minlevel, args.maxlevel, args.refine_threshold)
    sim = F.Hipclaw46(opt, stream=_STREAM[0].cuda_stream)
    if leaves is None:
        sim.initialize("humps", CF.c5_humps(args.seed))
    else:
        sim.init_leaves(leaves["level"], leaves["gx"], leaves["gy"])
        sim.set_q(q)
    return sim


def run_steps(sim, nsteps, dt, t0=0.0):
    import torch
    t = t0
    for _ in range(nsteps):
        sim.advance(t, dt)
        t += dt
    torch.cuda.synchronize()
    return t


def timed(sim, args, dt, graph, plain=False):
    """Warm-up steps, then args.steps timed coarse steps, then the same number
    of eager steps with HIP events around every patch-step launch for the
    kernel time.  Batched: the fixed-dt steps go to the device as
    hipclaw_advance_steps calls of up to 256 steps (what the fixed-dt run
    loop does, Hipclaw46.run), replayed from a hipGraph when graph is set; the
    graph of the timed call is captured after the warm-up, untimed
    (hipclaw_prepare_steps).  plain: the unbatched path (one call and one
    step launch per level and level step, same-level ghost cells copied by
    the ghost kernel, eager launches, the Courant number read back after
    every coarse step)."""
    import torch
    from forestclaw_amd import hipclaw as H
    sim.set_feature(H.FEATURE_GRAPH, graph and not plain)
    sim.set_feature(H.FEATURE_GHOST_PUSH, not (args.no_push or plain))
    sim.set_feature(H.FEATURE_MERGE_LEVELS, not (args.no_merge or plain))
    chunk = min(args.steps, 256)
    calls = [chunk] * (args.steps // chunk) + ([args.steps % chunk] if args.steps % chunk else [])
    t = 0.0
    if plain:
        t = run_steps(sim, args.warmup, dt)
    elif args.warmup > 0:
        sim.advance_steps(t, dt, args.warmup)
        for _ in range(args.warmup):
            t += dt
    if not plain:
        sim.prepare_steps(chunk)
    c0 = sim.cell_updates()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if plain:
        t = run_steps(sim, args.steps, dt, t0=t)
    else:
        for k in calls:
            sim.advance_steps(t, dt, k)
            for _ in range(k):
                t += dt
        torch.cuda.synchronize()
    t1 = time.perf_counter()
    cells = sim.cell_updates() - c0
    caps, reps = sim.graph_stats()
    # kernel timing pass (eager, events on the context's stream)
    sim.profile(True)
    t = run_steps(sim, args.steps, dt, t0=t)
    kms, launches, kcells = sim.profile_read()
    sim.profile(False)
    return dict(seconds=t1 - t0, cells=cells, kernel_ms=kms, launches=launches, kcells=kcells,
                t_end=t, graph_captures=caps, graph_replays=reps)


REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_swe.so")


def _ref_worker(args):
    """One process: the reference's own Fortran patch step (CLAWPACK46_STEP2_WRAP
    of src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_step2_wrap.f, built
    unchanged from the reference sources into oracle/_ref) over its share of
    the sample until the time budget is spent.  Separate processes, since the
    Fortran solvers keep their state in common blocks."""
    import ctypes
    path, tiles, dx, dy, dt, budget = args
    L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    ci, cd = ctypes.c_int, ctypes.c_double
    cd.in_dll(L, "cparam_").value = 1.0
    mx = my = 32
    mbc, meqn, mwaves, maux = 2, 3, 3, 0
    maxm = 32
    mwork = (maxm + 2 * mbc) * (12 * meqn + (meqn + 1) * mwaves + 3 * maux + 2)
    work = np.zeros(mwork)
    size = meqn * (mx + 2 * mbc) * (my + 2 * mbc)
    fp, fm, gp, gm = (np.zeros(size) for _ in range(4))
    method = (ci * 7)(0, 2, 2, 0, 0, 0, 0)
    mthlim = (ci * 3)(4, 4, 4)
    bcc = (ci * 4)(0, 0, 0, 0)
    rpn2 = ctypes.cast(L.clawpack46_rpn2_, ctypes.c_void_p)
    rpt2 = ctypes.cast(L.clawpack46_rpt2_, ctypes.c_void_p)
    flux2 = ctypes.cast(L.clawpack46_flux2_, ctypes.c_void_p)
    aux = np.zeros(1)
    cfl, ierr = cd(0.0), ci(0)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    cells = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget:
        for k in range(tiles.shape[0]):
            q = tiles[k].copy()
            L.clawpack46_set_block_(ctypes.byref(ci(0)))
            L.clawpack46_step2_wrap_(
                ctypes.byref(ci(maxm)), ctypes.byref(ci(meqn)), ctypes.byref(ci(maux)),
                ctypes.byref(ci(mbc)), method, mthlim, ctypes.byref(ci(0)), ctypes.byref(ci(mwaves)),
                ctypes.byref(ci(mx)), ctypes.byref(ci(my)), vp(q), vp(aux), ctypes.byref(cd(dx)),
                ctypes.byref(cd(dy)), ctypes.byref(cd(dt)), ctypes.byref(cfl), vp(work),
                ctypes.byref(ci(mwork)), ctypes.byref(cd(0.0)), ctypes.byref(cd(0.0)),
                ctypes.byref(ci(0)), ctypes.byref(cd(0.0)), vp(fp), vp(fm), vp(gp), vp(gm), rpn2,
                rpt2, rpn2, rpt2, flux2, bcc, ctypes.byref(ierr), ctypes.byref(ci(0)))
            L.clawpack46_unset_block_()
            cells += mx * my
    return cells, time.perf_counter() - t0


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(sim64, budget_s):
    """CPU rate of the same patch step on the GPU box's host cores, on a sample
    of level-7 patches of the workload.  With oracle/_ref built: the reference
    Fortran step in one process per core ("reference"); the oracle port with
    OpenMP over patches is reported beside it (or alone when the reference
    library is absent, as "port")."""
    from oracle import oracle as O
    info = sim64.patch_info()
    lev = info["level"]
    q = sim64.get_q()
    idx = np.nonzero(lev == lev.min())[0][:256]
    sample = np.ascontiguousarray(q[idx])
    del q
    cfg = O.make_cfg(O.SOLVER_SWE_ROE, 32, 32, 2, grav=1.0)
    dx, dy = info["dx"][idx], info["dy"][idx]
    dt = 0.5 * dx.min()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cells = 0
    t0 = time.perf_counter()
    reps = 0
    while True:
        work = sample.copy()
        O.step2_batch(cfg, work, None, dx, dy, dt)
        cells += len(idx) * 32 * 32
        reps += 1
        if time.perf_counter() - t0 > budget_s / 2:
            break
    el = time.perf_counter() - t0
    port = dict(value=cells / el, unit="cell-updates/s", cores=threads, kind="port",
                sample=f"{len(idx)} level-7 patches (32x32, meqn 3) x {reps} patch-steps each, "
                       f"oracle step2 FP64 (C restatement), OpenMP over patches, {el:.1f} s")
    if not os.path.exists(REF_LIB):
        return port, None
    import multiprocessing as mp
    nproc = threads
    shares = np.array_split(sample, nproc)
    args = [(REF_LIB, sh, float(dx[0]), float(dy[0]), dt, budget_s / 2) for sh in shares if len(sh)]
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(len(args)) as pool:
        res = pool.map(_ref_worker, args)
        # the same step on one core, over the whole sample
        one = pool.apply(_ref_worker, ((REF_LIB, sample, float(dx[0]), float(dy[0]), dt, budget_s / 4),))
    wall = time.perf_counter() - t0
    rcells = sum(r[0] for r in res)
    busy = max(r[1] for r in res)
    ref = dict(value=rcells / busy, unit="cell-updates/s", cores=len(args), kind="reference",
               sample=f"{len(idx)} level-7 patches (32x32, meqn 3) split over {len(args)} processes, "
                      f"the reference Fortran CLAWPACK46_STEP2_WRAP (oracle/_ref, amdflang -O2) "
                      f"repeated for {busy:.1f} s per process ({wall:.1f} s wall with start-up)",
               one_core={"value": one[0] / one[1], "unit": "cell-updates/s", "cores": 1,
                         "seconds": one[1]},
               cpu_model=_cpu_model(), host_threads=os.cpu_count())
    return ref, port


def host_path(sim64, n, dt):
    """PCIe-inclusive rate of the host-resident drop-in path: the deferred
    hipclaw_single_step_update protocol over n host tiles of the finest level
    (FP64 arrays, as ForestClaw stores them; one Python call per patch, a
    batch of up to 4096 patches per flush), FP32 and FP64 device modes."""
    import torch
    from forestclaw_amd import hipclaw as H
    info = sim64.patch_info()
    lev = info["level"]
    idx = np.nonzero(lev == lev.max())[0][:n]
    q = sim64.get_q()
    tiles = [np.ascontiguousarray(q[i]) for i in idx]
    del q
    dx, dy = float(info["dx"][idx[0]]), float(info["dy"][idx[0]])
    dtf = dt / (1 << int(lev.max() - lev.min()))
    out = {}
    for name, prec in (("fp32", H.FP32), ("fp64", H.FP64)):
        b = H.Batch(H.physics(H.SOLVER_SWE_ROE, 32, 32, precision=prec, grav=1.0),
                    min(len(tiles), 4096), _STREAM[0].cuda_stream)
        rates = []
        for rep in range(3):  # the first pass warms up
            t0 = time.perf_counter()
            for it, tq in enumerate(tiles):
                b.single_step_update(tq, None, dx, dy, rep * dtf, dtf, it, len(tiles))
            torch.cuda.synchronize()
            rates.append(len(tiles) * 32 * 32 / (time.perf_counter() - t0))
        b.close()
        out[name] = max(rates[1:])
    return {"value_fp32": out["fp32"], "value_fp64": out["fp64"], "unit": "cell-updates/s",
            "sample": f"{len(tiles)} level-{int(lev.max())} host tiles (36x36x3 float64) per pass, "
                      "hipclaw_single_step_update with batches of up to 4096, best of 2 passes"}


# algorithmic FLOPs per 36x36 tile of the refinement network (3x3 convolutions
# 8->16->32->32->16): forward 2*1296*9*sum(Cin*Cout); training adds the data
# gradient of layers 1-3 and the weight gradient of all four
CNN_FWD_FLOP = 2 * 1296 * 9 * (8 * 16 + 16 * 32 + 32 * 32 + 32 * 16)
CNN_TRAIN_FLOP = CNN_FWD_FLOP + 2 * 1296 * 9 * (16 * 32 + 32 * 32 + 32 * 16) + CNN_FWD_FLOP
BF16_PEAK = 2.5e15


def small_configs(args):
    """The other configurations (forestclaw_amd.configs C1 swirl, C2 radial
    dam, C3 shock-bubble) on their meshes after the initial refinement, no
    regrid, fixed dt: FP32 batched (the library defaults: merged launches,
    ghost push, graph replays; the steps go to the device in calls of up to
    64, as the fixed-dt run loop takes them, Hipclaw46.run) against FP64
    plain (one call and one launch per level step, eager), each timed over
    its stated run length after untimed warm-up calls of the same sizes."""
    import torch
    from forestclaw_amd import configs as CF
    from forestclaw_amd import hipclaw as H
    lengths = {"C1": 200, "C2": 100, "C3": 30}
    if getattr(args, "config", None):
        lengths = {k: v for k, v in lengths.items() if k in args.config}
    chunk = 64
    if not _STREAM:
        _STREAM.append(torch.cuda.Stream())
    out = {}
    for name, nsteps in lengths.items():
        res = {}
        for label, prec, plain in (("fp32_batched", H.FP32, False), ("fp64_plain", H.FP64, True)):
            sim, dt = CF.build(name, prec, stream=_STREAM[0].cuda_stream)
            if plain:  # batched keeps the defaults
                sim.set_feature(H.FEATURE_GRAPH, False)
                sim.set_feature(H.FEATURE_GHOST_PUSH, False)
                sim.set_feature(H.FEATURE_MERGE_LEVELS, False)
            calls = [1] * nsteps if plain else \
                [chunk] * (nsteps // chunk) + ([nsteps % chunk] if nsteps % chunk else [])
            t = 0.0

            def go(k, t):
                if plain:
                    sim.advance(t, dt)
                else:
                    sim.advance_steps(t, dt, k)
                for _ in range(k):
                    t += dt
                return t
            for k in sorted(set(calls)) if not plain else [1, 1]:
                t = go(k, t)
            c0 = sim.cell_updates()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for k in calls:
                t = go(k, t)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            res[label] = {"cell_updates_per_s": (sim.cell_updates() - c0) / (t1 - t0),
                          "ms_per_step": 1e3 * (t1 - t0) / nsteps, "patches": sim.num_patches(),
                          "steps_per_call": max(calls), "graph_replays": sim.graph_stats()[1]}
            sim.close()
        res["speedup"] = res["fp32_batched"]["cell_updates_per_s"] / res["fp64_plain"]["cell_updates_per_s"]
        res["steps"] = nsteps
        out[name] = res
    return out


def advection(level, steps=10, warmup=3, fixed=False):
    """Scalar advection with the swirl velocity field (period 8) on a uniform
    level of 32x32 patches, FP32 and FP64: coarse steps with the velocities of
    each step (b4step2), and with fixed velocities when fixed is set.
    The step kernel's rate is HIP-event time of its launches; algorithmic
    bytes per patch: read q of the 36x36 tile (and the two velocity fields
    when they are read from aux: fixed velocities; with per-step velocities
    the step kernel evaluates them itself), write the 32x32 interior."""
    import torch
    from forestclaw_amd import forest as F
    from forestclaw_amd import hipclaw as H
    out = {}
    for prec, name, w in [(H.FP32, "fp32", 4), (H.FP64, "fp64", 8)]:
        for tperiod in ((8.0, 0.0) if fixed else (8.0,)):
            opt = F.Options(minlevel=level, maxlevel=level, mx=32, my=32, solver=H.SOLVER_ADV, precision=prec)
            sim = F.Hipclaw46(opt)
            sim.set_swirl(8.0)
            sim.init_uniform(level)
            sim.qinit("swirl", (0.5,), nsub=1)
            sim.ghost_update(level, level)
            dt = 0.5 / (32 << level)
            if tperiod == 0.0:
                # fixed velocities: store the field of t = 0 in aux with one
                # unfused step, then keep it
                sim.set_feature(H.FEATURE_FUSE_AUX, False)
                sim.advance(0.0, dt)
                sim.set_feature(H.FEATURE_FUSE_AUX, True)
            sim.set_swirl(tperiod)
            shape, sweep = sim.autotune(level, 3)
            for s in range(warmup):
                sim.advance(s * dt, dt)
            torch.cuda.synchronize()
            sim.profile(True)
            t0 = time.perf_counter()
            for s in range(steps):
                sim.advance((warmup + s) * dt, dt)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            ms, nl, _ = sim.profile_read()
            sim.profile(False)
            npatch = sim.num_patches()
            algo = w * (36 * 36 * (1 if tperiod else 3) + 32 * 32) * npatch
            key = f"{name}_{'swirl' if tperiod else 'fixed_velocity'}"
            out[key] = {"patches": npatch, "ms_per_step": 1e3 * (t1 - t0) / steps,
                        "cell_updates_per_s": npatch * 1024 * steps / (t1 - t0),
                        "step_kernel_ms": ms / nl, "step_kernel_GBs": algo / (ms / nl * 1e-3) / 1e9,
                        "step_kernel_hbm_frac": algo / (ms / nl * 1e-3) / HBM_PEAK,
                        "launch": {"cols_per_lane": shape[0], "waves_per_group": shape[1],
                                   "sweep_ms": {f"cpl{c}_wpg{w}": round(v, 4) for (c, w), v in sweep.items()}}}
            sim.close()
    return out


def cnn_c4(nruns, batch=4096):
    """The learned-indicator configuration at its stated size: tiles of
    seeded dam-break runs (random centre, radius 0.3-0.8, h ratio 1.5-4) on
    uniform levels 4-6, labelled with the plain tags; one epoch of Adam at
    batch 4096 and one inference pass, both timed (BF16 MFMA kernels)."""
    import torch
    from forestclaw_amd import hipclaw as H
    from forestclaw_amd import learned_tagging as LT
    t0 = time.perf_counter()
    data = LT.generate(nruns=nruns, levels=(4, 5, 6), samples=8, lookahead=0, neg_keep=1.0, seed=11)
    torch.cuda.synchronize()
    tgen = time.perf_counter() - t0
    n = len(data)
    net = H.Cnn(seed=5)
    loss = torch.zeros(1, device="cuda")
    perm = torch.randperm(n, device="cuda", dtype=torch.int64).to(torch.int32)
    # one untimed batch (work buffers), then the timed epoch
    net.train_step(data.tiles, data.labels, perm[:batch], batch, 2e-3, loss)
    torch.cuda.synchronize()
    nb = n // batch
    t0 = time.perf_counter()
    for b in range(nb):
        net.train_step(data.tiles, data.labels, perm[b * batch:(b + 1) * batch], batch, 2e-3, loss)
    torch.cuda.synchronize()
    ttrain = time.perf_counter() - t0
    logits = torch.empty((n, 2), dtype=torch.float32, device="cuda")
    ib = 16384
    ids = torch.arange(0, n, dtype=torch.int32, device="cuda")
    net.forward(data.tiles, logits, idx=ids[:ib], n=ib)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for b0 in range(0, n, ib):
        m = min(ib, n - b0)
        net.forward(data.tiles, logits[b0:], idx=ids[b0:b0 + m], n=m)
    torch.cuda.synchronize()
    tinf = time.perf_counter() - t0
    pred = (logits[:, 1] > logits[:, 0]).to(torch.int32)
    acc = float((pred == data.labels).float().mean())
    train_rate = nb * batch / ttrain
    inf_rate = n / tinf
    net.close()
    return {"tiles": n, "positive": int(data.labels.sum()), "batch": batch,
            "generate_s": tgen, "train_tiles_per_s": train_rate, "infer_tiles_per_s": inf_rate,
            "train_tflops": train_rate * CNN_TRAIN_FLOP / 1e12,
            "infer_tflops": inf_rate * CNN_FWD_FLOP / 1e12,
            "mfma_peak_tflops": BF16_PEAK / 1e12,
            "train_frac": train_rate * CNN_TRAIN_FLOP / BF16_PEAK,
            "final_batch_loss": float(loss.item()), "train_accuracy_after_1_epoch": acc,
            "sample": f"{nruns} dam-break runs x levels 4-6 x 8 samples, all tiles kept"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--minlevel", type=int, default=7)
    ap.add_argument("--maxlevel", type=int, default=9)
    ap.add_argument("--refine-threshold", type=float, default=0.5)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="leave out the C1-C3 timings (key \"configs\" of the default line)")
    ap.add_argument("--no-graph", action="store_true", help="eager launches instead of hipGraph replay")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run 3 plain device copies of 1 GiB (float32), the reference "
                         "kernel for scaling hardware byte counters")
    ap.add_argument("--no-push", action="store_true",
                    help="same-level ghost copies by the ghost kernel instead of the step kernel")
    ap.add_argument("--no-merge", action="store_true",
                    help="one step launch per level instead of one per subcycled time")
    ap.add_argument("--cnn", type=int, default=0, metavar="NRUNS",
                    help="also time the refinement network on the learned-indicator "
                         "configuration built from NRUNS dam-break runs (50: about 2M tiles)")
    ap.add_argument("--advection", type=int, default=0, metavar="LEVEL",
                    help="only time scalar advection with the swirl field on a uniform level "
                         "(one JSON line of its own)")
    ap.add_argument("--configs", action="store_true",
                    help="only time configurations C1-C3, FP32 batched against FP64 plain "
                         "(one JSON line of its own)")
    ap.add_argument("--config", action="append", choices=["C1", "C2", "C3"],
                    help="with --configs: only this configuration (repeatable)")
    ap.add_argument("--advection-fixed", action="store_true",
                    help="with --advection: also time fixed velocities read from aux")
    ap.add_argument("--host-path", type=int, default=0, metavar="NPATCH",
                    help="also time the host-resident drop-in path (hipclaw_single_step_update "
                         "on NPATCH host tiles, PCIe included)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU path is "
              "a design in DESIGN.md, not built.", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.configs:
        print(json.dumps({"workload": "configurations C1-C3", "configs": small_configs(args)}))
        return
    if args.advection > 0:
        print(json.dumps({"workload": f"scalar advection, swirl field, uniform level {args.advection}, 32x32",
                          "advection": advection(args.advection, args.steps, args.warmup,
                                                 args.advection_fixed)}))
        return
    from forestclaw_amd import hipclaw as H
    assert torch.cuda.is_available(), "no GPU visible"

    # The mesh is built once, by the FP64 run (tagging sees FP64 data, as in
    # the reference); the FP32 run starts from the same leaves and the same
    # initial data rounded to FP32.
    log("building the mesh (FP64 initial refinement)")
    sim64 = make_sim(H.FP64, args)
    info = sim64.patch_info()
    q0 = sim64.get_q()
    sim32 = make_sim(H.FP32, args, leaves=info, q=q0)
    del q0
    info32 = sim32.patch_info()
    for key in ("level", "gx", "gy"):
        assert np.array_equal(info32[key], info[key]), "FP32 run is not on the FP64 mesh"
    levels = {int(L): int((info["level"] == L).sum()) for L in np.unique(info["level"])}
    from forestclaw_amd import configs as CF
    dt = CF.c5_dt(float(info["dx"].max()))  # CFL 0.9 at the coarsest level (subcycled)

    # launch shape of the patch-step kernel picked by the autotuner on the
    # level with the most patches (timed into a scratch buffer)
    log(f"mesh {levels}; autotuning")
    tuned = {}
    for name, sim in (("fp32", sim32), ("fp64", sim64)):
        best, sweep = sim.autotune(-1, 3)
        tuned[name] = {"cols_per_lane": best[0], "waves_per_group": best[1],
                       "sweep_ms": {f"cpl{c}_wpg{w}": round(v, 4) for (c, w), v in sweep.items()}}

    log("timing FP32 batched")
    r32 = timed(sim32, args, dt, not args.no_graph)
    r64 = r64p = None
    err = None
    if not args.no_fp64:
        log("timing FP64 batched")
        r64 = timed(sim64, args, dt, not args.no_graph)
        a = sim32.get_q()[:, :, 2:-2, 2:-2]
        b = sim64.get_q()[:, :, 2:-2, 2:-2]
        err = dict(rel_l1=float(np.abs(a - b).sum() / np.abs(b).sum()),
                   rel_linf=float(np.abs(a - b).max() / np.abs(b).max()),
                   all_finite=bool(np.isfinite(a).all()))
        del a, b
        # the plain FP64 path, with the launch shape tuned for its own kernel
        sim64.set_feature(H.FEATURE_GHOST_PUSH, False)
        sim64.set_feature(H.FEATURE_MERGE_LEVELS, False)
        best, sweep = sim64.autotune(-1, 3)
        tuned["fp64_plain"] = {"cols_per_lane": best[0], "waves_per_group": best[1],
                               "sweep_ms": {f"cpl{c}_wpg{w}": round(v, 4) for (c, w), v in sweep.items()}}
        log("timing FP64 plain")
        r64p = timed(sim64, args, dt, False, plain=True)

    cpu = port = None
    if not args.no_cpu:
        log("CPU baseline")
        cpu, port = cpu_baseline(sim64, args.cpu_seconds)
    hp = host_path(sim64, args.host_path, dt) if args.host_path > 0 else None
    small = None
    if not args.no_configs:
        log("configurations C1-C3")
        small = small_configs(args)

    w = 4
    T, N, meqn = 36 * 36, 32 * 32, 3
    b_upd = w * meqn * (T + N) / N  # algorithmic bytes per cell update (read tile, write interior)
    k_time = r32["kernel_ms"] / 1e3
    k_rate = r32["kcells"] / k_time if k_time > 0 else 0.0  # cell updates per second of step kernel
    hbm_achieved = b_upd * k_rate
    measured = levels == MEASURED_MESH
    flops = FLOP_PER_CELL * k_rate if (FLOP_PER_CELL is not None and measured) else None
    value = r32["cells"] / r32["seconds"]
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r32["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian humps, h=1 at rest)",
        "config": {"workload": "SWE Roe, unit square, 32x32 patches, base level 7 uniform "
                               "refined to level 9 around 8 seeded humps, subcycled",
                   "patches_per_level": levels, "dt_coarse": dt,
                   "cell_updates_per_step": r32["cells"] // args.steps,
                   "hipgraph": {"captures": r32["graph_captures"], "replays": r32["graph_replays"]},
                   "launch": tuned},
        # The step kernel is bound by VALU issue (SQ counters), not HBM: the
        # roofline is the FP32 vector peak, with the VALU busy fraction and
        # the HBM fraction beside it.
        "roofline": {"bound": "valu",
                     "achieved": flops / 1e12 if flops is not None else None,
                     "peak": FP32_VALU_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": flops / FP32_VALU_PEAK if flops is not None else None,
                     "flop_per_cell_update": FLOP_PER_CELL,
                     "valu_busy": VALU_ACTIVE if measured else None,
                     "valu_insts_per_cell_update": VALU_INSTS_PER_CELL if measured else None,
                     "valu_issue_frac": (VALU_INSTS_PER_CELL * k_rate / VALU_ISSUE_PEAK) if measured else None,
                     "traffic": TRAFFIC_PER_LAUNCH if measured else None,
                     "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE, WRITE_SIZE)",
                     "algorithmic_per_launch": b_upd * r32["kcells"] / max(r32["launches"], 1),
                     "hbm": {"achieved": hbm_achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                             "frac": hbm_achieved / HBM_PEAK, "bytes_per_cell_update": b_upd},
                     "kernel": "k_step2<float,Swe>",
                     "kernel_only_rate": k_rate,
                     "kernel_launches": r32["launches"]},
        "cpu_baseline": cpu,
    }
    if port is not None:
        res["cpu_port_openmp"] = port
    if small is not None:
        res["configs"] = small
    if hp is not None:
        res["host_path_pcie"] = hp
    if args.cnn > 0:
        sim32.close()
        sim64.close()
        res["cnn_c4"] = cnn_c4(args.cnn)
    if args.no_push or args.no_merge:
        res["config"]["features"] = {"ghost_push": not args.no_push, "merge_levels": not args.no_merge}
    if r64 is not None:
        def rate(r, what):
            k = r["kernel_ms"] / 1e3
            return {"value": r["cells"] / r["seconds"], "ms_per_step": 1e3 * r["seconds"] / args.steps,
                    "kernel_only_rate": r["kcells"] / k if k > 0 else None, "path": what}
        res["fp64_plain"] = rate(r64p, "FP64, one step launch per level and level step, ghost "
                                       "copies by the ghost kernel, eager launches")
        res["fp64_batched"] = rate(r64, "FP64 with the FP32 run's batching: one launch per "
                                        "subcycled time, ghost push, hipGraph replay")
        res["speedup_vs_fp64_plain"] = value / res["fp64_plain"]["value"]
        res["speedup_vs_fp64_batched"] = value / res["fp64_batched"]["value"]
        res["fp32_vs_fp64"] = err
    if args.copy_calibration:
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        res["copy_calibration_bytes"] = {"read": 1 << 30, "write": 1 << 30}
        del x, y
    print(json.dumps(res))


if __name__ == "__main__":
    main()
