"""Training data, training loop and evaluation of the learned refinement
indicator (hipclaw_cnn_*, include/hipclaw_cnn.h) on the shallow-water radial
dam break on [-2.5, 2.5]^2 (32x32 patches, g = 1, refine when |grad h| > 0.05,
the radialdam application's tagging).

Labels.  A tile's label is the plain refinement tag of its patch now or at
any level step within the next `lookahead` coarse steps (lookahead = 0: the
plain tag itself).  With richardson_tol set, a patch is also labelled for
refinement where its Richardson error estimate exceeds that tolerance: two
steps of the patch's level against one step of twice the size on the grid
coarsened by two, |average(q_2h) - q_H| / (2^2 - 1) for the second-order
method (the estimate AMRClaw's flag_richardson makes,
src/solvers/fc2d_clawpack5/fortran_source/clawpack5_amr_module.f90:273), taken
on the depth h.  The data come from uniform-level runs, where every patch
persists, so the future tags of a patch are known: for a level-L run the
coarse step of an AMR run with minlevel 2 is 2^(L-2) level steps.  With
lookahead K, a regrid every K coarse steps with the network's tags keeps the
moving front inside refined patches, which the plain tags (computed on the
current state only) cannot do at that regrid interval.

Everything runs on the device through the C ABI; this module only drives it.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import forest as F
from . import hipclaw as H

DOMAIN = (-2.5, 2.5, -2.5, 2.5)
THRESHOLD = 0.05
MINLEVEL = 2


def dam_case(rng):
    """Random centre, radius 0.3-0.8, h ratio 1.5-4 on h = 1 at rest."""
    return (float(rng.uniform(-1.2, 1.2)), float(rng.uniform(-1.2, 1.2)),
            float(rng.uniform(0.3, 0.8)), float(rng.uniform(1.5, 4.0)), 1.0)


def make_sim(minlevel, maxlevel, precision=H.FP32, threshold=THRESHOLD, time_sync=False):
    ax, bx, ay, by = DOMAIN
    opt = F.Options(minlevel=minlevel, maxlevel=maxlevel, mx=32, my=32, ax=ax, bx=bx, ay=ay, by=by,
                    solver=H.SOLVER_SWE_ROE, precision=precision, grav=1.0,
                    refine_threshold=threshold, coarsen_threshold=threshold / 4, subcycle=True,
                    time_sync=time_sync)
    return F.Hipclaw46(opt)


def coarse_dt(h_in, cfl=0.4):
    """Fixed coarse step: CFL `cfl` at level MINLEVEL for the dam's wave speed."""
    dx = (DOMAIN[1] - DOMAIN[0]) / (32 << MINLEVEL)
    return cfl * dx / (2.0 * math.sqrt(h_in))


@dataclass
class Dataset:
    tiles: object   # torch bf16 [n, 36, 36, 8] on the device
    labels: object  # torch int32 [n]

    def __len__(self):
        return int(self.labels.shape[0])


def generate(nruns=24, levels=(3, 4, 5), samples=12, lookahead=0, neg_keep=0.35, seed=0,
             precision=H.FP32, richardson_tol=None, stats=None) -> Dataset:
    """Tiles and labels from uniform-level dam-break runs (one sample per
    coarse step; all positive tiles and a fraction neg_keep of the others).
    richardson_tol: also label patches whose Richardson error estimate
    exceeds it (stats, a dict, then collects the counts of tag-only,
    estimate-only and both labels)."""
    import torch
    rng = np.random.default_rng(seed)
    tiles, labels = [], []
    for _ in range(nruns):
        case = dam_case(rng)
        dtc = coarse_dt(case[3])
        for L in levels:
            sub = 1 << (L - MINLEVEL)
            sim = make_sim(L, L, precision)
            sim.initialize("radial_dam", case, nsub=4)
            n = sim.num_patches()
            dt = dtc / sub
            nsteps = (samples + lookahead) * sub
            rich = Richardson(sim, precision) if richardson_tol is not None else None
            tag_hist = []
            saved = {}
            rflag = {}
            for s in range(nsteps + 1):
                tg = torch.empty(n, dtype=torch.int32, device="cuda")
                if s % sub == 0 and s // sub < samples:
                    tl = torch.empty((n, 36, 36, 8), dtype=torch.bfloat16, device="cuda")
                    sim.make_tiles(tl, tg)
                    saved[s // sub] = tl
                    if rich is not None:
                        rflag[s // sub] = rich.estimate(sim.get_q(), s * dt, dt) > richardson_tol
                else:
                    sim.make_tiles(_scratch(n), tg)
                tag_hist.append(tg)
                if s < nsteps:
                    sim.advance(s * dt, dt)
            sim.close()
            if rich is not None:
                rich.close()
            tags = torch.stack(tag_hist)  # (steps+1, n)
            for c in range(samples):
                lo, hi = c * sub, c * sub + lookahead * sub
                lab = tags[lo:hi + 1].amax(dim=0)
                if rich is not None:
                    rf = torch.from_numpy(rflag[c].astype(np.int32)).cuda()
                    if stats is not None:
                        for key, v in (("tag_only", (lab > 0) & (rf == 0)), ("estimate_only", (lab == 0) & (rf > 0)),
                                       ("both", (lab > 0) & (rf > 0)), ("patches", torch.ones_like(lab) > 0)):
                            stats[key] = stats.get(key, 0) + int(v.sum())
                    lab = torch.maximum(lab, rf)
                keep = (lab > 0) | (torch.rand(n, device="cuda") < neg_keep)
                tiles.append(saved[c][keep])
                labels.append(lab[keep])
    return Dataset(torch.cat(tiles).contiguous(), torch.cat(labels).contiguous())


class Richardson:
    """Richardson error estimate of the depth for every patch of a uniform
    level-L run: from the state q at time t, two level-L steps of dt against
    one step of 2 dt on the level L-1 grid whose interiors are the 2x2
    averages of q (ghost cells from a ghost update, extrapolation at the
    domain boundary); the estimate of a level-L patch is the largest
    |average(q_2 dt) - q_coarse| / 3 over the coarse cells it covers."""

    def __init__(self, sim, precision):
        self.info = sim.patch_info()
        self.L = int(self.info["level"][0])
        assert (self.info["level"] == self.L).all() and self.L >= 1
        self.fine = make_sim(self.L, self.L, precision)
        self.fine.init_leaves(self.info["level"], self.info["gx"], self.info["gy"])
        fi = self.fine.patch_info()
        assert (fi["gx"] == self.info["gx"]).all() and (fi["gy"] == self.info["gy"]).all()
        self.coarse = make_sim(self.L - 1, self.L - 1, precision)
        self.coarse.init_uniform(self.L - 1)
        ci = self.coarse.patch_info()
        self.cgx, self.cgy = ci["gx"].astype(int), ci["gy"].astype(int)

    def estimate(self, q, t, dt):
        L = self.L
        fine, coarse = self.fine, self.coarse
        fine.set_q(q)
        fine.ghost_update(L, L)
        fine.advance(t, dt)
        fine.advance(t + dt, dt)
        q2 = fine.get_q()
        # coarse interiors: 2x2 averages of the fine interiors, field by field
        N = 32 << L
        qc = np.zeros((len(self.cgx),) + coarse.tile_shape)
        for m in range(q.shape[1]):
            g = self._global(q, m, N).reshape(N // 2, 2, N // 2, 2).mean(axis=(1, 3))
            for k, (gx, gy) in enumerate(zip(self.cgx, self.cgy)):
                qc[k, m, 2:-2, 2:-2] = g[gy * 32:(gy + 1) * 32, gx * 32:(gx + 1) * 32]
        coarse.set_q(qc)
        coarse.ghost_update(L - 1, L - 1)
        coarse.advance(t, 2 * dt)
        qc1 = coarse.get_q()
        hc = np.zeros((N // 2, N // 2))
        for k, (gx, gy) in enumerate(zip(self.cgx, self.cgy)):
            hc[gy * 32:(gy + 1) * 32, gx * 32:(gx + 1) * 32] = qc1[k, 0, 2:-2, 2:-2]
        e = np.abs(self._global(q2, 0, N).reshape(N // 2, 2, N // 2, 2).mean(axis=(1, 3)) - hc) / 3.0
        out = np.zeros(len(q))
        for k, (gx, gy) in enumerate(zip(self.info["gx"], self.info["gy"])):
            out[k] = e[gy * 16:(gy + 1) * 16, gx * 16:(gx + 1) * 16].max()
        return out

    def _global(self, q, m, N):
        g = np.zeros((N, N))
        for k, (gx, gy) in enumerate(zip(self.info["gx"], self.info["gy"])):
            g[gy * 32:(gy + 1) * 32, gx * 32:(gx + 1) * 32] = q[k, m, 2:-2, 2:-2]
        return g

    def close(self):
        self.fine.close()
        self.coarse.close()


_SCRATCH = {}


def _scratch(n):
    import torch
    t = _SCRATCH.get("t")
    if t is None or t.shape[0] < n:
        t = torch.empty((n, 36, 36, 8), dtype=torch.bfloat16, device="cuda")
        _SCRATCH["t"] = t
    return t


def train(net: H.Cnn, data: Dataset, epochs=3, batch=1024, lr=2e-3, seed=0):
    """Adam on shuffled batches; returns the mean loss of each epoch."""
    import torch
    rng = np.random.default_rng(seed)
    n = len(data)
    loss = torch.zeros(1, device="cuda")
    hist = []
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(n).astype(np.int32)).cuda()
        tot, cnt = 0.0, 0
        for b0 in range(0, n - batch + 1, batch):
            idx = perm[b0:b0 + batch]
            net.train_step(data.tiles, data.labels, idx, batch, lr, loss)
            tot += float(loss.item())
            cnt += 1
        hist.append(tot / max(cnt, 1))
    return hist


def predict(net: H.Cnn, data: Dataset, batch=4096):
    import torch
    out = torch.empty((len(data), 2), dtype=torch.float32, device="cuda")
    for b0 in range(0, len(data), batch):
        nb = min(batch, len(data) - b0)
        idx = torch.arange(b0, b0 + nb, dtype=torch.int32, device="cuda")
        net.forward(data.tiles, out[b0:], idx=idx, n=nb)
    return out


def fine_reference(case, ncoarse, maxlevel=6, precision=H.FP64):
    """Uniform maxlevel solution after ncoarse coarse steps, as an array of h."""
    sim = make_sim(maxlevel, maxlevel, precision)
    sim.initialize("radial_dam", case, nsub=4)
    dtc = coarse_dt(case[3])
    sub = 1 << (maxlevel - MINLEVEL)
    for s in range(ncoarse * sub):
        sim.advance(s * dtc / sub, dtc / sub)
    H_ = _assemble(sim, maxlevel)
    sim.close()
    return H_


def _assemble(sim, level):
    info = sim.patch_info()
    q = sim.get_q()
    N = 32 << level
    out = np.zeros((N, N))
    for k in range(len(q)):
        gx, gy = int(info["gx"][k]), int(info["gy"][k])
        out[gy * 32:(gy + 1) * 32, gx * 32:(gx + 1) * 32] = q[k, 0, 2:-2, 2:-2]
    return out


def amr_error(sim, href, maxlevel=6):
    """Relative L1 error of h over the leaves against the fine reference
    averaged onto each leaf's cells."""
    info = sim.patch_info()
    q = sim.get_q()
    num = den = 0.0
    for k in range(len(q)):
        L, gx, gy = int(info["level"][k]), int(info["gx"][k]), int(info["gy"][k])
        r = 1 << (maxlevel - L)
        blk = href[gy * 32 * r:(gy + 1) * 32 * r, gx * 32 * r:(gx + 1) * 32 * r]
        avg = blk.reshape(32, r, 32, r).mean(axis=(1, 3))
        area = (r * r)
        num += np.abs(q[k, 0, 2:-2, 2:-2] - avg).sum() * area
        den += np.abs(avg).sum() * area
    return num / den


def run_amr(case, ncoarse, regrid_interval, net=None, margin=0.0, maxlevel=6, precision=H.FP32):
    """C2 run with plain (net None) or learned refinement tags; returns the
    simulation, the number of regrids and the mean tag statistics."""
    sim = make_sim(MINLEVEL, maxlevel, precision)
    sim.initialize("radial_dam", case, nsub=4)
    if net is not None:
        sim.set_tagging_cnn(net, margin)
    dtc = coarse_dt(case[3])
    nreg = 0
    stats = []
    for s in range(ncoarse):
        sim.advance(s * dtc, dtc)
        if (s + 1) % regrid_interval == 0 and s + 1 < ncoarse:
            sim.regrid()
            nreg += 1
            if net is not None:
                stats.append(sim.tag_stats())
    return sim, nreg, stats
