"""GPU tests of the host-side drop-in entry points.

* hipclaw_single_step_update (host-resident patches, the deferred
  patch_vt->single_step_update protocol of fclaw2d_update_single_step with
  iter / total_count): flushes when the buffer is full and on the last patch;
  every patch interior bitwise equal to the oracle step (FP64), the fmax of
  the returned values equal to the oracle's level CFL; FP32 within 1e-5.
* hipclaw_level_upload / hipclaw_level_download: per-patch host arrays of a
  level round-trip exactly and feed the same solution as hipclaw_set_q.
* hipclaw_output_ascii: a fort.t / fort.q frame of an AMR context lists every
  patch once in the reference's order with its geometry and interior values.
"""
import os

import numpy as np
import pytest

import cases as C
from forestclaw_amd import forest as F
from forestclaw_amd import hipclaw as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _level(nps, mx=16, seed=2):
    rng = np.random.default_rng(seed)
    qs, dts = [], []
    for _ in range(nps):
        q, dx, dy = C.swe_state(rng, mx, mx, 2)
        qs.append(np.ascontiguousarray(q))
        dts.append(C.swe_dt(q, dx, dy))
    return qs, dx, dy, min(dts)


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_single_step_update_protocol(dev, precision):
    import torch
    nps, cap = 23, 8
    qs, dx, dy, dt = _level(nps)
    cfg = O.make_cfg(O.SOLVER_SWE_ROE, 16, 16, 2, grav=1.0)
    qb = np.stack([q.copy() for q in qs])
    ocfl = O.step2_batch(cfg, qb, None, dx, dy, dt)
    b = H.Batch(H.physics(H.SOLVER_SWE_ROE, 16, 16, precision=precision), cap,
                torch.cuda.current_stream().cuda_stream)
    work = [q.copy() for q in qs]
    got = []
    for it in range(nps):
        got.append(b.single_step_update(work[it], None, dx, dy, 0.0, dt, it, nps))
        flushed = (it + 1) % cap == 0 or it == nps - 1
        assert (got[-1] > 0) == flushed, it
        if not flushed:  # staged patches are untouched until their flush
            assert np.array_equal(work[it], qs[it])
    b.close()
    w = np.stack(work)[:, :, 2:-2, 2:-2]
    o = qb[:, :, 2:-2, 2:-2]
    if precision == H.FP64:
        assert np.array_equal(w, o)
        assert max(got) == ocfl.max()
    else:
        assert np.abs(w - o).sum() / np.abs(o).sum() <= 1e-5
        assert max(got) == pytest.approx(ocfl.max(), rel=1e-5)
    # ghost cells are handed back as they were given
    assert np.array_equal(np.stack(work)[:, :, :2, :], np.stack(qs)[:, :, :2, :])


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_single_step_update_fwave_transport(dev, precision):
    """Mode A with the f-wave transport solver: host patches with their edge
    data (aux, 7 fields) staged per patch; FP64 bitwise against the oracle,
    FP32 within 1e-5."""
    import torch
    rng = np.random.default_rng(17)
    nps, cap = 11, 4
    qs, auxs = [], []
    for _ in range(nps):
        q, aux, dx, dy = C.fw_state(rng, 16, 16, 2, per_edge=True, lengths=True)
        qs.append(q)
        auxs.append(aux)
    dt = min(C.fw_dt(a, dx, dy) for a in auxs)
    qb = np.stack([q.copy() for q in qs])
    ocfl = O.step2_batch(O.make_cfg(O.SOLVER_TRANSPORT_FW, 16, 16, 2), qb, np.stack(auxs), dx, dy, dt)
    b = H.Batch(H.physics(H.SOLVER_TRANSPORT_FW, 16, 16, precision=precision), cap,
                torch.cuda.current_stream().cuda_stream)
    work = [q.copy() for q in qs]
    got = [b.single_step_update(work[it], auxs[it], dx, dy, 0.0, dt, it, nps) for it in range(nps)]
    b.close()
    w = np.stack(work)[:, :, 2:-2, 2:-2]
    o = qb[:, :, 2:-2, 2:-2]
    if precision == H.FP64:
        assert np.array_equal(w, o)
        assert max(got) == ocfl.max()
    else:
        assert np.abs(w - o).sum() / np.abs(o).sum() <= 1e-5
        assert max(got) == pytest.approx(ocfl.max(), rel=1e-5)


def test_single_step_update_rejects_bad_counters(dev):
    import torch
    qs, dx, dy, dt = _level(1)
    b = H.Batch(H.physics(H.SOLVER_SWE_ROE, 16, 16), 4, torch.cuda.current_stream().cuda_stream)
    with pytest.raises(H.HipclawError):
        b.single_step_update(qs[0], None, dx, dy, 0.0, dt, 3, 3)
    with pytest.raises(H.HipclawError):
        b.single_step_update(qs[0], None, dx, dy, 0.0, dt, 0, 2)
        b.single_step_update(qs[0], None, dx, dy, 0.0, 2 * dt, 1, 2)  # dt changed
    b.close()


def test_level_upload_download(dev):
    opt = F.Options(minlevel=2, maxlevel=3, mx=8, my=8, solver=H.SOLVER_SWE_ROE,
                    refine_threshold=0.5)
    sim = F.Hipclaw46(opt)
    sim.initialize("radial_dam", (0.3, 0.35, 0.15, 2.0, 1.0), nsub=4)
    info = sim.patch_info()
    q = sim.get_q()
    for L in sorted(set(info["level"])):
        idx = np.nonzero(info["level"] == L)[0]
        tiles = [np.zeros(sim.tile_shape) for _ in idx]
        sim.level_download(int(L), tiles)
        for t, k in zip(tiles, idx):
            assert np.array_equal(t, q[k])
        new = [np.ascontiguousarray(t * 1.5) for t in tiles]
        sim.level_upload(int(L), new)
        back = [np.zeros(sim.tile_shape) for _ in idx]
        sim.level_download(int(L), back)
        assert all(np.array_equal(a, b) for a, b in zip(new, back))
    with pytest.raises(H.HipclawError):
        sim.level_upload(int(info["level"].min()), [np.zeros(sim.tile_shape)])


def _parse_frame(directory, iframe):
    """(time, meqn, ngrids, patches) of a fort.t/fort.q frame."""
    t = open(os.path.join(directory, f"fort.t{iframe:04d}")).read().split("\n")
    time, meqn, ngrids = float(t[0].split()[0]), int(t[1].split()[0]), int(t[2].split()[0])
    lines = open(os.path.join(directory, f"fort.q{iframe:04d}")).read().split("\n")
    pos, patches = 0, []
    for _ in range(ngrids):
        hdr = [int(lines[pos + k].split()[0]) for k in range(6)]
        geo = [float(lines[pos + 6 + k].split()[0]) for k in range(4)]
        num, level, block, rank, mx, my = hdr
        assert lines[pos + 10] == ""
        pos += 11
        q = np.zeros((meqn, my, mx))
        for j in range(my):
            for i in range(mx):
                q[:, j, i] = [float(v) for v in lines[pos].split()]
                pos += 1
            assert lines[pos] == "  "
            pos += 1
        patches.append(dict(num=num, level=level, block=block, mx=mx, my=my, geo=geo, q=q))
    assert all(l == "" for l in lines[pos:])
    return time, meqn, ngrids, patches


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
def test_output_ascii_frame(dev, tmp_path, precision):
    """hipclaw_output_ascii writes every patch once, numbered 0.. in the
    reference's order (Morton order of the leaves), with the patch geometry and
    the interior values to the 16 digits of E26.16."""
    opt = F.Options(minlevel=2, maxlevel=4, mx=8, my=8, solver=H.SOLVER_SWE_ROE, precision=precision,
                    refine_threshold=0.5, ax=-1.0, bx=1.5, ay=0.0, by=2.5)
    sim = F.Hipclaw46(opt)
    sim.initialize("radial_dam", (0.0, 1.0, 0.3, 2.0, 1.0), nsub=4)
    sim.advance(0.0, 0.002)
    sim.output_ascii(7, str(tmp_path), time=0.25)
    info = sim.patch_info()
    q = sim.get_q()
    time, meqn, ngrids, patches = _parse_frame(str(tmp_path), 7)
    assert time == 0.25 and meqn == 3 and ngrids == len(q)
    assert [p["num"] for p in patches] == list(range(ngrids))
    # Morton order of the lower-left corners at the finest level
    maxl = int(info["level"].max())

    def mort(L, gx, gy):
        x, y, k = gx << (maxl - L), gy << (maxl - L), 0
        for b in range(16):
            k |= ((x >> b) & 1) << (2 * b) | ((y >> b) & 1) << (2 * b + 1)
        return k
    want = sorted(range(ngrids), key=lambda k: mort(int(info["level"][k]), int(info["gx"][k]), int(info["gy"][k])))
    for p, k in zip(patches, want):
        assert p["level"] == info["level"][k] and p["block"] == 0 and (p["mx"], p["my"]) == (8, 8)
        np.testing.assert_allclose(p["geo"], [info["xlower"][k], info["ylower"][k], info["dx"][k], info["dy"][k]],
                                   rtol=1e-15, atol=0)
        np.testing.assert_allclose(p["q"], q[k][:, 2:-2, 2:-2], rtol=1e-15, atol=1e-99)
