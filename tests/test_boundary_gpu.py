"""GPU tests of the host-side drop-in entry points.

* hipclaw_single_step_update (host-resident patches, the deferred
  patch_vt->single_step_update protocol of fclaw2d_update_single_step with
  iter / total_count): flushes when the buffer is full and on the last patch;
  every patch interior bitwise equal to the oracle step (FP64), the fmax of
  the returned values equal to the oracle's level CFL; FP32 within 1e-5.
* hipclaw_level_upload / hipclaw_level_download: per-patch host arrays of a
  level round-trip exactly and feed the same solution as hipclaw_set_q.
"""
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
