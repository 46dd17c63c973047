"""hipclaw_step2_batch_noalloc (the entry point that allocates nothing: the
per-tile Courant numbers live in the library's static device array) against
hipclaw_step2_batch with a cfl_patch workspace reduced by hipclaw_max_reduce:
the same tiles and the same maximum Courant number, bit for bit, on two calls
in a row, the second over fewer tiles (so stale values of the first call in
the static array must not leak into its maximum)."""
import numpy as np
import pytest

import cases as C
from forestclaw_amd import hipclaw as H

pytestmark = pytest.mark.gpu


def _tiles(kind, rng, n, mx):
    qs = []
    for _ in range(n):
        if kind == "swe":
            q, dx, dy = C.swe_state(rng, mx, mx, 2)
        else:
            q, dx, dy = C.euler_state(rng, mx, mx, 2)
        qs.append(q)
    return np.stack(qs), dx, dy


@pytest.mark.parametrize("precision", [H.FP64, H.FP32])
@pytest.mark.parametrize("kind", ["swe", "euler"])
def test_noalloc_equals_step2_batch_on_two_calls(dev, precision, kind):
    import torch
    mx = 32 if kind == "swe" else 16
    solver = H.SOLVER_SWE_ROE if kind == "swe" else H.SOLVER_EULER5_ROE
    p = H.physics(solver, mx, mx, precision=precision)
    tdt = torch.float64 if precision == H.FP64 else torch.float32
    rng = np.random.default_rng(17 if kind == "swe" else 18)
    # call 1: 300 tiles, all stepped; call 2: 24 of 60 tiles by a patch list,
    # at a smaller time step: values the first call left in the static array
    # are larger than the second call's maximum, so a stale one would show
    calls = [(300, None, 0.45), (60, rng.permutation(60)[:24].astype(np.int32), 0.1)]
    for ntiles, ids, frac in calls:
        print(f"{kind} precision {precision}: {ntiles} tiles, list {None if ids is None else len(ids)}",
              flush=True)
        q, dx, dy = _tiles(kind, rng, ntiles, mx)
        dt = frac * dx
        qin = torch.from_numpy(q).to(dev, tdt)
        dxdy = torch.tensor([[dx, dy]] * ntiles, dtype=torch.float64, device=dev)
        tids = None if ids is None else torch.from_numpy(ids).to(dev)
        npatch = ntiles if ids is None else len(ids)
        # existing entry point: per-tile values in a workspace, then the maximum
        ref_out = qin.clone()
        work = torch.zeros(ntiles, dtype=torch.float64, device=dev)
        ref_max = torch.zeros(1, dtype=torch.float64, device=dev)
        H.step2_batch(p, npatch, qin, ref_out, None, dxdy, dt, work, patch_ids=tids, ntiles=ntiles)
        H.max_reduce(work, ntiles, ref_max)
        # the entry point that allocates nothing
        out = qin.clone()
        got_max = torch.full((1,), -1.0, dtype=torch.float64, device=dev)
        H.step2_batch_noalloc(p, npatch, qin, out, None, dxdy, dt, got_max, patch_ids=tids, ntiles=ntiles)
        torch.cuda.synchronize()
        assert torch.equal(out, ref_out)
        assert got_max.item() == ref_max.item()
        assert got_max.item() > 0.0


def test_noalloc_rejects_too_many_tiles(dev):
    import torch
    p = H.physics(H.SOLVER_SWE_ROE, 8, 8)
    m = torch.zeros(1, dtype=torch.float64, device=dev)
    with pytest.raises(H.HipclawError):
        H.step2_batch_noalloc(p, 4, 1, 1, None, 1, 0.1, m, ntiles=H.NOALLOC_MAX_TILES + 1)
