"""GPU checks of the learned refinement indicator's BF16 MFMA kernels.

* forward logits and every parameter gradient (conv weights / biases via the
  MFMA weight- and data-gradient kernels, head) against the CPU oracle with
  the same bfloat16 rounding points (oracle/cnn.py): logits within 1e-4 of
  their scale, the loss within 1e-4 relative, each gradient tensor within
  1e-4 relative L2 (measured: 1e-5 for the logits, at most 7e-6 for the
  gradients; what remains is FP32 against FP64 accumulation);
* the batch index list gathers tiles exactly;
* Adam training on the GPU separates tiles with a sharp front from smooth
  ones (loss < 0.3, accuracy > 95%).
"""
import numpy as np
import pytest

from forestclaw_amd import hipclaw as H
from oracle import cnn as N

pytestmark = pytest.mark.gpu


def _tiles(rng, n, front=None):
    """Smooth random fields, optionally with a sharp front (normalised units)."""
    y, x = np.mgrid[0:36, 0:36] / 36.0
    t = np.zeros((n, 36, 36, 8))
    lab = np.zeros(n, dtype=np.int32)
    for k in range(n):
        a, b, c = rng.normal(size=3)
        for ch in range(3):
            t[k, :, :, ch] = 0.2 * rng.normal() * np.sin(2 * np.pi * (a * x + b * y + c + ch))
        f = rng.random() < 0.5 if front is None else front
        if f:
            cx, cy, r = rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.15, 0.4)
            t[k, :, :, 0] += 2.0 * (np.hypot(x - cx, y - cy) < r)
            lab[k] = 1
    return t, lab


def _dev(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).cuda()


def test_forward_and_gradients_match_oracle(dev):
    import torch
    rng = np.random.default_rng(7)
    net = H.Cnn(seed=11)
    p = net.get_params().astype(np.float64)
    tiles, lab = _tiles(rng, 6)
    tq = N.bf16(tiles)
    td = _dev(tq, torch.bfloat16)
    lg = torch.zeros(6, 2, dtype=torch.float32, device=dev)
    net.forward(td, lg)
    torch.cuda.synchronize()
    ref, _, _ = N.forward(p, tq)
    got = lg.cpu().numpy()
    scale = np.abs(ref).max()
    print(f"\nlogits max abs err {np.abs(got - ref).max():.3e} (scale {scale:.3e})")
    assert np.abs(got - ref).max() <= 1e-4 * scale

    loss, g = net.gradients(td, _dev(lab, torch.int32))
    rl, rg = N.loss_and_grad(p, tq, lab)
    assert loss == pytest.approx(rl, rel=1e-4)
    off, _ = N.layout()
    for name, (o, shape) in off.items():
        a = g[o:o + int(np.prod(shape))].astype(np.float64)
        b = rg[o:o + int(np.prod(shape))]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
        print(f"{name}: rel L2 {rel:.3e} cos {cos:.6f}")
        assert rel <= 1e-4 and cos >= 1 - 1e-8, name


def test_index_gather(dev):
    import torch
    rng = np.random.default_rng(3)
    net = H.Cnn(seed=5)
    tiles, _ = _tiles(rng, 10)
    td = _dev(N.bf16(tiles), torch.bfloat16)
    perm = np.array([7, 2, 9, 0, 4], dtype=np.int32)
    a = torch.zeros(5, 2, device=dev)
    b = torch.zeros(5, 2, device=dev)
    net.forward(td, a, idx=_dev(perm, torch.int32), n=5)
    net.forward(td[torch.from_numpy(perm).long().cuda()].contiguous(), b)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_training_separates_fronts(dev):
    import torch
    rng = np.random.default_rng(4)
    tiles, lab = _tiles(rng, 512)
    td = _dev(N.bf16(tiles), torch.bfloat16)
    ld = _dev(lab, torch.int32)
    net = H.Cnn(seed=9)
    loss = torch.zeros(1, device=dev)
    hist = []
    for step in range(80):
        idx = _dev(rng.permutation(512)[:128].astype(np.int32), torch.int32)
        net.train_step(td, ld, idx, 128, 3e-3, loss)
        hist.append(float(loss.item()))
    lg = torch.zeros(512, 2, device=dev)
    net.forward(td, lg)
    acc = float(((lg[:, 1] > lg[:, 0]).int().cpu().numpy() == lab).mean())
    print(f"\nloss {hist[0]:.3f} -> {np.mean(hist[-5:]):.3f}, accuracy {acc:.3f}")
    assert np.mean(hist[-5:]) < 0.3 and acc > 0.95


def test_tag_patches_of_host_arrays_match_the_regrid_pass(dev):
    """hipclaw_cnn_tag_patches (the batched tag4refinement hook for a host-
    resident forest) gives, from host copies of a context's patches, the
    refine decisions of the context's own pass: its tiles and the network's
    logits (p(refine) >= 0.5)."""
    import torch
    from forestclaw_amd import learned_tagging as LT
    sim = LT.make_sim(2, 4, precision=H.FP64)
    sim.initialize("radial_dam", (0.3, -0.2, 0.6, 2.5, 1.0), nsub=4)
    net = H.Cnn(seed=5)
    n = sim.num_patches()
    tiles = torch.empty((n, 36, 36, 8), dtype=torch.bfloat16, device=dev)
    sim.make_tiles(tiles)
    lg = torch.empty((n, 2), dtype=torch.float32, device=dev)
    net.forward(tiles, lg)
    torch.cuda.synchronize()
    lgd = lg.cpu().numpy().astype(np.float64)
    want = (lgd[:, 1] - lgd[:, 0] >= 0).astype(np.int32)
    q = sim.get_q()
    info = sim.patch_info()
    got = net.tag_patches(list(q), info["dx"], LT.THRESHOLD)
    assert len(set(info["level"])) >= 2 and 0 < want.sum() < n
    assert np.array_equal(got, want)
    assert len(net.tag_patches([], [], LT.THRESHOLD)) == 0
