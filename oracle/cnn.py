"""CPU restatement of the learned refinement indicator (the network of
forestclaw_amd/csrc/hc_cnn.hip).  TEST INFRASTRUCTURE ONLY: the checker of
the GPU forward and backward kernels, never on the product path.

Network: 36x36 tiles, channels last, 8 input channels -> conv3x3 8->16 ->
ReLU -> 16->32 -> ReLU -> 32->32 -> ReLU -> 32->16 -> ReLU -> global average
pool -> linear 16->2; mean softmax cross-entropy.  Convolutions use zero
'same' padding; weights are stored [Cout][3][3][Cin] (k = tap*Cin + ci, tap
= 3*(dy+1) + (dx+1)).

quantize=True rounds to bfloat16 (round to nearest even) where the GPU
stores bfloat16: input tiles, convolution weights, every convolution output,
the pooled-gradient broadcast and every activation gradient; everything
else is float64 here (FP32 on the GPU).  quantize=False is the exact
float64 network (used for finite-difference checks).
"""
from __future__ import annotations

import numpy as np

TW = 36
CIN = (8, 16, 32, 32)
COUT = (16, 32, 32, 16)


def layout():
    off, o = {}, 0
    for l in range(4):
        off[f"w{l}"] = (o, (COUT[l], 9, CIN[l]))
        o += COUT[l] * 9 * CIN[l]
        off[f"b{l}"] = (o, (COUT[l],))
        o += COUT[l]
    off["fcw"] = (o, (2, 16))
    o += 32
    off["fcb"] = (o, (2,))
    o += 2
    return off, o


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    off, _ = layout()
    return {k: np.asarray(params[o:o + int(np.prod(s))], dtype=np.float64).reshape(s)
            for k, (o, s) in off.items()}


def _shift(x, dy, dx):
    """x (n, H, W, C) -> y with y[:, Y, X] = x[:, Y+dy, X+dx] (zero outside)."""
    n, H, W, C = x.shape
    y = np.zeros_like(x)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    y[:, ys:ye, xs:xe] = x[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return y


TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def conv(x, W, b):
    """z[n,y,x,co] = sum_tap sum_ci x[n, y+dy, x+dx, ci] W[co, tap, ci] + b[co]."""
    z = np.zeros(x.shape[:3] + (W.shape[0],))
    for t, (dy, dx) in enumerate(TAPS):
        z += _shift(x, dy, dx) @ W[:, t, :].T
    return z + b


def forward(params, tiles, quantize=True):
    P = unpack(params)
    q = bf16 if quantize else (lambda v: np.asarray(v, dtype=np.float64))
    x = q(tiles)
    acts = [x]
    for l in range(4):
        W = q(P[f"w{l}"])
        x = q(np.maximum(conv(x, W, P[f"b{l}"]), 0.0))
        acts.append(x)
    feat = acts[-1].mean(axis=(1, 2))
    logits = feat @ P["fcw"].T + P["fcb"]
    return logits, acts, feat


def loss_and_grad(params, tiles, labels, quantize=True):
    """Mean softmax cross-entropy and its gradient (flat, same layout)."""
    P = unpack(params)
    q = bf16 if quantize else (lambda v: np.asarray(v, dtype=np.float64))
    logits, acts, feat = forward(params, tiles, quantize)
    n = len(labels)
    z = logits - logits.max(axis=1, keepdims=True)
    p = np.exp(z) / np.exp(z).sum(axis=1, keepdims=True)
    onehot = np.eye(2)[np.asarray(labels)]
    loss = -np.log(p[np.arange(n), labels]).mean()
    dlog = (p - onehot) / n
    g = {"fcw": dlog.T @ feat, "fcb": dlog.sum(0)}
    dfeat = dlog @ P["fcw"]
    da = q(np.broadcast_to(dfeat[:, None, None, :] / (TW * TW), acts[4].shape))
    for l in (3, 2, 1, 0):
        W = q(P[f"w{l}"])
        dz = da * (acts[l + 1] > 0)
        xin = acts[l]
        gw = np.zeros(W.shape)
        for t, (dy, dx) in enumerate(TAPS):
            gw[:, t, :] = np.einsum("nyxo,nyxi->oi", dz, _shift(xin, dy, dx))
        g[f"w{l}"] = gw
        g[f"b{l}"] = dz.sum(axis=(0, 1, 2))
        if l > 0:
            din = np.zeros(xin.shape)
            for t, (dy, dx) in enumerate(TAPS):
                din += _shift(dz, -dy, -dx) @ W[:, t, :]
            da = q(din)
    off, total = layout()
    flat = np.zeros(total)
    for k, (o, s) in off.items():
        flat[o:o + int(np.prod(s))] = g[k].ravel()
    return loss, flat
