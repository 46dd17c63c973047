"""CPU checks of the network oracle (oracle/cnn.py).

* one convolution layer equals scipy.signal.correlate2d summed over input
  channels (the textbook 'same' cross-correlation with zero padding);
* the analytic gradient of the exact float64 network equals central finite
  differences for entries of every parameter tensor;
* bfloat16 rounding is round-to-nearest-even on the float32 bit pattern.
"""
import numpy as np
import pytest
from scipy.signal import correlate2d

from oracle import cnn as N


def _params(seed):
    _, total = N.layout()
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, total)
    return p


def test_conv_layer_is_cross_correlation():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 36, 36, 8))
    W = rng.normal(size=(16, 9, 8))
    b = rng.normal(size=16)
    z = N.conv(x, W, b)
    for n in range(2):
        for co in (0, 7, 15):
            ref = sum(correlate2d(x[n, :, :, ci], W[co, :, ci].reshape(3, 3), mode="same")
                      for ci in range(8)) + b[co]
            np.testing.assert_allclose(z[n, :, :, co], ref, rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(2)
    tiles = rng.normal(size=(3, 36, 36, 8))
    tiles[..., 3:] = 0.0
    labels = np.array([0, 1, 1])
    loss, g = N.loss_and_grad(p, tiles, labels, quantize=False)
    off, _ = N.layout()
    eps = 1e-6
    for name, (o, shape) in off.items():
        size = int(np.prod(shape))
        for j in rng.choice(size, min(3, size), replace=False):
            e = o + j
            pp, pm = p.copy(), p.copy()
            pp[e] += eps
            pm[e] -= eps
            fd = (N.loss_and_grad(pp, tiles, labels, quantize=False)[0] -
                  N.loss_and_grad(pm, tiles, labels, quantize=False)[0]) / (2 * eps)
            assert g[e] == pytest.approx(fd, rel=1e-4, abs=1e-9), (name, j)


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -3.14159, 1e-40, 65504.0])
    got = N.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa -> 1
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -6
    assert got[3] == -3.140625
    assert got[4] == 2.0 ** -133  # nearest bfloat16 subnormal
    assert got[5] == 65536.0
