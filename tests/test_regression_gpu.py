"""GPU reproduction of the reference's own regression checks
(applications/clawpack/*/regressions.sh with --regression-check): each case
runs the options of the application's regression.ini through the solver's
initialisation, the run loop of fclaw2d_run outstyle 1 and a regrid after every
step, and must give the counters of the reference's regression_values CSV
(copied into tests/golden/regression/): advance calls, patches at the end and
regrids that changed the mesh.  These pin the AMR semantics end to end: the
tag criterion, refinement smoothing, 2:1 balance, the initial refinement,
the time-step counters of subcycling and the CFL-driven step control.

The radial-dam case starts from the reference's polygon cell average of the
dam (cellave2, qinit nsub = 0), checked here bitwise against the oracle's
restatement first.

The shock-bubble case adds the application's initial condition and its
axisymmetric source term after every patch step (src2); with its mthbc
(1 1 3 1) the application's bc2 reduces to the default routine.
"""
import os

import numpy as np
import pytest

from forestclaw_amd import configs as CF
from forestclaw_amd import forest as F
from forestclaw_amd import hipclaw as H
from oracle import cellave as CA

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "regression")


def expected(name):
    out = {}
    with open(os.path.join(GOLD, f"{name}_regression_values_0.csv")) as f:
        for line in f:
            k, v = line.strip().split(",")
            out[k] = int(v)
    return out


def test_swirl_regression_counts(dev):
    got = CF.swirl_regression()
    print("\nswirl", got)
    assert got == expected("swirl")


@pytest.mark.parametrize("prm", [(0.0, 0.0, 0.5, 2.0, 1.0), (0.31, -0.17, 0.83, 1.5, 0.25)])
def test_radialdam_polygon_qinit_bitwise(dev, prm):
    opt = F.Options(minlevel=2, maxlevel=3, mx=8, my=8, solver=H.SOLVER_SWE_ROE, ax=-2.5, bx=2.5, ay=-2.5,
                    by=2.5)
    sim = F.Hipclaw46(opt)
    try:
        lv, gx, gy = [], [], []
        for y in range(4):
            for x in range(4):
                if x in (1, 2) and y in (1, 2):
                    for c in range(4):
                        lv.append(3), gx.append(2 * x + (c & 1)), gy.append(2 * y + (c >> 1))
                else:
                    lv.append(2), gx.append(x), gy.append(y)
        sim.init_leaves(lv, gx, gy)
        sim.qinit("radial_dam", prm, nsub=0)
        q = sim.get_q()
        info = sim.patch_info()
        cut = 0
        for k in range(q.shape[0]):
            ref = np.array(CA.radialdam_patch(prm, info["xlower"][k], info["ylower"][k], info["dx"][k],
                                              info["dy"][k], opt.mx, opt.my))
            assert np.array_equal(q[k, 0], ref), k
            assert not q[k, 1:].any()
            cut += int(((ref > prm[4]) & (ref < prm[3])).sum())
        assert cut > 20
    finally:
        sim.close()


def test_radialdam_regression_counts(dev):
    got = CF.radialdam_regression()
    print("\nradialdam", got)
    assert got == expected("radialdam")


def test_shockbubble_regression_counts(dev):
    got = CF.shockbubble_regression()
    print("\nshockbubble", got)
    assert got == expected("shockbubble")
