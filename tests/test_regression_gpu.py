"""GPU reproduction of the reference's own regression checks
(applications/clawpack/*/regressions.sh with --regression-check): each case
runs the options of the application's regression.ini through the solver's
initialisation, the run loop of fclaw2d_run outstyle 1 and a regrid after every
step, and must give the counters of the reference's regression_values CSV
(copied into tests/golden/regression/): advance calls, patches at the end and
regrids that changed the mesh.  These pin the AMR semantics end to end: the
tag criterion, refinement smoothing, 2:1 balance, the initial refinement,
the time-step counters of subcycling and the CFL-driven step control.

The shock-bubble case needs the axisymmetric source term and the inflow
boundary routine of its application (src2, bc2 hooks), which are not built.
"""
import os

import pytest

from forestclaw_amd import configs as CF

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
