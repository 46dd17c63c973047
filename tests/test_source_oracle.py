"""Pins the oracle's restatement of the shock-bubble application's source term,
aux and initial condition (oracle/source.py):

* src2 and setaux bitwise against the reference Fortran
  (applications/clawpack/euler/2d/shockbubble/user_4.6/src2.f, setaux.f):
  committed golden vectors, and live random cases when oracle/_ref is built;
* a planted mistake in the restatement (the second stage's energy flux using
  the old energy) is caught by the golden vectors;
* the post-shock state satisfies the Rankine-Hugoniot conditions of a shock of
  pressure ratio pinf running into gas at rest with rho = p = 1;
* the qinit's bubble is the polygon average (cells wholly inside / outside
  take qin / qout exactly), the tracer is doubled only on interior rows left
  of x = 0.5, and the post-shock region is x_low < 0.2.
"""
import math
import os
import sys

import numpy as np
import pytest

from oracle import source as S

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "src2", "src2.npz")
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_src2.so")


def _cases():
    z = np.load(GOLD)
    n = len([k for k in z.files if k.startswith("geom")])
    return [(z[f"q{c}"], z[f"aux{c}"], z[f"out{c}"], z[f"geom{c}"]) for c in range(n)]


def test_src2_and_setaux_match_reference_golden():
    for q, aux, out, (ylower, dy, dt, gamma) in _cases():
        my, mx = q.shape[1] - 4, q.shape[2] - 4
        rad = S.setaux_rad(mx, my, 2, ylower, dy)
        assert np.array_equal(rad, aux)
        got = S.src2_axisym(q, aux[0], mbc=2, dt=dt, gamma=gamma)
        assert np.array_equal(got, out)
        assert np.array_equal(got[4], q[4])            # the tracer has no source
        assert np.array_equal(got[:, :2], q[:, :2])    # ghost cells untouched


def test_planted_mistake_is_caught():
    q, aux, out, (ylower, dy, dt, gamma) = _cases()[1]

    def wrong(q, rad, *, mbc, dt, gamma):
        good = S.src2_axisym(q, rad, mbc=mbc, dt=dt, gamma=gamma)
        # second stage with the old energy in the flux v (E + p)
        sl = (slice(mbc, -mbc), slice(mbc, -mbc))
        r = rad[sl]
        q1, q2, q3, q4 = (np.asarray(q[m])[sl] for m in range(4))
        bad = good.copy()
        bad[3][sl] = q4 - dt / r * (q3 / q1) * (q4 + 0.4 * (q4 - 0.5 * (q2 * q2 + q3 * q3) / q1))
        return bad

    assert not np.array_equal(wrong(q, aux[0], mbc=2, dt=dt, gamma=gamma), out)


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref/libref_src2.so not built")
def test_src2_matches_live_reference():
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import refdrive as R
    rng = np.random.default_rng(5)
    for case in range(6):
        mx, my = (int(v) for v in rng.integers(2, 20, 2))
        ylower = float(rng.choice([0.0, rng.uniform(0, 1)]))
        dy = float(rng.uniform(1e-3, 0.1))
        dt = float(rng.uniform(1e-4, 1e-2))
        gamma = float(rng.uniform(1.1, 1.7))
        aux = R.setaux_axisym(mx=mx, my=my, mbc=2, xlower=0.0, ylower=ylower, dx=dy, dy=dy)
        rho = rng.uniform(0.1, 3.0, (my + 4, mx + 4))
        u, v = rng.normal(0, 1.0, (2, my + 4, mx + 4))
        p = rng.uniform(0.5, 5.0, (my + 4, mx + 4))
        q = np.stack([rho, rho * u, rho * v, p / (gamma - 1) + 0.5 * rho * (u * u + v * v), rho])
        ref = R.src2_axisym(q, aux, mx=mx, my=my, mbc=2, xlower=0.0, ylower=ylower, dx=dy, dy=dy, t=0.0, dt=dt,
                            gamma=gamma)
        assert np.array_equal(S.setaux_rad(mx, my, 2, ylower, dy), aux), case
        assert np.array_equal(S.src2_axisym(q, aux[0], mbc=2, dt=dt, gamma=gamma), ref), case


@pytest.mark.parametrize("gamma,pinf", [(1.4, 5.0), (1.4, 1.5698), (1.67, 3.0)])
def test_post_shock_state_is_rankine_hugoniot(gamma, pinf):
    _, _, rinf, vinf, einf = S.shock_states(gamma, pinf)
    # shock moving into rho = 1, u = 0, p = 1 with speed s: mass, momentum and
    # energy jumps s [U] = [F(U)]
    r1, u1, p1 = 1.0, 0.0, 1.0
    r2, u2, p2 = rinf, vinf, pinf
    s = (r2 * u2 - r1 * u1) / (r2 - r1)
    e1 = p1 / (gamma - 1) + 0.5 * r1 * u1 * u1
    e2 = p2 / (gamma - 1) + 0.5 * r2 * u2 * u2
    assert abs(e2 - einf) < 1e-12 * e2
    assert abs(s * (r2 * u2 - r1 * u1) - ((r2 * u2 * u2 + p2) - (r1 * u1 * u1 + p1))) < 1e-12 * p2
    assert abs(s * (e2 - e1) - (u2 * (e2 + p2) - u1 * (e1 + p1))) < 1e-12 * e2 * abs(s)
    # the shock Mach number of that pressure ratio
    mach = math.sqrt(1 + (gamma + 1) / (2 * gamma) * (pinf - 1))
    assert abs(s / math.sqrt(gamma) - mach) < 1e-12


def test_qinit_regions():
    prm = (1.4, 0.5, 0.0, 0.2, 0.1, 5.0)
    mx = my = 16
    dx = dy = 0.5 / 32
    q = S.shockbubble_qinit(prm, 0.25, 0.0, dx, dy, mx, my)
    qin, qout, rinf, vinf, einf = S.shock_states(1.4, 5.0)
    xc_low = 0.25 + (np.arange(-2, mx + 2)) * dx
    yc_low = 0.0 + (np.arange(-2, my + 2)) * dy
    for ii, xl in enumerate(xc_low):
        for jj, yl in enumerate(yc_low):
            interior = 2 <= jj < my + 2
            corners = [(xl + a * dx - 0.5) ** 2 + (yl + b * dy) ** 2 < 0.04 for a in (0, 1) for b in (0, 1)]
            if xl < 0.2:
                assert q[0, jj, ii] == rinf and q[3, jj, ii] == einf
                assert q[4, jj, ii] == (2.0 if interior else 1.0)
            elif all(corners):
                assert q[0, jj, ii] == 0.1 and q[4, jj, ii] == (2.0 if interior and xl < 0.5 else 1.0)
            elif not any(corners):
                assert q[0, jj, ii] == 1.0 and q[4, jj, ii] == 0.0
            else:
                assert 0.1 < q[0, jj, ii] < 1.0
            assert q[1, jj, ii] == (rinf * vinf if xl < 0.2 else 0.0) and q[2, jj, ii] == 0.0
