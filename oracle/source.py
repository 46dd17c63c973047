"""CPU restatement of the shock-bubble application's source term and aux
(applications/clawpack/euler/2d/shockbubble/user_4.6/src2.f:1-77,
setaux.f:1-22) and of its initial condition (user_4.6/qinit.f:1-64 with
setprob.f:1-58 and the idisc = 2 circle of fdisc.f:36-38).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product path
never imports it.

numpy float64 element-wise operations round once each (no fused
multiply-add), and the expressions keep the reference's operation order, so
the results compare bitwise with the reference Fortran built without FMA
contraction and with the device kernels.

Pinned by: the reference's src2.f and setaux.f compiled unchanged into
oracle/_ref (golden vectors in tests/golden/src2/ and live random cases,
tests/test_source_oracle.py); the qinit's post-shock state by the
Rankine-Hugoniot conditions of a shock with pressure ratio pinf into gas at
rest (rho = p = 1); the bubble average by the polygon-area checks of
oracle/cellave.py.
"""
from __future__ import annotations

import math

import numpy as np

from . import cellave as CA


def setaux_rad(mx, my, mbc, ylower, dy):
    """setaux.f:14-19: aux(i,j,1) = ylower + (j - 0.5) dy over the whole tile."""
    j = np.arange(1 - mbc, my + mbc + 1, dtype=np.float64)
    yj = ylower + (j - 0.5) * dy
    return np.repeat(yj[:, None], mx + 2 * mbc, axis=1)[None]


def src2_axisym(q, rad, *, mbc, dt, gamma):
    """src2.f:21-72 on the interior cells of q (meqn >= 4, tile layout
    [meqn, my+2mbc, mx+2mbc]); rad: [my+2mbc, mx+2mbc].  Returns a new array."""
    q = np.array(q, dtype=np.float64)
    gamma1 = gamma - 1.0
    sl = (slice(mbc, q.shape[1] - mbc), slice(mbc, q.shape[2] - mbc))
    r = rad[sl]
    q1, q2, q3, q4 = (q[m][sl].copy() for m in range(4))
    dt2 = dt / 2.0
    ndim = 2
    rho = q1
    u = q2 / q1
    v = q3 / q1
    press = gamma1 * (q4 - 0.5 * rho * (u * u + v * v))
    f = dt2 * (ndim - 1)
    s1 = q1 - f / r * q3
    s2 = q2 - f / r * (rho * u * v)
    s3 = q3 - f / r * (rho * v * v)
    s4 = q4 - f / r * v * (q4 + press)
    rho = s1
    u = s2 / s1
    v = s3 / s1
    press = gamma1 * (s4 - 0.5 * rho * (u * u + v * v))
    g = dt * (ndim - 1)
    q[0][sl] = q1 - g / r * s3
    q[1][sl] = q2 - g / r * (rho * u * v)
    q[2][sl] = q3 - g / r * (rho * v * v)
    q[3][sl] = q4 - g / r * v * (s4 + press)
    return q


def shock_states(gamma, pinf):
    """setprob.f:40-58: (qin, qout, rinf, vinf, einf)."""
    gamma1 = gamma - 1.0
    qin = [None, 0.0, 0.0, 1.0 / gamma1, 1.0]
    qout = [1.0, 0.0, 0.0, 1.0 / gamma1, 0.0]
    rinf = (gamma1 + (gamma + 1) * pinf) / ((gamma + 1) + gamma1 * pinf)
    vinf = (1.0 / math.sqrt(gamma)) * (pinf - 1.0) / math.sqrt(
        0.5 * ((gamma + 1) / gamma) * pinf + 0.5 * gamma1 / gamma)
    einf = 0.5 * rinf * vinf * vinf + pinf / gamma1
    return qin, qout, rinf, vinf, einf


def shockbubble_qinit(prm, xlower, ylower, dx, dy, mx, my, mbc=2):
    """qinit.f:23-64 for idisc = 2, meqn = 5; prm = (gamma, x0, y0, r0, rhoin,
    pinf).  Returns [5, my+2mbc, mx+2mbc]."""
    gamma, x0, y0, r0, rhoin, pinf = prm
    qin, qout, rinf, vinf, einf = shock_states(gamma, pinf)
    qin[0] = rhoin
    q = np.zeros((5, my + 2 * mbc, mx + 2 * mbc))
    fd = lambda x, y: CA.fdisc(x0, y0, r0, x, y)  # noqa: E731
    for ii, i in enumerate(range(1 - mbc, mx + mbc + 1)):
        xclow = xlower + (i - 1) * dx
        for jj, j in enumerate(range(1 - mbc, my + mbc + 1)):
            yclow = ylower + (j - 1) * dy
            win = CA.cellave2(fd, xclow, yclow, dx, dy)
            for m in range(5):
                q[m, jj, ii] = win * qin[m] + (1.0 - win) * qout[m]
        if xclow < 0.2:
            q[:, :, ii] = np.array([rinf, rinf * vinf, 0.0, einf, 1.0])[:, None]
        if xclow < 0.5:
            q[4, mbc:mbc + my, ii] = 2.0 * q[4, mbc:mbc + my, ii]
    return q
