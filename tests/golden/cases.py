"""Seeded synthetic patch inputs shared by the golden-vector generator and the
parity tests.  Each case is a dict with the tile q (meqn, NY, NX), optional
aux, grid spacing and a time step at roughly CFL 0.9.

The states are built to visit every branch of the reference solvers:
transonic rarefactions (entropy fix), left/right-going shocks, contact and
shear waves, zero waves (limiter skip) and both sweep directions.
"""
from __future__ import annotations

import math

import numpy as np


def _grid(mx, my, mbc, xlower, ylower, dx, dy):
    i = np.arange(mx + 2 * mbc) - mbc + 0.5
    j = np.arange(my + 2 * mbc) - mbc + 0.5
    x = xlower + i * dx
    y = ylower + j * dy
    return np.meshgrid(x, y)  # (NY, NX)


def swe_state(rng, mx, my, mbc, *, dam=True, transonic=True):
    dx = dy = 1.0 / mx
    X, Y = _grid(mx, my, mbc, 0.0, 0.0, dx, dy)
    h = 1.0 + 0.3 * np.sin(2 * math.pi * (X + rng.uniform())) * np.cos(2 * math.pi * Y)
    h += 0.05 * rng.standard_normal(h.shape)
    if dam:
        cx, cy, r = rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7), rng.uniform(0.15, 0.35)
        inside = (X - cx) ** 2 + (Y - cy) ** 2 < r * r
        h = np.where(inside, 2.0 + 0.1 * rng.uniform(), h)
    if transonic:
        # very shallow strip next to deep water: transonic rarefactions
        h = np.where((X > 0.78) & (X < 0.9), 0.08, h)
        h = np.where((Y > 0.1) & (Y < 0.18), 0.06, h)
    u = 0.4 * rng.standard_normal(h.shape)
    v = 0.4 * rng.standard_normal(h.shape)
    q = np.stack([h, h * u, h * v])
    return q, dx, dy


def swe_dt(q, dx, dy, grav=1.0, cfl=0.9):
    h = q[0]
    u = np.abs(q[1] / h) + np.sqrt(grav * h)
    v = np.abs(q[2] / h) + np.sqrt(grav * h)
    return cfl * min(dx / u.max(), dy / v.max())


def euler_state(rng, mx, my, mbc, gamma=1.4):
    """Shock + bubble + random perturbation, plus a strong expansion region."""
    dx = 2.0 / (4 * mx)
    dy = 0.5 / my
    X, Y = _grid(mx, my, mbc, 0.05, 0.0, dx, dy)
    g1 = gamma - 1.0
    pinf = 1.5698
    rinf = (g1 + (gamma + 1) * pinf) / ((gamma + 1) + g1 * pinf)
    vinf = (1.0 / math.sqrt(gamma)) * (pinf - 1.0) / math.sqrt(
        0.5 * ((gamma + 1) / gamma) * pinf + 0.5 * g1 / gamma)
    rho = np.ones_like(X)
    u = np.zeros_like(X)
    v = np.zeros_like(X)
    p = np.ones_like(X)
    tr = np.zeros_like(X)
    shock = X < 0.12
    rho = np.where(shock, rinf, rho)
    u = np.where(shock, vinf, u)
    p = np.where(shock, pinf, p)
    bub = (X - 0.2) ** 2 + (Y - 0.25) ** 2 < 0.08 ** 2
    rho = np.where(bub, 0.138, rho)
    tr = np.where(bub, 1.0, tr)
    # strong expansion: fluid moving apart at high speed
    exp_ = (X > 0.26) & (X < 0.3)
    u = np.where(exp_, np.where(X < 0.28, -2.0, 2.0), u)
    rho = rho * (1.0 + 0.05 * rng.standard_normal(rho.shape))
    u = u + 0.1 * rng.standard_normal(u.shape)
    v = v + 0.1 * rng.standard_normal(v.shape)
    p = p * (1.0 + 0.05 * rng.standard_normal(p.shape))
    E = p / g1 + 0.5 * rho * (u * u + v * v)
    q = np.stack([rho, rho * u, rho * v, E, rho * tr])
    return q, dx, dy


def euler_dt(q, dx, dy, gamma=1.4, cfl=0.9):
    rho = q[0]
    u = q[1] / rho
    v = q[2] / rho
    p = (gamma - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v))
    c = np.sqrt(gamma * p / rho)
    return cfl * min(dx / (np.abs(u) + c).max(), dy / (np.abs(v) + c).max())


def adv_state(rng, mx, my, mbc, xlower=0.0, ylower=0.0):
    dx = dy = 1.0 / (8 * mx)
    X, Y = _grid(mx, my, mbc, xlower, ylower, dx, dy)
    q = np.where(X < xlower + 0.5 * 8 * mx * dx * rng.uniform(0.3, 0.7), 1.0, 0.0)
    q = q + 0.2 * rng.standard_normal(q.shape) * (rng.uniform(size=q.shape) < 0.2)
    q[:, : (mx + 2 * mbc) // 3] += 0.5 * np.sin(7 * Y[:, : (mx + 2 * mbc) // 3])
    return q[None], dx, dy


def adv_dt(aux, dx, dy, cfl=0.9):
    return cfl * min(dx / max(np.abs(aux[0]).max(), 1e-12),
                     dy / max(np.abs(aux[1]).max(), 1e-12))


def fw_state(rng, mx, my, mbc, *, per_edge=False, lengths=False):
    """Conservative transport with f-waves on a Cartesian grid: a density
    with a jump and a smooth part, aux (7, NY, NX) as the transport
    applications' setaux + b4step2 give it on an unmapped grid: capacity 1,
    the normal velocities of a cell's left/right/bottom/top edges (its
    rotating velocity (u, v), or independent values per edge with
    per_edge), edge lengths 1 (or perturbed with lengths)."""
    dx, dy = 1.0 / mx, 1.0 / my
    X, Y = _grid(mx, my, mbc, 0.0, 0.0, dx, dy)
    q = 0.5 + 0.3 * np.sin(2 * math.pi * (X + rng.uniform())) * np.cos(2 * math.pi * Y)
    q = np.where(X < rng.uniform(0.3, 0.7), q + 1.0, q)[None]
    u = np.sin(math.pi * X) ** 2 * np.sin(2 * math.pi * Y) + 0.2 * rng.standard_normal(X.shape)
    v = -np.sin(math.pi * Y) ** 2 * np.sin(2 * math.pi * X) + 0.2 * rng.standard_normal(X.shape)
    aux = np.zeros((7,) + X.shape)
    aux[0] = 1.0
    aux[1] = aux[2] = u
    aux[3] = aux[4] = v
    if per_edge:
        aux[2] = u + 0.1 * rng.standard_normal(X.shape)
        aux[4] = v + 0.1 * rng.standard_normal(X.shape)
    aux[5] = aux[6] = 1.0
    if lengths:
        aux[5] = rng.uniform(0.9, 1.1, X.shape)
        aux[6] = rng.uniform(0.9, 1.1, X.shape)
    return q, aux, dx, dy


def fw_dt(aux, dx, dy, cfl=0.9):
    smax = max(np.abs(aux[1:3]).max() * aux[5].max() / dx, np.abs(aux[3:5]).max() * aux[6].max() / dy)
    return cfl / smax
