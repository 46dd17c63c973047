"""CPU checks of the oracle's conservative time synchronisation
(oracle/claw46_timesync.c, store_flux / accumulate_waves in claw46_oracle.c,
the sync driver in oracle/amr.py).

The corrections exist to make subcycled AMR conservative: with wall
boundaries the total mass over the leaves must stay constant to rounding
when time sync is on, while without it mass drifts at the coarse/fine
interfaces.  A wrong sign, register, face or pair index breaks the first
property.  The SWE flux used for the edge registers is checked against the
textbook shallow-water flux.
"""
import numpy as np
import pytest

from oracle import amr as A
from oracle import oracle as O


def test_swe_flux_is_the_shallow_water_flux():
    rng = np.random.default_rng(3)
    g = 9.81
    for _ in range(20):
        h, hu, hv = rng.uniform(0.1, 3), rng.normal(), rng.normal()
        fx = O.swe_flux(0, g, [h, hu, hv])
        fy = O.swe_flux(1, g, [h, hu, hv])
        np.testing.assert_allclose(fx, [hu, hu * hu / h + g * h * h / 2, hu * hv / h], rtol=1e-15)
        np.testing.assert_allclose(fy, [hv, hu * hv / h, hv * hv / h + g * h * h / 2], rtol=1e-15)


def test_store_flux_registers_hold_dt_el_f_at_the_edge_cells():
    """Edge-flux blocks E0 / E1 after one step: dt*dy*f_x(q) at the first
    interior / ghost column, dt*dx*f_y(q) at the first interior / ghost row."""
    import cases as C
    mx, my, mbc = 8, 12, 2
    q, dx, dy = C.swe_state(np.random.default_rng(5), mx, my, mbc)
    q = np.ascontiguousarray(q)
    cfg = O.make_cfg(O.SOLVER_SWE_ROE, mx, my, mbc, grav=1.0)
    reg = np.zeros(O.reg_size(mx, my, 3))
    dt = 0.2 * dx
    q0 = q.copy()
    O.step2_ts(cfg, q, dx, dy, dt, reg)

    def block(face, b):
        L = my if face < 2 else mx
        off = sum(4 * 3 * (my if f < 2 else mx) for f in range(face)) + b * 3 * L
        return reg[off:off + 3 * L].reshape(3, L)

    def fx(col):
        return np.stack([O.swe_flux(0, 1.0, q0[:, mbc + j, col]) for j in range(my)], 1)

    def fy(row):
        return np.stack([O.swe_flux(1, 1.0, q0[:, row, mbc + i]) for i in range(mx)], 1)

    assert np.array_equal(block(0, 2), dt * dy * fx(mbc))
    assert np.array_equal(block(0, 3), dt * dy * fx(mbc - 1))
    assert np.array_equal(block(1, 2), dt * dy * fx(mbc + mx - 1))
    assert np.array_equal(block(1, 3), dt * dy * fx(mbc + mx))
    assert np.array_equal(block(2, 2), dt * dx * fy(mbc))
    assert np.array_equal(block(2, 3), dt * dx * fy(mbc - 1))
    assert np.array_equal(block(3, 2), dt * dx * fy(mbc + my - 1))
    assert np.array_equal(block(3, 3), dt * dx * fy(mbc + my))
    # fluctuation blocks are filled too
    assert all(np.abs(block(f, b)).max() > 0 for f in range(4) for b in (0, 1))


def _dam_amr(refine, *, time_sync, subcycle, mx=8, lo=2):
    leaves = []
    for gy in range(1 << lo):
        for gx in range(1 << lo):
            if (lo, gx, gy) in refine:
                leaves += [(lo + 1, 2 * gx + (c & 1), 2 * gy + (c >> 1)) for c in range(4)]
            else:
                leaves.append((lo, gx, gy))
    geom = np.array([[gx / (1 << L), gy / (1 << L), 1 / (1 << L) / mx, 1 / (1 << L) / mx]
                     for L, gx, gy in leaves])
    cfg = O.make_cfg(O.SOLVER_SWE_ROE, mx, mx, 2, grav=1.0)
    amr = A.AmrOracle(cfg, leaves, geom, (3, 3, 3, 3), subcycle=subcycle, time_sync=time_sync)
    for k in range(len(leaves)):
        x0, y0, dx, dy = geom[k]
        xc = x0 + (np.arange(-2, mx + 2) + 0.5) * dx
        yc = y0 + (np.arange(-2, mx + 2) + 0.5) * dy
        X, Y = np.meshgrid(xc, yc)
        amr.q[k][0] = 1.0 + 1.0 * (np.hypot(X - 0.45, Y - 0.4) < 0.2)
    amr.ghost_update(min(amr.levels), max(amr.levels), 0)
    return amr, geom


def _mass(amr, geom):
    return sum(amr.q[k][0, 2:-2, 2:-2].sum() * geom[k, 2] * geom[k, 3] for k in range(len(amr.q)))


@pytest.mark.parametrize("subcycle", [True, False])
@pytest.mark.parametrize("refine", [((2, 1, 1), (2, 2, 1)),
                                    ((2, 1, 1), (2, 2, 1), (2, 1, 2), (2, 2, 2))])
def test_time_sync_makes_amr_conservative(subcycle, refine):
    dm = {}
    for ts in (True, False):
        amr, geom = _dam_amr(refine, time_sync=ts, subcycle=subcycle)
        m0 = _mass(amr, geom)
        dt = 0.4 * geom[:, 2].max() / 1.5
        for s in range(6):
            amr.advance(s * dt, dt)
        dm[ts] = abs(_mass(amr, geom) - m0) / m0
    assert dm[True] <= 2e-15, dm
    assert dm[False] >= 1e-5, dm
    # all registers are cleared after a full coarse step (every face synchronised)
    amr, geom = _dam_amr(refine, time_sync=True, subcycle=subcycle)
    amr.advance(0.0, 0.4 * geom[:, 2].max() / 1.5)
    assert all(not r.any() for r in amr.reg)
