"""CPU checks of the oracle's HLLE Euler solver (orc_rpn2_hlle_face in
oracle/claw46_oracle.c), the two-wave Harten-Lax-van Leer solver with
Einfeldt's wave-speed bounds (LeVeque, Finite Volume Methods for Hyperbolic
Problems, section 15.3.7; Einfeldt, SIAM J. Numer. Anal. 25 (1988)).  The
reference has no HLLE solver (applications/clawpack/euler/2d/rp/README), so
these properties of the method are what pins it:

* consistency: equal states give zero waves and fluctuations, exactly;
* conservation: amdq + apdq = f(q_b) - f(q_a) for the textbook Euler flux;
* upwinding: supersonic flow to the right (left) gives amdq = 0 (apdq = 0);
* speeds: Einfeldt's s1 = min(u_a - c_a, u_roe - c_roe) and
  s2 = max(u_b + c_b, u_roe + c_roe), with the Roe average computed here;
* positivity: the HLL middle state of a strong symmetric rarefaction (the
  "1-2-3" problem, where a linearised Roe state has negative pressure) has
  positive density and pressure;
* symmetry: a y face is an x face with the two momenta swapped.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

GAMMA = 1.4


def prim(rho, u, v, p, phi=0.0, g=GAMMA):
    E = p / (g - 1) + 0.5 * rho * (u * u + v * v)
    return np.array([rho, rho * u, rho * v, E, rho * phi])


def flux_x(q, g=GAMMA):
    rho, mu, mv, E, t = q
    u = mu / rho
    p = (g - 1) * (E - 0.5 * (mu * mu + mv * mv) / rho)
    return np.array([mu, mu * u + p, mv * u, u * (E + p), t * u]), u, math.sqrt(g * p / rho)


def pressure(q, g=GAMMA):
    return (g - 1) * (q[3] - 0.5 * (q[1] ** 2 + q[2] ** 2) / q[0])


def random_state(rng):
    return prim(rng.uniform(0.2, 3), rng.normal(), rng.normal(), rng.uniform(0.2, 3), rng.uniform())


def test_equal_states_give_zero_waves():
    rng = np.random.default_rng(1)
    for _ in range(50):
        q = random_state(rng)
        for ixy in (1, 2):
            w, s, am, ap = O.rpn2_hlle_face(GAMMA, ixy, q, q)
            assert not w.any() and not am.any() and not ap.any()
            assert s[0] < s[1]


def test_fluctuations_sum_to_the_flux_difference():
    rng = np.random.default_rng(2)
    for _ in range(200):
        qa, qb = random_state(rng), random_state(rng)
        w, s, am, ap = O.rpn2_hlle_face(GAMMA, 1, qa, qb)
        df = flux_x(qb)[0] - flux_x(qa)[0]
        scale = np.abs(flux_x(qa)[0]).max() + np.abs(flux_x(qb)[0]).max()
        np.testing.assert_allclose(am + ap, df, rtol=0, atol=1e-14 * scale)
        # the waves add up to the jump and carry s1, s2
        np.testing.assert_allclose(w[0] + w[1], qb - qa, rtol=0, atol=1e-14 * np.abs(qb - qa).max())
        np.testing.assert_allclose(s[0] * w[0] + s[1] * w[1], df, rtol=0, atol=1e-13 * scale)


def test_supersonic_flow_is_upwinded():
    for u, side in ((3.0, "right"), (-3.0, "left")):
        qa, qb = prim(1.0, u, 0.2, 1.0, 0.3), prim(0.8, u * 1.1, -0.1, 0.9, 0.7)
        w, s, am, ap = O.rpn2_hlle_face(GAMMA, 1, qa, qb)
        df = flux_x(qb)[0] - flux_x(qa)[0]
        if side == "right":
            assert s[0] > 0 and not am.any()
            np.testing.assert_allclose(ap, df, rtol=1e-14, atol=1e-14)
        else:
            assert s[1] < 0 and not ap.any()
            np.testing.assert_allclose(am, df, rtol=1e-14, atol=1e-14)


def test_speeds_are_einfeldts_bounds():
    rng = np.random.default_rng(3)
    for _ in range(200):
        qa, qb = random_state(rng), random_state(rng)
        _, s, _, _ = O.rpn2_hlle_face(GAMMA, 1, qa, qb)
        _, ua, ca = flux_x(qa)
        _, ub, cb = flux_x(qb)
        # Roe average velocity and sound speed
        ra, rb = math.sqrt(qa[0]), math.sqrt(qb[0])
        u = (qa[1] / ra + qb[1] / rb) / (ra + rb)
        v = (qa[2] / ra + qb[2] / rb) / (ra + rb)
        H = ((qa[3] + pressure(qa)) / ra + (qb[3] + pressure(qb)) / rb) / (ra + rb)
        c = math.sqrt((GAMMA - 1) * (H - 0.5 * (u * u + v * v)))
        np.testing.assert_allclose(s, [min(ua - ca, u - c), max(ub + cb, u + c)], rtol=1e-14)
        assert s[0] < u - c + 1e-14 < u + c <= s[1] + 1e-14


@pytest.mark.parametrize("u", [2.0, 3.0, 5.0])
def test_strong_rarefaction_keeps_the_middle_state_positive(u):
    qa, qb = prim(1.0, -u, 0.0, 0.4), prim(1.0, u, 0.0, 0.4)
    w, s, am, ap = O.rpn2_hlle_face(GAMMA, 1, qa, qb)
    qm = qa + w[0]
    assert qm[0] > 0 and pressure(qm) > 0
    # a linearisation about the Roe average gives a negative pressure here
    # (Einfeldt et al., J. Comput. Phys. 92 (1991)): the intermediate Roe
    # state after the 1-wave, rho_1 = rho_a + alpha_1, with
    # alpha_1 = (dp - rho c du) / (2 c^2) at the Roe average
    ra = rb = 1.0
    Hm = ((qa[3] + pressure(qa)) + (qb[3] + pressure(qb))) / 2
    c2 = (GAMMA - 1) * (Hm - 0.0)
    c = math.sqrt(c2)
    alpha1 = (0.0 - 1.0 * c * (2 * u)) / (2 * c2)
    q1 = qa + alpha1 * np.array([1, -c, 0, Hm - 0 * c, 0])  # eigenvector (1, u - c, v, H - u c) at u = 0
    if u >= 3.0:
        assert q1[0] < 0 or pressure(q1) < 0


def test_y_face_is_the_x_face_with_momenta_swapped():
    rng = np.random.default_rng(4)
    sw = [0, 2, 1, 3, 4]
    for _ in range(30):
        qa, qb = random_state(rng), random_state(rng)
        wx, sx, amx, apx = O.rpn2_hlle_face(GAMMA, 1, qa, qb)
        wy, sy, amy, apy = O.rpn2_hlle_face(GAMMA, 2, qa[sw], qb[sw])
        assert np.array_equal(sx, sy)
        assert np.array_equal(wy[:, sw], wx) and np.array_equal(amy[sw], amx) and np.array_equal(apy[sw], apx)


def _roe_state(qa, qb):
    ra, rb = math.sqrt(qa[0]), math.sqrt(qb[0])
    u = (qa[1] / ra + qb[1] / rb) / (ra + rb)
    v = (qa[2] / ra + qb[2] / rb) / (ra + rb)
    H = ((qa[3] + pressure(qa)) / ra + (qb[3] + pressure(qb)) / rb) / (ra + rb)
    return u, v, H


def jacobian_y(u, v, H, g=GAMMA):
    """Textbook Jacobian of the y flux of the Euler equations in (rho, rho u,
    rho v, E), written with the velocity and the total enthalpy (Toro, Riemann
    Solvers and Numerical Methods for Fluid Dynamics, section 3.2)."""
    g1 = g - 1
    k = 0.5 * (u * u + v * v)
    return np.array([[0, 0, 1, 0],
                     [-u * v, v, u, 0],
                     [g1 * k - v * v, -g1 * u, (3 - g) * v, g1],
                     [v * (g1 * k - H), -g1 * u * v, H - g1 * v * v, g * v]])


def test_transverse_split_sums_to_the_roe_jacobian():
    """bm + bp = B asdq, B the y-flux Jacobian at the Roe average of an x
    face (the tracer is carried with the transverse velocity); subsonic
    splits send parts both ways, supersonic ones one way."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        qa, qb = random_state(rng), random_state(rng)
        d = rng.normal(size=5)
        bm, bp = O.rpt2_hlle_face(GAMMA, 1, qa, qb, d)
        u, v, H = _roe_state(qa, qb)
        Bd = np.append(jacobian_y(u, v, H) @ d[:4], v * d[4])
        np.testing.assert_allclose(bm + bp, Bd, rtol=0, atol=1e-12 * (1 + np.abs(Bd).max()))
    qa, qb = prim(1.0, 0.1, 3.0, 1.0), prim(1.1, 0.0, 3.2, 1.0)  # v > c: everything goes up
    bm, bp = O.rpt2_hlle_face(GAMMA, 1, qa, qb, rng.normal(size=5))
    assert not bm.any() and np.abs(bp).max() > 0


def test_contact_fluctuations_spread_transversely():
    """A fluctuation the Roe split moves with zero transverse speed (a
    density jump at rest) is spread both ways by the HLL form."""
    qa, qb = prim(1.0, 0.0, 0.0, 1.0), prim(0.2, 0.0, 0.0, 1.0)
    d = np.array([0.5, 0.0, 0.0, 0.0, 0.0])
    bm, bp = O.rpt2_hlle_face(GAMMA, 1, qa, qb, d)
    assert bm[0] < 0 < bp[0]


def test_shock_bubble_stays_stable_at_courant_0_7():
    """The C3 initial data (Mach 1.22 shock, density 0.138 bubble) on a
    uniform level-3 4x1 brick, 12 subcycled steps at C3's time step (Courant
    number 0.71): the HLLE pair stays bounded with positive density and
    pressure.  With the Roe transverse split paired to the HLLE normal solver
    the same run blows up after 7 steps (checked when the pair was built)."""
    import sys
    from oracle import amr as A
    from forestclaw_amd import configs as CF
    g, xs, rinf, vinf, pinf, x0, y0, r0, rhob = CF.c3_params()
    mx, L = 16, 3
    n = 1 << L
    leaves, geom = [], []
    for bx in range(4):
        for gy in range(n):
            for gx in range(n):
                h = 0.5 / n
                leaves.append((L, bx * n + gx, gy))
                geom.append([bx * 0.5 + gx * h, gy * h, h / mx, h / mx])
    geom = np.array(geom)
    cfg = O.make_cfg(O.SOLVER_EULER_HLLE, mx, mx, 2, gamma=g)
    amr = A.AmrOracle(cfg, leaves, geom, (1, 1, 1, 1), brick=(4, 1), minlevel=L, maxlevel=L)
    for k in range(len(leaves)):
        xl, yl, dx, dy = geom[k]
        X, Y = np.meshgrid(xl + (np.arange(-2, mx + 2) + 0.5) * dx, yl + (np.arange(-2, mx + 2) + 0.5) * dy)
        rho, u, p, tr = np.ones_like(X), np.zeros_like(X), np.ones_like(X), np.zeros_like(X)
        sh = X < xs
        rho[sh], u[sh], p[sh] = rinf, vinf, pinf
        bub = (~sh) & ((X - x0) ** 2 + (Y - y0) ** 2 < r0 * r0)
        rho[bub], tr[bub] = rhob, 1.0
        amr.q[k][...] = np.stack([rho, rho * u, 0 * rho, p / (g - 1) + 0.5 * rho * u * u, rho * tr])
    amr.ghost_update(L, L, 0)
    dt = CF.c3_dt(L)
    cfl = [amr.advance(s * dt, dt) for s in range(12)]
    assert max(cfl) < 0.75, cfl
    qs = np.stack(amr.q)[:, :, 2:-2, 2:-2]
    assert qs[:, 0].min() > 0.1 and np.array([pressure(q) for q in qs.transpose(0, 2, 3, 1).reshape(-1, 5)]).min() > 0.5
