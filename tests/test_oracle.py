"""CPU tests of the FP64 oracle (oracle/claw46_oracle.c).

The oracle is pinned three ways:
  * bitwise against golden vectors written by the reference's own Fortran
    (tests/golden/*.npz, see tests/golden/make_golden.py), and live against
    oracle/_ref when it is built in this container;
  * against closed forms and invariants that a plausible slip would break
    (constant states, lake at rest, limiter formulas, 1-D reduction,
    conservation of the flux-difference form);
  * planted mistakes (entropy fix off, wrong limiter, no transverse terms)
    must be detected by the golden vectors, which shows the fixtures visit
    those branches.
"""
import glob
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz")))


def _golden(prefix):
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, prefix + "*.npz")))


def _cfg_for(name, g, **over):
    solver = {"swe": O.SOLVER_SWE_ROE, "euler": O.SOLVER_EULER5_ROE, "adv": O.SOLVER_ADV,
              "fw": O.SOLVER_TRANSPORT_FW}[name.split("_")[0]]
    kw = dict(order=tuple(int(x) for x in g["order"]), mthlim=[int(x) for x in g["mthlim"]])
    if solver == O.SOLVER_SWE_ROE:
        kw["grav"] = float(g["grav"])
    if solver == O.SOLVER_EULER5_ROE:
        kw["gamma"] = float(g["gamma"])
    kw.update(over)
    return O.make_cfg(solver, int(g["mx"]), int(g["my"]), int(g["mbc"]), **kw)


STEP_CASES = (_golden("swe_step") + _golden("euler_step") + _golden("adv_step")
              + _golden("fw_step"))


def test_golden_fixtures_present():
    assert len(_golden("swe_step")) >= 6 and len(_golden("euler_step")) >= 2
    assert len(_golden("adv_step")) >= 3 and len(_golden("fw_step")) >= 4


@pytest.mark.parametrize("name", STEP_CASES)
def test_oracle_matches_reference_golden_bitwise(name):
    g = _load(name)
    cfg = _cfg_for(name, g)
    qn, cfl = O.step2(cfg, g["q"], g.get("aux"), float(g["dx"]), float(g["dy"]), float(g["dt"]))
    assert np.array_equal(qn, g["qout"]), np.abs(qn - g["qout"]).max()
    assert cfl == float(g["cfl"])


def test_planted_no_entropy_fix_is_detected():
    hit = 0
    for name in _golden("swe_step") + _golden("euler_step"):
        g = _load(name)
        qn, _ = O.step2(_cfg_for(name, g, efix=0), g["q"], None, float(g["dx"]), float(g["dy"]),
                        float(g["dt"]))
        hit += not np.array_equal(qn, g["qout"])
    assert hit >= 6  # the fixtures contain transonic rarefactions


def test_planted_wrong_limiter_is_detected():
    for name in _golden("swe_step")[:2]:
        g = _load(name)
        qn, _ = O.step2(_cfg_for(name, g, mthlim=[1, 1, 1]), g["q"], None, float(g["dx"]),
                        float(g["dy"]), float(g["dt"]))
        assert not np.array_equal(qn, g["qout"])


def test_planted_dropped_transverse_is_detected():
    g = _load("swe_step_0")
    qn, _ = O.step2(_cfg_for("swe_step_0", g, order=(2, 0)), g["q"], None, float(g["dx"]),
                    float(g["dy"]), float(g["dt"]))
    assert not np.array_equal(qn, g["qout"])


def test_bc2_matches_reference_golden():
    g = _load("bc2")
    for mthbc, q, qout in zip(g["mthbc"], g["q"], g["qout"]):
        got = O.bc2(q, int(g["mx"]), int(g["my"]), int(g["mbc"]), [int(x) for x in mthbc])
        assert np.array_equal(got, qout)


def test_swirl_aux_matches_reference_golden():
    for name in _golden("adv_step"):
        g = _load(name)
        mx, my, mbc = int(g["mx"]), int(g["my"]), int(g["mbc"])
        aux = O.swirl_aux(mx, my, mbc, float(g["xlower"]), float(g["ylower"]), float(g["dx"]),
                          float(g["dy"]), float(g["t"]), float(g["auxdt"]), float(g["tperiod"]))
        assert np.array_equal(aux, g["aux"])


@pytest.mark.parametrize("mthlim", [1, 2, 3, 4, 5, 6])
def test_limiter_closed_forms(mthlim):
    th = float(np.float32(1.3))
    ref = {
        1: lambda r: max(0.0, min(1.0, r)),
        2: lambda r: max(0.0, min(1.0, 2 * r), min(2.0, r)),
        3: lambda r: (r + abs(r)) / (1 + abs(r)),
        4: lambda r: max(0.0, min((1 + r) / 2, 2.0, 2 * r)),
        5: lambda r: r,
        6: lambda r: max(0.0, min(th * r, (1 + r) / 2, th)),
    }[mthlim]
    for r in [-3.0, -0.5, 0.0, 0.25, 0.5, 1.0, 1.7, 2.5, 3.0, 10.0]:
        assert O.limiter_phi(mthlim, r) == pytest.approx(ref(r), rel=0, abs=1e-15)
    assert th != 1.3 and abs(th - 1.3) < 1e-7  # the single-precision literal


def test_constant_states_are_preserved():
    rng = np.random.default_rng(3)
    for solver, state in [(O.SOLVER_SWE_ROE, [1.7, 0.3, -0.2]),
                          (O.SOLVER_EULER5_ROE, [1.2, 0.3, -0.1, 3.0, 0.4])]:
        cfg = O.make_cfg(solver, 12, 10, 2)
        q = np.ones((len(state), 14, 16)) * np.array(state)[:, None, None]
        qn, cfl = O.step2(cfg, q, None, 0.1, 0.1, 0.01)
        assert np.array_equal(qn[:, 2:-2, 2:-2], q[:, 2:-2, 2:-2])
        assert cfl > 0
    # advection with a discretely divergence-free edge field from psi
    cfg = O.make_cfg(O.SOLVER_ADV, 16, 16, 2)
    aux = O.swirl_aux(16, 16, 2, 0.25, 0.25, 1 / 64, 1 / 64, 0.0, 0.0, 4.0)
    q = np.full((1, 20, 20), 0.75)
    qn, _ = O.step2(cfg, q, aux, 1 / 64, 1 / 64, 0.005)
    assert np.abs(qn[:, 2:-2, 2:-2] - 0.75).max() < 1e-14


def test_lake_at_rest_is_bitwise_steady():
    cfg = O.make_cfg(O.SOLVER_SWE_ROE, 16, 16, 2, grav=9.81)
    q = np.zeros((3, 20, 20))
    q[0] = 2.5
    qn, _ = O.step2(cfg, q, None, 0.05, 0.05, 0.004)
    assert np.array_equal(qn[:, 2:-2, 2:-2], q[:, 2:-2, 2:-2])


def test_advection_1d_reduction_matches_textbook_upwind_mc():
    """Uniform positive velocity u, v = 0, q constant in y: the step is the 1-D
    high-resolution upwind method with the MC limiter (LeVeque, Finite Volume
    Methods for Hyperbolic Problems, sec. 6.9-6.12), written out here from
    the textbook formula."""
    mx, mbc = 24, 2
    u, dx, dt = 0.8, 0.1, 0.09
    nu = u * dt / dx
    rng = np.random.default_rng(5)
    prof = np.concatenate([np.zeros(8), np.ones(10), rng.uniform(0, 1, mx + 2 * mbc - 18)])
    q = np.tile(prof, (mx + 2 * mbc, 1))[None]
    aux = np.zeros((2, mx + 2 * mbc, mx + 2 * mbc))
    aux[0] = u
    cfg = O.make_cfg(O.SOLVER_ADV, mx, mx, mbc)
    qn, cfl = O.step2(cfg, q, aux, dx, dx, dt)
    assert cfl == pytest.approx(nu, rel=1e-15)

    Q = prof
    dQ = np.diff(Q)  # dQ[k] = Q[k+1]-Q[k] = wave at face k+1

    def phi(r):
        return max(0.0, min((1 + r) / 2, 2.0, 2 * r))

    ref = Q.copy()
    N = len(Q)
    for i in range(mbc, N - mbc):
        def F(face):  # limited flux correction at face between face-1 and face
            w = dQ[face - 1]
            if w == 0:
                return 0.5 * u * (1 - nu) * w
            r = dQ[face - 2] / w
            return 0.5 * u * (1 - nu) * phi(r) * w
        ref[i] = Q[i] - nu * (Q[i] - Q[i - 1]) - dt / dx * (F(i + 1) - F(i))
    np.testing.assert_allclose(qn[0, 5, mbc:N - mbc], ref[mbc:N - mbc], rtol=0, atol=1e-14)


def test_flux_difference_form_is_conservative_without_boundary_flux():
    """With extrapolated (zero-gradient) ghost values around a compactly
    supported bump, the interior total of q is conserved to rounding."""
    mx = 16
    cfg = O.make_cfg(O.SOLVER_SWE_ROE, mx, mx, 2)
    q = np.zeros((3, mx + 4, mx + 4))
    x = np.arange(mx + 4) - 2 + 0.5
    X, Y = np.meshgrid(x, x)
    q[0] = 1.0 + 0.3 * np.exp(-((X - 8) ** 2 + (Y - 8) ** 2) / 2)
    dt = 0.02
    qn, _ = O.step2(cfg, q, None, 0.1, 0.1, dt)
    assert abs(qn[0, 2:-2, 2:-2].sum() - q[0, 2:-2, 2:-2].sum()) < 1e-11


def test_rejects_periodic_and_user_bc():
    q = np.zeros((3, 12, 12))
    with pytest.raises(ValueError):
        O.bc2(q, 8, 8, 2, [2, 1, 1, 1])
    with pytest.raises(ValueError):
        O.bc2(q, 8, 8, 2, [1, 0, 1, 1])


refdrive = pytest.importorskip("refdrive")


@pytest.mark.skipif(not refdrive.available(), reason="oracle/_ref not built (needs /root/reference)")
def test_oracle_matches_live_reference_random():
    import cases as C
    rng = np.random.default_rng(11)
    for k in range(4):
        mx, my = int(rng.integers(6, 20)), int(rng.integers(6, 20))
        q, dx, dy = C.swe_state(rng, mx, my, 2)
        dt = C.swe_dt(q, dx, dy)
        lim = [int(x) for x in rng.integers(0, 7, 3)]
        order = [(2, 2), (2, 1), (1, 1), (2, 0)][k]
        refdrive.set_params("swe", grav=1.0)
        qr, cr = refdrive.step2("swe", q, None, mx=mx, my=my, mbc=2, dx=dx, dy=dy, dt=dt,
                                order=order, mthlim=lim)
        qo, co = O.step2(O.make_cfg(O.SOLVER_SWE_ROE, mx, my, 2, order=order, mthlim=lim), q,
                         None, dx, dy, dt)
        assert np.array_equal(qr, qo) and cr == co


# ---- conservative transport with f-waves ---------------------------------

def _fw(q, aux, dx, dy, dt, order=(2, 2), lim=4):
    ny, nx = q.shape[1:]
    cfg = O.make_cfg(O.SOLVER_TRANSPORT_FW, nx - 4, ny - 4, 2, order=order, mthlim=[lim])
    return O.step2(cfg, q, aux, dx, dy, dt)


@pytest.mark.parametrize("u,v", [(0.7, 0.4), (-0.6, 0.5), (0.45, -0.8), (-0.3, -0.9)])
def test_fw_constant_velocity_equals_wave_propagation(u, v):
    """With one velocity everywhere the f-wave is s times the jump, so the
    f-wave step equals the wave-propagation advection step (pinned above
    against the reference and the textbook) to rounding, for every sign of
    (u, v): a correction weighted by |s| instead of sign(s), or a transverse
    velocity taken from the wrong side, fails this."""
    rng = np.random.default_rng(int(100 * (u + 2 * v) + 500))
    q = rng.uniform(0, 1, (1, 20, 20))
    aux = np.zeros((7, 20, 20))
    aux[0] = 1.0
    aux[1] = aux[2] = u
    aux[3] = aux[4] = v
    aux[5] = aux[6] = 1.0
    dt = 0.8 * 0.05 / max(abs(u), abs(v))
    for order in [(2, 2), (2, 1), (1, 2)]:
        qf, cf = _fw(q, aux, 0.05, 0.05, dt, order)
        qa, ca = O.step2(O.make_cfg(O.SOLVER_ADV, 16, 16, 2, order=order), q,
                         np.stack([np.full((20, 20), u), np.full((20, 20), v)]), 0.05, 0.05, dt)
        np.testing.assert_allclose(qf[:, 2:-2, 2:-2], qa[:, 2:-2, 2:-2], rtol=0, atol=1e-14)
        assert cf == pytest.approx(ca, rel=1e-15)


def test_fw_1d_reduction_matches_bale_leveque_fwave_formula():
    """v = 0, q and u constant in y, u > 0 varying cell to cell: the step is the
    1-D f-wave method of Bale, LeVeque, Mitran and Rossmanith (SIAM J. Sci.
    Comput. 24, 2002), eqs. 2.10-2.14, with the MC limiter, written out here:
    Z_{i-1/2} = u_i Q_i - u_{i-1} Q_{i-1}, s = (u_{i-1} + u_i)/2,
    Q_i -= dt/dx (Z_{i-1/2} + F_{i+1/2} - F_{i-1/2}),
    F = 1/2 sgn(s) (1 - dt/dx |s|) phi(r) Z."""
    mx, mbc, dx = 24, 2, 0.1
    N = mx + 2 * mbc
    rng = np.random.default_rng(8)
    uc = rng.uniform(0.3, 0.9, N)
    Q = np.concatenate([np.zeros(8), np.ones(10), rng.uniform(0, 1, N - 18)])
    dt = 0.8 * dx / uc.max()
    q = np.tile(Q, (N, 1))[None]
    aux = np.zeros((7, N, N))
    aux[0] = 1.0
    aux[1] = aux[2] = np.tile(uc, (N, 1))
    aux[5] = aux[6] = 1.0
    qn, cfl = _fw(q, aux, dx, dx, dt)
    Z = uc[1:] * Q[1:] - uc[:-1] * Q[:-1]   # Z[k] at face between cells k and k+1
    s = 0.5 * (uc[1:] + uc[:-1])
    assert cfl == pytest.approx(dt / dx * s[1:-1].max(), rel=1e-15)

    def phi(r):
        return max(0.0, min((1 + r) / 2, 2.0, 2 * r))

    def F(k):  # face between cells k and k+1, upwind neighbour face k-1 (s > 0)
        w = Z[k]
        r = Z[k - 1] / w if w != 0 else 0.0
        return 0.5 * (1 - dt / dx * s[k]) * (phi(r) if w != 0 else 1.0) * w

    ref = Q.copy()
    for i in range(mbc, N - mbc):
        ref[i] = Q[i] - dt / dx * (Z[i - 1] + F(i) - F(i - 1))
    np.testing.assert_allclose(qn[0, 5, mbc:N - mbc], ref[mbc:N - mbc], rtol=0, atol=1e-14)


def test_fw_conserves_with_cell_velocities():
    """With one velocity per cell (uL = uR, vB = vT) and unit edge lengths the
    f-waves telescope: the interior total of a compactly supported q is
    conserved to rounding, for any velocity field."""
    import sys
    sys.path.insert(0, GOLD)
    import cases as C
    rng = np.random.default_rng(21)
    q, aux, dx, dy = C.fw_state(rng, 16, 16, 2)
    x = np.arange(20) - 2 + 0.5
    X, Y = np.meshgrid(x, x)
    q[0] = np.where((X - 8) ** 2 + (Y - 8) ** 2 < 16, q[0], 0.0)
    qn, _ = _fw(q, aux, dx, dy, C.fw_dt(aux, dx, dy))
    assert abs(qn[0, 2:-2, 2:-2].sum() - q[0, 2:-2, 2:-2].sum()) < 1e-13


@pytest.mark.parametrize("perm,which", [([0, 2, 1, 3, 4, 5, 6], "fw_step_1"),
                                        ([0, 1, 2, 4, 3, 5, 6], "fw_step_3"),
                                        ([0, 1, 2, 3, 4, 6, 5], "fw_step_2")])
def test_fw_planted_swapped_aux_is_detected(perm, which):
    """Left/right (bottom/top) velocities and x/y edge lengths swapped in the
    input change the result, so the fixtures tell those components apart."""
    g = _load(which)
    qn, _ = O.step2(_cfg_for(which, g), g["q"], g["aux"][perm], float(g["dx"]), float(g["dy"]),
                    float(g["dt"]))
    assert np.abs(qn - g["qout"]).max() > 1e-6


@pytest.mark.skipif(not refdrive.available(), reason="oracle/_ref not built (needs /root/reference)")
def test_fw_oracle_matches_live_reference_random():
    import cases as C
    rng = np.random.default_rng(13)
    for k in range(4):
        mx, my = int(rng.integers(6, 20)), int(rng.integers(6, 20))
        q, aux, dx, dy = C.fw_state(rng, mx, my, 2, per_edge=k % 2 == 1, lengths=k >= 2)
        dt = C.fw_dt(aux, dx, dy)
        lim = [int(rng.integers(0, 7))]
        order = [(2, 2), (2, 1), (1, 1), (2, 0)][k]
        qr, cr = refdrive.step2("fw", q, aux, mx=mx, my=my, mbc=2, dx=dx, dy=dy, dt=dt,
                                order=order, mthlim=lim)
        qo, co = O.step2(O.make_cfg(O.SOLVER_TRANSPORT_FW, mx, my, 2, order=order, mthlim=lim),
                         q, aux, dx, dy, dt)
        assert np.array_equal(qr, qo) and cr == co
