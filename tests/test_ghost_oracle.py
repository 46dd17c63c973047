"""CPU tests of the oracle's ghost-fill, regrid and time-interpolation kernels
(oracle/claw46_ghost.c) and of the Python AMR driver restatement
(oracle/amr.py).

Pinning: time interpolation bitwise against the reference Fortran golden
vector; copies against brute-force index shifts; averages and
interpolations against data they must reproduce exactly (constants and
linear fields with dyadic values); conservation of average2coarse after
interpolate2fine; ghost updates on uniform forests against the exact
values of a global array.  The patch-pair kernels themselves cannot run
from the reference here (they call ForestClaw's transform callbacks, which
need p4est), so those checks are what pins them.
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import amr as A
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
_dp = ctypes.POINTER(ctypes.c_double)


def P(a):
    return a.ctypes.data_as(_dp)


@pytest.fixture(scope="module")
def L():
    return A._setup_sigs()


def tile(meqn, mx, my, mbc=2):
    return np.zeros((meqn, my + 2 * mbc, mx + 2 * mbc))


def test_timeinterp_matches_reference_golden(L):
    g = dict(np.load(os.path.join(GOLD, "timeinterp.npz")))
    mx, my, mbc = int(g["mx"]), int(g["my"]), int(g["mbc"])
    qi = np.zeros_like(g["qcurr"])
    L.orc_timeinterp(mx, my, mbc, 3, P(g["qcurr"]), P(g["qlast"]), P(qi), float(g["alpha"]))
    assert np.array_equal(qi, g["qinterp"])


def test_copy_face_and_corner_are_index_shifts(L):
    rng = np.random.default_rng(1)
    mx = my = 8
    nb = rng.standard_normal((2, my + 4, mx + 4))
    for f, (sx, sy) in enumerate([(mx, 0), (-mx, 0), (0, my), (0, -my)]):
        q = np.full((2, my + 4, mx + 4), np.nan)
        L.orc_copy_face(mx, my, 2, 2, P(q), P(nb), f, sx, sy)
        written = ~np.isnan(q)
        assert written.sum() == 2 * 2 * 8
        jj, ii = np.nonzero(written[0])
        assert np.array_equal(q[0][jj, ii], nb[0][jj + sy, ii + sx])
    for c, (sx, sy) in enumerate([(mx, my), (-mx, my), (mx, -my), (-mx, -my)]):
        q = np.full((2, my + 4, mx + 4), np.nan)
        L.orc_copy_corner(mx, my, 2, 2, P(q), P(nb), c, sx, sy)
        jj, ii = np.nonzero(~np.isnan(q[0]))
        assert len(ii) == 4
        assert np.array_equal(q[1][jj, ii], nb[1][jj + sy, ii + sx])


def _linear(mx, my, level, gx, gy, a=0.25, b=-0.5, c=3.0):
    """Values of a*X + b*Y + c at cell centres in global fine units (dyadic, exact)."""
    h = 2.0 ** (-level)
    i = np.arange(mx + 4) - 2 + 0.5
    j = np.arange(my + 4) - 2 + 0.5
    X = (gx * mx + i) * h
    Y = (gy * my + j) * h
    XX, YY = np.meshgrid(X, Y)
    return (a * XX + b * YY + c)[None]


def test_average_and_interpolate_reproduce_linear_fields(L):
    mx = my = 8
    # coarse patch at level 1 (gx=0, gy=0); fine neighbours at level 2 to the right
    qc_exact = _linear(mx, my, 1, 0, 0)
    for fy in (0, 1):
        fine = _linear(mx, my, 2, 2, fy)
        ox, oy = (2 * 0 - 2) * mx, (2 * 0 - fy) * my
        qc = qc_exact.copy()
        qc[:, :, mx + 2:] = 0  # right ghost columns to be filled
        L.orc_average_face(mx, my, 2, 1, P(qc), P(fine), 1, ox, oy)
        rows = slice(2 + fy * my // 2, 2 + (fy + 1) * my // 2)
        assert np.array_equal(qc[:, rows, mx + 2:], qc_exact[:, rows, mx + 2:])
        # interpolation from the coarse patch into the fine patch's left ghosts
        qf = fine.copy()
        qf[:, :, :2] = -7
        L.orc_interpolate_face(mx, my, 2, 1, P(qc_exact), P(qf), 1, ox, oy)
        assert np.array_equal(qf[:, :, :2], fine[:, :, :2]) or np.array_equal(
            qf[:, 2:-2, :2], fine[:, 2:-2, :2])


def test_interpolate_constant_and_corner(L):
    mx = my = 8
    qc = np.full((2, my + 4, mx + 4), 1.25)
    qf = np.zeros((2, my + 4, mx + 4))
    # fine corner neighbour at the coarse upper-right corner
    L.orc_interpolate_corner(mx, my, 2, 2, P(qc), P(qf), 3, (2 * 0 - 2) * mx, (2 * 0 - 2) * my)
    assert np.all(qf[:, :2, :2] == 1.25)
    assert np.count_nonzero(qf) == 2 * 4


def test_interpolate2fine_then_average2coarse_is_identity_on_linear_data(L):
    mx = my = 8
    qc = _linear(mx, my, 1, 1, 0)
    back = np.zeros_like(qc)
    for ig in range(4):
        qf = np.zeros_like(qc)
        L.orc_interpolate2fine(mx, my, 2, 1, P(qc), P(qf), ig)
        fx, fy = 2 * 1 + (ig & 1), ig >> 1
        exact = _linear(mx, my, 2, fx, fy)
        assert np.array_equal(qf[:, 2:-2, 2:-2], exact[:, 2:-2, 2:-2])
        L.orc_average2coarse(mx, my, 2, 1, P(back), P(qf), ig)
    assert np.array_equal(back[:, 2:-2, 2:-2], qc[:, 2:-2, 2:-2])


def _minmod(*a):
    """Textbook minmod: the argument of least magnitude when all share a
    sign, else 0."""
    if all(x > 0 for x in a):
        return min(a)
    if all(x < 0 for x in a):
        return max(a)
    return 0.0


def _mc(sl, sr):
    """Monotonized-central slope (van Leer 1977; LeVeque, Finite Volume
    Methods for Hyperbolic Problems, eq. 6.29): minmod(2 sl, 2 sr, (sl+sr)/2)."""
    return _minmod(2 * sl, 2 * sr, (sl + sr) / 2)


def _slopes():
    s = O.lib().orc_compute_slopes
    s.argtypes = [ctypes.c_double, ctypes.c_double]
    s.restype = ctypes.c_double
    return s


def test_interpolation_slope_is_the_mc_slope():
    """The AMRClaw slope of the interpolation (mth = 5) is the MC slope.
    The table separates it from minmod and from the unlimited centred slope:
    (1, 3) gives 2 (minmod 1), (1, 5) gives 2 (centred 3), opposite signs
    give 0 (centred nonzero)."""
    s = _slopes()
    table = {(1.0, 3.0): 2.0, (3.0, 1.0): 2.0, (1.0, 5.0): 2.0, (5.0, 1.0): 2.0,
             (-1.0, -5.0): -2.0, (-1.0, 2.0): 0.0, (2.0, -1.0): 0.0, (0.0, 5.0): 0.0,
             (1.0, 1.0): 1.0, (-2.0, -0.5): -1.0, (0.25, 0.5): 0.375}
    for (sl, sr), want in table.items():
        assert s(sl, sr) == want, (sl, sr)
        assert _mc(sl, sr) == want, (sl, sr)
    rng = np.random.default_rng(9)
    for sl, sr in rng.standard_normal((2000, 2)) * np.array([1.0, 3.0]):
        assert s(sl, sr) == _mc(sl, sr)


def _centres(level, gx, gy, mx, my):
    """Cell-centre coordinates (X, Y) of a tile (ghost cells included) of the
    level-`level` patch (gx, gy) of the unit square."""
    h = 2.0 ** (-level)
    x = gx * h + (np.arange(mx + 4) - 2 + 0.5) * (h / mx)
    y = gy * h + (np.arange(my + 4) - 2 + 0.5) * (h / my)
    return np.meshgrid(x, y)


def _expected_interp(qc, level, cg, X, Y, mx, my, slope):
    """Value at fine-cell centre (X, Y): the coarse cell containing it, plus
    the limited x and y slopes times the offset in coarse-cell units."""
    h = 2.0 ** (-level)
    hc = h / mx
    fi = (X - cg[0] * h) / hc      # position in coarse cells from the patch corner
    fj = (Y - cg[1] * h) / hc
    ic, jc = int(np.floor(fi)), int(np.floor(fj))  # 0-based interior index
    ox, oy = fi - (ic + 0.5), fj - (jc + 0.5)      # +-1/4
    I, J = ic + 2, jc + 2                          # tile index
    v = qc[J, I]
    gx = slope(v - qc[J, I - 1], qc[J, I + 1] - v)
    gy = slope(v - qc[J - 1, I], qc[J + 1, I] - v)
    return v + gx * ox + gy * oy


@pytest.mark.parametrize("side", range(8))
def test_interpolation_of_nonlinear_data_by_coordinates(L, side):
    """Coarse patch (level 1, (1, 1)) with random data; a half-size
    neighbour across one of its faces (sides 0-3, the lower or left half)
    or corners (4-7).  Every fine ghost cell the interpolation writes equals
    the coarse cell that contains its centre plus MC-limited slopes times its
    offset, all located from coordinates.  The same data with minmod or
    unlimited slopes gives different values, so the check separates the
    limiters."""
    mx = my = 8
    cg = (1, 1)
    if side < 4:
        iface = side
        fg = [(1, 2), (4, 2), (2, 1), (2, 4)][iface]  # level 2, across coarse face iface
    else:
        icorner = side - 4
        fg = [(1, 1), (4, 1), (1, 4), (4, 4)][icorner]
    ox, oy = (2 * cg[0] - fg[0]) * mx, (2 * cg[1] - fg[1]) * my
    X, Y = _centres(2, fg[0], fg[1], mx, my)
    # the fine patch's ghost band facing the coarse patch
    band = {0: (slice(2, -2), slice(-2, None)), 1: (slice(2, -2), slice(0, 2)),
            2: (slice(-2, None), slice(2, -2)), 3: (slice(0, 2), slice(2, -2)),
            4: (slice(-2, None), slice(-2, None)), 5: (slice(-2, None), slice(0, 2)),
            6: (slice(0, 2), slice(-2, None)), 7: (slice(0, 2), slice(0, 2))}[side]
    diff = {"minmod": 0, "centred": 0}
    for seed in range(6):
        qc = np.random.default_rng(30 + 8 * seed + side).standard_normal((1, my + 4, mx + 4))
        qf = np.full((1, my + 4, mx + 4), np.nan)
        if side < 4:
            L.orc_interpolate_face(mx, my, 2, 1, P(qc), P(qf), iface, ox, oy)
        else:
            L.orc_interpolate_corner(mx, my, 2, 1, P(qc), P(qf), icorner, ox, oy)
        written = ~np.isnan(qf[0])
        assert written[band].all()
        assert not written[2:-2, 2:-2].any()  # interior untouched
        for j, i in zip(*np.nonzero(written)):
            want = _expected_interp(qc[0], 1, cg, X[j, i], Y[j, i], mx, my, _mc)
            assert qf[0, j, i] == want, (side, j, i)
            diff["minmod"] += qf[0, j, i] != _expected_interp(
                qc[0], 1, cg, X[j, i], Y[j, i], mx, my, lambda a, b: _minmod(a, b))
            diff["centred"] += qf[0, j, i] != _expected_interp(
                qc[0], 1, cg, X[j, i], Y[j, i], mx, my, lambda a, b: (a + b) / 2)
    assert diff["minmod"] > 0 and diff["centred"] > 0, diff


def _uniform_oracle(level, mx=8, meqn=3):
    cfg = O.make_cfg(O.SOLVER_SWE_ROE, mx, mx, 2)
    n = 1 << level
    leaves = [(level, x, y) for y in range(n) for x in range(n)]
    h = 1.0 / n
    geom = np.array([[x * h, y * h, h / mx, h / mx] for (_, x, y) in leaves])
    return A.AmrOracle(cfg, leaves, geom, [1, 1, 1, 1]), leaves


def test_uniform_ghost_update_equals_global_array():
    amr, leaves = _uniform_oracle(2)
    mx = 8
    N = 4 * mx
    rng = np.random.default_rng(4)
    G = rng.standard_normal((3, N, N))
    for k, (_, x, y) in enumerate(leaves):
        amr.q[k][:, 2:-2, 2:-2] = G[:, y * mx:(y + 1) * mx, x * mx:(x + 1) * mx]
    amr.ghost_update(2, 2, 0)
    Gp = np.pad(G, ((0, 0), (2, 2), (2, 2)), mode="edge")  # extrapolation BC at the boundary
    for k, (_, x, y) in enumerate(leaves):
        assert np.array_equal(amr.q[k], Gp[:, y * mx:y * mx + mx + 4, x * mx:x * mx + mx + 4])


def test_two_level_ghost_update_reproduces_linear_field():
    """Coarse level 1 with one refined quadrant at level 2: after a ghost
    update every ghost cell holds the exact linear field (averages and
    limited interpolation are exact on linear data)."""
    mx = 8
    cfg = O.make_cfg(O.SOLVER_SWE_ROE, mx, mx, 2)
    leaves = [(1, 1, 0), (1, 0, 1), (1, 1, 1)] + [(2, x, y) for y in (0, 1) for x in (0, 1)]
    geom = []
    for (L_, x, y) in leaves:
        h = 2.0 ** (-L_)
        geom.append([x * h, y * h, h / mx, h / mx])
    amr = A.AmrOracle(cfg, leaves, np.array(geom), [1, 1, 1, 1])
    exact = []
    for k, (L_, x, y) in enumerate(leaves):
        e = np.concatenate([_linear(mx, mx, L_ + 2, x, y)] * 3)  # fine units of level L+2
        exact.append(e)
        amr.q[k][:, 2:-2, 2:-2] = e[:, 2:-2, 2:-2]
    amr.ghost_update(1, 2, 0)
    hc = 1.0 / 16  # coarse cell size
    for k, (L_, x, y) in enumerate(leaves):
        h = 2.0 ** (-L_) / mx
        c = (np.arange(mx + 4) - 2 + 0.5) * h
        XC, YC = np.meshgrid(x * 2.0 ** (-L_) + c, y * 2.0 ** (-L_) + c)
        # ghost values next to the physical boundary come from extrapolation
        # and limited slopes there are not exact; compare away from it
        far = np.minimum(np.minimum(XC, 1 - XC), np.minimum(YC, 1 - YC)) > 2.5 * hc
        got, ex = amr.q[k][:, far], exact[k][:, far]
        assert far.sum() > 0
        assert np.array_equal(got, ex), (k, leaves[k])
