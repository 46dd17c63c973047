"""Discriminating CPU checks of the oracle's time-sync pieces: the register
accumulation of a patch step (store_flux / accumulate_waves), the same-size
correction and the fine-to-coarse correction (oracle/claw46_oracle.c,
oracle/claw46_timesync.c; reference
src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_time_sync.f:37-226,
254-498, 528-704).

None of these checks restates the oracle's correction formulas.  They use
three independent facts:

* Flux form of a wave-propagation step.  At a face the numerical flux seen
  from the cell on its low side is F = fm + f(q_low) and from the cell on its
  high side F = fp + f(q_high); the two agree because amdq + apdq =
  f(q_high) - f(q_low) (the Roe property), and the flux-difference update
  telescopes, so the interior mass change of a patch equals the sum of its
  boundary face fluxes.  With the register meanings of accumulate_waves and
  store_flux (P = -dt el fp, M = +dt el fm, E0 = dt el f(q) at the first
  interior cell, E1 at the first ghost cell) every face flux of a step can be
  read off the registers, from the inside (E0) or from the ghost side (E1).
* A conservative correction makes two patches that share a face agree on
  the flux through it: the same-size correction moves both sides to the mean
  of their two fluxes, the fine-to-coarse correction moves the coarse cell to
  the sum of the fine fluxes through its face segment.  A union of patches
  is then conservative: its mass change is the sum of its outer face fluxes.
* Geometry: a fine register entry belongs to the coarse cell whose extent
  along the face contains the fine segment, found here from coordinates.

Both corrections are affine in (q, registers), so probing every register
entry one at a time pins them completely.
"""
import ctypes
import math

import numpy as np
import pytest

import cases as C
from oracle import amr as A
from oracle import oracle as O

_dp = ctypes.POINTER(ctypes.c_double)
MBC = 2
GRAV = 1.0


def P(a):
    return a.ctypes.data_as(_dp)


@pytest.fixture(scope="module")
def L():
    return A._setup_sigs()


def blk(reg, mx, my, meqn, face, b):
    """View of register block b (0 P, 1 M, 2 E0, 3 E1) of face f (0 left, 1
    right, 2 bottom, 3 top): meqn x length, the layout the ABI documents."""
    lens = [my, my, mx, mx]
    off = sum(4 * meqn * lens[f] for f in range(face)) + b * meqn * lens[face]
    return reg[off:off + meqn * lens[face]].reshape(meqn, lens[face])


def face_flux(reg, mx, my, meqn, face, view):
    """Time-integrated flux through each cell face of a patch side, in the
    direction of increasing x (faces 0, 1) or y (faces 2, 3), read from the
    registers.  view 'in': with the patch's own edge cell as the state on its
    side (E0); view 'ghost': with the ghost cell as the state on the far side
    (E1).  Left / bottom faces are low faces of the patch: F = fp + f(q_in)
    = -P + E0, or seen with the ghost cell on the low side F = fm + f(q_ghost)
    = M + E1.  Right / top faces: F = fm + f(q_in) = M + E0, or
    F = fp + f(q_ghost) = -P + E1."""
    Pb, Mb = blk(reg, mx, my, meqn, face, 0), blk(reg, mx, my, meqn, face, 1)
    E0, E1 = blk(reg, mx, my, meqn, face, 2), blk(reg, mx, my, meqn, face, 3)
    low = face in (0, 2)
    if view == "in":
        return (-Pb + E0) if low else (Mb + E0)
    return (Mb + E1) if low else (-Pb + E1)


def budget(reg, mx, my, meqn):
    """Mass change of a patch implied by its boundary fluxes (per field)."""
    f = [face_flux(reg, mx, my, meqn, k, "in").sum(1) for k in range(4)]
    return f[0] - f[1] + f[2] - f[3]


def interior_mass(q, dx, dy):
    return q[:, MBC:-MBC, MBC:-MBC].sum(axis=(1, 2)) * dx * dy


def _cfg(mx, my):
    return O.make_cfg(O.SOLVER_SWE_ROE, mx, my, MBC, grav=GRAV)


def _state(seed, nx, ny, mx):
    """A smooth-plus-noise SWE state over nx x ny patches (with a 2-cell
    border), so that neighbouring tiles cut from it have consistent ghosts."""
    rng = np.random.default_rng(seed)
    q, _, _ = C.swe_state(rng, nx * mx, ny * mx, MBC, transonic=False)
    return np.ascontiguousarray(q)


def _step(q, mx, my, dx, dy, dt, reg=None):
    q = np.ascontiguousarray(q.copy())
    reg = np.zeros(O.reg_size(mx, my, 3)) if reg is None else reg
    O.step2_ts(_cfg(mx, my), q, dx, dy, dt, reg)
    return q, reg


def _scale(reg):
    return np.abs(reg).max()


def test_registers_close_the_mass_budget_of_a_patch():
    """Interior mass change of a step = sum of the boundary face fluxes
    read from the registers; the ghost-side reading of every face agrees
    with the inside reading (Roe property).  A wrong dt / edge-length factor,
    sign or block in store_flux or accumulate_waves breaks this."""
    mx, my = 8, 12
    for seed in range(4):
        q, dx, _ = C.swe_state(np.random.default_rng(seed), mx, my, MBC)
        q = np.ascontiguousarray(q)
        dy = 1.37 * dx  # non-square cells: x and y edge lengths differ
        dt = 0.5 * C.swe_dt(q, dx, dy)
        qn, reg = _step(q, mx, my, dx, dy, dt)
        dm = interior_mass(qn, dx, dy) - interior_mass(q, dx, dy)
        tol = 1e-13 * _scale(reg) * (mx + my)
        np.testing.assert_allclose(dm, budget(reg, mx, my, 3), rtol=0, atol=tol)
        for face in range(4):
            np.testing.assert_allclose(face_flux(reg, mx, my, 3, face, "ghost"),
                                       face_flux(reg, mx, my, 3, face, "in"), rtol=0, atol=tol)
        assert np.abs(dm).max() > 1e3 * tol  # the data moves mass


@pytest.mark.parametrize("axis", [0, 1])
def test_register_rows_agree_across_a_same_size_face(axis):
    """Two same-size patches cut from one array: per cell of the shared face,
    the flux the low patch records equals the flux the high patch records,
    and each side's ghost-cell flux is the other's interior-cell flux
    exactly.  A row shift or transposition in the registers breaks it
    (the data varies along the face)."""
    mx = my = 8
    nx, ny = (2, 1) if axis == 0 else (1, 2)
    G = _state(11 + axis, nx, ny, mx)
    dx, dy = 1.0 / (nx * mx), 1.25 / (nx * mx)
    dt = 0.5 * C.swe_dt(G, dx, dy)
    tiles = []
    for k in range(2):
        ox, oy = (k * mx, 0) if axis == 0 else (0, k * my)
        tiles.append(G[:, oy:oy + my + 2 * MBC, ox:ox + mx + 2 * MBC])
    (_, ra), (_, rb) = (_step(t, mx, my, dx, dy, dt) for t in tiles)
    hi, lo = (1, 0) if axis == 0 else (3, 2)  # the shared face seen by A (hi) and B (lo)
    tol = 1e-13 * max(_scale(ra), _scale(rb)) * mx
    np.testing.assert_allclose(face_flux(ra, mx, my, 3, hi, "in"), face_flux(rb, mx, my, 3, lo, "in"),
                               rtol=0, atol=tol)
    assert np.array_equal(blk(ra, mx, my, 3, hi, 3), blk(rb, mx, my, 3, lo, 2))
    assert np.array_equal(blk(rb, mx, my, 3, lo, 3), blk(ra, mx, my, 3, hi, 2))
    # the face data is not uniform along the face (the check discriminates rows)
    assert np.ptp(face_flux(ra, mx, my, 3, hi, "in")[0]) > 1e3 * tol


def _samesize(L, q, face, area, rt, rn, mx, my):
    L.orc_time_sync_samesize(mx, my, MBC, 3, face, area, P(q), P(rt), P(rn))


@pytest.mark.parametrize("axis", [0, 1])
def test_samesize_corrections_conserve_a_planted_mismatch(L, axis):
    """Patch B steps with perturbed ghost cells on the face it shares with A,
    so the two sides disagree on the face flux by a non-rounding amount.
    After the same-size correction of both sides the pair is conservative:
    its mass change equals its outer face fluxes.  Only the face's edge cells
    change, both by half the mismatch of their row.  Without the correction,
    and with a sign error, the pair budget is off by the mismatch and by
    twice the mismatch."""
    mx = my = 8
    nx, ny = (2, 1) if axis == 0 else (1, 2)
    G = _state(21 + axis, nx, ny, mx)
    dx, dy = 1.0 / (nx * mx), 0.8 / (nx * mx)
    dt = 0.4 * C.swe_dt(G, dx, dy)
    tiles = []
    for k in range(2):
        ox, oy = (k * mx, 0) if axis == 0 else (0, k * my)
        tiles.append(G[:, oy:oy + my + 2 * MBC, ox:ox + mx + 2 * MBC].copy())
    qa, qb = tiles
    ramp = 1.0 + 0.5 * np.sin(np.arange(my + 2 * MBC))  # varies along the face
    if axis == 0:
        qb[0, :, :MBC] += 0.3 * ramp[:, None]  # B's left ghost columns
    else:
        qb[0, :MBC, :] += 0.3 * ramp[None, :]  # B's bottom ghost rows
    na, ra = _step(qa, mx, my, dx, dy, dt)
    nb, rb = _step(qb, mx, my, dx, dy, dt)
    hi, lo = (1, 0) if axis == 0 else (3, 2)
    area = dx * dy
    m0 = interior_mass(qa, dx, dy) + interior_mass(qb, dx, dy)
    # outer faces of the pair: every face flux except the shared one
    outer = budget(ra, mx, my, 3) + budget(rb, mx, my, 3)
    outer = outer + face_flux(ra, mx, my, 3, hi, "in").sum(1) - face_flux(rb, mx, my, 3, lo, "in").sum(1)
    mismatch = face_flux(ra, mx, my, 3, hi, "in") - face_flux(rb, mx, my, 3, lo, "in")
    tol = 1e-13 * max(_scale(ra), _scale(rb)) * (mx + my)
    assert np.abs(mismatch.sum(1)).max() > 1e4 * tol

    def corrected(sign):
        ca, cb = na.copy(), nb.copy()
        _samesize(L, ca, hi, sign * area, ra, rb, mx, my)
        _samesize(L, cb, lo, sign * area, rb, ra, mx, my)
        return ca, cb

    ca, cb = corrected(1.0)
    err = interior_mass(ca, dx, dy) + interior_mass(cb, dx, dy) - m0 - outer
    np.testing.assert_allclose(err, 0.0, atol=tol)
    # the uncorrected pair misses by the mismatch, a sign error doubles it
    err0 = interior_mass(na, dx, dy) + interior_mass(nb, dx, dy) - m0 - outer
    np.testing.assert_allclose(err0, -mismatch.sum(1), atol=tol)
    fa, fb = corrected(-1.0)
    errf = interior_mass(fa, dx, dy) + interior_mass(fb, dx, dy) - m0 - outer
    np.testing.assert_allclose(errf, -2 * mismatch.sum(1), atol=tol)
    # locality: only the two edge lines change, each by half its row's mismatch
    da, db = (ca - na) * area, (cb - nb) * area
    if axis == 0:
        ea, eb = da[:, MBC:-MBC, MBC + mx - 1].copy(), db[:, MBC:-MBC, MBC].copy()
        da[:, MBC:-MBC, MBC + mx - 1] = 0
        db[:, MBC:-MBC, MBC] = 0
    else:
        ea, eb = da[:, MBC + my - 1, MBC:-MBC].copy(), db[:, MBC, MBC:-MBC].copy()
        da[:, MBC + my - 1, MBC:-MBC] = 0
        db[:, MBC, MBC:-MBC] = 0
    assert not da.any() and not db.any()
    np.testing.assert_allclose(ea, 0.5 * mismatch, atol=tol)
    np.testing.assert_allclose(eb, 0.5 * mismatch, atol=tol)


def _probe(fn, shape_regs):
    """Response of an affine map of register arrays to a unit value at every
    entry of each array in turn (the affine part subtracted)."""
    base = fn(*[np.zeros(n) for n in shape_regs])
    for a, n in enumerate(shape_regs):
        for e in range(n):
            regs = [np.zeros(n_) for n_ in shape_regs]
            regs[a][e] = 1.0
            yield a, e, fn(*regs) - base


def _entry(mx, my, meqn, e):
    """(face, block, field, index along the face) of register entry e."""
    lens = [my, my, mx, mx]
    for f in range(4):
        n = 4 * meqn * lens[f]
        if e < n:
            b, r = divmod(e, meqn * lens[f])
            m, k = divmod(r, lens[f])
            return f, b, m, k
        e -= n
    raise IndexError(e)


def test_samesize_response_to_every_register_entry(L):
    """Every register entry of this patch and its neighbour, probed alone:
    the edge cell of the same row moves by +-1/2 * (change of the
    neighbour's face flux minus this side's) / area, with the sign of the
    side's inflow; entries of other faces and blocks move nothing."""
    mx, my, meqn = 4, 6, 3
    area = 0.375
    nreg = O.reg_size(mx, my, meqn)
    other = {0: 1, 1: 0, 2: 3, 3: 2}
    for face in range(4):
        def fn(rt, rn):
            q = np.zeros((meqn, my + 2 * MBC, mx + 2 * MBC))
            L.orc_time_sync_samesize(mx, my, MBC, meqn, face, area, P(q), P(rt), P(rn))
            return q
        for arr, e, dq in _probe(fn, [nreg, nreg]):
            f, b, m, k = _entry(mx, my, meqn, e)
            # change of this side's and of the neighbour's flux through the face
            dthis = dnbr = 0.0
            if arr == 0 and f == face:
                r = np.zeros(nreg); r[e] = 1.0
                dthis = face_flux(r, mx, my, meqn, f, "in")[m, k]
            if arr == 1 and f == other[face]:
                r = np.zeros(nreg); r[e] = 1.0
                dnbr = face_flux(r, mx, my, meqn, f, "ghost")[m, k]
            inflow = 1.0 if face in (0, 2) else -1.0  # flux in +x / +y enters a low face
            want = 0.5 * inflow * (dnbr - dthis) / area
            expect = np.zeros_like(dq)
            if want != 0.0:
                i = {0: 0, 1: mx - 1}.get(face, k)
                j = {2: 0, 3: my - 1}.get(face, k)
                expect[m, MBC + j, MBC + i] = want
            assert np.array_equal(dq, expect), (face, arr, f, b, m, k)


def test_f2c_response_to_every_register_entry(L):
    """Fine-to-coarse correction, every register entry probed alone: a fine
    entry moves only the coarse edge cell whose extent along the face
    contains the fine cell (found from coordinates), by the change of the
    fine flux into the coarse cell over the cell area; a coarse entry moves
    its own edge cell by minus its flux change when this fine half borders
    that cell, else nothing; entries of faces and blocks that carry no flux
    through this face move nothing."""
    mx, my, meqn = 4, 4, 3  # square patches (the reference's half-size transform)
    area = 0.0625
    nreg = O.reg_size(mx, my, meqn)
    fine_face = {0: 1, 1: 0, 2: 3, 3: 2}
    for iface in range(4):
        length = my if iface < 2 else mx
        for igrid in (0, 1):
            def fn(rc, rf):
                q = np.zeros((meqn, my + 2 * MBC, mx + 2 * MBC))
                L.orc_time_sync_f2c(mx, my, MBC, meqn, iface, igrid, area, P(q), P(rc), P(rf))
                return q
            for arr, e, dq in _probe(fn, [nreg, nreg]):
                f, b, m, k = _entry(mx, my, meqn, e)
                r = np.zeros(nreg); r[e] = 1.0
                inflow = 1.0 if iface in (0, 2) else -1.0
                expect = np.zeros_like(dq)
                covered = 0.5 * igrid <= (k + 0.5) / length < 0.5 * igrid + 0.5
                if arr == 0 and f == iface and covered:  # coarse register, cell next to this half
                    c = k
                    want = -inflow * face_flux(r, mx, my, meqn, f, "in")[m, k] / area
                elif arr == 1 and f == fine_face[iface]:  # fine register facing the coarse patch
                    # fine patch igrid covers [igrid/2, igrid/2 + 1/2] of the coarse face
                    s = 0.5 * igrid + (k + 0.5) * 0.5 / length
                    c = math.floor(s * length)
                    want = inflow * face_flux(r, mx, my, meqn, f, "ghost")[m, k] / area
                else:
                    want = 0.0
                if want != 0.0:
                    i = {0: 0, 1: mx - 1}.get(iface, c)
                    j = {2: 0, 3: my - 1}.get(iface, c)
                    expect[m, MBC + j, MBC + i] = want
                np.testing.assert_allclose(dq, expect, rtol=1e-15, atol=0,
                                           err_msg=str((iface, igrid, arr, f, b, m, k)))


@pytest.mark.parametrize("iface", [0, 1, 2, 3])
def test_f2c_makes_a_coarse_fine_union_conservative(L, iface):
    """A coarse patch and its two half-size neighbours across face iface,
    each with its own (arbitrary) ghost data: the fine patches take two
    steps of dt/2, the coarse one step of dt, all accumulating registers.
    After the fine-to-coarse correction of both halves the union's mass
    change equals its outer face fluxes; before it, it does not."""
    mx = my = 8
    rng = np.random.default_rng(40 + iface)
    hc = 1.0 / mx
    hf = 0.5 / mx
    dt = 0.25 * hf
    qc, _, _ = C.swe_state(rng, mx, my, MBC, transonic=False)
    qc = np.ascontiguousarray(qc)
    nc, rc = _step(qc, mx, my, hc, hc, dt)
    fines = []
    for igrid in (0, 1):
        qf, _, _ = C.swe_state(rng, mx, my, MBC, transonic=False)
        qf = np.ascontiguousarray(qf)
        nf, rf = _step(qf, mx, my, hf, hf, dt / 2)
        nf2 = nf.copy()
        ghost = np.ones_like(nf, dtype=bool)
        ghost[:, MBC:-MBC, MBC:-MBC] = False
        nf2[ghost] = qf[ghost]  # ghost cells refilled between the two fine steps
        nf2, rf = _step(nf2, mx, my, hf, hf, dt / 2, rf)
        fines.append((qf, nf2, rf))
    m0 = interior_mass(qc, hc, hc) + sum(interior_mass(f[0], hf, hf) for f in fines)
    ff = {0: 1, 1: 0, 2: 3, 3: 2}[iface]
    outer = budget(rc, mx, my, 3) + sum(budget(f[2], mx, my, 3) for f in fines)
    sgn = 1.0 if iface in (0, 2) else -1.0  # shared-face flux term of the coarse budget
    outer = outer - sgn * face_flux(rc, mx, my, 3, iface, "in").sum(1)
    outer = outer + sgn * sum(face_flux(f[2], mx, my, 3, ff, "in").sum(1) for f in fines)
    tol = 1e-13 * max(_scale(rc), max(_scale(f[2]) for f in fines)) * (mx + my)
    m1 = interior_mass(nc, hc, hc) + sum(interior_mass(f[1], hf, hf) for f in fines)
    assert np.abs(m1 - m0 - outer).max() > 1e4 * tol  # coarse and fine fluxes differ
    cc = nc.copy()
    for igrid in (0, 1):
        L.orc_time_sync_f2c(mx, my, MBC, 3, iface, igrid, hc * hc, P(cc), P(rc), P(fines[igrid][2]))
    m2 = interior_mass(cc, hc, hc) + sum(interior_mass(f[1], hf, hf) for f in fines)
    np.testing.assert_allclose(m2 - m0, outer, rtol=0, atol=tol)
    changed = np.argwhere((cc != nc).any(0))
    edge = {0: (None, MBC), 1: (None, MBC + mx - 1), 2: (MBC, None), 3: (MBC + my - 1, None)}[iface]
    for j, i in changed:
        assert (edge[0] is None or j == edge[0]) and (edge[1] is None or i == edge[1])
