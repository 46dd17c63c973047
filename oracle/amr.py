"""Pure-Python restatement of the ForestClaw drivers around the patch step:
ghost update, subcycled advance and time interpolation, driving the C
oracle's per-patch kernels.  TEST INFRASTRUCTURE ONLY (small cases).

Follows, loop for loop:
  fclaw2d_ghost_update_async   src/fclaw2d_ghost_fill.c:611-830 (one rank: the
                               ghost-only passes are empty)
  copy2ghost / average2ghost / interpolate2ghost / setphysical
                               src/fclaw2d_ghost_fill.c:144-304
  cb_face_fill / cb_corner_fill src/fclaw2d_face_neighbors.c:195-450,
                               src/fclaw2d_corner_neighbors.c:334-535
  advance                      src/fclaw2d_advance.c:54-357
  timeinterp                   src/fclaw2d_timeinterp.c:30-57
  time_sync                    src/fclaw2d_time_sync.c:118-202 with the reset
                               rules of fclaw2d_clawpatch_conservation.c:207-280
                               and the dispatch of fclaw2d_face_neighbors.c:320-390
Neighbour relations are derived here independently from the leaf
coordinates (level, gx, gy) on the brick's global index grid.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import oracle as O
from . import source as SRC

_dp = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


def _setup_sigs():
    L = O.lib()
    if getattr(L, "_amr_sigs", False):
        return L
    i4 = [ctypes.c_int] * 4
    for name in ("orc_copy_face", "orc_copy_corner", "orc_average_face", "orc_average_corner",
                 "orc_interpolate_face", "orc_interpolate_corner"):
        fn = getattr(L, name)
        fn.argtypes = i4 + [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        fn.restype = None
    for name in ("orc_interpolate2fine", "orc_average2coarse"):
        fn = getattr(L, name)
        fn.argtypes = i4 + [_dp, _dp, ctypes.c_int]
        fn.restype = None
    L.orc_timeinterp.argtypes = i4 + [_dp, _dp, _dp, ctypes.c_double]
    L.orc_time_sync_f2c.argtypes = i4 + [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp,
                                         _dp]
    L.orc_time_sync_f2c.restype = None
    L.orc_time_sync_samesize.argtypes = i4 + [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
    L.orc_time_sync_samesize.restype = None
    L.orc_time_sync_reset_face.argtypes = [ctypes.c_int] * 4 + [_dp]
    L.orc_time_sync_reset_face.restype = None
    L.orc_tag_gradient.argtypes = [ctypes.c_int] * 3 + [_dp, ctypes.c_int, ctypes.c_double,
                                                         ctypes.c_double, ctypes.c_double]
    L._amr_sigs = True
    return L


class AmrOracle:
    def __init__(self, cfg: O.OrcCfg, leaves, geom, mthbc, brick=(1, 1), subcycle=True,
                 swirl_tperiod=0.0, time_sync=False, minlevel=None, maxlevel=None, src2_gamma=None):
        """leaves: list of (level, gx, gy); geom: array (n, 4) of xlower, ylower, dx, dy.
        minlevel / maxlevel: the options' level range (default: the leaves' range),
        from which the time-step counters count.  src2_gamma: apply the
        shock-bubble axisymmetric source after every patch step
        (clawpack46_update, fc2d_clawpack46.cpp:434-441) with this gamma."""
        self.L = _setup_sigs()
        self.cfg = cfg
        self.mx, self.my, self.mbc, self.meqn = cfg.mx, cfg.my, cfg.mbc, cfg.meqn
        self.leaves = [tuple(int(v) for v in l) for l in leaves]
        self.geom = np.asarray(geom, dtype=np.float64)
        self.mthbc = list(mthbc)
        self.mi, self.mj = brick
        self.subcycle = subcycle
        self.swirl_tperiod = swirl_tperiod
        self.index = {l: k for k, l in enumerate(self.leaves)}
        shape = (self.meqn, self.my + 2 * self.mbc, self.mx + 2 * self.mbc)
        n = len(self.leaves)
        self.q = [np.zeros(shape) for _ in range(n)]
        self.qlast = [np.zeros(shape) for _ in range(n)]
        self.qti = [np.zeros(shape) for _ in range(n)]
        self.levels = sorted(set(l[0] for l in self.leaves))
        self.minlevel = self.levels[0] if minlevel is None else minlevel
        self.maxlevel = self.levels[-1] if maxlevel is None else maxlevel
        self.cell_updates = 0
        self.time_sync = time_sync
        self.src2_gamma = src2_gamma
        self.reg = [np.zeros(O.reg_size(self.mx, self.my, self.meqn)) for _ in range(n)]

    # ---------------- topology ------------------------------------------
    def inside(self, L, gx, gy):
        return 0 <= gx < (self.mi << L) and 0 <= gy < (self.mj << L)

    def face(self, k, f):
        L, gx, gy = self.leaves[k]
        dx, dy = [(-1, 0), (1, 0), (0, -1), (0, 1)][f]
        nx, ny = gx + dx, gy + dy
        if not self.inside(L, nx, ny):
            return ("bdry", None)
        if (L, nx, ny) in self.index:
            return ("same", self.index[(L, nx, ny)])
        if (L - 1, nx >> 1, ny >> 1) in self.index:
            return ("coarse", self.index[(L - 1, nx >> 1, ny >> 1)])
        if f == 0:
            kids = [(L + 1, 2 * nx + 1, 2 * ny + t) for t in (0, 1)]
        elif f == 1:
            kids = [(L + 1, 2 * nx, 2 * ny + t) for t in (0, 1)]
        elif f == 2:
            kids = [(L + 1, 2 * nx + t, 2 * ny + 1) for t in (0, 1)]
        else:
            kids = [(L + 1, 2 * nx + t, 2 * ny) for t in (0, 1)]
        return ("fine", [self.index[c] for c in kids])

    def corner(self, k, c):
        L, gx, gy = self.leaves[k]
        dx = 1 if c & 1 else -1
        dy = 1 if c & 2 else -1
        nx, ny = gx + dx, gy + dy
        if not self.inside(L, nx, ny):
            return ("bdry", None)
        if (L, nx, ny) in self.index:
            return ("same", self.index[(L, nx, ny)])
        kid = (L + 1, 2 * nx + (0 if dx > 0 else 1), 2 * ny + (0 if dy > 0 else 1))
        if kid in self.index:
            return ("fine", self.index[kid])
        vx, vy = gx + (1 if dx > 0 else 0), gy + (1 if dy > 0 else 0)
        if vx % 2 == 0 and vy % 2 == 0 and (L - 1, nx >> 1, ny >> 1) in self.index:
            return ("coarse", self.index[(L - 1, nx >> 1, ny >> 1)])
        return ("none", None)

    def has_fine(self, k):
        return any(self.face(k, f)[0] == "fine" for f in range(4)) or any(
            self.corner(k, c)[0] == "fine" for c in range(4))

    def patches_at(self, L):
        return [k for k, l in enumerate(self.leaves) if l[0] == L]

    def _off(self, kc, kf):
        (Lc, xc, yc), (Lf, xf, yf) = self.leaves[kc], self.leaves[kf]
        return (2 * xc - xf) * self.mx, (2 * yc - yf) * self.my

    # ---------------- ghost fill phases ----------------------------------
    def _args(self):
        return self.mx, self.my, self.mbc, self.meqn

    def copy2ghost(self, L, ti):
        for k in self.patches_at(L):
            qa = self.qti if ti else self.q
            for f in range(4):
                kind, nb = self.face(k, f)
                if kind == "same":
                    (_, xp, yp), (_, xn, yn) = self.leaves[k], self.leaves[nb]
                    self.L.orc_copy_face(*self._args(), _p(qa[k]), _p(qa[nb]), f,
                                         (xp - xn) * self.mx, (yp - yn) * self.my)
        for k in self.patches_at(L):
            qa = self.qti if ti else self.q
            for c in range(4):
                kind, nb = self.corner(k, c)
                if kind == "same":
                    (_, xp, yp), (_, xn, yn) = self.leaves[k], self.leaves[nb]
                    self.L.orc_copy_corner(*self._args(), _p(qa[k]), _p(qa[nb]), c,
                                           (xp - xn) * self.mx, (yp - yn) * self.my)

    def average2ghost(self, L, ti):
        for k in self.patches_at(L):
            qc = self.qti[k] if ti else self.q[k]
            for f in range(4):
                kind, nb = self.face(k, f)
                if kind == "fine":
                    for kf in nb:
                        ox, oy = self._off(k, kf)
                        self.L.orc_average_face(*self._args(), _p(qc), _p(self.q[kf]), f, ox, oy)
        for k in self.patches_at(L):
            qc = self.qti[k] if ti else self.q[k]
            for c in range(4):
                kind, nb = self.corner(k, c)
                if kind == "fine":
                    ox, oy = self._off(k, nb)
                    self.L.orc_average_corner(*self._args(), _p(qc), _p(self.q[nb]), c, ox, oy)

    def interpolate2ghost(self, L, ti):
        for k in self.patches_at(L):
            qc = self.qti[k] if ti else self.q[k]
            for f in range(4):
                kind, nb = self.face(k, f)
                if kind == "fine":
                    for kf in nb:
                        ox, oy = self._off(k, kf)
                        self.L.orc_interpolate_face(*self._args(), _p(qc), _p(self.q[kf]), f, ox, oy)
        for k in self.patches_at(L):
            qc = self.qti[k] if ti else self.q[k]
            for c in range(4):
                kind, nb = self.corner(k, c)
                if kind == "fine":
                    ox, oy = self._off(k, nb)
                    self.L.orc_interpolate_corner(*self._args(), _p(qc), _p(self.q[nb]), c, ox, oy)

    def setphysical(self, L, ti):
        for k in self.patches_at(L):
            mthbc = []
            for f in range(4):
                kind, _ = self.face(k, f)
                mthbc.append(self.mthbc[f] if kind == "bdry" else -1)
            qa = self.qti if ti else self.q
            qa[k][...] = O.bc2(qa[k], self.mx, self.my, self.mbc, mthbc)

    def fill_physical(self, lo, hi, ti):
        for L in range(hi, lo - 1, -1):
            self.setphysical(L, 0)
        if ti and lo - 1 >= 0:
            self.setphysical(lo - 1, 1)

    def ghost_update(self, minlevel, maxlevel, time_interp):
        mincoarse, maxcoarse = minlevel, maxlevel - 1
        tl = minlevel - 1
        for L in range(maxcoarse, mincoarse - 1, -1):
            self.average2ghost(L, 0)
        if time_interp:
            self.average2ghost(tl, 1)
        for L in range(maxlevel, minlevel - 1, -1):
            self.copy2ghost(L, 0)
        if time_interp:
            self.copy2ghost(tl, 1)
        self.fill_physical(mincoarse, maxcoarse, time_interp)
        for L in range(maxcoarse, mincoarse - 1, -1):
            self.interpolate2ghost(L, 0)
        if time_interp:
            self.interpolate2ghost(tl, 1)
        self.fill_physical(mincoarse + 1, maxlevel, time_interp)
        self.fill_physical(minlevel, maxlevel, time_interp)

    # ---------------- time stepping --------------------------------------
    def update_level(self, L, t, dt):
        ks = self.patches_at(L)
        if not ks:
            return 0.0
        for k in ks:
            self.qlast[k][...] = self.q[k]
        if self.time_sync:
            cfl = 0.0
            for k in ks:
                cfl = max(cfl, O.step2_ts(self.cfg, self.q[k], self.geom[k, 2], self.geom[k, 3],
                                          dt, self.reg[k]))
            self._src2(ks, dt)
            self.cell_updates += len(ks) * self.mx * self.my
            return cfl
        qb = np.stack([self.q[k] for k in ks])
        aux = None
        if self.cfg.solver == O.SOLVER_ADV:
            aux = np.stack([O.swirl_aux(self.mx, self.my, self.mbc, *self.geom[k, :4], t, dt,
                                        self.swirl_tperiod) for k in ks])
        cfl = O.step2_batch(self.cfg, qb, aux, self.geom[ks, 2], self.geom[ks, 3], dt)
        for i, k in enumerate(ks):
            self.q[k][...] = qb[i]
        self._src2(ks, dt)
        self.cell_updates += len(ks) * self.mx * self.my
        return float(cfl.max())

    def _src2(self, ks, dt):
        if self.src2_gamma is None:
            return
        for k in ks:
            rad = SRC.setaux_rad(self.mx, self.my, self.mbc, self.geom[k, 1], self.geom[k, 3])[0]
            self.q[k][...] = SRC.src2_axisym(self.q[k], rad, mbc=self.mbc, dt=dt, gamma=self.src2_gamma)

    # ---------------- conservative time synchronisation ----------------
    def _reset(self, k, kinds):
        for f in range(4):
            if self.face(k, f)[0] in kinds:
                self.L.orc_time_sync_reset_face(self.mx, self.my, self.meqn, f, _p(self.reg[k]))

    def sync(self, minlevel, maxlevel):
        """fclaw2d_time_sync(minlevel, maxlevel) -> correct_coarse_cells."""
        a = self._args()
        for L in range(maxlevel, minlevel - 1, -1):
            for k in self.patches_at(L):
                area = self.geom[k, 2] * self.geom[k, 3]
                for f in range(4):
                    kind, nb = self.face(k, f)
                    if kind == "same":
                        self.L.orc_time_sync_samesize(*a, f, area, _p(self.q[k]),
                                                      _p(self.reg[k]), _p(self.reg[nb]))
            for k in self.patches_at(L):
                self._reset(k, ("same",))
            if L < maxlevel:
                for k in self.patches_at(L):
                    area = self.geom[k, 2] * self.geom[k, 3]
                    for f in range(4):
                        kind, nb = self.face(k, f)
                        if kind == "fine":
                            for g in (0, 1):
                                self.L.orc_time_sync_f2c(*a, f, g, area, _p(self.q[k]),
                                                         _p(self.reg[k]), _p(self.reg[nb[g]]))
                for k in self.patches_at(L):
                    self._reset(k, ("fine",))
                for k in self.patches_at(L + 1):
                    self._reset(k, ("coarse",))
        for L in range(maxlevel, minlevel - 1, -1):
            for k in self.patches_at(L):
                self._reset(k, ("bdry",))

    def timeinterp(self, L, alpha):
        for k in self.patches_at(L):
            if self.has_fine(k):
                self.L.orc_timeinterp(*self._args(), _p(self.q[k]), _p(self.qlast[k]),
                                      _p(self.qti[k]), alpha)

    def advance(self, t, dt):
        """fclaw2d_advance_all_levels (src/fclaw2d_advance.c:240-357) with the
        counters of initialize_timestep_counters (:54-118), which count from
        the options' minlevel m0: dt is the step of level m0, level L steps
        2^(L - m0) times (subcycling), or every level 2^(maxlevel - minlevel)
        times (no subcycling)."""
        lo, hi = min(self.levels), max(self.levels)
        m0 = min(self.minlevel, lo)
        rf = 1 << max(self.maxlevel - self.minlevel, hi - lo)
        c = {}
        d = dt
        for L in range(m0, hi + 1):
            if self.subcycle:
                c[L] = dict(last=0, inc=1 << (hi - L), total=1 << (L - m0), t=t, dt=d)
                d /= 2
            else:
                c[L] = dict(last=0, inc=1, total=rf, t=t, dt=dt / rf)
        maxcfl = 0.0

        def advance_level(level, curr):
            nonlocal maxcfl
            maxcfl = max(maxcfl, self.update_level(level, c[level]["t"], c[level]["dt"]))
            c[level]["last"] += c[level]["inc"]
            c[level]["t"] += c[level]["dt"]
            if level > lo and c[level - 1]["last"] == curr:
                advance_level(level - 1, c[level - 1]["last"])
                if self.subcycle:
                    inc = c[level - 1]["inc"]
                    self.timeinterp(level - 1, (c[level]["last"] % inc) / inc)

        for nf in range(c[hi]["total"]):
            advance_level(hi, nf)
            if c[hi]["last"] < c[hi]["total"]:
                if self.subcycle:
                    tl = hi - 1
                    while c[hi]["last"] % c[tl]["inc"] == 0:
                        tl -= 1
                    if tl < lo:  # all present levels synchronised
                        self.ghost_update(lo, hi, 0)
                        if self.time_sync:
                            self.sync(lo, hi)
                    else:
                        self.ghost_update(tl + 1, hi, 1)
                        if self.time_sync:
                            self.sync(tl + 1, hi)
                else:
                    self.ghost_update(lo, hi, 0)
                    if self.time_sync:
                        self.sync(lo, hi)
        self.ghost_update(lo, hi, 0)
        if self.time_sync:
            self.sync(lo, hi)
        return maxcfl
