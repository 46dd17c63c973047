"""Host front end of the device-resident AMR solver (hipclaw_ctx_* C ABI).

Mirrors the ForestClaw objects a user of the reference works with:
  Options              [Options] / [clawpatch] / [clawpack46] groups
                       (src/fclaw_options.c, fc2d_clawpack46_options.c)
  Hipclaw46.initialize fclaw2d_initialize (qinit + initial refinement)
  Hipclaw46.advance    fclaw2d_advance_all_levels
  Hipclaw46.run        fclaw2d_run outstyle 1 (dt from the CFL number,
                       step rejection when maxcfl > max_cfl, regrid cadence)
Every numerical operation runs in libhipclaw.so on the GPU.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import hipclaw as H


Domain = H.Domain
QINIT = {"const": 0, "swirl": 1, "radial_dam": 2, "humps": 3, "shock_bubble": 4, "shockbubble_app": 5}
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p


def _lib():
    return H.lib()


@dataclass
class Options:
    """The option names of the reference's .ini sections that the hot path uses."""
    # [Options]
    minlevel: int = 0
    maxlevel: int = 0
    regrid_interval: int = 1
    refine_threshold: float = 0.5
    coarsen_threshold: float = 0.1
    smooth_refine: bool = False
    smooth_level: int = 0
    init_ghostcell: bool = False
    subcycle: bool = True
    time_sync: bool = False
    max_cfl: float = 1.0
    desired_cfl: float = 0.9
    use_fixed_dt: bool = False
    initial_dt: float = 0.01
    ax: float = 0.0
    bx: float = 1.0
    ay: float = 0.0
    by: float = 1.0
    mi: int = 1
    mj: int = 1
    # [clawpatch]
    mx: int = 32
    my: int = 32
    mbc: int = 2
    threshold_variable: int = 0   # 0-based (the reference's threshold-variable is 1-based)
    refinement_criteria: str = "gradient"   # value | difference | minmax | gradient
    # [clawpack46]
    solver: int = H.SOLVER_SWE_ROE
    order: tuple = (2, 2)
    mthlim: tuple | None = None
    mthbc: tuple = (1, 1, 1, 1)
    src2: int = H.SRC2_NONE       # the application's fort_src2 when src_term = 1
    claw_version: int = 4         # [user] claw-version: 4 (fc2d_clawpack46) or 5 (fc2d_clawpack5)
    grav: float = 1.0
    gamma: float = 1.4
    precision: int = H.FP64


def _check(rc):
    H._check(rc)


class Hipclaw46:
    """One device-resident simulation (one hipclaw_ctx)."""

    def __init__(self, opt: Options, stream: int | None = None):
        self.opt = opt
        L = _lib()
        if opt.claw_version not in (4, 5):
            raise ValueError("claw_version must be 4 or 5")
        solver = opt.solver | (H.CLAW5 if opt.claw_version == 5 else 0)
        self.phys = H.physics(solver, opt.mx, opt.my, precision=opt.precision, order=opt.order,
                              mthlim=opt.mthlim, mthbc=opt.mthbc, grav=opt.grav, gamma=opt.gamma)
        self.dom = Domain(opt.mi, opt.mj, opt.ax, opt.bx, opt.ay, opt.by, opt.minlevel,
                          opt.maxlevel, 1 if opt.subcycle else 0)
        h = _VP()
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream
        _check(L.hipclaw_ctx_create(ctypes.byref(self.phys), ctypes.byref(self.dom), stream,
                                    ctypes.byref(h)))
        self.h = h
        _check(L.hipclaw_set_tagging(self.h, opt.threshold_variable, opt.refine_threshold,
                                     opt.coarsen_threshold))
        _check(L.hipclaw_set_tag_criterion(self.h, H.TAG_CRITERIA[opt.refinement_criteria]))
        _check(L.hipclaw_set_refinement(self.h, int(opt.smooth_refine), opt.smooth_level, 0))
        _check(L.hipclaw_set_feature(self.h, H.FEATURE_INIT_GHOSTCELL, int(opt.init_ghostcell)))
        if opt.time_sync:
            _check(L.hipclaw_set_time_sync(self.h, 1))
        if opt.src2:
            _check(L.hipclaw_set_src2(self.h, opt.src2))
        self.t = 0.0
        self.n_inner = 0
        # the reference's regression counters (src/fclaw2d_advance.c:354,
        # src/fclaw2d_regrid.c:315): advance calls (retaken steps included)
        # and regrids that changed the mesh
        self.count_amr_advance = 0
        self.count_amr_new_domain = 0
        self.retakes = 0
        self.dt_history = []

    def close(self):
        if self.h:
            _lib().hipclaw_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- forest / data ----------------------------------------------------
    @property
    def tile_shape(self):
        """Shape of one host patch array: (meqn, NY, NX) for Clawpack 4.6,
        q(i, j, m) in Fortran order; (NY, NX, meqn) for Clawpack 5, q(m, i, j)."""
        ny, nx = self.opt.my + 2 * self.opt.mbc, self.opt.mx + 2 * self.opt.mbc
        if self.opt.claw_version == 5:
            return (ny, nx, self.phys.meqn)
        return (self.phys.meqn, ny, nx)

    def num_patches(self) -> int:
        n = ctypes.c_int()
        _check(_lib().hipclaw_num_patches(self.h, ctypes.byref(n)))
        return n.value

    def patch_info(self):
        n = self.num_patches()
        lev, gx, gy = (np.zeros(n, np.int32) for _ in range(3))
        xl, yl, dx, dy = (np.zeros(n) for _ in range(4))
        ip = lambda a: a.ctypes.data_as(_IP)  # noqa: E731
        dp = lambda a: a.ctypes.data_as(_DP)  # noqa: E731
        _check(_lib().hipclaw_patch_info(self.h, ip(lev), ip(gx), ip(gy), dp(xl), dp(yl), dp(dx),
                                         dp(dy)))
        return dict(level=lev, gx=gx, gy=gy, xlower=xl, ylower=yl, dx=dx, dy=dy)

    def init_uniform(self, level: int):
        _check(_lib().hipclaw_init_uniform(self.h, level))

    def init_leaves(self, level, gx, gy):
        """Forest from an explicit leaf list; patches come back renumbered."""
        a = [np.ascontiguousarray(v, dtype=np.int32) for v in (level, gx, gy)]
        if not (len(a[0]) == len(a[1]) == len(a[2])):
            raise ValueError("level, gx, gy must have equal lengths")
        _check(_lib().hipclaw_init_leaves(self.h, len(a[0]), *(v.ctypes.data_as(_IP) for v in a)))

    def set_time_sync(self, on: bool = True):
        """Conservative time synchronisation (SWE): FP64 registers per patch,
        corrections after every synchronising ghost update of advance()."""
        _check(_lib().hipclaw_set_time_sync(self.h, int(bool(on))))

    def set_feature(self, feature: int, on: bool):
        """H.FEATURE_MERGE_LEVELS / H.FEATURE_GHOST_PUSH (both on by default)."""
        _check(_lib().hipclaw_set_feature(self.h, feature, int(bool(on))))

    def graph_stats(self):
        """(coarse steps captured as graphs, graph replays)."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        _check(_lib().hipclaw_graph_stats(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def time_sync(self, minlevel: int, maxlevel: int):
        _check(_lib().hipclaw_time_sync(self.h, minlevel, maxlevel))

    def get_registers(self) -> np.ndarray:
        n = self.num_patches()
        r = np.zeros((n, 8 * self.phys.meqn * (self.opt.mx + self.opt.my)))
        _check(_lib().hipclaw_get_registers(self.h, r.ctypes.data_as(_DP)))
        return r

    def make_tiles(self, tiles, tags=None):
        """Network input of every patch (torch bf16 [n, 36, 36, 8] on the device);
        tags (torch int32 [n], optional) receives the plain refinement tags."""
        _check(_lib().hipclaw_make_tiles(self.h, H._ptr(tiles), H._ptr(tags)))

    def set_tagging_cnn(self, net, margin: float = 0.0):
        """Refinement tags of regrid from the network (None: plain tags)."""
        _check(_lib().hipclaw_set_tagging_cnn(self.h, None if net is None else net.h, margin))

    def tag_stats(self):
        """patches, network refines, plain refines, misses, false refines, fallbacks."""
        st = (ctypes.c_int * 6)()
        _check(_lib().hipclaw_tag_stats(self.h, st))
        return dict(zip(("patches", "net_refine", "plain_refine", "misses", "false_refines",
                         "fallbacks"), list(st)))

    def output_ascii(self, iframe: int, directory: str = ".", time: float | None = None):
        """fort.tXXXX / fort.qXXXX frame of the current solution (the reference's
        ascii-out output)."""
        _check(_lib().hipclaw_output_ascii(self.h, iframe, self.t if time is None else time,
                                           directory.encode()))

    def set_launch(self, cols_per_lane: int | None = None, waves_per_group: int | None = None,
                   row_bands: int = -1):
        """Patch-step launch shape for this context (None: library default).
        row_bands -1 keeps the library's shape for launches of fewer than
        2,048 patches; 0 sweeps whole patches; 2..8 bands."""
        if cols_per_lane is None:
            _check(_lib().hipclaw_set_launch(self.h, None))
        else:
            lc = H.Launch(cols_per_lane, waves_per_group, row_bands)
            _check(_lib().hipclaw_set_launch(self.h, ctypes.byref(lc)))

    def get_launch(self):
        """(cols_per_lane, waves_per_group, row_bands); cols_per_lane 0 = library default."""
        lc = H.Launch()
        _check(_lib().hipclaw_get_launch(self.h, ctypes.byref(lc)))
        return lc.cols_per_lane, lc.waves_per_group, lc.row_bands

    def autotune(self, level: int = -1, reps: int = 3):
        """Time the six launch shapes on one level and keep the fastest.
        Returns ((cols_per_lane, waves_per_group), {shape: ms})."""
        best = H.Launch()
        ms = (ctypes.c_float * 6)()
        _check(_lib().hipclaw_autotune(self.h, level, reps, ctypes.byref(best), ms))
        shapes = [(c, w) for c in (1, 2) for w in (1, 2, 4)]
        return (best.cols_per_lane, best.waves_per_group), dict(zip(shapes, list(ms)))

    def level_upload(self, level: int, tiles, aux_tiles=None):
        """tiles: one C-contiguous float64 array per patch of the level, slot order."""
        n = len(tiles)
        arr = (_DP * n)(*[t.ctypes.data_as(_DP) for t in tiles])
        aarr = None
        if aux_tiles is not None:
            aarr = (_DP * n)(*[t.ctypes.data_as(_DP) for t in aux_tiles])
        _check(_lib().hipclaw_level_upload(self.h, level, n, arr, aarr))

    def level_download(self, level: int, tiles):
        n = len(tiles)
        arr = (_DP * n)(*[t.ctypes.data_as(_DP) for t in tiles])
        _check(_lib().hipclaw_level_download(self.h, level, n, arr))

    def set_q(self, q: np.ndarray):
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.shape == (self.num_patches(),) + self.tile_shape
        _check(_lib().hipclaw_set_q(self.h, q.ctypes.data_as(_DP)))

    def get_q(self) -> np.ndarray:
        q = np.zeros((self.num_patches(),) + self.tile_shape)
        _check(_lib().hipclaw_get_q(self.h, q.ctypes.data_as(_DP)))
        return q

    def set_swirl(self, tperiod: float):
        _check(_lib().hipclaw_set_swirl(self.h, tperiod))

    def qinit(self, kind: str, params=(), nsub: int = 4):
        p = np.ascontiguousarray(params, dtype=np.float64)
        _check(_lib().hipclaw_qinit(self.h, QINIT[kind], p.ctypes.data_as(_DP), len(p), nsub))

    def initialize(self, kind: str, params=(), nsub: int = 4):
        """fclaw2d_initialize: qinit at minlevel, refine while tagged, qinit again."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        _check(_lib().hipclaw_initialize(self.h, QINIT[kind], p.ctypes.data_as(_DP), len(p), nsub))

    # ---- time stepping ----------------------------------------------------
    def ghost_update(self, minlevel, maxlevel, time_interp=0, t=0.0):
        _check(_lib().hipclaw_ghost_update(self.h, minlevel, maxlevel, t, time_interp))

    def update_level(self, level, t, dt) -> float:
        c = ctypes.c_double()
        _check(_lib().hipclaw_update_level(self.h, level, t, dt, ctypes.byref(c)))
        return c.value

    def advance(self, t: float, dt: float, want_cfl: bool = True) -> float:
        c = ctypes.c_double(0.0)
        _check(_lib().hipclaw_advance(self.h, t, dt, ctypes.byref(c) if want_cfl else None))
        self.count_amr_advance += 1
        return c.value

    def advance_steps(self, t: float, dt: float, nsteps: int) -> np.ndarray:
        """nsteps fixed-dt coarse steps from t (one graph launch when graphs
        are on); the maximum Courant number of each, read back once."""
        c = np.zeros(nsteps)
        _check(_lib().hipclaw_advance_steps(self.h, t, dt, nsteps, c.ctypes.data_as(_DP)))
        self.count_amr_advance += nsteps
        return c

    def prepare_steps(self, nsteps: int):
        """Capture the graph of advance_steps(nsteps) from the current state
        without running it."""
        _check(_lib().hipclaw_prepare_steps(self.h, nsteps))

    def regrid(self, init_flag: int = 0) -> bool:
        ch = ctypes.c_int()
        _check(_lib().hipclaw_regrid(self.h, init_flag, ctypes.byref(ch)))
        self.count_amr_new_domain += int(ch.value != 0)
        return bool(ch.value)

    def save_step(self):
        """save_time_step: device copy of the current solution."""
        _check(_lib().hipclaw_save_step(self.h))

    def restore_step(self):
        """restore_time_step: back to the last saved solution (same mesh)."""
        _check(_lib().hipclaw_restore_step(self.h))

    def mass(self) -> np.ndarray:
        s = np.zeros(self.phys.meqn)
        _check(_lib().hipclaw_mass(self.h, s.ctypes.data_as(_DP)))
        return s

    def cell_updates(self) -> int:
        n = ctypes.c_longlong()
        _check(_lib().hipclaw_cell_updates(self.h, ctypes.byref(n)))
        return n.value

    def profile(self, enable: bool = True):
        _check(_lib().hipclaw_profile(self.h, 1 if enable else 0))

    def profile_read(self):
        ms = ctypes.c_double()
        nl, nc = ctypes.c_longlong(), ctypes.c_longlong()
        _check(_lib().hipclaw_profile_read(self.h, ctypes.byref(ms), ctypes.byref(nl),
                                           ctypes.byref(nc)))
        return ms.value, nl.value, nc.value

    def run(self, tfinal: float, nout: int = 1, dt0: float | None = None, max_steps: int = 10**9,
            on_step=None, batch: int = 64):
        """fclaw2d_run outstyle 1 (src/fclaw2d_run.c:119-297): variable dt from the
        CFL number unless use_fixed_dt; reject and retake a step whose maxcfl
        exceeds max_cfl (the solution is restored from the device copy that
        save_step made before the step, src/fclaw2d_run.c:155-158, 220-236).
        With use_fixed_dt no step depends on the previous one's Courant number,
        so the steps up to the next output time or regrid (at most batch, and
        one at a time when on_step is given) go to the device as one call.
        self.retakes counts rejected steps; self.dt_history records every
        attempted (dt, Courant number)."""
        o = self.opt
        dt_min = dt0 if dt0 is not None else o.initial_dt
        t0 = self.t
        dt_outer = (tfinal - t0) / nout
        steps = 0
        regrid_on = o.regrid_interval > 0 and o.maxlevel > o.minlevel
        for n in range(nout):
            tend = t0 + (n + 1) * dt_outer
            while self.t < tend and steps < max_steps:
                if o.use_fixed_dt and on_step is None and batch > 1:
                    # the steps the loop below would take before leaving
                    # the interval, reaching max_steps or a regrid
                    k, tt = 0, self.t
                    while k < min(batch, max_steps - steps) and tt < tend:
                        k += 1
                        tt += dt_min
                        if regrid_on and (self.n_inner + k) % o.regrid_interval == 0:
                            break
                    cfls = self.advance_steps(self.t, dt_min, k)
                    self.dt_history.extend((dt_min, float(c)) for c in cfls)
                    self.n_inner += k
                    steps += k
                    self.t = tt
                    if regrid_on and self.n_inner % o.regrid_interval == 0:
                        self.regrid()
                    continue
                if not o.use_fixed_dt:
                    self.save_step()
                dt_step = dt_min
                tol = 1e-2 * dt_step
                took_small = False
                if not o.use_fixed_dt:
                    small = tend - (self.t + dt_step)
                    if small < tol:
                        dt_step = tend - self.t
                        took_small = small < 0
                cfl = self.advance(self.t, dt_step)
                self.dt_history.append((dt_step, cfl))
                if cfl > o.max_cfl and not o.use_fixed_dt:
                    self.restore_step()
                    self.retakes += 1
                    dt_min = dt_min * o.desired_cfl / cfl
                    continue
                self.n_inner += 1
                steps += 1
                self.t += dt_step
                if not o.use_fixed_dt and not took_small:
                    dt_min = dt_min * o.desired_cfl / cfl
                if on_step:
                    on_step(self, cfl, dt_step)
                if regrid_on and self.n_inner % o.regrid_interval == 0:
                    self.regrid()
        return dt_min
