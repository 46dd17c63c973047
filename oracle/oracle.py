"""ctypes front end of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; the product path
(forestclaw_amd) never imports it.

Arrays use the clawpatch tile layout: shape (meqn, my+2mbc, mx+2mbc), C order,
i.e. q(i, j, m) with i fastest (src/patches/clawpatch/fclaw2d_clawpatch.cpp:305-335).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SOLVER_ADV, SOLVER_SWE_ROE, SOLVER_EULER5_ROE, SOLVER_EULER_HLLE, SOLVER_TRANSPORT_FW = 0, 1, 2, 3, 4


class OrcCfg(ctypes.Structure):
    _fields_ = [
        ("solver", ctypes.c_int),
        ("mx", ctypes.c_int), ("my", ctypes.c_int), ("mbc", ctypes.c_int),
        ("meqn", ctypes.c_int), ("maux", ctypes.c_int), ("mwaves", ctypes.c_int),
        ("order2", ctypes.c_int), ("transverse", ctypes.c_int),
        ("mthlim", ctypes.c_int * 8),
        ("efix", ctypes.c_int),
        ("grav", ctypes.c_double), ("gamma", ctypes.c_double),
        ("claw5", ctypes.c_int),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_step2.argtypes = [ctypes.POINTER(OrcCfg), dp, dp, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double, dp]
        L.orc_step2_batch.argtypes = [ctypes.POINTER(OrcCfg), ctypes.c_int, dp, dp, dp,
                                      dp, ctypes.c_double, dp]
        L.orc_swirl_aux.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 7 + [dp]
        L.orc_swirl_aux.restype = None
        L.orc_bc2.argtypes = [ctypes.c_int] * 4 + [ip, dp]
        L.orc_limiter_phi.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_limiter_phi.restype = ctypes.c_double
        _lib = L
    return _lib


_DPT = ctypes.POINTER(ctypes.c_double)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_DPT)


def make_cfg(solver: int, mx: int, my: int, mbc: int = 2, *, maux: int | None = None,
             order=(2, 2), mthlim=None, grav: float = 1.0, gamma: float = 1.4,
             efix: int = 1, claw5: bool = False) -> OrcCfg:
    meqn = {SOLVER_ADV: 1, SOLVER_SWE_ROE: 3, SOLVER_EULER5_ROE: 5, SOLVER_EULER_HLLE: 5,
            SOLVER_TRANSPORT_FW: 1}[solver]
    mwaves = 2 if solver == SOLVER_EULER_HLLE else meqn
    if maux is None:
        maux = {SOLVER_ADV: 2, SOLVER_TRANSPORT_FW: 7}.get(solver, 0)
    c = OrcCfg()
    c.solver, c.mx, c.my, c.mbc = solver, mx, my, mbc
    c.meqn, c.maux, c.mwaves = meqn, maux, mwaves
    c.order2, c.transverse = order
    lim = list(mthlim) if mthlim is not None else [4] * mwaves
    for k in range(8):
        c.mthlim[k] = lim[k] if k < len(lim) else 0
    c.efix = efix
    c.grav, c.gamma = grav, gamma
    c.claw5 = int(claw5)
    return c


def step2(cfg: OrcCfg, q: np.ndarray, aux: np.ndarray | None, dx: float, dy: float,
          dt: float):
    """One patch step on a copy of q; returns (q_new, cfl)."""
    qn = np.ascontiguousarray(q, dtype=np.float64).copy()
    a = None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    cfl = ctypes.c_double(0.0)
    rc = lib().orc_step2(ctypes.byref(cfg), _dp(qn), None if a is None else _dp(a),
                         dx, dy, dt, ctypes.byref(cfl))
    if rc != 0:
        raise ValueError("oracle rejected the configuration")
    return qn, cfl.value


def step2_batch(cfg: OrcCfg, q: np.ndarray, aux: np.ndarray | None, dx, dy, dt: float):
    """Step npatch tiles in place (q shape (npatch, meqn, NY, NX)); returns per-patch CFL."""
    assert q.dtype == np.float64 and q.flags.c_contiguous
    n = q.shape[0]
    dxa = np.broadcast_to(np.asarray(dx, dtype=np.float64), (n,)).copy()
    dya = np.broadcast_to(np.asarray(dy, dtype=np.float64), (n,)).copy()
    cfl = np.zeros(n)
    a = None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
    rc = lib().orc_step2_batch(ctypes.byref(cfg), n, _dp(q), None if a is None else _dp(a),
                               _dp(dxa), _dp(dya), dt, _dp(cfl))
    if rc != 0:
        raise ValueError("oracle rejected the configuration")
    return cfl


def reg_size(mx, my, meqn) -> int:
    """Doubles of time-sync registers per patch (4 faces x 4 blocks x meqn x length)."""
    return 8 * meqn * (mx + my)


def swe_flux(idir: int, grav: float, qv) -> np.ndarray:
    L = lib()
    L.orc_swe_flux.argtypes = [ctypes.c_int, ctypes.c_double, _DPT, _DPT]
    L.orc_swe_flux.restype = None
    qv = np.ascontiguousarray(qv, dtype=np.float64)
    out = np.zeros(qv.shape[0])
    L.orc_swe_flux(idir, grav, _dp(qv), _dp(out))
    return out


def step2_ts(cfg: OrcCfg, q: np.ndarray, dx: float, dy: float, dt: float, reg: np.ndarray):
    """One SWE patch step in place that also accumulates the time-sync
    registers (store_flux before, accumulate_waves after); returns the CFL."""
    L = lib()
    L.orc_step2_ts.argtypes = [ctypes.c_void_p, _DPT, _DPT, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, _DPT, _DPT]
    L.orc_step2_ts.restype = ctypes.c_int
    assert q.dtype == np.float64 and q.flags.c_contiguous
    assert reg.dtype == np.float64 and reg.flags.c_contiguous
    cfl = np.zeros(1)
    rc = L.orc_step2_ts(ctypes.byref(cfg), _dp(q), None, dx, dy, dt, _dp(cfl), _dp(reg))
    if rc != 0:
        raise ValueError("oracle rejected the configuration (time sync needs SWE)")
    return float(cfl[0])


def swirl_aux(mx, my, mbc, xlower, ylower, dx, dy, t, dt, tperiod):
    aux = np.zeros((2, my + 2 * mbc, mx + 2 * mbc))
    lib().orc_swirl_aux(mx, my, mbc, xlower, ylower, dx, dy, t, dt, tperiod, _dp(aux))
    return aux


def bc2(q: np.ndarray, mx, my, mbc, mthbc):
    qn = np.ascontiguousarray(q, dtype=np.float64).copy()
    mb = (ctypes.c_int * 4)(*mthbc)
    rc = lib().orc_bc2(mx, my, mbc, qn.shape[0], mb, _dp(qn))
    if rc != 0:
        raise ValueError("bc2: mthbc 0 (user) and 2 (periodic) are rejected")
    return qn


def rpn2_hlle_face(gamma: float, ixy: int, qa, qb):
    """HLLE at one face (qa low side, qb high side; ixy 1 x, 2 y): returns
    (waves (2, 5), speeds (2,), amdq, apdq)."""
    L = lib()
    L.orc_rpn2_hlle_face.argtypes = [ctypes.c_double, ctypes.c_int] + [_DPT] * 7
    L.orc_rpn2_hlle_face.restype = None
    qa = np.ascontiguousarray(qa, dtype=np.float64)
    qb = np.ascontiguousarray(qb, dtype=np.float64)
    w, sp, am, ap, r = np.zeros(10), np.zeros(2), np.zeros(5), np.zeros(5), np.zeros(7)
    L.orc_rpn2_hlle_face(gamma, ixy, _dp(qa), _dp(qb), _dp(w), _dp(sp), _dp(am), _dp(ap), _dp(r))
    return w.reshape(2, 5), sp, am, ap


def rpt2_hlle_face(gamma: float, ixy: int, qa, qb, asdq):
    """Transverse split of the HLLE pair of the face (qa, qb): (bmasdq, bpasdq)
    of the fluctuation asdq."""
    L = lib()
    L.orc_rpn2_hlle_face.argtypes = [ctypes.c_double, ctypes.c_int] + [_DPT] * 7
    L.orc_rpt2_hlle_face.argtypes = [ctypes.c_int] + [_DPT] * 4
    L.orc_rpt2_hlle_face.restype = None
    qa = np.ascontiguousarray(qa, dtype=np.float64)
    qb = np.ascontiguousarray(qb, dtype=np.float64)
    d = np.ascontiguousarray(asdq, dtype=np.float64)
    w, sp, am, ap, r = np.zeros(10), np.zeros(2), np.zeros(5), np.zeros(5), np.zeros(7)
    L.orc_rpn2_hlle_face(gamma, ixy, _dp(qa), _dp(qb), _dp(w), _dp(sp), _dp(am), _dp(ap), _dp(r))
    bm, bp = np.zeros(5), np.zeros(5)
    L.orc_rpt2_hlle_face(ixy, _dp(r), _dp(d), _dp(bm), _dp(bp))
    return bm, bp


def limiter_phi(mthlim: int, r: float) -> float:
    return lib().orc_limiter_phi(mthlim, r)
