"""Drive the reference's own Fortran patch step (compiled unchanged from
/root/reference into oracle/_ref/ by oracle/Makefile) through ctypes.

Used only in the build container, by make_golden.py to write fixtures and by
the CPU tests to cross-check the oracle when oracle/_ref exists.  Nothing here
runs on a GPU box unless the libraries were built.

Entry points driven (reference file:line):
  CLAWPACK46_STEP2_WRAP  src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_step2_wrap.f:1
  (work sizes as fc2d_clawpack46.cpp:345-352, method[] as fc2d_clawpack46_options.c:93-99)
  clawpack46_b4step2     applications/clawpack/advection/2d/swirl/user_4.6/b4step2.f:1
  clawpack46_bc2_default src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_bc2_default.f:1
  fclaw2d_clawpatch46_fort_timeinterp src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_timeinterp.f:11
  fclaw2d_clawpatch46_fort_header_ascii / _output_ascii
                         src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_output_ascii.f:13,51
  clawpack46_src2 / clawpack46_setaux of the shock-bubble application
                         applications/clawpack/euler/2d/shockbubble/user_4.6/src2.f:1, setaux.f:1
Common blocks (/cparam/, /compi/, /comvt/) are written directly as data.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_DIR = os.path.join(REPO, "oracle", "_ref")

_ci = ctypes.c_int
_cd = ctypes.c_double
_libs: dict[str, ctypes.CDLL] = {}

RP = {  # physics -> (library, rpn2 symbol, rpt2 symbol)
    "adv": ("libref_adv.so", "clawpack46_rpn2adv_", "clawpack46_rpt2adv_"),
    "swe": ("libref_swe.so", "clawpack46_rpn2_", "clawpack46_rpt2_"),
    "euler": ("libref_euler.so", "clawpack46_rpn2_euler5_", "clawpack46_rpt2_euler5_"),
}
# the f-wave step: clawpack46_flux2fw with the transport applications'
# f-wave pair (applications/clawpack/transport/2d/all/clawpack46_rpn2cons_fw_manifold.f,
# clawpack46_rpt2cons_manifold.f), passed as rpn2fw / rpt2fw with use_fwaves = 1
RP_FW = ("libref_fw.so", "clawpack46_rpn2cons_fw_manifold_", "clawpack46_rpt2cons_manifold_")


def available() -> bool:
    return all(os.path.exists(os.path.join(REF_DIR, RP[k][0])) for k in RP)


def _lib(name: str) -> ctypes.CDLL:
    if name not in _libs:
        _libs[name] = ctypes.CDLL(os.path.join(REF_DIR, name), mode=ctypes.RTLD_LOCAL)
    return _libs[name]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_params(physics: str, *, grav=1.0, gamma=1.4, tperiod=4.0):
    L = _lib(RP[physics][0])
    if physics == "swe":
        _cd.in_dll(L, "cparam_").value = grav
    elif physics == "euler":
        arr = (_cd * 2).in_dll(L, "cparam_")
        arr[0] = gamma
        arr[1] = gamma - 1.0
    else:
        pi = 4.0 * math.atan(1.0)
        arr = (_cd * 2).in_dll(L, "compi_")
        arr[0] = pi
        arr[1] = 2.0 * pi
        _cd.in_dll(L, "comvt_").value = tperiod


def step2(physics: str, q: np.ndarray, aux: np.ndarray | None, *, mx, my, mbc, dx, dy,
          dt, order=(2, 2), mthlim=None, xlower=0.0, ylower=0.0, t=0.0):
    """Run CLAWPACK46_STEP2_WRAP on a copy of q (tile layout (meqn, NY, NX)).
    physics "fw": the f-wave step (use_fwaves = 1, flux2fw).  Returns
    (q_after, cfl)."""
    fw = physics == "fw"
    libname, rpn2_sym, rpt2_sym = RP_FW if fw else RP[physics]
    L = _lib(libname)
    meqn = q.shape[0]
    mwaves = meqn
    maux = 0 if aux is None else aux.shape[0]
    qf = np.ascontiguousarray(q, dtype=np.float64).copy()
    auxf = (np.zeros(1) if aux is None else np.ascontiguousarray(aux, dtype=np.float64).copy())
    maxm = max(mx, my)
    mwork = (maxm + 2 * mbc) * (12 * meqn + (meqn + 1) * mwaves + 3 * maux + 2)
    work = np.zeros(mwork)
    size = meqn * (mx + 2 * mbc) * (my + 2 * mbc)
    fp, fm, gp, gm = (np.zeros(size) for _ in range(4))
    method = (_ci * 7)(0, order[0], order[1], 0, 0, 0, 0)
    lim = list(mthlim) if mthlim is not None else [4] * mwaves
    mthlim_c = (_ci * mwaves)(*lim)
    bcc = (_ci * 4)(0, 0, 0, 0)
    cfl = _cd(0.0)
    ierr = _ci(0)
    rpn2 = ctypes.cast(getattr(L, rpn2_sym), ctypes.c_void_p)
    rpt2 = ctypes.cast(getattr(L, rpt2_sym), ctypes.c_void_p)
    flux2 = ctypes.cast(L.clawpack46_flux2fw_ if fw else L.clawpack46_flux2_, ctypes.c_void_p)
    L.clawpack46_set_block_(ctypes.byref(_ci(0)))
    L.clawpack46_step2_wrap_(
        ctypes.byref(_ci(maxm)), ctypes.byref(_ci(meqn)), ctypes.byref(_ci(maux)),
        ctypes.byref(_ci(mbc)), method, mthlim_c, ctypes.byref(_ci(0)),
        ctypes.byref(_ci(mwaves)), ctypes.byref(_ci(mx)), ctypes.byref(_ci(my)),
        _p(qf), _p(auxf), ctypes.byref(_cd(dx)), ctypes.byref(_cd(dy)),
        ctypes.byref(_cd(dt)), ctypes.byref(cfl), _p(work), ctypes.byref(_ci(mwork)),
        ctypes.byref(_cd(xlower)), ctypes.byref(_cd(ylower)), ctypes.byref(_ci(0)),
        ctypes.byref(_cd(t)), _p(fp), _p(fm), _p(gp), _p(gm), rpn2, rpt2, rpn2, rpt2,
        flux2, bcc, ctypes.byref(ierr), ctypes.byref(_ci(1 if fw else 0)))
    L.clawpack46_unset_block_()
    if ierr.value != 0:
        raise RuntimeError(f"step2_wrap ierror={ierr.value}")
    return qf, cfl.value


def swirl_aux(*, mx, my, mbc, xlower, ylower, dx, dy, t, dt, tperiod):
    L = _lib("libref_adv.so")
    set_params("adv", tperiod=tperiod)
    aux = np.zeros((2, my + 2 * mbc, mx + 2 * mbc))
    q = np.zeros((1, my + 2 * mbc, mx + 2 * mbc))
    L.clawpack46_b4step2_(
        ctypes.byref(_ci(mx)), ctypes.byref(_ci(my)), ctypes.byref(_ci(mbc)),
        ctypes.byref(_ci(mx)), ctypes.byref(_ci(my)), ctypes.byref(_ci(1)), _p(q),
        ctypes.byref(_cd(xlower)), ctypes.byref(_cd(ylower)), ctypes.byref(_cd(dx)),
        ctypes.byref(_cd(dy)), ctypes.byref(_cd(t)), ctypes.byref(_cd(dt)),
        ctypes.byref(_ci(2)), _p(aux))
    return aux


def bc2(q: np.ndarray, *, mx, my, mbc, mthbc):
    L = _lib("libref_swe.so")
    meqn = q.shape[0]
    qf = np.ascontiguousarray(q, dtype=np.float64).copy()
    aux = np.zeros(1)
    mb = (_ci * 4)(*mthbc)
    L.clawpack46_bc2_default_(
        ctypes.byref(_ci(mx)), ctypes.byref(_ci(my)), ctypes.byref(_ci(meqn)),
        ctypes.byref(_ci(mbc)), ctypes.byref(_ci(mx)), ctypes.byref(_ci(my)),
        ctypes.byref(_cd(0.0)), ctypes.byref(_cd(0.0)), ctypes.byref(_cd(1.0)),
        ctypes.byref(_cd(1.0)), _p(qf), ctypes.byref(_ci(0)), _p(aux),
        ctypes.byref(_cd(0.0)), ctypes.byref(_cd(0.0)), mb)
    return qf


def timeinterp(qcurr, qlast, alpha, *, mx, my, mbc):
    L = _lib("libref_timeinterp.so")
    meqn = qcurr.shape[0]
    qc = np.ascontiguousarray(qcurr, dtype=np.float64)
    ql = np.ascontiguousarray(qlast, dtype=np.float64)
    qi = np.zeros_like(qc)
    mint = mbc
    psize = (mx * my - (mx - 2 * mint) * (my - 2 * mint)) * meqn
    ierr = _ci(0)
    L.fclaw2d_clawpatch46_fort_timeinterp_(
        ctypes.byref(_ci(mx)), ctypes.byref(_ci(my)), ctypes.byref(_ci(mbc)),
        ctypes.byref(_ci(meqn)), ctypes.byref(_ci(psize)), _p(qc), _p(ql), _p(qi),
        ctypes.byref(_cd(alpha)), ctypes.byref(ierr))
    return qi, ierr.value


def ascii_frame(directory: str, iframe: int, time: float, meqn: int, maux: int, patches) -> None:
    """fort.tXXXX / fort.qXXXX written by the reference's Fortran writers in
    directory.  patches: dicts with q [meqn, my+2mbc, mx+2mbc], mbc, xlower,
    ylower, dx, dy, num, level, block (the writer zeroes tiny values of its
    q argument in place, so it gets a copy)."""
    L = _lib("libref_ascii.so")
    m1 = ctypes.create_string_buffer(f"fort.q{iframe:04d}".encode(), 11)
    m2 = ctypes.create_string_buffer(f"fort.t{iframe:04d}".encode(), 11)
    cwd = os.getcwd()
    os.chdir(directory)
    try:
        L.fclaw2d_clawpatch46_fort_header_ascii_(m1, m2, ctypes.byref(_cd(time)), ctypes.byref(_ci(meqn)),
                                                 ctypes.byref(_ci(maux)), ctypes.byref(_ci(len(patches))),
                                                 ctypes.c_size_t(10), ctypes.c_size_t(10))
        for p in patches:
            q = np.asfortranarray(np.transpose(np.array(p["q"], dtype=np.float64), (2, 1, 0)))
            nm, ny, nx = p["q"].shape
            mbc = p["mbc"]
            L.fclaw2d_clawpatch46_fort_output_ascii_(
                m1, ctypes.byref(_ci(nx - 2 * mbc)), ctypes.byref(_ci(ny - 2 * mbc)), ctypes.byref(_ci(nm)),
                ctypes.byref(_ci(mbc)), ctypes.byref(_cd(p["xlower"])), ctypes.byref(_cd(p["ylower"])),
                ctypes.byref(_cd(p["dx"])), ctypes.byref(_cd(p["dy"])), q.ctypes.data_as(ctypes.c_void_p),
                ctypes.byref(_ci(p["num"])), ctypes.byref(_ci(p["level"])), ctypes.byref(_ci(p["block"])),
                ctypes.byref(_ci(0)), ctypes.c_size_t(10))
    finally:
        os.chdir(cwd)



def src2_axisym(q: np.ndarray, aux: np.ndarray, *, mx, my, mbc, xlower, ylower, dx, dy, t, dt, gamma):
    """The shock-bubble application's clawpack46_src2
    (applications/clawpack/euler/2d/shockbubble/user_4.6/src2.f:1), on a copy."""
    L = _lib("libref_src2.so")
    arr = (_cd * 2).in_dll(L, "cparam_")
    arr[0] = gamma
    arr[1] = gamma - 1.0
    meqn, maux = q.shape[0], aux.shape[0]
    qf = np.ascontiguousarray(q, dtype=np.float64).copy()
    af = np.ascontiguousarray(aux, dtype=np.float64)
    L.clawpack46_src2_(
        ctypes.byref(_ci(mx)), ctypes.byref(_ci(my)), ctypes.byref(_ci(meqn)), ctypes.byref(_ci(mbc)),
        ctypes.byref(_ci(mx)), ctypes.byref(_ci(my)), ctypes.byref(_cd(xlower)), ctypes.byref(_cd(ylower)),
        ctypes.byref(_cd(dx)), ctypes.byref(_cd(dy)), _p(qf), ctypes.byref(_ci(maux)), _p(af),
        ctypes.byref(_cd(t)), ctypes.byref(_cd(dt)))
    return qf


def setaux_axisym(*, mx, my, mbc, xlower, ylower, dx, dy):
    """The shock-bubble application's clawpack46_setaux (user_4.6/setaux.f:1): rad."""
    L = _lib("libref_src2.so")
    aux = np.zeros((1, my + 2 * mbc, mx + 2 * mbc))
    L.clawpack46_setaux_(
        ctypes.byref(_ci(mx)), ctypes.byref(_ci(my)), ctypes.byref(_ci(mbc)), ctypes.byref(_ci(mx)),
        ctypes.byref(_ci(my)), ctypes.byref(_cd(xlower)), ctypes.byref(_cd(ylower)), ctypes.byref(_cd(dx)),
        ctypes.byref(_cd(dy)), ctypes.byref(_ci(1)), _p(aux))
    return aux


RP5 = {  # physics -> (library, rpn2 symbol, rpt2 symbol) of the Clawpack 5 step
    "adv": ("libref5_adv.so", "clawpack5_rpn2adv_", "clawpack5_rpt2adv_"),
    "swe": ("libref5_swe.so", "clawpack5_rpn2_", "clawpack5_rpt2_"),
    "euler": ("libref5_euler.so", "clawpack5_rpn2_euler5_", "clawpack5_rpt2_euler5_"),
}


def available5() -> bool:
    return all(os.path.exists(os.path.join(REF_DIR, RP5[k][0])) for k in RP5)


def set_params5(physics: str, *, grav=1.0, gamma=1.4, tperiod=4.0):
    L = _lib(RP5[physics][0])
    if physics == "swe":
        _cd.in_dll(L, "cparam_").value = grav
    elif physics == "euler":
        _cd.in_dll(L, "cparam_").value = gamma
    else:
        pi = 4.0 * math.atan(1.0)
        arr = (_cd * 2).in_dll(L, "compi_")
        arr[0] = pi
        arr[1] = 2.0 * pi
        _cd.in_dll(L, "comvt_").value = tperiod


def step2_claw5(physics: str, q5: np.ndarray, aux5: np.ndarray | None, *, mx, my, mbc, dx, dy, dt,
                order=(2, 2), mthlim=None, xlower=0.0, ylower=0.0, t=0.0):
    """Run clawpack5_step2_wrap (src/solvers/fc2d_clawpack5/fortran_source/
    clawpack5_step2_wrap.f:1) on a copy of q5, in the Clawpack 5 layout
    q(meqn, i, j): numpy shape (NY, NX, meqn).  The module parameters are set
    through clawpack5_set_amr_module as fc2d_clawpack5.cpp does.  Returns
    (q_after, cfl)."""
    libname, rpn2_sym, rpt2_sym = RP5[physics]
    L = _lib(libname)
    meqn = q5.shape[2]
    mwaves = meqn
    maux = 0 if aux5 is None else aux5.shape[2]
    qf = np.ascontiguousarray(q5, dtype=np.float64).copy()
    auxf = (np.zeros(1) if aux5 is None else np.ascontiguousarray(aux5, dtype=np.float64).copy())
    maxm = max(mx, my)
    mwork = (maxm + 2 * mbc) * (12 * meqn + (meqn + 1) * mwaves + 3 * maux + 2)
    work = np.zeros(mwork)
    size = meqn * (mx + 2 * mbc) * (my + 2 * mbc)
    fp, fm, gp, gm = (np.zeros(size) for _ in range(4))
    method = (_ci * 7)(0, order[0], order[1], 0, 0, 0, 0)
    lim = list(mthlim) if mthlim is not None else [4] * mwaves
    mthlim_c = (_ci * mwaves)(*lim)
    L.clawpack5_set_amr_module_(ctypes.byref(_ci(mwaves)), ctypes.byref(_ci(0)), mthlim_c, method,
                                ctypes.byref(_ci(0)))
    bcc = (_ci * 4)(0, 0, 0, 0)
    cfl = _cd(0.0)
    ierr = _ci(0)
    rpn2 = ctypes.cast(getattr(L, rpn2_sym), ctypes.c_void_p)
    rpt2 = ctypes.cast(getattr(L, rpt2_sym), ctypes.c_void_p)
    flux2 = ctypes.cast(L.clawpack5_flux2_, ctypes.c_void_p)
    L.clawpack5_step2_wrap_(
        ctypes.byref(_ci(maxm)), ctypes.byref(_ci(meqn)), ctypes.byref(_ci(maux)),
        ctypes.byref(_ci(mbc)), method, mthlim_c, ctypes.byref(_ci(0)),
        ctypes.byref(_ci(mwaves)), ctypes.byref(_ci(mx)), ctypes.byref(_ci(my)),
        _p(qf), _p(auxf), ctypes.byref(_cd(dx)), ctypes.byref(_cd(dy)),
        ctypes.byref(_cd(dt)), ctypes.byref(cfl), _p(work), ctypes.byref(_ci(mwork)),
        ctypes.byref(_cd(xlower)), ctypes.byref(_cd(ylower)), ctypes.byref(_ci(0)),
        ctypes.byref(_cd(t)), _p(fp), _p(fm), _p(gp), _p(gm), rpn2, rpt2, flux2, bcc, ctypes.byref(ierr))
    if ierr.value != 0:
        raise RuntimeError(f"clawpack5_step2_wrap ierror={ierr.value}")
    return qf, cfl.value


def swirl_aux5(*, mx, my, mbc, xlower, ylower, dx, dy, t, dt, tperiod):
    """The swirl application's Clawpack 5 b4step2 (user_5.0/b4step2.f90:1):
    aux in the (NY, NX, maux) layout."""
    L = _lib("libref5_adv.so")
    set_params5("adv", tperiod=tperiod)
    aux = np.zeros((my + 2 * mbc, mx + 2 * mbc, 2))
    q = np.zeros((my + 2 * mbc, mx + 2 * mbc, 1))
    L.clawpack5_b4step2_(
        ctypes.byref(_ci(mbc)), ctypes.byref(_ci(mx)), ctypes.byref(_ci(my)), ctypes.byref(_ci(1)), _p(q),
        ctypes.byref(_cd(xlower)), ctypes.byref(_cd(ylower)), ctypes.byref(_cd(dx)), ctypes.byref(_cd(dy)),
        ctypes.byref(_cd(t)), ctypes.byref(_cd(dt)), ctypes.byref(_ci(2)), _p(aux))
    return aux
