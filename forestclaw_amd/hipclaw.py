"""ctypes binding of libhipclaw.so (include/hipclaw.h).

The product path: every call goes to the HIP library.  If the library is not
built, or no GPU is present, calls raise; there is no CPU fallback.
Device buffers are passed as raw device addresses (torch tensors on cuda are
accepted and converted with .data_ptr()).
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libhipclaw.so")

SOLVER_ADV, SOLVER_SWE_ROE, SOLVER_EULER5_ROE, SOLVER_EULER_HLLE = 0, 1, 2, 3
SOLVER_TRANSPORT_FW = 4
FP64, FP32 = 0, 1
FEATURE_MERGE_LEVELS, FEATURE_GHOST_PUSH, FEATURE_GRAPH, FEATURE_INIT_GHOSTCELL = 1, 2, 3, 4
FEATURE_FUSE_AUX = 5
NOALLOC_MAX_TILES = 262144  # HIPCLAW_NOALLOC_MAX_TILES
TAG_VALUE, TAG_DIFFERENCE, TAG_MINMAX, TAG_GRADIENT = 0, 1, 2, 3
SRC2_NONE, SRC2_AXISYMMETRIC = 0, 1
CLAW5 = 0x100   # solver flag: the Clawpack 5 solver and host layout q(meqn, i, j)
TAG_CRITERIA = {"value": TAG_VALUE, "difference": TAG_DIFFERENCE, "minmax": TAG_MINMAX,
                "gradient": TAG_GRADIENT}
ERRORS = {1: "invalid argument", 2: "HIP runtime error", 3: "unsupported",
          4: "out of device memory", 5: "call out of sequence"}


class Physics(ctypes.Structure):
    _fields_ = [
        ("solver", ctypes.c_int),
        ("mx", ctypes.c_int), ("my", ctypes.c_int), ("mbc", ctypes.c_int),
        ("meqn", ctypes.c_int), ("maux", ctypes.c_int), ("mwaves", ctypes.c_int),
        ("order", ctypes.c_int * 2),
        ("mthlim", ctypes.c_int * 5),
        ("mthbc", ctypes.c_int * 4),
        ("efix", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("grav", ctypes.c_double),
        ("gamma", ctypes.c_double),
    ]


class Launch(ctypes.Structure):
    _fields_ = [("cols_per_lane", ctypes.c_int), ("waves_per_group", ctypes.c_int),
                ("row_bands", ctypes.c_int)]


class Domain(ctypes.Structure):
    _fields_ = [("mi", ctypes.c_int), ("mj", ctypes.c_int),
                ("ax", ctypes.c_double), ("bx", ctypes.c_double),
                ("ay", ctypes.c_double), ("by", ctypes.c_double),
                ("minlevel", ctypes.c_int), ("maxlevel", ctypes.c_int),
                ("subcycle", ctypes.c_int)]


class HipclawError(RuntimeError):
    pass


_lib = None

_VP = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_SIGS = {
    "hipclaw_abi_version": ([], ctypes.c_int),
    "hipclaw_last_error": ([], ctypes.c_char_p),
    "hipclaw_physics_default": ([ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.POINTER(Physics)], ctypes.c_int),
    "hipclaw_step2_batch": ([ctypes.POINTER(Physics), ctypes.c_int, _VP, _VP, _VP, _VP, _VP,
                             ctypes.c_int, ctypes.c_double, _VP, ctypes.POINTER(Launch), _VP],
                            ctypes.c_int),
    "hipclaw_step2_batch_sync": ([ctypes.POINTER(Physics), ctypes.c_int, _VP, _VP, _VP, _VP,
                                  _VP, ctypes.c_int, ctypes.c_double, _VP, _VP,
                                  ctypes.POINTER(Launch), _VP], ctypes.c_int),
    "hipclaw_max_reduce": ([_VP, ctypes.c_int, _VP, _VP], ctypes.c_int),
    "hipclaw_step2_batch_noalloc": ([ctypes.POINTER(Physics), ctypes.c_int, _VP, _VP, _VP, _VP, _VP,
                                     ctypes.c_int, ctypes.c_double, _VP, ctypes.POINTER(Launch), _VP],
                                    ctypes.c_int),
    "hipclaw_launch_shape": ([ctypes.POINTER(Physics), ctypes.c_int, ctypes.c_void_p,
                              ctypes.POINTER(Launch)], ctypes.c_int),
    "hipclaw_cnn_create": ([_VP, ctypes.POINTER(_VP)], ctypes.c_int),
    "hipclaw_cnn_destroy": ([_VP], ctypes.c_int),
    "hipclaw_cnn_num_params": ([_VP, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "hipclaw_cnn_init_params": ([_VP, ctypes.c_ulonglong], ctypes.c_int),
    "hipclaw_cnn_set_params": ([_VP, _VP], ctypes.c_int),
    "hipclaw_cnn_get_params": ([_VP, _VP], ctypes.c_int),
    "hipclaw_cnn_forward": ([_VP, _VP, _VP, ctypes.c_int, _VP], ctypes.c_int),
    "hipclaw_cnn_tag_patches": ([_VP, ctypes.c_int, _VP, _DP, ctypes.c_int, ctypes.c_double, _VP],
                                ctypes.c_int),
    "hipclaw_cnn_gradients": ([_VP, _VP, _VP, _VP, ctypes.c_int, _VP, _VP], ctypes.c_int),
    "hipclaw_cnn_train_step": ([_VP, _VP, _VP, _VP, ctypes.c_int, ctypes.c_float, _VP],
                               ctypes.c_int),
    "hipclaw_make_tiles": ([_VP, _VP, _VP], ctypes.c_int),
    "hipclaw_set_tagging_cnn": ([_VP, _VP, ctypes.c_float], ctypes.c_int),
    "hipclaw_tag_stats": ([_VP, _IP], ctypes.c_int),
    "hipclaw_batch_create": ([ctypes.POINTER(Physics), ctypes.c_int, _VP, ctypes.POINTER(_VP)],
                             ctypes.c_int),
    "hipclaw_batch_destroy": ([_VP], ctypes.c_int),
    "hipclaw_single_step_update": ([_VP, _DP, _DP, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    _DP], ctypes.c_int),
    "hipclaw_level_upload": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_DP),
                              ctypes.POINTER(_DP)], ctypes.c_int),
    "hipclaw_level_download": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_DP)],
                               ctypes.c_int),
    "hipclaw_swirl_aux_batch": ([ctypes.POINTER(Physics), ctypes.c_int, _VP, _VP,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, _VP, _VP],
                                ctypes.c_int),
    "hipclaw_bc2_batch": ([ctypes.POINTER(Physics), ctypes.c_int, _VP, _VP, ctypes.c_int, _VP,
                           _VP], ctypes.c_int),
    "hipclaw_src2_batch": ([ctypes.POINTER(Physics), ctypes.c_int, ctypes.c_int, _VP, _VP, _VP, _VP,
                            ctypes.c_double, ctypes.c_double, _VP], ctypes.c_int),
    "hipclaw_batch_set_src2": ([_VP, ctypes.c_int], ctypes.c_int),
}


_SIGS.update({
    "hipclaw_ctx_create": ([ctypes.POINTER(Physics), ctypes.POINTER(Domain), _VP,
                            ctypes.POINTER(_VP)], ctypes.c_int),
    "hipclaw_ctx_destroy": ([_VP], ctypes.c_int),
    "hipclaw_init_uniform": ([_VP, ctypes.c_int], ctypes.c_int),
    "hipclaw_init_leaves": ([_VP, ctypes.c_int, _IP, _IP, _IP], ctypes.c_int),
    "hipclaw_set_launch": ([_VP, ctypes.c_void_p], ctypes.c_int),
    "hipclaw_get_launch": ([_VP, ctypes.POINTER(Launch)], ctypes.c_int),
    "hipclaw_save_step": ([_VP], ctypes.c_int),
    "hipclaw_restore_step": ([_VP], ctypes.c_int),
    "hipclaw_set_time_sync": ([_VP, ctypes.c_int], ctypes.c_int),
    "hipclaw_set_src2": ([_VP, ctypes.c_int], ctypes.c_int),
    "hipclaw_set_feature": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "hipclaw_graph_stats": ([_VP, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)],
                            ctypes.c_int),
    "hipclaw_time_sync": ([_VP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "hipclaw_get_registers": ([_VP, _DP], ctypes.c_int),
    "hipclaw_autotune": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p],
                         ctypes.c_int),
    "hipclaw_num_patches": ([_VP, _IP], ctypes.c_int),
    "hipclaw_patch_info": ([_VP, _IP, _IP, _IP, _DP, _DP, _DP, _DP], ctypes.c_int),
    "hipclaw_set_q": ([_VP, _DP], ctypes.c_int),
    "hipclaw_get_q": ([_VP, _DP], ctypes.c_int),
    "hipclaw_set_swirl": ([_VP, ctypes.c_double], ctypes.c_int),
    "hipclaw_set_aux": ([_VP, _DP], ctypes.c_int),
    "hipclaw_qinit": ([_VP, ctypes.c_int, _DP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "hipclaw_ghost_update": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int],
                             ctypes.c_int),
    "hipclaw_update_level": ([_VP, ctypes.c_int, ctypes.c_double, ctypes.c_double, _DP],
                             ctypes.c_int),
    "hipclaw_timeinterp": ([_VP, ctypes.c_int, ctypes.c_double], ctypes.c_int),
    "hipclaw_advance": ([_VP, ctypes.c_double, ctypes.c_double, _DP], ctypes.c_int),
    "hipclaw_advance_steps": ([_VP, ctypes.c_double, ctypes.c_double, ctypes.c_int, _DP], ctypes.c_int),
    "hipclaw_prepare_steps": ([_VP, ctypes.c_int], ctypes.c_int),
    "hipclaw_set_tagging": ([_VP, ctypes.c_int, ctypes.c_double, ctypes.c_double], ctypes.c_int),
    "hipclaw_regrid": ([_VP, ctypes.c_int, _IP], ctypes.c_int),
    "hipclaw_initialize": ([_VP, ctypes.c_int, _DP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "hipclaw_mass": ([_VP, _DP], ctypes.c_int),
    "hipclaw_set_tag_criterion": ([_VP, ctypes.c_int], ctypes.c_int),
    "hipclaw_set_refinement": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "hipclaw_output_ascii": ([_VP, ctypes.c_int, ctypes.c_double, ctypes.c_char_p], ctypes.c_int),
    "hipclaw_ascii_header": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int], ctypes.c_int),
    "hipclaw_ascii_patch": ([ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             _DP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "hipclaw_compare_frames": ([ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                _DP, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "hipclaw_cell_updates": ([_VP, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "hipclaw_profile": ([_VP, ctypes.c_int], ctypes.c_int),
    "hipclaw_profile_read": ([_VP, _DP, ctypes.POINTER(ctypes.c_longlong),
                              ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
})


def lib() -> ctypes.CDLL:
    """Load libhipclaw.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HipclawError(
                f"{LIB_PATH} is missing: build it with `python -m forestclaw_amd.build`")
        # torch ships its own HIP runtime under the same soname
        # (libamdhip64.so.7).  Loading torch first lets the dynamic linker bind
        # this library to that runtime, so the process has exactly one HIP
        # runtime and device pointers, streams and allocations are shared.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().hipclaw_last_error().decode(errors="replace")
        raise HipclawError(f"hipclaw error {rc} ({ERRORS.get(rc, '?')}): {msg}")


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _stream(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return stream


def physics(solver: int, mx: int, my: int, *, precision: int = FP64, order=(2, 2),
            mthlim=None, mthbc=(1, 1, 1, 1), grav: float = 1.0, gamma: float = 1.4,
            efix: int = 1, maux: int | None = None) -> Physics:
    p = Physics()
    _check(lib().hipclaw_physics_default(solver, mx, my, ctypes.byref(p)))
    if maux is not None:
        p.maux = maux
    p.precision = precision
    p.order[0], p.order[1] = order
    if mthlim is not None:
        for k in range(5):
            p.mthlim[k] = mthlim[k] if k < len(mthlim) else 0
    for k in range(4):
        p.mthbc[k] = mthbc[k]
    p.grav, p.gamma, p.efix = grav, gamma, efix
    return p


def step2_batch(phys: Physics, npatch: int, q_in, q_out, aux, dxdy, dt: float, cfl=None,
                patch_ids=None, launch: tuple[int, int] | None = None, stream=None,
                ntiles: int | None = None) -> None:
    """npatch tiles (or the npatch entries of patch_ids) of buffers holding
    ntiles tiles (default npatch)."""
    if ntiles is None:
        ntiles = npatch
    ln = None
    if launch is not None:
        ln = Launch(*launch)
    _check(lib().hipclaw_step2_batch(
        ctypes.byref(phys), npatch, _ptr(q_in), _ptr(q_out), _ptr(aux), _ptr(dxdy),
        _ptr(patch_ids), int(ntiles), float(dt), _ptr(cfl), ctypes.byref(ln) if ln is not None else None,
        _stream(stream)))


def step2_batch_noalloc(phys: Physics, npatch: int, q_in, q_out, aux, dxdy, dt: float, maxcfl,
                        patch_ids=None, launch=None, stream=None, ntiles: int | None = None) -> None:
    """step2_batch that allocates nothing and writes the batch's maximum
    Courant number to maxcfl (one device double)."""
    if ntiles is None:
        ntiles = npatch
    ln = Launch(*launch) if launch is not None else None
    _check(lib().hipclaw_step2_batch_noalloc(
        ctypes.byref(phys), npatch, _ptr(q_in), _ptr(q_out), _ptr(aux), _ptr(dxdy),
        _ptr(patch_ids), int(ntiles), float(dt), _ptr(maxcfl),
        ctypes.byref(ln) if ln is not None else None, _stream(stream)))


def step2_batch_sync(phys: Physics, npatch: int, q_in, q_out, aux, dxdy, dt: float, cfl,
                     registers, launch: tuple[int, int] | None = None, stream=None) -> None:
    """step2_batch that also accumulates the FP64 time-sync registers
    (8*meqn*(mx+my) doubles per tile)."""
    ln = Launch(*launch) if launch is not None else None
    _check(lib().hipclaw_step2_batch_sync(
        ctypes.byref(phys), npatch, _ptr(q_in), _ptr(q_out), _ptr(aux), _ptr(dxdy), None,
        int(npatch), float(dt), _ptr(cfl), _ptr(registers),
        ctypes.byref(ln) if ln is not None else None, _stream(stream)))


class Cnn:
    """The learned refinement indicator (hipclaw_cnn_*): parameters live on
    the device; tiles are device bf16 [n, 36, 36, 8] (torch.bfloat16)."""

    def __init__(self, stream=None, seed: int | None = None):
        h = _VP()
        _check(lib().hipclaw_cnn_create(_stream(stream), ctypes.byref(h)))
        self.h = h
        n = ctypes.c_longlong()
        _check(lib().hipclaw_cnn_num_params(self.h, ctypes.byref(n)))
        self.num_params = n.value
        if seed is not None:
            _check(lib().hipclaw_cnn_init_params(self.h, seed))

    def get_params(self):
        import numpy as np
        out = np.zeros(self.num_params, dtype=np.float32)
        _check(lib().hipclaw_cnn_get_params(self.h, out.ctypes.data))
        return out

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        _check(lib().hipclaw_cnn_set_params(self.h, p.ctypes.data))

    def forward(self, tiles, logits, idx=None, n=None):
        n = tiles.shape[0] if n is None else n
        _check(lib().hipclaw_cnn_forward(self.h, _ptr(tiles), _ptr(idx), n, _ptr(logits)))

    def gradients(self, tiles, labels, idx=None, n=None):
        import numpy as np
        n = tiles.shape[0] if n is None else n
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _check(lib().hipclaw_cnn_gradients(self.h, _ptr(tiles), _ptr(labels), _ptr(idx), n,
                                           g.ctypes.data, loss.ctypes.data))
        return float(loss[0]), g

    def tag_patches(self, patches, dx, threshold):
        """Refinement tags of host patches (a list of FP64 (meqn, 36, 36)
        arrays, 32x32 patches) from the network; returns an int array."""
        import numpy as np
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in patches]
        n = len(arrs)
        ptrs = (ctypes.c_void_p * max(n, 1))(*[a.ctypes.data for a in arrs])
        dxa = np.ascontiguousarray(dx, dtype=np.float64)
        tags = np.zeros(max(n, 1), dtype=np.int32)
        meqn = arrs[0].shape[0] if n else 1
        _check(lib().hipclaw_cnn_tag_patches(self.h, n, ctypes.addressof(ptrs), dxa.ctypes.data_as(_DP),
                                             meqn, float(threshold), tags.ctypes.data))
        return tags[:n]

    def train_step(self, tiles, labels, idx, n, lr, loss_dev=None):
        _check(lib().hipclaw_cnn_train_step(self.h, _ptr(tiles), _ptr(labels), _ptr(idx), n,
                                            float(lr), _ptr(loss_dev)))

    def close(self):
        if self.h:
            lib().hipclaw_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Deferred single-step update for host-resident patches
    (hipclaw_batch_create / hipclaw_single_step_update)."""

    def __init__(self, phys: Physics, capacity: int, stream=None):
        h = _VP()
        _check(lib().hipclaw_batch_create(ctypes.byref(phys), capacity, _stream(stream),
                                          ctypes.byref(h)))
        self.h = h

    def single_step_update(self, q, aux, dx, dy, t, dt, it, total) -> float:
        """q, aux: C-contiguous float64 numpy tiles, updated in place on a flush."""
        c = ctypes.c_double()
        _check(lib().hipclaw_single_step_update(
            self.h, q.ctypes.data_as(_DP), None if aux is None else aux.ctypes.data_as(_DP),
            dx, dy, t, dt, it, total, ctypes.byref(c)))
        return c.value

    def set_src2(self, kind: int) -> None:
        _check(lib().hipclaw_batch_set_src2(self.h, kind))

    def close(self):
        if self.h:
            lib().hipclaw_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def launch_shape(phys: Physics, npatch: int, launch=None):
    """(cols_per_lane, waves_per_group, row_bands) a patch-step launch of
    npatch patches really uses for the shape `launch` (None: library
    default).  Host-only."""
    ln = Launch(*launch) if launch is not None else None
    out = Launch()
    _check(lib().hipclaw_launch_shape(ctypes.byref(phys), npatch,
                                      ctypes.byref(ln) if ln is not None else None, ctypes.byref(out)))
    return out.cols_per_lane, out.waves_per_group, out.row_bands


def max_reduce(x, n: int, out, stream=None) -> None:
    _check(lib().hipclaw_max_reduce(_ptr(x), n, _ptr(out), _stream(stream)))


def swirl_aux_batch(phys: Physics, npatch: int, xylower, dxdy, t: float, dt: float,
                    tperiod: float, aux, stream=None) -> None:
    _check(lib().hipclaw_swirl_aux_batch(ctypes.byref(phys), npatch, _ptr(xylower), _ptr(dxdy),
                                         t, dt, tperiod, _ptr(aux), _stream(stream)))


def src2_batch(phys: Physics, kind: int, npatch: int, q, aux=None, xylower=None, dxdy=None, t: float = 0.0,
               dt: float = 0.0, stream=None) -> None:
    _check(lib().hipclaw_src2_batch(ctypes.byref(phys), kind, npatch, _ptr(q), _ptr(aux), _ptr(xylower),
                                    _ptr(dxdy), t, dt, _stream(stream)))


def bc2_batch(phys: Physics, npatch: int, q, mthbc, patch_ids=None, stream=None,
              ntiles: int | None = None) -> None:
    if ntiles is None:
        ntiles = npatch
    _check(lib().hipclaw_bc2_batch(ctypes.byref(phys), npatch, _ptr(q), _ptr(patch_ids),
                                   int(ntiles), _ptr(mthbc), _stream(stream)))


def ascii_header(directory: str, iframe: int, time: float, meqn: int, maux: int, ngrids: int) -> None:
    """fort.tXXXX of a frame (and an empty fort.qXXXX) in directory."""
    _check(lib().hipclaw_ascii_header(directory.encode(), iframe, time, meqn, maux, ngrids))


def ascii_patch(directory: str, iframe: int, q, xlower: float, ylower: float, dx: float, dy: float,
                patch_num: int, level: int, blockno: int = 0, mpirank: int = 0, mbc: int = 2) -> None:
    """Append one patch (q: float64 tile [meqn, my+2mbc, mx+2mbc]) to fort.qXXXX."""
    import numpy as np
    q = np.ascontiguousarray(q, dtype=np.float64)
    meqn, ny, nx = q.shape
    _check(lib().hipclaw_ascii_patch(directory.encode(), iframe, nx - 2 * mbc, ny - 2 * mbc, meqn, mbc, xlower,
                                     ylower, dx, dy, q.ctypes.data_as(_DP), patch_num, level, blockno,
                                     mpirank))

