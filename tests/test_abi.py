"""CPU tests of the C-ABI library: it loads, exports every entry point that
include/hipclaw.h declares, and rejects bad configurations with an error code
before touching the device (no compute calls here)."""
import ctypes
import os
import re

import pytest

from forestclaw_amd import hipclaw as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(hipclaw_\w+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = H.lib()
    names = declared_symbols()
    assert len(names) >= 6
    for n in names:
        assert hasattr(L, n), n


def test_python_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) <= set(H.exported_symbols())


def test_abi_version():
    assert H.lib().hipclaw_abi_version() == 1


def test_physics_defaults_follow_reference_options():
    p = H.physics(H.SOLVER_SWE_ROE, 32, 32)
    assert (p.meqn, p.mwaves, p.maux, p.mbc) == (3, 3, 0, 2)
    assert list(p.order) == [2, 2] and list(p.mthlim)[:3] == [4, 4, 4]
    assert p.efix == 1
    p = H.physics(H.SOLVER_ADV, 16, 16)
    assert (p.meqn, p.maux) == (1, 2)


@pytest.mark.parametrize("field,value", [("mbc", 3), ("precision", 7), ("meqn", 2),
                                         ("solver", 9), ("grav", -1.0)])
def test_invalid_physics_is_rejected_before_any_device_call(field, value):
    p = H.physics(H.SOLVER_SWE_ROE, 32, 32)
    setattr(p, field, value)
    with pytest.raises(H.HipclawError):
        H.step2_batch(p, 4, 1, 1, None, 1, 0.1, stream=0)


def test_bad_limiter_and_launch_rejected():
    p = H.physics(H.SOLVER_SWE_ROE, 32, 32, mthlim=(4, 4, 9))
    with pytest.raises(H.HipclawError):
        H.step2_batch(p, 4, 1, 1, None, 1, 0.1, stream=0)
    p = H.physics(H.SOLVER_SWE_ROE, 32, 32)
    with pytest.raises(H.HipclawError):
        H.step2_batch(p, 4, 1, 1, None, 1, 0.1, launch=(3, 4), stream=0)


def test_launch_shape_defaults_and_tuned_shapes():
    """Host-side launch shapes: the library default bands small launches and
    gives FP32 launches of 2,048 wide patches or more two columns per lane; a
    tuned shape with row_bands -1 applies only to launches of 2,048 patches
    or more; two columns per lane never band and keep their waves per group."""
    p64 = H.physics(H.SOLVER_SWE_ROE, 32, 32)
    p32 = H.physics(H.SOLVER_SWE_ROE, 32, 32, precision=H.FP32)
    assert H.launch_shape(p64, 200) == (1, 8, 8)
    assert H.launch_shape(p64, 400) == (1, 4, 4)
    assert H.launch_shape(p64, 1500) == (1, 2, 2)
    assert H.launch_shape(p64, 5000) == (1, 4, 1)
    assert H.launch_shape(p32, 5000) == (2, 4, 1)
    for n in (200, 1500):
        assert H.launch_shape(p64, n, (2, 2, -1)) == H.launch_shape(p64, n)
    assert H.launch_shape(p64, 5000, (2, 2, -1)) == (2, 2, 1)
    assert H.launch_shape(p64, 200, (2, 4, 4)) == (2, 4, 1)
    assert H.launch_shape(p64, 200, (1, 1, 0)) == (1, 1, 1)
    assert H.launch_shape(p64, 200, (1, 1, 3)) == (1, 3, 3)
    # bands of at least two rows: an 8-row patch takes at most 4
    p8 = H.physics(H.SOLVER_SWE_ROE, 8, 8)
    assert H.launch_shape(p8, 10) == (1, 4, 4)
    with pytest.raises(H.HipclawError):
        H.launch_shape(p64, 10, (1, 1, -2))


def test_patch_list_over_no_tiles_rejected():
    """A launch reads tile 0 from lanes that own no patch, so a patch list
    over buffers of no tiles is rejected before any device call."""
    p = H.physics(H.SOLVER_SWE_ROE, 32, 32)
    ids = (ctypes.c_int * 2)(0, 1)
    with pytest.raises(H.HipclawError, match="ntiles"):
        H.step2_batch(p, 2, 1, 1, None, 1, 0.1, patch_ids=ctypes.addressof(ids), ntiles=0, stream=0)


def test_fwave_transport_physics_and_rejections():
    """f-wave transport: meqn 1, maux 7 by default; other aux counts, the
    Clawpack 5 flag and the AMR context are refused before any device call."""
    p = H.physics(H.SOLVER_TRANSPORT_FW, 16, 16)
    assert (p.meqn, p.mwaves, p.maux) == (1, 1, 7)
    p.maux = 2
    with pytest.raises(H.HipclawError, match="maux=7"):
        H.step2_batch(p, 4, 1, 1, 1, 1, 0.1, stream=0)
    with pytest.raises(H.HipclawError, match="Clawpack 5"):
        H.step2_batch(H.physics(H.SOLVER_TRANSPORT_FW | H.CLAW5, 16, 16), 4, 1, 1, 1, 1, 0.1,
                      stream=0)
    dom = H.Domain(1, 1, 0.0, 1.0, 0.0, 1.0, 2, 3, 0)
    h = ctypes.c_void_p()
    rc = H.lib().hipclaw_ctx_create(ctypes.byref(H.physics(H.SOLVER_TRANSPORT_FW, 16, 16)),
                                    ctypes.byref(dom), None, ctypes.byref(h))
    assert rc != 0 and b"hipclaw_step2_batch" in H.lib().hipclaw_last_error()
