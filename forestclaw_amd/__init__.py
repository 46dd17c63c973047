"""forestclaw_amd -- MI355X-native (gfx950) Clawpack-4.6 wave-propagation patch
solver for ForestClaw: a C-ABI HIP library (libhipclaw.so, include/hipclaw.h)
plus the host-side driver that mirrors the fc2d_clawpack46 solver interface.
"""
from . import hipclaw  # noqa: F401

__all__ = ["hipclaw"]
