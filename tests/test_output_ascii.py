"""CPU tests of the ASCII frame writer (hipclaw_ascii_header / hipclaw_ascii_patch,
forestclaw_amd/csrc/hc_output.cpp) against the reference's own Fortran writers
(src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_output_ascii.f).

* golden: tests/golden/ascii/fort.t0003 / fort.q0003 were written by the
  reference writers (compiled unchanged into oracle/_ref, see
  tests/golden/make_golden.py) from the patches in patches.npz; the library
  must reproduce both files byte for byte.  The values cover zero, values
  below 1e-99 (written as zero), a rounding carry, negative values and
  three-digit exponents.
* live: when oracle/_ref/libref_ascii.so is built, random patches, sizes and
  times written by both writers compare equal.
"""
import os

import numpy as np
import pytest

from forestclaw_amd import hipclaw as H

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "ascii")
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_ascii.so")


def _patches_from_npz():
    z = np.load(os.path.join(GOLD, "patches.npz"))
    n = len([k for k in z.files if k.startswith("q")])
    out = []
    for i in range(n):
        out.append({k: z[f"{k}{i}"] for k in ("q", "mbc", "xlower", "ylower", "dx", "dy", "num", "level",
                                             "block")})
    return out


def _write(directory, iframe, time, meqn, maux, patches):
    H.ascii_header(str(directory), iframe, time, meqn, maux, len(patches))
    for p in patches:
        H.ascii_patch(str(directory), iframe, p["q"], float(p["xlower"]), float(p["ylower"]), float(p["dx"]),
                      float(p["dy"]), int(p["num"]), int(p["level"]), int(p["block"]), 0, int(p["mbc"]))


def _read(path):
    with open(path, "rb") as f:
        return f.read()


def test_frame_matches_reference_golden(tmp_path):
    patches = _patches_from_npz()
    _write(tmp_path, 3, 0.1 + 1e-17, 3, 0, patches)
    for name in ("fort.t0003", "fort.q0003"):
        assert _read(tmp_path / name) == _read(os.path.join(GOLD, name)), name


def test_tiny_values_are_written_as_zero_and_input_is_untouched(tmp_path):
    q = np.zeros((1, 5, 5))
    q[0, 2, 2] = 1e-100
    q0 = q.copy()
    H.ascii_header(str(tmp_path), 0, 0.0, 1, 0, 1)
    H.ascii_patch(str(tmp_path), 0, q, 0.0, 0.0, 1.0, 1.0, 0, 0)
    text = (tmp_path / "fort.q0000").read_text()
    assert "0.0000000000000000E+00" in text and "E-100" not in text and "-100" not in text
    assert np.array_equal(q, q0)


def test_rejects_bad_frames(tmp_path):
    with pytest.raises(H.HipclawError):
        H.ascii_header(str(tmp_path), 10000, 0.0, 1, 0, 1)
    with pytest.raises(H.HipclawError):
        H.ascii_patch(str(tmp_path), 0, np.zeros((21, 5, 5)), 0.0, 0.0, 1.0, 1.0, 0, 0)


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref/libref_ascii.so not built")
def test_frame_matches_live_reference(tmp_path):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import refdrive as R
    rng = np.random.default_rng(11)
    for case in range(4):
        meqn = int(rng.integers(1, 6))
        patches = []
        for k in range(int(rng.integers(1, 4))):
            mx, my = (int(v) for v in rng.integers(1, 9, 2))
            q = rng.standard_normal((meqn, my + 4, mx + 4)) * 10.0 ** rng.uniform(-30, 30, (meqn, my + 4, mx + 4))
            patches.append(dict(q=q, mbc=2, xlower=float(rng.uniform(-3, 3)), ylower=float(rng.uniform(-3, 3)),
                                dx=float(rng.uniform(1e-4, 1)), dy=float(rng.uniform(1e-4, 1)), num=k,
                                level=int(rng.integers(0, 10)), block=int(rng.integers(0, 4))))
        t = float(rng.uniform(0, 100))
        ours, ref = tmp_path / f"o{case}", tmp_path / f"r{case}"
        ours.mkdir()
        ref.mkdir()
        _write(ours, case, t, meqn, 0, patches)
        R.ascii_frame(str(ref), case, t, meqn, 0, patches)
        for name in (f"fort.t{case:04d}", f"fort.q{case:04d}"):
            assert _read(ours / name) == _read(ref / name), (case, name)
