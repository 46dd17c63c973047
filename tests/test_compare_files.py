"""Frame diff (hipclaw_compare_frames, forestclaw_amd/compare_files.py)
against the reference's own tool, scripts/compare_files/compare_files.f90,
compiled unchanged into oracle/_ref/compare_files: on frames written by the
Clawpack 4.6 ASCII writers the difference frame fort.qXXXX and
compare_summary.XXXX are byte-identical to the tool's, and fort.tXXXX holds
what the tool's format 1000 writes.  (The tool opens its fort.t on an unset
unit number, matunit2 = 52 being commented out, so the runtime writes that
record to an implicit fort.<unit> file in the working directory; the test
compares against that file.)  Host code only: no device is used."""
import ctypes
import glob
import os
import subprocess

import numpy as np
import pytest

from forestclaw_amd import compare_files as CF
from forestclaw_amd import hipclaw as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_TOOL = os.path.join(ROOT, "oracle", "_ref", "compare_files")
_DP = ctypes.POINTER(ctypes.c_double)


def write_frame(d, iframe, t, grids, meqn, mx=6, my=5):
    os.makedirs(d, exist_ok=True)
    L = H.lib()
    assert L.hipclaw_ascii_header(d.encode(), iframe, ctypes.c_double(t), meqn, 0, len(grids)) == 0
    for g, (q, xl, yl, lev) in enumerate(grids):
        q = np.ascontiguousarray(q)
        assert L.hipclaw_ascii_patch(d.encode(), iframe, mx, my, meqn, 2, ctypes.c_double(xl), ctypes.c_double(yl),
                                     ctypes.c_double(0.1), ctypes.c_double(0.2), q.ctypes.data_as(_DP), g, lev,
                                     g // 2, 0) == 0


def grids(rng, meqn, n=4, mx=6, my=5, eps=0.0, seed=0):
    out = []
    for g in range(n):
        q = np.random.default_rng(seed + g).normal(size=(meqn, my + 4, mx + 4))
        q[0, 3, 3] = 1e-120  # written as zero
        q += eps * rng.normal(size=q.shape) * (g + 1)
        out.append((q, 0.6 * g, -0.25 * g, 1 + g % 3))
    return out


@pytest.mark.parametrize("meqn", [1, 3, 5])
def test_matches_the_reference_tool(tmp_path, meqn):
    if not os.path.exists(REF_TOOL):
        pytest.skip("oracle/_ref/compare_files not built (the reference is not mounted)")
    rng = np.random.default_rng(meqn)
    iframe = 7
    write_frame(str(tmp_path / "run_00001"), iframe, 0.5, grids(rng, meqn), meqn)
    write_frame(str(tmp_path / "run_00002"), iframe, 0.5, grids(rng, meqn, eps=1e-7), meqn)
    os.makedirs(tmp_path / "compare_00001_00002")
    subprocess.run([REF_TOOL, "1", "2", str(iframe)], cwd=tmp_path, check=True, capture_output=True)
    mine = str(tmp_path / "mine")
    os.makedirs(mine)
    gmax, gwhere = CF.compare_frames(str(tmp_path / "run_00001"), str(tmp_path / "run_00002"), mine, iframe)
    ref = tmp_path / "compare_00001_00002"
    for name in (f"fort.q{iframe:04d}", f"compare_summary.{iframe:04d}"):
        assert open(os.path.join(mine, name), "rb").read() == open(ref / name, "rb").read(), name
    implicit = [f for f in glob.glob(str(tmp_path / "fort.*")) if os.path.basename(f)[5:].isdigit()]
    assert len(implicit) == 1
    assert open(os.path.join(mine, f"fort.t{iframe:04d}"), "rb").read() == open(implicit[0], "rb").read()
    assert gmax.shape == (meqn,) and (gmax > 0).all() and (gwhere >= 0).all()


def test_identical_frames_and_refusals(tmp_path):
    rng = np.random.default_rng(3)
    g = grids(rng, 3)
    write_frame(str(tmp_path / "a"), 2, 1.0, g, 3)
    write_frame(str(tmp_path / "b"), 2, 1.0, g, 3)
    os.makedirs(tmp_path / "c")
    gmax, gwhere = CF.compare_frames(str(tmp_path / "a"), str(tmp_path / "b"), str(tmp_path / "c"), 2)
    assert (gmax == 0).all() and (gwhere == 0).all()
    summary = open(tmp_path / "c" / "compare_summary.0002").read().splitlines()
    assert summary[0].split()[0] == "0.0000000000000000E+00" and len(summary) == 3 + 1 + 4
    # the differences the reference tool stops on
    cases = {
        "meqn": lambda d: write_frame(d, 2, 1.0, grids(rng, 2), 2),
        "ngrids": lambda d: write_frame(d, 2, 1.0, g[:3], 3),
        "time": lambda d: write_frame(d, 2, 1.001, g, 3),
        "xlow": lambda d: write_frame(d, 2, 1.0, [(q, xl + 0.5, yl, lv) for q, xl, yl, lv in g], 3),
    }
    for name, make in cases.items():
        d = str(tmp_path / name)
        make(d)
        with pytest.raises(H.HipclawError):
            CF.compare_frames(str(tmp_path / "a"), d, str(tmp_path / "c"), 2)


def test_maxima_are_those_of_the_written_values(tmp_path):
    """Per field, the largest |q1 - q2| over the interiors as written (16
    significant digits) and the first grid where it occurs."""
    rng = np.random.default_rng(5)
    g1, g2 = grids(rng, 3, n=5), grids(rng, 3, n=5, eps=1e-3, seed=0)
    write_frame(str(tmp_path / "a"), 0, 0.0, g1, 3)
    write_frame(str(tmp_path / "b"), 0, 0.0, g2, 3)
    os.makedirs(tmp_path / "c")
    gmax, gwhere = CF.compare_frames(str(tmp_path / "a"), str(tmp_path / "b"), str(tmp_path / "c"), 0)

    def rounded(q):
        v = np.array([float(f"{x:.15e}") for x in q.ravel()]).reshape(q.shape)
        v[np.abs(v) < 1e-99] = 0.0
        return v
    d = np.stack([np.abs(rounded(a[0][:, 2:-2, 2:-2]) - rounded(b[0][:, 2:-2, 2:-2])) for a, b in zip(g1, g2)])
    want = d.max(axis=(0, 2, 3))
    assert np.allclose(gmax, want, rtol=1e-14, atol=0)
    for m in range(3):
        assert gwhere[m] == int(np.argmax(d[:, m].reshape(5, -1).max(axis=1)))
