"""Frame diff of two runs (scripts/compare_files/compare_files.f90): the
largest |q1 - q2| per field over all grids of one output frame, the grid it
occurs on, and the difference frame, through hipclaw_compare_frames.

    python -m forestclaw_amd.compare_files D1 D2 IFRAME

compares run_0000D1/fort.qIFRAME with run_0000D2/fort.qIFRAME and writes
compare_0000D1_0000D2/ (created if missing), as the reference tool does with
its three command-line arguments.
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

from . import hipclaw as H


def compare_frames(dir1: str, dir2: str, dir3: str, iframe: int):
    """Returns (global_max, grid_max): per field the largest |q1 - q2| and the
    grid number where it first occurs; raises HipclawError where the
    reference tool stops (different meqn, grid counts or geometry, times more
    than 1e-5 apart)."""
    gmax = np.zeros(10)
    gwhere = np.zeros(10, dtype=np.int32)
    H._check(H.lib().hipclaw_compare_frames(dir1.encode(), dir2.encode(), dir3.encode(), int(iframe), 10,
                                            gmax.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            gwhere.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    with open(os.path.join(dir1, f"fort.t{iframe:04d}")) as f:
        f.readline()
        meqn = int(f.readline().split()[0])
    return gmax[:meqn], gwhere[:meqn]


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 3:
        print(__doc__.strip().splitlines()[-6])
        return 2
    d1, d2, iframe = (int(a) for a in argv)
    dir1, dir2 = f"run_{d1:05d}", f"run_{d2:05d}"
    dir3 = f"compare_{d1:05d}_{d2:05d}"
    os.makedirs(dir3, exist_ok=True)
    gmax, gwhere = compare_frames(dir1, dir2, dir3, iframe)
    for m, (v, g) in enumerate(zip(gmax, gwhere)):
        print(f"{v:30.16E}     occurs in field{m + 1:5d} on grid {g:5d}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
