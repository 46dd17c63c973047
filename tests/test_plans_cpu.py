"""CPU bounds / coverage check of the ghost-fill plans and regrid transfers
(tests/cpp/check_plans.cpp replays the GPU kernels' index arithmetic over
plans built by the library's own host code for random balanced forests)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_plans_stay_in_bounds_and_cover_every_ghost_cell(tmp_path):
    exe = tmp_path / "check_plans"
    csrc = os.path.join(ROOT, "forestclaw_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", csrc,
                    os.path.join(ROOT, "tests", "cpp", "check_plans.cpp"),
                    os.path.join(csrc, "hc_forest.cpp"), os.path.join(csrc, "hc_plan.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "0 failures" in r.stdout
