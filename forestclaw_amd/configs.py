"""The workloads the solver is measured and checked on, as runnable setups,
and the FP32-vs-FP64 error report of the precision mode.

  C1  swirl advection, unit square, uniform level 3 (64 patches of 32x32),
      edge velocities from the stream function with period 8 (cos(pi t/4)),
      q = 1 for x < 0.5, MC limiter, fixed dt at CFL 0.9, 200 steps
      (applications/clawpack/advection/2d/swirl).
  C2  shallow-water radial dam on [-2.5, 2.5]^2, AMR levels 2-6, 32x32,
      g = 1, h = 2 inside r < 0.5 + 0.05 N(0,1), refine when |grad h| > 0.05,
      subcycled, time sync (applications/clawpack/shallow/2d/radialdam).
  C3  Euler shock-bubble, gamma 1.4, [0, 2] x [0, 0.5] as a 4x1 brick, AMR
      levels 3-7, 16x16 patches, Mach 1.22 shock at x = 0.2 meeting a bubble
      of density 0.138, radius 0.2 at (0.5, 0.25) (applications/clawpack/euler/2d/shockbubble);
      Roe (C3) or HLLE (C3-HLLE) normal solver.
  C5  shallow water on the unit square, level 7 uniform refined to level 9
      around 8 seeded Gaussian humps (the throughput workload of bench.py).

precision_report() runs a configuration in FP64 and in FP32 on the same
mesh (the FP64 run's, no regrid during the comparison) from the same
initial data and returns the relative L1 and L-inf errors of the FP32
interior against FP64 after each of a few checkpoints.
"""
from __future__ import annotations

import math

import numpy as np

from . import forest as F
from . import hipclaw as H


def c1_options(precision=H.FP64):
    return F.Options(minlevel=3, maxlevel=3, mx=32, my=32, solver=H.SOLVER_ADV, precision=precision,
                     mthbc=(1, 1, 1, 1))


def c1_dt():
    # |u| = |sin^2(pi x) sin(2 pi y)| <= 1 at t = 0; dx = 1/256
    dx = 1.0 / (32 << 3)
    return 0.9 * dx


def c2_case(seed=0):
    r = 0.5 + 0.05 * float(np.random.default_rng(seed).standard_normal())
    return (0.0, 0.0, r, 2.0, 1.0)


def c2_options(precision=H.FP64, time_sync=True):
    return F.Options(minlevel=2, maxlevel=6, mx=32, my=32, ax=-2.5, bx=2.5, ay=-2.5, by=2.5,
                     solver=H.SOLVER_SWE_ROE, precision=precision, grav=1.0, refine_threshold=0.05,
                     coarsen_threshold=0.0125, subcycle=True, time_sync=time_sync)


def c2_dt(coarsest=2):
    # fastest wave of the h = 2 : 1 dam about 1.7 (measured Courant number)
    dx = 5.0 / (32 << coarsest)
    return 0.8 * dx / 1.7


def c3_params(gamma=1.4):
    pinf = 1.5698
    rinf = ((gamma - 1) + (gamma + 1) * pinf) / ((gamma + 1) + (gamma - 1) * pinf)
    vinf = (1 / math.sqrt(gamma)) * (pinf - 1) / math.sqrt(
        0.5 * ((gamma + 1) / gamma) * pinf + 0.5 * (gamma - 1) / gamma)
    return (gamma, 0.2, rinf, vinf, pinf, 0.5, 0.25, 0.2, 0.138)


def c3_options(precision=H.FP64, solver=H.SOLVER_EULER5_ROE):
    return F.Options(minlevel=3, maxlevel=7, mx=16, my=16, mi=4, mj=1, ax=0.0, bx=2.0, ay=0.0,
                     by=0.5, solver=solver, precision=precision, gamma=1.4,
                     refine_threshold=0.1, coarsen_threshold=0.025, subcycle=True)


def c3_dt(coarsest=3):
    dx = 0.5 / (16 << coarsest)
    return 0.8 * dx / 3.6  # bubble sound speed sqrt(1.4/0.138) = 3.2 plus the flow


def c5_humps(seed=7, n=8):
    """qinit parameters of C5: h = 1 + sum of n Gaussian humps with seeded
    centres U(0.15, 0.85)^2, amplitudes U(0.1, 0.5) and widths U(0.02, 0.1)."""
    rng = np.random.default_rng(seed)
    p = [1.0, n]
    for _ in range(n):
        p += [rng.uniform(0.15, 0.85), rng.uniform(0.15, 0.85), rng.uniform(0.1, 0.5),
              rng.uniform(0.02, 0.1)]
    return p


def c5_options(precision=H.FP64, minlevel=7, maxlevel=9, refine_threshold=0.5):
    return F.Options(minlevel=minlevel, maxlevel=maxlevel, mx=32, my=32, solver=H.SOLVER_SWE_ROE,
                     precision=precision, grav=1.0, refine_threshold=refine_threshold,
                     coarsen_threshold=0.0, subcycle=True)


def c5_dt(dx_coarse):
    # CFL 0.9 at the coarsest level for |u| + sqrt(g h) <= 1 + sqrt(1.6)
    return 0.9 * dx_coarse / (1.0 + math.sqrt(1.6))


def build(name, precision=H.FP64, stream=None):
    """(simulation, coarse dt) with initial data and the initial refinement;
    stream: the HIP stream of the simulation (default: torch's current one;
    coarse steps are replayed as graphs only on a stream other than the null
    stream)."""
    if name == "C1":
        sim = F.Hipclaw46(c1_options(precision), stream=stream)
        sim.set_swirl(8.0)
        sim.initialize("swirl", (0.5,))
        return sim, c1_dt()
    if name == "C2":
        sim = F.Hipclaw46(c2_options(precision), stream=stream)
        sim.initialize("radial_dam", c2_case(), nsub=4)
        return sim, c2_dt(int(sim.patch_info()["level"].min()))
    if name in ("C3", "C3-HLLE"):
        solver = H.SOLVER_EULER_HLLE if name == "C3-HLLE" else H.SOLVER_EULER5_ROE
        sim = F.Hipclaw46(c3_options(precision, solver), stream=stream)
        sim.initialize("shock_bubble", c3_params(), nsub=4)
        return sim, c3_dt(int(sim.patch_info()["level"].min()))
    if name == "C5":
        sim = F.Hipclaw46(c5_options(precision), stream=stream)
        sim.initialize("humps", c5_humps())
        return sim, c5_dt(float(sim.patch_info()["dx"].max()))
    raise ValueError(name)


def _follow(sim64, precision):
    """A simulation on sim64's mesh with its data, in another precision."""
    o = sim64.opt
    opt = F.Options(**{**o.__dict__, "precision": precision})
    sim = F.Hipclaw46(opt)
    info = sim64.patch_info()
    sim.init_leaves(info["level"], info["gx"], info["gy"])
    if o.solver == H.SOLVER_ADV:
        sim.set_swirl(8.0)
    sim.set_q(sim64.get_q())
    return sim


def precision_report(name, checkpoints=(1, 10, 50)):
    sim64, dt = build(name, H.FP64)
    sim32 = _follow(sim64, H.FP32)
    out = []
    done = 0
    for c in checkpoints:
        for s in range(done, c):
            cfl = sim64.advance(s * dt, dt)
            sim32.advance(s * dt, dt)
            if not cfl <= 1.0:
                raise RuntimeError(f"{name}: Courant number {cfl} at step {s}")
        done = c
        a = sim32.get_q()[:, :, 2:-2, 2:-2]
        b = sim64.get_q()[:, :, 2:-2, 2:-2]
        out.append(dict(steps=c, rel_l1=float(np.abs(a - b).sum() / np.abs(b).sum()),
                        rel_linf=float(np.abs(a - b).max() / np.abs(b).max()),
                        patches=int(a.shape[0])))
    sim64.close()
    sim32.close()
    return out


# ---------------------------------------------------------------------------
# The reference's regression cases (applications/clawpack/*/regressions.sh):
# the options in effect for each run, run to tfinal with the run loop of
# fclaw2d_run outstyle 1, giving the three counters its regression_values CSV
# files hold.  The options in effect are the application's fclaw_options.ini
# overlaid by regression.ini: every option package is registered with
# fclaw_options.ini, which fclaw_app_options_parse loads before the command
# line's -F regression.ini (src/fclaw_base.c:759-792).  claw_version = 5 runs
# the same cases through the Clawpack 5 solver (regression_values_1.csv).
# ---------------------------------------------------------------------------
def swirl_regression(precision=H.FP64, claw_version=4):
    """applications/clawpack/advection/2d/swirl/regression.ini (claw-version 4):
    8x8 patches, levels 2-6, minmax tagging 0.25 / 0.05, smooth refinement,
    van Leer limiter, swirl period 4, tfinal 0.1 from initial_dt 0.005."""
    opt = F.Options(minlevel=2, maxlevel=6, mx=8, my=8, solver=H.SOLVER_ADV, precision=precision,
                    mthbc=(1, 1, 1, 1), mthlim=(3,), order=(2, 2), refine_threshold=0.25,
                    coarsen_threshold=0.05, smooth_refine=True, smooth_level=0,
                    refinement_criteria="minmax", regrid_interval=1, use_fixed_dt=False, initial_dt=0.005,
                    max_cfl=1.0, desired_cfl=0.9, subcycle=True, claw_version=claw_version)
    sim = F.Hipclaw46(opt)
    sim.set_swirl(4.0)
    sim.initialize("swirl", (0.5,), nsub=1)
    sim.run(0.1, nout=1)
    out = dict(count_amr_advance=sim.count_amr_advance, global_num_patches=sim.num_patches(),
               count_amr_new_domain=sim.count_amr_new_domain)
    sim.close()
    return out


def radialdam_regression(precision=H.FP64, claw_version=4):
    """applications/clawpack/shallow/2d/radialdam/regression.ini (example 0,
    claw-version 4): 16x16 patches on [-2.5, 2.5]^2, levels 1-5, gradient tagging
    (refinement-criteria of fclaw_options.ini) 0.5 / 0.25, smooth refinement
    from level 4, MC limiter, time-sync on, tfinal 0.4 from initial_dt 0.0625;
    the dam (r0 0.5, hin 2, hout 1) initialised by the reference's polygon cell
    average (nsub 0)."""
    opt = F.Options(minlevel=1, maxlevel=5, mx=16, my=16, solver=H.SOLVER_SWE_ROE, precision=precision,
                    mthbc=(1, 1, 1, 1), mthlim=(4, 4, 4), order=(2, 2), refine_threshold=0.5,
                    coarsen_threshold=0.25, smooth_refine=True, smooth_level=4, refinement_criteria="gradient",
                    regrid_interval=1, use_fixed_dt=False, initial_dt=0.0625, max_cfl=1.0, desired_cfl=0.9,
                    subcycle=True, time_sync=True, ax=-2.5, bx=2.5, ay=-2.5, by=2.5, grav=1.0,
                    claw_version=claw_version)
    sim = F.Hipclaw46(opt)
    sim.initialize("radial_dam", (0.0, 0.0, 0.5, 2.0, 1.0), nsub=0)
    sim.run(0.4, nout=1)
    out = dict(count_amr_advance=sim.count_amr_advance, global_num_patches=sim.num_patches(),
               count_amr_new_domain=sim.count_amr_new_domain)
    sim.close()
    return out



def shockbubble_regression(precision=H.FP64, claw_version=4):
    """applications/clawpack/euler/2d/shockbubble/regression.ini (claw-version 4):
    4x1 brick on [0, 2] x [0, 0.5], 16x16 patches, levels 0-3, difference
    tagging (refinement-criteria of fclaw_options.ini) 0.5 / 0.2 on density,
    smooth refinement, Euler + tracer Roe with limiters
    MC x4 and superbee, walls at y = 0 (the axis), the axisymmetric source,
    tfinal 0.25 from initial_dt 2.5e-3; the application's qinit (bubble r0 0.2
    at (0.5, 0), rhoin 0.1, pinf 5, gamma 1.4)."""
    opt = F.Options(minlevel=0, maxlevel=3, mx=16, my=16, mi=4, mj=1, ax=0.0, bx=2.0, ay=0.0, by=0.5,
                    solver=H.SOLVER_EULER5_ROE, precision=precision, gamma=1.4, mthbc=(1, 1, 3, 1),
                    mthlim=(4, 4, 4, 4, 2), order=(2, 2), refine_threshold=0.5, coarsen_threshold=0.2,
                    smooth_refine=True, smooth_level=0, refinement_criteria="difference", regrid_interval=1,
                    use_fixed_dt=False, initial_dt=2.5e-3, max_cfl=1.0, desired_cfl=0.9, subcycle=True,
                    src2=H.SRC2_AXISYMMETRIC, claw_version=claw_version)
    sim = F.Hipclaw46(opt)
    sim.initialize("shockbubble_app", (1.4, 0.5, 0.0, 0.2, 0.1, 5.0), nsub=0)
    sim.run(0.25, nout=1)
    out = dict(count_amr_advance=sim.count_amr_advance, global_num_patches=sim.num_patches(),
               count_amr_new_domain=sim.count_amr_new_domain)
    sim.close()
    return out
