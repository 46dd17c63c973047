"""Pins the oracle's Clawpack 5 variant (orc_cfg.claw5) against the reference's
own Clawpack 5 patch step (src/solvers/fc2d_clawpack5/fortran_source/
clawpack5_step2_wrap.f, clawpack5_step2.f90, clawpack5_flux2.f,
clawpack5_inlinelimiter.f with the applications' clawpack5_rpn2 / rpt2),
compiled unchanged into oracle/_ref by oracle/Makefile:

* committed golden vectors (tests/golden/claw5/, Clawpack 5 layout
  q(meqn, i, j)): interiors and Courant numbers bitwise;
* live random cases when oracle/_ref is built;
* the variant matters: the 4.6 arithmetic differs from the golden Clawpack 5
  results in the last bits for SWE and Euler (operation order of the wave
  strengths and of the transverse split), and limiter 6 is minmod there.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "claw5", "step2_claw5.npz")
SOL = {"swe": O.SOLVER_SWE_ROE, "euler": O.SOLVER_EULER5_ROE, "adv": O.SOLVER_ADV}


def from5(a):
    return np.ascontiguousarray(np.transpose(a, (2, 0, 1)))


def to5(a):
    return np.ascontiguousarray(np.transpose(a, (1, 2, 0)))


def cases():
    z = np.load(GOLD)
    n = len([k for k in z.files if k.startswith("dims")])
    out = []
    for k in range(n):
        d = dict(phys=str(z[f"phys{k}"]), q=from5(z[f"q{k}"]), qout=from5(z[f"qout{k}"]), cfl=float(z[f"cfl{k}"]),
                 mx=int(z[f"dims{k}"][0]), my=int(z[f"dims{k}"][1]), dx=float(z[f"geom{k}"][0]),
                 dy=float(z[f"geom{k}"][1]), dt=float(z[f"geom{k}"][2]), mthlim=[int(v) for v in z[f"mthlim{k}"]])
        d["aux"] = from5(z[f"aux{k}"]) if f"aux{k}" in z.files else None
        out.append(d)
    return out


def _oracle(c, claw5):
    cfg = O.make_cfg(SOL[c["phys"]], c["mx"], c["my"], 2, mthlim=c["mthlim"], claw5=claw5)
    return O.step2(cfg, c["q"], c["aux"], c["dx"], c["dy"], c["dt"])


def test_oracle_claw5_matches_reference_golden():
    for k, c in enumerate(cases()):
        q, cfl = _oracle(c, True)
        assert np.array_equal(q[:, 2:-2, 2:-2], c["qout"][:, 2:-2, 2:-2]), (k, c["phys"])
        assert cfl == c["cfl"], k


def test_the_variant_matters():
    differs = set()
    for c in cases():
        q, _ = _oracle(c, False)
        if not np.array_equal(q[:, 2:-2, 2:-2], c["qout"][:, 2:-2, 2:-2]):
            differs.add((c["phys"], 6 in c["mthlim"]))
    # SWE and Euler round differently; advection differs only through limiter 6
    assert {("swe", False), ("euler", False), ("adv", True)} <= differs
    assert ("adv", False) not in differs


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref5_swe.so")),
                    reason="oracle/_ref Clawpack 5 libraries not built")
def test_oracle_claw5_matches_live_reference():
    sys.path.insert(0, os.path.join(HERE, "golden"))
    import cases as C
    import refdrive as R
    rng = np.random.default_rng(17)
    for k in range(9):
        phys = ("swe", "euler", "adv")[k % 3]
        mx, my = (int(v) for v in rng.integers(4, 24, 2))
        aux = None
        if phys == "swe":
            q, dx, dy = C.swe_state(rng, mx, my, 2)
            dt = C.swe_dt(q, dx, dy)
            R.set_params5("swe", grav=1.0)
            lim = [int(v) for v in rng.integers(0, 7, 3)]
        elif phys == "euler":
            q, dx, dy = C.euler_state(rng, mx, my, 2)
            dt = C.euler_dt(q, dx, dy)
            R.set_params5("euler", gamma=1.4)
            lim = [int(v) for v in rng.integers(0, 7, 5)]
        else:
            q, dx, dy = C.adv_state(rng, mx, my, 2)
            aux = rng.normal(0, 1, (2, my + 4, mx + 4))
            dt = C.adv_dt(aux, dx, dy)
            lim = [int(rng.integers(0, 7))]
        order = [(2, 2), (2, 1), (1, 2), (1, 0)][k % 4]
        ref, cref = R.step2_claw5(phys, to5(q), None if aux is None else to5(aux), mx=mx, my=my, mbc=2, dx=dx,
                                  dy=dy, dt=dt, order=order, mthlim=lim)
        cfg = O.make_cfg(SOL[phys], mx, my, 2, mthlim=lim, order=order, claw5=True)
        got, cgot = O.step2(cfg, q, aux, dx, dy, dt)
        assert np.array_equal(got[:, 2:-2, 2:-2], from5(ref)[:, 2:-2, 2:-2]), (k, phys, lim, order)
        assert cgot == cref, k
