"""Write golden input/output vectors from the reference's own Fortran.

Run in the build container (needs /root/reference and `make -C oracle ref`):
    python tests/golden/make_golden.py
Writes tests/golden/*.npz (inputs, reference outputs and parameters).  The
fixtures are data only; the reference itself never leaves this container.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import cases as C  # noqa: E402
import refdrive as R  # noqa: E402

SEED = 20261203


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


def main():
    if not R.available():
        raise SystemExit("oracle/_ref not built: run `make -C oracle ref` first")
    rng = np.random.default_rng(SEED)

    # ---- shallow water, Roe + entropy fix -------------------------------
    swe = []
    for (mx, my), order, lim in [((16, 12), (2, 2), (4, 4, 4)),
                                 ((32, 32), (2, 2), (4, 4, 4)),
                                 ((16, 16), (2, 1), (1, 2, 3)),
                                 ((16, 16), (1, 0), (4, 4, 4)),
                                 ((12, 16), (2, 2), (5, 6, 4)),
                                 ((16, 16), (1, 2), (4, 4, 4))]:
        mbc = 2
        q, dx, dy = C.swe_state(rng, mx, my, mbc)
        dt = C.swe_dt(q, dx, dy)
        R.set_params("swe", grav=1.0)
        qo, cfl = R.step2("swe", q, None, mx=mx, my=my, mbc=mbc, dx=dx, dy=dy, dt=dt,
                          order=order, mthlim=lim)
        swe.append(dict(q=q, qout=qo, cfl=cfl, mx=mx, my=my, mbc=mbc, dx=dx, dy=dy, dt=dt,
                        order=np.array(order), mthlim=np.array(lim), grav=1.0))
    for k, c in enumerate(swe):
        save(f"swe_step_{k}", **c)

    # ---- Euler + tracer, Roe + entropy fix ------------------------------
    for k, ((mx, my), order) in enumerate([((16, 16), (2, 2)), ((16, 12), (2, 1))]):
        mbc = 2
        q, dx, dy = C.euler_state(rng, mx, my, mbc)
        dt = C.euler_dt(q, dx, dy)
        R.set_params("euler", gamma=1.4)
        qo, cfl = R.step2("euler", q, None, mx=mx, my=my, mbc=mbc, dx=dx, dy=dy, dt=dt,
                          order=order, mthlim=[4] * 5)
        save(f"euler_step_{k}", q=q, qout=qo, cfl=cfl, mx=mx, my=my, mbc=mbc, dx=dx, dy=dy,
             dt=dt, order=np.array(order), mthlim=np.array([4] * 5), gamma=1.4)

    # ---- advection with the swirl velocity field ------------------------
    for k, ((mx, my), order, t) in enumerate([((16, 16), (2, 2), 0.3),
                                              ((32, 32), (2, 2), 1.7),
                                              ((16, 16), (2, 1), 2.9)]):
        mbc = 2
        q, dx, dy = C.adv_state(rng, mx, my, mbc)
        xlower, ylower = 2 * mx * dx, 3 * my * dy
        aux = R.swirl_aux(mx=mx, my=my, mbc=mbc, xlower=xlower, ylower=ylower, dx=dx, dy=dy,
                          t=t, dt=0.01, tperiod=8.0)
        dt = C.adv_dt(aux, dx, dy)
        R.set_params("adv", tperiod=8.0)
        qo, cfl = R.step2("adv", q, aux, mx=mx, my=my, mbc=mbc, dx=dx, dy=dy, dt=dt,
                          order=order, mthlim=[4])
        save(f"adv_step_{k}", q=q, aux=aux, qout=qo, cfl=cfl, mx=mx, my=my, mbc=mbc, dx=dx,
             dy=dy, dt=dt, order=np.array(order), mthlim=np.array([4]), xlower=xlower,
             ylower=ylower, t=t, auxdt=0.01, tperiod=8.0)

    # ---- physical boundary conditions -----------------------------------
    bcs = []
    for mthbc in ([1, 1, 1, 1], [3, 3, 3, 3], [1, 3, -1, 3], [-1, -1, 1, 3], [3, -1, 3, -1]):
        mx, my, mbc = 10, 8, 2
        q = rng.standard_normal((3, my + 2 * mbc, mx + 2 * mbc))
        bcs.append((np.array(mthbc), q, R.bc2(q, mx=mx, my=my, mbc=mbc, mthbc=mthbc)))
    save("bc2", mthbc=np.stack([b[0] for b in bcs]), q=np.stack([b[1] for b in bcs]),
         qout=np.stack([b[2] for b in bcs]), mx=10, my=8, mbc=2)

    # ---- time interpolation band ----------------------------------------
    mx, my, mbc = 12, 10, 2
    qc = rng.standard_normal((3, my + 2 * mbc, mx + 2 * mbc))
    ql = rng.standard_normal((3, my + 2 * mbc, mx + 2 * mbc))
    qi, ierr = R.timeinterp(qc, ql, 0.375, mx=mx, my=my, mbc=mbc)
    assert ierr == 0
    save("timeinterp", qcurr=qc, qlast=ql, qinterp=qi, alpha=0.375, mx=mx, my=my, mbc=mbc)
    fw_golden()


def fw_golden():
    """f-wave step fixtures (python tests/golden/make_golden.py fw)."""
    rng = np.random.default_rng(SEED + 11)
    # ---- conservative transport with f-waves (flux2fw) ------------------
    for k, ((mx, my), order, lim, opts) in enumerate([((16, 16), (2, 2), 4, {}),
                                                      ((16, 12), (2, 2), 1, {"per_edge": True}),
                                                      ((12, 16), (2, 1), 3, {"lengths": True}),
                                                      ((16, 16), (1, 2), 4, {"per_edge": True, "lengths": True})]):
        mbc = 2
        q, aux, dx, dy = C.fw_state(rng, mx, my, mbc, **opts)
        dt = C.fw_dt(aux, dx, dy)
        qo, cfl = R.step2("fw", q, aux, mx=mx, my=my, mbc=mbc, dx=dx, dy=dy, dt=dt, order=order,
                          mthlim=[lim])
        save(f"fw_step_{k}", q=q, aux=aux, qout=qo, cfl=cfl, mx=mx, my=my, mbc=mbc, dx=dx, dy=dy,
             dt=dt, order=np.array(order), mthlim=np.array([lim]))


def ascii_golden():
    """fort.t0003 / fort.q0003 written by the reference's Fortran writers for a
    few patches whose values exercise the E26.16 edge cases (zero, tiny values
    that the writer zeroes, negative values, rounding carries, exponents of
    three digits), with the inputs, under tests/golden/ascii/."""
    rng = np.random.default_rng(SEED + 7)
    out = os.path.join(HERE, "ascii")
    os.makedirs(out, exist_ok=True)
    patches = []
    for k, (mx, my, meqn, level, block) in enumerate([(5, 3, 3, 4, 0), (4, 4, 1, 6, 0), (3, 2, 5, 2, 1)]):
        q = rng.standard_normal((meqn, my + 4, mx + 4)) * 10.0 ** rng.integers(-8, 9, (meqn, my + 4, mx + 4))
        q[0, 2, 2] = 0.0
        q[0, 2, 3] = 3e-120            # |q| < 1e-99: written as zero
        q[0, 3, 2] = -0.99999999999999999
        q[-1, 2, 4] = 9.999999999999999e99
        q[-1, 3, 3] = 1.5e150          # three-digit exponent
        q[-1, 3, 4] = -2.5e-77
        h = 1.0 / (mx << level)
        patches.append(dict(q=q, mbc=2, xlower=-2.5 + 5 * (k + 1) * h, ylower=0.125 * k, dx=5 * h, dy=5 * h,
                            num=3 * k + 1, level=level, block=block))
    R.ascii_frame(out, 3, 0.1 + 1e-17, 3, 0, patches)
    save(os.path.join("ascii", "patches"), **{f"{key}{i}": np.asarray(p[key]) for i, p in enumerate(patches)
                                              for key in p})


def src2_golden():
    """The shock-bubble application's src2 and setaux (reference Fortran) on
    random Euler states near the axis and away from it."""
    rng = np.random.default_rng(SEED + 7)
    out = {}
    for c, (mx, my, ylower, dt) in enumerate([(8, 8, 0.0, 1e-3), (16, 12, 0.0, 4e-3), (6, 10, 0.3, 2e-2)]):
        mbc = 2
        dy = 0.5 / (4 * my)
        aux = R.setaux_axisym(mx=mx, my=my, mbc=mbc, xlower=0.1, ylower=ylower, dx=dy, dy=dy)
        rho = rng.uniform(0.1, 3.0, (my + 4, mx + 4))
        u, v = rng.normal(0, 1.0, (2, my + 4, mx + 4))
        p = rng.uniform(0.5, 5.0, (my + 4, mx + 4))
        q = np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v),
                      rng.uniform(0, 2, (my + 4, mx + 4))])
        qo = R.src2_axisym(q, aux, mx=mx, my=my, mbc=mbc, xlower=0.1, ylower=ylower, dx=dy, dy=dy, t=0.0,
                           dt=dt, gamma=1.4)
        out.update({f"q{c}": q, f"aux{c}": aux, f"out{c}": qo, f"geom{c}": np.array([ylower, dy, dt, 1.4])})
    os.makedirs(os.path.join(HERE, "src2"), exist_ok=True)
    save(os.path.join("src2", "src2"), **out)


def claw5_golden():
    """The Clawpack 5 patch step (clawpack5_step2_wrap with the applications'
    clawpack5_rpn2 / rpt2), inputs and outputs in the Clawpack 5 layout
    q(meqn, i, j) (numpy (NY, NX, meqn)), including limiter 6, which that
    step's limiter treats as minmod."""
    if not R.available5():
        raise SystemExit("oracle/_ref Clawpack 5 libraries not built: run `make -C oracle ref` first")
    rng = np.random.default_rng(SEED + 5)
    t5 = lambda a: np.ascontiguousarray(np.transpose(a, (1, 2, 0)))  # noqa: E731
    out = {}
    k = 0
    for phys, (mx, my), lim in [("swe", (16, 12), (4, 4, 4)), ("swe", (16, 16), (5, 6, 3)),
                                ("euler", (16, 16), (4, 4, 4, 4, 2)), ("euler", (12, 16), (1, 2, 6, 4, 3)),
                                ("adv", (16, 16), (4,)), ("adv", (12, 12), (6,))]:
        mbc = 2
        aux = None
        if phys == "swe":
            q, dx, dy = C.swe_state(rng, mx, my, mbc)
            dt = C.swe_dt(q, dx, dy)
            R.set_params5("swe", grav=1.0)
        elif phys == "euler":
            q, dx, dy = C.euler_state(rng, mx, my, mbc)
            dt = C.euler_dt(q, dx, dy)
            R.set_params5("euler", gamma=1.4)
        else:
            q, dx, dy = C.adv_state(rng, mx, my, mbc)
            aux = t5(R.swirl_aux(mx=mx, my=my, mbc=mbc, xlower=0.0, ylower=0.0, dx=dx, dy=dy, t=0.4,
                                 dt=0.01, tperiod=8.0))
            dt = C.adv_dt(np.transpose(aux, (2, 0, 1)), dx, dy)
        q5 = t5(q)
        qo, cfl = R.step2_claw5(phys, q5, aux, mx=mx, my=my, mbc=mbc, dx=dx, dy=dy, dt=dt, mthlim=lim)
        out.update({f"q{k}": q5, f"qout{k}": qo, f"cfl{k}": cfl, f"dims{k}": np.array([mx, my]),
                    f"geom{k}": np.array([dx, dy, dt]), f"mthlim{k}": np.array(lim),
                    f"phys{k}": np.array(phys)})
        if aux is not None:
            out[f"aux{k}"] = aux
        k += 1
    os.makedirs(os.path.join(HERE, "claw5"), exist_ok=True)
    save(os.path.join("claw5", "step2_claw5"), **out)


if __name__ == "__main__":
    if sys.argv[1:] == ["ascii"]:
        ascii_golden()
    elif sys.argv[1:] == ["src2"]:
        src2_golden()
    elif sys.argv[1:] == ["claw5"]:
        claw5_golden()
    elif sys.argv[1:] == ["fw"]:
        fw_golden()
    else:
        main()
