"""CPU restatement of the reference's polygon cell average of a two-state
initial condition (cellave2 + fss + zeroin, src/fortran_source2d/cellave2.f:1-241)
with the radial-dam discontinuity (applications/clawpack/shallow/2d/radialdam/fdisc.f:1-27,
unmapped branch) and its qinit (radialdam/user_4.6/qinit.f:25-37).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product path
never imports it.

Python floats are IEEE doubles evaluated one operation at a time (no fused
multiply-add), in the reference's operation order, so the values compare
bitwise with a device restatement built without FMA contraction.

Pinned independently of this restatement (tests/test_cellave_oracle.py): the
crossing points of the cell edges with the circle have a closed form (the
roots of a quadratic along each edge), so the polygon area built from them
must agree to ~1e-13; cells wholly inside / outside give 1 / 0; the average
over a grid covering the disc converges to pi r0^2 at second order; and the
four-fold symmetry of a centred disc holds cell by cell.
"""
from __future__ import annotations

import math


def fdisc(x0, y0, r0, x, y):
    """radialdam/fdisc.f:21-24 (no mapping): r^2 - r0^2."""
    r2 = (x - x0) * (x - x0) + (y - y0) * (y - y0)
    return r2 - r0 * r0


def _eps():
    # cellave2.f:161-164: the relative machine precision by halving
    eps = 1.0
    while True:
        eps = eps / 2.0
        if not (1.0 + eps > 1.0):
            return eps


EPS = _eps()


def zeroin(f, ax, bx, tol):
    """cellave2.f:121-241: Brent's zero finder (Forsythe, Malcolm and Moler's
    ZEROIN) with the reference's step rule b - xm for xm < 0."""
    a, b = ax, bx
    fa, fb = f(a), f(b)
    restart = True
    c = fc = d = e = 0.0
    while True:
        if restart:                                   # label 20
            c, fc = a, fa
            d = b - a
            e = d
        if abs(fc) < abs(fb):                         # label 30
            a, b, c = b, c, b
            fa, fb, fc = fb, fc, fb
        tol1 = 2.0 * EPS * abs(b) + 0.5 * tol         # label 40
        xm = 0.5 * (c - b)
        if abs(xm) <= tol1 or fb == 0.0:
            return b
        bisect = abs(e) < tol1 or abs(fa) <= abs(fb)
        if not bisect:
            if a == c:
                s = fb / fa
                p = 2.0 * xm * s
                q = 1.0 - s
            else:
                q = fa / fc
                r = fb / fc
                s = fb / fa
                p = s * (2.0 * xm * q * (q - r) - (b - a) * (r - 1.0))
                q = (q - 1.0) * (r - 1.0) * (s - 1.0)
            if p > 0.0:
                q = -q
            p = abs(p)
            if (2.0 * p) >= (3.0 * xm * q - abs(tol1 * q)) or p >= abs(0.5 * e * q):
                bisect = True
            else:
                e = d
                d = p / q
        if bisect:                                    # label 70
            d = xm
            e = d
        a, fa = b, fb                                 # label 80
        if abs(d) > tol1:
            b = b + d
        if abs(d) <= tol1:
            b = b + tol1 if xm >= 0 else b - xm
        fb = f(b)
        restart = (fb * (fc / abs(fc))) > 0.0


def cellave2(fd, xlow, ylow, dx, dy):
    """cellave2.f:1-95: fraction of the cell [xlow, xlow+dx] x [ylow, ylow+dy]
    on the side fd < 0, as the area of the polygon through the inside corners
    and the edge crossings."""
    xx = [xlow, xlow, xlow + dx, xlow + dx]
    yy = [ylow, ylow + dy, ylow + dy, ylow]
    xx.append(xx[0])
    yy.append(yy[0])
    fl = [fd(xx[i], yy[i]) < 0.0 for i in range(4)]
    fl.append(fl[0])
    if all(fl[:4]):
        return 1.0
    if not any(fl[:4]):
        return 0.0
    x, y = [], []
    for i in range(4):
        if fl[i]:
            x.append(xx[i])
            y.append(yy[i])
        if fl[i] != fl[i + 1]:
            x0, y0, x1, y1 = xx[i], yy[i], xx[i + 1], yy[i + 1]
            ss = zeroin(lambda s: fd(x0 + s * (x1 - x0), y0 + s * (y1 - y0)), 0.0, 1.0, 1e-14)
            x.append(xx[i] + ss * (xx[i + 1] - xx[i]))
            y.append(yy[i] + ss * (yy[i + 1] - yy[i]))
    x.append(x[0])
    y.append(y[0])
    area = 0.0
    for i in range(len(x) - 1):
        area = area + 0.5 * (y[i] + y[i + 1]) * (x[i + 1] - x[i])
    return area / (dx * dy)


def radialdam_h(x0, y0, r0, hin, hout, xlow, ylow, dx, dy):
    """radialdam/user_4.6/qinit.f:30-31: h = hin*w + hout*(1-w)."""
    w = cellave2(lambda x, y: fdisc(x0, y0, r0, x, y), xlow, ylow, dx, dy)
    return hin * w + hout * (1.0 - w)


def radialdam_patch(prm, xlower, ylower, dx, dy, mx, my, mbc=2):
    """h over the whole patch, ghost cells included, shape (my+2mbc, mx+2mbc):
    xlow = xlower + (i-1)*dx for i = 1-mbc..mx+mbc (qinit.f:25-29)."""
    x0, y0, r0, hin, hout = prm
    out = []
    for j in range(1 - mbc, my + mbc + 1):
        ylow = ylower + (j - 1.0) * dy
        row = []
        for i in range(1 - mbc, mx + mbc + 1):
            xlow = xlower + (i - 1.0) * dx
            row.append(radialdam_h(x0, y0, r0, hin, hout, xlow, ylow, dx, dy))
        out.append(row)
    return out


def polygon_fraction_closed_form(x0, y0, r0, xlow, ylow, dx, dy):
    """Independent check: the same polygon with edge crossings from the
    quadratic |p0 + s (p1 - p0) - c|^2 = r0^2 solved in closed form."""
    xx = [xlow, xlow, xlow + dx, xlow + dx, xlow]
    yy = [ylow, ylow + dy, ylow + dy, ylow, ylow]
    ins = [(xx[i] - x0) ** 2 + (yy[i] - y0) ** 2 < r0 ** 2 for i in range(4)]
    ins.append(ins[0])
    if all(ins[:4]):
        return 1.0
    if not any(ins[:4]):
        return 0.0
    pts = []
    for i in range(4):
        if ins[i]:
            pts.append((xx[i], yy[i]))
        if ins[i] != ins[i + 1]:
            ex, ey = xx[i + 1] - xx[i], yy[i + 1] - yy[i]
            fx, fy = xx[i] - x0, yy[i] - y0
            A = ex * ex + ey * ey
            B = 2.0 * (fx * ex + fy * ey)
            C = fx * fx + fy * fy - r0 * r0
            disc = math.sqrt(max(B * B - 4 * A * C, 0.0))
            roots = [(-B - disc) / (2 * A), (-B + disc) / (2 * A)]
            s = min(roots, key=lambda t: abs(t - 0.5) if 0.0 <= t <= 1.0 else 9.0)
            pts.append((xx[i] + s * ex, yy[i] + s * ey))
    area = 0.0
    for k in range(len(pts)):
        (xa, ya), (xb, yb) = pts[k], pts[(k + 1) % len(pts)]
        area += 0.5 * (ya + yb) * (xb - xa)
    return area / (dx * dy)
