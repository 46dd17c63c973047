"""Pins the oracle's restatement of the reference's polygon cell average
(oracle/cellave.py; src/fortran_source2d/cellave2.f) by checks that do not
retype it:

* the edge crossings have a closed form (a quadratic along each edge); the
  polygon built from them must give the same fraction to ~1e-13 on cells cut
  by the circle in every way (one, two or three corners inside);
* cells wholly inside / outside the disc give exactly 1 / 0;
* summed over a grid that covers the disc, the fractions converge to the disc
  area pi r0^2 at second order (the chords cut off O(h^3) per cut cell);
* a centred disc gives the same fraction on the four mirror cells;
* zeroin finds the roots of polynomials with a known root to its tolerance.
"""
import math

import numpy as np
import pytest

from oracle import cellave as C


def test_zeroin_known_roots():
    for root in (0.1, 0.5, 0.73, 1e-3, 0.999):
        s = C.zeroin(lambda t: (t - root) * (t + 2.0), 0.0, 1.0, 1e-14)
        assert abs(s - root) < 1e-13, root
    s = C.zeroin(lambda t: t * t * t - 0.2, 0.0, 1.0, 1e-14)
    assert abs(s - 0.2 ** (1 / 3)) < 1e-13


def test_whole_cells():
    fd = lambda x, y: C.fdisc(0.0, 0.0, 1.0, x, y)
    assert C.cellave2(fd, -0.1, -0.1, 0.2, 0.2) == 1.0
    assert C.cellave2(fd, 2.0, 2.0, 0.2, 0.2) == 0.0
    # a cell that straddles the circle but whose corners are all outside is
    # counted empty by the corner test, as in the reference
    assert C.cellave2(fd, 0.99, -0.3, 0.21, 0.6) == 0.0


@pytest.mark.parametrize("seed", range(4))
def test_against_closed_form_crossings(seed):
    rng = np.random.default_rng(seed)
    x0, y0, r0 = rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), rng.uniform(0.3, 0.8)
    fd = lambda x, y: C.fdisc(x0, y0, r0, x, y)
    kinds = set()
    for _ in range(400):
        h = rng.uniform(0.01, 0.2)
        th = rng.uniform(0, 2 * math.pi)
        xl = x0 + r0 * math.cos(th) - rng.uniform(0, h)
        yl = y0 + r0 * math.sin(th) - rng.uniform(0, h)
        ins = sum(fd(xl + a * h, yl + b * h) < 0 for a in (0, 1) for b in (0, 1))
        kinds.add(ins)
        w = C.cellave2(fd, xl, yl, h, h)
        wc = C.polygon_fraction_closed_form(x0, y0, r0, xl, yl, h, h)
        assert abs(w - wc) < 1e-11, (xl, yl, h, w, wc)
        assert 0.0 <= w <= 1.0
    assert {1, 2, 3} <= kinds


def _disc_area(r0, n):
    L = 1.25
    h = 2 * L / n
    fd = lambda x, y: C.fdisc(0.0, 0.0, r0, x, y)
    tot = 0.0
    for j in range(n):
        for i in range(n):
            tot += C.cellave2(fd, -L + i * h, -L + j * h, h, h)
    return tot * h * h


def test_area_converges_second_order():
    r0 = 1.0
    e1 = abs(_disc_area(r0, 40) - math.pi * r0 * r0)
    e2 = abs(_disc_area(r0, 80) - math.pi * r0 * r0)
    assert e1 < 5e-3 and e2 < e1 / 3.0, (e1, e2)


def test_mirror_symmetry_of_centred_disc():
    fd = lambda x, y: C.fdisc(0.0, 0.0, 0.5, x, y)
    h = 0.0625
    for j in range(-10, 10):
        for i in range(-10, 10):
            xl, yl = i * h, j * h
            w = C.cellave2(fd, xl, yl, h, h)
            assert abs(w - C.cellave2(fd, -xl - h, yl, h, h)) < 1e-12
            assert abs(w - C.cellave2(fd, xl, -yl - h, h, h)) < 1e-12
            assert abs(w - C.cellave2(fd, yl, xl, h, h)) < 1e-12
