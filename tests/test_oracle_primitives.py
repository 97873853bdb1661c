"""Pins the oracle's primitives: Philox-4x32-10 against the published
Random123 known-answer vectors, the Box-Muller normals statistically, and
the Lucy kernel against its closed forms (normalisation, W(0), F = -W'/r)."""
import json
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden", "philox4x32_10_kat.json")


def test_philox_known_answers():
    vec = json.load(open(GOLD))["vectors"]
    assert len(vec) == 3
    for v in vec:
        ctr = [int(x, 16) for x in v["ctr"]]
        key = [int(x, 16) for x in v["key"]]
        out = [int(x, 16) for x in v["out"]]
        assert list(orc.philox(ctr, key)) == out


def test_pair_normals_symmetric_and_standard():
    # same pair -> same normals whatever the order; moments of N(0, 1)
    assert np.array_equal(orc.pair_normals(7, 3, 10, 20), orc.pair_normals(7, 3, 20, 10))
    xs = np.array([orc.pair_normals(1, c, i, i + 1) for c in range(40) for i in range(500)])
    assert xs.shape == (20000, 3)
    m, s = xs.mean(axis=0), xs.std(axis=0)
    assert np.all(np.abs(m) < 0.03), m
    assert np.all(np.abs(s - 1) < 0.03), s
    # the three components are uncorrelated
    c = np.corrcoef(xs.T)
    assert np.all(np.abs(c - np.eye(3)) < 0.03)
    # a different counter gives different numbers
    assert not np.array_equal(orc.pair_normals(1, 0, 1, 2), orc.pair_normals(1, 1, 1, 2))


@pytest.mark.parametrize("h", [1.0, 1.5, 2.7])
def test_lucy_closed_forms(h):
    # W(0) = 105 / (16 pi h^3)
    assert orc.lucy_w(0.0, h) == pytest.approx(105.0 / (16.0 * math.pi * h ** 3), rel=1e-14)
    # unit normalisation: 4 pi int_0^h W r^2 dr = 1 (Gauss-Legendre)
    x, w = np.polynomial.legendre.leggauss(64)
    r = 0.5 * h * (x + 1)
    integ = 0.5 * h * sum(wi * 4 * math.pi * ri * ri * orc.lucy_w(ri, h) for ri, wi in zip(r, w))
    assert integ == pytest.approx(1.0, rel=1e-12)
    # compact support and F = -W'(r)/r by central differences
    assert orc.lucy_w(h, h) == 0.0 and orc.lucy_w(1.3 * h, h) == 0.0
    for q in (0.1, 0.35, 0.6, 0.9):
        rr, d = q * h, 1e-6 * h
        dw = (orc.lucy_w(rr + d, h) - orc.lucy_w(rr - d, h)) / (2 * d)
        assert orc.lucy_f(rr, h) == pytest.approx(-dw / rr, rel=1e-7)
