"""Pins the oracle's cell list: grid sizing against the reference's default
domain, FP32 keys against an independent float32 evaluation, the sort
against numpy's lexicographic sort of (key, id), BeginEndCube against binary
search, and the neighbour search against brute force."""
import numpy as np
import pytest

from oracle import pyoracle as orc
from tests import util


def test_config_domain_reference_default():
    # main.cuh:11-12: CutRadius 1, BBox {0,20,-10,10,-10,10}; Particle.cu:551-560
    # pads y and z by one cutoff: 20 x 22 x 22 cubes.
    nc, outer, inv, K = orc.config_domain([0, 20, -10, 10, -10, 10], 1.0, [1, 0, 0])
    assert list(nc) == [20, 22, 22]
    assert list(outer) == [20.0, 22.0, 22.0]
    assert K == 20 * 22 * 22
    assert np.allclose(inv, [1.0, 1.0, 1.0])
    with pytest.raises(ValueError):
        orc.config_domain([0, 2, 0, 10, 0, 10], 1.0, [1, 0, 0])  # periodic axis < 3 cubes


def test_preconlist_keys_float32():
    rng = np.random.default_rng(3)
    bbox = np.array([0, 31.7, -10, 10, -7.5, 7.5], np.float32)
    per = [1, 0, 0]
    cut = np.float32(1.5)
    nc, outer, inv, K = orc.config_domain(bbox, cut, per)
    n = 5000
    x = np.empty((n, 3), np.float32)
    x[:, 0] = rng.uniform(0, 31.7, n)
    x[:, 1] = rng.uniform(-11.5, 11.5, n)   # some beyond the padding: clamped
    x[:, 2] = rng.uniform(-9, 9, n)
    sp = rng.permutation(n).astype(np.int32)
    cc, key, old = orc.preconlist(x, cut, per, bbox, inv, nc, sp)
    for d in range(3):
        off = np.float32(0.0) if per[d] else cut
        t = (x[:, d] - bbox[2 * d]).astype(np.float32)
        u = (t + off).astype(np.float32)
        f = (u * inv[d]).astype(np.float32)
        k = np.clip(np.floor(f).astype(np.int64), 0, nc[d] - 1)
        assert np.array_equal(cc[:, d], k)
    assert np.array_equal(key, cc[:, 0] + nc[0] * (cc[:, 1] + nc[1] * cc[:, 2]))
    assert np.array_equal(old[sp], np.arange(n))
    assert key.min() >= 0 and key.max() < K


def test_sort_and_begin_end():
    rng = np.random.default_rng(5)
    n, K = 20000, 3000
    keys = rng.integers(0, K, n).astype(np.int32)
    vals = rng.permutation(n).astype(np.int32)
    k2, v2 = orc.sort(keys, vals)
    o = np.lexsort((vals, keys))
    assert np.array_equal(k2, keys[o]) and np.array_equal(v2, vals[o])
    be = orc.begin_end(k2, K)
    kk = np.arange(K)
    assert np.array_equal(be[:, 0], np.searchsorted(k2, kk, "left"))
    assert np.array_equal(be[:, 1], np.searchsorted(k2, kk, "right"))


@pytest.mark.parametrize("n", [800, 4000])  # brute-force and binned search paths
def test_density_neighbour_search_against_brute_force(n):
    L = 12.0 if n < 1000 else 16.0
    pos = util.random_fluid(n, L, seed=n)
    ph = util.replace_physics(h=1.5)
    case = util.mini_case(pos, L=(L, L, L), physics=ph)
    o = orc.Oracle(case)
    rho = np.zeros(n)
    o.density(pos.astype(np.float32).astype(np.float64), rho)
    p = pos.astype(np.float32).astype(np.float64)
    d = p[:, None, :] - p[None, :, :]
    d -= L * np.round(d / L)
    r = np.sqrt((d ** 2).sum(-1))
    q = r / ph.h
    W = np.where(q < 1, 105 / (16 * np.pi * ph.h ** 3) * (1 + 3 * q) * (1 - q) ** 3, 0.0)
    assert np.allclose(rho, ph.mass * W.sum(axis=1), rtol=1e-12, atol=0)


def test_density_walled_axes():
    # walls (non-periodic y, z): no image contributions across them
    n, L = 3500, 15.0
    pos = util.random_fluid(n, L, seed=11)
    ph = util.replace_physics(h=1.5)
    case = util.mini_case(pos, L=(L, L, L), physics=ph, periodic=(1, 0, 0))
    o = orc.Oracle(case)
    rho = np.zeros(n)
    p = pos.astype(np.float32).astype(np.float64)
    o.density(p, rho)
    d = p[:, None, :] - p[None, :, :]
    d[..., 0] -= L * np.round(d[..., 0] / L)
    r = np.sqrt((d ** 2).sum(-1))
    q = r / ph.h
    W = np.where(q < 1, 105 / (16 * np.pi * ph.h ** 3) * (1 + 3 * q) * (1 - q) ** 3, 0.0)
    assert np.allclose(rho, W.sum(axis=1), rtol=1e-12, atol=0)
