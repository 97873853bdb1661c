"""Pins the oracle's membrane and Morse terms:

* geometry: exact area and volume of the regular icosahedron and of a cube;
* forces = -grad E by FP64 central differences on a deformed RBC, for each
  energy term (WLC-POW, bending, local/global area, volume);
* translation and rotation invariance: sum F = 0, sum r x F = 0;
* the reference state is force-free (k_p balances the WLC spring at l0);
* unwrapping across a periodic face changes nothing;
* Morse: -dU/dr by central differences, zero force at r0, only between
  different cells, Newton's third law."""
import math

import numpy as np
import pytest

from gpucell_amd import casegen
from oracle import pyoracle as orc
from tests import util


def _cell_case(V, F, ph, L=40.0, center=20.0, periodic=(1, 1, 1), n_cells=1, shift=None):
    verts = [V + center + (0 if shift is None else shift * c) for c in range(n_cells)]
    pos = np.concatenate(verts)
    return util.mini_case(pos, nm=pos.shape[0], L=(L, L, L), periodic=periodic, physics=ph,
                          cells=(F, n_cells))


def test_geometry_icosahedron_and_cube():
    V, F = casegen.icosphere(0)
    ph = util.replace_physics()
    case = _cell_case(V, F, ph, center=0.0)
    o = orc.Oracle(case)
    g = o.geo(-1, case.coor.astype(np.float64), np.zeros((12, 3), np.int32))
    a = 4.0 / math.sqrt(10 + 2 * math.sqrt(5))  # edge of the unit-circumradius icosahedron
    # the float32 vertex coordinates differ from the exact ones by < 6e-8
    assert g["cell_area"][0] == pytest.approx(5 * math.sqrt(3) * a * a, rel=5e-7)
    assert g["cell_vol"][0] == pytest.approx(5.0 / 12.0 * (3 + math.sqrt(5)) * a ** 3, rel=5e-7)
    assert g["par_area"].sum() == pytest.approx(g["cell_area"][0], rel=1e-12)
    # cube of side 2: 8 vertices, 12 outward triangles
    C = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], float)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    T = []
    for a0, b0, c0, d0 in quads:
        T += [(a0, b0, c0), (a0, c0, d0)]
    T = np.array(T)
    n = np.cross(C[T[:, 1]] - C[T[:, 0]], C[T[:, 2]] - C[T[:, 0]])
    flip = np.einsum("ij,ij->i", n, C[T].mean(1)) < 0
    T[flip] = T[flip][:, [0, 2, 1]]
    case = _cell_case(C, T, ph, center=0.0)
    o = orc.Oracle(case)
    g = o.geo(-1, case.coor.astype(np.float64), np.zeros((8, 3), np.int32))
    assert g["cell_area"][0] == pytest.approx(24.0, rel=1e-12)
    assert g["cell_vol"][0] == pytest.approx(8.0, rel=1e-12)
    assert np.allclose(np.linalg.norm(g["tri_nor"], axis=1), 1.0)


def _deformed_rbc(seed=0, scale=0.08, subdiv=2, D0=6.0, **phys):
    V, F = casegen.icosphere(subdiv)
    S = casegen.biconcave(V, D0)
    ph = util.replace_physics(**phys)
    case = _cell_case(S, F, ph)
    o = orc.Oracle(case)
    pos0 = case.coor.astype(np.float64)
    img = np.zeros((pos0.shape[0], 3), np.int32)
    o.geo(-1, pos0, img)
    o.mech_para(pos0, img)
    rng = np.random.default_rng(seed)
    pos = pos0 + rng.normal(0, scale, size=pos0.shape)
    return o, case, pos0, pos, img


@pytest.mark.parametrize("term", ["all", "spring", "bending", "local", "global", "volume"])
def test_membrane_force_is_energy_gradient(term):
    k = dict(mu0=1.0, kc=0.5, kd=3.0, ka=20.0, kv=40.0, kT=0.02)
    off = {"spring": dict(kc=0.0, kd=0.0, ka=0.0, kv=0.0),
           "bending": dict(mu0=1e30, kd=0.0, ka=0.0, kv=0.0),
           "local": dict(mu0=1e30, kc=0.0, ka=0.0, kv=0.0),
           "global": dict(mu0=1e30, kc=0.0, kd=0.0, kv=0.0),
           "volume": dict(mu0=1e30, kc=0.0, kd=0.0, ka=0.0)}.get(term, {})
    k.update(off)
    o, case, pos0, pos, img = _deformed_rbc(**k)
    o.geo(0, pos, img)
    F = o.def_force(pos, img)
    assert np.abs(F).max() > 1e-6
    rng = np.random.default_rng(7)
    eps = 1e-6
    for v in rng.choice(pos.shape[0], 10, replace=False):
        g = np.zeros(3)
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[v, d] += eps
            pm[v, d] -= eps
            g[d] = (o.membrane_energy(pp, img) - o.membrane_energy(pm, img)) / (2 * eps)
        assert np.allclose(F[v], -g, rtol=2e-5, atol=2e-6 * np.abs(F).max()), (term, v, F[v], -g)


def test_membrane_invariances_and_reference_state():
    o, case, pos0, pos, img = _deformed_rbc(mu0=1.0, kc=0.5, kd=3.0, ka=20.0, kv=40.0)
    o.geo(0, pos, img)
    F = o.def_force(pos, img)
    scale = np.abs(F).sum()
    assert np.abs(F.sum(0)).max() < 1e-10 * scale
    c = pos.mean(0)
    assert np.abs(np.cross(pos - c, F).sum(0)).max() < 1e-9 * scale
    # reference state: every term vanishes
    o.geo(0, pos0, img)
    F0 = o.def_force(pos0, img)
    assert np.abs(F0).max() < 1e-9 * np.abs(F).max()
    # a force-free state is the energy minimum (WLC and POW cancel in slope
    # at l0, not in value)
    assert o.membrane_energy(pos0, img) < o.membrane_energy(pos, img)


def test_unwrapping_across_periodic_face():
    o, case, pos0, pos, img = _deformed_rbc(mu0=1.0, kc=0.5, kd=3.0, ka=20.0, kv=40.0)
    o.geo(0, pos, img)
    F = o.def_force(pos, img)
    g = o.geo(0, pos, img)
    # shift the whole cell so it straddles x = 0 and wrap it into the box
    L = 40.0
    p2 = pos.copy()
    p2[:, 0] -= 20.0
    img2 = np.zeros_like(img)
    neg = p2[:, 0] < 0
    p2[neg, 0] += L
    img2[neg, 0] = -1
    g2 = o.geo(0, p2, img2)
    F2 = o.def_force(p2, img2)
    assert g2["cell_area"][0] == pytest.approx(g["cell_area"][0], rel=1e-12)
    assert g2["cell_vol"][0] == pytest.approx(g["cell_vol"][0], rel=1e-10)
    assert np.allclose(F2, F, rtol=1e-8, atol=1e-10 * np.abs(F).max())


def test_morse_pair():
    ph = util.replace_physics(De=0.7, r0=0.4, beta=1.5, mcut=1.5, h=1.5)
    V, F = casegen.icosphere(0)
    # two small cells close to each other; vertex 0 of each faces the other
    case = _cell_case(0.3 * V, F, ph, n_cells=2, shift=np.array([1.2, 0.0, 0.0]))
    o = orc.Oracle(case)
    pos = case.coor.astype(np.float64)
    Fm = o.agg_force(pos)
    assert np.abs(Fm.sum(0)).max() < 1e-12 * np.abs(Fm).sum()
    eps = 1e-6
    for v in range(0, 24, 5):
        g = np.zeros(3)
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[v, d] += eps
            pm[v, d] -= eps
            g[d] = (o.morse_energy(pp) - o.morse_energy(pm)) / (2 * eps)
        assert np.allclose(Fm[v], -g, rtol=1e-6, atol=1e-9)
    # only inter-cell pairs act: one cell alone feels nothing
    case1 = _cell_case(0.3 * V, F, ph)
    o1 = orc.Oracle(case1)
    assert np.abs(o1.agg_force(case1.coor.astype(np.float64))).max() == 0.0
    # zero force at r0 for an isolated pair
    two = np.array([[5.0, 5, 5], [5.4, 5, 5]])
    pair = util.mini_case(two, nm=2, L=(10, 10, 10), physics=ph)
    pair.celltab = np.array([0, 1], np.int32)
    pair.beginendtri = np.zeros((2, 2), np.int32)
    o2 = orc.Oracle(pair)
    # r and r0 both pass through float32 (|r - r0| < 1e-7): |F| << 2 beta De
    assert np.abs(o2.agg_force(two.astype(np.float32).astype(np.float64))).max() < 1e-6
    two[1, 0] = 5.6
    assert o2.agg_force(two)[0, 0] > 0  # attractive beyond r0: pulled toward +x
