"""Shared helpers for the tests: small hand-made cases and error metrics."""
from __future__ import annotations

import dataclasses

import numpy as np

from gpucell_amd import abi, casegen


def mini_case(pos, nb1=0, nb2=0, nm=0, *, L=(16.0, 16.0, 16.0), lo=(0.0, 0.0, 0.0),
              periodic=(1, 1, 1), physics=None, noise=False, vel=None, cells=None,
              force_mode=0, force_split=0.0, name="mini"):
    """Case from explicit positions (load order Bound1|Bound2|fluid|membrane).
    `cells`: optional (tri0 (T,3) 0-based, n_cells) for membranes made of
    n_cells copies of one mesh with nm/n_cells vertices each."""
    ph = physics or casegen.Physics()
    pos = np.asarray(pos, dtype=np.float32)
    N = pos.shape[0]
    nfl = N - nb1 - nb2 - nm
    ext = abi.default_ext()
    for d in range(3):
        ext.periodic[d] = int(periodic[d])
    ext.noise_on = int(noise)
    ext.force_mode = force_mode
    ext.force_split = force_split
    bbox = np.array([lo[0], lo[0] + L[0], lo[1], lo[1] + L[1], lo[2], lo[2] + L[2]], np.float32)
    consts = casegen.make_consts(ph, nb1, nb2, nfl, nm)
    if cells is not None:
        tri0, ncell = cells
        nv = nm // ncell
        E0 = casegen.edge_records(tri0)
        T, E = tri0.shape[0], E0.shape[0]
        tri = np.concatenate([tri0 + 1 + c * nv for c in range(ncell)]).astype(np.int32)
        ed = []
        for c in range(ncell):
            e = E0.copy()
            e[:, :4] += 1 + c * nv
            e[:, 4:] += 1 + c * T
            ed.append(e)
        edge = np.concatenate(ed).astype(np.int32)
        bet = np.array([(c * T, (c + 1) * T - 1) for c in range(ncell)], np.int32)
        celltab = np.repeat(np.arange(ncell, dtype=np.int32), nv)
    else:
        tri = np.zeros((0, 3), np.int32)
        edge = np.zeros((0, 6), np.int32)
        bet = np.zeros((0, 2), np.int32)
        celltab = np.zeros(0, np.int32)
        T = E = 0
    ncell = bet.shape[0]
    if vel is None:
        vel = np.zeros((N, 3), np.float32)
    return casegen.Case(name=name, coor=pos, vel=np.asarray(vel, np.float32), nBound1=nb1,
                        nBound2=nb2, nFluid=nfl, bbox=bbox, consts=consts, ext=ext, tri=tri,
                        edge=edge, celltab=celltab, beginendtri=bet,
                        shearmod=np.full(ncell * E, ph.mu0, np.float32),
                        bendmod=np.full(ncell * E, ph.kc, np.float32),
                        hlarea=np.full(ncell * T, ph.kd, np.float32),
                        hgarea=np.full(ncell * T, ph.ka, np.float32),
                        hvol=np.full(ncell * T, ph.kv, np.float32))


def random_fluid(n, L, seed=0, margin=0.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(margin, L - margin, size=(n, 3))


def b11_error(F, Fref, abs_sum, fext=0.0):
    """Per-particle error |F - F*| / (sum_j |f_ij*| + |F_ext|); returns
    (max, global relative L2)."""
    den = abs_sum + abs(fext)
    den = np.where(den > 0, den, 1.0)
    e = np.linalg.norm(F - Fref, axis=1) / den
    l2 = np.linalg.norm(F - Fref) / max(np.linalg.norm(Fref), 1e-300)
    return float(e.max()), float(l2)


def replace_physics(**kw):
    return dataclasses.replace(casegen.Physics(), **kw)


def outer_walls_as_bound1(case, physics, half_width: float):
    """Channel case whose outermost wall layer (beyond half_width + 2 from
    the axis) becomes Bound1 (repulsion only); the other wall layers stay
    Bound2.  Rebuilds the load order and the parameter block's id ranges."""
    c = dataclasses.replace(case)
    walls = c.coor[:c.nBound2]
    outer = np.maximum(np.abs(walls[:, 1]), np.abs(walls[:, 2])) > half_width + 2.0
    b1, b2 = walls[outer], walls[~outer]
    rest = c.coor[c.nBound2:]
    c.vel = np.concatenate([np.zeros((walls.shape[0], 3), np.float32), c.vel[c.nBound2:]])
    c.coor = np.concatenate([b1, b2, rest]).astype(np.float32)
    c.nBound1, c.nBound2 = int(b1.shape[0]), int(b2.shape[0])
    c.consts = casegen.make_consts(physics, c.nBound1, c.nBound2, c.nFluid, c.nCellPar)
    return c
