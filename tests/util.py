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


def inner_walls_as_bound1(case, physics, half_width: float):
    """Channel case whose innermost wall layer (the one next to the fluid,
    max(|y|, |z|) < half_width + 1) is split in a checkerboard: half of it
    becomes Bound1 (repulsion only), within h of the first fluid layer, and
    the rest of the walls stay Bound2 (some of them with fluid neighbours,
    so the no-slip wall velocity acts as well).  Rebuilds the load order
    and the parameter block's id ranges."""
    c = dataclasses.replace(case)
    walls = c.coor[:c.nBound2]
    inner = np.maximum(np.abs(walls[:, 1]), np.abs(walls[:, 2])) < half_width + 1.0
    checker = (np.floor(walls).astype(np.int64).sum(axis=1) % 2) == 0
    pick = inner & checker
    b1, b2 = walls[pick], walls[~pick]
    rest = c.coor[c.nBound2:]
    c.vel = np.concatenate([np.zeros((walls.shape[0], 3), np.float32), c.vel[c.nBound2:]])
    c.coor = np.concatenate([b1, b2, rest]).astype(np.float32)
    c.nBound1, c.nBound2 = int(b1.shape[0]), int(b2.shape[0])
    c.consts = casegen.make_consts(physics, c.nBound1, c.nBound2, c.nFluid, c.nCellPar)
    return c


def stacked_pair_channel(length: int = 24, width: int = 12, diameter: float = 5.0, subdiv: int = 2,
                         gap: float = 0.6, physics=None, noise: bool = True, seed: int = 3,
                         name: str = "stacked_pair"):
    """Square duct (periodic along x, three Bound2 wall layers) holding two
    biconcave RBCs face to face along the flow axis, their rims `gap` apart
    (the closest pair of vertices of different cells), so that a Morse
    potential with MaxCut > gap acts between many vertex pairs from the
    first step (a rouleau of two)."""
    from scipy.spatial import cKDTree
    ph = physics or casegen.Physics()
    rng = np.random.default_rng(seed)
    W, Lx = float(width), float(length)
    half = 0.5 * W
    bbox = np.array([0, Lx, -half, half, -half, half], np.float32)
    ext = abi.default_ext()
    ext.noise_on = int(noise)
    pts = casegen._lattice([0, -half - 3, -half - 3], [Lx, half + 3, half + 3], 1.0)
    sd = np.maximum(np.abs(pts[:, 1]), np.abs(pts[:, 2])) - half
    fluid, walls = pts[sd < 0], pts[(sd >= 0) & (sd < 3)]
    V0, F0 = casegen.icosphere(subdiv)
    disc = casegen.biconcave(V0, diameter)
    # axis of the disc along x
    disc = disc @ np.array([[0.0, 0.0, 1.0], [0.0, 1.0, 0.0], [-1.0, 0.0, 0.0]]).T
    tree = cKDTree(disc)
    lo, hi = 0.0, diameter
    for _ in range(60):   # centre separation s with min vertex distance = gap
        s = 0.5 * (lo + hi)
        d = tree.query(disc + np.array([s, 0.0, 0.0]), k=1)[0].min()
        lo, hi = (s, hi) if d < gap else (lo, s)
    s = 0.5 * (lo + hi)
    c0 = np.array([0.5 * Lx - 0.5 * s, 0.0, 0.0])
    cells = [disc + c0, disc + c0 + np.array([s, 0.0, 0.0])]
    E0 = casegen.edge_records(F0)
    return casegen._assemble(name, ph, ext, bbox, np.zeros((0, 3)), walls, fluid, cells, F0, E0, rng,
                             0.1, 1.0)


def morse_pairs(case, pos=None):
    """Number of membrane-vertex pairs of different cells closer than MaxCut."""
    from scipy.spatial import cKDTree
    nc = int(case.consts.nCellStart)
    X = (case.coor if pos is None else pos)[nc:].astype(np.float64)
    pairs = cKDTree(X).query_pairs(float(case.consts.MaxCut), output_type="ndarray")
    ct = case.celltab
    return int(np.count_nonzero(ct[pairs[:, 0]] != ct[pairs[:, 1]]))
