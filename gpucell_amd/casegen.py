"""Seeded synthetic cases in the reference's data model.

A case holds what the reference's four loaders produce (Particle::LoadParInfo,
src/Particle.cu:29-427): particle coordinates in load order
(Bound1 | Bound2 | fluid | membrane), membrane triangles and edge records
with 1-based ids offset per cell, per-edge and per-triangle moduli, the
bounding box and the 96-byte parameter block.

Geometry:
  * fluid: simple-cubic lattice of spacing dx with uniform jitter;
  * channel walls: the same lattice continued 3 layers past the fluid
    region (frozen Bound2 particles);
  * RBC: icosahedron subdivided 3x (642 vertices, 1280 triangles,
    1920 edges) mapped onto the Evans & Fung (1972) biconcave profile
    z = +-D0 sqrt(1 - 4 r^2/D0^2) (a0 + a1 r^2/D0^2 + a2 r^4/D0^4),
    a0 = 0.0518, a1 = 2.0026, a2 = -4.491;
  * RBC placement: random centres and orientations from the seed,
    rejected when any vertex comes closer than `gap` to another cell's or
    to a wall;
  * fluid particles closer than `clear` to a membrane vertex are removed.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import abi


# ---------------------------------------------------------------------------
# membrane mesh
def icosphere(subdiv: int = 3):
    """Unit icosphere with outward-oriented triangles (0-based)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts)
    F = np.array(faces, dtype=np.int64)
    # orient outward
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = np.einsum("ij,ij->i", n, V[F].mean(axis=1)) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    # Mesh neighbours get nearby numbers (reverse Cuthill-McKee on the vertex
    # graph; bandwidth 612 -> 41 at subdiv 3) and triangles are ordered by
    # their smallest vertex, so the edges and triangles of a run of
    # consecutive vertices lie close together in their arrays.
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    nv = V.shape[0]
    pairs = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    A = coo_matrix((np.ones(pairs.shape[0]), (pairs[:, 0], pairs[:, 1])), shape=(nv, nv)).tocsr()
    perm = np.asarray(reverse_cuthill_mckee(A, symmetric_mode=False), dtype=np.int64)
    inv = np.empty(nv, dtype=np.int64)
    inv[perm] = np.arange(nv)
    V, F = V[perm], inv[F]
    F = F[np.argsort(F.min(axis=1), kind="stable")]
    return V, F


def edge_records(F: np.ndarray) -> np.ndarray:
    """(a, b, c, d, t1, t2) per edge, 0-based: (a, b) shared, c the third
    vertex of t1 and d the third vertex of t2 (Particle.cuh:35)."""
    owner = {}
    for t, (a, b, c) in enumerate(F):
        for (p, q, r) in ((a, b, c), (b, c, a), (c, a, b)):
            owner.setdefault((min(p, q), max(p, q)), []).append((t, r))
    rec = []
    for (p, q), lst in sorted(owner.items()):
        assert len(lst) == 2, "mesh must be closed and manifold"
        (t1, c), (t2, d) = lst
        rec.append((p, q, c, d, t1, t2))
    return np.array(rec, dtype=np.int64)


def biconcave(V: np.ndarray, D0: float) -> np.ndarray:
    """Maps the unit sphere onto the Evans-Fung profile of diameter D0."""
    a0, a1, a2 = 0.0518, 2.0026, -4.491
    rs2 = V[:, 0] ** 2 + V[:, 1] ** 2
    u = rs2 / 4.0  # (rho / D0)^2 with rho = rs * D0 / 2
    out = np.empty_like(V)
    out[:, 0] = 0.5 * D0 * V[:, 0]
    out[:, 1] = 0.5 * D0 * V[:, 1]
    out[:, 2] = D0 * V[:, 2] * (a0 + a1 * u + a2 * u * u)
    return out


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Case:
    name: str
    coor: np.ndarray            # [N,3] float32, load order
    vel: np.ndarray             # [N,3] float32
    nBound1: int
    nBound2: int
    nFluid: int
    bbox: np.ndarray            # [6] float32
    consts: abi.Structconst
    ext: abi.gc_ext_consts
    # membrane (1-based ids as the reference's loader stores them)
    tri: np.ndarray             # [T,3] int32
    edge: np.ndarray            # [E,6] int32
    celltab: np.ndarray         # [Nm] int32
    beginendtri: np.ndarray     # [C,2] int32 inclusive
    shearmod: np.ndarray        # [E] float32
    bendmod: np.ndarray         # [E] float32
    hlarea: np.ndarray          # [T]
    hgarea: np.ndarray          # [T]
    hvol: np.ndarray            # [T]

    @property
    def nCell(self) -> int:
        return int(self.beginendtri.shape[0])

    @property
    def nCellPar(self) -> int:
        return int(self.celltab.shape[0])

    @property
    def N(self) -> int:
        return int(self.coor.shape[0])

    def desc(self) -> "tuple[abi.gc_case_desc, list]":
        """gc_case_desc pointing at this case's arrays (keep the list alive)."""
        keep = []

        def p(a, dt):
            if a is None or a.size == 0:
                return None
            b = np.ascontiguousarray(a, dtype=dt)
            keep.append(b)
            return b.ctypes.data

        d = abi.gc_case_desc()
        d.nBound1Par, d.nBound2Par, d.nFluidPar = self.nBound1, self.nBound2, self.nFluid
        d.nCell, d.nAllTri, d.nAllEdge = self.nCell, int(self.tri.shape[0]), int(self.edge.shape[0])
        d.nCellPar = self.nCellPar
        d.coor = p(self.coor, np.float32)
        d.vel = p(self.vel, np.float32)
        d.tri = p(self.tri, np.int32)
        d.edge = p(self.edge, np.int32)
        d.celltab = p(self.celltab, np.int32)
        d.beginendtri = p(self.beginendtri, np.int32)
        d.shearmod_para = p(self.shearmod, np.float32)
        d.bendingmod_para = p(self.bendmod, np.float32)
        d.hlarea_para = p(self.hlarea, np.float32)
        d.hgarea_para = p(self.hgarea, np.float32)
        d.hvol_para = p(self.hvol, np.float32)
        for i in range(6):
            d.bbox[i] = float(self.bbox[i])
        d.consts = self.consts
        d.ext = self.ext
        return d, keep


@dataclasses.dataclass
class Physics:
    h: float = 1.5
    mass: float = 1.0
    rho0: float = 1.0
    c1: float = 4.0
    c2: float = 4.0
    eta: float = 1.0
    zeta: float = 0.0
    kT: float = 0.01
    dt: float = 0.005
    fext: float = 0.02
    lam: float = 0.65
    seed: int = 1
    # membrane moduli (per RBC type)
    mu0: float = 1.0
    kc: float = 0.1
    ka: float = 20.0
    kd: float = 5.0
    kv: float = 50.0
    # Morse
    De: float = 0.0
    r0: float = 0.3
    beta: float = 1.5
    mcut: float = 1.5


def make_consts(ph: Physics, nb1: int, nb2: int, nfl: int, nm: int) -> abi.Structconst:
    c = abi.Structconst()
    c.nBound2Start = nb1
    c.nFluidStart = nb1 + nb2
    c.nCellStart = nb1 + nb2 + nfl
    c.nTotalPar = nb1 + nb2 + nfl + nm
    c.nCellPar = nm
    c.iSeed = ph.seed
    c.CutRadius = ph.h
    c.Mass = ph.mass
    c.SoundSpeed1 = ph.c1
    c.SoundSpeed2 = ph.c2
    c.DenInit = ph.rho0
    c.VVsigma = ph.lam
    c.ShearVis = ph.eta
    c.BulkVis = ph.zeta
    c.ConstExtForce = ph.fext * ph.mass
    c.TimeStep = ph.dt
    c.surEnergy = ph.De
    c.MaxCut = ph.mcut
    c.zeroLen = ph.r0
    c.scaFac = ph.beta
    c.Temp = ph.kT
    return c


def _lattice(lo, hi, dx):
    axes = [np.arange(lo[d] + 0.5 * dx, hi[d], dx) for d in range(3)]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
    return g


def _axis_rotation(axis, angle):
    """Rotation matrix by `angle` about the unit vector `axis` (Rodrigues)."""
    a = np.asarray(axis, dtype=np.float64)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1.0 - np.cos(angle)) * (K @ K)


def _place_cells(rng, n_cells, shape, inside, gap, lo, hi, max_tries=200000, period=None,
                 dense=False):
    """Random sequential placement of rigid copies of `shape` (Nv,3): a
    candidate (uniform centre, uniform random rotation) is kept when it lies
    inside the domain and no vertex comes within `gap` of a placed cell.
    `period` (3 floats or None per axis) makes the distance test periodic;
    coordinates stay unwrapped (as the reference's loader keeps them)."""
    from scipy.spatial import cKDTree
    big = 1e6
    box = np.array([p if p else big for p in (period or (None, None, None))], dtype=np.float64)
    shift = np.where(box == big, big / 2, 0.0)

    def wrap(X):
        return np.mod(X + shift, box)

    placed = []
    # one periodic KD tree per placed cell plus its centre and radius: a
    # candidate is tested only against cells whose centres lie within the sum
    # of the radii plus `gap` (periodic distance; by the triangle inequality
    # no other cell can come within `gap`), the same test as one tree over
    # every placed vertex at a cost independent of the number placed
    trees, centres, radii = [], np.zeros((0, 3)), np.zeros(0)
    per_axes = np.array([p is not None and p != 0 for p in (period or (None, None, None))])
    tries = 0
    # Random candidates for up to `max_tries`; then (or from the start with
    # `dense`) a lattice of slots sized to the cell's extent, with the discs
    # nearly aligned (tilt <= 0.15 rad, random spin): random sequential
    # addition of randomly oriented discs jams below ~25 % volume fraction.
    random_tries = 0 if dense else max_tries
    # lattice discs lie with their axis along x (the flow axis of a channel)
    R0 = _axis_rotation(np.array([0.0, 1.0, 0.0]), 0.5 * np.pi)
    ext = (shape @ R0.T).max(axis=0) - (shape @ R0.T).min(axis=0)
    slots = None

    def candidate():
        nonlocal slots
        if tries <= random_tries:
            return lo + rng.random(3) * (hi - lo), random_rotation(rng)
        if slots is None:
            # half-extent steps across the disc (neighbouring slots overlap;
            # the distance test keeps a dense non-overlapping subset)
            step = [ext[0] + gap, 0.5 * (ext[1] + gap), 0.5 * (ext[2] + gap)]
            axes = [np.arange(lo[d] + 0.5 * step[d], hi[d], step[d]) for d in range(3)]
            # raster order (x slowest): greedy acceptance then packs the
            # cross-section regularly, which random order does not
            slots = list(np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)[::-1])
        if not slots:
            return None, None
        spin = rng.uniform(0.0, 2.0 * np.pi)
        axis = rng.normal(size=3)
        axis[0] = 0.0
        axis /= max(np.linalg.norm(axis), 1e-12)
        tilt = rng.uniform(0.0, 0.15)
        return slots.pop(), (_axis_rotation(axis, tilt) @ _axis_rotation(np.array([1.0, 0.0, 0.0]), spin)
                             @ R0)

    while len(placed) < n_cells:
        tries += 1
        if tries > max_tries + random_tries:
            raise RuntimeError(f"could place only {len(placed)} of {n_cells} cells")
        c, R = candidate()
        if c is None:
            raise RuntimeError(f"could place only {len(placed)} of {n_cells} cells")
        X = shape @ R.T + c
        if not inside(X):
            continue
        c = X.mean(axis=0)
        rad = float(np.sqrt(((X - c) ** 2).sum(axis=1).max()))
        if centres.shape[0]:
            dc = c - centres
            dc[:, per_axes] -= box[per_axes] * np.round(dc[:, per_axes] / box[per_axes])
            near = np.nonzero(np.sqrt((dc ** 2).sum(axis=1)) < radii + rad + gap)[0]
            WX = wrap(X)
            hit = False
            for k in near:
                d, _ = trees[k].query(WX, k=1, distance_upper_bound=gap)
                if np.isfinite(d).any():
                    hit = True
                    break
            if hit:
                continue
        placed.append(X)
        trees.append(cKDTree(wrap(X), boxsize=box))
        centres = np.vstack([centres, c[None, :]])
        radii = np.append(radii, rad)
    return placed


def _remove_near(points: np.ndarray, verts: np.ndarray, dist: float) -> np.ndarray:
    """Mask of points farther than `dist` from every vertex."""
    if verts.shape[0] == 0:
        return np.ones(points.shape[0], dtype=bool)
    from scipy.spatial import cKDTree
    d, _ = cKDTree(verts).query(points, k=1, distance_upper_bound=dist)
    return ~np.isfinite(d)


# Fluid sites closer than this (in lattice spacings) to a membrane vertex are
# removed.  The membrane vertices count in every neighbour's number density,
# so the fluid next to a membrane must thin out by about as many particles
# as the vertices add: with 0.5 the fluid inside a 642-vertex cell of
# diameter 16-20 starts ~10 % denser than outside and the cell inflates by
# several per cent over the first ~100 steps (FP64 oracle runs of the 32^3
# box); with 0.7 the inside and outside densities agree to ~1 % and the
# volume stays within ~1 % of its initial value.
MEMBRANE_CLEARANCE = 0.7


def _assemble(name, ph, ext, bbox, bound1, bound2, fluid, cells, tri0, edge0, rng, jitter, dx):
    nb1, nb2 = bound1.shape[0], bound2.shape[0]
    if jitter > 0:
        fluid = fluid + rng.uniform(-jitter * dx, jitter * dx, size=fluid.shape)
    mverts = np.concatenate(cells, axis=0) if cells else np.zeros((0, 3))
    if cells:
        keep = _remove_near(fluid, mverts, MEMBRANE_CLEARANCE * dx)
        fluid = fluid[keep]
    nfl = fluid.shape[0]
    nv = cells[0].shape[0] if cells else 0
    nm = mverts.shape[0]
    coor = np.concatenate([bound1, bound2, fluid, mverts], axis=0).astype(np.float32)
    N = coor.shape[0]
    vel = np.zeros((N, 3), dtype=np.float64)
    if ph.kT > 0:
        mv = slice(nb1 + nb2, N)
        vel[mv] = rng.normal(0.0, math.sqrt(ph.kT / ph.mass), size=(N - nb1 - nb2, 3))
        vel[mv] -= vel[mv].mean(axis=0)
    ncell = len(cells)
    T, E = tri0.shape[0], edge0.shape[0]
    tri = np.zeros((ncell * T, 3), dtype=np.int32)
    edge = np.zeros((ncell * E, 6), dtype=np.int32)
    bet = np.zeros((ncell, 2), dtype=np.int32)
    for c in range(ncell):
        tri[c * T:(c + 1) * T] = tri0 + 1 + c * nv
        e = edge0.copy()
        e[:, :4] += 1 + c * nv
        e[:, 4:] += 1 + c * T
        edge[c * E:(c + 1) * E] = e
        bet[c] = (c * T, (c + 1) * T - 1)
    celltab = np.repeat(np.arange(ncell, dtype=np.int32), nv)
    consts = make_consts(ph, nb1, nb2, nfl, nm)
    return Case(name=name, coor=coor, vel=vel.astype(np.float32), nBound1=nb1, nBound2=nb2,
                nFluid=nfl, bbox=np.asarray(bbox, dtype=np.float32), consts=consts, ext=ext,
                tri=tri, edge=edge, celltab=celltab, beginendtri=bet,
                shearmod=np.full(ncell * E, ph.mu0, np.float32),
                bendmod=np.full(ncell * E, ph.kc, np.float32),
                hlarea=np.full(ncell * T, ph.kd, np.float32),
                hgarea=np.full(ncell * T, ph.ka, np.float32),
                hvol=np.full(ncell * T, ph.kv, np.float32))


def periodic_box(n: int = 32, n_rbc: int = 1, rbc_diameter: float = 16.0, seed: int = 20261203,
                 physics: Physics | None = None, jitter: float = 0.1, subdiv: int = 3,
                 reverse_poiseuille: bool = True, noise: bool = True) -> Case:
    """Fully periodic n^3 box of unit-spaced fluid with `n_rbc` RBCs
    (n=32, one RBC: the 32x32x32 / one-RBC reverse-Poiseuille case).  The
    default diameter 16 puts the 642 vertices of a 3x subdivided mesh about
    one lattice spacing apart; a finer membrane counts for more than the
    fluid it displaces in its neighbours' densities, and a 12-wide cell
    (spacing ~0.7) swells its area by ~30 % in the first 150 steps."""
    ph = physics or Physics()
    rng = np.random.default_rng(seed)
    L = float(n)
    bbox = np.array([0, L, 0, L, 0, L], dtype=np.float32)
    ext = abi.default_ext()
    ext.periodic[0] = ext.periodic[1] = ext.periodic[2] = 1
    ext.force_mode = 1 if reverse_poiseuille else 0
    ext.force_split = 0.5 * L
    ext.noise_on = 1 if noise else 0
    fluid = _lattice([0, 0, 0], [L, L, L], 1.0)
    cells = []
    V0, F0 = icosphere(subdiv)
    shape = biconcave(V0, rbc_diameter)
    if n_rbc == 1:
        cells = [shape + np.array([L / 2, L / 2, L / 2])]
    elif n_rbc > 1:
        cells = _place_cells(rng, n_rbc, shape, lambda X: True, 1.0,
                             np.zeros(3), np.full(3, L))
    E0 = edge_records(F0)
    return _assemble(f"periodic{n}", ph, ext, bbox, np.zeros((0, 3)), np.zeros((0, 3)), fluid,
                     cells, F0, E0, rng, jitter, 1.0)


def couette(n: int = 24, n_rbc: int = 0, rbc_diameter: float = 8.0, shear_rate: float = 0.005,
            seed: int = 20261203, physics: Physics | None = None, jitter: float = 0.1, subdiv: int = 2,
            wall_layers: int = 3, noise: bool = True) -> Case:
    """Plane Couette shear: an n^3 box of unit-spaced fluid, periodic in x
    and z, between `wall_layers` frozen Bound2 layers below y = 0 and above
    y = n that move along x at -U and +U (U = shear_rate n / 2; positions
    stay fixed, the wall velocity enters through the no-slip image), no body
    force, `n_rbc` RBCs placed at random."""
    ph = dataclasses.replace(physics or Physics(), fext=0.0)
    rng = np.random.default_rng(seed)
    L = float(n)
    bbox = np.array([0, L, 0, L, 0, L], dtype=np.float32)
    ext = abi.default_ext()
    ext.periodic[0], ext.periodic[1], ext.periodic[2] = 1, 0, 1
    ext.force_mode = 0
    ext.noise_on = 1 if noise else 0
    U = 0.5 * shear_rate * L
    ext.wall_split, ext.wall_u_lo, ext.wall_u_hi = 0.5 * L, -U, U
    nl = wall_layers
    pts = _lattice([0, -nl, 0], [L, L + nl, L], 1.0)
    fluid = pts[(pts[:, 1] >= 0) & (pts[:, 1] < L)]
    walls = pts[(pts[:, 1] < 0) | (pts[:, 1] >= L)]
    V0, F0 = icosphere(subdiv)
    shape = biconcave(V0, rbc_diameter)
    gap = 1.0

    def inside(X):
        return bool(np.all((X[:, 1] > gap) & (X[:, 1] < L - 1 - gap)))

    cells = _place_cells(rng, n_rbc, shape, inside, gap, np.zeros(3), np.full(3, L),
                         period=(L, None, L)) if n_rbc else []
    cells = [X - L * np.floor(X.mean(0) / L) * np.array([1.0, 0, 1.0]) for X in cells]
    E0 = edge_records(F0)
    case = _assemble(f"couette{n}", ph, ext, bbox, np.zeros((0, 3)), walls, fluid, cells, F0, E0, rng,
                     jitter, 1.0)
    case.shear_rate = shear_rate
    return case


def channel(length: int, width: int, n_rbc: int, hematocrit: float | None = None,
            rbc_diameter: float | None = None, seed: int = 20261203,
            physics: Physics | None = None, jitter: float = 0.1, wall_layers: int = 3,
            subdiv: int = 3, noise: bool = True, stenosis: float = 0.0,
            name: str = "channel", dense: bool = False) -> Case:
    """Square duct, periodic along x, `wall_layers` frozen Bound2 layers.

    With `hematocrit` the RBC diameter is chosen so that n_rbc cells fill
    that fraction of the fluid volume.  `stenosis` in [0, 1) narrows the
    duct with a cosine constriction: the wall sits where the signed distance
    max(|y|, |z|) - w(x) changes sign, w(x) = W/2 (1 - s/2 (1 + cos(2 pi
    (x - L/2) / L))) within the middle half of the length.  `dense` places
    the cells on aligned lattice slots instead of at random."""
    ph = physics or Physics()
    rng = np.random.default_rng(seed)
    W, Lx = float(width), float(length)
    half = 0.5 * W
    bbox = np.array([0, Lx, -half, half, -half, half], dtype=np.float32)
    ext = abi.default_ext()
    ext.noise_on = 1 if noise else 0
    nl = wall_layers
    pts = _lattice([0, -half - nl, -half - nl], [Lx, half + nl, half + nl], 1.0)

    def wall_half(x):
        if stenosis <= 0:
            return np.full_like(x, half)
        w = np.full_like(x, half)
        m = np.abs(x - Lx / 2) < Lx / 4
        w[m] = half * (1 - 0.5 * stenosis * (1 + np.cos(2 * np.pi * (x[m] - Lx / 2) / (Lx / 2))))
        return w

    sd = np.maximum(np.abs(pts[:, 1]), np.abs(pts[:, 2])) - wall_half(pts[:, 0])
    fluid = pts[sd < 0]
    walls = pts[(sd >= 0) & (sd < nl)]
    V0, F0 = icosphere(subdiv)
    if rbc_diameter is None:
        if hematocrit is None or n_rbc == 0:
            rbc_diameter = 8.0
        else:
            unit = biconcave(V0, 1.0)
            vol_unit = _mesh_volume(unit, F0)
            rbc_diameter = (hematocrit * fluid.shape[0] / (n_rbc * vol_unit)) ** (1.0 / 3.0)
    shape = biconcave(V0, rbc_diameter)
    gap = 0.8

    def inside(X):
        w = wall_half(np.mod(X[:, 0], Lx))
        return bool(np.all(np.maximum(np.abs(X[:, 1]), np.abs(X[:, 2])) < w - gap))

    cells = _place_cells(rng, n_rbc, shape, inside, gap, np.array([0, -half, -half]),
                         np.array([Lx, half, half]), period=(Lx, None, None),
                         dense=dense) if n_rbc else []
    # keep each cell whole but put its centre in [0, Lx): vertices may lie
    # outside the box in x; the first step wraps them (as the reference)
    cells = [X - Lx * np.floor(X.mean(0)[0] / Lx) * np.array([1.0, 0, 0]) for X in cells]
    E0 = edge_records(F0)
    case = _assemble(name, ph, ext, bbox, np.zeros((0, 3)), walls, fluid, cells, F0, E0, rng,
                     jitter, 1.0)
    case.rbc_diameter = rbc_diameter
    return case


def _mesh_volume(V, F):
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def mesh_volume(V, F):
    return _mesh_volume(V, F)


def suspension_1m(seed: int = 20261203, physics: Physics | None = None, noise: bool = True) -> Case:
    """1M fluid+wall particles (1,015,188 of 1,097,712 lattice sites remain
    around the cells), 100 RBCs at 20 % hematocrit (diameter 21.0), 3 wall
    layers: a 252 x 60 x 60 duct (66 x 66 with walls)."""
    c = channel(252, 60, 100, hematocrit=0.20, seed=seed, physics=physics, noise=noise,
                name="suspension_1m_100rbc")
    c.hematocrit = 0.20
    return c


def stenosis_2m(seed: int = 20261203, physics: Physics | None = None, noise: bool = True) -> Case:
    """2M fluid+wall particles (2,062,828) and 300 RBCs, 50 % cosine
    stenosis, Morse aggregation De = 0.3 kT, r0 = 0.3, beta = 1.5: a
    380 x 80 x 80 duct (86 x 86 with walls).  The hematocrit is the highest
    of 30 %, 27 %, ... at which the discs fit on aligned lattice slots
    (15 % with the default seed: the throat limits the slots)."""
    ph = physics or Physics()
    ph = dataclasses.replace(ph, De=0.3 * ph.kT, r0=0.3, beta=1.5, mcut=min(1.5, ph.h))
    # 300 rigid biconcave discs pack (with a 0.8 gap, throat included) only
    # below ~30 % in this duct: take the highest hematocrit that places them
    last = None
    for hct in (0.30, 0.27, 0.24, 0.21, 0.18, 0.15, 0.12):
        try:
            c = channel(380, 80, 300, hematocrit=hct, seed=seed, physics=ph, noise=noise,
                        stenosis=0.5, name="stenosis_2m_300rbc", dense=True)
            c.hematocrit = hct
            return c
        except RuntimeError as e:
            last = e
    raise last


def suspension_10m(seed: int = 20261203, physics: Physics | None = None, noise: bool = True,
                   n_rbc: int = 2000, hematocrit: float = 0.18) -> Case:
    """10M fluid+wall particles (10,109,183 of 11,486,592 lattice sites
    remain around the cells) and 2000 RBCs x 642 vertices at 18 %
    hematocrit (diameter 17.0; random placement of 2000 discs jams near
    19 % in this duct), 3 wall layers: a 472 x 150 x 150 duct (156 x 156
    with walls), Poiseuille body force, aggregation off."""
    c = channel(472, 150, n_rbc, hematocrit=hematocrit, seed=seed, physics=physics,
                noise=noise, name=f"suspension_10m_{n_rbc}rbc")
    c.hematocrit = hematocrit
    return c


def ensemble64(seed: int = 20261203, nrep: int = 64, nmin: int = 20000, nmax: int = 400000,
               rbc_max: int = 40, subdiv: int | None = None, shear: bool = True) -> "list[Case]":
    """Independent suspensions of uneven size for a batched run: particle
    counts log-uniform in [nmin, nmax] (cubic cells, unit lattice), 1..rbc_max
    RBCs each (log-uniform), about 10 % hematocrit, and per replica a shear
    rate (log-uniform in [0.002, 0.01]: plane Couette cells between walls
    moving at -U and +U, wall Mach number <= 0.1), Morse D_e (0 or 0.3 kT)
    and noise seed drawn from `seed`; with shear=False fully periodic boxes
    driven by a reverse-Poiseuille body force instead.  The membrane mesh is
    refined with the cell size so that vertices stay about one lattice
    spacing apart (a denser membrane than fluid drives the SDPD pressure
    unstable): 642 vertices from D = 16, 162 from D = 8; D >= 8 caps the cell
    count of small boxes at 20 % hematocrit."""
    rng = np.random.default_rng(seed)
    base = Physics()
    cases = []
    vol = 0.195   # biconcave mesh volume / D^3
    for r in range(nrep):
        N = float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        n = max(16, int(round(N ** (1.0 / 3.0))))
        n_rbc = int(np.clip(round(float(np.exp(rng.uniform(0.0, np.log(rbc_max))))), 1, rbc_max))
        n_rbc = max(1, min(n_rbc, int(0.20 * n ** 3 / (vol * 8.0 ** 3))))
        d = min(0.45 * n, max(8.0, (0.10 * n ** 3 / (n_rbc * vol)) ** (1.0 / 3.0)))
        sub = 3 if d >= 16.0 else 2
        De = float(rng.choice([0.0, 0.3])) * base.kT
        fext = float(rng.uniform(0.01, 0.05))
        rate = float(np.exp(rng.uniform(np.log(0.002), np.log(0.01))))
        ph = dataclasses.replace(base, fext=fext, De=De, r0=0.3, beta=1.5, mcut=min(1.5, base.h),
                                 seed=int(rng.integers(1, 2 ** 31 - 1)))
        cseed = int(rng.integers(1, 2 ** 31 - 1))
        if shear:
            c = couette(n=n, n_rbc=n_rbc, rbc_diameter=d, shear_rate=rate, seed=cseed, physics=ph,
                        subdiv=sub if subdiv is None else subdiv)
        else:
            c = periodic_box(n=n, n_rbc=n_rbc, rbc_diameter=d, seed=cseed, physics=ph,
                             subdiv=sub if subdiv is None else subdiv)
        c.name = f"ensemble_replica_{r}"
        cases.append(c)
    return cases
