"""Analysis of engine states: the quantities the reference writes per cell
(Particle::OutPutCellGeo / OutPutCellForce, src/Particle.cu:1010-1246) and
the flow diagnostics used to check thermal runs.

* deformation_index: Taylor D = (L - B)/(L + B) of the ellipsoid with the
  same volume second moments as the closed membrane (exact for a closed
  triangulated surface: sum of signed tetrahedra to the centroid);
* cell_series: per cell |mean vertex velocity|, area, volume, mean
  membrane and Morse force (Geo_Cell<i>.plt / TF_Cell<i>.plt content);
* velocity_profile: binned mean velocity along a transverse coordinate;
* membrane_rdf: radial distribution of membrane vertices of different
  cells (periodic minimum image on periodic axes).
"""
from __future__ import annotations

import numpy as np


def _tet_second_moments(a, b, c):
    """sum over tetrahedra (0, a, b, c) of the integral of x x^T dV
    (closed form: V/20 (sum_i p_i p_i^T + (sum_i p_i)(sum_i p_i)^T) with the
    fourth vertex at the origin)."""
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0
    s = a + b + c
    M = (np.einsum("i,ij,ik->jk", vol, a, a) + np.einsum("i,ij,ik->jk", vol, b, b) +
         np.einsum("i,ij,ik->jk", vol, c, c) + np.einsum("i,ij,ik->jk", vol, s, s)) / 20.0
    return vol.sum(), M


def ellipsoid_axes(verts: np.ndarray, tri0: np.ndarray) -> np.ndarray:
    """Semi-axes (descending) of the solid ellipsoid with the membrane's
    volume and volume second moments; tri0 is 0-based and outward."""
    x = verts - verts.mean(axis=0)
    a, b, c = x[tri0[:, 0]], x[tri0[:, 1]], x[tri0[:, 2]]
    vol = np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0
    V = vol.sum()
    cen = (np.einsum("i,ij->j", vol, a + b + c) / 4.0) / V
    a, b, c = a - cen, b - cen, c - cen
    V, M = _tet_second_moments(a, b, c)
    lam = np.linalg.eigvalsh(M / V)[::-1]
    # a solid ellipsoid with semi-axes s_k has <x_k^2> = s_k^2 / 5
    return np.sqrt(5.0 * np.maximum(lam, 0.0))


def deformation_index(verts: np.ndarray, tri0: np.ndarray) -> float:
    s = ellipsoid_axes(verts, tri0)
    return float((s[0] - s[-1]) / (s[0] + s[-1]))


def unwrap_cells(state: dict, case) -> list[np.ndarray]:
    """Unwrapped vertex positions of each cell from an engine state."""
    nc = case.N - case.nCellPar
    L = np.array([case.bbox[1] - case.bbox[0], case.bbox[3] - case.bbox[2],
                  case.bbox[5] - case.bbox[4]], dtype=np.float64)
    per = np.array(list(case.ext.periodic), dtype=np.float64)
    pos = state["pos"][nc:] + state["nthrobound"][nc:] * (L * per)
    out = []
    for c in range(case.nCell):
        out.append(pos[case.celltab == c])
    return out


def cell_tri0(case, c: int) -> np.ndarray:
    """0-based, cell-local triangles of cell c."""
    t0, t1 = case.beginendtri[c]
    vb = int(np.argmax(case.celltab == c))
    return case.tri[t0:t1 + 1] - 1 - vb


def cell_series(state: dict, case) -> dict:
    """Per-cell rows of Geo_Cell<i>.plt and TF_Cell<i>.plt."""
    nc = case.N - case.nCellPar
    v = state["velrhop"][nc:, :3].astype(np.float64)
    out = {"vel_mag": [], "area": state.get("area"), "vol": state.get("vol"),
           "mean_def": [], "mean_agg": [], "DI": []}
    cells = unwrap_cells(state, case)
    for c in range(case.nCell):
        m = case.celltab == c
        out["vel_mag"].append(float(np.linalg.norm(v[m].mean(axis=0))))
        if "defforce" in state:
            out["mean_def"].append(state["defforce"][m].astype(np.float64).mean(axis=0))
            out["mean_agg"].append(state["aggforce"][m].astype(np.float64).mean(axis=0))
        out["DI"].append(deformation_index(cells[c], cell_tri0(case, c)))
    return out


def velocity_profile(pos: np.ndarray, vel: np.ndarray, axis: int = 1, comp: int = 0,
                     bins: int = 20, lo: float | None = None, hi: float | None = None):
    """Mean velocity component `comp` in `bins` slabs along `axis`."""
    x = pos[:, axis]
    lo = float(x.min()) if lo is None else lo
    hi = float(x.max()) if hi is None else hi
    edges = np.linspace(lo, hi, bins + 1)
    idx = np.clip(np.digitize(x, edges) - 1, 0, bins - 1)
    s = np.bincount(idx, weights=vel[:, comp], minlength=bins)
    n = np.bincount(idx, minlength=bins)
    centres = 0.5 * (edges[1:] + edges[:-1])
    with np.errstate(invalid="ignore", divide="ignore"):
        return centres, s / n, n


def membrane_rdf(pos: np.ndarray, cell_of: np.ndarray, L, periodic, rmax: float,
                 bins: int = 50):
    """g(r) of vertex pairs in different cells, normalised by the ideal-gas
    pair count of the same number density in the box volume."""
    from scipy.spatial import cKDTree
    L = np.asarray(L, dtype=np.float64)
    per = np.asarray(periodic, dtype=bool)
    box = np.where(per, L, 0.0)
    if per.all():
        tree = cKDTree(np.mod(pos, L), boxsize=L)
        pairs = tree.query_pairs(rmax, output_type="ndarray")
        d = pos[pairs[:, 0]] - pos[pairs[:, 1]]
        d -= L * np.round(d / L)
    else:
        tree = cKDTree(pos)
        pairs = tree.query_pairs(rmax, output_type="ndarray")
        d = pos[pairs[:, 0]] - pos[pairs[:, 1]]
        d -= box * np.round(d / np.where(box > 0, box, 1.0))
    keep = cell_of[pairs[:, 0]] != cell_of[pairs[:, 1]]
    r = np.linalg.norm(d[keep], axis=1)
    edges = np.linspace(0.0, rmax, bins + 1)
    h, _ = np.histogram(r, bins=edges)
    n = pos.shape[0]
    V = float(np.prod(L))
    shell = 4.0 / 3.0 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    # expected pairs of distinct cells at uniform density
    same = np.bincount(cell_of).astype(np.float64)
    npairs = 0.5 * (n * n - (same * same).sum())
    ideal = npairs * shell / V
    with np.errstate(invalid="ignore", divide="ignore"):
        return 0.5 * (edges[1:] + edges[:-1]), h / ideal
