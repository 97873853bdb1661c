"""Thermal-run statistics shared by the oracle fixture generator
(tests/golden/make_thermal.py) and the GPU comparison
(tests/test_gpu_thermal.py): per-cell deformation index, the binned
streamwise velocity profile of the fluid and the inter-cell membrane-vertex
RDF (the per-cell series of Particle::OutPutCellGeo, src/Particle.cu:1010-1246,
and the flow diagnostics of a reverse-Poiseuille box).

The case is a 32^3 periodic box (the size of the one-RBC benchmark box) at
kT = 0.01 with reverse-Poiseuille forcing (+-f in the two halves in y) and
eight 162-vertex cells of diameter 10 (mesh spacing ~1.25 lattice units),
so that the RDF has pairs of cells to count."""
from __future__ import annotations

import hashlib

import numpy as np

from gpucell_amd import analysis, casegen
from tests import util

STEPS = 2000           # steps of each run
SAMPLE_FROM = 1000     # first sampled step
SAMPLE_EVERY = 20      # sampling interval (51 samples)
VEL_BINS = 8           # slabs in y for the x-velocity profile
RDF_RMAX, RDF_BINS = 8.0, 16


def thermal_case():
    ph = util.replace_physics(kT=0.01)
    return casegen.periodic_box(n=32, n_rbc=8, rbc_diameter=10.0, subdiv=2, seed=424242, physics=ph)


def fingerprint(case) -> dict:
    """Identity of the case a fixture was made from."""
    h = hashlib.sha256(np.ascontiguousarray(case.coor, np.float32).tobytes())
    h.update(np.ascontiguousarray(case.vel, np.float32).tobytes())
    return {"N": int(case.N), "nCell": int(case.nCell), "nCellPar": int(case.nCellPar),
            "sha256": h.hexdigest()[:32], "dt": float(case.consts.TimeStep),
            "kT": float(case.consts.Temp), "h": float(case.consts.CutRadius)}


def sample_steps():
    return list(range(SAMPLE_FROM, STEPS + 1, SAMPLE_EVERY))


def stats(state: dict, case) -> dict:
    """Statistics of one state (engine layout: pos FP64 [N,3] in the box,
    nthrobound [N,3] image counters, velrhop [N,4], all in id order)."""
    L = np.array([case.bbox[1] - case.bbox[0], case.bbox[3] - case.bbox[2],
                  case.bbox[5] - case.bbox[4]], dtype=np.float64)
    lo = np.array([case.bbox[0], case.bbox[2], case.bbox[4]], dtype=np.float64)
    cells = analysis.unwrap_cells(state, case)
    di = [analysis.deformation_index(cells[c], analysis.cell_tri0(case, c)) for c in range(case.nCell)]
    f0, f1 = int(case.consts.nFluidStart), int(case.consts.nCellStart)
    pos = lo + np.mod(state["pos"][f0:f1] - lo, L)
    vel = state["velrhop"][f0:f1, :3].astype(np.float64)
    _, prof, _ = analysis.velocity_profile(pos, vel, axis=1, comp=0, bins=VEL_BINS,
                                           lo=float(case.bbox[2]), hi=float(case.bbox[3]))
    mpos = np.concatenate(cells, axis=0)
    _, g = analysis.membrane_rdf(mpos, case.celltab, L, [1, 1, 1], RDF_RMAX, bins=RDF_BINS)
    return {"di": [float(x) for x in di], "vel": [float(x) for x in prof], "rdf": [float(x) for x in g]}


def block_mean_err(series: np.ndarray, blocks: int = 5):
    """Mean over samples (axis 0) and its standard error from `blocks`
    consecutive block means (the samples are correlated in time)."""
    s = np.asarray(series, dtype=np.float64)
    n = (s.shape[0] // blocks) * blocks
    bm = s[:n].reshape(blocks, n // blocks, *s.shape[1:]).mean(axis=1)
    return s.mean(axis=0), bm.std(axis=0, ddof=1) / np.sqrt(blocks)
