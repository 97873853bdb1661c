"""Thermal run of the engine against the FP64 oracle's (a check of the
deformation, flow and aggregation statistics): the 32^3 reverse-Poiseuille
box with eight cells at kT = 0.01 (tests.thermal_util), 2000 steps, 51
samples from step 1000.  The oracle's samples are the fixture
tests/golden/thermal_box32.json (tests/golden/make_thermal.py).  The two
runs start from the same state with the same noise streams; FP32 and FP64
trajectories separate over the run, so the comparison is of time averages:

* deformation index of every cell: |mean difference| <= max(0.01,
  4 x the block-average standard error of the per-sample difference);
* x-velocity in 8 slabs along y: |mean difference| <= max(10 % of the
  largest oracle slab speed, 4 x that standard error);
* inter-cell vertex RDF (16 bins to r = 8): |mean difference| <= max(0.05,
  4 x that standard error) in every bin.

tests/test_thermal_fixture.py shows on the CPU that the oracle with a 25 %
larger body force fails the velocity bar."""
import json
import os

import numpy as np
import pytest

from gpucell_amd.engine import Engine
from tests import thermal_util as tu

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "thermal_box32.json")


def compare(ref: dict, run: dict) -> list:
    """Failures of `run`'s samples against `ref`'s (both {di, vel, rdf}: one
    row per sample, the same sample steps).  The standard error is that of
    the per-sample difference (block averages): the runs share their noise
    streams, so the difference is far quieter than either series."""
    bad = []
    for key in ("di", "vel", "rdf"):
        o, g = np.array(ref[key]), np.array(run[key])
        mo, _ = tu.block_mean_err(o)
        mg, _ = tu.block_mean_err(g)
        _, ed = tu.block_mean_err(g - o)
        err = 4.0 * ed
        if key == "di":
            tol = np.maximum(0.01, err)
        elif key == "vel":
            tol = np.maximum(0.1 * np.abs(mo).max(), err)
        else:
            tol = np.maximum(0.05, err)
        d = np.abs(mg - mo)
        for i in np.nonzero(d > tol)[0]:
            bad.append(f"{key}[{i}]: {mg[i]:.4f} vs oracle {mo[i]:.4f} (tol {tol[i]:.4f})")
    return bad


def test_thermal_run_statistics_match_oracle(gpu):
    with open(FIXTURE) as f:
        fx = json.load(f)
    case = tu.thermal_case()
    assert fx["case"] == tu.fingerprint(case), "fixture made from another case: rerun make_thermal"
    assert fx["sample_steps"] == tu.sample_steps()
    e = Engine(case)
    run = {"di": [], "vel": [], "rdf": []}
    done = 0
    for k in tu.sample_steps():
        e.step(k - done)
        done = k
        s = tu.stats(e.state(), case)
        for key in run:
            run[key].append(s[key])
    assert e.info().fault == 0
    e.close()
    for key in run:
        m, _ = tu.block_mean_err(np.array(run[key]))
        mo, _ = tu.block_mean_err(np.array(fx["oracle"][key]))
        print(f"{key}: engine {np.round(m, 4)}\n{'':{len(key)}}  oracle {np.round(mo, 4)}", flush=True)
    bad = compare(fx["oracle"], run)
    assert not bad, "; ".join(bad)
