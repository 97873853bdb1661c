"""Writes tests/golden/thermal_box32.json: the thermal-run statistics of the
FP64 oracle (oracle/oracle.c, one reference step per orc_step) on
tests.thermal_util.thermal_case(), sampled every SAMPLE_EVERY steps from
SAMPLE_FROM to STEPS.  tests/test_gpu_thermal.py runs the engine on the same
case and compares.  Run from the repository root:

    python -m tests.golden.make_thermal [threads]

(about 10 minutes on 8 threads)."""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

from oracle import pyoracle as orc
from tests import thermal_util as tu

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "thermal_box32.json")


def main(threads: int = 8):
    case = tu.thermal_case()
    o = orc.Oracle(case, threads=threads)
    st = o.initial_state()
    samples = {"di": [], "vel": [], "rdf": []}
    want = set(tu.sample_steps())
    t0 = time.time()
    for k in range(1, tu.STEPS + 1):
        o.step(st)
        if k in want:
            s = tu.stats({"pos": st["pos"], "nthrobound": st["img"],
                          "velrhop": np.concatenate([st["vel"], st["rho"][:, None]], axis=1)}, case)
            for key in samples:
                samples[key].append(s[key])
        if k % 100 == 0:
            print(f"step {k} ({time.time() - t0:.0f} s)", flush=True)
    out = {"case": tu.fingerprint(case), "steps": tu.STEPS, "sample_steps": tu.sample_steps(),
           "vel_bins": tu.VEL_BINS, "rdf_rmax": tu.RDF_RMAX, "rdf_bins": tu.RDF_BINS,
           "oracle": samples}
    with open(OUT, "w") as f:
        json.dump(out, f)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
