"""One engine step against the FP64 oracle at the workload of every
benchmark configuration (full size, default launch shapes):

* the 32^3 periodic box with one 642-vertex RBC (reverse-Poiseuille
  forcing), three consecutive steps checked;
* the 1M-particle channel with 100 RBCs at 20 % hematocrit, three
  consecutive steps checked;
* the ~1.9M-particle stenosed channel with 300 RBCs and Morse aggregation
  (Morse pairs between cells asserted present);
* the batch of 64 sheared replicas of uneven size stepped together: the
  smallest, the largest and a Morse replica checked;
* the 10M-particle channel with 2000 RBCs, every particle checked.

The bars are those of the engine tests: FP64 positions to 1e-9, densities
to 2e-6, every particle's force within 1e-5 of the oracle's (per-particle
error against the sum of its pair-term magnitudes), cell area and volume
to 2e-6."""
import numpy as np
import pytest

from gpucell_amd import casegen
from gpucell_amd.engine import Engine, Ensemble
from oracle import pyoracle as orc
from tests import gpu_util as gu
from tests import util

pytestmark = pytest.mark.gpu


def _run(case, label, warm=4, checked=1):
    e = Engine(case)
    o = orc.Oracle(case)
    o.initial_state()
    e.step(warm)
    out = []
    for k in range(checked):
        S, info = e.state(), e.info()
        e.step(1)
        S1 = e.state()
        emax, l2, extra = gu.step_parity(o, case, S, S1, info, label=f"{label} step {info.step}")
        print(f"{label} N={case.N} step {info.step}: max per-particle force error {emax:.3e}, "
              f"relative L2 {l2:.3e}", flush=True)
        out.append((S1, extra))
    assert e.info().fault == 0
    e.close()
    return out


def test_box_32_one_rbc(gpu):
    case = casegen.periodic_box(n=32)
    assert case.nCellPar == 642 and case.tri.shape[0] == 1280 and case.edge.shape[0] == 1920
    _run(case, "32^3 box + RBC", checked=3)


@pytest.mark.timeout(600)
def test_suspension_1m(gpu):
    case = casegen.suspension_1m()
    assert case.nCell == 100 and case.N > 1_000_000
    _run(case, "1M suspension", checked=3)


@pytest.mark.timeout(600)
def test_stenosis_morse(gpu):
    case = casegen.stenosis_2m()
    assert case.nCell == 300 and case.consts.surEnergy > 0
    (S1, extra), = _run(case, "stenosis + Morse")
    assert util.morse_pairs(case, S1["coor"]) > 1000
    agg = extra["agg"]
    assert np.count_nonzero(np.linalg.norm(agg, axis=1) > 0) > 1000
    assert np.allclose(S1["aggforce"], agg, rtol=1e-4, atol=2e-5 * np.abs(agg).max())


@pytest.mark.timeout(900)
def test_ensemble64_replicas(gpu):
    cases = casegen.ensemble64()
    sizes = np.array([c.N for c in cases])
    morse = [r for r, c in enumerate(cases) if c.consts.surEnergy > 0]
    pick = sorted({int(sizes.argmin()), int(sizes.argmax()), morse[0]})
    ens = Ensemble(cases)
    ens.step(4)
    info = ens.info()
    before = {r: ens.state(r) for r in pick}
    ens.step(1)
    assert ens.info().fault == 0
    for r in pick:
        S1 = ens.state(r)
        o = orc.Oracle(cases[r])
        o.initial_state()
        emax, l2, _ = gu.step_parity(o, cases[r], before[r], S1, info, label=f"replica {r}")
        print(f"ensemble replica {r} N={cases[r].N} De={cases[r].consts.surEnergy}: max per-particle "
              f"force error {emax:.3e}, relative L2 {l2:.3e}", flush=True)
    ens.close()


@pytest.mark.timeout(900)
def test_suspension_10m(gpu):
    case = casegen.suspension_10m()
    assert case.nCell == 2000 and case.N > 10_000_000
    _run(case, "10M suspension", warm=2)
