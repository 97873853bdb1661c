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
* the 10M-particle channel with 2000 RBCs, every particle checked;
* the same in the engine's Verlet skin mode (the mode the benchmark runs),
  past at least one list rebuild: the 1M and 10M suspensions, the
  stenosed channel with Morse, and three replicas of the batch.

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


def _run(case, label, warm=4, checked=1, skin=0.0):
    e = Engine(case, skin=skin)
    o = orc.Oracle(case)
    o.initial_state()
    e.step(warm)
    while skin and e.info().nrebuild == 0 and e.info().step < 80:   # check after a list rebuild
        e.step(1)
    out = []
    for k in range(checked):
        S, info = e.state(), e.info()
        e.step(1)
        S1 = e.state()
        emax, l2, extra = gu.step_parity(o, case, S, S1, info, label=f"{label} step {info.step}")
        print(f"{label} N={case.N} step {info.step}: max per-particle force error {emax:.3e}, "
              f"relative L2 {l2:.3e}", flush=True)
        out.append((S1, extra))
    info = e.info()
    assert info.fault == 0
    if skin:
        print(f"{label}: skin {skin}, {info.nrebuild} list rebuilds in {info.step} steps", flush=True)
        assert info.nrebuild >= 1
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


# Verlet skin mode at the benchmark's skin (0.2): the start-up transient
# of these cases rebuilds the lists every ~10-30 steps; the checked steps
# come after the first rebuild (and between two).
@pytest.mark.timeout(600)
def test_suspension_1m_skin(gpu):
    _run(casegen.suspension_1m(), "1M suspension", warm=14, checked=2, skin=0.2)


def test_suspension_1m_skin_long_run(gpu):
    """3000 steps in the skin mode (past the start-up transient, with
    rebuilds every ~10 steps first and further apart later), graph replay:
    the state stays finite, no list overflows, and the step after is still
    the FP64 oracle's."""
    case = casegen.suspension_1m()
    e = Engine(case, skin=0.2)
    e.step(3000, graph=True)
    info = e.info()
    print(f"1M suspension, skin 0.2: {info.nrebuild} list rebuilds in {info.step} steps, "
          f"{info.nl_overflow} list overflows", flush=True)
    assert info.fault == 0 and info.nl_overflow == 0
    assert 3000 // 40 <= info.nrebuild <= 3000 // 4
    o = orc.Oracle(case)
    o.initial_state()
    S = e.state()
    assert np.isfinite(S["pos"]).all() and np.isfinite(S["velrhop"]).all()
    e.step(1)
    emax, l2, _ = gu.step_parity(o, case, S, e.state(), info, label="1M suspension step 3001")
    print(f"1M suspension step 3001: max per-particle force error {emax:.3e}, relative L2 {l2:.3e}")
    e.close()


@pytest.mark.timeout(600)
def test_stenosis_morse_skin(gpu):
    case = casegen.stenosis_2m()
    (S1, extra), = _run(case, "stenosis + Morse", warm=14, skin=0.2)
    assert util.morse_pairs(case, S1["coor"]) > 1000
    agg = extra["agg"]
    assert np.allclose(S1["aggforce"], agg, rtol=1e-4, atol=2e-5 * np.abs(agg).max())


@pytest.mark.timeout(900)
def test_ensemble64_replicas_skin(gpu):
    cases = casegen.ensemble64()
    sizes = np.array([c.N for c in cases])
    morse = [r for r, c in enumerate(cases) if c.consts.surEnergy > 0]
    pick = sorted({int(sizes.argmin()), int(sizes.argmax()), morse[0]})
    ens = Ensemble(cases, skin=0.2)
    ens.step(14)
    while ens.info().nrebuild == 0 and ens.info().step < 80:
        ens.step(1)
    info = ens.info()
    assert info.nrebuild >= 1
    before = {r: ens.state(r) for r in pick}
    ens.step(1)
    assert ens.info().fault == 0
    for r in pick:
        S1 = ens.state(r)
        o = orc.Oracle(cases[r])
        o.initial_state()
        emax, l2, _ = gu.step_parity(o, cases[r], before[r], S1, info, label=f"skin replica {r}")
        print(f"ensemble (skin 0.2) replica {r} N={cases[r].N}: max per-particle force error {emax:.3e}, "
              f"relative L2 {l2:.3e}", flush=True)
    ens.close()


@pytest.mark.timeout(900)
def test_suspension_10m_skin(gpu):
    case = casegen.suspension_10m()
    _run(case, "10M suspension", warm=14, skin=0.2)
