"""Failure detection of the engine (the reference only aborts on runtime
errors through CHECK, src/erro.cuh:6-20): conditions forced from the host,
every kernel still runs to its end.

* a non-finite velocity makes the next step's position non-finite: the
  fault word gets GC_FAULT_NONFINITE, export and checkpoint return
  GC_ERR_NONFINITE, and a clean SetState clears it;
* a membrane vertex pulled out so that its edges pass l_max (the WLC
  spring's singularity) raises GC_FAULT_WLC;
* a rejected gc_EngineSetMembraneMode call (bad normalisation vector)
  leaves the engine exactly as it was."""
import numpy as np
import pytest

from gpucell_amd import abi, casegen
from gpucell_amd.engine import Engine
from tests import util

pytestmark = pytest.mark.gpu

FAULT_NONFINITE, FAULT_WLC = 1, 2


def _case():
    ph = util.replace_physics(kT=0.01, fext=0.05)
    return casegen.channel(24, 12, 2, rbc_diameter=5.0, subdiv=2, seed=8, physics=ph)


def test_nonfinite_state_is_reported_and_cleared(gpu, tmp_path):
    case = _case()
    e = Engine(case)
    e.step(2)
    S, info = e.state(), e.info()
    assert info.fault == 0
    bad = S["velrhop"].copy()
    i = int(case.consts.nFluidStart) + 17
    bad[i, 0] = np.nan
    e.set_state({"velrhop": bad}, info.step, info.noise_counter)
    e.step(1)
    assert e.info().fault & FAULT_NONFINITE
    with pytest.raises(abi.GcError, match="non-finite"):
        e.state()
    with pytest.raises(abi.GcError, match="non-finite"):
        e.save(tmp_path / "bad.ck")
    assert not (tmp_path / "bad.ck").exists()
    # the fault is sticky until a state is set; a clean state clears it
    e.step(1)
    assert e.info().fault & FAULT_NONFINITE
    e.set_state({k: S[k] for k in ("pos", "velrhop", "velhalf", "force", "nthrobound")}, info.step,
                info.noise_counter)
    assert e.info().fault == 0
    e.step(2)
    S2 = e.state()
    assert np.isfinite(S2["pos"]).all() and e.info().fault == 0


def test_edge_past_lmax_is_reported(gpu):
    case = _case()
    e = Engine(case)
    e.step(1)
    S, info = e.state(), e.info()
    # pull vertex v of cell 0 outward by three times its longest edge: its
    # edges reach ~3 l0 > l_max = 2.2 l0
    nc = int(case.consts.nCellStart)
    v = 0
    cell = S["pos"][nc:nc + case.nCellPar // case.nCell]
    tri = case.tri[:case.tri.shape[0] // case.nCell] - 1
    nb = np.unique(tri[np.any(tri == v, axis=1)])
    nb = nb[nb != v]
    lmax = np.linalg.norm(cell[nb] - cell[v], axis=1).max()
    out = cell[v] - cell.mean(0)
    pos = S["pos"].copy()
    pos[nc + v] += 3.0 * lmax * out / np.linalg.norm(out)
    e.set_state({"pos": pos}, info.step, info.noise_counter)
    e.step(1)
    f = e.info().fault
    assert f & FAULT_WLC, f
    with pytest.raises(abi.GcError, match="l_max"):
        e.state()


def test_rejected_membrane_mode_leaves_the_engine_unchanged(gpu):
    case = _case()
    a, b = Engine(case), Engine(case)
    a.step(2)
    b.step(2)
    bad_norm = np.zeros(2 * (36 + 24), np.float32)   # standard deviations of zero
    with pytest.raises(abi.GcError, match="std must be > 0"):
        a.set_membrane_mode(1, k=4, norm=bad_norm)
    a.step(3)
    b.step(3)
    sa, sb = a.state(), b.state()
    for k in ("pos", "velrhop", "force", "defforce"):
        assert np.array_equal(sa[k], sb[k]), k
