"""Batched ensembles (gc_EnsembleCreate): several independent suspensions of
different sizes, boxes, boundary conditions and physics stepped by one
engine.  Each replica must reproduce its solo gc_EngineCreate run bit for
bit (same slot order within its cubes, same summation order, pair noise
from its own ids and seed), eagerly and through the graph.  The solo runs
are themselves pinned to the oracle by test_gpu_engine.py."""
import numpy as np
import pytest

from gpucell_amd import abi, casegen
from gpucell_amd.engine import Engine, Ensemble
from tests import util

pytestmark = pytest.mark.gpu


def _cases():
    ph_a = util.replace_physics(kT=0.01, fext=0.05)
    ph_b = util.replace_physics(kT=0.02, De=0.02, r0=0.3, beta=1.5, mcut=1.5, fext=0.03, seed=9)
    ph_c = util.replace_physics(kT=0.005, eta=2.0, seed=3)
    return [
        util.inner_walls_as_bound1(casegen.channel(24, 12, 2, rbc_diameter=5.0, subdiv=2,
                                                   seed=5, physics=ph_a), ph_a, 6.0),
        util.stacked_pair_channel(30, 14, diameter=5.0, subdiv=2, gap=0.6, physics=ph_b, seed=7),
        casegen.periodic_box(n=14, rbc_diameter=6.0, subdiv=2, seed=11, physics=ph_c),
        casegen.periodic_box(n=10, n_rbc=0, seed=2),   # fluid only
        casegen.couette(n=12, n_rbc=1, rbc_diameter=5.0, shear_rate=0.01, seed=4,
                        physics=util.replace_physics(kT=0.01, De=0.02, r0=0.3, beta=1.5, mcut=1.5)),
    ]


KEYS = ("pos", "velrhop", "velhalf", "force", "nthrobound", "coor", "area", "vol",
        "defforce", "aggforce")


def _same(a, b, tag):
    for k in KEYS:
        if k in a:
            assert np.array_equal(a[k], b[k]), f"{tag}: {k} differs"


@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_replicas_match_solo_runs(gpu, graph):
    cases = _cases()
    nsteps = 7
    ens = Ensemble(cases)
    assert ens.nrep == len(cases)
    ens.step(nsteps, graph=graph)
    for r, c in enumerate(cases):
        solo = Engine(c)
        solo.step(nsteps, graph=graph)
        _same(ens.state(r), solo.state(), f"replica {r}")
        solo.close()
    info = ens.info()
    assert info.step == nsteps
    assert info.nTotalPar == sum(c.N for c in cases)


def test_ensemble_rejects_single_replica_calls(gpu):
    cases = _cases()[:2]
    ens = Ensemble(cases)
    lib = ens.lib
    hs = abi.gc_host_state()
    assert lib.gc_EngineGetState(ens.h, hs) == abi.GC_ERR_INVALID
    assert lib.gc_EngineSave(ens.h, b"/nonexistent") == abi.GC_ERR_INVALID
    assert lib.gc_EnsembleGetState(ens.h, 2, hs) == abi.GC_ERR_INVALID
    assert lib.gc_EnsembleGetState(ens.h, -1, hs) == abi.GC_ERR_INVALID
