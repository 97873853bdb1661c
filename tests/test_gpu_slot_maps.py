"""The engine's slot maps (sp, CurPos of the membrane vertices, act) pass
gc_EngineCheckSlotMaps after steps of every kind of case: fluid channel,
Morse pair, membranes past the 64-entry list, moving walls (act = every
slot), a cell in the three-pass membrane form, a batched ensemble, after an
autotune (the state is restored between variants) and after a checkpoint
resume with Morse on.  GPUCELL_CHECK_SLOTS=1 (set before the engines are
created) also checks them after every eager re-sort of those runs."""
import pytest

from gpucell_amd.engine import Engine, Ensemble
from tests import test_gpu_engine as ge

pytestmark = pytest.mark.gpu


@pytest.fixture
def check_every_step(monkeypatch):
    monkeypatch.setenv("GPUCELL_CHECK_SLOTS", "1")


@pytest.mark.parametrize("name", sorted(ge.CASES))
def test_slot_maps_after_steps(gpu, check_every_step, name):
    e = Engine(ge.CASES[name]())
    e.step(3)
    e.check_slot_maps()
    e.step(2, graph=True)
    e.check_slot_maps()
    e.close()


def test_slot_maps_after_autotune_and_resume(gpu, check_every_step, tmp_path):
    case = ge.morse_case()
    e = Engine(case)
    e.step(2)
    e.autotune(steps=2)
    e.check_slot_maps()
    e.step(2)
    e.save(tmp_path / "m.ck")
    e.step(2)
    e.load(tmp_path / "m.ck")
    e.step(1)
    e.check_slot_maps()
    e.close()


def test_slot_maps_of_an_ensemble(gpu, check_every_step):
    cases = [ge.CASES["channel"](), ge.morse_case(), ge.CASES["box"]()]
    ens = Ensemble(cases)
    ens.step(3)
    ens.check_slot_maps()
    ens.close()
