"""CPU checks of the thermal-run fixture (tests/golden/thermal_box32.json):
it was made from the case tests.thermal_util builds today, and the bars of
tests/test_gpu_thermal.py catch a plausible physics error: the same run with
a 25 % larger body force (the oracle with ConstExtForce x 1.25, a mistake
planted in the oracle, e.g. a force taken per mass instead of per particle)
fails the velocity-profile bar, while the fixture passes against itself.
(In this window a 25 % larger shear viscosity changes the slab speeds by
less than 1 %: the flow is still dominated by its start-up.)"""
import ctypes as C
import json
import os

import numpy as np

from oracle import pyoracle as orc
from tests import thermal_util as tu
from tests.test_gpu_thermal import FIXTURE, compare


def _fixture():
    with open(FIXTURE) as f:
        return json.load(f)


def test_fixture_is_of_the_current_case():
    fx = _fixture()
    case = tu.thermal_case()
    assert fx["case"] == tu.fingerprint(case)
    assert fx["sample_steps"] == tu.sample_steps()
    for key, width in (("di", case.nCell), ("vel", tu.VEL_BINS), ("rdf", tu.RDF_BINS)):
        a = np.array(fx["oracle"][key])
        assert a.shape == (len(tu.sample_steps()), width) and np.isfinite(a).all()
    assert compare(fx["oracle"], fx["oracle"]) == []


def test_wrong_body_force_fails_the_velocity_bar():
    fx = _fixture()
    case = tu.thermal_case()
    consts = type(case.consts).from_buffer_copy(case.consts)
    consts.ConstExtForce = C.c_float(case.consts.ConstExtForce * 1.25).value
    o = orc.Oracle(case, consts=consts, threads=os.cpu_count() or 1)
    st = o.initial_state()
    run = {"di": [], "vel": [], "rdf": []}
    want = set(tu.sample_steps())
    for k in range(1, tu.STEPS + 1):
        o.step(st)
        if k in want:
            s = tu.stats({"pos": st["pos"], "nthrobound": st["img"],
                          "velrhop": np.concatenate([st["vel"], st["rho"][:, None]], axis=1)}, case)
            for key in run:
                run[key].append(s[key])
    bad = compare(fx["oracle"], run)
    assert any(b.startswith("vel") for b in bad), bad
