"""Membrane substepping and the learned substepper's machinery
(gc_EngineSetMembraneMode, include/gpucell/surrogate.h).

* plain substepping (mode 1) against a velocity-Verlet loop on the FP64
  oracle's membrane force, with the step's fluid force held fixed;
* the surrogate (mode 2) with a threshold no prediction meets falls back
  on every step and reproduces mode 1 bit for bit;
* recorded training rows hold the substep displacements (checked through
  their rotation-invariant norm);
* leaving substepping restores the reference's membrane update;
* replaying the substepped steps from a HIP graph gives the eager bits, in
  mode 1 and in mode 2 with the fallback decided on the device."""
import numpy as np
import pytest

from gpucell_amd import casegen
from gpucell_amd.engine import Engine
from gpucell_amd.surrogate import Mlp
from oracle import pyoracle as orc
from tests import gpu_util as gu
from tests import util

pytestmark = pytest.mark.gpu


def small_case(seed=4):
    ph = util.replace_physics(fext=0.05, mu0=4.0, ka=40.0, kv=80.0)
    return casegen.periodic_box(n=14, rbc_diameter=6.0, subdiv=2, seed=seed, physics=ph, noise=False)


def _L(case):
    b = case.bbox
    per = np.array(list(case.ext.periodic), np.float64)
    return np.array([b[1] - b[0], b[3] - b[2], b[5] - b[4]], np.float64) * per


def _unwrap(S, case):
    return S["pos"] + S["nthrobound"] * _L(case)


@pytest.mark.parametrize("k", [1, 4])
def test_plain_substeps_match_oracle(gpu, k):
    case = small_case()
    e = Engine(case)
    o = orc.Oracle(case)
    o.initial_state()
    e.step(2)
    e.set_membrane_mode(1, k=k)
    e.step(2)
    S0 = e.state()
    e.step(1)
    S1 = e.state()
    nc, m = case.N - case.nCellPar, float(case.consts.Mass)
    dt = float(case.consts.TimeStep)
    dts = dt / k
    x = _unwrap(S0, case)
    img = np.zeros_like(S0["nthrobound"])
    v = S0["velrhop"][:, :3].astype(np.float64)
    fe = S1["force"][nc:].astype(np.float64)     # the step's fluid force on the membrane

    def fmem(xx):
        o.geo(int(S0["velrhop"].shape[0]), xx, img)
        return o.def_force(xx, img)

    fm = fmem(x)
    x0 = x[nc:].copy()
    for _ in range(k):
        v[nc:] += 0.5 * dts * (fe + fm) / m
        x[nc:] += dts * v[nc:]
        fm = fmem(x)
        v[nc:] += 0.5 * dts * (fe + fm) / m
    xg = _unwrap(S1, case)[nc:]
    vg = S1["velrhop"][nc:, :3].astype(np.float64)
    dx = np.linalg.norm(x[nc:] - x0, axis=1)
    ex = np.linalg.norm(xg - x[nc:], axis=1).max() / dx.max()
    ev = np.linalg.norm(vg - v[nc:], axis=1).max() / np.linalg.norm(v[nc:] - S0["velrhop"][nc:, :3], axis=1).max()
    print(f"k={k}: position error / displacement {ex:.2e}, velocity error / change {ev:.2e}")
    assert ex < 1e-4
    assert ev < 1e-3
    assert np.allclose(S1["defforce"], fm, rtol=1e-4, atol=1e-4 * np.abs(fm).max())
    # fluid particles still take the reference step: nothing membrane-side leaks
    assert np.all(np.isfinite(S1["velrhop"]))


def test_surrogate_fallback_reproduces_plain(gpu):
    case = small_case(seed=6)
    a, b = Engine(case), Engine(case)
    a.set_membrane_mode(1, k=8)
    net = Mlp(seed=2)
    b.set_membrane_mode(2, k=8, mlp=net, threshold=1e-30)
    a.step(3)
    b.step(3)
    sa, sb = a.state(), b.state()
    for key in ("pos", "velrhop", "force", "nthrobound", "defforce", "area", "vol"):
        assert np.array_equal(sa[key], sb[key]), key
    st = b.membrane_stats()
    assert st["steps"] == 3 and st["fallbacks"] == 3
    assert st["last_residual"] > 0


def test_recorded_samples_hold_substep_displacements(gpu):
    import torch
    case = small_case(seed=7)
    e = Engine(case)
    e.set_membrane_mode(1, k=8)
    e.step(1)
    nm, nc = case.nCellPar, case.N - case.nCellPar
    X = torch.zeros((2 * nm, 36), dtype=torch.float32, device="cuda")
    Y = torch.zeros((2 * nm, 24), dtype=torch.float32, device="cuda")
    S0 = e.state()
    e.record_samples(X, Y)
    e.step(1)
    S1 = e.state()
    assert e.membrane_stats()["samples"] == nm
    x, y = gu.host(X), gu.host(Y)
    assert np.all(np.isfinite(x[:nm])) and np.all(np.isfinite(y[:nm]))
    assert np.all(x[nm:] == 0)
    # rest edge length per vertex (mean over its edges at creation)
    E = case.edge[:, :2] - 1
    l0 = np.linalg.norm(case.coor[nc + E[:, 0]] - case.coor[nc + E[:, 1]], axis=1)
    lsum = np.bincount(E.ravel(), weights=np.repeat(l0, 2), minlength=nm)
    lref = lsum / np.bincount(E.ravel(), minlength=nm)
    dt = float(case.consts.TimeStep)
    disp = _unwrap(S1, case)[nc:] - _unwrap(S0, case)[nc:] - dt * S0["velrhop"][nc:, :3]
    got = np.linalg.norm(y[:nm, 21:24], axis=1) * lref
    assert np.allclose(got, np.linalg.norm(disp, axis=1), rtol=2e-3, atol=1e-7)
    # features: ring edge vectors have the rest length scale, degree flag 0/1
    assert np.all((x[:nm, 33] == 0) | (x[:nm, 33] == 1))
    assert 0.5 < np.median(np.linalg.norm(x[:nm, 0:3], axis=1)) < 1.5
    e.record_samples(None, None)


def test_leaving_substepping_restores_reference_update(gpu):
    case = small_case(seed=8)
    e = Engine(case)
    e.set_membrane_mode(1, k=2)
    e.step(2)
    e.set_membrane_mode(0)
    e.step(2)
    S = e.state()
    assert np.all(np.isfinite(S["pos"])) and np.all(np.isfinite(S["velrhop"]))
    assert e.membrane_stats()["steps"] == 0


_KEYS = ("pos", "velrhop", "force", "nthrobound", "defforce", "area", "vol")


@pytest.mark.parametrize("k", [2, 8])
def test_plain_substeps_graph_replay_bit_identical(gpu, k):
    case = small_case(seed=10)
    a, b = Engine(case), Engine(case)
    for e in (a, b):
        e.set_membrane_mode(1, k=k)
    a.step(5)
    b.step(5, graph=True)
    sa, sb = a.state(), b.state()
    for key in _KEYS:
        assert np.array_equal(sa[key], sb[key]), key
    assert a.membrane_stats()["steps"] == b.membrane_stats()["steps"] == 5


@pytest.mark.parametrize("variant", [0, 3])
def test_surrogate_fallback_graph_replay_reproduces_plain(gpu, variant):
    """Variant 3 (per-cell membrane kernel) decides the fallback on the
    device and replays from the graph; variant 0 (three passes) reads the
    residual on the host and steps eagerly.  Both match plain substepping."""
    case = small_case(seed=11)
    a, b = Engine(case), Engine(case)
    for e in (a, b):
        t = e.tuning()
        t[3] = variant
        e.set_tuning(t)
    a.set_membrane_mode(1, k=8)
    b.set_membrane_mode(2, k=8, mlp=Mlp(seed=4), threshold=1e-30)
    a.step(5)
    b.step(5, graph=True)
    sa, sb = a.state(), b.state()
    for key in _KEYS:
        assert np.array_equal(sa[key], sb[key]), key
    st = b.membrane_stats()
    assert st["steps"] == 5 and st["fallbacks"] == 5
