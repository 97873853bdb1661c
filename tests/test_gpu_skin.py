"""The engine's Verlet skin mode (gc_case_desc.skin > 0): neighbour lists of
every pair within h + skin, re-sorted and rebuilt only when some pair could
have moved apart by more than the skin since the last rebuild (decided on
the device from the displacements), density over the lists, force over the
neighbour bits the density pass sets.

* one step from an exported state against the FP64 oracle on every engine
  case, with a skin that rebuilds every few steps (so the checked step comes
  after rebuilds, between two) and one that never rebuilds in the run: FP64
  positions to 1e-9, densities to 2e-6, every particle's force within 1e-5;
* a longer run with many rebuilds still checks against the oracle;
* the same trajectory as the per-step mode within FP32 rounding (different
  summation order, noise off);
* graph replay == eager launches, bit for bit, rebuilds included;
* each replica of a batch == its solo run, bit for bit (each replica decides
  its own rebuilds), also with replicas only a few waves long;
* membrane substepping under the skin mode (its displacements join the
  rebuild bound);
* particles that start outside a periodic face (the first lists are built
  from the wrapped positions)."""
import numpy as np
import pytest

from gpucell_amd import casegen
from gpucell_amd.engine import Engine, Ensemble
from oracle import pyoracle as orc
from tests import gpu_util as gu
from tests import util
from tests.test_gpu_engine import CASES
from tests.test_gpu_ensemble import KEYS, _cases as ensemble_cases

pytestmark = pytest.mark.gpu


def _same(a, b, tag):
    for k in KEYS:
        if k in a:
            assert np.array_equal(a[k], b[k]), f"{tag}: {k} differs"


@pytest.mark.parametrize("skin", [0.02, 0.3])
@pytest.mark.parametrize("name", list(CASES))
def test_skin_step_parity(gpu, name, skin):
    case = CASES[name]()
    e = Engine(case, skin=skin)
    o = orc.Oracle(case)
    o.initial_state()
    e.step(9)
    S, info = e.state(), e.info()
    e.step(1)
    S1 = e.state()
    emax, l2, _ = gu.step_parity(o, case, S, S1, info, label=f"{name} skin {skin}")
    i1 = e.info()
    print(f"{name} skin {skin}: max per-particle force error {emax:.3e}, relative L2 {l2:.3e}, "
          f"rebuilds {i1.nrebuild}, list overflows {i1.nl_overflow}")
    assert abs(i1.skin - skin) < 1e-7
    if name == "long_channel":   # the skin mode's wider cubes are tiled too
        from tests.test_gpu_engine import tiled
        assert tiled(i1.ncube)
    if skin == 0.02:
        assert i1.nrebuild >= 1
    assert i1.fault == 0


def test_skin_many_rebuilds(gpu):
    case = CASES["channel_morse"]()
    e = Engine(case, skin=0.03)
    o = orc.Oracle(case)
    o.initial_state()
    e.step(59, graph=True)
    S, info = e.state(), e.info()
    e.step(1)
    S1 = e.state()
    emax, l2, _ = gu.step_parity(o, case, S, S1, info, label="channel_morse skin 0.03, step 60")
    i1 = e.info()
    print(f"60 steps, skin 0.03: {i1.nrebuild} rebuilds, force error {emax:.3e}")
    assert i1.nrebuild >= 5
    assert i1.fault == 0


@pytest.mark.parametrize("name", ["channel", "box", "couette"])
def test_skin_tracks_per_step_mode(gpu, name):
    """Noise off: the skin mode sums the same pair terms in another order,
    so the two trajectories differ by FP32 rounding only."""
    if name == "channel":
        ph = util.replace_physics(kT=0.01, fext=0.05)
        case = casegen.channel(24, 12, 2, rbc_diameter=5.0, subdiv=2, seed=5, physics=ph, noise=False)
    else:
        case = CASES[name]()
        case.ext.noise_on = 0
    a = Engine(case)
    b = Engine(case, skin=0.05)
    a.step(30)
    b.step(30)
    Sa, Sb = a.state(), b.state()
    assert b.info().nrebuild >= 1
    dpos = np.abs(Sa["pos"] - Sb["pos"]).max()
    dvel = np.abs(Sa["velrhop"] - Sb["velrhop"]).max()
    scale = np.abs(Sa["velrhop"][:, :3]).max()
    print(f"{name}: 30 steps, per-step vs skin: max |dx| {dpos:.2e}, max |dv| {dvel:.2e} (|v| <= {scale:.2e})")
    assert np.array_equal(Sa["nthrobound"], Sb["nthrobound"])
    assert dpos < 1e-5
    assert dvel < 1e-4 * max(scale, 1.0)


def test_skin_graph_equals_eager(gpu):
    case = CASES["channel_morse"]()
    a = Engine(case, skin=0.04)
    b = Engine(case, skin=0.04)
    a.step(16)
    b.step(16, graph=True)
    ia, ib = a.info(), b.info()
    assert ia.nrebuild == ib.nrebuild >= 1
    _same(a.state(), b.state(), "graph vs eager")


@pytest.mark.parametrize("graph", [False, True])
def test_skin_ensemble_replicas_match_solo_runs(gpu, graph):
    cases = ensemble_cases()
    nsteps = 14
    ens = Ensemble(cases, skin=0.04)
    ens.step(nsteps, graph=graph)
    total = 0
    for r, c in enumerate(cases):
        solo = Engine(c, skin=0.04)
        solo.step(nsteps, graph=graph)
        _same(ens.state(r), solo.state(), f"replica {r}")
        total += solo.info().nrebuild
        solo.close()
    info = ens.info()
    print(f"ensemble: {info.nrebuild} rebuilds over {len(cases)} replicas")
    assert info.nrebuild == total >= len(cases)


@pytest.mark.parametrize("graph", [False, True])
def test_skin_batch_of_small_replicas_matches_solo_runs(gpu, graph):
    """Replicas of 216-512 particles (a few waves each) beside a larger one:
    many 64-slot chunks and waves hold two replicas, one rebuilding in a
    step and the other not (the batch re-sorts and rebuilds the lists of the
    rebuilding ones only, lane by lane in such a wave), and each replica
    still evolves bit for bit as its solo run."""
    cases = [casegen.periodic_box(n=6, n_rbc=0, seed=21, physics=util.replace_physics(kT=0.02, seed=21)),
             casegen.periodic_box(n=7, n_rbc=0, seed=22, physics=util.replace_physics(kT=0.005, seed=22)),
             ensemble_cases()[0],
             casegen.periodic_box(n=8, n_rbc=0, seed=23, physics=util.replace_physics(kT=0.03, seed=23)),
             casegen.periodic_box(n=6, n_rbc=0, seed=24, physics=util.replace_physics(kT=0.01, seed=24))]
    nsteps, skin = 24, 0.03
    ens = Ensemble(cases, skin=skin)
    ens.step(nsteps, graph=graph)
    counts = []
    for r, c in enumerate(cases):
        solo = Engine(c, skin=skin)
        solo.step(nsteps, graph=graph)
        _same(ens.state(r), solo.state(), f"replica {r}")
        counts.append(solo.info().nrebuild)
        solo.close()
    print(f"small replicas: rebuilds per replica {counts}")
    assert ens.info().nrebuild == sum(counts)
    assert min(counts) >= 1 and len(set(counts)) > 1   # the replicas rebuild at different steps


@pytest.mark.parametrize("mode", [1])
def test_skin_membrane_substeps(gpu, mode):
    case = casegen.periodic_box(n=14, rbc_diameter=6.0, subdiv=2, seed=11)
    case.ext.noise_on = 0
    a = Engine(case)
    b = Engine(case, skin=0.03)
    for e in (a, b):
        e.set_membrane_mode(mode, k=4)
        e.step(20)
    Sa, Sb = a.state(), b.state()
    assert b.info().nrebuild >= 1
    assert b.info().fault == 0
    dpos = np.abs(Sa["pos"] - Sb["pos"]).max()
    print(f"substeps k=4, 20 steps: per-step vs skin max |dx| {dpos:.2e}, rebuilds {b.info().nrebuild}")
    assert dpos < 1e-5


def test_skin_start_outside_periodic_face(gpu):
    """Moving particles that start just outside the periodic x faces (as
    generated cells crossing a face do): the first lists must come from the
    wrapped positions, while the exported start state keeps the given ones."""
    case = CASES["channel_morse"]()
    mov = np.arange(case.N) >= case.consts.nFluidStart
    L = float(case.bbox[1] - case.bbox[0])
    x = case.coor[:, 0]
    shift = np.where(mov & (x < 0.7), -0.75, np.where(mov & (x > L - 0.7), 0.75, 0.0)).astype(np.float32)
    case.coor[:, 0] = x + shift
    n_out = int(np.sum((case.coor[:, 0] < 0) | (case.coor[:, 0] >= L)))
    assert n_out > 50
    e = Engine(case, skin=0.3)
    S0 = e.state()
    assert np.array_equal(S0["coor"], case.coor)   # the start state as given
    o = orc.Oracle(case)
    o.initial_state()
    for k in range(2):
        S, info = e.state(), e.info()
        e.step(1)
        S1 = e.state()
        emax, _, _ = gu.step_parity(o, case, S, S1, info, label=f"start outside the box, step {k + 1}")
        print(f"{n_out} particles start outside the box: step {k + 1} force error {emax:.3e}")
    assert e.info().nrebuild == 0
