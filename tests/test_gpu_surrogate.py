"""The learned membrane substepper end to end on the GPU: rows recorded
from plain substepping, the network trained by the library's own
BF16-MFMA kernels, and the surrogate run against plain substepping (the
reference for its error) on a held-out suspension.  What the trained network gains over the
method without it: its predictions pass the residual check, so steps need
no plain substeps, while an untrained network is rejected on every step."""
import numpy as np
import pytest

from gpucell_amd import casegen, substepper
from gpucell_amd.surrogate import Mlp
from tests import util

pytestmark = pytest.mark.gpu

THRESHOLD = 2e-5   # residual, in rest edge lengths


def _box(seed):
    ph = util.replace_physics(fext=0.05, mu0=4.0, ka=40.0, kv=80.0)
    return casegen.periodic_box(n=20, n_rbc=3, rbc_diameter=7.0, subdiv=3, seed=seed, physics=ph)


def test_trained_surrogate_tracks_plain_substepping(gpu):
    X, Y = substepper.collect([_box(1), _box(2)], steps=120, every=2, warmup=5)
    assert X.shape[0] > 200000
    net, norm, loss, secs = substepper.train(X, Y, epochs=12, batch=2048, lr=2e-3)
    print(f"rows {X.shape[0]}, loss {loss[0]:.3g} -> {loss[-1]:.3g} in {secs:.2f} s")
    assert loss[-1] < 0.2 * loss[0]
    held = _box(9)
    r = substepper.compare(held, net, norm, THRESHOLD, steps=30, relax=0)
    print("trained:", r)
    assert r["fallback_rate"] < 0.2
    assert r["di_rel_err_max"] < 1e-2          # deformation index vs plain
    assert r["vol_rel_diff_max"] < 5e-3        # cell volume vs plain
    # the device-decided fallback replayed from a graph: the eager bits
    from gpucell_amd.engine import Engine
    runs = []
    for graph in (False, True):
        e = Engine(held)
        e.set_membrane_mode(2, k=8, mlp=net, norm=norm, threshold=THRESHOLD)
        e.step(9, graph=graph)
        runs.append((e.state(), e.membrane_stats()))
        e.close()
    for key in ("pos", "velrhop", "defforce", "area", "vol"):
        assert np.array_equal(runs[0][0][key], runs[1][0][key]), key
    assert runs[0][1]["fallbacks"] == runs[1][1]["fallbacks"]
    assert runs[0][1]["steps"] == runs[1][1]["steps"] == 9
    u = substepper.compare(held, Mlp(seed=5), norm, THRESHOLD, steps=5, relax=0)
    print("untrained:", u)
    assert u["fallback_rate"] == 1.0
    # with every step falling back, the untrained network reproduces plain
    assert u["di_rel_err_max"] == 0.0 and u["vol_rel_diff_max"] == 0.0
