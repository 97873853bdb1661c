"""Autotuner (gc_EngineAutotune): every launch-shape variant of every tuned
kernel family computes bit-identical states (the membrane family includes
the fused per-cell kernel against the three-pass form), and tuning in the
middle of a run leaves the trajectory unchanged (the tuner restores the
state it started from)."""
import numpy as np
import pytest

from gpucell_amd import casegen
from gpucell_amd.engine import Engine
from tests import util

pytestmark = pytest.mark.gpu

NVAR = [3, 12, 3, 4]
KEYS = ("pos", "velrhop", "velhalf", "force", "nthrobound", "defforce", "area", "vol")


def case():
    ph = util.replace_physics(kT=0.01, De=0.02, r0=0.3, beta=1.5, mcut=1.5, fext=0.05)
    return casegen.channel(24, 12, 2, rbc_diameter=5.0, subdiv=2, seed=12, physics=ph)


def _run(c, tuning, steps=3):
    e = Engine(c)
    e.set_tuning(tuning)
    e.step(steps)
    return e.state()


def test_every_variant_is_bit_identical(gpu):
    c = case()
    ref_e = Engine(c)
    default = ref_e.tuning()
    ref_e.step(3)
    ref = ref_e.state()
    for fam, nv in enumerate(NVAR):
        for v in range(nv):
            t = list(default)
            t[fam] = v
            st = _run(c, t)
            for k in KEYS:
                assert np.array_equal(st[k], ref[k]), (fam, v, k)


def test_autotune_keeps_trajectory_and_reports_each_variant(gpu):
    c = case()
    a, b = Engine(c), Engine(c)
    a.step(2)
    rows = a.autotune(steps=2)
    a.step(3)
    b.step(5)
    sa, sb = a.state(), b.state()
    for k in KEYS:
        assert np.array_equal(sa[k], sb[k]), k
    assert a.info().step == 5
    fams = {}
    for r in rows:
        fams.setdefault(r["family"], []).append(r)
        assert r["us"] > 0
    assert sorted(len(v) for v in fams.values()) == sorted(NVAR)
    for fam, rs in fams.items():
        assert sum(r["chosen"] for r in rs) == 1, fam
        best = min(rs, key=lambda r: r["us"])
        assert best["chosen"], fam
    # the chosen variants are installed
    chosen = {r["family"]: r["variant"] for r in rows if r["chosen"]}
    t = a.tuning()
    assert t == [chosen["integrate"], chosen["density"], chosen["sdpd force"], chosen["membrane force"]]


def test_membrane_substeps_per_cell_kernel_bit_identical(gpu):
    """Plain membrane substepping evaluates the membrane force 8 times per
    step: the per-cell kernel (variant 3) and the three-pass form (variant
    0) give the same bits, and the same residual of the exact trajectory
    (the per-cell kernel forms x8 - 2 x7 + x6 itself)."""
    c = case()
    outs, res = [], []
    for v in (0, 3):
        e = Engine(c)
        t = e.tuning()
        t[3] = v
        e.set_tuning(t)
        e.set_membrane_mode(1, k=8)
        e.step(3)
        outs.append(e.state())
        res.append(e.membrane_stats()["plain_residual"])
        e.close()
    for k in KEYS:
        assert np.array_equal(outs[0][k], outs[1][k]), k
    assert res[0] > 0 and res[0] == res[1]
