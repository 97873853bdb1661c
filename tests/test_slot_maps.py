"""gc_CheckSlotMaps, the host-side check of an engine's slot maps (sp,
CurPos of the membrane vertices, the force pass's act list), on the CPU:
consistent maps pass, and each planted mistake is named -- an id outside
[0, N), an id in two slots, a membrane vertex's CurPos pointing at another
particle, act missing a moving particle, out of order, or holding a wall.
These are the index arrays whose mistakes faulted the GPU (a lane reading
another wave's neighbour-list block; a gather from a stale old slot)."""
import numpy as np
import pytest

from gpucell_amd import abi


def _maps(seed=3, n=400, nf=100, nc=300):
    rng = np.random.default_rng(seed)
    sp = rng.permutation(n).astype(np.int32)           # slot -> id
    curpos = np.zeros(n, np.int32)
    curpos[sp] = np.arange(n, dtype=np.int32)          # id -> slot
    act = np.nonzero(sp >= nf)[0].astype(np.int32)     # moving slots in order
    return n, nf, nc, sp, curpos, act


def _check(n, nf, nc, sp, curpos, act, nmov=None):
    lib = abi.load()
    nmov = len(act) if nmov is None else nmov
    st = lib.gc_CheckSlotMaps(n, nf, nc, nmov, sp.ctypes.data, curpos.ctypes.data, act.ctypes.data)
    return st, lib.gc_GetLastError().decode()


def test_consistent_maps_pass():
    assert _check(*_maps())[0] == 0
    # moving walls: act lists every slot in order
    n, nf, nc, sp, cp, _ = _maps()
    assert _check(n, nf, nc, sp, cp, np.arange(n, dtype=np.int32))[0] == 0
    # only membrane CurPos entries are checked (the step keeps no others)
    cp2 = cp.copy()
    cp2[:nc] = -7
    assert _check(n, nf, nc, sp, cp2, _maps()[5])[0] == 0


@pytest.mark.parametrize("mistake, needle", [
    ("id_out_of_range", "outside"),
    ("duplicate_id", "two slots"),
    ("curpos_wrong", "CurPos of membrane vertex"),
    ("curpos_out_of_range", "CurPos of membrane vertex"),
    ("act_missing", "act holds"),
    ("act_unordered", "in order"),
    ("act_wall", "holds wall"),
])
def test_planted_mistakes_are_named(mistake, needle):
    n, nf, nc, sp, cp, act = _maps()
    sp, cp, act = sp.copy(), cp.copy(), act.copy()
    if mistake == "id_out_of_range":
        sp[17] = n + 5
    elif mistake == "duplicate_id":
        sp[17] = sp[18]
    elif mistake == "curpos_wrong":
        v = nc + 4
        cp[v] = (cp[v] + 1) % n
    elif mistake == "curpos_out_of_range":
        cp[nc + 9] = n
    elif mistake == "act_missing":
        act = act[:-1]
    elif mistake == "act_unordered":
        act[[3, 4]] = act[[4, 3]]
    elif mistake == "act_wall":
        wall_slot = int(np.nonzero(sp < nf)[0][0])
        act[0] = wall_slot
        act.sort()
    st, msg = _check(n, nf, nc, sp, cp, act)
    assert st == -1 and needle in msg, msg


def test_bad_arguments_rejected():
    n, nf, nc, sp, cp, act = _maps()
    lib = abi.load()
    assert lib.gc_CheckSlotMaps(n, nf, nc, len(act), None, cp.ctypes.data, act.ctypes.data) == -1
    assert lib.gc_CheckSlotMaps(n, nc + 1, nc, len(act), sp.ctypes.data, cp.ctypes.data, act.ctypes.data) == -1
