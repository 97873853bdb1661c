"""The reference's four input decks (Particle::LoadParInfo, src/Particle.cu:
29-427): a hand-written deck in the 20-column format reads back to the
values the reference's loader would produce, and write -> read round-trips
a generated suspension exactly (host code only, no GPU)."""
import ctypes as C
import os

import numpy as np
import pytest

from gpucell_amd import abi, casegen


def col(*vals):
    return "".join(f"{v:<20}" for v in vals) + "\n"


def write_tiny_deck(d):
    # two Bound2 walls, three fluid particles, one tetrahedral "cell"
    with open(os.path.join(d, "SDPDParCoor.txt"), "w") as f:
        f.write("header\n" + col(0, 2, 3) + "Bound1\nBound2\n")
        f.write(col(0.5, -9.5, 0.0) + col(1.5, -9.5, 0.0))
        f.write("Fluid\n" + col(1.0, 0.0, 0.0) + col(2.0, 1.0, -1.0) + col(3.25, 2.5, 4.0))
    with open(os.path.join(d, "SDPDPhyPara.txt"), "w") as f:
        f.write("header\n" + col(10.0) + col(12.0) + "eta\n" + col(0.5) + "zeta\n" + col(0.1)
                + "rho0\n" + col(1.0) + "fext\n" + col(0.25))
    tet = [(1, 3, 2), (1, 2, 4), (1, 4, 3), (2, 3, 4)]
    edges = [(1, 2, 3, 4, 1, 2), (1, 3, 2, 4, 1, 3), (1, 4, 2, 3, 2, 3),
             (2, 3, 1, 4, 1, 4), (2, 4, 1, 3, 2, 4), (3, 4, 1, 2, 3, 4)]
    with open(os.path.join(d, "CELLParStruct.txt"), "w") as f:
        f.write("header\n" + col(1) + "cell 0\n" + col(1, 4, 4, 6) + "xyz\n")
        for p in [(5, 5, 5), (6, 5, 5), (5, 6, 5), (5, 5, 6)]:
            f.write(col(*p))
        f.write("tri\n")
        for t in tet:
            f.write(col(*t))
        f.write("edges\n")
        for e in edges:
            f.write(col(*e))
    with open(os.path.join(d, "CELLPhyPara.txt"), "w") as f:
        f.write("header\n" + col(1) + "ShearMod\n" + col(6.3) + "BendingMod\n" + col(0.2)
                + "HGarea\n" + col(4900) + "HLarea\n" + col(100) + "HVol\n" + col(5000)
                + "MaxCut\n" + col(0.9) + "zeroLen\n" + col(0.3) + "scaFac\n" + col(1.5)
                + "surEnergy\n" + col(0.01))


def arr(ptr, ctype, n, shape):
    if n == 0:
        return np.zeros(shape)
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(n * int(np.prod(shape[1:])),)).reshape(shape).copy()


def read_case(path):
    lib = abi.load()
    d = abi.gc_case_desc()
    abi.check(lib.gc_ReadCase(path.encode(), C.byref(d)), "ReadCase")
    N = d.consts.nTotalPar
    out = {
        "counts": (d.nBound1Par, d.nBound2Par, d.nFluidPar, d.nCell, d.nAllTri, d.nAllEdge, d.nCellPar),
        "coor": arr(d.coor, C.c_float, N, (N, 3)),
        "tri": arr(d.tri, C.c_int, d.nAllTri, (d.nAllTri, 3)),
        "edge": arr(d.edge, C.c_int, d.nAllEdge, (d.nAllEdge, 6)),
        "celltab": arr(d.celltab, C.c_int, d.nCellPar, (d.nCellPar,)),
        "bet": arr(d.beginendtri, C.c_int, d.nCell, (d.nCell, 2)),
        "mu": arr(d.shearmod_para, C.c_float, d.nAllEdge, (d.nAllEdge,)),
        "kc": arr(d.bendingmod_para, C.c_float, d.nAllEdge, (d.nAllEdge,)),
        "kd": arr(d.hlarea_para, C.c_float, d.nAllTri, (d.nAllTri,)),
        "ka": arr(d.hgarea_para, C.c_float, d.nAllTri, (d.nAllTri,)),
        "kv": arr(d.hvol_para, C.c_float, d.nAllTri, (d.nAllTri,)),
        "consts": abi.Structconst.from_buffer_copy(d.consts),
        "bbox": list(d.bbox),
    }
    lib.gc_FreeCase(C.byref(d))
    return out


def test_reads_reference_format(tmp_path):
    write_tiny_deck(str(tmp_path))
    r = read_case(str(tmp_path))
    assert r["counts"] == (0, 2, 3, 1, 4, 6, 4)
    assert np.allclose(r["coor"][2], [1.0, 0.0, 0.0]) and np.allclose(r["coor"][4], [3.25, 2.5, 4.0])
    assert np.allclose(r["coor"][5], [5, 5, 5])          # CoorMerge: membrane after SDPD
    c = r["consts"]
    # LoadSDPDPara + ConstantDataUp + AllocCpuMemory (Particle.cu:124-172,432-437,656-660)
    assert (c.SoundSpeed1, c.SoundSpeed2) == (10.0, 12.0)
    assert c.ShearVis == pytest.approx(0.5) and c.BulkVis == pytest.approx(0.1)
    assert c.ConstExtForce == pytest.approx(0.25 * c.Mass)
    assert c.VVsigma == pytest.approx(0.65) and c.iSeed == 1 and c.TimeStep == pytest.approx(1e-3)
    assert (c.nBound2Start, c.nFluidStart, c.nCellStart, c.nTotalPar, c.nCellPar) == (0, 2, 5, 9, 4)
    assert r["bbox"] == [0, 20, -10, 10, -10, 10]
    # ids stay 1-based with the per-cell offsets (Particle.cu:252,268-284)
    assert r["tri"][0].tolist() == [1, 3, 2]
    assert r["edge"][0].tolist() == [1, 2, 3, 4, 1, 2]
    assert r["bet"].tolist() == [[0, 3]]
    # per-type -> per-edge / per-triangle broadcast (Particle.cu:405-425)
    assert np.allclose(r["mu"], 6.3) and np.allclose(r["kc"], 0.2)
    assert np.allclose(r["ka"], 4900) and np.allclose(r["kd"], 100) and np.allclose(r["kv"], 5000)
    assert (c.MaxCut, c.zeroLen, c.scaFac) == pytest.approx((0.9, 0.3, 1.5))
    assert c.surEnergy == pytest.approx(0.01)


def test_write_read_round_trip(tmp_path):
    case = casegen.channel(24, 12, 3, rbc_diameter=5.0, subdiv=2, seed=4)
    d, keep = case.desc()
    lib = abi.load()
    abi.check(lib.gc_WriteCase(str(tmp_path).encode(), C.byref(d)), "WriteCase")
    for name in ("SDPDParCoor.txt", "SDPDPhyPara.txt", "CELLParStruct.txt", "CELLPhyPara.txt"):
        lines = open(os.path.join(tmp_path, name)).read().splitlines()
        # every numeric field leaves blanks inside its 20 columns
        for ln in lines[1:40]:
            if ln and ln[0] in "-0123456789":
                assert all(len(ln) <= 20 * k or ln[20 * k - 1] == " " for k in range(1, 7))
    r = read_case(str(tmp_path))
    assert r["counts"] == (case.nBound1, case.nBound2, case.nFluid, case.nCell, case.tri.shape[0],
                           case.edge.shape[0], case.nCellPar)
    assert np.array_equal(r["coor"], case.coor)   # %.9g round-trips float32 exactly
    assert np.array_equal(r["tri"], case.tri) and np.array_equal(r["edge"], case.edge)
    assert np.array_equal(r["celltab"], case.celltab)
    assert np.array_equal(r["bet"], case.beginendtri)
    for k, a in (("mu", case.shearmod), ("kc", case.bendmod), ("kd", case.hlarea),
                 ("ka", case.hgarea), ("kv", case.hvol)):
        assert np.array_equal(r[k], a), k


def test_missing_deck_is_an_error(tmp_path):
    lib = abi.load()
    d = abi.gc_case_desc()
    assert lib.gc_ReadCase(str(tmp_path).encode(), C.byref(d)) == -1
    assert b"SDPDParCoor.txt" in lib.gc_GetLastError()
