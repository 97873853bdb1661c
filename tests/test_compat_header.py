"""The drop-in header include/gpucell/compat.hpp compiles against host code
typed like the reference's class Particle (member types of
src/Particle.cuh:72-116) and links against libgpucell_hip.so: every
cupar:: / cudiv:: wrapper is called with the argument types of its
reference call site (src/Particle.cu:476-479, 662, 668-677, 1331-1391).
Nothing is executed (no GPU here): the program returns before its first
call unless given an argument."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "gpucell_amd")

SRC = r"""
#include "compat.hpp"

// The device members of a Particle-like host class, with the reference's
// pointer types; null here (the calls below are compiled and linked, not run).
struct Host {
    float *d_OuterRegion = nullptr, *d_invWid = nullptr, *d_BBoxSDPD = nullptr, *d_Press = nullptr;
    int *d_nCube = nullptr, *d_CubePart = nullptr, *d_SortPart = nullptr, *d_OldPos = nullptr;
    int *d_CurPos = nullptr, *d_CellTab = nullptr;
    float3 *d_TotalCoor = nullptr, *d_VelHalf = nullptr, *d_TotalForce = nullptr;
    float3 *d_TriNorDir = nullptr, *d_TriCenter = nullptr, *d_DefForce = nullptr, *d_AggForce = nullptr;
    float4* d_Velrhop = nullptr;
    int2 *d_BeginEndCube = nullptr, *d_BeginEndTri = nullptr;
    int3 *d_CC = nullptr, *d_nThroBound = nullptr, *d_Tri = nullptr;
    int6* d_Edge = nullptr;
    float *d_TriArea = nullptr, *d_RefTriArea = nullptr, *d_RefTriAngle = nullptr, *d_RefEdgeLen = nullptr;
    float *d_MaxRefEdgeLen = nullptr, *d_BendingMod_Para = nullptr, *d_ShearMod_Para = nullptr;
    float *d_PerLenWLC = nullptr, *d_HPow = nullptr, *d_HLarea_Para = nullptr, *d_HGarea_Para = nullptr;
    float *d_HVol_Para = nullptr, *d_Vol = nullptr, *d_Area = nullptr, *d_Area_Para = nullptr;
    float *d_RefArea_Para = nullptr, *d_Vol_Para = nullptr, *d_RefVol_Para = nullptr;
    float *d_PowerIndex = nullptr, *d_HBending = nullptr, *d_ParArea = nullptr;
    tfloat3 *h_SDPDForce = nullptr, *h_DefForce = nullptr;
    Structconst consts{};
    int nCell = 0, nAllTri = 0, nAllEdge = 0, SizeCube = 0;
    long StepCount = -1;

    void set_par()
    {
        cupar::CalGeoPara(StepCount, nAllTri, nCell, d_BBoxSDPD, d_CurPos, d_Tri, d_BeginEndTri, d_nThroBound,
                          d_TotalCoor, d_TriNorDir, d_TriArea, d_RefTriArea, d_TriCenter, d_RefArea_Para,
                          d_RefVol_Para, d_Area_Para, d_Vol_Para, d_Area, d_Vol, d_ParArea);
        cupar::SetMechPara(nCell, nAllEdge, d_Edge, d_CurPos, d_RefEdgeLen, d_MaxRefEdgeLen, d_BendingMod_Para,
                           d_ShearMod_Para, d_HBending, d_PowerIndex, d_PerLenWLC, d_HPow, d_RefTriAngle,
                           d_TriNorDir, d_TriCenter, d_TotalCoor);
        cupar::ConstsDataUp(&consts);
    }
    void con_cube_list()
    {
        cudiv::PreConList(consts.nTotalPar, consts.CutRadius, d_TotalCoor, d_BBoxSDPD, d_invWid, d_nCube, d_CC,
                          d_CubePart, d_SortPart, d_OldPos);
        cudiv::Sort(consts.nTotalPar, d_CubePart, d_SortPart);
        cudiv::CalBeginEndCube(consts.nTotalPar, SizeCube, d_CubePart, d_BeginEndCube);
        if (nCell > 0)
            cudiv::SortArraysWithCells(consts.nTotalPar, d_SortPart, d_OldPos, d_CurPos, d_TotalCoor, d_CC,
                                       d_Velrhop, d_VelHalf, d_TotalForce, d_Press, d_nThroBound);
        else
            cudiv::SortBasicArrays(consts.nTotalPar, d_SortPart, d_OldPos, d_TotalCoor, d_CC, d_Velrhop,
                                   d_VelHalf, d_TotalForce, d_Press);
    }
    void update_par(int drive_on)
    {
        const int n = consts.nTotalPar;
        cupar::UpdateFluidCoor(n, d_SortPart, d_TotalForce, d_TotalCoor, d_Velrhop);
        cupar::AppFluidPeriodicBC(n, d_SortPart, d_TotalCoor, d_nThroBound, d_BBoxSDPD);
        con_cube_list();
        cupar::CalFluidDen(n, d_TotalCoor, d_Velrhop, d_CC, d_nCube, d_OuterRegion, d_SortPart, d_BeginEndCube);
        cupar::PredictFluidVel(n, d_SortPart, d_Velrhop, d_VelHalf, d_TotalForce);
        cupar::CalBound2Vel(n, d_TotalCoor, d_Velrhop, d_CC, d_nCube, d_OuterRegion, d_SortPart, d_BeginEndCube);
        cupar::EOS(n, d_SortPart, d_Velrhop, d_Press);
        cupar::CalFluidSDPDForce(n, h_SDPDForce, d_TotalCoor, d_Velrhop, d_CC, d_nCube, d_SortPart,
                                 d_OuterRegion, d_BeginEndCube, d_TotalForce, d_Press);
        cupar::CalFluidRepForce(n, d_TotalCoor, d_CC, d_nCube, d_SortPart, d_OuterRegion, d_BeginEndCube,
                                d_TotalForce);
        cupar::SetFluidExtForce(drive_on, n, d_SortPart, d_TotalForce);
        cupar::CalGeoPara(StepCount, nAllTri, nCell, d_BBoxSDPD, d_CurPos, d_Tri, d_BeginEndTri, d_nThroBound,
                          d_TotalCoor, d_TriNorDir, d_TriArea, d_RefTriArea, d_TriCenter, d_RefArea_Para,
                          d_RefVol_Para, d_Area_Para, d_Vol_Para, d_Area, d_Vol, d_ParArea);
        cupar::CalCellDefForce(nCell, nAllEdge, nAllTri, n, consts.nCellPar, d_BBoxSDPD, d_Edge, d_Tri,
                               d_TotalCoor, h_DefForce, d_DefForce, d_nThroBound, d_MaxRefEdgeLen,
                               d_PowerIndex, d_PerLenWLC, d_HPow, d_TriNorDir, d_TriCenter, d_RefTriAngle,
                               d_HBending, d_TriArea, d_RefTriArea, d_HLarea_Para, d_HGarea_Para,
                               d_Area_Para, d_RefArea_Para, d_HVol_Para, d_Vol_Para, d_RefVol_Para,
                               d_TotalForce, d_SortPart, d_CurPos);
        cupar::CalCellAggForce(nCell, n, consts.nCellPar, d_TotalCoor, d_AggForce, d_CC, d_nCube, d_SortPart,
                               d_CellTab, d_OuterRegion, d_BeginEndCube, d_TotalForce, d_ParArea);
        cupar::CorrectFliudVel(n, d_SortPart, d_TotalForce, d_VelHalf, d_Velrhop, d_TotalCoor, d_CC, d_nCube,
                               d_OuterRegion, d_BeginEndCube);
        cupar::CalBound2Vel(n, d_TotalCoor, d_Velrhop, d_CC, d_nCube, d_OuterRegion, d_SortPart, d_BeginEndCube);
    }
};

int main(int argc, char**)
{
    if (argc < 2) return 0;   // compiled and linked only
    Host h;
    h.set_par();
    h.update_par(1);
    return 0;
}
"""


def test_compat_header_compiles_and_links(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    lib = os.path.join(LIBDIR, "libgpucell_hip.so")
    assert os.path.exists(lib), "build libgpucell_hip.so first"
    src = tmp_path / "particle_like.cpp"
    src.write_text(SRC)
    exe = tmp_path / "particle_like"
    cmd = [hipcc, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include", "gpucell"),
           str(src), "-o", str(exe), "-L", LIBDIR, "-lgpucell_hip", "-Wl,--no-undefined",
           f"-Wl,-rpath,{LIBDIR}"]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert p.returncode == 0, p.stdout
    assert exe.exists()
    # every gc_ entry point the wrappers call is resolved from the library
    nm = subprocess.run(["nm", "-D", "--undefined-only", str(exe)], stdout=subprocess.PIPE, text=True).stdout
    called = {l.split()[-1] for l in nm.splitlines() if " gc_" in l}
    expect = {"gc_PreConList", "gc_Sort", "gc_CalBeginEndCube", "gc_SortBasicArrays", "gc_SortArraysWithCells",
              "gc_ConstsDataUp", "gc_UpdateFluidCoor", "gc_AppFluidPeriodicBC", "gc_CalFluidDen",
              "gc_PredictFluidVel", "gc_CalBound2Vel", "gc_EOS", "gc_CalFluidSDPDForce", "gc_CalFluidRepForce",
              "gc_SetFluidExtForce", "gc_CorrectFliudVel", "gc_CalGeoPara", "gc_SetMechPara",
              "gc_CalCellDefForce", "gc_CalCellAggForce", "gc_GetLastError"}
    assert expect <= called, expect - called
