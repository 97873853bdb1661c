"""ctypes view of libgpucell_hip.so (include/gpucell/types.h, gpucell.h).

The library is the product; this module only declares its structs and
prototypes.  Loading fails loudly when the built library is absent: there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpucell_hip.so")


class tint2(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int)]


class tint3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


class int6(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("x", "y", "z", "k", "v", "w")]


class tfloat3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class tfloat4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class Structconst(C.Structure):
    """src/TypesDef.cuh:49-78 (96 bytes)."""
    _fields_ = [
        ("nFluidStart", C.c_uint), ("nBound2Start", C.c_uint), ("nCellStart", C.c_uint),
        ("nTotalPar", C.c_uint), ("nCellPar", C.c_uint), ("iSeed", C.c_long),
        ("CutRadius", C.c_float), ("Mass", C.c_float), ("SoundSpeed1", C.c_float),
        ("SoundSpeed2", C.c_float), ("DenInit", C.c_float), ("VVsigma", C.c_float),
        ("ShearVis", C.c_float), ("BulkVis", C.c_float), ("ConstExtForce", C.c_float),
        ("TimeStep", C.c_float), ("surEnergy", C.c_float), ("MaxCut", C.c_float),
        ("zeroLen", C.c_float), ("scaFac", C.c_float), ("Temp", C.c_float),
    ]


class gc_ext_consts(C.Structure):
    _fields_ = [
        ("periodic", C.c_int * 3), ("force_mode", C.c_int), ("force_split", C.c_float),
        ("noise_on", C.c_int), ("wlc_ratio", C.c_float), ("pow_index", C.c_float),
        ("rep_scale", C.c_float), ("wall_split", C.c_float), ("wall_u_lo", C.c_float),
        ("wall_u_hi", C.c_float), ("noise_counter", C.c_uint64),
    ]


class gc_case_desc(C.Structure):
    _fields_ = [
        ("nBound1Par", C.c_int), ("nBound2Par", C.c_int), ("nFluidPar", C.c_int),
        ("nCell", C.c_int), ("nAllTri", C.c_int), ("nAllEdge", C.c_int), ("nCellPar", C.c_int),
        ("coor", C.c_void_p), ("vel", C.c_void_p), ("tri", C.c_void_p), ("edge", C.c_void_p),
        ("celltab", C.c_void_p), ("beginendtri", C.c_void_p), ("shearmod_para", C.c_void_p),
        ("bendingmod_para", C.c_void_p), ("hlarea_para", C.c_void_p),
        ("hgarea_para", C.c_void_p), ("hvol_para", C.c_void_p), ("bbox", C.c_float * 6),
        ("consts", Structconst), ("ext", gc_ext_consts), ("skin", C.c_float),
    ]


class gc_engine_info(C.Structure):
    _fields_ = [
        ("step", C.c_long), ("noise_counter", C.c_uint64), ("nTotalPar", C.c_int),
        ("nCellPar", C.c_int), ("nCell", C.c_int), ("sizecube", C.c_int),
        ("ncube", C.c_int * 3), ("outerregion", C.c_float * 3), ("invwid", C.c_float * 3),
        ("fault", C.c_int), ("nl_overflow", C.c_int), ("nrebuild", C.c_int), ("skin", C.c_float),
    ]


class gc_host_state(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "pos", "coor", "velrhop", "velhalf", "force", "press", "nthrobound", "area", "vol",
        "defforce", "aggforce")]


class gc_host_cells(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("sortpart", "cubepart", "cc", "curpos", "beginendcube")]


class gc_membrane_stats(C.Structure):
    _fields_ = [("steps", C.c_long), ("fallbacks", C.c_long), ("samples", C.c_long),
                ("last_residual", C.c_float), ("plain_residual", C.c_float)]


class gc_tune_report(C.Structure):
    _fields_ = [("n", C.c_int), ("family", (C.c_char * 24) * 64), ("variant", C.c_int * 64),
                ("shape", (C.c_char * 32) * 64), ("us", C.c_float * 64), ("chosen", C.c_int * 64)]


class gc_engine_profile(C.Structure):
    _fields_ = [("nkernels", C.c_int), ("names", (C.c_char * 32) * 16), ("us", C.c_float * 16)]


assert C.sizeof(Structconst) == 96
assert C.sizeof(gc_ext_consts) == 56

P = C.c_void_p
I = C.c_int
U = C.c_uint
L = C.c_long

# name -> (restype, argtypes); the reference entry points first
PROTOTYPES = {
    "gc_GetLastError": (C.c_char_p, []),
    "gc_Version": (C.c_char_p, []),
    "gc_SetStream": (I, [P]),
    "gc_Synchronize": (I, []),
    "gc_ConstsDataUp": (I, [C.POINTER(Structconst)]),
    "gc_DefaultExtConsts": (None, [C.POINTER(gc_ext_consts)]),
    "gc_SetExtConsts": (I, [C.POINTER(gc_ext_consts)]),
    "gc_GetExtConsts": (I, [C.POINTER(gc_ext_consts)]),
    "gc_ConfigDomain": (I, [P, C.c_float, P, P, P, P, P]),
    "gc_PreConList": (I, [I, I, P, P, P, P, P, P, P, P]),
    "gc_Sort": (I, [I, P, P]),
    "gc_SortNoAlloc": (I, [I, P, P]),
    "gc_CalBeginEndCube": (I, [I, I, P, P]),
    "gc_SortBasicArrays": (I, [I, P, P, P, P, P, P, P, P]),
    "gc_SortArraysWithCells": (I, [I, P, P, P, P, P, P, P, P, P, P]),
    "gc_UpdateFluidCoor": (I, [I, P, P, P, P]),
    "gc_AppFluidPeriodicBC": (I, [I, P, P, P, P]),
    "gc_CalFluidDen": (I, [I, P, P, P, P, P, P, P]),
    "gc_PredictFluidVel": (I, [I, P, P, P, P]),
    "gc_CalBound2Vel": (I, [I, P, P, P, P, P, P, P]),
    "gc_EOS": (I, [I, P, P, P]),
    "gc_CalFluidSDPDForce": (I, [I, P, P, P, P, P, P, P, P, P, P]),
    "gc_CalFluidRepForce": (I, [I, P, P, P, P, P, P, P]),
    "gc_SetFluidExtForce": (I, [I, I, P, P]),
    "gc_SetFluidExtForcePos": (I, [I, I, P, P, P]),
    "gc_CorrectFliudVel": (I, [I, P, P, P, P, P, P, P, P, P]),
    "gc_CalGeoPara": (I, [L, I, I] + [P] * 17),
    "gc_SetMechPara": (I, [I, I] + [P] * 14),
    "gc_CalCellDefForce": (I, [I, U, U, U, U] + [P] * 27),
    "gc_CalCellAggForce": (I, [U, U, U] + [P] * 10),
    "gc_InvalidateTopology": (I, []),
    "gc_EngineCreate": (P, [C.POINTER(gc_case_desc), C.POINTER(C.c_int)]),
    "gc_EngineDestroy": (None, [P]),
    "gc_EngineStep": (I, [P, I, I]),
    "gc_EngineStepGraph": (I, [P, I, I]),
    "gc_EngineInfo": (I, [P, C.POINTER(gc_engine_info)]),
    "gc_CheckSlotMaps": (I, [I, I, I, I, P, P, P]),
    "gc_EngineCheckSlotMaps": (I, [P]),
    "gc_EngineGetState": (I, [P, C.POINTER(gc_host_state)]),
    "gc_EngineSetState": (I, [P, C.POINTER(gc_host_state), L, C.c_uint64]),
    "gc_EngineGetCells": (I, [P, C.POINTER(gc_host_cells)]),
    "gc_EngineProfile": (I, [P, I, I, C.POINTER(gc_engine_profile)]),
    "gc_EngineGetTriangles": (I, [P, P, P]),
    "gc_EngineRanges": (I, [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "gc_ReadCase": (I, [C.c_char_p, C.POINTER(gc_case_desc)]),
    "gc_FreeCase": (None, [C.POINTER(gc_case_desc)]),
    "gc_WriteCase": (I, [C.c_char_p, C.POINTER(gc_case_desc)]),
    "gc_EngineWriteTecplot": (I, [P, C.c_char_p, L, L, C.c_float, L, P, P, P]),
    "gc_EngineSave": (I, [P, C.c_char_p]),
    "gc_EngineLoad": (I, [P, C.c_char_p]),
    "gc_EngineAutotune": (I, [P, I, I, C.POINTER(gc_tune_report)]),
    "gc_EngineGetTuning": (I, [P, P]),
    "gc_EngineSetTuning": (I, [P, P]),
    "gc_EnsembleCreate": (P, [C.POINTER(gc_case_desc), I, C.POINTER(C.c_int)]),
    "gc_EnsembleGetState": (I, [P, I, C.POINTER(gc_host_state)]),
    "gc_EnsembleCount": (I, [P]),
    # learned membrane substepper (gpucell/surrogate.h)
    "gc_MlpParamCount": (I, []),
    "gc_MlpCreate": (P, [C.c_uint64, C.POINTER(C.c_int)]),
    "gc_MlpDestroy": (None, [P]),
    "gc_MlpSetParams": (I, [P, P, I]),
    "gc_MlpGetParams": (I, [P, P, I]),
    "gc_MlpForward": (I, [P, P, P, I]),
    "gc_MlpGradient": (I, [P, P, P, I, P]),
    "gc_MlpTrain": (I, [P, P, P, I, I, I, C.c_float, C.c_uint64, P]),
    "gc_EngineSetMembraneMode": (I, [P, I, I, P, P, C.c_float]),
    "gc_EngineRecordSamples": (I, [P, P, P, I]),
    "gc_EngineMembraneStats": (I, [P, C.POINTER(gc_membrane_stats)]),
}

# Entry points that replace the reference's cupar/cudiv functions.
REFERENCE_ENTRY_POINTS = [
    "gc_ConstsDataUp", "gc_PreConList", "gc_Sort", "gc_CalBeginEndCube", "gc_SortBasicArrays",
    "gc_SortArraysWithCells", "gc_UpdateFluidCoor", "gc_AppFluidPeriodicBC", "gc_CalFluidDen",
    "gc_PredictFluidVel", "gc_CalBound2Vel", "gc_EOS", "gc_CalFluidSDPDForce",
    "gc_CalFluidRepForce", "gc_SetFluidExtForce", "gc_CorrectFliudVel", "gc_CalGeoPara",
    "gc_SetMechPara", "gc_CalCellDefForce", "gc_CalCellAggForce",
]

_lib = None

# status codes of gpucell.h
GC_OK, GC_ERR_INVALID, GC_ERR_NOCONSTS, GC_ERR_GRID = 0, -1, -2, -3
GC_ERR_NOMEM, GC_ERR_TOPOLOGY, GC_ERR_HIP, GC_ERR_NONFINITE = -4, -5, -6, -7


class GcError(RuntimeError):
    pass


def load(path: str | None = None) -> C.CDLL:
    """Loads the built library; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise GcError(f"{p} is not built: run `python -m gpucell_amd.build` "
                      "(there is no CPU fallback)")
    lib = C.CDLL(p)
    for name, (res, args) in PROTOTYPES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = load().gc_GetLastError().decode()
        raise GcError(f"{what or 'gpucell call'} failed with status {status}: {msg}")


def default_ext() -> gc_ext_consts:
    e = gc_ext_consts()
    load().gc_DefaultExtConsts(C.byref(e))
    return e


def ptr(a) -> int | None:
    """Address of a numpy array or torch tensor (None stays None)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"], "arrays handed to the library must be contiguous"
        return a.ctypes.data
    return a.data_ptr()
