"""Python handle on the C++ engine (gc_Engine* in libgpucell_hip.so).

Mirrors the reference's host class (class Particle, src/Particle.cuh:8-167):
create = InitPar/SetPar, step = DrivePar + UpdatePar, state = the arrays
ParticlesDataReturn copies back (un-sorted to id order).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


def _state_arrays(case) -> dict:
    N, M, Cn = case.N, case.nCellPar, case.nCell
    out = {
        "pos": np.zeros((N, 3), np.float64), "coor": np.zeros((N, 3), np.float32),
        "velrhop": np.zeros((N, 4), np.float32), "velhalf": np.zeros((N, 3), np.float32),
        "force": np.zeros((N, 3), np.float32), "press": np.zeros(N, np.float32),
        "nthrobound": np.zeros((N, 3), np.int32),
    }
    if Cn:
        out.update(area=np.zeros(Cn, np.float32), vol=np.zeros(Cn, np.float32),
                   defforce=np.zeros((M, 3), np.float32), aggforce=np.zeros((M, 3), np.float32))
    return out


def _host_state(out: dict) -> abi.gc_host_state:
    hs = abi.gc_host_state()
    for k, v in out.items():
        setattr(hs, k, v.ctypes.data)
    return hs


class Engine:
    """skin > 0: Verlet lists with that skin, rebuilt (with the re-sort)
    only when the particles have moved far enough (gc_case_desc.skin);
    0: the reference's per-step neighbour build."""

    def __init__(self, case, skin: float = 0.0):
        self.lib = abi.load()
        self.case = case
        desc, self._keep = case.desc()
        desc.skin = float(skin)
        self.skin = float(skin)
        st = C.c_int(0)
        self.h = self.lib.gc_EngineCreate(C.byref(desc), C.byref(st))
        if not self.h:
            raise abi.GcError(f"gc_EngineCreate failed ({st.value}): "
                              f"{self.lib.gc_GetLastError().decode()}")
        self.N = case.N
        self.nm = case.nCellPar
        self.ncell = case.nCell

    def close(self):
        if getattr(self, "h", None):
            self.lib.gc_EngineDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1, driveon: int = 1, graph: bool = False):
        f = self.lib.gc_EngineStepGraph if graph else self.lib.gc_EngineStep
        abi.check(f(self.h, int(n), int(driveon)), "EngineStep")

    def info(self) -> abi.gc_engine_info:
        i = abi.gc_engine_info()
        abi.check(self.lib.gc_EngineInfo(self.h, C.byref(i)), "EngineInfo")
        return i

    def check_slot_maps(self):
        """Raises GcError if the slot maps (sp, CurPos, act) are inconsistent."""
        abi.check(self.lib.gc_EngineCheckSlotMaps(self.h), "EngineCheckSlotMaps")

    def state(self) -> dict:
        out = _state_arrays(self.case)
        hs = _host_state(out)
        abi.check(self.lib.gc_EngineGetState(self.h, C.byref(hs)), "EngineGetState")
        return out

    def set_state(self, st: dict, step: int, counter: int):
        hs = abi.gc_host_state()
        keep = []
        for k, dt in (("pos", np.float64), ("velrhop", np.float32), ("velhalf", np.float32),
                      ("force", np.float32), ("press", np.float32), ("nthrobound", np.int32)):
            if k in st and st[k] is not None:
                a = np.ascontiguousarray(st[k], dtype=dt)
                keep.append(a)
                setattr(hs, k, a.ctypes.data)
        abi.check(self.lib.gc_EngineSetState(self.h, C.byref(hs), int(step), int(counter)),
                  "EngineSetState")

    def save(self, path: str):
        """Checkpoint (bitwise-resumable) to `path`."""
        abi.check(self.lib.gc_EngineSave(self.h, str(path).encode()), "EngineSave")

    def load(self, path: str):
        abi.check(self.lib.gc_EngineLoad(self.h, str(path).encode()), "EngineLoad")

    def write_tecplot(self, directory: str, stepdata: int, stepcount: int, stepoutput: int):
        """One frame of the reference's Tecplot output under directory/Result."""
        c = self.case
        self._topo = (np.ascontiguousarray(c.tri, np.int32),
                      np.ascontiguousarray(c.beginendtri, np.int32),
                      np.ascontiguousarray(c.celltab, np.int32))
        t, b, ct = self._topo
        abi.check(self.lib.gc_EngineWriteTecplot(
            self.h, str(directory).encode(), int(stepdata), int(stepcount),
            float(c.consts.TimeStep), int(stepoutput), abi.ptr(t) if t.size else None,
            abi.ptr(b) if b.size else None, abi.ptr(ct) if ct.size else None), "EngineWriteTecplot")

    def cells(self) -> dict:
        i = self.info()
        N, K = self.N, i.sizecube
        out = {"sortpart": np.zeros(N, np.int32), "cubepart": np.zeros(N, np.int32),
               "cc": np.zeros((N, 3), np.int32), "curpos": np.zeros(N, np.int32),
               "beginendcube": np.zeros((K, 2), np.int32)}
        hc = abi.gc_host_cells()
        for k, v in out.items():
            setattr(hc, k, v.ctypes.data)
        abi.check(self.lib.gc_EngineGetCells(self.h, C.byref(hc)), "EngineGetCells")
        return out

    def profile(self, nsteps: int = 10, driveon: int = 1) -> dict:
        p = abi.gc_engine_profile()
        abi.check(self.lib.gc_EngineProfile(self.h, int(nsteps), int(driveon), C.byref(p)),
                  "EngineProfile")
        return {p.names[k].value.decode(): float(p.us[k]) for k in range(p.nkernels)}

    # --- autotuner ---------------------------------------------------------
    def autotune(self, steps: int = 5, driveon: int = 1) -> list:
        """Times every launch-shape variant, keeps the fastest per kernel
        family and restores the state; returns the timing table."""
        r = abi.gc_tune_report()
        abi.check(self.lib.gc_EngineAutotune(self.h, int(steps), int(driveon), C.byref(r)),
                  "EngineAutotune")
        return [{"family": r.family[k].value.decode(), "variant": r.variant[k],
                 "shape": r.shape[k].value.decode(), "us": float(r.us[k]), "chosen": bool(r.chosen[k])}
                for k in range(r.n)]

    def tuning(self) -> list:
        v = (C.c_int * 4)()
        abi.check(self.lib.gc_EngineGetTuning(self.h, v), "EngineGetTuning")
        return list(v)

    def set_tuning(self, v):
        a = (C.c_int * 4)(*[int(x) for x in v])
        abi.check(self.lib.gc_EngineSetTuning(self.h, a), "EngineSetTuning")

    # --- membrane time stepping (include/gpucell/surrogate.h) ------------
    def set_membrane_mode(self, mode: int, k: int = 8, mlp=None, norm=None, threshold: float = 0.0):
        """0: membrane moves with the fluid step; 1: k plain substeps;
        2: the MLP (gpucell_amd.surrogate.Mlp) predicts the 8 substeps, with
        fallback to plain substepping when the residual exceeds threshold."""
        self._norm = None if norm is None else np.ascontiguousarray(norm, np.float32)
        abi.check(self.lib.gc_EngineSetMembraneMode(
            self.h, int(mode), int(k), mlp.h if mlp is not None else None,
            self._norm.ctypes.data if self._norm is not None else None, C.c_float(threshold)),
            "EngineSetMembraneMode")

    def record_samples(self, x=None, y=None):
        """Record mode-1 training rows into CUDA tensors x [n][36], y [n][24]
        (None stops)."""
        self._rec = (x, y)
        abi.check(self.lib.gc_EngineRecordSamples(
            self.h, x.data_ptr() if x is not None else None, y.data_ptr() if y is not None else None,
            int(x.shape[0]) if x is not None else 0), "EngineRecordSamples")

    def membrane_stats(self) -> dict:
        m = abi.gc_membrane_stats()
        abi.check(self.lib.gc_EngineMembraneStats(self.h, C.byref(m)), "EngineMembraneStats")
        return {"steps": m.steps, "fallbacks": m.fallbacks, "samples": m.samples,
                "last_residual": m.last_residual, "plain_residual": m.plain_residual}


class Ensemble:
    """Independent suspensions stepped together by one engine
    (gc_EnsembleCreate); replica r evolves bit for bit as Engine(cases[r])."""

    def __init__(self, cases, skin: float = 0.0):
        self.lib = abi.load()
        self.cases = list(cases)
        self.skin = float(skin)
        descs = (abi.gc_case_desc * len(self.cases))()
        self._keep = []
        for r, c in enumerate(self.cases):
            d, keep = c.desc()
            d.skin = float(skin)
            descs[r] = d
            self._keep.append(keep)
        st = C.c_int(0)
        self.h = self.lib.gc_EnsembleCreate(descs, len(self.cases), C.byref(st))
        if not self.h:
            raise abi.GcError(f"gc_EnsembleCreate failed ({st.value}): "
                              f"{self.lib.gc_GetLastError().decode()}")
        self.nrep = self.lib.gc_EnsembleCount(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.gc_EngineDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1, driveon: int = 1, graph: bool = False):
        f = self.lib.gc_EngineStepGraph if graph else self.lib.gc_EngineStep
        abi.check(f(self.h, int(n), int(driveon)), "EngineStep")

    def info(self) -> abi.gc_engine_info:
        i = abi.gc_engine_info()
        abi.check(self.lib.gc_EngineInfo(self.h, C.byref(i)), "EngineInfo")
        return i

    def check_slot_maps(self):
        """Raises GcError if the slot maps (sp, CurPos, act) are inconsistent."""
        abi.check(self.lib.gc_EngineCheckSlotMaps(self.h), "EngineCheckSlotMaps")

    def state(self, rep: int) -> dict:
        out = _state_arrays(self.cases[rep])
        hs = _host_state(out)
        abi.check(self.lib.gc_EnsembleGetState(self.h, int(rep), C.byref(hs)), "EnsembleGetState")
        return out

    def profile(self, nsteps: int = 10, driveon: int = 1) -> dict:
        p = abi.gc_engine_profile()
        abi.check(self.lib.gc_EngineProfile(self.h, int(nsteps), int(driveon), C.byref(p)),
                  "EngineProfile")
        return {p.names[k].value.decode(): float(p.us[k]) for k in range(p.nkernels)}

    def autotune(self, steps: int = 5, driveon: int = 1) -> list:
        return Engine.autotune(self, steps, driveon)
