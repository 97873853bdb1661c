"""ctypes binding of oracle/liboracle.so, the FP64 checker.

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and
by the CPU-baseline leg of bench.py.  State arrays are id-ordered float64.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_F = C.POINTER(C.c_float)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        lib = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "orc_philox4x32_10": (None, [vp, vp, vp]),
            "orc_pair_normals": (None, [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, vp]),
            "orc_lucy_w": (C.c_double, [C.c_double, C.c_double]),
            "orc_lucy_f": (C.c_double, [C.c_double, C.c_double]),
            "orc_config_domain": (C.c_int, [vp, C.c_float, vp, vp, vp, vp, vp]),
            "orc_preconlist": (None, [C.c_int, C.c_float, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            "orc_sort": (None, [C.c_int, vp, vp]),
            "orc_begin_end": (None, [C.c_int, C.c_int, vp, vp]),
            "orc_create": (vp, [vp, vp, vp, C.c_int, C.c_int, C.c_int] + [vp] * 9),
            "orc_destroy": (None, [vp]),
            "orc_set_threads": (None, [C.c_int]),
            "orc_update_coor": (None, [vp, vp, vp, vp]),
            "orc_periodic_bc": (None, [vp, vp, vp]),
            "orc_density": (None, [vp, vp, vp]),
            "orc_predict": (None, [vp, vp, vp, vp]),
            "orc_bound2_vel": (None, [vp, vp, vp]),
            "orc_eos": (None, [vp, vp, vp]),
            "orc_sdpd_force": (None, [vp, vp, vp, vp, vp, C.c_uint64, vp, vp]),
            "orc_rep_force": (None, [vp, vp, vp, vp]),
            "orc_ext_force": (None, [vp, vp, C.c_int, vp, vp]),
            "orc_correct": (None, [vp, vp, vp, vp]),
            "orc_geo": (None, [vp, C.c_long, vp, vp, vp, vp, vp, vp, vp, vp]),
            "orc_mech_para": (None, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            "orc_def_force": (None, [vp, vp, vp, vp, vp, vp]),
            "orc_membrane_energy": (C.c_double, [vp, vp, vp]),
            "orc_agg_force": (None, [vp, vp, vp, vp, vp]),
            "orc_morse_energy": (C.c_double, [vp, vp]),
            "orc_step": (None, [vp, vp, C.c_int]),
        }
        for k, (r, a) in sig.items():
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a
        _lib = lib
    return _lib


def _p(a):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("oracle arrays must be C-contiguous (got a strided view)")
    return a.ctypes.data


class OrcState(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("pos", "img", "vel", "rho", "velhalf", "force", "press",
                                          "defforce", "aggforce", "cell_area", "cell_vol",
                                          "abs_sum")] + [("step", C.c_long),
                                                         ("counter", C.c_uint64)]


def philox(ctr, key):
    lib = load()
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib.orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def pair_normals(seed, counter, a, b):
    o = np.zeros(3)
    load().orc_pair_normals(seed, counter, a, b, _p(o))
    return o


def lucy_w(r, h):
    return load().orc_lucy_w(r, h)


def lucy_f(r, h):
    return load().orc_lucy_f(r, h)


def config_domain(bbox, cut, periodic):
    bb = np.asarray(bbox, np.float32)
    per = np.asarray(periodic, np.int32)
    nc = np.zeros(3, np.int32)
    out = np.zeros(3, np.float32)
    inv = np.zeros(3, np.float32)
    K = np.zeros(1, np.int32)
    rc = load().orc_config_domain(_p(bb), float(cut), _p(per), _p(nc), _p(out), _p(inv), _p(K))
    if rc:
        raise ValueError("grid unusable")
    return nc, out, inv, int(K[0])


def preconlist(coor32, cut, periodic, bbox, invwid, ncube, sortpart):
    n = coor32.shape[0]
    co = np.ascontiguousarray(coor32, np.float32)
    cc = np.zeros((n, 3), np.int32)
    key = np.zeros(n, np.int32)
    sp = np.ascontiguousarray(sortpart, np.int32)
    old = np.zeros(n, np.int32)
    load().orc_preconlist(n, float(cut), _p(np.asarray(periodic, np.int32)), _p(co),
                          _p(np.asarray(bbox, np.float32)), _p(np.asarray(invwid, np.float32)),
                          _p(np.asarray(ncube, np.int32)), _p(cc), _p(key), _p(sp), _p(old))
    return cc, key, old


def sort(keys, values):
    k = np.array(keys, np.int32)
    v = np.array(values, np.int32)
    load().orc_sort(k.shape[0], _p(k), _p(v))
    return k, v


def begin_end(keys, K):
    k = np.ascontiguousarray(keys, np.int32)
    be = np.zeros((K, 2), np.int32)
    load().orc_begin_end(k.shape[0], K, _p(k), _p(be))
    return be


class Oracle:
    """One orc_sys built from a case (gpucell_amd.casegen.Case)."""

    def __init__(self, case, consts=None, ext=None, threads: int = 0):
        lib = load()
        self.case = case
        self.consts = consts if consts is not None else case.consts
        self.ext = ext if ext is not None else case.ext
        self._keep = []

        def arr(a, dt):
            if a is None or a.size == 0:
                return None
            b = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(b)
            return b.ctypes.data

        bb = np.ascontiguousarray(case.bbox, np.float32)
        self._keep.append(bb)
        self.s = lib.orc_create(C.byref(self.consts), C.byref(self.ext), bb.ctypes.data,
                                case.nCell, int(case.tri.shape[0]), int(case.edge.shape[0]),
                                arr(case.tri, np.int32), arr(case.edge, np.int32),
                                arr(case.celltab, np.int32), arr(case.beginendtri, np.int32),
                                arr(case.shearmod, np.float32), arr(case.bendmod, np.float32),
                                arr(case.hlarea, np.float32), arr(case.hgarea, np.float32),
                                arr(case.hvol, np.float32))
        if threads:
            lib.orc_set_threads(threads)
        self.N = int(self.consts.nTotalPar)
        self.nm = int(self.consts.nCellPar)
        self.ncell = case.nCell

    def __del__(self):
        try:
            if getattr(self, "s", None):
                load().orc_destroy(self.s)
        except Exception:
            pass

    # --- state -------------------------------------------------------
    def initial_state(self):
        """SetParticle (Particle.cu:485-546) + reference geometry/mechanics."""
        c = self.case
        N = self.N
        st = {
            "pos": c.coor.astype(np.float64).copy(),
            "img": np.zeros((N, 3), np.int32),
            "vel": c.vel.astype(np.float64).copy(),
            "rho": np.full(N, float(self.consts.DenInit)),
            "force": np.zeros((N, 3)),
            "press": np.zeros(N),
            "abs_sum": np.zeros(N),
            "step": 0,
            "counter": int(self.ext.noise_counter),
        }
        st["velhalf"] = st["vel"].copy()
        self.eos(st["rho"], st["press"])
        if self.ncell:
            st["defforce"] = np.zeros((self.nm, 3))
            st["aggforce"] = np.zeros((self.nm, 3))
            st["cell_area"] = np.zeros(self.ncell)
            st["cell_vol"] = np.zeros(self.ncell)
            self.geo(-1, st["pos"], st["img"])
            self.mech_para(st["pos"], st["img"])
        return st

    def _orcstate(self, st):
        s = OrcState()
        for k in ("pos", "img", "vel", "rho", "velhalf", "force", "press", "defforce",
                  "aggforce", "cell_area", "cell_vol", "abs_sum"):
            a = st.get(k)
            setattr(s, k, None if a is None else a.ctypes.data)
        s.step = int(st["step"])
        s.counter = int(st["counter"])
        return s

    def step(self, st, driveon=1):
        for k in ("pos", "vel", "velhalf", "force", "rho", "press", "abs_sum"):
            assert st[k].flags["C_CONTIGUOUS"] and st[k].dtype == np.float64
        s = self._orcstate(st)
        load().orc_step(self.s, C.byref(s), int(driveon))
        st["step"] = int(s.step)
        st["counter"] = int(s.counter)
        return st

    # --- single functions ---------------------------------------------
    def update_coor(self, pos, vel, force):
        load().orc_update_coor(self.s, _p(pos), _p(vel), _p(force))

    def periodic_bc(self, pos, img):
        load().orc_periodic_bc(self.s, _p(pos), _p(img))

    def density(self, pos, rho):
        load().orc_density(self.s, _p(pos), _p(rho))

    def predict(self, vel, velhalf, force):
        load().orc_predict(self.s, _p(vel), _p(velhalf), _p(force))

    def bound2_vel(self, pos, vel):
        load().orc_bound2_vel(self.s, _p(pos), _p(vel))

    def eos(self, rho, press):
        load().orc_eos(self.s, _p(rho), _p(press))

    def sdpd_force(self, pos, vel, rho, press, counter, force, abs_sum=None):
        load().orc_sdpd_force(self.s, _p(pos), _p(vel), _p(rho), _p(press), int(counter),
                              _p(force), _p(abs_sum))

    def rep_force(self, pos, force, abs_sum=None):
        load().orc_rep_force(self.s, _p(pos), _p(force), _p(abs_sum))

    def ext_force(self, pos, driveon, force, abs_sum=None):
        load().orc_ext_force(self.s, _p(pos), int(driveon), _p(force), _p(abs_sum))

    def correct(self, vel, velhalf, force):
        load().orc_correct(self.s, _p(vel), _p(velhalf), _p(force))

    def geo(self, stepcount, pos, img):
        T, Cn, M = int(self.case.tri.shape[0]), self.ncell, self.nm
        out = {"tri_nor": np.zeros((T, 3)), "tri_area": np.zeros(T), "tri_cen": np.zeros((T, 3)),
               "cell_area": np.zeros(Cn), "cell_vol": np.zeros(Cn), "par_area": np.zeros(M)}
        load().orc_geo(self.s, int(stepcount), _p(pos), _p(img), _p(out["tri_nor"]),
                       _p(out["tri_area"]), _p(out["tri_cen"]), _p(out["cell_area"]),
                       _p(out["cell_vol"]), _p(out["par_area"]))
        return out

    def mech_para(self, pos, img):
        E = int(self.case.edge.shape[0])
        out = {k: np.zeros(E) for k in ("l0", "lmax", "kb", "mexp", "p", "kp", "theta0")}
        load().orc_mech_para(self.s, _p(pos), _p(img), *[_p(out[k]) for k in
                                                         ("l0", "lmax", "kb", "mexp", "p", "kp", "theta0")])
        return out

    def def_force(self, pos, img, force=None, abs_sum=None):
        df = np.zeros((self.nm, 3))
        load().orc_def_force(self.s, _p(pos), _p(img), _p(df), _p(force), _p(abs_sum))
        return df

    def membrane_energy(self, pos, img):
        return load().orc_membrane_energy(self.s, _p(pos), _p(img))

    def agg_force(self, pos, force=None, abs_sum=None):
        af = np.zeros((self.nm, 3))
        load().orc_agg_force(self.s, _p(pos), _p(af), _p(force), _p(abs_sum))
        return af

    def morse_energy(self, pos):
        return load().orc_morse_energy(self.s, _p(pos))
