"""GPU-side helpers for the parity tests: device buffers (torch as plumbing),
C-ABI calls with status checks, and the oracle evaluations they compare to."""
from __future__ import annotations

import ctypes as C

import numpy as np

from gpucell_amd import abi


def dev(a, dtype=None):
    import torch
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(a.copy()).to("cuda")


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def call(name, *args):
    lib = abi.load()
    conv = [abi.ptr(a) if (hasattr(a, "data_ptr") or isinstance(a, np.ndarray)) else a
            for a in args]
    st = getattr(lib, name)(*conv)
    if st != 0:
        raise abi.GcError(f"{name}: status {st}: {lib.gc_GetLastError().decode()}")
    return st


def upload_consts(case):
    lib = abi.load()
    abi.check(lib.gc_ConstsDataUp(C.byref(case.consts)), "ConstsDataUp")
    abi.check(lib.gc_SetExtConsts(C.byref(case.ext)), "SetExtConsts")


class GpuCase:
    """Device copies of a case's arrays in the reference's layouts, slot order
    = load order (sortpart = identity) until a neighbour build reorders them."""

    def __init__(self, case, ncube, outer, invw, K):
        import torch
        N = case.N
        self.case = case
        self.N, self.K = N, K
        self.coor = dev(case.coor, np.float32)
        self.velrhop = dev(np.concatenate([case.vel, np.full((N, 1), case.consts.DenInit)], 1),
                           np.float32)
        self.velhalf = dev(case.vel, np.float32)
        self.force = torch.zeros((N, 3), dtype=torch.float32, device="cuda")
        self.press = torch.zeros(N, dtype=torch.float32, device="cuda")
        self.cc = torch.zeros((N, 3), dtype=torch.int32, device="cuda")
        self.cubepart = torch.zeros(N, dtype=torch.int32, device="cuda")
        self.sortpart = dev(np.arange(N), np.int32)
        self.oldpos = torch.zeros(N, dtype=torch.int32, device="cuda")
        self.curpos = dev(np.arange(N), np.int32)
        self.nthro = torch.zeros((N, 3), dtype=torch.int32, device="cuda")
        self.be = torch.zeros((K, 2), dtype=torch.int32, device="cuda")
        self.bbox = dev(case.bbox, np.float32)
        self.ncube = dev(ncube, np.int32)
        self.outer = dev(outer, np.float32)
        self.invw = dev(invw, np.float32)

    def con_cube_list(self):
        """Particle::ConCubeList (Particle.cu:666-678) through the C-ABI."""
        c = self.case
        call("gc_PreConList", self.N, int(c.consts.CutRadius), self.coor, self.bbox, self.invw,
             self.ncube, self.cc, self.cubepart, self.sortpart, self.oldpos)
        call("gc_Sort", self.N, self.cubepart, self.sortpart)
        call("gc_CalBeginEndCube", self.N, self.K, self.cubepart, self.be)
        if c.nCell > 0:
            call("gc_SortArraysWithCells", self.N, self.sortpart, self.oldpos, self.curpos,
                 self.coor, self.cc, self.velrhop, self.velhalf, self.force, self.press,
                 self.nthro)
        else:
            call("gc_SortBasicArrays", self.N, self.sortpart, self.oldpos, self.coor, self.cc,
                 self.velrhop, self.velhalf, self.force, self.press)


def id_order(slot_arr, sortpart):
    out = np.empty_like(slot_arr)
    out[sortpart] = slot_arr
    return out


def oracle_forces(o, case, pos, vel, rho_in, img, counter, driveon, want_rho=True):
    """Total force of one step at fixed inputs, as the oracle evaluates it:
    density (fluid+membrane), EOS, SDPD, Bound1 repulsion, body force,
    membrane and Morse; returns (F, abs_sum, rho, extras)."""
    N = pos.shape[0]
    rho = rho_in.copy()
    if want_rho:
        o.density(pos, rho)
    press = np.zeros(N)
    o.eos(rho, press)
    F = np.zeros((N, 3))
    ab = np.zeros(N)
    o.sdpd_force(pos, vel, rho, press, counter, F, ab)
    o.rep_force(pos, F, ab)
    o.ext_force(pos, driveon, F, ab)
    extras = {"press": press}
    if case.nCell:
        g = o.geo(0, pos, img)
        extras["geo"] = g
        extras["def"] = o.def_force(pos, img, F, ab)
        extras["agg"] = o.agg_force(pos, F, ab)
    return F, ab, rho, extras
