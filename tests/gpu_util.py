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


def oracle_state(case, S, info):
    """The oracle's step input from an engine state S (id order) and the
    engine's step / noise counter."""
    st = {
        "pos": S["pos"].copy(), "img": S["nthrobound"].copy(),
        "vel": S["velrhop"][:, :3].astype(np.float64), "rho": S["velrhop"][:, 3].astype(np.float64),
        "velhalf": S["velhalf"].astype(np.float64), "force": S["force"].astype(np.float64),
        "press": S["press"].astype(np.float64), "abs_sum": np.zeros(case.N),
        "step": int(info.step), "counter": int(info.noise_counter)}
    if case.nCell:
        st.update(defforce=np.zeros((case.nCellPar, 3)), aggforce=np.zeros((case.nCellPar, 3)),
                  cell_area=np.zeros(case.nCell), cell_vol=np.zeros(case.nCell))
    return st


def step_parity(o, case, S, S1, info, driveon=1, label=""):
    """One engine step (state S before, S1 after; info = the engine's
    step and noise counter before the step) against the FP64 oracle `o`
    (reference geometry and mechanics already set by o.initial_state()):

    * FP64 positions after the step to 1e-9 and the image counters exactly
      (the oracle advances S with S's own velocities and forces);
    * densities to 2e-6 relative and every particle's total force to the
      per-particle bound |F - F*| / (sum_j |f_ij*| + |F_ext|) <= 1e-5, both
      evaluated by the oracle at the FP32 positions the kernels read;
    * cell areas and volumes to 2e-6 relative; the corrector
      v = velhalf + dt/2 F/m.
    Returns (max error, relative L2, oracle extras)."""
    st = oracle_state(case, S, info)
    o.step(st, driveon=driveon)
    dpos = np.abs(S1["pos"] - st["pos"]).max()
    assert dpos < 1e-9, (label, dpos)
    assert np.array_equal(S1["nthrobound"], st["img"]), label
    mov = np.arange(case.N) >= case.consts.nFluidStart
    dt, lam, m = float(case.consts.TimeStep), float(case.consts.VVsigma), float(case.consts.Mass)
    vt = S["velrhop"][:, :3].astype(np.float64).copy()
    vt[mov] += lam * dt * S["force"][mov].astype(np.float64) / m
    pos32 = S1["coor"].astype(np.float64)
    o.bound2_vel(pos32, vt)
    rho = np.full(case.N, float(case.consts.DenInit))
    F_o, ab, rho, extra = oracle_forces(o, case, pos32, vt, rho, S1["nthrobound"], info.noise_counter,
                                        driveon)
    assert np.allclose(S1["velrhop"][:, 3], rho, rtol=2e-6), label
    den = ab + abs(float(case.consts.ConstExtForce))
    den = np.where(den > 0, den, 1.0)
    err = np.linalg.norm(S1["force"].astype(np.float64) - F_o, axis=1) / den
    emax = float(err.max())
    l2 = float(np.linalg.norm(S1["force"] - F_o) / max(np.linalg.norm(F_o), 1e-300))
    assert emax <= 1e-5, (label, emax, int(err.argmax()))
    if case.nCell:
        assert np.allclose(S1["area"], extra["geo"]["cell_area"], rtol=2e-6), label
        assert np.allclose(S1["vol"], extra["geo"]["cell_vol"], rtol=2e-6), label
    vh = S1["velhalf"].astype(np.float64)
    v_exp = vh + 0.5 * dt * S1["force"].astype(np.float64) / m
    assert np.allclose(S1["velrhop"][mov, :3], v_exp[mov], rtol=1e-6, atol=1e-7), label
    extra["pos32"] = pos32
    return emax, l2, extra
