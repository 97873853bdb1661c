"""Pins the oracle's SDPD physics without retyping its formulas:

* conservative force = -grad E with E = sum_i m c_i^2 ln d_i (+ the frozen
  walls' sum_ij (m c_w^2/d0) W(r_ij)): with P = c^2 rho the pressure term
  of the pair force is exactly the gradient of the isothermal free energy;
* Bound1 repulsion = -grad (K sum W);
* momentum conservation (sum F = 0) with all three pair terms;
* zero force on a perfect lattice at rest;
* dissipation: power <= 0, zero for uniform motion, and the continuum limit
  of the viscous term on a parabolic profile recovers eta (rho dv/dt =
  eta lap v for a divergence-free flow);
* noise: fluctuation-dissipation gives the equilibrium temperature kT;
* integrator: exact parabola under a constant force; periodic wrap keeps
  the unwrapped coordinate continuous."""
import math

import numpy as np
import pytest

from gpucell_amd import casegen
from oracle import pyoracle as orc
from tests import util


def _free_energy(o, ph, pos, nb1, nf, c2_of, L):
    rho = np.full(pos.shape[0], ph.rho0)
    o.density(pos, rho)
    d = rho[nf:] / ph.mass
    E = float((ph.mass * c2_of[nf:] * np.log(d)).sum())
    if nf > nb1:  # walls (Bound2) at fixed rest density
        d0 = ph.rho0 / ph.mass
        for i in range(nf, pos.shape[0]):
            for j in range(nb1, nf):
                dd = pos[i] - pos[j]
                dd -= L * np.round(dd / L)
                E += ph.mass * c2_of[j] / d0 * orc.lucy_w(np.linalg.norm(dd), ph.h)
    return E


@pytest.mark.parametrize("with_walls", [False, True])
def test_conservative_force_is_free_energy_gradient(with_walls):
    rng = np.random.default_rng(1)
    L = 7.0
    n = 260
    pos = rng.uniform(0, L, size=(n, 3))
    nb2 = 30 if with_walls else 0
    nm = 20  # membrane-kind particles: second sound speed, no membrane mesh
    ph = util.replace_physics(h=1.5, c1=3.0, c2=5.0)
    case = util.mini_case(pos, nb2=nb2, L=(L, L, L), physics=ph)
    case.consts.nCellStart = n - nm
    case.consts.nCellPar = nm
    o = orc.Oracle(case)
    nf = nb2
    c2_of = np.where(np.arange(n) < n - nm, ph.c1 ** 2, ph.c2 ** 2)
    rho = np.full(n, ph.rho0)
    o.density(pos, rho)
    press = np.zeros(n)
    o.eos(rho, press)
    F = np.zeros((n, 3))
    o.sdpd_force(pos, np.zeros((n, 3)), rho, press, 0, F)
    eps = 1e-6
    for i in rng.choice(np.arange(nf, n), 12, replace=False):
        g = np.zeros(3)
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, d] += eps
            pm[i, d] -= eps
            g[d] = (_free_energy(o, ph, pp, 0, nf, c2_of, L) -
                    _free_energy(o, ph, pm, 0, nf, c2_of, L)) / (2 * eps)
        assert np.allclose(F[i], -g, rtol=1e-6, atol=1e-6 * np.abs(g).max()), (i, F[i], -g)


def test_bound1_repulsion_is_gradient():
    rng = np.random.default_rng(2)
    L = 6.0
    n = 120
    pos = rng.uniform(0, L, size=(n, 3))
    nb1 = 40
    ph = util.replace_physics(h=1.5, c1=2.0)
    case = util.mini_case(pos, nb1=nb1, L=(L, L, L), physics=ph)
    case.ext.rep_scale = 0.7
    o = orc.Oracle(case)
    K = 2 * 0.7 * ph.c1 ** 2 * ph.rho0 / (ph.rho0 / ph.mass) ** 2

    def energy(p):
        E = 0.0
        for i in range(nb1, n):
            for j in range(nb1):
                dd = p[i] - p[j]
                dd -= L * np.round(dd / L)
                E += K * orc.lucy_w(np.linalg.norm(dd), ph.h)
        return E

    F = np.zeros((n, 3))
    o.rep_force(pos, F)
    for i in (nb1, nb1 + 7, n - 1):
        g = np.zeros(3)
        for d in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, d] += 1e-6
            pm[i, d] -= 1e-6
            g[d] = (energy(pp) - energy(pm)) / 2e-6
        assert np.allclose(F[i], -g, rtol=1e-6, atol=1e-9)
    assert np.all(F[:nb1] == 0)


@pytest.mark.parametrize("noise", [False, True])
def test_momentum_conservation(noise):
    rng = np.random.default_rng(3)
    L, n = 10.0, 1500
    pos = rng.uniform(0, L, size=(n, 3))
    vel = rng.normal(0, 0.3, size=(n, 3))
    ph = util.replace_physics(h=1.5, eta=2.0, zeta=0.5, kT=0.1)
    case = util.mini_case(pos, L=(L, L, L), physics=ph, noise=noise)
    o = orc.Oracle(case)
    rho = np.zeros(n)
    o.density(pos, rho)
    press = np.zeros(n)
    o.eos(rho, press)
    F = np.zeros((n, 3))
    ab = np.zeros(n)
    o.sdpd_force(pos, vel, rho, press, 5, F, ab)
    assert np.abs(F.sum(axis=0)).max() < 1e-9 * ab.sum()
    # dissipation alone: uniform motion -> zero; power <= 0
    Fd = np.zeros((n, 3))
    case2 = util.mini_case(pos, L=(L, L, L), physics=util.replace_physics(h=1.5, eta=2.0, zeta=0.5, c1=0.0, c2=0.0))
    o2 = orc.Oracle(case2)
    o2.sdpd_force(pos, vel, rho, np.zeros(n), 0, Fd)
    assert (Fd * vel).sum() < 0
    o2.sdpd_force(pos, np.tile([0.3, -0.2, 0.1], (n, 1)), rho, np.zeros(n), 0, Fd)
    assert np.abs(Fd).max() < 1e-12


def test_lattice_at_rest_has_zero_force():
    n = 8
    g = np.stack(np.meshgrid(*[np.arange(n) + 0.5] * 3, indexing="ij"), -1).reshape(-1, 3)
    ph = util.replace_physics(h=1.5)
    case = util.mini_case(g, L=(n, n, n), physics=ph)
    o = orc.Oracle(case)
    N = g.shape[0]
    rho = np.zeros(N)
    o.density(g, rho)
    assert np.allclose(rho, rho[0], rtol=1e-13)
    press = np.zeros(N)
    o.eos(rho, press)
    F = np.zeros((N, 3))
    o.sdpd_force(g, np.zeros((N, 3)), rho, press, 0, F)
    assert np.abs(F).max() < 1e-11


def test_viscous_term_recovers_shear_viscosity():
    # v_x = A y^2 on a fine lattice (h = 3 dx): rho dv/dt = eta * 2A
    dx, h, A, eta = 0.25, 0.75, 0.01, 1.7
    nx, ny = 12, 48
    axes = [np.arange(nx) * dx, np.arange(ny) * dx - ny * dx / 2, np.arange(nx) * dx]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3) + dx / 2
    N = g.shape[0]
    ph = util.replace_physics(h=h, eta=eta, zeta=0.0, c1=0.0, c2=0.0, mass=dx ** 3)
    case = util.mini_case(g, L=(nx * dx, ny * dx, nx * dx), lo=(0, -ny * dx / 2, 0),
                          periodic=(1, 1, 1), physics=ph)
    o = orc.Oracle(case)
    vel = np.zeros((N, 3))
    vel[:, 0] = A * g[:, 1] ** 2
    rho = np.zeros(N)
    o.density(g, rho)
    F = np.zeros((N, 3))
    o.sdpd_force(g, vel, rho, np.zeros(N), 0, F)
    interior = np.abs(g[:, 1]) < ny * dx / 2 - 2 * h   # away from the periodic seam
    got = (F[interior, 0] * rho[interior] / ph.mass)   # rho dv/dt
    assert np.allclose(got.mean(), 2 * A * eta, rtol=0.02), got.mean() / (2 * A * eta)
    # lateral components vanish
    assert np.abs(F[interior, 1:]).max() < 1e-3 * np.abs(F[interior, 0]).max()


@pytest.mark.slow
def test_noise_gives_equilibrium_temperature():
    # fluctuation-dissipation: a fluid at rest heats up to <m v^2> = 3 kT
    ph = util.replace_physics(h=1.5, kT=0.05, eta=1.0, c1=2.0, c2=2.0, dt=0.005)
    n = 12
    g = np.stack(np.meshgrid(*[np.arange(n) + 0.5] * 3, indexing="ij"), -1).reshape(-1, 3)
    case = util.mini_case(g, L=(n, n, n), physics=ph, noise=True)
    o = orc.Oracle(case)
    st = o.initial_state()
    T = []
    for k in range(1200):
        o.step(st, driveon=0)
        if k >= 400:
            v = st["vel"]
            T.append(ph.mass * ((v - v.mean(0)) ** 2).sum() / (3 * (v.shape[0] - 1)))
    Tm = float(np.mean(T))
    assert abs(Tm / ph.kT - 1) < 0.06, Tm / ph.kT


def test_constant_force_parabola_and_wrap():
    # one particle, no neighbours, uniform body force: GW is exact
    ph = util.replace_physics(h=1.0, fext=0.3, dt=0.01, mass=2.0)
    case = util.mini_case([[1.0, 1.0, 1.0]], L=(5, 5, 5), physics=ph, vel=[[0.7, 0.1, -0.2]])
    case.consts.ConstExtForce = 0.3
    o = orc.Oracle(case)
    st = o.initial_state()
    x0, v0 = st["pos"][0].copy(), st["vel"][0].copy()
    a = np.array([0.3 / 2.0, 0, 0])
    o.step(st, 1)  # first step moves with F(0) = 0
    x1 = x0 + v0 * 0.01
    v1 = v0 + 0.5 * a * 0.01  # corrector half kick
    assert np.allclose(st["pos"][0], x1, atol=1e-14)
    assert np.allclose(st["vel"][0], v1, atol=1e-14)
    unwrapped = []
    for k in range(400):
        o.step(st, 1)
        unwrapped.append(st["pos"][0] + st["img"][0] * 5.0)
        assert np.all((st["pos"][0] >= 0) & (st["pos"][0] < 5))
    t = 400 * 0.01
    # after the first step the force is constant: x(t) = x1 + v1 t + a t^2 / 2
    assert np.allclose(unwrapped[-1], x1 + v1 * t + 0.5 * a * t * t, atol=1e-9)
    assert st["img"][0][0] >= 1
    steps = np.diff(np.array(unwrapped), axis=0)
    assert np.abs(steps).max() < 0.1


def test_moving_wall_velocity_is_no_slip_image():
    """Couette walls: a Bound2 particle's velocity is 2 u_w - sum(W v)/sum(W)
    over its moving neighbours, u_w = -U below the mid-plane and +U above;
    with the fluid in uniform motion v0 that is exactly 2 u_w - v0, and a
    wall particle out of reach of the fluid (third layer) takes u_w."""
    case = casegen.couette(n=8, shear_rate=0.02, noise=False)
    o = orc.Oracle(case)
    N, nb2, nf = case.N, case.consts.nBound2Start, case.consts.nFluidStart
    pos = case.coor.astype(np.float64)
    v0 = np.array([0.03, -0.01, 0.02])
    vel = np.zeros((N, 3))
    vel[nf:] = v0
    o.bound2_vel(pos, vel)
    U = 0.5 * 0.02 * 8.0
    walls = np.arange(nb2, nf)
    uw = np.where(pos[walls, 1] >= 4.0, U, -U)
    # walls with a fluid particle within h (brute force, periodic in x and z)
    d = pos[walls][:, None, :] - pos[nf:][None, :, :]
    d[..., [0, 2]] -= 8.0 * np.round(d[..., [0, 2]] / 8.0)
    reach = (np.linalg.norm(d, axis=2) < case.consts.CutRadius).any(axis=1)
    exp = np.zeros((walls.size, 3))
    exp[reach] = -v0
    exp[reach, 0] += 2 * uw[reach]
    exp[~reach, 0] = uw[~reach]
    assert reach.any() and (~reach).any()
    assert np.allclose(vel[walls], exp, atol=1e-12)
