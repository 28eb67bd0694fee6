"""Workload builders: meshes, reference states and initial states of the
benchmark and parity cases.

All inputs are synthetic and seeded (splitmix64 streams inside libdgx), so
the same call produces bit-identical meshes and states on any host.

Physics follows the reference's semantics: no heat conduction and an
adiabatic-energy no-slip wall state (flux.cpp:156-202, bound_states.cpp:42-49);
Pr only names the case.  The reference has only a fixed time step; the
channel's dt is derived here from CFL 0.3 as
    dt = CFL * h_min / ((|u|_max + c) (2p + 1)).
"""
import math
from dataclasses import dataclass

import numpy as np

from . import dgx

TWO_PI = 2.0 * math.pi


@dataclass
class Case:
    name: str
    mesh: "dgx.Mesh"
    params: "dgx.Params"
    U0: np.ndarray
    scheme: str
    dt: float
    nsteps: int

    @property
    def dof(self):
        return self.mesh.nelem * self.mesh.nbf

    @property
    def stages(self):
        return 4 if self.scheme == "RK4" else 3


def periodic_bounds():
    return [dgx.PERIODIC] * 6


def tgv_case(n=8, p=2, seed=0, noise=1e-3, M=0.1, Re=1600.0, rho0=1.0, V0=1.0, L=1.0,
             dt=1e-3, nsteps=10, scheme="RK4", gamma=1.4, eta=6.0, nxyz=None, amp=0.0):
    """Taylor-Green-type state on the periodic cube [0, 2pi]^3 (q = 1).

    ``amp`` > 0 displaces interior nodes smoothly (non-affine elements) while
    keeping the periodic faces matched."""
    nx, ny, nz = nxyz if nxyz is not None else (n, n, n)
    nodes = dgx.box_nodes(1, nx, ny, nz, 0.0, 0.0, 0.0, TWO_PI, TWO_PI, TWO_PI, amp, seed)
    mesh = dgx.Mesh(p, 1, nx, ny, nz, periodic_bounds(), nodes)
    c = V0 / M
    mu = rho0 * V0 * L / Re
    prm = dgx.make_reference_state(gamma=gamma, R=287.0, mu=mu, c=c, rho=rho0, u=0.0, v=0.0,
                                   w=0.0, eta=eta)
    U0 = dgx.tgv_state(mesh, gamma, rho0, V0, prm.P, noise, seed)
    return Case("tgv_%dx%dx%d_p%d" % (nx, ny, nz, p), mesh, prm, U0, scheme, dt, nsteps)


def channel_case(nx=48, ny=24, nz=24, p=3, q=1, seed=1, amp=0.1, M=0.2, Re=3000.0, rho0=1.0,
                 Uc=1.0, frac=0.05, kmax=4, cfl=0.3, nsteps=200, scheme="tvdRK3",
                 gamma=1.4, eta=6.0, wall=dgx.NO_SLIP_WALL):
    """Channel [0,4pi]x[-1,1]x[0,4pi/3], periodic in x and z, walls at y=+-1,
    interior nodes displaced by a smooth field of amplitude amp*h."""
    lx, ly, lz = 4.0 * math.pi, 2.0, 4.0 * math.pi / 3.0
    nodes = dgx.box_nodes(q, nx, ny, nz, 0.0, -1.0, 0.0, lx, ly, lz, amp, seed)
    bounds = [dgx.PERIODIC, dgx.PERIODIC, wall, wall, dgx.PERIODIC, dgx.PERIODIC]
    mesh = dgx.Mesh(p, q, nx, ny, nz, bounds, nodes)
    c = Uc / M
    mu = rho0 * Uc * 1.0 / Re
    prm = dgx.make_reference_state(gamma=gamma, R=287.0, mu=mu, c=c, rho=rho0, u=0.0, v=0.0,
                                   w=0.0, eta=eta)
    U0 = dgx.channel_state(mesh, gamma, rho0, Uc, prm.P, frac, kmax, lx, lz, seed)
    hmin = min(lx / nx, ly / ny, lz / nz)
    dt = cfl * hmin / ((Uc + c) * (2 * p + 1))
    return Case("channel_%dx%dx%d_p%d" % (nx, ny, nz, p), mesh, prm, U0, scheme, dt, nsteps)


def config0():
    """8^3 periodic p=2 Taylor-Green, RK4 dt=1e-3, 10 steps."""
    return tgv_case(n=8, p=2, seed=0, dt=1e-3, nsteps=10, scheme="RK4")


def config1():
    """48x24x24 p=3 curved channel with no-slip walls, TVD-RK3, 200 steps."""
    return channel_case()


# time step of the 64^3 p=3 workload (half its nominal 5e-4, see config2)
CONFIG2_DT = 2.5e-4


def config2():
    """64^3 periodic p=3 Taylor-Green, RK4, 100 steps at dt = 2.5e-4.

    The nominal dt = 5e-4 exceeds the explicit stability limit of this
    scheme at this resolution: the same CFL (dt c / h) on the 4^3 mesh
    (dt = 8e-3) diverges after about 40 RK4 steps in the CPU restatement of
    the reference residual as well as on the device, while dt c / h halved
    stays bounded over 220 steps."""
    return tgv_case(n=64, p=3, seed=2, dt=CONFIG2_DT, nsteps=100, scheme="RK4")


def ensemble_cases(ncase=64, seed=3, nsteps=50, cfl=0.2):
    """Independent periodic Taylor-Green cases drawn from ``seed``: n uniform in
    {8, 12, 16, 24, 32} (n^3 elements), p uniform in {1, 2, 3}, Mach uniform in
    [0.05, 0.3], Re log-uniform in [400, 5000], each with seeded velocity
    noise; RK4 with dt = cfl h / ((V0 + c)(2p + 1)) per case.  cfl = 0.2 is
    the 64^3 p=3 workload's stable dt (2.5e-4) expressed in the same form."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(ncase):
        n = int(rng.choice([8, 12, 16, 24, 32]))
        p = int(rng.integers(1, 4))
        M = float(rng.uniform(0.05, 0.3))
        Re = float(math.exp(rng.uniform(math.log(400.0), math.log(5000.0))))
        c = 1.0 / M
        dt = cfl * (TWO_PI / n) / ((1.0 + c) * (2 * p + 1))
        out.append(tgv_case(n=n, p=p, seed=1000 * seed + i, M=M, Re=Re, dt=dt,
                            nsteps=nsteps, scheme="RK4"))
    return out


def config3():
    """64 independent periodic cubes (sizes, orders, Mach and Re from seed 3),
    FP64 RK4, 50 steps each."""
    return ensemble_cases()


def smooth_field_case(n=32, p=2, seed=4, M=0.1, Re=1600.0, kmax=3, dt=1e-3, nsteps=100,
                      gamma=1.4, eta=6.0):
    """Periodic cube [0, 2pi]^3 with a seeded random smooth velocity field:
    Fourier modes 0 < |k| <= kmax with uniform random amplitudes and phases,
    scaled to unit RMS over the nodes; rho = 1, p = 1 / (gamma M^2) (so
    c = 1/M), mu = 1/Re.  State sampled at the equispaced Lagrange nodes of
    the affine elements."""
    nodes = dgx.box_nodes(1, n, n, n, 0.0, 0.0, 0.0, TWO_PI, TWO_PI, TWO_PI, 0.0, seed)
    mesh = dgx.Mesh(p, 1, n, n, n, periodic_bounds(), nodes)
    c = 1.0 / M
    prm = dgx.make_reference_state(gamma=gamma, R=287.0, mu=1.0 / Re, c=c, rho=1.0, u=0.0,
                                   v=0.0, w=0.0, eta=eta)
    h = TWO_PI / n
    xi = np.linspace(0.0, 1.0, p + 1)
    ax = (np.arange(n)[:, None] + xi[None, :]) * h            # [elem index][node]
    # node coordinates [k][j][i][bk][bj][bi] -> element-major [e][bf]
    X = ax[None, None, :, None, None, :]
    Y = ax[None, :, None, None, :, None]
    Z = ax[:, None, None, :, None, None]
    shape = (n, n, n, p + 1, p + 1, p + 1)
    X, Y, Z = np.broadcast_to(X, shape), np.broadcast_to(Y, shape), np.broadcast_to(Z, shape)
    rng = np.random.default_rng(seed)
    vel = np.zeros((3,) + shape)
    for kx in range(-kmax, kmax + 1):
        for ky in range(-kmax, kmax + 1):
            for kz in range(-kmax, kmax + 1):
                if kx * kx + ky * ky + kz * kz == 0 or kx * kx + ky * ky + kz * kz > kmax * kmax:
                    continue
                amp = rng.uniform(-1.0, 1.0, 3)
                ph = rng.uniform(0.0, TWO_PI, 3)
                arg = kx * X + ky * Y + kz * Z
                for d in range(3):
                    vel[d] += amp[d] * np.cos(arg + ph[d])
    vel /= np.sqrt(np.mean(np.sum(vel ** 2, axis=0)))
    nel, nbf = n ** 3, (p + 1) ** 3
    u = vel.reshape(3, nel, nbf)
    P0 = prm.P
    U = np.zeros((nel, 5, nbf))
    U[:, 0] = 1.0
    U[:, 1:4] = np.transpose(u, (1, 0, 2))
    U[:, 4] = P0 / (gamma - 1.0) + 0.5 * np.sum(u ** 2, axis=0)
    return Case("smooth_%d3_p%d" % (n, p), mesh, prm, U.reshape(-1), "RK4", dt, nsteps)


def config4():
    """32^3 periodic p=2, random smooth velocity (|k| <= 3, unit RMS, seed 4),
    M = 0.1, RK4 dt = 1e-3, 100 steps: the adjoint workload."""
    return smooth_field_case()
