"""The benchmark workloads at their full sizes, through the C ABI.

Parity against the oracle needs meshes the oracle finishes in seconds; these
tests run the benchmark configurations themselves and check properties that
hold at any size:

* config 0 (8^3 p = 2 Taylor-Green, RK4, 10 steps) is compared with the
  reference's own outputs in tests/test_gpu_reference_golden.py (case
  "config0");
* config 1 (48x24x24 p = 3 curved channel, no-slip walls, TVD-RK3) and
  config 2 (64^3 p = 3 Taylor-Green, RK4): a few steps keep the state finite
  and conserve the total mass to rounding (interior fluxes cancel, the
  no-slip wall's boundary state carries no mass flux, and the residual sums
  to zero against the constant basis function, solve.cpp:484-794);
* config 3 (the 64-case ensemble): all 64 members built, two steps of the
  default layout (one launch per stage for all orders) replayed from the
  hipGraph equal each member's
  single-case steps to rounding (1e-13 relative: the same arithmetic in a
  differently compiled kernel);
* config 4 (32^3 p = 2 adjoint): the residual dot-product test and the
  gradient of the final kinetic energy after 5 RK4 steps against central
  differences along 3 random directions and in mu.
"""
import numpy as np
import pytest

from dgpr_amd import cases, dgx

pytestmark = pytest.mark.gpu

FD_TOL = 1e-6


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def gauss_and_basis(p):
    """Gauss points / weights on [0, 1] and B[i][m] = l_m(g_i) for the
    equispaced Lagrange basis of order p."""
    x, w = np.polynomial.legendre.leggauss(p + 1)
    g, w = 0.5 * (x + 1.0), 0.5 * w
    nodes = np.linspace(0.0, 1.0, p + 1)
    B = np.ones((p + 1, p + 1))
    for m in range(p + 1):
        for o in range(p + 1):
            if o != m:
                B[:, m] *= (g - nodes[o]) / (nodes[m] - nodes[o])
    return g, w, B


def total_mass(mesh, U):
    """sum_e integral of rho_h over element e with the (p+1)^3 Gauss rule
    (q = 1 trilinear geometry), i.e. the mass-matrix row sums applied to the
    density coefficients."""
    p, n = mesh.p, mesh.p + 1
    g, w, B = gauss_and_basis(p)
    rho = U.reshape(mesh.nelem, 5, n, n, n)[:, 0]                # [e][k][j][i]
    rq = np.einsum("ai,bj,ck,ekji->ecba", B, B, B, rho)          # at Gauss (c=k, b=j, a=i)
    X = mesh.nodes.reshape(mesh.nelem, 2, 2, 2, 3)               # [e][z][y][x][d]
    s = np.stack([1.0 - g, g])                                   # s[corner][pt]
    ds = np.array([-1.0, 1.0])
    J = np.zeros((mesh.nelem, n, n, n, 3, 3))                    # [e][k][j][i][d][c]
    for cz in range(2):
        for cy in range(2):
            for cx in range(2):
                x = X[:, cz, cy, cx, :]
                wx = np.einsum("i,j,k->kji", ds[cx] * np.ones(n), s[cy], s[cz])
                wy = np.einsum("i,j,k->kji", s[cx], ds[cy] * np.ones(n), s[cz])
                wz = np.einsum("i,j,k->kji", s[cx], s[cy], ds[cz] * np.ones(n))
                J[..., 0] += np.einsum("ed,kji->ekjid", x, wx)
                J[..., 1] += np.einsum("ed,kji->ekjid", x, wy)
                J[..., 2] += np.einsum("ed,kji->ekjid", x, wz)
    det = np.linalg.det(J)
    W = np.einsum("k,j,i->kji", w, w, w)
    return float(np.sum(det * W[None] * rq))


@pytest.mark.parametrize("cfg,nsteps", [("config1", 6), ("config2", 4)])
def test_full_size_steps_conserve_mass(cfg, nsteps):
    c = getattr(cases, cfg)()
    print("workload", c.name, c.scheme, "nelem", c.mesh.nelem, "steps", nsteps, flush=True)
    m0 = total_mass(c.mesh, c.U0)
    with dgx.Solver(c.mesh, c.params) as s:
        s.set_state(c.U0)
        norms = s.step(c.dt, c.scheme, nsteps)
        U = s.get_state()
    assert np.all(np.isfinite(norms)) and np.all(np.isfinite(U))
    m1 = total_mass(c.mesh, U)
    print("  mass %.17g -> %.17g (rel %.2e), norms %s" % (m0, m1, abs(m1 - m0) / m0, norms),
          flush=True)
    assert abs(m1 - m0) <= 1e-13 * abs(m0)
    assert rel(U, c.U0) > 1e-8  # the state did move


def test_config3_full_ensemble_matches_single_cases():
    cs = cases.config3()
    assert len(cs) == 64
    print("ensemble of %d members, %d DOF" % (len(cs), sum(c.dof for c in cs)), flush=True)
    dts = [c.dt for c in cs]
    with dgx.Ensemble([c.mesh for c in cs], [c.params for c in cs]) as en:
        for m, c in zip(en.members, cs):
            m.set_state(c.U0)
        en.step("RK4", dts, 2, graph=True)
        got = [m.get_state() for m in en.members]
    for i, c in enumerate(cs):
        with dgx.Solver(c.mesh, c.params) as s:
            s.tile = 0  # the ensemble kernel's units before tuning: one row
            s.set_state(c.U0)
            s.step(c.dt, "RK4", 2, norms=False)
            ref = s.get_state()
        assert np.all(np.isfinite(got[i]))
        e = rel(got[i], ref)
        assert e < 1e-13, "member %d (n=%d p=%d): %.3e" % (i, c.mesh.nx, c.mesh.p, e)
    print("  all 64 members equal single-case steps to 1e-13", flush=True)


def test_config4_adjoint_full_size():
    c = cases.config4()
    U0 = c.U0
    rng = np.random.default_rng(4)
    eps = 1e-6
    nsteps = 5
    with dgx.Solver(c.mesh, c.params) as s:
        # residual-level dot-product test at the initial state
        v = U0 * rng.uniform(-1.0, 1.0, U0.shape)
        w = rng.uniform(-1.0, 1.0, U0.shape)
        s.set_state(U0 + eps * v)
        _, fp = s.residual()
        s.set_state(U0 - eps * v)
        _, fm = s.residual()
        s.set_state(U0)
        gv = s.residual_vjp(w)
        lhs, rhs = float(w @ ((fp - fm) / (2 * eps))), float(v @ gv)
        err = abs(lhs - rhs) / abs(rhs)
        print("  residual dot test rel %.2e" % err, flush=True)
        assert err < FD_TOL

        s.set_state(U0)
        J, g, dmu = s.adjoint_rk4(c.dt, nsteps, 0)

        def energy(U, mu=None):
            if mu is not None:
                prm = c.params.copy()
                prm.mu = mu
                s.set_params(prm)
            s.set_state(U)
            s.step(c.dt, "RK4", nsteps, norms=False)
            e = s.kinetic_energy()
            if mu is not None:
                s.set_params(c.params)
            return e

        assert abs(energy(U0) - J) <= 1e-14 * abs(J)
        for k in range(3):
            d = U0 * rng.uniform(-1.0, 1.0, U0.shape)
            fd = (energy(U0 + eps * d) - energy(U0 - eps * d)) / (2 * eps)
            ad = float(g @ d)
            err = abs(fd - ad) / abs(ad)
            print("  direction %d: FD %.12e adjoint %.12e rel %.2e" % (k, fd, ad, err), flush=True)
            assert err < FD_TOL
        mu = c.params.mu
        fd = (energy(U0, mu * (1 + 1e-4)) - energy(U0, mu * (1 - 1e-4))) / (2e-4 * mu)
        err = abs(fd - dmu) / abs(dmu)
        print("  dJ/dmu: FD %.12e adjoint %.12e rel %.2e" % (fd, dmu, err), flush=True)
        assert err < FD_TOL
