"""Pins the oracle's pointwise physics (restated flux.cpp / bound_states.cpp).

* Roe flux: the reference's flux_test.cpp cases (consistency, antisymmetry,
  supersonic upwinding) plus an independent formulation
  F = (F_L + F_R).n/2 - |A_n(roe avg)| (U_R - U_L)/2 with |A| from a numerical
  eigen-decomposition of the complex-step Jacobian (Roe 1981).
* Viscous flux A: the Newtonian stress tensor written in tensor form.
* Boundary states: the invariants each state is defined by.
"""
import math

import numpy as np
import pytest

G = 1.4


def cons(rho, vel, p, gamma=G):
    vel = np.asarray(vel, dtype=float)
    return np.array([rho, *(rho * vel), p / (gamma - 1) + 0.5 * rho * vel @ vel])


def prim(U, gamma=G):
    rho = U[0]
    vel = U[1:4] / rho
    p = (gamma - 1) * (U[4] - 0.5 * rho * vel @ vel)
    return rho, vel, p


def euler_flux_tensor(U, gamma=G):
    rho, vel, p = prim(U, gamma)
    F = np.zeros((3, 5), dtype=np.result_type(U, 1.0))
    F[:, 0] = rho * vel
    F[:, 1:4] = rho * np.outer(vel, vel) + p * np.eye(3)
    F[:, 4] = vel * (U[4] + p)
    return F


# ---- the reference's flux_test.cpp fixture (flux_test.cpp:25-90) ------------
def roe_3d_fixture():
    f32 = np.float32
    deg = math.pi / 180.0

    def ang(ph, th):
        ph, th = ph * deg, th * deg
        return (float(np.cos(f32(ph))) * float(np.sin(f32(th))),
                float(np.sin(f32(ph))) * float(np.sin(f32(th))), float(np.cos(f32(th))))

    g, Minf, c = 1.4, 2.2, 1.0
    Ml, rl, Mr, rr = Minf + 0.01, 1.02, Minf - 0.02, 1.04
    vl = np.array(ang(5.0, 85.0)) * Ml
    vr = np.array(ang(3.0, 87.0)) * Mr
    n = np.array(ang(2.0, 83.0))
    Ul = np.array([rl, *(rl * vl), (rl * c * c) / (g * (g - 1.0)) + 0.5 * rl * Ml * Ml])
    Ur = np.array([rr, *(rr * vr), (rr * c * c) / (g * (g - 1.0)) + 0.5 * rr * Mr * Mr])
    return Ul, Ur, n


def test_roe_reference_flux_test(orc):
    """flux_test.cpp:92-134 (5-ULP comparisons there; 8e-15 relative here)."""
    Ul, Ur, n = roe_3d_fixture()
    Fl = euler_flux_tensor(Ul).T @ n
    f, _ = orc.roe(Ul, Ul, n, G)                       # consistency
    assert np.allclose(f, Fl, rtol=8e-15, atol=1e-15)
    ff, _ = orc.roe(Ul, Ur, n, G)                      # supersonic: upwind = left
    assert np.allclose(ff, Fl, rtol=8e-15, atol=1e-14)
    fr, _ = orc.roe(Ur, Ul, -n, G)                     # antisymmetry
    assert np.allclose(ff, -fr, rtol=8e-15, atol=1e-15)


def roe_eigen(UL, UR, n, gamma=G):
    """Roe flux from the eigen-decomposition of the Roe-averaged Jacobian."""
    rl, vl, pl = prim(UL, gamma)
    rr, vr, pr = prim(UR, gamma)
    Hl, Hr = (UL[4] + pl) / rl, (UR[4] + pr) / rr
    sl, sr = math.sqrt(rl), math.sqrt(rr)
    vel = (sl * vl + sr * vr) / (sl + sr)
    H = (sl * Hl + sr * Hr) / (sl + sr)
    c2 = (gamma - 1) * (H - 0.5 * vel @ vel)
    c = math.sqrt(c2)
    # any density with the Roe velocity and enthalpy gives the same Jacobian
    rho_t = 1.0
    Ut = cons(rho_t, vel, rho_t * c2 / gamma, gamma)
    J = np.zeros((5, 5))
    h = 1e-30
    for j in range(5):
        Uc = Ut.astype(complex)
        Uc[j] += 1j * h
        J[:, j] = (euler_flux_tensor(Uc, gamma).T @ n).imag / h
    lam, V = np.linalg.eig(J)
    lam = lam.real
    eps = 0.05 * c
    al = np.abs(lam)
    al = np.where(al < eps, (eps * eps + lam * lam) / (2 * eps), al)
    Aabs = (V @ np.diag(al) @ np.linalg.inv(V)).real
    return 0.5 * (euler_flux_tensor(UL, gamma).T @ n + euler_flux_tensor(UR, gamma).T @ n) \
        - 0.5 * Aabs @ (UR - UL)


@pytest.mark.parametrize("seed", range(8))
def test_roe_equals_eigendecomposition(orc, seed):
    rng = np.random.default_rng(seed)
    n = rng.standard_normal(3)
    n /= np.linalg.norm(n)
    base = rng.uniform(-0.6, 0.6, 3)
    UL = cons(rng.uniform(0.5, 2.0), base + rng.uniform(-0.3, 0.3, 3), rng.uniform(0.5, 2.0))
    UR = cons(rng.uniform(0.5, 2.0), base + rng.uniform(-0.3, 0.3, 3), rng.uniform(0.5, 2.0))
    if seed == 0:   # nearly stagnant across the face: exercises the entropy fix
        UR = UL.copy()
        UR[0] *= 1.01
        UL[1:4] = 0.001 * n
    f, _ = orc.roe(UL, UR, n, G)
    ref = roe_eigen(UL, UR, n)
    assert np.allclose(f, ref, rtol=1e-11, atol=1e-12 * np.max(np.abs(ref)))


def test_analytical_flux_tensor_form(orc):
    rng = np.random.default_rng(5)
    for _ in range(5):
        U = cons(rng.uniform(0.5, 2), rng.uniform(-1, 1, 3), rng.uniform(0.5, 2))
        assert np.allclose(orc.analytical_flux(U, G), euler_flux_tensor(U), rtol=1e-14, atol=1e-15)


def viscous_tensor(s, grad, mu):
    """tau = mu (grad u + grad u^T) - 2/3 mu (div u) I, Q = [0, tau, tau.u]."""
    rho = s[0]
    vel = s[1:4] / rho
    # grad[d, r]: d/dx_d of conserved r; velocity gradient by the quotient rule
    gu = (grad[:, 1:4] - np.outer(grad[:, 0], vel)) / rho      # gu[d, i] = d u_i / d x_d
    L = gu.T                                                    # L[i, d] = d u_i / d x_d
    tau = mu * (L + L.T) - (2.0 / 3.0) * mu * np.trace(L) * np.eye(3)
    Q = np.zeros((3, 5))
    Q[:, 1:4] = tau
    Q[:, 4] = tau @ vel
    return Q


def test_viscous_flux_tensor_form_and_linearity(orc):
    rng = np.random.default_rng(7)
    mu = 3.3e-3
    for _ in range(5):
        s = cons(rng.uniform(0.5, 2), rng.uniform(-1, 1, 3), rng.uniform(0.5, 2))
        g1 = rng.standard_normal((3, 5))
        g2 = rng.standard_normal((3, 5))
        Q1 = orc.viscous_A(s, g1[0], g1[1], g1[2], mu)
        assert np.allclose(Q1, viscous_tensor(s, g1, mu), rtol=1e-13, atol=1e-16)
        Q2 = orc.viscous_A(s, g2[0], g2[1], g2[2], mu)
        g3 = 2.5 * g1 - 0.75 * g2
        Q3 = orc.viscous_A(s, g3[0], g3[1], g3[2], mu)
        assert np.allclose(Q3, 2.5 * Q1 - 0.75 * Q2, rtol=1e-12, atol=1e-15)


# ---- boundary states (bound_states.cpp) ----------------------------------------
def test_wall_states(orc):
    prm = orc.make_params(gamma=G, mu=1e-3, c=1.0, rho=1.0, u=0.1)
    rng = np.random.default_rng(9)
    n = rng.standard_normal(3)
    n /= np.linalg.norm(n)
    UL = cons(1.2, [0.3, -0.2, 0.5], 0.9)
    st, Ub = orc.bound_state(-4, UL, n, prm)             # inviscid wall
    assert st == 0
    assert abs(Ub[1:4] @ n) < 1e-15
    t = UL[1:4] - (UL[1:4] @ n) * n
    assert np.allclose(Ub[1:4], t, atol=1e-15) and Ub[0] == UL[0] and Ub[4] == UL[4]
    st, Ub = orc.bound_state(-7, UL, n, prm)             # no-slip (adiabatic energy)
    assert st == 0 and np.array_equal(Ub, [UL[0], 0, 0, 0, UL[4]])
    st, Ub = orc.bound_state(-3, UL, n, prm)             # freestream
    assert np.array_equal(Ub, np.array(prm.Ufr[:]))


def test_subsonic_inflow_invariants(orc):
    gamma, R = G, 287.0
    prm = orc.make_params(gamma=gamma, R=R, mu=1e-3, c=1.0, rho=1.0, u=0.3)
    u = np.array(prm.Ufr[1:4]) / prm.Ufr[0]
    M = np.linalg.norm(u) / prm.c
    Tt = prm.T * (1 + 0.5 * (gamma - 1) * M * M)
    Pt = prm.P * (Tt / prm.T) ** (gamma / (gamma - 1))
    n = np.array([-1.0, 0.0, 0.0])                      # outward normal of an inlet at -x
    UL = cons(0.98, [0.28, 0.01, -0.02], prm.P * 1.01)
    st, Ub = orc.bound_state(-5, UL, n, prm, Pt, Tt)
    assert st == 0
    rl, vl, pl = prim(UL)
    rb, vb, pb = prim(Ub)
    cl, cb = math.sqrt(gamma * pl / rl), math.sqrt(gamma * pb / rb)
    Jl = vl @ n + 2 * cl / (gamma - 1)
    Jb = vb @ n + 2 * cb / (gamma - 1)
    assert abs(Jl - Jb) < 1e-13 * abs(Jl)                # outgoing Riemann invariant
    Tb = pb / (rb * R)
    Mb = np.linalg.norm(vb) / cb
    assert abs(Tb * (1 + 0.5 * (gamma - 1) * Mb * Mb) - Tt) < 1e-13 * Tt
    assert abs(pb * (1 + 0.5 * (gamma - 1) * Mb * Mb) ** (gamma / (gamma - 1)) - Pt) < 1e-13 * Pt
    assert np.allclose(np.cross(vb, n), 0, atol=1e-15) and vb @ n < 0


def test_subsonic_outflow_invariants(orc):
    gamma = G
    prm = orc.make_params(gamma=gamma, mu=1e-3, c=1.0, rho=1.0, u=0.2)
    n = np.array([0.6, 0.8, 0.0])
    UL = cons(1.05, [0.21, 0.05, 0.03], prm.P * 0.97)
    st, Ub = orc.bound_state(-6, UL, n, prm)
    assert st == 0
    rl, vl, pl = prim(UL)
    rb, vb, pb = prim(Ub)
    assert abs(pb - prm.P) < 1e-14 * prm.P                          # fixed pressure
    assert abs(pl / rl ** gamma - pb / rb ** gamma) < 1e-13 * pl    # entropy
    cl, cb = math.sqrt(gamma * pl / rl), math.sqrt(gamma * pb / rb)
    assert abs((vl @ n + 2 * cl / (gamma - 1)) - (vb @ n + 2 * cb / (gamma - 1))) < 1e-13
    tl, tb = vl - (vl @ n) * n, vb - (vb @ n) * n                   # tangential velocity
    assert np.allclose(tl, tb, atol=1e-15)


def test_inflow_without_root_reports(orc):
    prm = orc.make_params(gamma=G, mu=1e-3, c=1.0, rho=1.0, u=0.3)
    n = np.array([-1.0, 0.0, 0.0])
    UL = cons(1.0, [-200.0, 0.0, 0.0], 0.01)  # J+ so large both Mach roots are negative
    st, _ = orc.bound_state(-5, UL, n, prm, 1.0, 1.0)
    assert st != 0
