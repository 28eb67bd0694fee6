"""Pins the oracle's derived outputs (render/vrtxgen.cpp:25-59, eval_output;
output.cpp:39-49, update_mean) by closed forms: a state that is a
polynomial of degree <= p in the reference coordinates is reproduced exactly
at every sub-grid point, each output is the stated function of that state
(with the reference's Mach pressure (gamma - 1)(rho E - |u|^2 / 2)), and the
running mean of a sequence is its arithmetic mean."""
import numpy as np
import pytest

G = 1.4


def lagrange_nodes(p):
    return np.linspace(0.0, 1.0, p + 1)


def poly_state(p, nelem, seed):
    """Coefficients of rho, m, E that are random polynomials of degree <= p in
    each reference coordinate (so the equispaced interpolant is exact), with
    the evaluation function."""
    rng = np.random.default_rng(seed)
    n = p + 1
    cs = rng.uniform(-0.2, 0.2, (nelem, 5, n, n, n)) / n ** 3
    cs[:, 0, 0, 0, 0] += 1.0    # rho ~ 1
    cs[:, 4, 0, 0, 0] += 3.0    # E ~ 3
    def f(e, r, x, y, z):
        px = np.array([x ** i for i in range(n)])
        py = np.array([y ** i for i in range(n)])
        pz = np.array([z ** i for i in range(n)])
        return np.einsum("ijk,i,j,k->", cs[e, r], px, py, pz)
    x = lagrange_nodes(p)
    U = np.zeros((nelem, 5, n ** 3))
    for e in range(nelem):
        for r in range(5):
            for bk in range(n):
                for bj in range(n):
                    for bi in range(n):
                        U[e, r, (bk * n + bj) * n + bi] = f(e, r, x[bi], x[bj], x[bk])
    return U, f


@pytest.mark.parametrize("p,N", [(1, 2), (2, 3), (3, 4)])
def test_eval_output_reproduces_polynomial_states(orc, p, N):
    nelem = 3
    U, f = poly_state(p, nelem, 40 + p)
    out = {v: orc.eval_output(p, U, nelem, N, v, G) for v in ("mach", "rho", "u", "v", "w")}
    for e in range(nelem):
        for k in range(N + 1):
            for j in range(N + 1):
                for i in range(N + 1):
                    node = (N + 1) ** 2 * k + (N + 1) * j + i
                    s = [f(e, r, i / N, j / N, k / N) for r in range(5)]
                    u, v, w = s[1] / s[0], s[2] / s[0], s[3] / s[0]
                    sp = np.sqrt(u * u + v * v + w * w)
                    mach = sp / np.sqrt(G * (G - 1) * (s[4] - 0.5 * sp * sp) / s[0])
                    for name, ref in (("rho", s[0]), ("u", u), ("v", v), ("w", w), ("mach", mach)):
                        assert abs(out[name][e, node] - ref) <= 1e-12 * max(1.0, abs(ref)), name


def test_mach_keeps_reference_pressure(orc):
    """The reference's Mach drops rho in the kinetic energy: for rho != 1 it
    differs from the physical Mach number (planted-mistake guard)."""
    p, N = 1, 1
    rho, u, pr = 2.0, np.array([0.4, 0.0, 0.0]), 1.0
    E = pr / (G - 1) + 0.5 * rho * u @ u
    U = np.zeros((1, 5, 8))
    U[0, 0], U[0, 1:4], U[0, 4] = rho, (rho * u)[:, None], E
    m = orc.eval_output(p, U, 1, N, "mach", G)[0, 0]
    quirk = 0.4 / np.sqrt(G * (G - 1) * (E - 0.5 * 0.16) / rho)
    physical = 0.4 / np.sqrt(G * pr / rho)
    assert abs(m - quirk) < 1e-15 and abs(m - physical) > 1e-3


def test_update_mean_is_arithmetic_mean(orc):
    rng = np.random.default_rng(7)
    seq = rng.uniform(-1, 1, (6, 50))
    mean = np.zeros(50)          # the reference starts from a zeroed simstate (dg.cpp:123-124)
    start = 10
    for k, s in enumerate(seq):
        mean = orc.update_mean(s, mean, start + k, start)
        np.testing.assert_allclose(mean, seq[:k + 1].mean(axis=0), rtol=0, atol=1e-15)
