"""Pins the oracle's building blocks against the reference's own unit-test
known answers (restated as fixtures in tests/golden) and against library
routines / closed forms that a plausible mistake would violate."""
import json
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---- quadrature (quadrature_test.cpp:24-244) --------------------------------
# Test polynomials of the reference quadrature test (quadrature_test.cpp:24-76).
def ord2_1d(x): return 1.2785 * x * x + 4.382 * x + 0.3854
def ord5_1d(x): return 2.394 * x ** 5 + 8.371 * x ** 3 + 0.23444 * x + 3.223
def ord16_1d(x): return 5.76594 * x ** 16 + 3.7694 * x ** 9 + 1.7674 * x ** 2 + 4.11854
def ord32_1d(x): return 7.584 * x ** 32 + 2.69 * x ** 17 + 10.9473
def ord2_2d(x, y): return 1.2785 * x * x + 4.1854 * y * y + 9.43 * x * y + 4.382 * x + 2.333 * y + 0.3854
def ord16_2d(x, y):
    return (5.76594 * x ** 16 - 0.5542 * y ** 16 + 4.2833 * x ** 10 * y ** 8
            + 3.7694 * x ** 9 * y ** 3 + 1.7674 * x ** 2 + 4.11854)
def ord32_2d(x, y): return 10.3874 * x ** 32 * y ** 32 + 3.452 * x ** 4 * y ** 20 + 3.6654
def ord2_3d(x, y, z): return 5.2038 * x ** 2 * y ** 2 * z ** 2 - 2.3385 * y * z + 1.224
def ord16_3d(x, y, z):
    return (5.76594 * x ** 16 - 0.5542 * y ** 16 * z ** 16 + 4.2833 * x ** 10 * y ** 8 * z ** 6
            + 3.7694 * x ** 9 * y ** 3 + 1.7674 * x ** 2 * z ** 2 + 4.11854)


FUNCS = {k: v for k, v in globals().items() if k.startswith("ord")}


def rule1d(orc, order, a, b):
    n = int(math.ceil((order + 1.0) / 2.0))
    return orc.gauss_legendre(n, a, b)


def test_quadrature_reference_integrals(orc):
    """Every integral of quadrature_test.cpp:177-244.  The expected values are
    printed there with 15 significant digits, so they are compared to a
    relative 2e-14."""
    cases = json.load(open(os.path.join(GOLD, "reference_quadrature.json")))["cases"]
    for c in cases:
        f = FUNCS[c["func"]]
        a = c["args"]
        if c["dim"] == 1:
            x, w = rule1d(orc, a[0], a[1], a[2])
            val = float(np.sum(w * f(x)))
        elif c["dim"] == 2:
            x, wx = rule1d(orc, a[0], a[2], a[3])
            y, wy = rule1d(orc, a[1], a[4], a[5])
            X, Y = np.meshgrid(x, y, indexing="ij")
            val = float(np.sum(np.outer(wx, wy) * f(X, Y)))
        else:
            x, wx = rule1d(orc, a[0], a[3], a[4])
            y, wy = rule1d(orc, a[1], a[5], a[6])
            z, wz = rule1d(orc, a[2], a[7], a[8])
            X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
            W = wx[:, None, None] * wy[None, :, None] * wz[None, None, :]
            val = float(np.sum(W * f(X, Y, Z)))
        assert abs(val - c["expected"]) <= 2e-14 * abs(c["expected"]), (c, val)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 9, 17])
def test_gauss_legendre_vs_numpy(orc, n):
    """Nodes/weights equal numpy's leggauss mapped to [0,1] (weights to a few
    ulps: the reference's weight formula 1/((1-x^2) P_n'^2) loses a little
    accuracy as n grows)."""
    x, w = orc.gauss_legendre(n, 0.0, 1.0)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.allclose(x, 0.5 * (xr + 1.0), rtol=0, atol=4e-16)
    assert np.allclose(w, 0.5 * wr, rtol=0, atol=1.2e-15)
    # exactness on monomials up to degree 2n-1
    for k in range(2 * n):
        assert abs(np.sum(w * x ** k) - 1.0 / (k + 1)) < 1e-15


# ---- Lagrange basis (basis_test.cpp:20-476) ----------------------------------
PNTS = [0.223, 0.437, 0.945]   # basis_test.cpp:26-28
P1 = [lambda x: 1.0 - x, lambda x: x]
P2 = [lambda x: 2 * x * x - 3 * x + 1, lambda x: -4 * x * x + 4 * x, lambda x: 2 * x * x - x]
DP2 = [lambda x: 4 * x - 3, lambda x: -8 * x + 4, lambda x: 4 * x - 1]


def test_lagrange_closed_forms(orc):
    for x in PNTS:
        assert orc.lagrange1d(0, 0, x) == 1.0
        assert orc.lagrange1d_deriv(0, 0, x) == 0.0
        for i in range(2):
            assert abs(orc.lagrange1d(i, 1, x) - P1[i](x)) < 1e-15
            assert abs(orc.lagrange1d_deriv(i, 1, x) - (-1.0 if i == 0 else 1.0)) < 1e-15
        for i in range(3):
            assert abs(orc.lagrange1d(i, 2, x) - P2[i](x)) < 2e-15
            assert abs(orc.lagrange1d_deriv(i, 2, x) - DP2[i](x)) < 4e-15
    # Lagrange property on the equispaced nodes
    for p in range(1, 8):
        for i in range(p + 1):
            for j in range(p + 1):
                assert orc.lagrange1d(i, p, j / p) == (1.0 if i == j else 0.0)


@pytest.mark.parametrize("p", [0, 2, 7])
def test_partition_of_unity(orc, p):
    """basis_test.cpp:397-476: sum of basis = 1, sum of derivatives = 0."""
    for x in np.linspace(0, 1, 13):
        s = sum(orc.lagrange1d(i, p, x) for i in range(p + 1))
        d = sum(orc.lagrange1d_deriv(i, p, x) for i in range(p + 1))
        assert abs(s - 1.0) < 1e-12 and abs(d) < 1e-9


def test_lagrange_derivative_vs_numpy_polyfit(orc):
    """Derivative equals the derivative of numpy's interpolating polynomial."""
    for p in (3, 5):
        nodes = np.arange(p + 1) / p
        for i in range(p + 1):
            y = (np.arange(p + 1) == i).astype(float)
            coef = np.polynomial.polynomial.polyfit(nodes, y, p)
            dcoef = np.polynomial.polynomial.polyder(coef)
            for x in (0.1, 0.5, 0.77):
                assert abs(orc.lagrange1d(i, p, x) - np.polynomial.polynomial.polyval(x, coef)) < 1e-10
                assert abs(orc.lagrange1d_deriv(i, p, x)
                           - np.polynomial.polynomial.polyval(x, dcoef)) < 1e-8


# ---- dense algebra (math_test.cpp) --------------------------------------------
@pytest.fixture(scope="module")
def mdata():
    return dict(np.load(os.path.join(GOLD, "reference_math.npz")))


def test_det_inv_matmul_known_answers(orc, mdata):
    """math_test.cpp:166-202 (5-ULP comparisons there; relative 4e-15 here)."""
    for M, inv, det in (("A", "invA", "detA"), ("B", "invB", "detB")):
        d = orc.det3x3(mdata[M])
        assert abs(d - float(mdata[det])) <= 4e-15 * abs(float(mdata[det]))
        iv = orc.inv3x3(mdata[M])
        assert np.allclose(iv, mdata[inv], rtol=4e-15, atol=0)
    AB = orc.matmul(mdata["A"].reshape(3, 3), mdata["B"].reshape(3, 3))
    assert np.allclose(AB.ravel(), mdata["AxB"], rtol=4e-15, atol=1e-14)


def test_symmstore_mvmul_known_answer(orc, mdata):
    """math_test.cpp:204-249."""
    r = orc.symmstore_mvmul(10, mdata["symm_pack"], mdata["vec"])
    assert np.array_equal(r, mdata["sol"])


@pytest.mark.parametrize("name,n,diag_tol,off_tol", [("C", 10, 1e-14, 5e-15), ("D", 100, 4e-14, 5e-14)])
def test_gauss_jordan_invert(orc, mdata, name, n, diag_tol, off_tol):
    """math_test.cpp:251-307: M * inv(M) = I; also equals numpy.linalg.inv."""
    M = mdata[name].reshape(n, n)
    st, Mi = orc.invert(M)
    assert st == 0
    I = orc.matmul(M, Mi)   # the reference checks with its own matmul
    assert np.max(np.abs(np.diag(I) - 1.0)) < diag_tol
    assert np.max(np.abs(I - np.diag(np.diag(I)))) < off_tol
    ref = np.linalg.inv(M)
    assert np.max(np.abs(Mi - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_invert_detects_singular(orc):
    M = np.ones((4, 4))
    st, _ = orc.invert(M)
    assert st != 0


def test_spmm_equals_dense(orc):
    rng = np.random.default_rng(3)
    n, p = 7, 4
    A = rng.standard_normal((n, n))
    S = A + A.T
    packed = np.concatenate([S[r, r:] for r in range(n)])
    V = rng.standard_normal((p, n))
    out = np.zeros(n * p)
    lib = orc.lib()
    lib.orc_spmm(n, p, orc._p(packed), orc._p(np.ascontiguousarray(V.ravel())), -2.0, orc._p(out))
    assert np.allclose(out.reshape(p, n), -2.0 * V @ S, rtol=1e-13, atol=1e-13)
