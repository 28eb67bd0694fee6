"""Extract the known-answer DATA of the reference's unit tests into fixtures.

Run here (where /root/reference exists); it reads the reference test sources
as text and writes only numbers:

  tests/golden/reference_math.npz      math_test.cpp:28-160 matrices A, B,
                                       invA, invB, AxB, detA, detB, C (10x10),
                                       D (100x100), symmetric packed matrix,
                                       vec, sol (math_test.cpp:204-249)
  tests/golden/reference_quadrature.json  quadrature_test.cpp:177-244: rule
                                       arguments and expected integrals

The script needs the reference sources, so it runs only where they are;
tests load the committed fixtures.
"""
import json
import os
import re

import numpy as np

REF = "/root/reference/test"
HERE = os.path.dirname(os.path.abspath(__file__))

NUM = r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?"


def array_after(src, name):
    m = re.search(r"real\s+" + re.escape(name) + r"\s*\[[^\]]*\]\s*=\s*\{(.*?)\};", src, re.S)
    if not m:
        raise KeyError(name)
    return np.array([float(x) for x in re.findall(NUM, m.group(1))])


def scalar_after(src, name):
    m = re.search(r"real\s+" + re.escape(name) + r"\s*=\s*(" + NUM + r")\s*;", src)
    return float(m.group(1))


def math_fixture():
    src = open(os.path.join(REF, "math_test.cpp")).read()
    out = {}
    for nm in ("A", "B", "invA", "invB", "AxB", "C", "D"):
        out[nm] = array_after(src, nm)
    out["detA"] = np.array(scalar_after(src, "detA"))
    out["detB"] = np.array(scalar_after(src, "detB"))
    out["symm_full"] = array_after(src, "symm_full")
    out["symm_pack"] = array_after(src, "symm_pack")
    out["vec"] = array_after(src, "vec")
    out["sol"] = array_after(src, "sol")
    assert out["C"].size == 100 and out["D"].size == 10000 and out["symm_pack"].size == 55
    np.savez_compressed(os.path.join(HERE, "reference_math.npz"), **out)


def quadrature_fixture():
    src = open(os.path.join(REF, "quadrature_test.cpp")).read()
    cases = []
    pat = re.compile(r"rule\s*=\s*gauss_legendre_(\d)d\(([^)]*)\);\s*EXPECT_DOUBLE_EQ\((" + NUM +
                     r"),\s*int\dd\(rule,\s*(\w+)\),\s*(\d+)\)")
    for m in pat.finditer(src):
        dim = int(m.group(1))
        args = [float(a) for a in m.group(2).split(",")]
        cases.append({"dim": dim, "args": args, "expected": float(m.group(3)),
                      "func": m.group(4), "ulps": int(m.group(5))})
    assert len(cases) >= 18, len(cases)
    with open(os.path.join(HERE, "reference_quadrature.json"), "w") as f:
        json.dump({"source": "reference test/quadrature_test.cpp:177-244", "cases": cases}, f,
                  indent=1)


if __name__ == "__main__":
    math_fixture()
    quadrature_fixture()
    print("fixtures written")
