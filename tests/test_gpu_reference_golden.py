"""The gfx950 FP64 path against the reference's own outputs.

For every case of ``tests/golden/reference_residuals.npz`` (the reference's
residual() and RK4 / TVD-RK3 steps run on the same mesh and initial state;
the 8^3 p = 2 Taylor-Green workload at full size among them) the device
evaluates the
residual and takes the same steps through the C ABI:

* state after N steps and every step's RMS(R): 1e-10 relative;
* residual f = -Minv R and R: 1e-12 relative to the reference wherever the
  reference's own double evaluation lies within 5e-13 of the exact value of
  the discrete residual (the same algorithm in extended precision,
  oracle/liboracle_x.so).  On the Mach 0.1 Taylor-Green meshes the residual
  is ill-conditioned: the mass row cancels O(|m|/h) terms down to O(M^2 |m|
  /h), and the reference's double result is itself ~1e-12 from the exact
  value (8^3 p = 2 Taylor-Green: 1.03e-12); two double evaluations in different
  orders of operations cannot agree beyond their rounding floors there.  For
  those cases the bound is relative to that floor: the device must be within
  2x the reference's own distance from the exact value, and within 3x of it
  from the reference (config 0 measures 1.6e-12 and 1.9e-12: the literal
  1e-12 bar is not met there).  tests/test_oracle_reference_golden.py
  measures the floors.
"""
import numpy as np
import pytest

import reference_cases as rc
from dgpr_amd import dgx

pytestmark = pytest.mark.gpu

RES_TOL = 1e-12
SOL_TOL = 1e-10


def rel(a, b):
    a = np.asarray(a, dtype=np.longdouble)
    b = np.asarray(b, dtype=np.longdouble)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def mesh_and_params(c):
    m = dgx.Mesh(*c.mesh_args())
    kw = c.param_kwargs()
    prm = dgx.make_reference_state(gamma=kw["gamma"], R=287.0, mu=kw["mu"], c=kw["c"],
                                   rho=kw["rho"], u=kw["u"], v=kw["v"], w=kw["w"],
                                   eta=kw["eta"], source=kw["source"])
    return m, prm


@pytest.mark.parametrize("name", rc.names())
def test_residual_matches_reference(name, orc):
    c = rc.load(name)
    print("case", name, "nelem", c.nelem, "p", c.p, "q", c.q, flush=True)
    m, prm = mesh_and_params(c)
    with dgx.Solver(m, prm) as s:
        s.set_state(c.U0)
        R_g, f_g = s.residual()
    gx = orc.Geometry(*c.mesh_args(), extended=True)
    R_x, f_x = gx.residual(orc.make_params(**c.param_kwargs()), c.U0)
    for nm, g, ref, x in (("f", f_g, c.f, f_x), ("R", R_g, c.R, R_x)):
        e_ref, e_x, floor = rel(g, ref), rel(g, x), rel(ref, x)
        print("  %s: device vs reference %.2e, device vs exact %.2e, reference vs exact %.2e"
              % (nm, e_ref, e_x, floor), flush=True)
        if floor < 0.5 * RES_TOL:
            assert e_ref < RES_TOL and e_x < RES_TOL
        else:
            print("  ill-conditioned: bound relative to the rounding floor", flush=True)
            assert e_x <= 2.0 * floor
            assert e_ref <= 3.0 * floor


@pytest.mark.parametrize("name", rc.names())
def test_steps_match_reference(name):
    c = rc.load(name)
    print("case", name, c.scheme, "steps", c.nsteps, flush=True)
    m, prm = mesh_and_params(c)
    with dgx.Solver(m, prm) as s:
        s.set_state(c.U0)
        norms = s.step(c.dt, c.scheme, c.nsteps)
        U = s.get_state()
    eU = rel(U, c.UN)
    en = float(np.max(np.abs(norms - c.norms) / np.abs(c.norms)))
    print("  state %.2e  norms %.2e" % (eU, en), flush=True)
    assert eU < SOL_TOL
    assert en < SOL_TOL
