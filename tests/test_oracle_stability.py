"""The time step of the 64^3 p = 3 Taylor-Green workload.

The workload's nominal dt = 5e-4 at h = 2 pi / 64 puts dt c / h (c = 10 at
Mach 0.1) at 0.51.  The same dt c / h on the 4^3 mesh is dt = 8e-3: there the
oracle (the CPU restatement of the reference's residual and RK4, pinned
against the reference's own outputs) goes unstable (a NaN norm, which
RK4_distributed reports as "looks like things went unstable",
time_distributed.cpp:23-44) within 60 steps, while at half that CFL,
dt = 4e-3 (the reference's own 100-step run of the same case, held in the
golden vectors), the norm decays.  The workload therefore runs at
dt = 2.5e-4 (cases.config2), half the nominal step.
"""
import numpy as np
import pytest

import reference_cases as rc
from dgpr_amd import cases

TWO_PI = 2.0 * np.pi


def cfl(case):
    return case.dt * case.params.c / (TWO_PI / case.mesh.nx)


# dt c / h of the 64^3 workload (c = 1 / Mach = 10) at its nominal step and
# at the step it runs
NOMINAL_CFL = 5e-4 * 10.0 / (TWO_PI / 64)
RUN_CFL = 2.5e-4 * 10.0 / (TWO_PI / 64)


def test_nominal_dt_of_the_64_cube_is_unstable(orc):
    c = cases.tgv_case(n=4, p=3, seed=25, dt=8e-3, nsteps=60)
    assert c.params.c == 10.0
    assert cfl(c) == pytest.approx(NOMINAL_CFL, rel=1e-12)
    m = c.mesh
    geo = orc.Geometry(m.p, m.q, m.nx, m.ny, m.nz, m.bounds, m.nodes)
    prm = orc.params_from(c.params)
    U = c.U0.copy()
    for k in range(c.nsteps):
        try:
            U, nrm = geo.step(prm, U, c.dt, "RK4", 1)
        except RuntimeError:
            print("NaN norm at step %d" % k)
            break
        assert np.isfinite(nrm[0])
    else:
        pytest.fail("the nominal CFL stayed stable for %d steps" % c.nsteps)


def test_half_the_nominal_dt_is_stable():
    c = rc.load("tgv4_p3_long")
    assert c.dt * 10.0 / (TWO_PI / c.nx) == pytest.approx(RUN_CFL, rel=1e-12)
    assert c.nsteps == 100 and np.all(np.isfinite(c.norms))
    assert np.max(c.norms) <= c.norms[0] * (1.0 + 1e-12)
    assert c.norms[-1] < 0.5 * c.norms[0]
    assert cases.CONFIG2_DT == 2.5e-4
