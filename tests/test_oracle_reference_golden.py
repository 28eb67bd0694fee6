"""Pins the oracle against the reference's own outputs.

``tests/golden/reference_residuals.npz`` holds, for 32 cases (the 8^3 p = 2
Taylor-Green workload at full size, periodic Taylor-Green p = 1..3 up to 8^3 p = 3 at
Mach 0.1, curved q = 1..3
channels with no-slip and inviscid walls, no-slip walls under a moving
state, the reference's default bump with subsonic inflow / outflow and
inviscid walls, free-stream boundaries with a source term, single-element
and ragged periodic meshes, and the orders outside 1..3: p = 0 periodic
and walled, p = 4 Taylor-Green and default bump with q = 2, p = 5 and p = 2
q = 4 channels, a p = 6 q = 3 channel, a periodic p = 7 box and a p = 8
q = 8 channel), the reference's residual R and f = -Minv R at
the initial state and its state and RMS(R) norms after N RK4 / TVD-RK3
steps, produced by the reference's residual() (solve.cpp:150-807) and
RK4_distributed / tvdRK3_distributed (time_distributed.cpp:46-102) compiled
from its sources (tests/golden/make_reference_residuals.py).

* The oracle (oracle/dg_oracle.c) must reproduce them: R and f to 1e-12
  relative, the state after N steps and every norm to 1e-12.
* Planted mistakes in the BR2 penalty, the one part of the residual the
  invariants do not fix (interior lifting 1/2 dropped, boundary lifting
  given a 1/2, boundary lifting with A at the interior trace instead of the
  boundary state, interior QH penalty 1/2 dropped), each move f far outside
  that bar on a named case.
* The extended-precision build of the same algorithm measures how far each
  double evaluation lies from the exact discrete residual: the rounding
  floor.  On the p = 3 Mach 0.1 Taylor-Green cases the reference's own f is
  ~1e-12 from the exact value, so a residual bar of 1e-12 against the
  reference sits at the reference's own rounding level there.
"""
import numpy as np
import pytest

import reference_cases as rc

TOL = 1e-12


def rel(a, b):
    a = np.asarray(a, dtype=np.longdouble)
    b = np.asarray(b, dtype=np.longdouble)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def setup(orc, name, extended=False):
    c = rc.load(name)
    geo = orc.Geometry(*c.mesh_args(), extended=extended)
    prm = orc.make_params(**c.param_kwargs())
    return c, geo, prm


@pytest.mark.parametrize("name", rc.names())
def test_oracle_residual_matches_reference(orc, name):
    c, geo, prm = setup(orc, name)
    R, f = geo.residual(prm, c.U0)
    eR, ef = rel(R, c.R), rel(f, c.f)
    print("%s: R %.2e  f %.2e" % (name, eR, ef))
    assert eR < TOL and ef < TOL


# the long case's 100 p=3 steps take ~10 s in the oracle; the others are short
@pytest.mark.parametrize("name", rc.names())
def test_oracle_steps_match_reference(orc, name):
    c, geo, prm = setup(orc, name)
    U, norms = geo.step(prm, c.U0, c.dt, c.scheme, c.nsteps)
    eU = rel(U, c.UN)
    en = float(np.max(np.abs(norms - c.norms) / np.abs(c.norms)))
    print("%s: %d %s steps, U %.2e  norms %.2e" % (name, c.nsteps, c.scheme, eU, en))
    assert eU < TOL and en < TOL


# Cases whose state jumps across faces (per-node perturbations).  The
# channel cases sample a continuous field at the element nodes: their face
# traces agree from both sides and at the walls the fluid is at rest, so the
# lifting vanishes at their initial state and they cannot see these plants.
PLANTS = [
    # (plant, case whose f must move, what it changes)
    (4, "tgv3_p2", "interior lifting without 1/2 (solve.cpp:455-467)"),
    (4, "single_element_p3", "interior lifting without 1/2 (solve.cpp:455-467)"),
    (7, "tgv3_p2", "interior QH penalty eta*(dL+dR) without 1/2 (solve.cpp:583-585)"),
    (7, "bump_default_p2q2", "interior QH penalty eta*(dL+dR) without 1/2 (solve.cpp:583-585)"),
    (5, "moving_wall_p2", "boundary lifting with 1/2 (solve.cpp:676-678)"),
    (5, "bump_freestream_p2q1", "boundary lifting with 1/2 (solve.cpp:676-678)"),
    (6, "moving_wall_p2", "boundary lifting A(SL,.) instead of A(Sb,.) (solve.cpp:667)"),
    (6, "moving_wall_p3", "boundary lifting A(SL,.) instead of A(Sb,.) (solve.cpp:667)"),
    (6, "bump_default_p1q2", "boundary lifting A(SL,.) instead of A(Sb,.) (solve.cpp:667)"),
]


@pytest.mark.parametrize("plant,name,what", PLANTS)
def test_planted_br2_mistakes_fail_against_reference(orc, plant, name, what):
    c, geo, prm = setup(orc, name)
    with orc.planted(plant):
        _, f = geo.residual(prm, c.U0)
    _, f_ok = geo.residual(prm, c.U0)
    e_bad, e_ok = rel(f, c.f), rel(f_ok, c.f)
    print("%s on %s: planted %.2e, clean %.2e" % (what, name, e_bad, e_ok))
    assert e_ok < TOL
    assert e_bad > 1e-6


def test_boundary_plants_cannot_act_without_boundary_faces(orc):
    """Control: a fully periodic mesh has no boundary faces, so the
    boundary-lifting plants leave its f unchanged."""
    c, geo, prm = setup(orc, "tgv3_p2")
    for plant in (5, 6):
        with orc.planted(plant):
            _, f = geo.residual(prm, c.U0)
        assert rel(f, c.f) < TOL


@pytest.mark.parametrize("name", ["config0", "tgv4_p3_long", "tgv8_p3", "tgv3_p3",
                                  "moving_wall_p3", "channel_q3_p3"])
def test_extended_precision_floor(orc, name):
    """The double evaluations (reference and oracle) against the same
    algorithm in extended precision.  The Mach 0.1 Taylor-Green cases sit at
    1e-12 (4^3 p = 3, 8^3 p = 2) to 4e-12 (8^3 p = 3) from the exact
    discrete residual (the reference's own f is that far from exact: the
    floor grows with p and with 1/h), the walled and moving cases two orders
    lower."""
    c, geo, prm = setup(orc, name)
    _, gx, _ = setup(orc, name, extended=True)
    Rx, fx = gx.residual(prm, c.U0)
    _, fo = geo.residual(prm, c.U0)
    e_ref, e_orc = rel(c.f, fx), rel(fo, fx)
    print("%s: reference vs exact %.2e, oracle vs exact %.2e" % (name, e_ref, e_orc))
    bound = 6e-12 if name == "tgv8_p3" else 2e-12
    assert e_ref < bound and e_orc < bound
    assert rel(c.R, Rx) < (2e-12 if name == "tgv8_p3" else 1e-12)
    if name in ("config0", "tgv4_p3_long", "tgv8_p3"):
        assert e_ref > 5e-13  # the ill-conditioned cases really are at the floor
    if name == "moving_wall_p3":
        assert e_ref < 1e-13


def test_conditioning_of_the_p3_taylor_green(orc):
    """A relative perturbation of 1e-15 of the state moves the exact
    (extended-precision) f of the 4^3 p = 3 Mach 0.1 Taylor-Green case by
    ~1e-12: its residual is ill-conditioned (the acoustic terms, c / h ~ 64
    times the pressure ~ 71, cancel to a small f), so double evaluations of
    it cannot agree beyond ~1e-12 whatever their order of operations."""
    c, _, prm = setup(orc, "tgv4_p3_long")
    _, gx, _ = setup(orc, "tgv4_p3_long", extended=True)
    _, f0 = gx.residual(prm, c.U0)
    rng = np.random.default_rng(7)
    U1 = c.U0 * (1.0 + 1e-15 * rng.uniform(-1.0, 1.0, c.U0.shape))
    _, f1 = gx.residual(prm, U1)
    moved = rel(f1, f0)
    print("f moved by %.2e under a 1e-15 relative state perturbation" % moved)
    assert 2e-13 < moved < 1e-11
