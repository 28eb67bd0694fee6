"""Batched ensembles: members of different size, p, q, boundary conditions
and parameters advanced together (in every layout: one launch per stage for
all orders, one per order, or one batched launch per (p, q, affine) group;
optionally replayed from a captured hipGraph) must reproduce the
single-case stepper on each member, and the oracle.  The members include
affine and curved q = 1 meshes of the same p = 2, which form two groups
(constant-Jacobian and general kernels)."""
import numpy as np
import pytest

from dgpr_amd import cases, dgx

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def members():
    return [
        cases.tgv_case(n=3, p=3, seed=41, M=0.1),
        cases.tgv_case(n=4, p=1, seed=42, M=0.2, Re=400.0),
        cases.channel_case(3, 4, 2, p=2, seed=43),
        cases.tgv_case(nxyz=(4, 3, 3), p=2, seed=44, amp=0.2, M=0.15),
        cases.channel_case(3, 3, 2, p=2, q=2, seed=45),
        cases.tgv_case(n=2, p=3, seed=46, M=0.3, Re=5000.0),
        cases.tgv_case(n=3, p=2, seed=47, M=0.2),
    ]


@pytest.fixture(scope="module")
def built():
    dgx.lib()
    return True


def single_runs(cs, scheme, nsteps, tiles=None):
    """Single-case steps of every member with the stage tile the ensemble
    runs it with (tiles[i]; by default 0, the one-row kernel of the batched
    launches and of the ensemble stage kernel before tuning)."""
    out = []
    for i, c in enumerate(cs):
        with dgx.Solver(c.mesh, c.params) as s:
            s.tile = tiles[i] if tiles else 0
            s.set_state(c.U0)
            s.step(c.dt, scheme, nsteps, norms=False)
            out.append(s.get_state())
    return out


@pytest.mark.parametrize("scheme,layout", [("RK4", 1), ("tvdRK3", 1), ("RK4", 2), ("RK4", 0),
                                           ("tvdRK3", 0)])
def test_ensemble_matches_single_cases(built, scheme, layout):
    """Every layout: one launch for all orders (1), one per order
    (2), batched launches per (p, q, affine) (0)."""
    cs = members()
    nsteps = 3
    ref = single_runs(cs, scheme, nsteps)
    dts = [c.dt for c in cs]
    with dgx.Ensemble([c.mesh for c in cs], [c.params for c in cs]) as en:
        en.set_layout(layout)
        for i, c in enumerate(cs):
            p, t, w = en.member_tuning(i)
            assert p == (layout if c.mesh.q == 1 else 0)
        for m, c in zip(en.members, cs):
            m.set_state(c.U0)
        en.step(scheme, dts, nsteps, graph=False)
        for i, (m, r) in enumerate(zip(en.members, ref)):
            e = rel(m.get_state(), r)
            print("  %s member %d (p=%d q=%d) eager rel err %.3e" % (scheme, i, cs[i].mesh.p, cs[i].mesh.q, e), flush=True)
            assert e < 1e-14
        # graph replay: two calls with eager work on a member between them
        for m, c in zip(en.members, cs):
            m.set_state(c.U0)
        en.step(scheme, dts, 2, graph=True)
        en.members[0].residual()
        en.step(scheme, dts, nsteps - 2, graph=True)
        for i, (m, r) in enumerate(zip(en.members, ref)):
            e = rel(m.get_state(), r)
            print("  %s member %d graph rel err %.3e" % (scheme, i, e), flush=True)
            assert e < 1e-14


def test_ensemble_member_matches_oracle(orc, built):
    cs = members()
    c = cs[2]
    m = c.mesh
    geo = orc.Geometry(m.p, m.q, m.nx, m.ny, m.nz, m.bounds, m.nodes)
    U_o, _ = geo.step(orc.params_from(c.params), c.U0, c.dt, "RK4", 3)
    with dgx.Ensemble([x.mesh for x in cs], [x.params for x in cs]) as en:
        for mm, x in zip(en.members, cs):
            mm.set_state(x.U0)
        en.step("RK4", [x.dt for x in cs], 3, graph=True)
        e = rel(en.members[2].get_state(), U_o)
    print("  member 2 vs oracle rel err %.3e" % e, flush=True)
    assert e < 1e-10


def test_ensemble_autotune_keeps_results(built):
    """The autotuner (occupancy targets; tiles of the ensemble stage kernel)
    leaves the states alone, and its choice gives the single-case results
    of the same tile to rounding."""
    cs = members()
    dts = [c.dt for c in cs]
    with dgx.Ensemble([c.mesh for c in cs], [c.params for c in cs]) as en:
        for m, c in zip(en.members, cs):
            m.set_state(c.U0)
        w = en.autotune("RK4", dts, reps=2)
        tun = [en.member_tuning(i) for i in range(len(cs))]
        print("  autotuned waves per member", w, "tuning", tun, flush=True)
        assert all(x in (1, 2, 3) for x in w)
        assert all(t[1] in (0, 1) for t in tun)
        # one layout for all q = 1 members; q > 1 members always batched
        assert len({t[0] for t, c in zip(tun, cs) if c.mesh.q == 1}) == 1
        assert all(t[0] == 0 for t, c in zip(tun, cs) if c.mesh.q != 1)
        en.step("RK4", dts, 2, graph=True)
        got = [m.get_state() for m in en.members]
    ref = single_runs(cs, "RK4", 2, tiles=[t[1] for t in tun])
    for g, r in zip(got, ref):
        assert rel(g, r) < 1e-14


def test_member_set_params_between_ensemble_steps(built):
    """dgx_set_params on a member between ensemble steps: the next steps use
    the new physics (the item tables and the captured graph are rebuilt),
    as single-case steps with those parameters do (to rounding), with and
    without the graph."""
    cs = members()
    dts = [c.dt for c in cs]
    k = 2
    prm2 = cs[k].params.copy()
    prm2.mu = 4.0 * cs[k].params.mu
    with dgx.Solver(cs[k].mesh, cs[k].params) as s:
        s.tile = 0
        s.set_state(cs[k].U0)
        s.step(cs[k].dt, "RK4", 1, norms=False)
        s.set_params(prm2)
        s.step(cs[k].dt, "RK4", 2, norms=False)
        ref = s.get_state()
    for graph in (False, True):
        with dgx.Ensemble([c.mesh for c in cs], [c.params for c in cs]) as en:
            for m, c in zip(en.members, cs):
                m.set_state(c.U0)
            en.step("RK4", dts, 1, graph=graph)
            en.members[k].set_params(prm2)
            en.step("RK4", dts, 2, graph=graph)
            got = en.members[k].get_state()
        print("  graph %s: rel err %.3e" % (graph, rel(got, ref)), flush=True)
        assert rel(got, ref) < 1e-14
