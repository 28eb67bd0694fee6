"""The general-order kernels (stage tile 2): every order the reference
accepts outside the tuned set, 0 <= p <= 8 and 1 <= q <= 8 (MAX_ORDER,
constants.cpp:21, geometry_distributed.cpp:819-830).

* Their results against the reference's own residual and RK steps at p = 0,
  4, 5, 6, 7, 8 and q = 2, 3, 4, 8 are in tests/test_gpu_reference_golden.py
  (the golden cases named *_p0, *_p4 ... *_q8_p8).
* Here: the same kernels at the tuned orders against the tuned kernels
  (same trace records, same algebra: agreement to rounding), the stepping
  and the per-step norms against the oracle on curved meshes with walls,
  FP32 trace records, the sampled outputs, ensembles with members of these
  orders, the tile rules, and the path that refuses them (the adjoint)."""
import numpy as np
import pytest

from dgpr_amd import cases, dgx

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def built():
    dgx.lib()
    return True


TUNED = {
    "tgv_affine_p3": lambda: cases.tgv_case(n=3, p=3, seed=81, M=0.2),
    "tgv_curved_p2": lambda: cases.tgv_case(nxyz=(4, 3, 3), p=2, seed=82, amp=0.2, M=0.2),
    "channel_p1": lambda: cases.channel_case(4, 3, 2, p=1, seed=83),
    "channel_q2_p2": lambda: cases.channel_case(3, 3, 2, p=2, q=2, seed=84),
    "channel_q3_p3": lambda: cases.channel_case(2, 3, 2, p=3, q=3, seed=85),
}


@pytest.mark.parametrize("name", sorted(TUNED))
def test_general_kernel_matches_tuned_kernels(built, name):
    """Tile 2 against the context's default tile on the same state: the
    residual and three steps with per-step norms agree to rounding."""
    c = TUNED[name]()
    print("case", name, "p", c.mesh.p, "q", c.mesh.q, flush=True)
    out = {}
    for tile in (None, 2):
        with dgx.Solver(c.mesh, c.params) as s:
            if tile is not None:
                s.tile = tile
            s.set_state(c.U0)
            R, f = s.residual()
            norms = s.step(c.dt, c.scheme, 3)
            out[tile] = (R, f, norms, s.get_state())
    (R0, f0, n0, U0), (R2, f2, n2, U2) = out[None], out[2]
    ef, eR, eU = rel(f2, f0), rel(R2, R0), rel(U2, U0)
    en = float(np.max(np.abs(n2 - n0) / np.abs(n0)))
    print("  f %.2e  R %.2e  U %.2e  norms %.2e" % (ef, eR, eU, en), flush=True)
    assert ef < 1e-12 and eR < 1e-12
    assert eU < 1e-13 and en < 1e-12


GENERAL = {
    "channel_p4": lambda: cases.channel_case(2, 3, 2, p=4, seed=91, nsteps=4),
    "channel_q4_p3": lambda: cases.channel_case(2, 2, 2, p=3, q=4, seed=92, nsteps=3),
    "tgv_p0": lambda: cases.tgv_case(n=4, p=0, seed=93, M=0.2, nsteps=5),
    "channel_p0_inviscid": lambda: cases.channel_case(4, 4, 2, p=0, seed=94, nsteps=5,
                                                      wall=dgx.INVISCID_WALL),
    # several elements per workgroup with a partly filled last workgroup:
    # 18 elements in workgroups of 8 (p = 1, q = 4), 5 in workgroups of 2
    # (p = 2, q = 5)
    "channel_q4_p1": lambda: cases.channel_case(3, 3, 2, p=1, q=4, seed=95, nsteps=3),
    "channel_q5_p2": lambda: cases.channel_case(5, 1, 1, p=2, q=5, seed=95, nsteps=3),
}


@pytest.mark.parametrize("name", sorted(GENERAL))
def test_general_orders_step_like_the_oracle(orc, built, name):
    c = GENERAL[name]()
    m = c.mesh
    print("case", name, "p", m.p, "q", m.q, c.scheme, c.nsteps, flush=True)
    geo = orc.Geometry(m.p, m.q, m.nx, m.ny, m.nz, m.bounds, m.nodes)
    prm = orc.params_from(c.params)
    R_o, f_o = geo.residual(prm, c.U0)
    U_o, n_o = geo.step(prm, c.U0, c.dt, c.scheme, c.nsteps)
    with dgx.Solver(m, c.params) as s:
        assert s.tile == 2
        s.set_state(c.U0)
        R_g, f_g = s.residual()
        n_g = s.step(c.dt, c.scheme, c.nsteps)
        U_g = s.get_state()
    ef, eU = rel(f_g, f_o), rel(U_g, U_o)
    en = float(np.max(np.abs(n_g - n_o) / np.abs(n_o)))
    print("  f %.2e  U %.2e  norms %.2e" % (ef, eU, en), flush=True)
    assert ef < 1e-11
    assert eU < 1e-10 and en < 1e-10


def test_general_orders_fp32_records(built):
    """The mixed forms at a general order store the trace records in FP32
    (the contractions stay FP64): close to the FP64 run, not equal to it."""
    c = GENERAL["channel_p4"]()
    res = {}
    for prec in ("fp64", "mixed"):
        with dgx.Solver(c.mesh, c.params) as s:
            s.precision = prec
            s.set_state(c.U0)
            s.step(c.dt, c.scheme, 4, norms=False)
            res[prec] = s.get_state()
    e = rel(res["mixed"], res["fp64"])
    print("  FP32 records vs FP64 after 4 steps: %.2e" % e, flush=True)
    assert 0.0 < e < 1e-6


@pytest.mark.parametrize("name,N", [("channel_p4", 3), ("tgv_p0", 2), ("channel_q4_p3", 4)])
def test_general_orders_sample_output(orc, built, name, N):
    c = GENERAL[name]()
    m = c.mesh
    with dgx.Solver(m, c.params) as s:
        s.set_state(c.U0)
        for var in ("mach", "rho", "u", "v", "w"):
            g = s.sample_output(var, N)
            o = orc.eval_output(m.p, c.U0, m.nelem, N, var, c.params.gamma)
            err = float(np.max(np.abs(g - o)) / np.max(np.abs(o)))
            print("  %s %-4s max rel err %.3e" % (name, var, err), flush=True)
            assert err < 1e-13, var


@pytest.mark.parametrize("layout", [1, 2, 0])
def test_ensemble_with_general_order_members(built, layout):
    """Ensemble members of orders outside the tuned set (p = 4, p = 0, p = 3
    with q = 4) next to tuned ones: in every layout, eagerly and replayed
    from the captured graph, each member ends where its single-case steps
    do (the general-order members run the same kernel: bit for bit)."""
    cs = [GENERAL["channel_p4"](), GENERAL["tgv_p0"](), GENERAL["channel_q4_p3"](),
          TUNED["tgv_curved_p2"](), TUNED["tgv_affine_p3"]()]
    nsteps = 3
    ref = []
    for c in cs:
        with dgx.Solver(c.mesh, c.params) as s:
            if c.mesh.p in (1, 2, 3) and c.mesh.q <= 3:
                s.tile = 0
            s.set_state(c.U0)
            s.step(c.dt, "RK4", nsteps, norms=False)
            ref.append(s.get_state())
    dts = [c.dt for c in cs]
    with dgx.Ensemble([c.mesh for c in cs], [c.params for c in cs]) as en:
        en.set_layout(layout)
        for i in range(3):
            assert en.member_tuning(i)[1] == 2
        for graph in (False, True):
            for m, c in zip(en.members, cs):
                m.set_state(c.U0)
            en.step("RK4", dts, nsteps, graph=graph)
            for i, (m, r) in enumerate(zip(en.members, ref)):
                e = rel(m.get_state(), r)
                print("  layout %d graph %s member %d (p=%d q=%d) rel err %.3e"
                      % (layout, graph, i, cs[i].mesh.p, cs[i].mesh.q, e), flush=True)
                if i < 3:
                    assert e == 0.0
                else:
                    assert e < 1e-14


def test_general_order_rules(built):
    """Tile 2 is the only tile outside the tuned orders (0 and 1 refused);
    occupancy targets and autotuning leave it alone; the adjoint refuses
    these orders."""
    c = GENERAL["channel_p4"]()
    with dgx.Solver(c.mesh, c.params) as s:
        assert s.tile == 2
        for t in (0, 1):
            with pytest.raises(dgx.DgxError) as ei:
                s.tile = t
            assert ei.value.status == dgx.ERR_UNSUPPORTED
        s.stage_waves = 3
        tile, waves, us = s.autotune_tiles(reps=1)
        assert tile == 2 and np.all(us == 0.0)
        s.set_state(c.U0)
        with pytest.raises(dgx.DgxError) as ei:
            s.residual_vjp(np.ones(c.mesh.state_size))
        assert ei.value.status == dgx.ERR_UNSUPPORTED
    # p = 9 is beyond the reference's MAX_ORDER
    m = c.mesh
    with pytest.raises(dgx.DgxError) as ei:
        dgx.Solver(dgx.Mesh(9, 1, m.nx, m.ny, m.nz, m.bounds, m.nodes), c.params)
    assert ei.value.status == dgx.ERR_UNSUPPORTED
