"""The mixed-precision forms against the FP64 oracle, and their switches:
DGX_PREC_MIXED (FP32 face trace records; in the tiled p = 3 kernel the
volume gradient and flux projection in FP32 on the matrix cores) and
DGX_PREC_MIXED_BF16 (the same contractions as 3-term split-BF16 products)
and DGX_PREC_MIXED_RECORDS (FP32 records, every contraction in FP64).

Rounding the trace records to FP32 perturbs the face terms by about 6e-8 of
the trace values.  The residual error scales with the flux terms, not with
the residual itself, so where f is a small difference of large fluxes (low
Mach, near-steady flow) its relative error is large although the rounding
is not.  The residual bar is therefore relative to the conditioning of the
case: the change of the FP64 residual when the state itself is rounded to
FP32 (kappa).  Bars: residual and per-step RMS(R) norms within 8 kappa
(measured: 1.1-3.5 kappa), solution after N steps <= 1e-6 relative 2-norm
error; the FP64 form keeps its own 1e-12 / 1e-10 bars after switching back.
The mixed results must also differ from the FP64 ones (the FP32 records
are really in use)."""
import numpy as np
import pytest

from dgpr_amd import cases, dgx

pytestmark = pytest.mark.gpu

KAPPA_FACTOR = 8.0
SOL_TOL = 1e-6


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def bump(p, q):
    nx, ny, nz = 6, 4, 2
    nodes = dgx.bump_nodes(q, nx, ny, nz, 2.0, 1.0, 1.0, 0.3, 0.1, 1.0, 0.05)
    bounds = [dgx.SUBSONIC_INFLOW, dgx.SUBSONIC_OUTFLOW, dgx.NO_SLIP_WALL, dgx.INVISCID_WALL,
              dgx.PERIODIC, dgx.PERIODIC]
    mesh = dgx.Mesh(p, q, nx, ny, nz, bounds, nodes)
    prm = dgx.make_reference_state(u=0.2)
    U0 = dgx.turbulent_state(mesh, prm, 2.0)
    return cases.Case("bump_p%d_q%d" % (p, q), mesh, prm, U0, "tvdRK3", 5e-4, 6)


CASES = {
    "tgv_curved_p3": lambda: cases.tgv_case(n=3, p=3, seed=71, M=0.2, Re=800.0, amp=0.1,
                                            dt=1e-3, nsteps=6),
    # affine p = 3 (the 64^3 workload's kernels) at Mach 0.1
    "tgv4_p3": lambda: cases.tgv_case(n=4, p=3, seed=75, dt=4e-3, nsteps=10),
    "tgv4_p2": lambda: cases.tgv_case(n=4, p=2, seed=72, dt=2e-3, nsteps=6),
    "channel_p3": lambda: cases.channel_case(4, 3, 3, p=3, seed=73, nsteps=20),
    "channel_q2_p2": lambda: cases.channel_case(3, 3, 2, p=2, q=2, seed=74, nsteps=20),
    "bump_p3_q2": lambda: bump(3, 2),
    "bump_p1_q3": lambda: bump(1, 3),
}


@pytest.fixture(scope="module")
def built():
    dgx.lib()
    return True


def geometry(orc, m):
    return orc.Geometry(m.p, m.q, m.nx, m.ny, m.nz, m.bounds, m.nodes)


@pytest.mark.parametrize("form", ["mixed", "mixed_bf16", "mixed_records"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_residual_and_solution(name, form, orc, built):
    c = CASES[name]()
    m = c.mesh
    print("case", name, "nelem", m.nelem, "p", m.p, "q", m.q, c.scheme, "steps", c.nsteps,
          flush=True)
    geo = geometry(orc, m)
    oprm = orc.params_from(c.params)
    _, f_o = geo.residual(oprm, c.U0)
    U_o, n_o = geo.step(oprm, c.U0, c.dt, c.scheme, c.nsteps)
    with dgx.Solver(m, c.params) as s:
        s.set_state(c.U0.astype(np.float32).astype(np.float64))
        _, f_r = s.residual()
        s.set_state(c.U0)
        _, f64 = s.residual()
        s.precision = form
        assert s.precision == form
        _, fmx = s.residual()
        n_g = s.step(c.dt, c.scheme, c.nsteps)
        U_g = s.get_state()
    ef, eU = rel(fmx, f_o), rel(U_g, U_o)
    en = float(np.max(np.abs(n_g - n_o) / np.abs(n_o)))
    cond = rel(f_r, f64)
    print("  %s: rel err f %.3e  U %.3e  norms %.3e  (FP64 f %.3e; FP64 f of the state "
          "rounded to FP32 %.3e)" % (form, ef, eU, en, rel(f64, f_o), cond), flush=True)
    assert ef < KAPPA_FACTOR * cond and en < KAPPA_FACTOR * cond
    assert eU < SOL_TOL
    assert rel(fmx, f64) > 1e-12


def test_switching_back_to_fp64_restores_parity(orc, built):
    c = CASES["channel_p3"]()
    m = c.mesh
    geo = geometry(orc, m)
    oprm = orc.params_from(c.params)
    U_o, _ = geo.step(oprm, c.U0, c.dt, c.scheme, 4)
    with dgx.Solver(m, c.params) as s:
        s.set_state(c.U0)
        s.precision = "mixed"
        s.step(c.dt, c.scheme, 2)
        s.set_state(c.U0)
        s.precision = "fp64"
        s.step(c.dt, c.scheme, 4)
        U_g = s.get_state()
        _, f_g = s.residual()
    _, f_o = geo.residual(oprm, U_g)
    assert rel(U_g, U_o) < 1e-10
    assert rel(f_g, f_o) < 1e-12


def test_mixed_state_switch_mid_run(orc, built):
    """Switching form between steps recomputes the records of the current
    state: two mixed steps then two FP64 steps stay within the mixed bar of
    four FP64 steps."""
    c = CASES["tgv4_p2"]()
    m = c.mesh
    U_o, _ = geometry(orc, m).step(orc.params_from(c.params), c.U0, c.dt, c.scheme, 4)
    with dgx.Solver(m, c.params) as s:
        s.set_state(c.U0)
        s.precision = "mixed"
        s.step(c.dt, c.scheme, 2)
        s.precision = "fp64"
        s.step(c.dt, c.scheme, 2)
        assert rel(s.get_state(), U_o) < SOL_TOL


def test_mixed_form_refuses_adjoint_and_ensembles(built):
    c = cases.tgv_case(n=2, p=2, seed=75)
    with dgx.Solver(c.mesh, c.params) as s:
        s.set_state(c.U0)
        s.precision = "mixed"
        with pytest.raises(dgx.DgxError) as ei:
            s.adjoint_rk4(c.dt, 2)
        assert ei.value.status == dgx.ERR_UNSUPPORTED
        with pytest.raises(dgx.DgxError) as ei:
            s.residual_vjp(np.ones(s.nstate))
        assert ei.value.status == dgx.ERR_UNSUPPORTED
        with pytest.raises(KeyError):
            s.precision = "fp16"
    with dgx.Ensemble([c.mesh, c.mesh], [c.params, c.params]) as en:
        en.members[1].precision = "mixed"
        with pytest.raises(dgx.DgxError) as ei:
            en.step("RK4", [c.dt, c.dt], 1)
        assert ei.value.status == dgx.ERR_UNSUPPORTED


def test_matrix_core_forms_differ_at_p3(built):
    """At p = 3 the two mixed forms run different contraction arithmetic
    (FP32 vs split-BF16 matrix-core products): their residuals differ from
    each other and from the FP32-records-only one-row kernel, and, each being
    within KAPPA_FACTOR kappa of the FP64 residual, from one another by at
    most twice that."""
    c = CASES["tgv4_p3"]()
    with dgx.Solver(c.mesh, c.params) as s:
        s.set_state(c.U0.astype(np.float32).astype(np.float64))
        _, f_r = s.residual()
        s.set_state(c.U0)
        _, f64 = s.residual()
        out = {}
        for form in ("mixed", "mixed_bf16"):
            s.precision = form
            _, out[form] = s.residual()
        s.tile = 0
        s.precision = "mixed"
        _, rec_only = s.residual()
    cond = rel(f_r, f64)
    d1 = rel(out["mixed"], out["mixed_bf16"])
    d2 = rel(out["mixed"], rec_only)
    print("  mixed vs mixed_bf16 %.3e, mixed vs FP32 records only %.3e, kappa %.3e"
          % (d1, d2, cond), flush=True)
    for form in out:
        print("  %s vs FP64 %.3e" % (form, rel(out[form], f64)), flush=True)
    assert d1 > 0.0 and d2 > 0.0
    assert d1 < 2 * KAPPA_FACTOR * cond and d2 < 2 * KAPPA_FACTOR * cond
