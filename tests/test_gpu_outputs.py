"""Derived outputs on the device against the oracle: eval_output
(render/vrtxgen.cpp:25-59) of every output variable on equispaced sub-grids,
and the driver's running mean (output.cpp:39-49, dg.cpp:123-131) over
steps."""
import numpy as np
import pytest

from dgpr_amd import cases, dgx

pytestmark = pytest.mark.gpu

OUT_TOL = 1e-13


@pytest.fixture(scope="module")
def built():
    dgx.lib()
    return True


@pytest.mark.parametrize("name,N", [("channel_p2", 3), ("tgv_curved_p3", 5), ("channel_p1", 2)])
def test_sample_output_matches_oracle(orc, built, name, N):
    case = {"channel_p2": lambda: cases.channel_case(3, 4, 2, p=2, seed=16),
            "tgv_curved_p3": lambda: cases.tgv_case(nxyz=(3, 4, 2), p=3, seed=14, amp=0.15),
            "channel_p1": lambda: cases.channel_case(4, 4, 3, p=1, seed=15)}[name]()
    m = case.mesh
    print("case", name, "N", N, flush=True)
    with dgx.Solver(m, case.params) as s:
        s.set_state(case.U0)
        for var in ("mach", "rho", "u", "v", "w"):
            g = s.sample_output(var, N)
            o = orc.eval_output(m.p, case.U0, m.nelem, N, var, case.params.gamma)
            err = float(np.max(np.abs(g - o)) / np.max(np.abs(o)))
            print("  %-4s max rel err %.3e" % (var, err), flush=True)
            assert err < OUT_TOL, var


def test_running_mean_matches_oracle(orc, built):
    case = cases.channel_case(3, 4, 2, p=2, seed=16)
    m = case.mesh
    start, nsteps = 3, 6
    mean_o = np.zeros(m.state_size)
    with dgx.Solver(m, case.params) as s:
        s.set_state(case.U0)
        with pytest.raises(dgx.DgxError):
            s.get_mean()
        for t in range(nsteps):
            if t >= start:
                s.update_mean(t, start)
                mean_o = orc.update_mean(s.get_state(), mean_o, t, start)
            s.step(case.dt, case.scheme, 1, norms=False)
        M = s.get_mean()
        err = float(np.linalg.norm(M - mean_o) / np.linalg.norm(mean_o))
        print("  running mean rel err %.3e" % err, flush=True)
        assert err < 1e-14
        # restarting at a later mean_start begins a new average
        s.update_mean(10, 10)
        U = s.get_state()
        assert float(np.linalg.norm(s.get_mean() - U) / np.linalg.norm(U)) < 1e-15
        with pytest.raises(dgx.DgxError):
            s.update_mean(2, 5)
