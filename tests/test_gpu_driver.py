"""dgx_drive, the solver's time-stepping program (solver/cudg.cu:24-151), on
the device, against the oracle:
  * a run from the generated bump channel with the turbulent initial state
    prints one "tstep : norm" line per step and writes the checkpoints
    <ofile><tstep>.dg before each step with (tstep + 1) % chkint == 0 and
    <ofile>.dg at the end; states and norms match the oracle stepping the
    same initial state (relative 2-norm <= 1e-10, norms <= 1e-10);
  * a restart from a checkpoint continues from the step after the one the
    file holds (cudg.cu:67), so "-ifile run2" replays steps 3.. on the state
    stored at step 2 (the state after two steps);
  * a run that goes unstable prints the NaN norm, reports "looks like
    things went unstable!" and exits with status 1 without a final file.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from dgpr_amd import build, dgx

pytestmark = pytest.mark.gpu

TOL = 1e-10
LINE = re.compile(r"^(\d+) : (\S+)$")


@pytest.fixture(scope="module")
def driver():
    dgx.lib()
    if not os.path.exists(build.DRIVER):
        raise RuntimeError("dgx_drive is not built (python -m dgpr_amd.build)")
    return build.DRIVER


def run(driver, args, timeout=120):
    print("dgx_drive", " ".join(args), flush=True)
    r = subprocess.run([driver] + args, capture_output=True, text=True, timeout=timeout)
    norms = {}
    for ln in r.stdout.splitlines():
        m = LINE.match(ln)
        if m:
            norms[int(m.group(1))] = float(m.group(2))
    return r, norms


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASE = dict(p=2, q=2, nx=4, ny=3, nz=2, lx=2.0, ly=1.0, lz=1.0, a=0.0, bmp_h=0.1, bmp_c=1.0,
            bmp_v=0.05, dt=2e-4)


def case_args(c):
    return ["-p", str(c["p"]), "-q", str(c["q"]), "-nx", str(c["nx"]), "-ny", str(c["ny"]),
            "-nz", str(c["nz"]), "-dt", repr(c["dt"]), "-scheme", "RK4", "-turb"]


def oracle_reference(orc, c, nsteps):
    """The oracle's states after 0..nsteps steps and its per-step norms."""
    nodes = dgx.bump_nodes(c["q"], c["nx"], c["ny"], c["nz"], c["lx"], c["ly"], c["lz"], c["a"],
                           c["bmp_h"], c["bmp_c"], c["bmp_v"])
    bounds = [dgx.SUBSONIC_INFLOW, dgx.SUBSONIC_OUTFLOW, dgx.INVISCID_WALL, dgx.INVISCID_WALL,
              dgx.PERIODIC, dgx.PERIODIC]
    prm = dgx.make_reference_state()
    geo = orc.Geometry(c["p"], c["q"], c["nx"], c["ny"], c["nz"], bounds, nodes)
    oprm = orc.params_from(prm)
    U = geo.turbulent(oprm, c["lx"])
    states, norms = [U], []
    for _ in range(nsteps):
        U, n = geo.step(oprm, U, c["dt"], "RK4", 1)
        states.append(U)
        norms.append(n[0])
    return nodes, bounds, states, norms


def test_driver_run_checkpoints_and_restart(orc, driver, tmp_path):
    c = CASE
    out = str(tmp_path / "run")
    r, norms = run(driver, case_args(c) + ["-niter", "6", "-chkint", "3", "-ofile", out])
    assert r.returncode == 0, r.stdout + r.stderr
    nodes, bounds, states, onorms = oracle_reference(orc, c, 6)

    assert sorted(norms) == list(range(6))
    en = max(abs(norms[t] - onorms[t]) / onorms[t] for t in range(6))
    print("  norm rel err %.3e" % en, flush=True)
    assert en < TOL

    # checkpoints before steps 2 and 5, the final state after step 5
    for path, tstep, nsteps in ((out + "2.dg", 2, 2), (out + "5.dg", 5, 5), (out + ".dg", 5, 6)):
        ts, mesh, U = dgx.read_dg(path)
        assert ts == tstep
        assert (mesh.p, mesh.q, mesh.nx, mesh.ny, mesh.nz) == (c["p"], c["q"], c["nx"], c["ny"],
                                                               c["nz"])
        assert list(mesh.bounds) == bounds
        assert np.array_equal(mesh.nodes, nodes)
        e = rel(U, states[nsteps])
        print("  %s rel err %.3e" % (os.path.basename(path), e), flush=True)
        assert e < TOL

    # restart from the step-2 checkpoint: steps 3, 4, 5 on the state after
    # two steps, ending on the state the step-5 checkpoint holds
    out2 = str(tmp_path / "re")
    r, norms2 = run(driver, ["-ifile", out + "2", "-niter", "3", "-dt", repr(c["dt"]),
                             "-scheme", "RK4", "-ofile", out2])
    assert r.returncode == 0, r.stdout + r.stderr
    assert sorted(norms2) == [3, 4, 5]
    assert abs(norms2[3] - onorms[2]) / onorms[2] < TOL
    ts, mesh, U = dgx.read_dg(out2 + ".dg")
    assert ts == 5
    assert rel(U, states[5]) < TOL
    _, _, U5 = dgx.read_dg(out + "5.dg")
    assert rel(U, U5) < 1e-12


def test_driver_tvdrk3_uniform_free_stream(orc, driver, tmp_path):
    """The default scheme (tvdRK3) on the uniform state of a flat periodic
    box keeps the free stream: norms at round-off, state unchanged."""
    out = str(tmp_path / "fs")
    r, norms = run(driver, ["-p", "3", "-q", "1", "-nx", "3", "-ny", "3", "-nz", "3",
                            "-bmp_h", "0", "-bmx", "periodic", "-bpx", "periodic",
                            "-bmy", "periodic", "-bpy", "periodic", "-niter", "4",
                            "-ofile", out])
    assert r.returncode == 0, r.stdout + r.stderr
    assert sorted(norms) == [0, 1, 2, 3]
    assert max(norms.values()) < 1e-13
    ts, mesh, U = dgx.read_dg(out + ".dg")
    prm = dgx.make_reference_state()
    Ur = U.reshape(mesh.nelem, 5, mesh.nbf)
    for k in range(5):
        assert np.max(np.abs(Ur[:, k] - prm.Ufr[k])) < 1e-13


def test_driver_reports_instability(driver, tmp_path):
    """A time step far above the stable limit on the turbulent state of a
    periodic box: the NaN norm is printed, the run stops with status 1 and
    the final file is not written (residual_norm, time_distributed.cpp:38-41)."""
    out = str(tmp_path / "blow")
    r, norms = run(driver, ["-p", "2", "-q", "1", "-nx", "3", "-ny", "3", "-nz", "3",
                            "-bmp_h", "0", "-bmx", "periodic", "-bpx", "periodic",
                            "-bmy", "periodic", "-bpy", "periodic", "-turb", "-dt", "20.0",
                            "-niter", "200", "-ofile", out])
    assert r.returncode == 1, r.stdout
    assert "looks like things went unstable!" in r.stdout
    last = max(norms)
    assert np.isnan(norms[last]) and last < 199
    assert not os.path.exists(out + ".dg")


def test_driver_running_mean_and_restart(orc, driver, tmp_path):
    """-mean_start 2: from step 2 on the driver keeps the running mean of
    the state (update_mean after each step, output.cpp:39-49, dg.cpp:122-131)
    and writes it as the "mean" output next to "state"; the oracle's
    update_mean over its own states gives the same mean.  A restart from
    the checkpoint before step 4 reads the mean and continues it."""
    c = CASE
    out = str(tmp_path / "avg")
    r, _ = run(driver, case_args(c) + ["-niter", "6", "-chkint", "4", "-mean_start", "2",
                                       "-ofile", out])
    assert r.returncode == 0, r.stdout + r.stderr
    _, _, states, _ = oracle_reference(orc, c, 6)
    mean = np.zeros_like(states[0])
    means = {}
    for t in range(6):  # step t produces states[t + 1]
        if t >= 2:
            mean = orc.update_mean(states[t + 1], mean, t, 2)
        means[t] = mean.copy()
    ts, _, outs = dgx.read_dg_outputs(out + ".dg")
    assert ts == 5 and list(outs) == ["state", "mean"]
    e = rel(outs["mean"], means[5])
    print("  final mean rel err %.3e" % e, flush=True)
    assert e < TOL and rel(outs["state"], states[6]) < TOL
    # the checkpoint before step 3 holds the mean after steps 2 only
    ts, _, outs3 = dgx.read_dg_outputs(out + "3.dg")
    assert ts == 3 and rel(outs3["mean"], means[2]) < TOL
    # restart from it: the run resumes at step 4 (the step after the one the
    # file holds, cudg.cu:67) on the state after three steps, and steps 4
    # and 5 continue the mean read from the file with their own step numbers
    out2 = str(tmp_path / "avg2")
    r, _ = run(driver, ["-ifile", out + "3", "-niter", "2", "-dt", repr(c["dt"]),
                        "-scheme", "RK4", "-mean_start", "2", "-ofile", out2])
    assert r.returncode == 0, r.stdout + r.stderr
    m = orc.update_mean(states[4], means[2], 4, 2)
    m = orc.update_mean(states[5], m, 5, 2)
    ts, _, outs2 = dgx.read_dg_outputs(out2 + ".dg")
    assert ts == 5 and list(outs2) == ["state", "mean"]
    assert rel(outs2["state"], states[5]) < TOL and rel(outs2["mean"], m) < TOL


def test_driver_general_order_matches_reference_driver(driver, tmp_path):
    """dgx_drive -p 4 -q 2 (the general-order kernels) against the
    reference's own CPU driver run of the same options
    (tests/golden/reference_driver_outputs.*, case p4_run: turbulent start on
    a 3x2x2 bump channel, three RK4 steps): same norms, same final state
    (1e-10 relative)."""
    import json
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    meta = json.load(open(os.path.join(gdir, "reference_driver_outputs.json")))["cases"]["p4_run"]
    z = np.load(os.path.join(gdir, "reference_driver_outputs.npz"), allow_pickle=False)
    out = str(tmp_path / "p4")
    r, norms = run(driver, meta["args"] + ["-ofile", out])
    assert r.returncode == 0, r.stdout + r.stderr
    ref_norms = z["p4_run/norms"]
    assert sorted(norms) == list(range(len(ref_norms)))
    en = max(abs(norms[t] - ref_norms[t]) / ref_norms[t] for t in range(len(ref_norms)))
    ts, mesh, outs = dgx.read_dg_outputs(out + ".dg")
    eU = rel(outs["state"], z["p4_run/state"])
    print("  norms %.2e  state %.2e" % (en, eU), flush=True)
    assert ts == meta["tstep"] and list(outs) == meta["keys"] and mesh.p == 4 and mesh.q == 2
    assert en < TOL and eU < TOL


def test_driver_matches_reference_driver_with_mean(driver, tmp_path):
    """The options of the reference's own CPU driver run held in
    tests/golden/reference_driver_outputs.* (solver/dg.cpp: turbulent start on
    a stretched bump channel, three RK4 steps, running mean from step 1):
    dgx_drive prints the same norms and writes "state" and "mean" in the same
    order and to the same values (1e-10 relative)."""
    import json
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    meta = json.load(open(os.path.join(gdir, "reference_driver_outputs.json")))["cases"]
    z = np.load(os.path.join(gdir, "reference_driver_outputs.npz"), allow_pickle=False)
    out = str(tmp_path / "ref_opts")
    r, norms = run(driver, meta["mean_run"]["args"] + ["-ofile", out])
    assert r.returncode == 0, r.stdout + r.stderr
    ref_norms = z["mean_run/norms"]
    assert sorted(norms) == list(range(len(ref_norms)))
    en = max(abs(norms[t] - ref_norms[t]) / ref_norms[t] for t in range(len(ref_norms)))
    ts, mesh, outs = dgx.read_dg_outputs(out + ".dg")
    eU, eM = rel(outs["state"], z["mean_run/state"]), rel(outs["mean"], z["mean_run/mean"])
    print("  norms %.2e  state %.2e  mean %.2e" % (en, eU, eM), flush=True)
    assert ts == meta["mean_run"]["tstep"] and list(outs) == meta["mean_run"]["keys"]
    assert np.array_equal(mesh.nodes, z["mean_run/nodes"])
    assert en < TOL and eU < TOL and eM < TOL
