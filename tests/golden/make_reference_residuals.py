"""Golden vectors from the reference's own residual and RK steps.

Run here, where the reference sources exist, after ``python oracle/build.py``
has compiled ``oracle/_ref/ref_harness`` from them:

    python tests/golden/make_reference_residuals.py

For every case below it writes the mesh and initial state as a ``.dg`` file
(the reference's restart format, io.cpp:33-159), runs the harness on it
(read_state_distributed + precompute, residual() once, then N steps of
RK4_distributed / tvdRK3_distributed on one MPI rank) and stores the inputs
and outputs as data:

  tests/golden/reference_residuals.npz    <case>/nodes, <case>/U0 (inputs),
                                          <case>/R, <case>/f (residual at U0),
                                          <case>/UN (state after N steps),
                                          <case>/norms (RMS(R) per step, as
                                          residual_norm printed them)
  tests/golden/reference_residuals.json   per case: p, q, nx, ny, nz, bounds,
                                          free parameters, scheme, dt, N

The inputs come from the seeded generators of dgpr_amd.cases; the stored
copies make the fixtures self-contained.  The script needs the reference
sources, so it runs only where they are; the tests load the fixtures.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from dgpr_amd import cases, dgx  # noqa: E402
from oracle import build as obuild  # noqa: E402

NPZ = os.path.join(HERE, "reference_residuals.npz")
META = os.path.join(HERE, "reference_residuals.json")


def perturbed_uniform(mesh, prm, amp, seed):
    """The free stream of ``prm`` with seeded relative perturbations of
    density, momentum (scaled by c) and energy at every node."""
    rng = np.random.default_rng(seed)
    U = np.zeros((mesh.nelem, 5, mesh.nbf))
    for r in range(5):
        U[:, r, :] = prm.Ufr[r]
    U[:, 0, :] *= 1.0 + amp * rng.uniform(-1, 1, (mesh.nelem, mesh.nbf))
    for r in (1, 2, 3):
        U[:, r, :] += amp * U[:, 0, :] * prm.c * rng.uniform(-1, 1, (mesh.nelem, mesh.nbf))
    U[:, 4, :] *= 1.0 + amp * rng.uniform(-1, 1, (mesh.nelem, mesh.nbf))
    return U.reshape(-1)


def bump_case(p, q, bounds, seed, amp=0.02, mu=1e-3, source=None, name=None):
    """The reference's default bump channel (input.cpp:23-66) with a perturbed
    uniform state."""
    nx, ny, nz = 6, 4, 2
    nodes = dgx.bump_nodes(q, nx, ny, nz, 2.0, 1.0, 1.0, 0.0, 0.1, 1.0, 0.05)
    mesh = dgx.Mesh(p, q, nx, ny, nz, bounds, nodes)
    prm = dgx.make_reference_state(gamma=1.4, R=287.0, mu=mu, c=1.0, rho=1.0, u=0.1, v=0.0,
                                   w=0.0, eta=6.0, source=source)
    U0 = perturbed_uniform(mesh, prm, amp, seed)
    return cases.Case(name or "bump_p%d_q%d" % (p, q), mesh, prm, U0, "tvdRK3", 1e-3, 5)


def periodic_box_case(nxyz, p, seed):
    """Periodic [0, 2pi]^3 box of nx x ny x nz elements with a perturbed
    uniform state (Mach 0.2), RK4."""
    base = cases.tgv_case(nxyz=nxyz, p=p, seed=seed, M=0.2)
    U0 = perturbed_uniform(base.mesh, base.params, 0.05, seed)
    return cases.Case("box_%dx%dx%d_p%d" % (tuple(nxyz) + (p,)), base.mesh, base.params, U0,
                      "RK4", base.dt, 3)


def moving_wall_case(p, seed):
    """No-slip walls under a state that moves along them: the free stream
    (u = 0.3 c) with perturbations, so the wall trace has tangential velocity
    and the boundary lifting's A(Ub, .) differs from A(UL, .) in every row."""
    base = cases.channel_case(3, 3, 2, p=p, seed=seed, nsteps=3)
    prm = dgx.make_reference_state(gamma=1.4, R=287.0, mu=2e-2, c=base.params.c, rho=1.0,
                                   u=0.3 * base.params.c, v=0.0, w=0.0, eta=6.0)
    U0 = perturbed_uniform(base.mesh, prm, 0.05, seed)
    return cases.Case("moving_wall_p%d" % p, base.mesh, prm, U0, "tvdRK3", base.dt, 3)


DEFAULT_BOUNDS = [dgx.SUBSONIC_INFLOW, dgx.SUBSONIC_OUTFLOW, dgx.INVISCID_WALL,
                  dgx.INVISCID_WALL, dgx.PERIODIC, dgx.PERIODIC]

CASES = {
    # the 8^3 p=2 Taylor-Green workload at full size, RK4 dt=1e-3, 10 steps
    "config0": cases.config0,
    "tgv4_p1": lambda: cases.tgv_case(n=4, p=1, seed=11, nsteps=5),
    "tgv3_p2": lambda: cases.tgv_case(n=3, p=2, seed=12, nsteps=5),
    "tgv3_p3": lambda: cases.tgv_case(n=3, p=3, seed=13, nsteps=5),
    # p=3 Mach 0.1 Taylor-Green at the CFL of the 64^3 workload (the
    # ill-conditioned case): residual, and 100 RK4 steps
    "tgv4_p3_long": lambda: cases.tgv_case(n=4, p=3, seed=25, dt=4e-3, nsteps=100),
    # 8^3 p=3 Mach 0.1 Taylor-Green: the physics and order of the 64^3
    # workload on 1/512 of its elements, at its CFL (dt c / h), 10 RK4 steps
    "tgv8_p3": lambda: cases.tgv_case(n=8, p=3, seed=28, dt=2e-3, nsteps=10),
    "tgv_curved_p3": lambda: cases.tgv_case(nxyz=(3, 4, 2), p=3, seed=14, amp=0.15, nsteps=3),
    "tgv_curved_p2": lambda: cases.tgv_case(nxyz=(4, 3, 3), p=2, seed=24, amp=0.2, nsteps=3),
    "channel_p1": lambda: cases.channel_case(4, 4, 3, p=1, seed=15, nsteps=5),
    "channel_p2": lambda: cases.channel_case(3, 4, 2, p=2, seed=16, nsteps=4),
    "channel_p3": lambda: cases.channel_case(3, 3, 2, p=3, seed=17, nsteps=3),
    "channel_q2_p2": lambda: cases.channel_case(3, 3, 2, p=2, q=2, seed=18, nsteps=3),
    "channel_q3_p3": lambda: cases.channel_case(2, 3, 2, p=3, q=3, seed=19, nsteps=2),
    "channel_inviscid_p2": lambda: cases.channel_case(3, 3, 2, p=2, seed=20, nsteps=3,
                                                      wall=dgx.INVISCID_WALL),
    "moving_wall_p2": lambda: moving_wall_case(2, 26),
    "moving_wall_p3": lambda: moving_wall_case(3, 27),
    "bump_default_p1q2": lambda: bump_case(1, 2, DEFAULT_BOUNDS, 21),
    "bump_default_p2q2": lambda: bump_case(2, 2, DEFAULT_BOUNDS, 22),
    "bump_freestream_p2q1": lambda: bump_case(
        2, 1, [dgx.FREESTREAM, dgx.FREESTREAM, dgx.NO_SLIP_WALL, dgx.FREESTREAM,
               dgx.INVISCID_WALL, dgx.SUBSONIC_OUTFLOW], 23, source=[0.0, 0.3, -0.1, 0.05, 0.2]),
    "single_element_p2": lambda: periodic_box_case((1, 1, 1), 2, 61),
    "single_element_p3": lambda: periodic_box_case((1, 1, 1), 3, 62),
    "ragged_1x2x1_p3": lambda: periodic_box_case((1, 2, 1), 3, 63),
    "ragged_2x1x3_p1": lambda: periodic_box_case((2, 1, 3), 1, 64),
    # orders outside 1..3 (p = 0 .. 8, q up to 8): the general-order kernels
    "box_4x3x2_p0": lambda: periodic_box_case((4, 3, 2), 0, 71),
    "channel_p0": lambda: cases.channel_case(4, 4, 2, p=0, seed=72, nsteps=3),
    "tgv2_p4": lambda: cases.tgv_case(n=2, p=4, seed=73, nsteps=3),
    "bump_default_p4q2": lambda: bump_case(4, 2, DEFAULT_BOUNDS, 74),
    "channel_p5": lambda: cases.channel_case(2, 3, 2, p=5, seed=75, nsteps=2),
    "channel_q4_p2": lambda: cases.channel_case(2, 3, 2, p=2, q=4, seed=76, nsteps=2),
    "channel_q3_p6": lambda: cases.channel_case(2, 2, 2, p=6, q=3, seed=77, nsteps=2),
    "box_1x2x1_p7": lambda: periodic_box_case((1, 2, 1), 7, 78),
    "channel_q8_p8": lambda: cases.channel_case(1, 2, 1, p=8, q=8, seed=79, nsteps=1),
}


def free_parameters(prm):
    """Invert make_reference_state (solver_params.cpp:40-70) for the inputs
    the harness takes: rho and velocity from Ufr, c, gamma, mu, eta, source."""
    rho = prm.Ufr[0]
    return {"gamma": prm.gamma, "mu": prm.mu, "c": prm.c, "rho": rho,
            "u": prm.Ufr[1] / rho, "v": prm.Ufr[2] / rho, "w": prm.Ufr[3] / rho,
            "eta": prm.eta, "source": [prm.source[i] for i in range(5)]}


def run_case(name, case, tmp):
    fp = free_parameters(case.params)
    # the harness rebuilds the parameters with make_reference_state: check
    # that they reproduce ours exactly
    chk = dgx.make_reference_state(gamma=fp["gamma"], R=287.0, mu=fp["mu"], c=fp["c"],
                                   rho=fp["rho"], u=fp["u"], v=fp["v"], w=fp["w"],
                                   eta=fp["eta"], source=fp["source"])
    for f in ("gamma", "R", "mu", "c", "T", "P", "eta"):
        assert getattr(chk, f) == getattr(case.params, f), (name, f)
    for i in range(5):
        assert chk.Ufr[i] == case.params.Ufr[i] and chk.source[i] == case.params.source[i]
    dg = os.path.join(tmp, name + ".dg")
    out = os.path.join(tmp, name + ".bin")
    dgx.write_dg(dg, 0, case.mesh, case.U0)
    args = [obuild.REF_HARNESS, dg, out, case.scheme, repr(case.dt), str(case.nsteps)]
    args += [repr(float(fp[k])) for k in ("gamma", "mu", "c", "rho", "u", "v", "w", "eta")]
    args += [repr(float(s)) for s in fp["source"]]
    r = subprocess.run(args, capture_output=True, text=True, check=True)
    norms = [float(line.split(":")[1]) for line in r.stdout.splitlines() if " : " in line]
    assert len(norms) == case.nsteps, (name, r.stdout)
    raw = np.fromfile(out, dtype=np.float64)
    n = int(raw[:2].view(np.uint64)[0])
    assert n == case.mesh.state_size and raw.size == 2 + 3 * n
    R, f, UN = raw[2:2 + n], raw[2 + n:2 + 2 * n], raw[2 + 2 * n:]
    m = case.mesh
    meta = {"p": m.p, "q": m.q, "nx": m.nx, "ny": m.ny, "nz": m.nz,
            "bounds": [int(b) for b in m.bounds], "params": fp, "scheme": case.scheme,
            "dt": case.dt, "nsteps": case.nsteps}
    arrays = {"nodes": m.nodes, "U0": case.U0, "R": R, "f": f, "UN": UN,
              "norms": np.array(norms)}
    return meta, arrays


def main():
    if not os.path.exists(obuild.REF_HARNESS):
        obuild.build_reference(verbose=True)
    metas, arrays = {}, {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, make in CASES.items():
            meta, arr = run_case(name, make(), tmp)
            metas[name] = meta
            for k, v in arr.items():
                arrays[name + "/" + k] = v
            print("%-22s nelem %4d p %d q %d  %s x %d" % (
                name, meta["nx"] * meta["ny"] * meta["nz"], meta["p"], meta["q"],
                meta["scheme"], meta["nsteps"]), flush=True)
    np.savez_compressed(NPZ, **arrays)
    with open(META, "w") as fh:
        json.dump({"source": "the reference's residual() (solve.cpp:150-807) and "
                             "RK4_distributed / tvdRK3_distributed "
                             "(time_distributed.cpp:46-102), one MPI rank, via "
                             "oracle/ref_harness.cpp",
                   "cases": metas}, fh, indent=1)
    print("wrote", NPZ, os.path.getsize(NPZ), "bytes")


if __name__ == "__main__":
    main()
