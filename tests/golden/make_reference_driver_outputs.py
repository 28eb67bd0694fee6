"""Golden outputs of the reference's own CPU driver (solver/dg.cpp).

Run here, where the reference sources exist, after ``python oracle/build.py``
has compiled ``oracle/_ref/dg`` from them:

    python tests/golden/make_reference_driver_outputs.py

For each case below it runs the driver with the given options on one MPI
rank, reads the .dg file it writes and stores, as data:

  tests/golden/reference_driver_outputs.npz   <case>/nodes (generate_bump_nodes,
                                              geometry_common.cpp:494-576),
                                              <case>/state (uniform or
                                              turbulent initial state, or the
                                              state after N steps), <case>/mean
                                              (the running mean, dg.cpp:122-131),
                                              <case>/norms (the printed
                                              "tstep : norm" lines),
                                              <case>/file (the .dg file's
                                              bytes, for the layout check)
  tests/golden/reference_driver_outputs.json  per case: options, tstep, keys

The script needs the reference sources, so it runs only where they are;
the tests load the fixtures.
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

from dgpr_amd import dgx  # noqa: E402
from oracle import build as obuild  # noqa: E402

NPZ = os.path.join(HERE, "reference_driver_outputs.npz")
META = os.path.join(HERE, "reference_driver_outputs.json")

CASES = {
    # the driver's defaults (input.cpp:20-66): p=1, q=2, 6x4x4 bump channel,
    # uniform state; no steps
    "bump_default": ["-niter", "0"],
    # stretched wall-normal spacing, a taller bump, q=3 geometry, the
    # turbulent initial state (solver_init.cpp:30-73); no steps
    "bump_stretched_turb": ["-p", "2", "-q", "3", "-nx", "5", "-ny", "4", "-nz", "2",
                            "-a", "0.4", "-bmp_h", "0.15", "-turb", "-niter", "0"],
    # three RK4 steps with the running mean from step 1 on: "state" and "mean"
    "mean_run": ["-p", "2", "-q", "2", "-nx", "4", "-ny", "3", "-nz", "2", "-a", "0.4",
                 "-turb", "-scheme", "RK4", "-dt", "2e-4", "-niter", "3", "-mean_start", "1",
                 "-chkint", "1000"],
    # orders outside 1..3: the p = 0 turbulent start, and three RK4 steps at
    # p = 4 on q = 2 geometry
    "p0_turb": ["-p", "0", "-q", "1", "-nx", "4", "-ny", "3", "-nz", "2", "-turb",
                "-niter", "0"],
    "p4_run": ["-p", "4", "-q", "2", "-nx", "3", "-ny", "2", "-nz", "2", "-turb", "-scheme",
               "RK4", "-dt", "2e-4", "-niter", "3", "-chkint", "1000"],
}


def main():
    if not os.path.exists(obuild.REF_DG):
        obuild.build_reference(verbose=True)
    metas, arrays = {}, {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, args in CASES.items():
            out = os.path.join(tmp, name)
            r = subprocess.run([obuild.REF_DG] + args + ["-ofile", out], capture_output=True,
                               text=True, check=True, cwd=tmp)
            norms = [float(ln.split(":")[1]) for ln in r.stdout.splitlines() if " : " in ln]
            path = out + ".dg"
            ts, mesh, outs = dgx.read_dg_outputs(path)
            arrays[name + "/nodes"] = mesh.nodes
            for k, v in outs.items():
                arrays[name + "/" + k] = v
            arrays[name + "/norms"] = np.array(norms)
            if name == "mean_run":
                arrays[name + "/file"] = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
            metas[name] = {"args": args, "tstep": ts, "keys": list(outs),
                           "p": mesh.p, "q": mesh.q, "nx": mesh.nx, "ny": mesh.ny, "nz": mesh.nz,
                           "bounds": [int(b) for b in mesh.bounds]}
            print(name, "tstep", ts, "outputs", list(outs), "norms", len(norms), flush=True)
    np.savez_compressed(NPZ, **arrays)
    with open(META, "w") as fh:
        json.dump({"source": "the reference's CPU driver solver/dg.cpp, one MPI rank, "
                             "built by oracle/build.py", "cases": metas}, fh, indent=1)
    print("wrote", NPZ, os.path.getsize(NPZ), "bytes")


if __name__ == "__main__":
    main()
