"""Host-side helpers against the reference's own CPU driver output files
(tests/golden/reference_driver_outputs.*, written by solver/dg.cpp through
tests/golden/make_reference_driver_outputs.py):

* dgx_bump_nodes reproduces generate_bump_nodes (geometry_common.cpp:494-576)
  bit for bit, with and without wall-normal stretching, at q = 2 and 3;
* the initial states: uniform (solver_init.cpp:22-28) exactly, turbulent
  (solver_init.cpp:30-73) to 1e-15 of its scale, also at p = 0 (one node
  per element, at the reference origin);
* dgx_write_dg_outputs, given the driver's "state" and "mean", writes the
  driver's file byte for byte (io.cpp:33-89, outputs in the map's order).
The GPU driver's run of the same options is compared in
tests/test_gpu_driver.py."""
import json
import os

import numpy as np
import pytest

from dgpr_amd import dgx

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(HERE, "reference_driver_outputs.json")))["cases"]
Z = np.load(os.path.join(HERE, "reference_driver_outputs.npz"), allow_pickle=False)


def opts(name):
    """The driver's defaults (input.cpp:20-66) overridden by the case's options."""
    d = dict(p=1, q=2, nx=6, ny=4, nz=4, lx=2.0, ly=1.0, lz=1.0, a=0.0, bmp_h=0.1, bmp_c=1.0,
             bmp_v=0.05, turb=False)
    args = META[name]["args"]
    i = 0
    while i < len(args):
        k = args[i].lstrip("-")
        if k == "turb":
            d["turb"] = True
            i += 1
            continue
        d[k] = args[i + 1]
        i += 2
    for k in ("p", "q", "nx", "ny", "nz"):
        d[k] = int(d[k])
    for k in ("lx", "ly", "lz", "a", "bmp_h", "bmp_c", "bmp_v"):
        d[k] = float(d[k])
    return d


def mesh_of(name):
    d = opts(name)
    nodes = dgx.bump_nodes(d["q"], d["nx"], d["ny"], d["nz"], d["lx"], d["ly"], d["lz"], d["a"],
                           d["bmp_h"], d["bmp_c"], d["bmp_v"])
    return d, dgx.Mesh(d["p"], d["q"], d["nx"], d["ny"], d["nz"], META[name]["bounds"], nodes)


@pytest.mark.parametrize("name", sorted(META))
def test_bump_nodes_match_reference_bitwise(name):
    d, m = mesh_of(name)
    ref = Z[name + "/nodes"]
    assert m.nodes.shape == ref.shape
    assert np.array_equal(m.nodes, ref), float(np.max(np.abs(m.nodes - ref)))


@pytest.mark.parametrize("name", ["bump_default", "bump_stretched_turb", "p0_turb"])
def test_initial_state_matches_reference(name):
    d, m = mesh_of(name)
    prm = dgx.make_reference_state()
    ref = Z[name + "/state"]
    if d["turb"]:
        U = dgx.turbulent_state(m, prm, d["lx"])
        err = float(np.max(np.abs(U - ref)) / np.max(np.abs(ref)))
        print("turbulent state max rel err %.2e" % err)
        assert err < 1e-15
    else:
        U = np.tile(np.repeat(np.array(prm.Ufr), m.nbf), m.nelem)
        assert np.array_equal(U, ref)


def test_two_output_file_matches_reference_bytes(tmp_path):
    d, m = mesh_of("mean_run")
    path = str(tmp_path / "mine.dg")
    # given in the other order: the writer orders them as the reference's map
    dgx.write_dg_outputs(path, META["mean_run"]["tstep"], m,
                         {"mean": Z["mean_run/mean"], "state": Z["mean_run/state"]})
    mine = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
    assert np.array_equal(mine, Z["mean_run/file"])
