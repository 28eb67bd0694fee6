"""libdgx: the C ABI loads and exports every entry point declared in
include/dgx.h; host-side helpers (reference state, generators, .dg files)
behave as the reference's counterparts.  No GPU needed."""
import ctypes as C
import math
import os
import re
import struct

import numpy as np
import pytest

from dgpr_amd import cases, dgx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    src = open(os.path.join(ROOT, "include", "dgx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dgx_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(dgx.LIB_PATH)
    names = declared_entry_points()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_make_reference_state():
    """make_reference_state (solver_params.cpp:40-70)."""
    p = dgx.make_reference_state(gamma=1.4, R=287.0, mu=2e-3, c=3.0, rho=1.3, u=0.2, v=-0.1,
                                 w=0.05, eta=6.0)
    P = 3.0 ** 2 * 1.3 / 1.4
    assert p.P == pytest.approx(P, rel=1e-15)
    assert p.T == pytest.approx(P / (1.3 * 287.0), rel=1e-15)
    assert list(p.Ufr) == pytest.approx([1.3, 0.26, -0.13, 0.065,
                                         P / 0.4 + 0.5 * 1.3 * (0.04 + 0.01 + 0.0025)], rel=1e-15)


def test_dg_file_layout_and_roundtrip(tmp_path):
    """.dg layout (io.cpp:33-89): u64 tstep; u64 nx,ny,nz,p,q; s64 bounds[6];
    u64 noutput; f64 nodes; then per output u64 keylen, key, f64 state."""
    c = cases.tgv_case(n=3, p=2, seed=3)
    path = str(tmp_path / "s.dg")
    dgx.write_dg(path, 17, c.mesh, c.U0)
    raw = open(path, "rb").read()
    ne = 27
    nn = ne * 8 * 3
    ns = ne * 5 * 27
    assert len(raw) == 8 * 6 + 8 * 6 + 8 + 8 * nn + 8 + 5 + 8 * ns
    hdr = struct.unpack("<6Q6qQ", raw[:104])
    assert hdr[:6] == (17, 3, 3, 3, 2, 1)
    assert hdr[6:12] == tuple([-2] * 6) and hdr[12] == 1
    assert raw[104 + 8 * nn: 104 + 8 * nn + 13] == struct.pack("<Q", 5) + b"state"
    ts, mesh, U = dgx.read_dg(path)
    assert ts == 17 and mesh.p == 2 and mesh.q == 1 and (mesh.nx, mesh.ny, mesh.nz) == (3, 3, 3)
    assert np.array_equal(mesh.nodes, c.mesh.nodes) and np.array_equal(U, c.U0)


def test_dg_read_rejects_truncated(tmp_path):
    c = cases.tgv_case(n=2, p=1, seed=3)
    path = str(tmp_path / "t.dg")
    dgx.write_dg(path, 0, c.mesh, c.U0)
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-100])
    with pytest.raises(dgx.DgxError):
        dgx.read_dg(path)


def test_bump_nodes_geometry():
    """generate_bump_nodes (geometry_common.cpp:494-576): uniform x/z
    spacing, top wall flat at ly, bottom following the Gaussian bump."""
    q, nx, ny, nz = 2, 6, 4, 2
    X = dgx.bump_nodes(q, nx, ny, nz, 2.0, 1.0, 1.0, 0.0, 0.1, 1.0, 0.05).reshape(
        nz, ny, nx, q + 1, q + 1, q + 1, 3)
    top = X[:, ny - 1, :, :, q, :, 1]
    assert np.allclose(top, 1.0, atol=1e-15)
    bot = X[:, 0, :, :, 0, :, :]
    bump = 0.1 * np.exp(-((bot[..., 0] - 1.0) ** 2) / (2 * 0.05))
    assert np.allclose(bot[..., 1], bump, atol=1e-15)
    xs = X[0, 0, :, 0, 0, :, 0]
    assert np.allclose(np.diff(np.concatenate([xs[:, :q].ravel(), [2.0]])), 2.0 / (nx * q))


def test_box_nodes_periodic_and_walls():
    q, nx, ny, nz = 1, 4, 3, 5
    lx, ly, lz = 4 * math.pi, 2.0, 4 * math.pi / 3
    X = dgx.box_nodes(q, nx, ny, nz, 0.0, -1.0, 0.0, lx, ly, lz, 0.1, 7).reshape(
        nz, ny, nx, 2, 2, 2, 3)
    # x-periodic faces coincide up to the period
    assert np.allclose(X[:, :, nx - 1, :, :, 1, :] - [lx, 0, 0], X[:, :, 0, :, :, 0, :], atol=1e-13)
    assert np.allclose(X[nz - 1, :, :, 1, :, :, :] - [0, 0, lz], X[0, :, :, 0, :, :, :], atol=1e-13)
    # walls stay planar
    assert np.allclose(X[:, 0, :, :, 0, :, 1], -1.0, atol=1e-15)
    assert np.allclose(X[:, ny - 1, :, :, 1, :, 1], 1.0, atol=1e-15)
    # shared corners of neighbouring elements are identical
    assert np.array_equal(X[:, :, 1:, :, :, 0, :], X[:, :, :-1, :, :, 1, :])


def test_tgv_state_formula():
    c = cases.tgv_case(n=2, p=1, seed=0, noise=0.0)
    U = c.U0.reshape(8, 5, 8)
    P0 = c.params.P
    # node 0 of element 0 sits at the origin: u = v = w = 0, p = P0 + 4/16
    assert U[0, 1, 0] == 0.0 and U[0, 2, 0] == 0.0 and U[0, 3, 0] == 0.0
    p = P0 + (1.0 / 16.0) * 2.0 * 3.0
    assert U[0, 0, 0] == pytest.approx(p / P0, rel=1e-15)
    assert U[0, 4, 0] == pytest.approx(p / 0.4, rel=1e-15)
    cn = cases.tgv_case(n=2, p=1, seed=0, noise=1e-3)
    du = (cn.U0.reshape(8, 5, 8)[:, 1:4] / cn.U0.reshape(8, 5, 8)[:, :1]
          - U[:, 1:4] / U[:, :1])
    assert 0 < np.max(np.abs(du)) <= 1e-3 * (1 + 1e-12)


def test_channel_state_profile_and_noise_level():
    c = cases.channel_case(4, 6, 3, p=2, seed=1, nsteps=1)
    m = c.mesh
    U = c.U0.reshape(m.nelem, 5, m.nbf)
    assert np.allclose(U[:, 0], 1.0)
    # remove the laminar part at each node and check the 5% RMS perturbation
    vel = U[:, 1:4] / U[:, :1]
    assert np.all(np.abs(vel) < 2.0)
    # CFL 0.3 on h_min = 2/6 with |u| + c = 1 + 5 and 2p + 1 = 5
    assert c.dt == pytest.approx(0.3 * (2.0 / 6) / ((1.0 + 5.0) * 5))


def test_create_without_device_fails_loudly():
    """On a host without a HIP device, creating a context is an error (no CPU
    fallback)."""
    try:
        hip = C.CDLL("libamdhip64.so.7")
        n = C.c_int(0)
        has_dev = hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        has_dev = False
    if has_dev:
        pytest.skip("a HIP device is present")
    c = cases.tgv_case(n=2, p=1)
    with pytest.raises(dgx.DgxError) as ei:
        dgx.Solver(c.mesh, c.params)
    assert ei.value.status == 6


def test_create_rejects_bad_input():
    c = cases.tgv_case(n=2, p=1)
    bad = dgx.Mesh(1, 1, 2, 2, 2, [dgx.PERIODIC, dgx.NO_SLIP_WALL, -2, -2, -2, -2], c.mesh.nodes)
    with pytest.raises(dgx.DgxError) as ei:
        dgx.Solver(bad, c.params)
    assert ei.value.status == 1
    # beyond the reference's MAX_ORDER = 8 (constants.cpp:21); checked before
    # any device is touched
    hi = dgx.Mesh(9, 1, 2, 2, 2, [-2] * 6, c.mesh.nodes)
    with pytest.raises(dgx.DgxError) as ei:
        dgx.Solver(hi, c.params)
    assert ei.value.status == 2


def test_entry_points_reject_null_arguments():
    """Every context-taking entry point returns DGX_ERR_INVALID for a null
    context (or null required buffer) instead of touching memory; the
    generators reject a null mesh."""
    L = dgx.lib()
    buf = (C.c_double * 8)()
    i = C.c_int()
    J = C.c_double()
    assert L.dgx_set_precision(None, dgx.PREC_MIXED) == dgx.ERR_INVALID
    assert L.dgx_get_precision(None, C.byref(i)) == dgx.ERR_INVALID
    assert L.dgx_set_tuning(None, 2) == dgx.ERR_INVALID
    assert L.dgx_step(None, dgx.RK4, 1e-3, 1, None) == dgx.ERR_INVALID
    assert L.dgx_kinetic_energy(None, C.byref(J)) == dgx.ERR_INVALID
    assert L.dgx_residual_vjp(None, buf, buf) == dgx.ERR_INVALID
    assert L.dgx_adjoint_rk4(None, 1e-3, 2, 0, C.byref(J), buf, None) == dgx.ERR_INVALID
    assert L.dgx_adjoint_release(None) == dgx.ERR_INVALID
    assert L.dgx_turbulent_state(None, None, 1.0, buf) == dgx.ERR_INVALID
    c = cases.tgv_case(n=2, p=1)
    cm = c.mesh.cstruct()
    prm = c.params.copy()
    assert L.dgx_turbulent_state(C.byref(cm), C.byref(prm), 0.0, buf) == dgx.ERR_INVALID
    assert L.dgx_turbulent_state(C.byref(cm), None, 1.0, buf) == dgx.ERR_INVALID


def test_affine_classifier_scales_with_element_size():
    """dgx_mesh_is_affine measures an element's twist against its own size:
    a box of small elements far from the origin (1e4, element size 1e-3) with
    a twist of 5e-8 of the element size (25 ulps of the coordinates, 5e-15 of
    their magnitude) is curved, the same box untwisted is affine, and q = 2
    is never affine."""
    n, h, x0 = 2, 1e-3, 1e4
    nodes = dgx.box_nodes(1, n, n, n, x0, x0, x0, n * h, n * h, n * h, 0.0, 0)
    straight = dgx.Mesh(2, 1, n, n, n, [dgx.PERIODIC] * 6, nodes)
    assert straight.affine
    twisted = nodes.reshape(-1, 8, 3).copy()
    twisted[:, 7, 0] += 5e-8 * h  # corner (1,1,1) moved: a trilinear term
    assert not dgx.Mesh(2, 1, n, n, n, [dgx.PERIODIC] * 6, twisted.reshape(-1)).affine
    q2 = dgx.box_nodes(2, n, n, n, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.0, 0)
    assert not dgx.Mesh(2, 2, n, n, n, [dgx.PERIODIC] * 6, q2).affine


def test_dg_file_with_several_outputs(tmp_path):
    """write_state of the map {"state", "mean"} (io.cpp:33-89): outputs in the
    order the reference's map visits them, djb2(key) % 256 bucket order
    (data_structures.cpp:537-673: "state" in bucket 6 before "mean" in 102),
    whatever order they are given in; read back by key and in file order;
    the single-output reader still finds "state"."""
    def bucket(k):
        h = 5381
        for ch in k.encode():
            h = (h * 33 + ch) & 0xFFFFFFFF
        return h % 256

    assert bucket("state") == 6 and bucket("mean") == 102
    c = cases.tgv_case(n=2, p=1, seed=5)
    M = np.random.default_rng(1).uniform(size=c.mesh.state_size)
    path = str(tmp_path / "two.dg")
    dgx.write_dg_outputs(path, 9, c.mesh, {"mean": M, "state": c.U0})
    raw = open(path, "rb").read()
    ne, nn, ns = 8, 8 * 8 * 3, 8 * 5 * 8
    assert struct.unpack("<Q", raw[96:104])[0] == 2
    o = 104 + 8 * nn
    assert raw[o:o + 13] == struct.pack("<Q", 5) + b"state"
    assert np.array_equal(np.frombuffer(raw[o + 13:o + 13 + 8 * ns], "<f8"), c.U0)
    o += 13 + 8 * ns
    assert raw[o:o + 12] == struct.pack("<Q", 4) + b"mean"
    assert np.array_equal(np.frombuffer(raw[o + 12:], "<f8"), M)
    ts, mesh, outs = dgx.read_dg_outputs(path)
    assert ts == 9 and list(outs) == ["state", "mean"] and ne == mesh.nelem
    assert np.array_equal(outs["mean"], M) and np.array_equal(outs["state"], c.U0)
    _, _, U = dgx.read_dg(path)
    assert np.array_equal(U, c.U0)
    with pytest.raises(dgx.DgxError):
        dgx.write_dg_outputs(path, 0, c.mesh, {"state": c.U0, "": M})
