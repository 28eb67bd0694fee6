"""The allocation-free residual adjoint (dgx_residual_vjp_device_noalloc)
against dgx_residual_vjp on the same contexts: two calls in a row, the second
on a smaller mesh (so the fixed workspace holds the first call's data when
the second starts), each equal bit for bit to the allocating entry point;
meshes beyond the fixed workspace are rejected with DGX_ERR_INVALID."""
import ctypes as C

import numpy as np
import pytest

from dgpr_amd import cases, dgx

pytestmark = pytest.mark.gpu


def hip():
    lib = C.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return lib


def noalloc_vjp(h, s, w):
    n = w.size
    wd, gd = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(wd), 8 * n) == 0 and h.hipMalloc(C.byref(gd), 8 * n) == 0
    try:
        assert h.hipMemcpy(wd, w.ctypes.data_as(C.c_void_p), 8 * n, 1) == 0  # host to device
        s.residual_vjp_device_noalloc(wd.value, gd.value)
        g = np.zeros(n)
        assert h.hipMemcpy(g.ctypes.data_as(C.c_void_p), gd, 8 * n, 2) == 0  # device to host
    finally:
        h.hipFree(wd)
        h.hipFree(gd)
    return g


def test_noalloc_vjp_matches_allocating_form_twice():
    h = hip()
    big = cases.tgv_case(n=4, p=3, seed=71, M=0.2)
    small = cases.tgv_case(nxyz=(2, 3, 2), p=2, seed=72, M=0.25, amp=0.1)
    for c in (big, small):
        print("case", c.name, "nelem", c.mesh.nelem, "p", c.mesh.p, flush=True)
        w = np.random.default_rng(c.mesh.nelem).uniform(-1.0, 1.0, c.mesh.state_size)
        with dgx.Solver(c.mesh, c.params) as s:
            s.set_state(c.U0)
            g_ref = s.residual_vjp(w)
            g = noalloc_vjp(h, s, w)
        assert np.all(np.isfinite(g))
        print("  max |diff| %.3e (|g| max %.3e)" % (np.max(np.abs(g - g_ref)),
                                                    np.max(np.abs(g_ref))), flush=True)
        assert np.array_equal(g, g_ref)


def test_noalloc_vjp_rejects_meshes_beyond_its_workspace():
    h = hip()
    c = cases.tgv_case(n=16, p=3, seed=73)  # 4096 elements: 3 states + 2 record sets > 64 MiB
    with dgx.Solver(c.mesh, c.params) as s:
        s.set_state(c.U0)
        with pytest.raises(dgx.DgxError) as ei:
            noalloc_vjp(h, s, np.zeros(c.mesh.state_size))
        assert ei.value.status == dgx.ERR_INVALID
        assert "too large" in str(ei.value)
