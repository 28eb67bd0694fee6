"""Python binding of libdgx (the gfx950 DG residual / RK library) over its C ABI.

This mirrors the reference's host-side GPU API (solver/cudg.cu,
solve_cuda.cu, time_cuda.cu) on numpy arrays:

=====================================  =====================================
reference                              here
=====================================  =====================================
``parameters`` / make_reference_state ``Params`` / ``make_reference_state``
``core_geometry`` nodes + bounds       ``Mesh``
``bump_shared`` node generation         ``bump_nodes``
``cuda_device_geometry`` + workspace    ``Solver(mesh, params)``
shuffle_state + cudaMemcpy              ``Solver.set_state`` / ``get_state``
``cuda_residual``                       ``Solver.residual``
``RK4_cuda`` / ``tvdRK3_cuda``          ``Solver.step``
``write_state`` / ``read_state``        ``write_dg`` / ``read_dg``
=====================================  =====================================

The library must exist (built by ``dgpr_amd/build.py``); there is no CPU
fallback: every compute call goes to the HIP kernels and raises on failure.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdgx.so")

D = C.c_double
PD = C.POINTER(C.c_double)
U64 = C.c_uint64

# btype codes (boundaries.cpp:26-35)
INTERIOR, PERIODIC, FREESTREAM, INVISCID_WALL = -1, -2, -3, -4
SUBSONIC_INFLOW, SUBSONIC_OUTFLOW, NO_SLIP_WALL = -5, -6, -7
BTYPES = {"periodic": PERIODIC, "freestream": FREESTREAM, "inviscid_wall": INVISCID_WALL,
          "subsonic_inflow": SUBSONIC_INFLOW, "subsonic_outflow": SUBSONIC_OUTFLOW,
          "no_slip_wall": NO_SLIP_WALL}
# time_scheme (time.cpp:23-27)
RK4, TVDRK3 = 0, 1
SCHEMES = {"RK4": RK4, "tvdRK3": TVDRK3}
PREC_FP64, PREC_MIXED = 0, 1
PREC_MIXED_BF16, PREC_MIXED_RECORDS = 2, 3
PRECISIONS = {"fp64": PREC_FP64, "mixed": PREC_MIXED, "mixed_bf16": PREC_MIXED_BF16,
              "mixed_records": PREC_MIXED_RECORDS}

STATUS = {0: "ok", 1: "invalid argument", 2: "unsupported", 3: "HIP error",
          4: "NaN residual norm", 5: "I/O error", 6: "no HIP device"}
OK, ERR_INVALID, ERR_UNSUPPORTED, ERR_HIP, ERR_NAN, ERR_IO, ERR_NODEVICE = range(7)
# output_type (output.cpp:23-30)
OUTPUTS = {"mach": 0, "rho": 1, "u": 2, "v": 3, "w": 4}


class DgxError(RuntimeError):
    def __init__(self, status, msg=""):
        super().__init__("dgx status %d (%s)%s" % (status, STATUS.get(status, "?"),
                                                   (": " + msg) if msg else ""))
        self.status = status


class Params(C.Structure):
    """``parameters`` (solver_params.cpp:22-36)."""
    _fields_ = [("gamma", D), ("R", D), ("mu", D), ("c", D), ("T", D), ("P", D),
                ("Ufr", D * 5), ("source", D * 5), ("eta", D)]

    def copy(self):
        p = Params()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(Params))
        return p


class _CMesh(C.Structure):
    _fields_ = [("p", U64), ("q", U64), ("nx", U64), ("ny", U64), ("nz", U64),
                ("bounds", C.c_int64 * 6), ("nodes", PD)]


_lib = None


def lib():
    """Load libdgx.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DgxError(2, "libdgx.so not built: run python -m dgpr_amd.build")
        L = C.CDLL(LIB_PATH)
        st = C.c_int
        P_ctx = C.c_void_p
        L.dgx_create.argtypes = [C.POINTER(_CMesh), C.POINTER(Params), C.POINTER(P_ctx)]
        L.dgx_destroy.argtypes = [P_ctx]
        L.dgx_set_params.argtypes = [P_ctx, C.POINTER(Params)]
        L.dgx_set_state.argtypes = [P_ctx, PD]
        L.dgx_get_state.argtypes = [P_ctx, PD]
        L.dgx_set_state_device.argtypes = [P_ctx, C.c_void_p]
        L.dgx_get_state_device.argtypes = [P_ctx, C.c_void_p]
        L.dgx_residual.argtypes = [P_ctx, PD, PD]
        L.dgx_residual_device.argtypes = [P_ctx, C.c_void_p, C.c_void_p, C.c_void_p]
        L.dgx_step.argtypes = [P_ctx, C.c_int, D, U64, PD]
        L.dgx_step_async.argtypes = [P_ctx, C.c_int, D, U64]
        L.dgx_step_norms_async.argtypes = [P_ctx, C.c_int, D, U64, C.c_void_p]
        L.dgx_device_copy.argtypes = [P_ctx, C.c_void_p, C.c_void_p, U64]
        L.dgx_sync.argtypes = [P_ctx]
        L.dgx_get_stream.argtypes = [P_ctx, C.POINTER(C.c_void_p)]
        L.dgx_get_sizes.argtypes = [P_ctx, C.POINTER(U64)]
        L.dgx_last_error.argtypes = [P_ctx]
        L.dgx_last_error.restype = C.c_char_p
        L.dgx_make_reference_state.argtypes = [D, D, D, D, D, D, D, D, D, PD, C.POINTER(Params)]
        L.dgx_bump_nodes.argtypes = [U64, U64, U64, U64, D, D, D, D, D, D, D, PD]
        L.dgx_box_nodes.argtypes = [U64, U64, U64, U64, D, D, D, D, D, D, D, U64, PD]
        L.dgx_tgv_state.argtypes = [C.POINTER(_CMesh), D, D, D, D, D, U64, PD]
        L.dgx_channel_state.argtypes = [C.POINTER(_CMesh), D, D, D, D, D, C.c_int, D, D, U64, PD]
        L.dgx_turbulent_state.argtypes = [C.POINTER(_CMesh), C.POINTER(Params), D, PD]
        L.dgx_write_dg.argtypes = [C.c_char_p, U64, C.POINTER(_CMesh), PD]
        L.dgx_read_dg_header.argtypes = [C.c_char_p, C.POINTER(U64), C.POINTER(_CMesh)]
        L.dgx_read_dg.argtypes = [C.c_char_p, PD, PD]
        L.dgx_set_tuning.argtypes = [P_ctx, C.c_int]
        L.dgx_mesh_is_affine.argtypes = [C.POINTER(_CMesh), C.POINTER(C.c_int)]
        L.dgx_kinetic_energy.argtypes = [P_ctx, PD]
        L.dgx_residual_vjp.argtypes = [P_ctx, PD, PD]
        L.dgx_adjoint_rk4.argtypes = [P_ctx, D, U64, U64, PD, PD, PD]
        L.dgx_adjoint_release.argtypes = [P_ctx]
        L.dgx_adjoint_set_store_limit.argtypes = [P_ctx, U64]
        L.dgx_residual_vjp_device_noalloc.argtypes = [P_ctx, C.c_void_p, C.c_void_p]
        L.dgx_ensemble_create.argtypes = [C.c_int, C.POINTER(_CMesh), C.POINTER(Params),
                                          C.POINTER(C.c_void_p)]
        L.dgx_ensemble_destroy.argtypes = [C.c_void_p]
        L.dgx_ensemble_member.argtypes = [C.c_void_p, C.c_int, C.POINTER(P_ctx)]
        L.dgx_ensemble_step.argtypes = [C.c_void_p, C.c_int, PD, U64, C.c_int]
        L.dgx_ensemble_sync.argtypes = [C.c_void_p]
        L.dgx_ensemble_get_stream.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.dgx_ensemble_autotune.argtypes = [C.c_void_p, C.c_int, PD, C.c_int,
                                            C.POINTER(C.c_int)]
        L.dgx_ensemble_last_error.argtypes = [C.c_void_p]
        L.dgx_ensemble_set_layout.argtypes = [C.c_void_p, C.c_int]
        L.dgx_ensemble_member_tuning.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int),
                                                 C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.dgx_ensemble_last_error.restype = C.c_char_p
        L.dgx_sample_output.argtypes = [P_ctx, C.c_int, C.c_int, PD]
        L.dgx_sample_output_device.argtypes = [P_ctx, C.c_int, C.c_int, C.c_void_p]
        L.dgx_update_mean.argtypes = [P_ctx, U64, U64]
        L.dgx_get_mean.argtypes = [P_ctx, PD]
        L.dgx_set_mean.argtypes = [P_ctx, PD]
        L.dgx_write_dg_outputs.argtypes = [C.c_char_p, U64, C.POINTER(_CMesh), U64,
                                           C.POINTER(C.c_char_p), C.POINTER(PD)]
        L.dgx_read_dg_output_key.argtypes = [C.c_char_p, U64, C.c_char_p, C.POINTER(U64)]
        L.dgx_read_dg_output.argtypes = [C.c_char_p, C.c_char_p, PD]
        L.dgx_get_tuning.argtypes = [P_ctx, C.POINTER(C.c_int)]
        L.dgx_set_precision.argtypes = [P_ctx, C.c_int]
        L.dgx_get_precision.argtypes = [P_ctx, C.POINTER(C.c_int)]
        L.dgx_autotune.argtypes = [P_ctx, C.c_int, C.POINTER(C.c_int), PD]
        L.dgx_set_tile.argtypes = [P_ctx, C.c_int]
        L.dgx_get_tile.argtypes = [P_ctx, C.POINTER(C.c_int)]
        L.dgx_autotune_tiles.argtypes = [P_ctx, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), PD]
        for nm in ("dgx_create", "dgx_destroy", "dgx_set_params", "dgx_set_state",
                   "dgx_get_state", "dgx_set_state_device", "dgx_get_state_device",
                   "dgx_residual", "dgx_residual_device", "dgx_step", "dgx_step_async",
                   "dgx_step_norms_async", "dgx_device_copy",
                   "dgx_sync", "dgx_get_stream", "dgx_get_sizes", "dgx_make_reference_state",
                   "dgx_bump_nodes", "dgx_box_nodes", "dgx_tgv_state", "dgx_channel_state",
                   "dgx_turbulent_state",
                   "dgx_write_dg", "dgx_read_dg_header", "dgx_read_dg", "dgx_set_tuning",
                   "dgx_mesh_is_affine",
                   "dgx_get_tuning", "dgx_autotune", "dgx_set_precision", "dgx_get_precision",
                   "dgx_set_tile", "dgx_get_tile", "dgx_autotune_tiles",
                   "dgx_sample_output",
                   "dgx_sample_output_device", "dgx_update_mean", "dgx_get_mean",
                   "dgx_set_mean", "dgx_write_dg_outputs", "dgx_read_dg_output_key",
                   "dgx_read_dg_output",
                   "dgx_ensemble_create", "dgx_ensemble_destroy", "dgx_ensemble_member",
                   "dgx_ensemble_step", "dgx_ensemble_sync", "dgx_ensemble_get_stream",
                   "dgx_ensemble_autotune", "dgx_ensemble_set_layout",
                   "dgx_ensemble_member_tuning", "dgx_kinetic_energy", "dgx_residual_vjp",
                   "dgx_adjoint_rk4", "dgx_adjoint_release", "dgx_adjoint_set_store_limit",
                   "dgx_residual_vjp_device_noalloc"):
            getattr(L, nm).restype = st
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(PD)


def _chk(status, ctx=None):
    if status != 0:
        msg = lib().dgx_last_error(ctx)
        raise DgxError(status, msg.decode() if msg else "")


def make_reference_state(gamma=1.4, R=287.0, mu=1e-3, c=1.0, rho=1.0, u=0.1, v=0.0, w=0.0,
                         eta=6.0, source=None):
    """``make_reference_state`` (solver_params.cpp:40-70)."""
    prm = Params()
    src = None
    if source is not None:
        src = np.ascontiguousarray(source, dtype=np.float64)
    _chk(lib().dgx_make_reference_state(gamma, R, mu, c, rho, u, v, w, eta,
                                        _p(src) if src is not None else None, C.byref(prm)))
    return prm


class Mesh:
    """Persistent part of ``core_geometry``: orders, element counts, boundary
    codes (-x +x -y +y -z +z) and element nodes [e][(q+1)^3][3]."""

    def __init__(self, p, q, nx, ny, nz, bounds, nodes):
        self.p, self.q, self.nx, self.ny, self.nz = int(p), int(q), int(nx), int(ny), int(nz)
        self.bounds = np.asarray(bounds, dtype=np.int64).copy()
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1)
        nq = self.q + 1
        if self.nodes.size != self.nelem * nq ** 3 * 3:
            raise ValueError("nodes has %d entries, expected %d"
                             % (self.nodes.size, self.nelem * nq ** 3 * 3))

    @property
    def nelem(self):
        return self.nx * self.ny * self.nz

    @property
    def nbf(self):
        return (self.p + 1) ** 3

    @property
    def state_size(self):
        return self.nelem * 5 * self.nbf

    @property
    def affine(self):
        """Whether libdgx runs its constant-Jacobian kernels on this mesh
        (dgx_mesh_is_affine)."""
        cm = self.cstruct()
        a = C.c_int()
        _chk(lib().dgx_mesh_is_affine(C.byref(cm), C.byref(a)))
        return bool(a.value)

    def cstruct(self):
        m = _CMesh()
        m.p, m.q, m.nx, m.ny, m.nz = self.p, self.q, self.nx, self.ny, self.nz
        for i in range(6):
            m.bounds[i] = int(self.bounds[i])
        m.nodes = _p(self.nodes)
        return m


def bump_nodes(q, nx, ny, nz, lx, ly, lz, a=0.0, bmp_height=0.0, bmp_center=1.0,
               bmp_variance=0.05):
    """Nodes of ``bump_shared`` (geometry_common.cpp:494-576)."""
    out = np.zeros(nx * ny * nz * (q + 1) ** 3 * 3)
    _chk(lib().dgx_bump_nodes(q, nx, ny, nz, lx, ly, lz, a, bmp_height, bmp_center,
                              bmp_variance, _p(out)))
    return out


def box_nodes(q, nx, ny, nz, x0, y0, z0, lx, ly, lz, amp=0.0, seed=0):
    """Box nodes with an optional smooth, x/z-periodic displacement that
    vanishes on the y planes (relative amplitude ``amp`` of the element size)."""
    out = np.zeros(nx * ny * nz * (q + 1) ** 3 * 3)
    _chk(lib().dgx_box_nodes(q, nx, ny, nz, x0, y0, z0, lx, ly, lz, amp, seed, _p(out)))
    return out


def tgv_state(mesh, gamma=1.4, rho0=1.0, V0=1.0, P0=None, noise=1e-3, seed=0):
    """Taylor-Green-type state at the Lagrange nodes plus seeded velocity noise."""
    if P0 is None:
        raise ValueError("P0 required")
    out = np.zeros(mesh.state_size)
    cm = mesh.cstruct()
    _chk(lib().dgx_tgv_state(C.byref(cm), gamma, rho0, V0, P0, noise, seed, _p(out)))
    return out


def channel_state(mesh, gamma, rho0, Uc, P0, frac, kmax, lx, lz, seed):
    """Poiseuille profile plus divergence-free Fourier noise."""
    out = np.zeros(mesh.state_size)
    cm = mesh.cstruct()
    _chk(lib().dgx_channel_state(C.byref(cm), gamma, rho0, Uc, P0, frac, kmax, lx, lz, seed,
                                 _p(out)))
    return out


def turbulent_state(mesh, params, l):
    """turbulent (solver_init.cpp:30-73): the free stream of ``params`` plus
    four sine-product modes of wavelength l/2 .. l/16 on each velocity."""
    out = np.zeros(mesh.state_size)
    cm = mesh.cstruct()
    _chk(lib().dgx_turbulent_state(C.byref(cm), C.byref(params), l, _p(out)))
    return out


def write_dg(path, tstep, mesh, U):
    U = np.ascontiguousarray(U, dtype=np.float64)
    cm = mesh.cstruct()
    _chk(lib().dgx_write_dg(path.encode(), tstep, C.byref(cm), _p(U)))


def read_dg(path):
    """Returns (tstep, Mesh, U) of a .dg file's "state" output."""
    ts = U64()
    cm = _CMesh()
    _chk(lib().dgx_read_dg_header(path.encode(), C.byref(ts), C.byref(cm)))
    nq = cm.q + 1
    ne = cm.nx * cm.ny * cm.nz
    nodes = np.zeros(ne * nq ** 3 * 3)
    U = np.zeros(ne * 5 * (cm.p + 1) ** 3)
    _chk(lib().dgx_read_dg(path.encode(), _p(nodes), _p(U)))
    mesh = Mesh(cm.p, cm.q, cm.nx, cm.ny, cm.nz, list(cm.bounds), nodes)
    return int(ts.value), mesh, U


def write_dg_outputs(path, tstep, mesh, outputs):
    """write_state (io.cpp:33-89) of several named simstates, a dict
    {key: state}; written in the reference's map order."""
    keys = list(outputs)
    arrs = [np.ascontiguousarray(outputs[k], dtype=np.float64).reshape(-1) for k in keys]
    kk = (C.c_char_p * len(keys))(*[k.encode() for k in keys])
    pp = (PD * len(keys))(*[_p(a) for a in arrs])
    cm = mesh.cstruct()
    _chk(lib().dgx_write_dg_outputs(path.encode(), tstep, C.byref(cm), len(keys), kk, pp))


def read_dg_outputs(path):
    """read_state (io.cpp:91-159) of every output: (tstep, mesh, {key: state})
    with the keys in file order."""
    ts, mesh, _ = read_dg(path)
    nout = U64()
    buf = C.create_string_buffer(1025)
    _chk(lib().dgx_read_dg_output_key(path.encode(), 0, buf, C.byref(nout)))
    out = {}
    for i in range(int(nout.value)):
        _chk(lib().dgx_read_dg_output_key(path.encode(), i, buf, C.byref(nout)))
        key = buf.value.decode()
        a = np.zeros(mesh.state_size)
        _chk(lib().dgx_read_dg_output(path.encode(), key.encode(), _p(a)))
        out[key] = a
    return ts, mesh, out


class Solver:
    """Device context: geometry, state and stage buffers on the current HIP
    device (``cuda_device_geometry`` + ``cuda_residual_workspace``)."""

    def __init__(self, mesh, params):
        self.mesh = mesh
        self.params = params.copy()
        self._cm = mesh.cstruct()
        ctx = C.c_void_p()
        st = lib().dgx_create(C.byref(self._cm), C.byref(self.params), C.byref(ctx))
        _chk(st, None)
        self.ctx = ctx
        s = (U64 * 4)()
        _chk(lib().dgx_get_sizes(self.ctx, s), self.ctx)
        self.nelem, self.nbf, self.nstate, self.device_bytes = [int(v) for v in s]

    def close(self):
        if getattr(self, "ctx", None):
            lib().dgx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_params(self, params):
        self.params = params.copy()
        _chk(lib().dgx_set_params(self.ctx, C.byref(self.params)), self.ctx)

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1)
        if U.size != self.nstate:
            raise ValueError("state has %d entries, expected %d" % (U.size, self.nstate))
        _chk(lib().dgx_set_state(self.ctx, _p(U)), self.ctx)

    def get_state(self):
        U = np.zeros(self.nstate)
        _chk(lib().dgx_get_state(self.ctx, _p(U)), self.ctx)
        return U

    def set_state_device(self, ptr):
        """Load the state from a device buffer (simstate layout, nstate
        doubles) given as an integer address; returns after completion."""
        _chk(lib().dgx_set_state_device(self.ctx, C.c_void_p(int(ptr))), self.ctx)

    def get_state_device(self, ptr):
        """Store the state into a device buffer given as an integer address."""
        _chk(lib().dgx_get_state_device(self.ctx, C.c_void_p(int(ptr))), self.ctx)

    def residual_device(self, U_ptr, R_ptr, f_ptr):
        """dgx_residual_device on device buffers given as integer addresses
        (R_ptr may be 0); returns after completion."""
        _chk(lib().dgx_residual_device(self.ctx, C.c_void_p(int(U_ptr)),
                                       C.c_void_p(int(R_ptr)) if R_ptr else None,
                                       C.c_void_p(int(f_ptr))), self.ctx)

    def residual(self):
        """(R, f) of the current state in simstate layout (solve.cpp:150-807)."""
        R = np.zeros(self.nstate)
        f = np.zeros(self.nstate)
        _chk(lib().dgx_residual(self.ctx, _p(R), _p(f)), self.ctx)
        return R, f

    def step(self, dt, scheme="RK4", nsteps=1, norms=True):
        """Advance nsteps; returns per-step RMS(R) norms (or None)."""
        sc = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
        if norms:
            out = np.zeros(nsteps)
            _chk(lib().dgx_step(self.ctx, sc, dt, nsteps, _p(out)), self.ctx)
            return out
        _chk(lib().dgx_step(self.ctx, sc, dt, nsteps, None), self.ctx)
        return None

    def step_async(self, dt, scheme="RK4", nsteps=1):
        sc = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
        _chk(lib().dgx_step_async(self.ctx, sc, dt, nsteps), self.ctx)

    def step_norms_async(self, dt, scheme, nsteps, norms_ptr):
        """Enqueue nsteps steps; step s writes its RMS(R) to the device
        double at norms_ptr + 8 s (an integer address)."""
        sc = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
        _chk(lib().dgx_step_norms_async(self.ctx, sc, dt, nsteps, C.c_void_p(int(norms_ptr))),
             self.ctx)

    def sync(self):
        _chk(lib().dgx_sync(self.ctx), self.ctx)

    def stream(self):
        s = C.c_void_p()
        _chk(lib().dgx_get_stream(self.ctx, C.byref(s)), self.ctx)
        return s.value

    @property
    def stage_waves(self):
        """Occupancy target (waves per SIMD) of the stage kernel."""
        w = C.c_int()
        _chk(lib().dgx_get_tuning(self.ctx, C.byref(w)), self.ctx)
        return w.value

    @stage_waves.setter
    def stage_waves(self, w):
        _chk(lib().dgx_set_tuning(self.ctx, int(w)), self.ctx)

    def kinetic_energy(self):
        """Volume integral of |m|^2 / (2 rho) of the current state."""
        J = C.c_double()
        _chk(lib().dgx_kinetic_energy(self.ctx, C.byref(J)), self.ctx)
        return J.value

    def residual_vjp(self, w):
        """(df/dU)^T w at the current state (periodic meshes)."""
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
        g = np.zeros(self.nstate)
        _chk(lib().dgx_residual_vjp(self.ctx, _p(w), _p(g)), self.ctx)
        return g

    def residual_vjp_device_noalloc(self, w_ptr, g_ptr):
        """dgx_residual_vjp_device_noalloc on device buffers given as integer
        addresses: (df/dU)^T w into g with the library's fixed workspace."""
        _chk(lib().dgx_residual_vjp_device_noalloc(self.ctx, C.c_void_p(int(w_ptr)),
                                                   C.c_void_p(int(g_ptr))), self.ctx)

    def adjoint_rk4(self, dt, nsteps, interval=0, want_mu=True):
        """Final kinetic energy after nsteps RK4 steps and its gradients with
        respect to the initial state and to mu (discrete adjoint)."""
        J = C.c_double()
        g = np.zeros(self.nstate)
        dmu = C.c_double()
        _chk(lib().dgx_adjoint_rk4(self.ctx, dt, int(nsteps), int(interval), C.byref(J), _p(g),
                                   C.byref(dmu) if want_mu else None), self.ctx)
        return J.value, g, (dmu.value if want_mu else None)

    def adjoint_release(self):
        """Free the stored-stage buffer adjoint_rk4 keeps between calls."""
        _chk(lib().dgx_adjoint_release(self.ctx), self.ctx)

    def adjoint_store_limit(self, nbytes):
        """Caps the stored-stage buffer of adjoint_rk4 at nbytes (0: no cap);
        a schedule whose store would exceed it falls to the next one."""
        _chk(lib().dgx_adjoint_set_store_limit(self.ctx, int(nbytes)), self.ctx)

    def sample_output(self, var, N):
        """eval_output (render/vrtxgen.cpp:25-59) of the current state on the
        (N+1)^3 equispaced sub-grid of every element: [nelem][(N+1)^3]."""
        v = OUTPUTS[var] if isinstance(var, str) else int(var)
        out = np.zeros(self.mesh.nelem * (N + 1) ** 3)
        _chk(lib().dgx_sample_output(self.ctx, v, int(N), _p(out)), self.ctx)
        return out.reshape(self.mesh.nelem, (N + 1) ** 3)

    def update_mean(self, time_step, mean_start):
        """update_mean (output.cpp:39-49) on the current state."""
        _chk(lib().dgx_update_mean(self.ctx, int(time_step), int(mean_start)), self.ctx)

    def get_mean(self):
        M = np.zeros(self.nstate)
        _chk(lib().dgx_get_mean(self.ctx, _p(M)), self.ctx)
        return M

    def set_mean(self, M):
        M = np.ascontiguousarray(M, dtype=np.float64).reshape(-1)
        _chk(lib().dgx_set_mean(self.ctx, _p(M)), self.ctx)

    @property
    def precision(self):
        """'fp64' or 'mixed' (FP32 face trace records, the rest FP64)."""
        v = C.c_int()
        _chk(lib().dgx_get_precision(self.ctx, C.byref(v)), self.ctx)
        return {b: a for a, b in PRECISIONS.items()}[v.value]

    @precision.setter
    def precision(self, name):
        _chk(lib().dgx_set_precision(self.ctx, PRECISIONS[name]), self.ctx)

    # what the mixed forms compute below FP64 (the rest stays FP64)
    MIXED_FORMS = {
        "mixed": "FP32 face trace records; p=3 tiled kernel: volume gradient and flux "
                 "projection in FP32 on the matrix cores (v_mfma_f32_4x4x1f32); geometry, "
                 "Roe, pressure/energy, viscous physics, inverse mass, RK in FP64",
        "mixed_bf16": "as mixed, the contractions as 3-term split-BF16 products "
                      "(v_mfma_f32_4x4x4bf16_1k, FP32 accumulation)",
        "mixed_records": "FP32 face trace records only; contractions, geometry, fluxes, "
                         "Roe, inverse mass, RK in FP64",
    }

    # rows per tile of the tiled stage kernel at p = 1, 2, 3
    TILE_ROWS = {1: 1, 2: 3, 3: 2}

    def stage_kernel_name(self, waves=None, tile=None):
        """Name (template arguments P, Q[, rows], W) of the stage kernel this
        context launches."""
        w = self.stage_waves if waves is None else waves
        t = self.tile if tile is None else tile
        if t == 2:
            return "dgx_gen_stage_kernel<%d,double>" % self.mesh.p
        if t:
            return "dgx_tile_stage_kernel<%d,%d,%d,%d>" % (self.mesh.p, self.mesh.q,
                                                           self.TILE_ROWS[self.mesh.p], w)
        return "dgx_rhs_stage_kernel<%d,%d,%d>" % (self.mesh.p, self.mesh.q, w)

    @property
    def tile(self):
        """Stage tile: 0 (one row of elements per wavefront, mixed-direction
        face passes), 1 (tiled, one direction per face subpass) or 2 (the
        general-order kernel, one Gauss node per lane: the only one
        outside 1 <= p, q <= 3, and available at every order)."""
        t = C.c_int()
        _chk(lib().dgx_get_tile(self.ctx, C.byref(t)), self.ctx)
        return t.value

    @tile.setter
    def tile(self, t):
        _chk(lib().dgx_set_tile(self.ctx, int(t)), self.ctx)

    def autotune_tiles(self, reps=3):
        """Times every (tile, waves) pair, keeps the fastest; returns
        (tile, waves, us[2][3]) with 0 where a pair is not compiled."""
        t, w = C.c_int(), C.c_int()
        us = np.zeros(6)
        _chk(lib().dgx_autotune_tiles(self.ctx, reps, C.byref(t), C.byref(w), _p(us)), self.ctx)
        return t.value, w.value, us.reshape(2, 3)

    def autotune(self, reps=5):
        """Times each compiled occupancy target on this mesh, keeps the
        fastest; returns (chosen, [us at 1, 2, 3 waves/SIMD])."""
        w = C.c_int()
        us = np.zeros(3)
        _chk(lib().dgx_autotune(self.ctx, int(reps), C.byref(w), _p(us)), self.ctx)
        return w.value, us


class _Member(Solver):
    """A Solver view of an ensemble member's context (owned by the ensemble)."""

    def __init__(self, ctx, mesh, params):  # noqa: D401 -- no dgx_create here
        self.ctx = ctx
        self.mesh = mesh
        self.params = params.copy()
        self.nelem, self.nbf = mesh.nelem, mesh.nbf
        self.nstate = mesh.nelem * 5 * mesh.nbf

    def close(self):
        self.ctx = None


class Ensemble:
    """Independent cases advanced together (dgx_ensemble_*): members with the
    same (p, q) share one batched launch per RK stage; ``step(graph=True)``
    replays one captured step of the whole ensemble."""

    def __init__(self, meshes, params):
        n = len(meshes)
        cm = (_CMesh * n)(*[m.cstruct() for m in meshes])
        cp = (Params * n)(*[p.copy() for p in params])
        h = C.c_void_p()
        st = lib().dgx_ensemble_create(n, cm, cp, C.byref(h))
        if st != 0:
            raise DgxError(st, lib().dgx_ensemble_last_error(None).decode())
        self.h = h
        self.meshes = list(meshes)
        self.members = []
        for i, (m, p) in enumerate(zip(meshes, params)):
            c = C.c_void_p()
            self._chk(lib().dgx_ensemble_member(self.h, i, C.byref(c)))
            self.members.append(_Member(c, m, p))

    def _chk(self, st):
        if st != 0:
            raise DgxError(st, lib().dgx_ensemble_last_error(self.h).decode())

    def step(self, scheme, dt, nsteps, graph=True, sync=True):
        sc = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
        dts = np.ascontiguousarray(dt, dtype=np.float64)
        self._chk(lib().dgx_ensemble_step(self.h, sc, _p(dts), int(nsteps), int(bool(graph))))
        if sync:
            self.sync()

    def sync(self):
        self._chk(lib().dgx_ensemble_sync(self.h))

    def stream(self):
        s = C.c_void_p()
        self._chk(lib().dgx_ensemble_get_stream(self.h, C.byref(s)))
        return s.value

    def autotune(self, scheme, dt, reps=3):
        sc = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
        dts = np.ascontiguousarray(dt, dtype=np.float64)
        w = (C.c_int * len(self.members))()
        self._chk(lib().dgx_ensemble_autotune(self.h, sc, _p(dts), int(reps), w))
        return list(w)

    def set_layout(self, layout):
        """Layout of the q = 1 members' launches: 1 one launch per stage for
        all orders (default); 2 one launch per order; 0 one batched launch
        per (p, q, affine) group."""
        self._chk(lib().dgx_ensemble_set_layout(self.h, int(layout)))

    def member_tuning(self, i):
        """(layout, tile, waves) of the launch member i runs in (layout 0:
        a batched launch)."""
        p, t, w = C.c_int(), C.c_int(), C.c_int()
        self._chk(lib().dgx_ensemble_member_tuning(self.h, int(i), C.byref(p), C.byref(t),
                                                   C.byref(w)))
        return p.value, t.value, w.value

    def close(self):
        if getattr(self, "h", None):
            for m in self.members:
                m.close()
            lib().dgx_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
