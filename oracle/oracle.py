"""ctypes front end of the oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module.  It wraps ``liboracle.so`` (the C
restatement of the reference's CPU residual and RK drivers, built by
``oracle/build.py``) and exposes it on numpy arrays in the reference's own
layouts: state ``U[e][r][bf]``, nodes ``[e][(q+1)^3][3]``.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
# the same C source built with -DORC_EXTENDED: x87 extended precision
LIBX_PATH = os.path.join(HERE, "liboracle_x.so")

_lib = None
_libx = None

D = C.c_double
PD = C.POINTER(C.c_double)
DX = C.c_longdouble
PDX = C.POINTER(C.c_longdouble)
PI64 = C.POINTER(C.c_int64)
U64 = C.c_uint64


class Params(C.Structure):
    """Mirror of the reference's ``parameters`` (solver_params.cpp:22-36)."""
    _fields_ = [("gamma", D), ("R", D), ("mu", D), ("c", D), ("T", D),
                ("P", D), ("Ufr", D * 5), ("source", D * 5), ("eta", D)]


class ParamsX(C.Structure):
    """``parameters`` of the extended-precision build (long double fields)."""
    _fields_ = [("gamma", DX), ("R", DX), ("mu", DX), ("c", DX), ("T", DX),
                ("P", DX), ("Ufr", DX * 5), ("source", DX * 5), ("eta", DX)]


def _register(L, D, PD, P):
    L.orc_gauss_legendre.argtypes = [D, D, C.c_int, PD, PD]
    L.orc_gauss_legendre.restype = C.c_int
    L.orc_lagrange1d.argtypes = [C.c_int, C.c_int, D]
    L.orc_lagrange1d.restype = D
    L.orc_lagrange1d_deriv.argtypes = [C.c_int, C.c_int, D]
    L.orc_lagrange1d_deriv.restype = D
    L.orc_det3x3.argtypes = [PD]
    L.orc_det3x3.restype = D
    L.orc_inv3x3.argtypes = [PD, PD]
    L.orc_matmul.argtypes = [C.c_int, C.c_int, C.c_int, D, PD, PD, D, PD]
    L.orc_symmstore_mvmul.argtypes = [C.c_int, PD, PD, D, PD]
    L.orc_spmm.argtypes = [C.c_int, C.c_int, PD, PD, D, PD]
    L.orc_invert.argtypes = [PD, C.c_int]
    L.orc_invert.restype = C.c_int
    L.orc_analytical_flux.argtypes = [PD, D, PD, PD, PD]
    L.orc_roe.argtypes = [PD, PD, PD, D, PD, PD]
    L.orc_A.argtypes = [PD, PD, PD, PD, D, PD, PD, PD]
    L.orc_bound_state.argtypes = [C.c_int64, PD, PD, C.POINTER(P), D, D, PD]
    L.orc_bound_state.restype = C.c_int
    L.orc_geom_create.argtypes = [C.c_int, C.c_int, U64, U64, U64, PI64, PD]
    L.orc_geom_create.restype = C.c_void_p
    L.orc_geom_destroy.argtypes = [C.c_void_p]
    L.orc_geom_sizes.argtypes = [C.c_void_p, C.POINTER(U64)]
    for nm in ("vJ", "Minv", "vgrad", "fgrad", "n", "fJ", "ih", "bh", "vw",
               "fw", "veval", "feval"):
        fn = getattr(L, "orc_geom_" + nm)
        fn.argtypes = [C.c_void_p]
        fn.restype = PD
    for nm in ("ifl", "bfl"):
        fn = getattr(L, "orc_geom_" + nm)
        fn.argtypes = [C.c_void_p]
        fn.restype = PI64
    L.orc_residual.argtypes = [C.c_void_p, C.POINTER(P), PD, PD, PD]
    L.orc_residual.restype = C.c_int
    L.orc_step.argtypes = [C.c_void_p, C.POINTER(P), C.c_int, D, PD, PD]
    L.orc_step.restype = C.c_int
    L.orc_residual_norm.argtypes = [PD, U64]
    L.orc_residual_norm.restype = D
    L.orc_set_plant.argtypes = [C.c_int]
    L.orc_eval_output.argtypes = [C.c_int, U64, PD, C.c_int, C.c_int, D, PD]
    L.orc_update_mean.argtypes = [PD, PD, U64, U64, U64]
    L.orc_turbulent.argtypes = [C.c_void_p, C.POINTER(P), D, PD]
    L.RHS = C.CFUNCTYPE(None, PD, PD, C.c_void_p)
    L.orc_step_generic.argtypes = [C.c_int, D, U64, PD, L.RHS, C.c_void_p]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from oracle import build as _b
            _b.build()
        _lib = C.CDLL(LIB_PATH)
        _register(_lib, D, PD, Params)
    return _lib


def libx():
    """The extended-precision build (same algorithm, long double)."""
    global _libx
    if _libx is None:
        if not os.path.exists(LIBX_PATH):
            from oracle import build as _b
            _b.build()
        _libx = C.CDLL(LIBX_PATH)
        _register(_libx, DX, PDX, ParamsX)
    return _libx


def _p(a):
    return a.ctypes.data_as(PD)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def make_params(gamma=1.4, R=287.0, mu=1e-3, c=1.0, rho=1.0, u=0.1, v=0.0,
                w=0.0, eta=6.0, source=(0, 0, 0, 0, 0)):
    """Reference state from inputs (make_reference_state, solver_params.cpp:40-70)."""
    P = c * c * (rho / gamma)
    prm = Params()
    prm.gamma, prm.R, prm.mu, prm.c = gamma, R, mu, c
    prm.T = P / (rho * R)
    prm.P = P
    prm.eta = eta
    ufr = [rho, rho * u, rho * v, rho * w,
           (P / (gamma - 1.0)) + 0.5 * rho * (u * u + v * v + w * w)]
    for i in range(5):
        prm.Ufr[i] = ufr[i]
        prm.source[i] = float(source[i])
    return prm


# ---- pointwise functions ----------------------------------------------------

def gauss_legendre(n, a=0.0, b=1.0):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(a, b, n, _p(x), _p(w))
    return x, w


def lagrange1d(i, p, x):
    return lib().orc_lagrange1d(i, p, x)


def lagrange1d_deriv(i, p, x):
    return lib().orc_lagrange1d_deriv(i, p, x)


def det3x3(m):
    m = _f64(m).ravel()
    return lib().orc_det3x3(_p(m))


def inv3x3(m):
    m = _f64(m).ravel()
    out = np.zeros(9)
    lib().orc_inv3x3(_p(m), _p(out))
    return out


def matmul(A, B):
    A = _f64(A)
    B = _f64(B)
    n, m = A.shape
    p = B.shape[1]
    Cm = np.zeros((n, p))
    lib().orc_matmul(n, m, p, 1.0, _p(A), _p(B), 0.0, _p(Cm))
    return Cm


def symmstore_mvmul(n, packed, v, c=1.0):
    packed = _f64(packed)
    v = _f64(v)
    r = np.zeros(n)
    lib().orc_symmstore_mvmul(n, _p(packed), _p(v), c, _p(r))
    return r


def invert(M):
    M = _f64(M).copy()
    st = lib().orc_invert(_p(M), M.shape[0])
    return st, M


def analytical_flux(U, gamma):
    U = _f64(U)
    F = np.zeros((3, 5))
    lib().orc_analytical_flux(_p(U), gamma, _p(F[0]), _p(F[1]), _p(F[2]))
    return F


def roe(UL, UR, n, gamma):
    UL, UR, n = _f64(UL), _f64(UR), _f64(n)
    f = np.zeros(5)
    smax = np.zeros(1)
    lib().orc_roe(_p(UL), _p(UR), _p(n), gamma, _p(f), _p(smax))
    return f, smax[0]


def viscous_A(s, sx, sy, sz, mu):
    s, sx, sy, sz = _f64(s), _f64(sx), _f64(sy), _f64(sz)
    Q = np.zeros((3, 5))
    lib().orc_A(_p(s), _p(sx), _p(sy), _p(sz), mu, _p(Q[0]), _p(Q[1]), _p(Q[2]))
    return Q


def bound_state(bt, UL, n, prm, Pt=0.0, Tt=0.0):
    UL, n = _f64(UL), _f64(n)
    Ub = np.zeros(5)
    st = lib().orc_bound_state(int(bt), _p(UL), _p(n), C.byref(prm), Pt, Tt, _p(Ub))
    return st, Ub


# ---- geometry + residual ----------------------------------------------------

class Geometry:
    """Reference geometry with its full precompute (core_geometry +
    shared_geometry), built from element nodes.  ``extended=True`` uses the
    extended-precision build: inputs are taken as given (doubles), every
    operation runs in long double and results come back as np.longdouble."""

    def __init__(self, p, q, nx, ny, nz, bounds, nodes, extended=False):
        self.extended = bool(extended)
        self._L = libx() if self.extended else lib()
        self._dt = np.longdouble if self.extended else np.float64
        self._PD = PDX if self.extended else PD
        L = self._L
        self.p, self.q, self.nx, self.ny, self.nz = p, q, nx, ny, nz
        self.bounds = np.asarray(bounds, dtype=np.int64)
        self.nodes = self._in(nodes)
        self.h = L.orc_geom_create(p, q, nx, ny, nz,
                                   self.bounds.ctypes.data_as(PI64), self._ptr(self.nodes))
        if not self.h:
            raise ValueError("oracle geometry construction failed")
        s = (U64 * 8)()
        L.orc_geom_sizes(self.h, s)
        (self.nelem, self.nbf, self.nvq, self.nfq, self.nface, self.nifc,
         self.nbfc, self.nminv) = [int(v) for v in s]

    def _in(self, a):
        return np.ascontiguousarray(a, dtype=self._dt)

    def _ptr(self, a):
        return a.ctypes.data_as(self._PD)

    def _prm(self, prm):
        return params_from(prm, self.extended)

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_geom_destroy(self.h)
            self.h = None

    def _arr(self, name, count, dtype=None):
        ptr = getattr(self._L, "orc_geom_" + name)(self.h)
        if dtype is None:
            dtype = np.float64
        return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype).copy()

    @property
    def vJ(self):
        return self._arr("vJ", self.nelem * self.nvq).reshape(self.nelem, self.nvq)

    @property
    def Minv(self):
        return self._arr("Minv", self.nelem * self.nminv).reshape(self.nelem, self.nminv)

    @property
    def vgrad(self):
        return self._arr("vgrad", self.nelem * self.nvq * self.nbf * 3).reshape(
            self.nelem, self.nvq, self.nbf, 3)

    @property
    def fgrad(self):
        return self._arr("fgrad", self.nelem * 6 * self.nfq * self.nbf * 3).reshape(
            self.nelem, 6, self.nfq, self.nbf, 3)

    @property
    def normals(self):
        return self._arr("n", self.nface * self.nfq * 3).reshape(self.nface, self.nfq, 3)

    @property
    def fJ(self):
        return self._arr("fJ", self.nface * self.nfq).reshape(self.nface, self.nfq)

    @property
    def interior_h(self):
        return self._arr("ih", self.nifc * 2).reshape(self.nifc, 2)

    @property
    def boundary_h(self):
        return self._arr("bh", self.nbfc)

    @property
    def interior_face_list(self):
        return self._arr("ifl", self.nifc * 5, np.int64).reshape(self.nifc, 5)

    @property
    def boundary_face_list(self):
        return self._arr("bfl", self.nbfc * 4, np.int64).reshape(self.nbfc, 4)

    @property
    def vw(self):
        return self._arr("vw", self.nvq)

    @property
    def fw(self):
        return self._arr("fw", self.nfq)

    @property
    def veval(self):
        return self._arr("veval", self.nvq * self.nbf).reshape(self.nvq, self.nbf)

    @property
    def feval(self):
        return self._arr("feval", 6 * self.nfq * self.nbf).reshape(6, self.nfq, self.nbf)

    def minv_full(self, e):
        """Unpack element e's packed upper-triangular inverse mass matrix."""
        pk = self.Minv[e]
        n = self.nbf
        M = np.zeros((n, n))
        idx = 0
        for r in range(n):
            for c in range(r, n):
                M[r, c] = M[c, r] = pk[idx]
                idx += 1
        return M

    def state_size(self):
        return self.nelem * 5 * self.nbf

    def residual(self, prm, U):
        U = self._in(U)
        R = np.zeros(self.state_size(), dtype=self._dt)
        f = np.zeros(self.state_size(), dtype=self._dt)
        st = self._L.orc_residual(self.h, C.byref(self._prm(prm)), self._ptr(U),
                                  self._ptr(R), self._ptr(f))
        if st:
            raise RuntimeError("oracle residual failed (status %d)" % st)
        return R, f

    def turbulent(self, prm, l):
        """turbulent (solver_init.cpp:30-73) on this mesh; returns U."""
        U = np.zeros(self.state_size(), dtype=self._dt)
        self._L.orc_turbulent(self.h, C.byref(self._prm(prm)), l, self._ptr(U))
        return U

    def step(self, prm, U, dt, scheme="RK4", nsteps=1):
        """Advance a copy of U by nsteps; returns (U, per-step RMS(R) norms)."""
        U = self._in(U).copy()
        sc = 0 if scheme == "RK4" else 1
        norms = []
        nrm = np.zeros(1, dtype=self._dt)
        cp = self._prm(prm)
        for _ in range(nsteps):
            st = self._L.orc_step(self.h, C.byref(cp), sc, dt, self._ptr(U), self._ptr(nrm))
            if st:
                raise RuntimeError("oracle step failed (status %d)" % st)
            norms.append(nrm[0])
        return U, np.array(norms)


def step_generic(scheme, dt, U, rhs):
    """Apply one RK4 / TVD-RK3 step with a Python right-hand side."""
    L = lib()
    U = _f64(U).copy()
    n = U.size

    def _cb(u_ptr, f_ptr, _ctx):
        u = np.ctypeslib.as_array(u_ptr, shape=(n,))
        f = np.ctypeslib.as_array(f_ptr, shape=(n,))
        f[:] = rhs(u.copy())

    cb = L.RHS(_cb)
    L.orc_step_generic(0 if scheme == "RK4" else 1, dt, n, _p(U), cb, None)
    return U


class planted:
    """Context manager planting a deliberate mistake in the oracle (tests of
    the tests only): 1 = R-side dual-consistency sign, 2 = R side drops the
    Roe flux, 3 = RK4 stage 3 with dt/2."""

    def __init__(self, kind):
        self.kind = kind

    def __enter__(self):
        lib().orc_set_plant(self.kind)
        return self

    def __exit__(self, *a):
        lib().orc_set_plant(0)


def params_from(prm, extended=False):
    """oracle Params (ParamsX for the extended build) from any object with
    the reference ``parameters`` fields."""
    o = ParamsX() if extended else Params()
    for f in ("gamma", "R", "mu", "c", "T", "P", "eta"):
        setattr(o, f, getattr(prm, f))
    for i in range(5):
        o.Ufr[i] = prm.Ufr[i]
        o.source[i] = prm.source[i]
    return o


OUTPUTS = {"mach": 0, "rho": 1, "u": 2, "v": 3, "w": 4}


def eval_output(p, U, nelem, N, var, gamma):
    """eval_output (render/vrtxgen.cpp:25-59) on the (N+1)^3 equispaced
    sub-grid of every element; returns [nelem][(N+1)^3]."""
    U = _f64(U)
    out = np.zeros(nelem * (N + 1) ** 3)
    lib().orc_eval_output(p, nelem, _p(U), N, OUTPUTS.get(var, var), gamma, _p(out))
    return out.reshape(nelem, (N + 1) ** 3)


def update_mean(state, mean, time_step, mean_start):
    """update_mean (output.cpp:39-49), in place on a copy; returns it."""
    state = _f64(state)
    mean = _f64(mean).copy()
    lib().orc_update_mean(_p(state), _p(mean), state.size, time_step, mean_start)
    return mean
