"""natrium_amd: MI355X-native hot path of a semi-Lagrangian / SEDG lattice Boltzmann solver.

Python binding of the C ABI declared in include/natrium_amd.h (libnatrium_amd.so, built
for gfx950 by __graft_entry__.build()). There is no CPU fallback: every call goes to the
HIP library, and a missing library raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnatrium_amd.so")

STENCILS = {"D2Q9": 0, "D3Q19": 1, "D3Q27": 2, "D2Q25H": 3}
SCHEMES = {"SL": 0, "SEDG": 1, "SEDG_CENTRAL": 2}
PRECISIONS = {"fp64": 0, "fp32": 1}
BACKENDS = {"auto": 0, "structured": 1, "general": 2}
BACKEND_NAMES = {1: "structured", 2: "general"}
ERRORS = {1: "ARG", 2: "UNSUPPORTED", 3: "HIP", 4: "DENSITY", 5: "STATE", 6: "NOMEM"}


class NatError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"natrium_amd error {ERRORS.get(code, code)}: {msg}")
        self.code = code


class DensityZeroError(NatError):
    pass


class _Desc(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("stencil", ctypes.c_int), ("order", ctypes.c_int),
        ("ncells", ctypes.c_int * 3), ("vertices", ctypes.POINTER(ctypes.c_double)),
        ("cartesian", ctypes.c_int), ("periodic", ctypes.c_int * 3), ("scaling", ctypes.c_double),
        ("viscosity", ctypes.c_double), ("cfl", ctypes.c_double), ("scheme", ctypes.c_int),
        ("precision", ctypes.c_int), ("equilibrium", ctypes.c_int), ("compressible", ctypes.c_int),
        ("gamma", ctypes.c_double), ("rho0", ctypes.c_double), ("backend", ctypes.c_int),
    ]


EXPORTS = [
    "nat_create", "nat_destroy", "nat_last_error", "nat_n_dofs", "nat_n_stored_nodes", "nat_Q", "nat_dt",
    "nat_tau", "nat_time", "nat_backend", "nat_node_coords", "nat_set_dt", "nat_init_equilibrium",
    "nat_upload_f", "nat_download_f", "nat_upload_g", "nat_download_g", "nat_collide", "nat_stream",
    "nat_step", "nat_step_timed", "nat_get_macro", "nat_sums", "nat_sync",
]

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    """Load libnatrium_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.nat_create.restype = ctypes.c_int
    L.nat_create.argtypes = [ctypes.POINTER(_Desc), ctypes.POINTER(vp)]
    L.nat_destroy.argtypes = [vp]
    L.nat_last_error.restype = ctypes.c_char_p
    L.nat_last_error.argtypes = [vp]
    L.nat_n_dofs.restype = ctypes.c_int64
    L.nat_n_dofs.argtypes = [vp]
    L.nat_n_stored_nodes.restype = ctypes.c_int
    L.nat_n_stored_nodes.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    for nm in ("nat_dt", "nat_tau", "nat_time"):
        getattr(L, nm).restype = ctypes.c_double
        getattr(L, nm).argtypes = [vp]
    for nm in ("nat_Q", "nat_backend", "nat_collide", "nat_stream", "nat_sync"):
        getattr(L, nm).restype = ctypes.c_int
        getattr(L, nm).argtypes = [vp]
    L.nat_node_coords.restype = ctypes.c_int
    L.nat_node_coords.argtypes = [vp, _dp]
    L.nat_set_dt.restype = ctypes.c_int
    L.nat_set_dt.argtypes = [vp, ctypes.c_double]
    L.nat_init_equilibrium.restype = ctypes.c_int
    L.nat_init_equilibrium.argtypes = [vp, _dp, _dp, _dp]
    for nm in ("nat_upload_f", "nat_download_f", "nat_upload_g", "nat_download_g"):
        getattr(L, nm).restype = ctypes.c_int
        getattr(L, nm).argtypes = [vp, _dp]
    L.nat_step.restype = ctypes.c_int
    L.nat_step.argtypes = [vp, ctypes.c_int]
    L.nat_step_timed.restype = ctypes.c_int
    L.nat_step_timed.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.nat_get_macro.restype = ctypes.c_int
    L.nat_get_macro.argtypes = [vp, _dp, _dp, _dp]
    L.nat_sums.restype = ctypes.c_int
    L.nat_sums.argtypes = [vp, _dp, _dp, _dp]
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(_dp)


class Solver:
    """One solve on the device: CFDSolver-like interface over the C ABI.

    Loop semantics follow CFDSolver::run (solver/CFDSolver.cpp:877-902): call
    collide() once after init_equilibrium(), then step(n) runs n x (stream; collide).
    """

    def __init__(self, stencil, order, ncells, vertices, cartesian=True, periodic=None, scaling=1.0,
                 viscosity=1.0, cfl=0.4, scheme="SL", precision="fp64", backend="auto", rho0=1.0,
                 compressible=False, gamma=1.4, equilibrium="BGK"):
        L = lib()
        self.stencil = stencil
        self.D = 3 if stencil in ("D3Q19", "D3Q27") else 2
        self.p = order
        self._verts = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, self.D)
        d = _Desc()
        d.dim = self.D
        d.stencil = STENCILS[stencil]
        d.order = order
        for a in range(3):
            d.ncells[a] = int(ncells[a]) if a < len(ncells) else 1
            d.periodic[a] = 1 if periodic is None or a >= len(periodic) else int(periodic[a])
        d.vertices = _ptr(self._verts)
        d.cartesian = int(cartesian)
        d.scaling = scaling
        d.viscosity = viscosity
        d.cfl = cfl
        d.scheme = SCHEMES[scheme]
        d.precision = PRECISIONS[precision]
        d.equilibrium = 0 if equilibrium == "BGK" else 1
        d.compressible = int(compressible)
        d.gamma = gamma
        d.rho0 = rho0
        d.backend = BACKENDS[backend]
        h = ctypes.c_void_p()
        rc = L.nat_create(ctypes.byref(d), ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = L.nat_last_error(h).decode() if h else "allocation failed"
            L.nat_destroy(h)
            self._h = None
            raise NatError(rc, msg)
        self.n = L.nat_n_dofs(h)
        self.Q = L.nat_Q(h)
        self.compressible = compressible

    def close(self):
        if getattr(self, "_h", None):
            lib().nat_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            msg = lib().nat_last_error(self._h).decode()
            if rc == 4:
                raise DensityZeroError(rc, msg)
            raise NatError(rc, msg)

    @property
    def dt(self):
        return lib().nat_dt(self._h)

    @property
    def tau(self):
        return lib().nat_tau(self._h)

    @property
    def time(self):
        return lib().nat_time(self._h)

    @property
    def backend_name(self):
        return BACKEND_NAMES.get(lib().nat_backend(self._h), "?")

    @property
    def n_stored(self):
        v = ctypes.c_int64()
        lib().nat_n_stored_nodes(self._h, ctypes.byref(v))
        return v.value

    def coords(self):
        x = np.zeros((self.n, 3))
        self._check(lib().nat_node_coords(self._h, _ptr(x)))
        return x[:, :self.D].copy()

    def set_dt(self, dt):
        self._check(lib().nat_set_dt(self._h, dt))

    def init_equilibrium(self, rho, u, T=None):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.D, self.n)
        Tp = None
        if T is not None:
            T = np.ascontiguousarray(T, dtype=np.float64)
            Tp = _ptr(T)
        self._check(lib().nat_init_equilibrium(self._h, _ptr(rho), _ptr(u), Tp))

    def upload_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == (self.Q, self.n)
        self._check(lib().nat_upload_f(self._h, _ptr(f)))

    def download_f(self):
        f = np.zeros((self.Q, self.n))
        self._check(lib().nat_download_f(self._h, _ptr(f)))
        return f

    def upload_g(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        self._check(lib().nat_upload_g(self._h, _ptr(g)))

    def download_g(self):
        g = np.zeros((self.Q, self.n))
        self._check(lib().nat_download_g(self._h, _ptr(g)))
        return g

    def collide(self):
        self._check(lib().nat_collide(self._h))

    def stream(self):
        self._check(lib().nat_stream(self._h))

    def step(self, n=1):
        self._check(lib().nat_step(self._h, int(n)))

    def step_timed(self, n):
        """Run n fused steps; returns the device time in ms (HIP events on the library stream)."""
        ms = ctypes.c_double()
        self._check(lib().nat_step_timed(self._h, int(n), ctypes.byref(ms)))
        return ms.value

    def macro(self):
        rho = np.zeros(self.n)
        u = np.zeros((self.D, self.n))
        T = np.zeros(self.n)
        self._check(lib().nat_get_macro(self._h, _ptr(rho), _ptr(u), _ptr(T)))
        return rho, u, T

    def sums(self):
        m = np.zeros(1)
        mom = np.zeros(3)
        e = np.zeros(1)
        self._check(lib().nat_sums(self._h, _ptr(m), _ptr(mom), _ptr(e)))
        return float(m[0]), mom[:self.D].copy(), float(e[0])

    def sync(self):
        self._check(lib().nat_sync(self._h))
