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
        ("prandtl", ctypes.c_double), ("sutherland", ctypes.c_int),
        ("wall_velocity", (ctypes.c_double * 3) * 6),
    ]


class _Stats(ctypes.Structure):
    _fields_ = [("iteration", ctypes.c_int64)] + [(nm, ctypes.c_double) for nm in (
        "time", "max_u", "mass", "kin_e", "enstrophy", "entropy", "max_p", "min_p", "res_rho", "res_u", "mean_ux",
        "enstrophy_sq")]


EXPORTS = [
    "nat_create", "nat_destroy", "nat_last_error", "nat_n_dofs", "nat_n_stored_nodes", "nat_Q", "nat_dt",
    "nat_tau", "nat_time", "nat_backend", "nat_node_coords", "nat_set_dt", "nat_set_vertices", "nat_set_stabilizer",
    "nat_init_equilibrium",
    "nat_upload_f", "nat_download_f", "nat_upload_g", "nat_download_g", "nat_collide", "nat_collide_init",
    "nat_stream", "nat_stabilize",
    "nat_step", "nat_sl_assemble", "nat_sl_bgk_step", "nat_sedg_rk5_step", "nat_compressible_step",
    "nat_step_timed", "nat_get_macro", "nat_get_knudsen", "nat_sums", "nat_sums_noalloc", "nat_sync",
    "nat_stats", "nat_write_stats", "nat_write_vtu", "nat_checkpoint_write", "nat_checkpoint_read",
    "nat_struct_tile", "nat_pack_key", "nat_step_packed", "nat_autotune", "nat_set_general_form",
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
    for nm in ("nat_Q", "nat_backend", "nat_collide", "nat_stream", "nat_stabilize", "nat_sync"):
        getattr(L, nm).restype = ctypes.c_int
        getattr(L, nm).argtypes = [vp]
    L.nat_collide_init.restype = ctypes.c_int
    L.nat_collide_init.argtypes = [vp, _dp]
    L.nat_node_coords.restype = ctypes.c_int
    L.nat_node_coords.argtypes = [vp, _dp]
    L.nat_set_dt.restype = ctypes.c_int
    L.nat_set_dt.argtypes = [vp, ctypes.c_double]
    L.nat_stats.restype = ctypes.c_int
    L.nat_stats.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(_Stats)]
    L.nat_write_stats.restype = ctypes.c_int
    L.nat_write_stats.argtypes = [vp, ctypes.c_char_p, ctypes.c_int64, ctypes.c_int]
    L.nat_write_vtu.restype = ctypes.c_int
    L.nat_write_vtu.argtypes = [vp, ctypes.c_char_p, ctypes.c_int64]
    L.nat_checkpoint_write.restype = ctypes.c_int
    L.nat_checkpoint_write.argtypes = [vp, ctypes.c_char_p, ctypes.c_int64]
    L.nat_checkpoint_read.restype = ctypes.c_int
    L.nat_checkpoint_read.argtypes = [vp, ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    L.nat_set_vertices.restype = ctypes.c_int
    L.nat_set_vertices.argtypes = [vp, _dp]
    L.nat_set_stabilizer.restype = ctypes.c_int
    L.nat_set_stabilizer.argtypes = [vp, ctypes.c_int]
    L.nat_init_equilibrium.restype = ctypes.c_int
    L.nat_init_equilibrium.argtypes = [vp, _dp, _dp, _dp]
    for nm in ("nat_upload_f", "nat_download_f", "nat_upload_g", "nat_download_g"):
        getattr(L, nm).restype = ctypes.c_int
        getattr(L, nm).argtypes = [vp, _dp]
    L.nat_step.restype = ctypes.c_int
    L.nat_step.argtypes = [vp, ctypes.c_int]
    L.nat_sl_assemble.restype = ctypes.c_int
    L.nat_sl_assemble.argtypes = [vp, ctypes.c_double]
    for name in ("nat_sl_bgk_step", "nat_sedg_rk5_step", "nat_compressible_step"):
        getattr(L, name).restype = ctypes.c_int
        getattr(L, name).argtypes = [vp, ctypes.c_int]
    L.nat_step_timed.restype = ctypes.c_int
    L.nat_step_timed.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.nat_get_macro.restype = ctypes.c_int
    L.nat_get_macro.argtypes = [vp, _dp, _dp, _dp]
    L.nat_get_knudsen.restype = ctypes.c_int
    L.nat_get_knudsen.argtypes = [vp, _dp]
    for nm in ("nat_sums", "nat_sums_noalloc"):
        getattr(L, nm).restype = ctypes.c_int
        getattr(L, nm).argtypes = [vp, _dp, _dp, _dp]
    L.nat_struct_tile.restype = ctypes.c_int
    L.nat_struct_tile.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.nat_pack_key.restype = ctypes.c_int
    L.nat_pack_key.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.nat_step_packed.restype = ctypes.c_int
    L.nat_step_packed.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int]
    L.nat_autotune.restype = ctypes.c_int
    L.nat_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.nat_set_general_form.restype = ctypes.c_int
    L.nat_set_general_form.argtypes = [vp, ctypes.c_int]
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
                 compressible=False, gamma=1.4, equilibrium=None, prandtl=0.0, sutherland=False,
                 wall_velocity=None, stabilizer=False):
        """periodic: flag per axis (default: all periodic). Both faces of a non-periodic axis
        are velocity-neq bounce-back walls (SL only); wall_velocity maps a box face
        2*axis+side (side 0 = lower) to its constant velocity (default: at rest)."""
        L = lib()
        self.stencil = stencil
        self.D = 3 if stencil in ("D3Q19", "D3Q27") else 2
        if equilibrium is None:
            equilibrium = "QUARTIC" if compressible else "BGK"
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
        d.prandtl = prandtl
        d.sutherland = int(sutherland)
        for f, v in (wall_velocity or {}).items():
            for k, x in enumerate(v):
                d.wall_velocity[int(f)][k] = float(x)
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
        if stabilizer:
            self.set_stabilizer(True)

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

    def set_vertices(self, vertices):
        """Move the mesh (same lattice) and reassemble the operator on the device."""
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, self.D)
        if v.shape != self._verts.shape:
            raise ValueError("vertex array of a different lattice")
        self._check(lib().nat_set_vertices(self._h, _ptr(v)))
        self._verts = v

    def set_stabilizer(self, on=True):
        """Pseudo-entropic stabilizer after every step's collision (D2Q9 / D3Q19)."""
        self._check(lib().nat_set_stabilizer(self._h, int(on)))

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

    def collide_init(self, u=None):
        """Collision with the velocity read from the stored field (the reference's init == true);
        u (dim x n, physical units) replaces that field first."""
        up = None
        if u is not None:
            u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.D, self.n)
            up = _ptr(u)
        self._check(lib().nat_collide_init(self._h, up))

    def stream(self):
        """CFDSolver::stream alone (SL operator with wall values, or one RK5 step of SEDG)."""
        self._check(lib().nat_stream(self._h))

    def stabilize(self):
        """The pseudo-entropic stabilizer as a separate data-processor pass (D2Q9 / D3Q19)."""
        self._check(lib().nat_stabilize(self._h))

    def step(self, n=1):
        self._check(lib().nat_step(self._h, int(n)))

    def step_timed(self, n):
        """Run n fused steps; returns the device time in ms (HIP events on the library stream)."""
        ms = ctypes.c_double()
        self._check(lib().nat_step_timed(self._h, int(n), ctypes.byref(ms)))
        return ms.value

    def macro(self):
        """rho, u (physical units) and T (None for incompressible contexts) of the last collision."""
        rho = np.zeros(self.n)
        u = np.zeros((self.D, self.n))
        T = np.zeros(self.n) if self.compressible else None
        self._check(lib().nat_get_macro(self._h, _ptr(rho), _ptr(u), _ptr(T) if T is not None else None))
        return rho, u, T

    def knudsen(self):
        """Knudsen-number estimate (shock sensor) of the last compressible collision."""
        kn = np.zeros(self.n)
        self._check(lib().nat_get_knudsen(self._h, _ptr(kn)))
        return kn

    def sums(self, noalloc=False):
        """(mass, momentum, E_kin) of the last collision's fields. noalloc: through
        nat_sums_noalloc (fixed per-GPU scratch, no device allocation)."""
        m = np.zeros(1)
        mom = np.zeros(3)
        e = np.zeros(1)
        fn = lib().nat_sums_noalloc if noalloc else lib().nat_sums
        self._check(fn(self._h, _ptr(m), _ptr(mom), _ptr(e)))
        return float(m[0]), mom[:self.D].copy(), float(e[0])

    def stats(self, iteration):
        """One results-table row as a dict (column names of the reference's table)."""
        st = _Stats()
        self._check(lib().nat_stats(self._h, int(iteration), ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _Stats._fields_}

    def write_stats(self, path, iteration, truncate=False):
        self._check(lib().nat_write_stats(self._h, os.fsencode(path), int(iteration), int(truncate)))

    def write_vtu(self, directory, iteration):
        self._check(lib().nat_write_vtu(self._h, os.fsencode(directory), int(iteration)))

    def checkpoint_write(self, directory, iteration):
        self._check(lib().nat_checkpoint_write(self._h, os.fsencode(directory), int(iteration)))

    def checkpoint_read(self, directory, iteration=1):
        """Load checkpoint `iteration` (1: the most recent); returns the stored iteration."""
        it = ctypes.c_int64(0)
        self._check(lib().nat_checkpoint_read(self._h, os.fsencode(directory), int(iteration), ctypes.byref(it)))
        return it.value

    def sync(self):
        self._check(lib().nat_sync(self._h))

    def struct_tile(self):
        """Node tile (x, y, z) of one workgroup of the structured step kernel on this mesh."""
        t = (ctypes.c_int * 3)()
        self._check(lib().nat_struct_tile(self._h, t))
        return tuple(t)

    def pack_key(self):
        """Packing key (nat_pack_key): solves with equal keys can share packed launches."""
        k = ctypes.c_int64(0)
        self._check(lib().nat_pack_key(self._h, ctypes.byref(k)))
        return k.value

    def autotune(self, nsteps=10):
        """Time the SL kernel's launch shapes on this mesh and keep the fastest (nat_autotune);
        returns the chosen tile (structured backend) or (DoFs per block, waves per SIMD,
        z-planes in flight) of the chosen form (general backend)."""
        t = (ctypes.c_int * 3)()
        self._check(lib().nat_autotune(self._h, int(nsteps), t))
        return tuple(t)

    def set_general_form(self, form):
        """Pin the general backend's kernel form (nat_set_general_form): -1 default rule,
        0 six-wave form, 1 DEEP form (FP32 storage, 3D)."""
        self._check(lib().nat_set_general_form(self._h, int(form)))


def step_packed(solvers, nsteps):
    """nsteps fused steps of each solver as packed launches (nat_step_packed): one launch per
    step advances all of them; each result is bit-identical to its own step(nsteps)."""
    solvers = list(solvers)
    hs = (ctypes.c_void_p * len(solvers))(*[s._h for s in solvers])
    rc = lib().nat_step_packed(hs, len(solvers), int(nsteps))
    if rc:
        msg = lib().nat_last_error(solvers[0]._h).decode() if solvers else ""
        if rc == 4:
            raise DensityZeroError(rc, msg)
        raise NatError(rc, msg)


# ----------------------------------------------------------------------------- request server
PROBLEMS = {"TGV2D": 0, "TGV3D": 1, "SHOCK_TUBE": 2}


class _Request(ctypes.Structure):
    _fields_ = [
        ("problem", ctypes.c_int), ("stencil", ctypes.c_int), ("refinement", ctypes.c_int), ("order", ctypes.c_int),
        ("reynolds", ctypes.c_double), ("mach", ctypes.c_double), ("cfl", ctypes.c_double),
        ("nsteps", ctypes.c_int), ("precision", ctypes.c_int),
    ]


class _Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int), ("n_dofs", ctypes.c_int64), ("queue_ms", ctypes.c_double),
        ("run_ms", ctypes.c_double), ("latency_ms", ctypes.c_double), ("mass", ctypes.c_double),
        ("ekin", ctypes.c_double), ("footprint_bytes", ctypes.c_int64), ("error", ctypes.c_char * 256),
    ]


class _ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_active", ctypes.c_int), ("chunk_steps", ctypes.c_int),
                ("use_graphs", ctypes.c_int), ("pack", ctypes.c_int), ("autotune", ctypes.c_int),
                ("setup_threads", ctypes.c_int)]


class _ServerStats(ctypes.Structure):
    _fields_ = [("queued", ctypes.c_int64), ("active", ctypes.c_int64), ("used_bytes", ctypes.c_int64),
                ("node_updates", ctypes.c_int64), ("busy_ms", ctypes.c_double), ("packed_chunks", ctypes.c_int64)]


SERVER_EXPORTS = ["nat_server_create", "nat_server_submit", "nat_server_wait", "nat_server_stats",
                  "nat_server_destroy"]
EXPORTS = EXPORTS + SERVER_EXPORTS


def _server_lib():
    L = lib()
    if not getattr(L, "_server_bound", False):
        vp = ctypes.c_void_p
        L.nat_server_create.restype = ctypes.c_int
        L.nat_server_create.argtypes = [ctypes.POINTER(_ServerOpts), ctypes.POINTER(vp)]
        L.nat_server_submit.restype = ctypes.c_int
        L.nat_server_submit.argtypes = [vp, ctypes.POINTER(_Request), ctypes.POINTER(ctypes.c_int64)]
        L.nat_server_wait.restype = ctypes.c_int
        L.nat_server_wait.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(_Result), _dp, ctypes.c_int64]
        L.nat_server_stats.restype = ctypes.c_int
        L.nat_server_stats.argtypes = [vp, ctypes.POINTER(_ServerStats)]
        L.nat_server_destroy.argtypes = [vp]
        L._server_bound = True
    return L


class Server:
    """Continuous-batching request server over the C ABI (one GPU).

    submit() returns an id at once; wait(id) blocks until that solve is retired and
    returns its timings, conservation sums and (optionally) its velocity field.
    """

    def __init__(self, hbm_budget_bytes=0, max_active=8, chunk_steps=50, use_graphs=True, pack=True, autotune=True,
                 setup_threads=4):
        L = _server_lib()
        o = _ServerOpts(int(hbm_budget_bytes), int(max_active), int(chunk_steps), int(bool(use_graphs)),
                        int(bool(pack)), int(bool(autotune)), int(setup_threads))
        h = ctypes.c_void_p()
        rc = L.nat_server_create(ctypes.byref(o), ctypes.byref(h))
        if rc != 0:
            raise NatError(rc, "server creation failed")
        self._h = h

    def submit(self, problem, stencil, refinement, order, reynolds, mach, nsteps, precision="fp64", cfl=0.4):
        r = _Request(PROBLEMS[problem], STENCILS[stencil], int(refinement), int(order), float(reynolds), float(mach),
                     float(cfl), int(nsteps), PRECISIONS[precision])
        i = ctypes.c_int64()
        rc = _server_lib().nat_server_submit(self._h, ctypes.byref(r), ctypes.byref(i))
        if rc != 0:
            raise NatError(rc, "request rejected")
        return i.value

    def wait(self, rid, want_u=False, n_dofs=None, dim=None):
        res = _Result()
        u = None
        if want_u:
            u = np.zeros(int(dim) * int(n_dofs))
        rc = _server_lib().nat_server_wait(self._h, int(rid), ctypes.byref(res), _ptr(u) if u is not None else None,
                                           0 if u is None else u.size)
        out = {k: getattr(res, k) for k, _ in _Result._fields_}
        out["error"] = res.error.decode(errors="replace")
        out["status"] = rc
        if u is not None:
            out["u"] = u.reshape(int(dim), int(n_dofs))
        return out

    def stats(self):
        st = _ServerStats()
        _server_lib().nat_server_stats(self._h, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            _server_lib().nat_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request_dofs(problem, stencil, refinement, order):
    """FE_Q DoF count and dimension of a server request's mesh."""
    N = 2 ** refinement
    if problem == "TGV3D":
        return (N * order + 1) ** 3, 3
    if problem == "TGV2D":
        return (N * order + 1) ** 2, 2
    M = max(N // 4, 1)
    return (N * order + 1) * (M * order + 1), 2
