"""ctypes wrapper of the oracle (FP64 CPU restatement of the NATriuM hot path).

TEST INFRASTRUCTURE ONLY. Imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker and the CPU baseline. The product
package (natrium_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

STENCILS = {"D2Q9": (0, 2, 9), "D3Q19": (1, 3, 19), "D3Q27": (2, 3, 27), "D2Q25H": (3, 2, 25)}
SCHEMES = {"SL": 0, "SEDG": 1, "SEDG_CENTRAL": 2}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_int64)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _dp, ctypes.c_int, _ip,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double]
        L.orc_error.restype = ctypes.c_char_p
        L.orc_error.argtypes = [ctypes.c_void_p]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for name, rt in [("orc_n_dofs", ctypes.c_int64), ("orc_dt", ctypes.c_double), ("orc_time", ctypes.c_double),
                         ("orc_tau", ctypes.c_double), ("orc_Q", ctypes.c_int), ("orc_collide", ctypes.c_int),
                         ("orc_stream", ctypes.c_int), ("orc_op_identity_rows", ctypes.c_int64)]:
            getattr(L, name).restype = rt
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.orc_assemble.restype = ctypes.c_int
        L.orc_assemble.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_steps.restype = ctypes.c_int
        L.orc_steps.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_node_coords.argtypes = [ctypes.c_void_p, _dp]
        L.orc_op_nnz.restype = ctypes.c_int64
        L.orc_op_nnz.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_op_get.argtypes = [ctypes.c_void_p, ctypes.c_int, _lp, _lp, _dp]
        L.orc_locate.restype = ctypes.c_int
        L.orc_locate.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, _dp]
        L.orc_init.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        for nm in ("orc_get_f", "orc_set_f", "orc_get_g", "orc_set_g"):
            getattr(L, nm).argtypes = [ctypes.c_void_p, _dp]
        L.orc_get_macro.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_sums.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_face_crossed_first.restype = ctypes.c_int
        L.orc_face_crossed_first.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_stencil.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _ip, _dp]
        L.orc_gll.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_lagrange.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_quartic_feq.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_double, _dp]
        L.orc_rk5_scalar.restype = ctypes.c_double
        L.orc_rk5_scalar.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_set_compressible_options.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
        L.orc_set_walls.argtypes = [ctypes.c_void_p, _dp]
        L.orc_set_stabilizer.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_stabilizer_matrix.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        L.orc_op_hits.argtypes = [ctypes.c_void_p, _lp, _lp]
        L.orc_op_src.restype = ctypes.c_int
        L.orc_op_src.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int8)]
        L.orc_op_bvec.restype = ctypes.c_int
        L.orc_op_bvec.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_op_opp.restype = ctypes.c_int
        L.orc_op_opp.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_collide_init.restype = ctypes.c_int
        L.orc_collide_init.argtypes = [ctypes.c_void_p, _dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def stencil(name, scaling=1.0):
    sid, D, Q = STENCILS[name]
    dirs = np.zeros((Q, 3))
    w = np.zeros(Q)
    opp = np.zeros(Q, dtype=np.int32)
    cs2 = np.zeros(1)
    lib().orc_stencil(sid, scaling, _p(dirs), _p(w), _p(opp, _ip), _p(cs2))
    return dirs[:, :D].copy(), w, opp, float(cs2[0])


def gll(p):
    x = np.zeros(p + 1)
    w = np.zeros(p + 1)
    lib().orc_gll(p, _p(x), _p(w))
    return x, w


def lagrange(p, j, xi):
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    v = np.zeros_like(xi)
    d = np.zeros_like(xi)
    lib().orc_lagrange(p, j, xi.size, _p(xi), _p(v), _p(d))
    return v, d


def quartic_feq(name, scaling, rho, u, T):
    sid, D, Q = STENCILS[name]
    uu = np.zeros(3)
    uu[:D] = u
    out = np.zeros(Q)
    lib().orc_quartic_feq(sid, scaling, rho, _p(uu), T, _p(out))
    return out


def stabilizer_matrix(name, scaling=1.0):
    sid, D, Q = STENCILS[name]
    P = np.zeros((Q, Q))
    lib().orc_stabilizer_matrix(sid, scaling, _p(P))
    return P


def rk5_scalar(lam, f0, dt, nsteps):
    return lib().orc_rk5_scalar(lam, f0, dt, nsteps)


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


class Oracle:
    """CFDSolver-like FP64 CPU solver (SL or SEDG, incompressible or compressible)."""

    def __init__(self, stencil, p, ncells, vertices, cartesian=True, periodic=None, scaling=1.0, viscosity=1.0,
                 cfl=0.4, scheme="SL", compressible=False, gamma=1.4, prandtl=0.0, sutherland=False,
                 wall_velocity=None, stabilizer=False):
        """periodic: flag per axis (default all periodic). Faces of non-periodic axes are
        velocity-neq bounce-back walls; wall_velocity maps a box face 2*axis+side to its
        velocity (default at rest)."""
        sid, D, Q = STENCILS[stencil]
        self.D, self.Q, self.p = D, Q, p
        nc = np.array(list(ncells) + [1] * (3 - len(ncells)), dtype=np.int32)
        per = np.array([1, 1, 1] if periodic is None else list(periodic) + [1] * (3 - len(periodic)), dtype=np.int32)
        self._verts = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, D)
        self.h = lib().orc_create(D, sid, p, _p(nc, _ip), _p(self._verts), int(cartesian), _p(per, _ip), scaling,
                                  viscosity, cfl, SCHEMES[scheme], int(compressible), gamma)
        err = lib().orc_error(self.h).decode()
        if err:
            raise RuntimeError(err)
        self.n = lib().orc_n_dofs(self.h)
        self.compressible = compressible
        if compressible:
            lib().orc_set_compressible_options(self.h, float(prandtl), int(sutherland))
        wu = np.zeros((6, 3))
        for f, v in (wall_velocity or {}).items():
            wu[int(f), :len(v)] = v
        lib().orc_set_walls(self.h, _p(wu))
        lib().orc_set_stabilizer(self.h, int(stabilizer))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(lib().orc_error(self.h).decode())

    @property
    def dt(self):
        return lib().orc_dt(self.h)

    @property
    def tau(self):
        return lib().orc_tau(self.h)

    @property
    def time(self):
        return lib().orc_time(self.h)

    def coords(self):
        x = np.zeros((self.n, 3))
        lib().orc_node_coords(self.h, _p(x))
        return x[:, :self.D].copy()

    def assemble(self, dt=0.0):
        self._check(lib().orc_assemble(self.h, dt))

    def operator(self, alpha):
        import scipy.sparse as sp
        nnz = lib().orc_op_nnz(self.h, alpha)
        rp = np.zeros(self.n + 1, dtype=np.int64)
        col = np.zeros(nnz, dtype=np.int64)
        val = np.zeros(nnz)
        lib().orc_op_get(self.h, alpha, _p(rp, _lp), _p(col, _lp), _p(val))
        return sp.csr_matrix((val, col, rp), shape=(self.n, self.n))

    def identity_rows(self):
        return lib().orc_op_identity_rows(self.h)

    def operator_src(self, alpha):
        """Source direction of every entry of block alpha (None: all read alpha)."""
        nnz = lib().orc_op_nnz(self.h, alpha)
        src = np.zeros(nnz, dtype=np.int8)
        if lib().orc_op_src(self.h, alpha, _p(src, ctypes.POINTER(ctypes.c_int8))):
            return src
        return None

    def opposite_coupling(self, alpha):
        """SEDG walls: diagonal of the operator block (alpha, opposite alpha), or None."""
        d = np.zeros(self.n)
        return d if lib().orc_op_opp(self.h, alpha, _p(d)) else None

    def boundary_values(self, alpha):
        """Values added to streamed f_alpha at wall hits (None: no hit)."""
        b = np.zeros(self.n)
        return b if lib().orc_op_bvec(self.h, alpha, _p(b)) else None

    def hits(self):
        """(hits, cells holding hits) of the boundary handler's hit list."""
        h = np.zeros(1, dtype=np.int64)
        c = np.zeros(1, dtype=np.int64)
        lib().orc_op_hits(self.h, _p(h, _lp), _p(c, _lp))
        return int(h[0]), int(c[0])

    def locate(self, dof, alpha):
        xi = np.zeros(3)
        c = lib().orc_locate(self.h, dof, alpha, _p(xi))
        return c, xi[:self.D].copy()

    def init(self, rho, u, T=None):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.D, self.n)
        Tp = None
        if T is not None:
            T = np.ascontiguousarray(T, dtype=np.float64)
            Tp = _p(T)
        lib().orc_init(self.h, _p(rho), _p(u), Tp)

    def collide(self):
        self._check(lib().orc_collide(self.h))

    def collide_init(self, u=None):
        """Initialisation collision: the stored velocity (u if given) drives the relaxation."""
        up = None
        if u is not None:
            u = np.ascontiguousarray(u, dtype=np.float64).reshape(self.D, self.n)
            up = _p(u)
        self._check(lib().orc_collide_init(self.h, up))

    def stream(self):
        self._check(lib().orc_stream(self.h))

    def steps(self, n):
        self._check(lib().orc_steps(self.h, n))

    def get_f(self):
        f = np.zeros((self.Q, self.n))
        lib().orc_get_f(self.h, _p(f))
        return f

    def set_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        lib().orc_set_f(self.h, _p(f))

    def get_g(self):
        g = np.zeros((self.Q, self.n))
        lib().orc_get_g(self.h, _p(g))
        return g

    def set_g(self, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        lib().orc_set_g(self.h, _p(g))

    def macro(self):
        rho = np.zeros(self.n)
        u = np.zeros((self.D, self.n))
        T = np.zeros(self.n)
        m = np.zeros(self.n)
        lib().orc_get_macro(self.h, _p(rho), _p(u), _p(T), _p(m))
        return rho, u, T, m

    def sums(self):
        m = np.zeros(1)
        e = np.zeros(1)
        lib().orc_sums(self.h, _p(m), _p(e))
        return float(m[0]), float(e[0])

    def face_crossed_first(self, cell, pin, pout, without_mapping=False):
        a = np.zeros(3)
        b = np.zeros(3)
        a[:self.D] = pin
        b[:self.D] = pout
        pb = np.zeros(3)
        lam = np.zeros(1)
        r = lib().orc_face_crossed_first(self.h, cell, _p(a), _p(b), _p(pb), _p(lam), int(without_mapping))
        return r, pb[:self.D].copy(), float(lam[0])
