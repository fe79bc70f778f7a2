"""Problem descriptions on the hot path: lattice meshes and Taylor-Green initial fields.

Host-side mirror of the reference benchmarks used by the solver configurations:
  benchmarks/TaylorGreenVortex2D.cpp:36-97  (fields, hyper_cube [0, L]^2, all periodic)
  benchmarks/TaylorGreenVortex3D.cpp:38-119 (fields, [0, 2 pi]^3, all periodic)
and of the solver parameter conventions
  utilities/CFDSolverUtilities.cpp:92-99     dt = CFL h_min / (|c|max p^2)
  collision_advanced/AuxiliaryCollisionFunctions.h:63-68  tau = nu / (dt cs2) + 1/2
"""
import math

import numpy as np


def box_vertices(dim, ncells, lo, hi):
    """Vertices of a lattice box mesh, lexicographic (x fastest), shape (nv, dim)."""
    axes = [lo[a] + (hi[a] - lo[a]) * np.arange(ncells[a] + 1, dtype=np.float64) / ncells[a] for a in range(dim)]
    grids = np.meshgrid(*axes, indexing="ij")
    # lexicographic with x fastest: transpose so the last index varies slowest
    cols = [g.transpose(tuple(reversed(range(dim)))).reshape(-1) for g in grids]
    return np.stack(cols, axis=1)


def perturbed_vertices(dim, ncells, lo, hi, amplitude=0.1, seed=1234):
    """Box vertices with every interior vertex moved uniformly by +-amplitude*h per component.

    Generator: numpy.random.default_rng(seed).uniform(-1, 1) drawn in vertex order, one value per
    (vertex, component), for all vertices; boundary vertices discard their draws so periodic faces match.
    """
    v = box_vertices(dim, ncells, lo, hi)
    h = np.array([(hi[a] - lo[a]) / ncells[a] for a in range(dim)])
    rng = np.random.default_rng(seed)
    d = rng.uniform(-1.0, 1.0, size=v.shape) * amplitude * h
    idx = np.stack(np.unravel_index(np.arange(v.shape[0]), tuple(ncells[a] + 1 for a in reversed(range(dim)))), 1)
    idx = idx[:, ::-1]  # (i, j[, k]) with x fastest
    interior = np.ones(v.shape[0], dtype=bool)
    for a in range(dim):
        interior &= (idx[:, a] > 0) & (idx[:, a] < ncells[a])
    v[interior] += d[interior]
    return v


def lbm_parameters(U=1.0, Ma=0.1):
    """Reference driver convention: cs = U/Ma, stencil scaling s = sqrt(3) cs."""
    cs = U / Ma
    return cs, math.sqrt(3.0) * cs


def tgv2d_fields(x, viscosity, cs, L=2 * math.pi, t=0.0, rho_analytic=True):
    """TaylorGreenVortex2D::AnalyticVelocity / AnalyticDensity (TaylorGreenVortex2D.cpp:36-58)."""
    X, Y = x[:, 0], x[:, 1]
    k = (2 * math.pi)
    dec = math.exp(-2 * viscosity * t * k / L)
    u = np.empty((2, x.shape[0]))
    u[0] = np.sin(X * k / L) * np.cos(Y * k / L) * dec
    u[1] = -np.cos(X * k / L) * np.sin(Y * k / L) * dec
    if rho_analytic:
        p = 1.0 / 4.0 * (np.cos(2 * (X * k / L)) + np.cos(2 * Y * k / L)) * math.exp(-4 * viscosity * t * k / L)
        rho = 1.0 + p / (cs * cs)
    else:
        rho = np.ones(x.shape[0])
    return rho, u


def tgv3d_fields(x, cs, rho_analytic=True):
    """TaylorGreenVortex3D::InitialVelocity / InitialDensity (TaylorGreenVortex3D.cpp:38-59)."""
    X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
    u = np.empty((3, x.shape[0]))
    u[0] = np.sin(X) * np.cos(Y) * np.cos(Z)
    u[1] = -np.cos(X) * np.sin(Y) * np.cos(Z)
    u[2] = 0.0
    if rho_analytic:
        p = 1.0 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2)
        rho = 1.0 + p / (cs * cs)
    else:
        rho = np.ones(x.shape[0])
    return rho, u


def riemann_quadrant_fields(x, states, lo=(0.0, 0.0), hi=(1.0, 1.0)):
    """Four-quadrant piecewise-constant (rho, u, v, p) -> (rho, u, T = p/rho)."""
    xm = 0.5 * (lo[0] + hi[0])
    ym = 0.5 * (lo[1] + hi[1])
    q = (x[:, 0] >= xm).astype(int) + 2 * (x[:, 1] >= ym).astype(int)
    st = np.asarray(states, dtype=np.float64)  # 4 x (rho, u, v, p): quadrants SW, SE, NW, NE
    rho = st[q, 0]
    u = np.stack([st[q, 1], st[q, 2]])
    T = st[q, 3] / st[q, 0]
    return rho, u, T


def riemann_states(seed=7):
    """Seeded quadrant states: rho in [0.5,1.5], p in [0.3,1.5], |u| <= 0.7."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(4):
        rho = rng.uniform(0.5, 1.5)
        p = rng.uniform(0.3, 1.5)
        ang = rng.uniform(0, 2 * math.pi)
        mag = rng.uniform(0, 0.7)
        out.append((rho, mag * math.cos(ang), mag * math.sin(ang), p))
    return out


def divergence_free_field(x, modes=8, amplitude=0.05, seed=1234, L=2 * math.pi):
    """Seeded random divergence-free Fourier velocity field (3D): sum of curl-free-projected modes."""
    rng = np.random.default_rng(seed)
    u = np.zeros((3, x.shape[0]))
    kk = 2 * math.pi / L
    for _ in range(modes * 3):
        k = rng.integers(-modes, modes + 1, size=3).astype(np.float64)
        if not k.any():
            continue
        a = rng.normal(size=3)
        a -= k * (a @ k) / (k @ k)  # project out the longitudinal part
        ph = rng.uniform(0, 2 * math.pi)
        s = np.sin(kk * (x @ k) + ph)
        u += np.outer(a, s)
    nrm = np.sqrt((u ** 2).sum(0).mean())
    return u * (amplitude / max(nrm, 1e-300))


def couette2d_velocity(y, t, U, L=1.0, viscosity=1.0):
    """CouetteFlow2D::AnalyticVelocity (benchmarks/CouetteFlow2D.cpp:87-125): x-velocity of the
    unsteady plane Couette flow on [0, L]^2 (bottom wall at rest, top wall moving with U,
    started from rest at t = 0), as its sine series truncated where exp(-nu lambda^2 t) < 1e-20."""
    y = np.asarray(y, dtype=np.float64)
    if t < 0.00001:
        return np.where(y < L - 0.00001, 0.0, U)
    s = U * y / L
    for i in range(1, 10001):
        lam = i * math.pi / L
        e = math.exp(-viscosity * lam * lam * t)
        s = s + (1.0 if i % 2 == 0 else -1.0) * 2 * U / (lam * L) * e * np.sin(lam * y)
        if e < 1e-20:
            break
    return s


def serving_mix(n, seed, full=True):
    """Seeded serving workload: 2D or 3D Taylor-Green vortex, or a shock tube on a 2D strip;
    refinement 3-5, p in {2, 3, 4, 6}, D2Q9 (2D) or D3Q19 / D3Q27 (3D), Re log-uniform in
    [10, 1600], Ma uniform in [0.05, 0.2], step count log-uniform in [100, 5000].
    full=False caps refinement at 4 (3 for 3D at p = 6) and steps at 1000."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        prob = ["TGV2D", "TGV3D", "SHOCK_TUBE"][rng.integers(3)]
        ref = int(rng.integers(3, 6 if full else 5))
        p = int([2, 3, 4, 6][rng.integers(4)])
        st = "D2Q9" if prob != "TGV3D" else ["D3Q19", "D3Q27"][rng.integers(2)]
        if prob == "TGV3D" and not full:
            ref = min(ref, 3 if p == 6 else 4)
        re_ = float(math.exp(rng.uniform(math.log(10), math.log(1600))))
        ma = float(rng.uniform(0.05, 0.2))
        steps = int(math.exp(rng.uniform(math.log(100), math.log(5000 if full else 1000))))
        out.append(dict(problem=prob, stencil=st, refinement=ref, order=p, reynolds=re_, mach=ma, nsteps=steps))
    return out
