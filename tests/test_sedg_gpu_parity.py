"""GPU parity: matrix-free SEDG operator in the low-storage RK5 integrator vs the oracle.

The oracle assembles the reference's block sparse SEDG matrix row by row
(advection/SEDGMinLee.cpp:321-536) and runs RK5 operation by operation
(timeintegration/RungeKutta5LowStorage.cpp:56-105); the GPU evaluates the same
operator matrix-free. Tolerances: FP64 populations within 1e-12 relative
(max |f_gpu - f_oracle| / max |f_oracle|) after every checked step; FP32-shifted
velocities within the bound stated in the test.
"""
import math

import numpy as np
import pytest

import natrium_amd as nat
from natrium_amd.problems import box_vertices, lbm_parameters, perturbed_vertices, tgv2d_fields, tgv3d_fields
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _pair(stencil, p, N, verts, scheme, scaling, nu, precision="fp64", cfl=0.4):
    g = nat.Solver(stencil, p, N, verts, cartesian=False, scaling=scaling, viscosity=nu, cfl=cfl,
                   scheme=scheme, precision=precision)
    o = Oracle(stencil, p, N, verts, cartesian=False, scaling=scaling, viscosity=nu, cfl=cfl, scheme=scheme)
    o.assemble()
    return g, o


def _setup_2d(p, N, scheme, perturb, precision="fp64"):
    cs, s = lbm_parameters(1.0, 0.05)
    L = 2 * math.pi
    nu = L / 10.0
    if perturb:
        v = perturbed_vertices(2, N, [0, 0], [L, L], amplitude=0.1, seed=1234)
    else:
        v = box_vertices(2, N, [0, 0], [L, L])
    g, o = _pair("D2Q9", p, N, v, scheme, s, nu, precision=precision)
    x = g.coords()
    np.testing.assert_allclose(x, o.coords(), rtol=0, atol=1e-14)
    rho, u = tgv2d_fields(x, nu, cs, L)
    g.init_equilibrium(rho, u)
    o.init(rho, u)
    return g, o


@pytest.mark.parametrize("p,N,scheme,perturb", [
    (3, [4, 4], "SEDG", False),
    (2, [5, 3], "SEDG", True),
    (4, [3, 3], "SEDG", False),
    (3, [3, 4], "SEDG_CENTRAL", True),
    (4, [3, 3], "SEDG_CENTRAL", False),
])
def test_sedg_2d_fp64_per_step(p, N, scheme, perturb):
    g, o = _setup_2d(p, N, scheme, perturb)
    assert g.n == o.n == N[0] * N[1] * (p + 1) ** 2
    assert abs(g.dt - o.dt) < 1e-18
    g.collide()
    o.collide()
    assert _rel(g.download_f(), o.get_f()) < FP64_TOL
    # the central flux does not preserve constants and the TGV state collapses after ~8 steps
    # (test_sedg_central_flux_density_collapse_matches_oracle); compare the first steps
    nsteps = 10 if scheme == "SEDG" else 4
    for step in range(1, nsteps + 1):
        g.step(1)
        o.steps(1)
        e = _rel(g.download_f(), o.get_f())
        assert e < FP64_TOL, f"step {step}: {e}"


def test_sedg_central_flux_density_collapse_matches_oracle():
    """The reference's central flux (+1/2 of the face term on the diagonal and on the
    neighbour, SEDGMinLee.cpp:515-518) is not conservative: on the 2D TGV at p=4 the density
    falls below the collision's 1e-10 limit within ten steps. The GPU reports the
    DensityZeroException error at the same step as the oracle, with populations matching the
    oracle until then."""
    g, o = _setup_2d(4, [3, 3], "SEDG_CENTRAL", False)
    g.collide()
    o.collide()
    fail_g = fail_o = None
    for step in range(1, 16):
        try:
            g.step(1)
        except nat.DensityZeroError:
            fail_g = step
        try:
            o.steps(1)
        except RuntimeError:
            fail_o = step
        if fail_g or fail_o:
            break
        rel = _rel(g.download_f(), o.get_f())
        assert rel < 1e-10, (step, rel)
    assert fail_o is not None and fail_g == fail_o, (fail_g, fail_o)


@pytest.mark.parametrize("p,scheme", [(3, "SEDG"), (4, "SEDG_CENTRAL")])
def test_sedg_stream_only_matches_rk5(p, scheme):
    """nat_stream = one RK5 step of the SEDG operator without collision."""
    g, o = _setup_2d(p, [4, 3], scheme, True)
    for _ in range(3):
        g.stream()
        o.stream()
        e = _rel(g.download_f(), o.get_f())
        assert e < FP64_TOL, e
    assert abs(g.time - o.time) < 1e-15


@pytest.mark.parametrize("stencil,p,N,perturb", [("D3Q19", 3, [3, 3, 3], True), ("D3Q27", 2, [3, 2, 4], False),
                                                 ("D3Q19", 4, [2, 2, 2], False), ("D3Q27", 3, [2, 3, 2], True)])
def test_sedg_3d_fp64_parity(stencil, p, N, perturb):
    cs, s = lbm_parameters(1.0, 0.05)
    L = 2 * math.pi
    if perturb:
        v = perturbed_vertices(3, N, [0] * 3, [L] * 3, amplitude=0.1, seed=1234)
    else:
        v = box_vertices(3, N, [0] * 3, [L] * 3)
    g, o = _pair(stencil, p, N, v, "SEDG", s, 2 * math.pi / 100)
    x = g.coords()
    np.testing.assert_allclose(x, o.coords(), rtol=0, atol=1e-14)
    rho, u = tgv3d_fields(x, cs)
    g.init_equilibrium(rho, u)
    o.init(rho, u)
    g.collide()
    o.collide()
    for _ in range(3):
        g.step(2)
        o.steps(2)
        e = _rel(g.download_f(), o.get_f())
        assert e < FP64_TOL, e


def test_sedg_sums_and_macro():
    g, o = _setup_2d(3, [4, 4], "SEDG", True)
    g.collide()
    o.collide()
    g.step(5)
    o.steps(5)
    m_g, mom_g, e_g = g.sums()
    m_o, e_o = o.sums()
    assert abs(m_g - m_o) / m_o < 1e-12
    assert abs(e_g - e_o) / e_o < 1e-10
    rg, ug, _ = g.macro()
    ro, uo, _, _ = o.macro()
    assert _rel(rg, ro) < FP64_TOL
    assert _rel(ug, uo) < 1e-10


def test_sedg_fp32_shifted_velocity_error():
    """FP32-shifted SEDG (upwind flux) vs the FP64 oracle, 50 steps on a perturbed 2D mesh.

    Tolerance 1e-6 on the relative velocity L2 error. (The reference's central flux adds
    +1/2 of the face term on both sides and does not preserve constants, so it is only
    checked step by step over a few steps above.)"""
    g, o = _setup_2d(3, [4, 4], "SEDG", True, precision="fp32")
    g.collide()
    o.collide()
    g.step(50)
    o.steps(50)
    _, ug, _ = g.macro()
    _, uo, _, _ = o.macro()
    err = np.sqrt(((ug - uo) ** 2).sum() / (uo ** 2).sum())
    assert err < 1e-6, err


def test_sedg_3d_fp32_shifted_velocity_error():
    """The C3 SEDG configuration in FP32-shifted storage against the FP64 oracle, on a 4^3-cell
    copy of its mesh: [0, 2 pi]^3 with interior vertices moved by U(-0.1 h, 0.1 h) (seed 1234),
    p = 3, D3Q19, Ma = 0.05, nu = 0.01, seeded divergence-free Fourier velocity (8 modes,
    amplitude 0.05 c_s), upwind flux, 50 steps. Tolerance 1e-6 on the relative velocity L2
    error (sqrt(sum |u_gpu - u_oracle|^2 / sum |u_oracle|^2))."""
    from natrium_amd.problems import divergence_free_field

    cs, s = lbm_parameters(1.0, 0.05)
    L = 2 * math.pi
    v = perturbed_vertices(3, [4] * 3, [0] * 3, [L] * 3, amplitude=0.1, seed=1234)
    g, o = _pair("D3Q19", 3, [4] * 3, v, "SEDG", s, 0.01, precision="fp32")
    x = g.coords()
    np.testing.assert_allclose(x, o.coords(), rtol=0, atol=1e-14)
    u = divergence_free_field(x, modes=8, amplitude=0.05 * cs, seed=1234)
    g.init_equilibrium(np.ones(g.n), u)
    o.init(np.ones(o.n), u)
    g.collide()
    o.collide()
    g.step(50)
    o.steps(50)
    _, ug, _ = g.macro()
    _, uo, _, _ = o.macro()
    err = float(np.sqrt(((ug - uo) ** 2).sum() / (uo ** 2).sum()))
    print(f"C3-shaped SEDG FP32 velocity rel L2 vs oracle after 50 steps: {err:.3e}", flush=True)
    assert err < 1e-6, err
