"""GPU parity: compressible D2Q25H (f, g, quartic equilibrium, FP64 temperature).

SL streaming of f and g with one set of departure records, fused with the
compressible BGK collision (collision_advanced/CollisionOperator.h:113-224), vs the
FP64 oracle. Tolerances: FP64 f and g within 1e-12 relative (max abs diff / max abs)
after every checked step; temperature and Knudsen sensor within 1e-11 relative;
FP32-shifted temperature within the bound stated in the test.
"""
import math

import numpy as np
import pytest

import natrium_amd as nat
from natrium_amd.problems import box_vertices, lbm_parameters, perturbed_vertices, riemann_quadrant_fields, riemann_states
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _riemann_pair(N, p, precision="fp64", perturb=False, prandtl=0.0, sutherland=False, viscosity=1e-3):
    # unit square, Riemann quadrants (periodic box for the parity run), |u| <= 0.7, Ma ~ 1
    s = 2.0
    if perturb:
        v = perturbed_vertices(2, N, [0, 0], [1, 1], amplitude=0.1, seed=1234)
    else:
        v = box_vertices(2, N, [0, 0], [1, 1])
    kw = dict(scaling=s, viscosity=viscosity, cfl=0.4, compressible=True, gamma=1.4, prandtl=prandtl,
              sutherland=sutherland)
    g = nat.Solver("D2Q25H", p, N, v, cartesian=not perturb, precision=precision, **kw)
    o = Oracle("D2Q25H", p, N, v, cartesian=not perturb, **kw)
    o.assemble()
    x = g.coords()
    np.testing.assert_allclose(x, o.coords(), rtol=0, atol=1e-14)
    rho, u, T = riemann_quadrant_fields(x, riemann_states(7))
    g.init_equilibrium(rho, u, T)
    o.init(rho, u, T)
    return g, o


def test_compressible_init_matches_oracle():
    g, o = _riemann_pair([6, 6], 2)
    assert _rel(g.download_f(), o.get_f()) < 1e-15
    assert _rel(g.download_g(), o.get_g()) < 1e-15


@pytest.mark.parametrize("N,p,perturb,prandtl,sutherland", [
    ([8, 8], 2, False, 0.0, False),
    ([6, 5], 3, False, 0.71, False),
    ([5, 6], 2, True, 0.0, True),
    ([4, 4], 4, True, 0.71, True),
])
def test_compressible_fp64_per_step(N, p, perturb, prandtl, sutherland):
    g, o = _riemann_pair(N, p, perturb=perturb, prandtl=prandtl, sutherland=sutherland)
    g.collide()
    o.collide()
    assert _rel(g.download_f(), o.get_f()) < FP64_TOL
    assert _rel(g.download_g(), o.get_g()) < FP64_TOL
    for step in range(1, 9):
        g.step(1)
        o.steps(1)
        ef = _rel(g.download_f(), o.get_f())
        eg = _rel(g.download_g(), o.get_g())
        assert ef < FP64_TOL and eg < FP64_TOL, (step, ef, eg)
    rg, ug, Tg = g.macro()
    ro, uo, To, ko = o.macro()
    assert _rel(rg, ro) < FP64_TOL
    assert _rel(Tg, To) < 1e-11
    assert _rel(ug, uo) < 1e-10
    kg = g.knudsen()
    assert np.abs(kg - ko).max() <= 1e-11 * max(np.abs(ko).max(), 1e-3)


def test_compressible_fp32_shifted_temperature_error():
    """FP32-shifted f and g vs the FP64 oracle, 100 steps: relative L2 error of T below 1e-5."""
    g, o = _riemann_pair([8, 8], 2, precision="fp32")
    g.collide()
    o.collide()
    g.step(100)
    o.steps(100)
    _, _, Tg = g.macro()
    _, _, To, _ = o.macro()
    err = np.sqrt(((Tg - To) ** 2).sum() / (To ** 2).sum())
    assert err < 1e-5, err
