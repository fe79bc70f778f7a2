"""nat_sums_noalloc, the allocation-free variant of nat_sums, through the C ABI.

Its partial sums go to a fixed per-GPU scratch instead of the context's buffer; the two
entry points run the same kernels, so their results must agree bit for bit. Two calls in
a row on different contexts, the second on a smaller mesh, check that nothing of the first
call's scratch contents leaks into the second; a mesh beyond the scratch is rejected with
an error before any launch.
"""
import math

import numpy as np
import pytest

import natrium_amd as nat
from natrium_amd.problems import box_vertices, lbm_parameters, perturbed_vertices, tgv2d_fields, tgv3d_fields

pytestmark = pytest.mark.gpu


def _tgv3d(precision, N=8, p=3, backend="structured"):
    cs, s = lbm_parameters(1.0, 0.1)
    L = 2 * math.pi
    v = box_vertices(3, [N] * 3, [0] * 3, [L] * 3)
    g = nat.Solver("D3Q19", p, [N] * 3, v, cartesian=True, scaling=s, viscosity=0.01, cfl=0.4,
                   precision=precision, backend=backend)
    rho, u = tgv3d_fields(g.coords(), cs)
    g.init_equilibrium(rho, u)
    g.collide()
    g.step(3)
    return g


def _tgv2d(precision, N=(4, 3), p=2, scheme="SL"):
    cs, s = lbm_parameters(1.0, 0.1)
    L = 2 * math.pi
    v = perturbed_vertices(2, list(N), [0, 0], [L, L], amplitude=0.1, seed=3)
    g = nat.Solver("D2Q9", p, list(N), v, cartesian=False, scaling=s, viscosity=0.05, cfl=0.4,
                   precision=precision, scheme=scheme)
    rho, u = tgv2d_fields(g.coords(), 0.05, cs, L)
    g.init_equilibrium(rho, u)
    g.collide()
    g.step(2)
    return g


def _same(a, b):
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])
    assert a[2] == b[2]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_sums_equal_allocating_sums(precision):
    big = _tgv3d(precision)                    # 8^3 cells, p = 3: 32,768 (cell, node) pairs
    small = _tgv2d(precision)                  # 4 x 3 cells, p = 2: 108 pairs
    dg = _tgv2d(precision, scheme="SEDG")      # DG layout
    r_big = big.sums(noalloc=True)
    r_small = small.sums(noalloc=True)         # second call, smaller input, same scratch
    r_dg = dg.sums(noalloc=True)
    _same(r_big, big.sums())
    _same(r_small, small.sums())
    _same(r_dg, dg.sums())
    # again after the allocating calls and in the other order
    _same(small.sums(noalloc=True), r_small)
    _same(big.sums(noalloc=True), r_big)
    assert r_big[0] > 0 and r_small[0] > 0 and r_dg[0] > 0


def test_noalloc_sums_reject_mesh_beyond_scratch():
    """2100^2 cells at p = 1: 17,640,000 (cell, node) pairs > 16,777,216."""
    cs, s = lbm_parameters(1.0, 0.1)
    n = 2100
    v = box_vertices(2, [n, n], [0, 0], [1.0, 1.0])
    g = nat.Solver("D2Q9", 1, [n, n], v, cartesian=True, scaling=s, viscosity=1e-4, cfl=0.4,
                   precision="fp32", backend="structured")
    g.init_equilibrium(np.ones(g.n), np.zeros((2, g.n)))
    g.collide()
    with pytest.raises(nat.NatError, match="exceed the fixed scratch"):
        g.sums(noalloc=True)
    # the allocating entry point still serves it: unit density on the unit square
    m, _, _ = g.sums()
    assert m == pytest.approx(1.0, rel=1e-6)
