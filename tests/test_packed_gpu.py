"""Packed structured steps and the tile autotuner (no reference counterpart).

nat_step_packed advances several solves with one launch per step; each member's tiles run
exactly as in its own launch, so every member's populations and fields must equal those of
its own nat_step bit for bit, and the FP64 ones must match the oracle's stream + collide
loop (CFDSolver.cpp:884-888) at 1e-12. nat_autotune times the tile shapes without changing
the state; results do not depend on the tile it picks.
"""
import math

import numpy as np
import pytest

import natrium_amd as nat
from natrium_amd.problems import box_vertices, lbm_parameters, tgv2d_fields, tgv3d_fields
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _tgv2d(N, p, Re, Ma, precision="fp64"):
    cs, s = lbm_parameters(1.0, Ma)
    L = 2 * math.pi
    nu = L / Re
    v = box_vertices(2, [N, N], [0, 0], [L, L])
    sol = nat.Solver("D2Q9", p, [N, N], v, scaling=s, viscosity=nu, cfl=0.4, precision=precision)
    rho, u = tgv2d_fields(sol.coords(), nu, cs, L)
    sol.init_equilibrium(rho, u)
    sol.collide()
    return sol, (v, s, nu, rho, u)


def _tgv3d(N, p, stencil, Re, Ma, precision="fp64"):
    cs, s = lbm_parameters(1.0, Ma)
    L = 2 * math.pi
    v = box_vertices(3, [N] * 3, [0] * 3, [L] * 3)
    sol = nat.Solver(stencil, p, [N] * 3, v, scaling=s, viscosity=1.0 / Re, cfl=0.4, precision=precision)
    rho, u = tgv3d_fields(sol.coords(), cs)
    sol.init_equilibrium(rho, u)
    sol.collide()
    return sol


def _state(sol):
    rho, u, _ = sol.macro()
    return sol.download_f(), rho, u


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_packed_2d_members_equal_their_own_steps_and_the_oracle(precision):
    cfgs = [(8, 4, 10.0, 0.05), (12, 4, 100.0, 0.1), (8, 4, 1600.0, 0.2)]
    packed = [_tgv2d(*c, precision=precision)[0] for c in cfgs]
    alone = [_tgv2d(*c, precision=precision)[0] for c in cfgs]
    keys = {s.pack_key() for s in packed}
    assert len(keys) == 1, keys
    print(f"pack of {len(packed)} D2Q9 p=4 {precision}, tile {packed[0].struct_tile()}", flush=True)
    for n in (7, 24):  # odd and even step counts, then further steps from the packed state
        nat.step_packed(packed, n)
        for s in alone:
            s.step(n)
        for a, b in zip(packed, alone):
            fa, ra, ua = _state(a)
            fb, rb, ub = _state(b)
            np.testing.assert_array_equal(fa, fb)
            np.testing.assert_array_equal(ra, rb)
            np.testing.assert_array_equal(ua, ub)
            assert a.time == b.time
    if precision == "fp64":
        sol, (v, s, nu, rho, u) = _tgv2d(*cfgs[1])
        o = Oracle("D2Q9", 4, [12, 12], v, cartesian=True, scaling=s, viscosity=nu, cfl=0.4)
        o.assemble()
        o.init(rho, u)
        o.collide()
        o.steps(31)
        fo = o.get_f()
        assert np.abs(packed[1].download_f() - fo).max() / np.abs(fo).max() < 1e-12


def test_packed_3d_headline_tile_members_equal_their_own_steps():
    """Two 16^3-cell p = 4 D3Q27 FP32 solves (the 16x8x8 tile of the headline kernel) and their
    packed launches."""
    a = [_tgv3d(16, 4, "D3Q27", Re, 0.1, "fp32") for Re in (400.0, 800.0)]
    b = [_tgv3d(16, 4, "D3Q27", Re, 0.1, "fp32") for Re in (400.0, 800.0)]
    assert a[0].struct_tile() == (16, 8, 8) and a[0].pack_key() == a[1].pack_key()
    nat.step_packed(a, 9)
    for s in b:
        s.step(9)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.download_f(), y.download_f())


def test_packed_rejects_mixed_keys_and_unpackable_contexts():
    s64 = _tgv3d(4, 3, "D3Q19", 100.0, 0.05, "fp64")
    s32 = _tgv3d(4, 3, "D3Q19", 100.0, 0.05, "fp32")
    assert s64.pack_key() != s32.pack_key()
    f0 = s64.download_f()
    with pytest.raises(nat.NatError):
        nat.step_packed([s64, s32], 3)
    with pytest.raises(nat.NatError):
        nat.step_packed([s64, s64], 3)
    np.testing.assert_array_equal(s64.download_f(), f0)  # nothing ran
    p6 = _tgv3d(2, 6, "D3Q19", 100.0, 0.05)
    with pytest.raises(nat.NatError):
        p6.pack_key()
    cs, s = lbm_parameters(1.0, 0.05)
    gen = nat.Solver("D2Q9", 3, [4, 4], box_vertices(2, [4, 4], [0, 0], [1, 1]), scaling=s, viscosity=0.01,
                     backend="general")
    with pytest.raises(nat.NatError):
        gen.pack_key()


@pytest.mark.parametrize("stencil,N,p", [("D3Q27", 8, 4), ("D3Q19", 8, 2), ("D2Q9", 16, 3)])
def test_autotune_keeps_state_and_results(stencil, N, p):
    if stencil == "D2Q9":
        a, _ = _tgv2d(N, p, 100.0, 0.1, "fp32")
        b, _ = _tgv2d(N, p, 100.0, 0.1, "fp32")
    else:
        a = _tgv3d(N, p, stencil, 400.0, 0.1, "fp32")
        b = _tgv3d(N, p, stencil, 400.0, 0.1, "fp32")
    f0, r0, u0 = _state(a)
    t0 = a.time
    tile = a.autotune(5)
    print(f"{stencil} {N} cells p={p}: autotuned tile {tile}", flush=True)
    assert tile == a.struct_tile()
    f1, r1, u1 = _state(a)
    np.testing.assert_array_equal(f0, f1)
    np.testing.assert_array_equal(u0, u1)
    assert a.time == t0
    a.step(12)
    b.step(12)
    np.testing.assert_array_equal(a.download_f(), b.download_f())
    # the choice is cached per (instantiation, mesh): a second context gets it without timing
    c = _tgv3d(N, p, stencil, 400.0, 0.1, "fp32") if stencil != "D2Q9" else _tgv2d(N, p, 100.0, 0.1, "fp32")[0]
    assert c.autotune(5) == tile
