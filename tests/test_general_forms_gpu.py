"""GPU tests of the general SL kernel's two FP32 3D forms (no reference counterpart).

The six-wave form and the DEEP form (4 waves, more source rows and the next departure
record in flight) run the same arithmetic in the same order, so every population must be
bit-identical between them, on periodic and walled perturbed meshes and at the C3 size
(389,017 DoFs, where the default rule takes the six-wave form). The DEEP form, which small
meshes take by default, is also checked against the FP64 oracle on the FP32 budget. The
launch-shape autotuner must leave populations, fields and time untouched.
"""
import math

import numpy as np
import pytest

import natrium_amd as nat
from natrium_amd.problems import (box_vertices, divergence_free_field, lbm_parameters, perturbed_vertices,
                                  tgv3d_fields)
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def _solver(stencil, p, N, v, periodic=None, precision="fp32", form=None):
    cs, s = lbm_parameters(1.0, 0.05)
    g = nat.Solver(stencil, p, N, v, cartesian=False, scaling=s, viscosity=0.01, cfl=0.4, precision=precision,
                   backend="general", periodic=periodic)
    if form is not None:
        g.set_general_form(form)
    u = divergence_free_field(g.coords(), modes=4, amplitude=0.05 * cs, seed=11)
    if periodic is not None:
        u[2] = 0.0
    g.init_equilibrium(np.ones(g.n), u)
    g.collide()
    return g


@pytest.mark.parametrize("stencil,p,walls", [("D3Q19", 3, False), ("D3Q27", 2, False), ("D3Q27", 4, False),
                                             ("D3Q19", 3, True), ("D3Q27", 3, True)])
def test_general_fp32_forms_bit_identical(stencil, p, walls):
    L = 2 * math.pi
    N = [5, 5, 4]
    v = perturbed_vertices(3, N, [0] * 3, [L] * 3, amplitude=0.1, seed=5)
    periodic = [1, 1, 0] if walls else None
    runs = {}
    for form in (-1, 0, 1):
        g = _solver(stencil, p, N, v, periodic=periodic, form=form)
        g.step(15)
        runs[form] = (g.download_f(), g.macro()[1], g.time)
    print(f"{stencil} p={p} walls={walls}: {runs[0][0].shape[1]} DoFs", flush=True)
    for form in (-1, 1):
        np.testing.assert_array_equal(runs[form][0], runs[0][0])
        np.testing.assert_array_equal(runs[form][1], runs[0][1])
        assert runs[form][2] == runs[0][2]


def test_general_fp32_deep_form_against_oracle():
    """DEEP form (the default on a mesh this small) vs the FP64 oracle: relative velocity L2
    error below 1e-6 after 40 steps (the FP32-shifted budget)."""
    cs, s = lbm_parameters(1.0, 0.05)
    L = 2 * math.pi
    N = [4, 4, 4]
    v = perturbed_vertices(3, N, [0] * 3, [L] * 3, amplitude=0.1, seed=1234)
    g = nat.Solver("D3Q19", 3, N, v, cartesian=False, scaling=s, viscosity=0.01, cfl=0.4, precision="fp32",
                   backend="general")
    g.set_general_form(1)
    o = Oracle("D3Q19", 3, N, v, cartesian=False, scaling=s, viscosity=0.01, cfl=0.4)
    o.assemble()
    rho, u = tgv3d_fields(g.coords(), cs)
    g.init_equilibrium(rho, u)
    o.init(rho, u)
    g.collide()
    o.collide()
    g.step(40)
    o.steps(40)
    _, ug, _ = g.macro()
    _, uo, _, _ = o.macro()
    err = float(np.sqrt(((ug - uo) ** 2).sum() / (uo ** 2).sum()))
    print(f"DEEP form vs oracle, 40 steps: velocity L2 {err:.3e}", flush=True)
    assert err < 1e-6


def test_general_fp32_forms_bit_identical_at_c3_size():
    """24^3 perturbed cells, D3Q19, p=3 (389,017 DoFs, 1520 blocks of 256): both forms on the
    C3 mesh give the same populations after 10 steps."""
    L = 2 * math.pi
    N = [24] * 3
    v = perturbed_vertices(3, N, [0] * 3, [L] * 3, amplitude=0.1, seed=1234)
    out = []
    for form in (0, 1):
        g = _solver("D3Q19", 3, N, v, form=form)
        assert g.n == 389017
        g.step(10)
        out.append(g.download_f())
        del g
    np.testing.assert_array_equal(out[0], out[1])


def test_general_autotune_keeps_state_and_results():
    L = 2 * math.pi
    N = [6, 6, 6]
    v = perturbed_vertices(3, N, [0] * 3, [L] * 3, amplitude=0.1, seed=9)
    a = _solver("D3Q27", 3, N, v)
    b = _solver("D3Q27", 3, N, v)
    f0 = a.download_f()
    _, u0, _ = a.macro()
    t0 = a.time
    shape = a.autotune(5)
    print(f"general D3Q27 p=3 {N}: autotuned form {shape}", flush=True)
    assert shape[0] == 256 and shape[1] in (4, 6) and 1 <= shape[2] <= 4
    np.testing.assert_array_equal(a.download_f(), f0)
    np.testing.assert_array_equal(a.macro()[1], u0)
    assert a.time == t0
    a.step(12)
    b.step(12)
    np.testing.assert_array_equal(a.download_f(), b.download_f())
    # cached per (instantiation, DoF count): a second context gets the same form
    c = _solver("D3Q27", 3, N, v)
    assert c.autotune(5) == shape
    # FP64 contexts have one form; the DoF-block, wave and plane figures still come back
    d = _solver("D3Q19", 3, [3, 3, 3], perturbed_vertices(3, [3] * 3, [0] * 3, [L] * 3, seed=3), precision="fp64")
    assert d.autotune(3) == (128, 4, 2)


def test_general_form_setter_rejects_what_does_not_exist():
    L = 2 * math.pi
    v = perturbed_vertices(3, [3] * 3, [0] * 3, [L] * 3, seed=3)
    g64 = _solver("D3Q19", 3, [3] * 3, v, precision="fp64")
    g64.set_general_form(0)
    g64.set_general_form(-1)
    with pytest.raises(nat.NatError):
        g64.set_general_form(1)  # FP64 has only the six-wave/LDS form
    g32 = _solver("D3Q19", 3, [3] * 3, v)
    with pytest.raises(nat.NatError):
        g32.set_general_form(2)
    cs, s = lbm_parameters(1.0, 0.05)
    st = nat.Solver("D3Q19", 2, [4] * 3, box_vertices(3, [4] * 3, [0] * 3, [L] * 3), scaling=s, viscosity=0.01,
                    precision="fp32", backend="structured")
    with pytest.raises(nat.NatError):
        st.set_general_form(0)


@pytest.mark.parametrize("stencil,p,N", [("D3Q19", 1, [2, 3, 2]), ("D3Q27", 2, [3, 2, 3]), ("D3Q19", 3, [5, 4, 3])])
def test_general_grouped_thread_order_against_oracle(stencil, p, N):
    """The six-wave and FP64 LDS forms give threads the x-interior DoFs of each group of 64
    lattice rows before its x-face DoFs. Edge shapes of that order: nfe0 = 3 (one interior
    DoF per row), row counts that leave a partial last group (12, 35, 130 rows). FP64: every
    population within 1e-12 of the oracle after each of 5 steps; FP32 (six-wave form pinned):
    velocity L2 within 1e-6 after 20 steps."""
    cs, s = lbm_parameters(1.0, 0.05)
    L = 2 * math.pi
    v = perturbed_vertices(3, N, [0] * 3, [L] * 3, amplitude=0.1, seed=21)
    o = Oracle(stencil, p, N, v, cartesian=False, scaling=s, viscosity=0.01, cfl=0.4)
    o.assemble()
    g64 = nat.Solver(stencil, p, N, v, cartesian=False, scaling=s, viscosity=0.01, cfl=0.4, backend="general")
    g32 = nat.Solver(stencil, p, N, v, cartesian=False, scaling=s, viscosity=0.01, cfl=0.4, backend="general",
                     precision="fp32")
    g32.set_general_form(0)
    nfe0 = N[0] * p + 1
    print(f"{stencil} p={p} {N}: nfe0 {nfe0}, {g64.n // nfe0} rows", flush=True)
    rho, u = tgv3d_fields(g64.coords(), cs)
    for g in (g64, g32):
        g.init_equilibrium(rho, u)
        g.collide()
    o.init(rho, u)
    o.collide()
    for step in range(1, 6):
        g64.step(1)
        o.steps(1)
        fo = o.get_f()
        e = float(np.abs(g64.download_f() - fo).max() / np.abs(fo).max())
        assert e < 1e-12, (step, e)
    g32.step(5)
    o.steps(15)
    g32.step(15)
    _, ug, _ = g32.macro()
    _, uo, _, _ = o.macro()
    err = float(np.sqrt(((ug - uo) ** 2).sum() / (uo ** 2).sum()))
    assert err < 1e-6, err
