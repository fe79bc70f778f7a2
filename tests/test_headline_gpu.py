"""The headline workload at its full size against the FP64 oracle.

3D Taylor-Green vortex, [0, 2 pi]^3, 16^3 affine hexes, p = 4 (274,625 FE_Q DoFs, 262,144
stored nodes), D3Q27, Re = 400, Ma = 0.1, CFL 0.4, structured backend: the kernel
instantiation the benchmark times (16 x 8 x 8 node tiles, 256 workgroups).

- FP64 populations within 1e-12 relative of the oracle after each of the first 3 steps
  (max |f_gpu - f_oracle| / max |f_oracle|);
- FP32-shifted velocity within 1e-6 relative L2 of the oracle after 100 steps
  (sqrt(sum |u_gpu - u_oracle|^2 / sum |u_oracle|^2));
- the FP32-shifted kinetic energy after 100 steps within 2.5 x that velocity error of the
  oracle's (E_kin is quadratic in u).

The oracle assembles the reference's CSR operator (365 M entries, about 6 GB of host memory,
10-15 s) once for the module.
"""
import math

import numpy as np
import pytest

import natrium_amd as nat
from natrium_amd.problems import box_vertices, lbm_parameters, tgv3d_fields
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N, P = 16, 4
MA, RE = 0.1, 400.0


@pytest.fixture(scope="module")
def c2():
    cs, s = lbm_parameters(1.0, MA)
    L = 2 * math.pi
    v = box_vertices(3, [N] * 3, [0] * 3, [L] * 3)
    kw = dict(cartesian=True, scaling=s, viscosity=1.0 / RE, cfl=0.4)
    o = O.Oracle("D3Q27", P, [N] * 3, v, **kw)
    o.assemble()
    x = o.coords()
    rho, u = tgv3d_fields(x, cs)
    return dict(o=o, v=v, kw=kw, rho=rho, u=u, x=x)


def _gpu(c2, precision):
    g = nat.Solver("D3Q27", P, [N] * 3, c2["v"], precision=precision, backend="structured", **c2["kw"])
    assert g.backend_name == "structured"
    assert g.n == (N * P + 1) ** 3 and g.n_stored == (N * P) ** 3
    np.testing.assert_array_equal(g.coords(), c2["x"])
    g.init_equilibrium(c2["rho"], c2["u"])
    g.collide()
    return g


def _reset_oracle(c2):
    o = c2["o"]
    o.init(c2["rho"], c2["u"])
    o.collide()
    return o


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_headline_tile_dispatched(c2, precision):
    g = _gpu(c2, precision)
    assert g.struct_tile() == (16, 8, 8)


def test_headline_fp64_per_step(c2):
    g = _gpu(c2, "fp64")
    o = _reset_oracle(c2)
    fo = o.get_f()
    assert float(np.abs(g.download_f() - fo).max() / np.abs(fo).max()) < 1e-12
    for k in range(1, 4):
        g.step(1)
        o.steps(1)
        fo = o.get_f()
        e = float(np.abs(g.download_f() - fo).max() / np.abs(fo).max())
        print(f"C2 FP64 step {k}: max rel diff vs oracle {e:.3e}", flush=True)
        assert e < 1e-12, f"step {k}: {e}"


def test_headline_fp32_velocity_100_steps(c2):
    g = _gpu(c2, "fp32")
    o = _reset_oracle(c2)
    _, e0 = o.sums()
    g.step(100)
    o.steps(100)
    _, ug, _ = g.macro()
    _, uo, _, _ = o.macro()
    l2 = float(np.sqrt(((ug - uo) ** 2).sum() / (uo ** 2).sum()))
    print(f"C2 FP32-shifted velocity rel L2 vs oracle after 100 steps: {l2:.3e}", flush=True)
    assert l2 < 1e-6
    _, _, eg = g.sums()
    _, e1 = o.sums()
    # E_kin is quadratic in u: |sum(ug^2 - uo^2)| <= ||ug - uo|| ||ug + uo|| ~ 2 l2 ||uo||^2
    # (Cauchy-Schwarz), so its relative error is bounded by about twice the velocity's
    # relative L2 error; 2.5 l2 leaves room for the density weight (FP32 rho, |drho| < 1%).
    print(f"C2 FP32-shifted E_kin rel diff vs oracle after 100 steps: {abs(eg / e1 - 1):.3e}", flush=True)
    assert abs(eg / e1 - 1) <= 2.5 * l2
