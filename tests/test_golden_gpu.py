"""GPU paths against the committed golden fixtures (no oracle at run time).

FP64 populations within 1e-12 relative of the fixture (max abs diff / max abs).
"""
import math
import os

import numpy as np
import pytest

import natrium_amd as nat
from natrium_amd.problems import box_vertices, lbm_parameters, perturbed_vertices

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-12


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("backend", ["structured", "general"])
def test_c1_fixture(backend):
    g = np.load(os.path.join(HERE, "c1_tgv2d.npz"))
    cs, s = lbm_parameters(1.0, 0.05)
    L = 2 * math.pi
    sol = nat.Solver("D2Q9", 4, [8, 8], box_vertices(2, [8, 8], [0, 0], [L, L]), scaling=s, viscosity=L / 10.0,
                     cfl=0.4, backend=backend)
    sol.init_equilibrium(g["rho_init"], g["u_init"])
    sol.collide()
    sol.step(1)
    r, u, _ = sol.macro()
    assert _rel(r, g["rho_1"]) < TOL and _rel(u, g["u_1"]) < 1e-10
    sol.step(9)
    assert _rel(sol.download_f(), g["f_10"]) < TOL
    sol.step(90)
    r, u, _ = sol.macro()
    assert _rel(r, g["rho_100"]) < TOL and _rel(u, g["u_100"]) < 1e-10


@pytest.mark.parametrize("backend", ["structured", "general"])
def test_tgv3d_fixture(backend):
    g = np.load(os.path.join(HERE, "tgv3d_2c.npz"))
    cs, s = lbm_parameters(1.0, 0.1)
    L = 2 * math.pi
    sol = nat.Solver("D3Q27", 4, [2, 2, 2], box_vertices(3, [2, 2, 2], [0] * 3, [L] * 3), scaling=s,
                     viscosity=1.0 / 400, cfl=0.4, backend=backend)
    sol.init_equilibrium(g["rho_init"], g["u_init"])
    sol.collide()
    sol.step(5)
    assert _rel(sol.download_f(), g["f_5"]) < TOL


def test_sedg_fixture():
    g = np.load(os.path.join(HERE, "sedg2d.npz"))
    cs, s = lbm_parameters(1.0, 0.05)
    L = 2 * math.pi
    sol = nat.Solver("D2Q9", 3, [3, 3], g["vertices"], cartesian=False, scaling=s, viscosity=L / 10.0, cfl=0.4,
                     scheme="SEDG")
    sol.init_equilibrium(g["rho_init"], g["u_init"])
    sol.collide()
    sol.step(3)
    assert _rel(sol.download_f(), g["f_3"]) < TOL


def test_compressible_fixture():
    g = np.load(os.path.join(HERE, "comp2d.npz"))
    sol = nat.Solver("D2Q25H", 2, [4, 4], box_vertices(2, [4, 4], [0, 0], [1, 1]), scaling=2.0, viscosity=1e-3,
                     cfl=0.4, compressible=True, gamma=1.4)
    sol.init_equilibrium(g["rho_init"], g["u_init"], g["T_init"])
    sol.collide()
    sol.step(3)
    assert _rel(sol.download_f(), g["f_3"]) < TOL
    assert _rel(sol.download_g(), g["g_3"]) < TOL
    _, _, T = sol.macro()
    assert _rel(T, g["T_3"]) < 1e-11
