"""Output and restart formats through the C ABI: results-table rows, VTU/PVTU files, checkpoints.

Table quantities (solver/SolverStats.h:102-175, solver/PhysicalProperties.cpp:29-420) are checked
against closed forms of the Taylor-Green fields they are evaluated on: in 2D the enstrophy is
int (dv/dx - du/dy)^2 = int (2 sin x sin y)^2 = 4 pi^2 and its square 9 pi^2 over [0, 2 pi]^2; in
3D the integral of |grad u|_F^2 is 6 pi^3 over [0, 2 pi]^3 (relative tolerance 1e-5: p = 4
interpolation and GLL quadrature). The VTU patches must reproduce fields the FE space holds
exactly (polynomials of degree <= p per axis) at every patch point to 1e-12. Checkpoints must
restart bit-identically.
"""
import math
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import natrium_amd as nat
from natrium_amd.problems import box_vertices, lbm_parameters, perturbed_vertices, tgv2d_fields, tgv3d_fields

pytestmark = pytest.mark.gpu


def _tgv2d(precision="fp64", backend="auto", N=8, p=4):
    cs, s = lbm_parameters(1.0, 0.05)
    L = 2 * math.pi
    v = box_vertices(2, [N, N], [0, 0], [L, L])
    g = nat.Solver("D2Q9", p, [N, N], v, scaling=s, viscosity=0.01, cfl=0.4, precision=precision, backend=backend)
    rho, u = tgv2d_fields(g.coords(), 0.01, cs, L)
    g.init_equilibrium(rho, u)
    g.collide()
    return g, cs


@pytest.mark.parametrize("backend", ["structured", "general"])
def test_stats_tgv2d_closed_forms(backend):
    g, cs = _tgv2d(backend=backend)
    st = g.stats(0)
    assert st["iteration"] == 0 and st["time"] == 0.0
    assert abs(st["enstrophy"] / (4 * math.pi ** 2) - 1) < 1e-5
    assert abs(st["enstrophy_sq"] / (9 * math.pi ** 2) - 1) < 1e-5
    assert abs(st["max_u"] - 1.0) < 1e-12
    # rho = 1 + (cos 2x + cos 2y) / (4 cs^2): (rho - 1) cs^2 spans [-1/2, 1/2]
    assert abs(st["max_p"] - 0.5) < 1e-10 and abs(st["min_p"] + 0.5) < 1e-10
    assert abs(st["mass"] / (2 * math.pi) ** 2 - 1) < 1e-12
    m, _, e = g.sums()  # the device reduction order varies between calls
    assert abs(st["mass"] - m) <= 1e-14 * m and abs(st["kin_e"] - e) <= 1e-14 * e
    assert abs(st["mean_ux"]) < 1e-12


def test_stats_tgv3d_gradient_norm():
    cs, s = lbm_parameters(1.0, 0.1)
    L = 2 * math.pi
    v = box_vertices(3, [8, 8, 8], [0] * 3, [L] * 3)
    g = nat.Solver("D3Q27", 4, [8, 8, 8], v, scaling=s, viscosity=0.01, cfl=0.4, precision="fp32")
    rho, u = tgv3d_fields(g.coords(), cs)
    g.init_equilibrium(rho, u)
    g.collide()
    st = g.stats(0)
    assert abs(st["enstrophy"] / (6 * math.pi ** 3) - 1) < 1e-5


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_stats_entropy_mean_velocity_on_perturbed_mesh(precision):
    """f = w rho at rest with rho = 2: entropy = -(volume) (p+1)^D 2 ln 2; a uniform velocity is
    its own volume-weighted mean."""
    N, p = [3, 4], 3
    v = perturbed_vertices(2, N, [0, 0], [1.0, 1.5], amplitude=0.1, seed=1234)
    g = nat.Solver("D2Q9", p, N, v, cartesian=False, scaling=1.0, viscosity=0.01, precision=precision, rho0=2.0)
    g.init_equilibrium(np.full(g.n, 2.0), np.zeros((2, g.n)))
    g.collide()
    st = g.stats(0)
    assert abs(st["entropy"] / (-1.5 * (p + 1) ** 2 * 2 * math.log(2)) - 1) < 1e-12
    u = np.zeros((2, g.n))
    u[0], u[1] = 0.03, -0.01
    g.init_equilibrium(np.ones(g.n), u)
    g.collide()
    # FP32 storage rounds the populations the velocity is recovered from
    assert abs(g.stats(0)["mean_ux"] - 0.03) < (1e-12 if precision == "fp64" else 1e-8)


def test_stats_residuals_and_table_file(tmp_path):
    g, cs = _tgv2d(N=4, p=3)
    path = tmp_path / "results_table.txt"
    g.write_stats(str(path), 0, truncate=True)
    rho0, u0, _ = g.macro()
    g.step(100)
    g.write_stats(str(path), 100)
    rho1, u1, _ = g.macro()
    lines = path.read_text().splitlines()
    assert lines[0].startswith("#") and len(lines) == 3
    cols = [float(c) for c in lines[2].split()]
    assert len(cols) == 13 and cols[0] == 100
    assert abs(cols[1] - g.time) <= 1e-5 * g.time
    # residuals against the fields of iteration 0
    assert abs(cols[9] - np.abs(rho1 - rho0).max()) <= 1e-5 * cols[9]
    assert abs(cols[10] - np.sqrt(((u1 - u0) ** 2).sum(0).max())) <= 1e-5 * cols[10]
    st = g.stats(105)  # not a multiple of 10: the residuals are repeated
    assert abs(st["res_rho"] - cols[9]) <= 1e-5 * cols[9]


def _vtk_array(el):
    """Decode a DataArray written as deal.II's write_vtu writes it with zlib (format="binary",
    compressor vtkZLibDataCompressor): base64 of the UInt32 block header (blocks, block size,
    last block size, compressed sizes), then base64 of the zlib blocks."""
    import base64
    import zlib

    assert el.get("format") == "binary"
    text = el.text.strip()
    hdr = np.frombuffer(base64.b64decode(text[:24]), dtype="<u4")
    assert hdr[0] == 1 and hdr[1] == hdr[2]
    raw = zlib.decompress(base64.b64decode(text[24:]))
    assert len(raw) == hdr[1]
    dt = {"Float64": "<f8", "Int32": "<i4", "UInt8": "u1"}[el.get("type")]
    return np.frombuffer(raw, dtype=dt)


@pytest.mark.parametrize("D,p", [(2, 1), (2, 3), (3, 2)])
def test_vtu_patches_reproduce_polynomial_fields(tmp_path, D, p):
    N = [3, 2, 2][:D]
    L = [1.0, 1.5, 1.2][:D]
    v = perturbed_vertices(D, N, [0] * D, L, amplitude=0.1, seed=1234)
    stencil = "D2Q9" if D == 2 else "D3Q19"
    g = nat.Solver(stencil, p, N, v, cartesian=False, scaling=1.0, viscosity=0.01)
    x = g.coords()

    # degree <= 1 per axis is exact in the FE space on any Q1 cell
    def rho_f(X):
        return 1.0 + 0.01 * X[:, 0] * X[:, 1]

    def ux_f(X):
        return 0.01 * X[:, 0] + 0.02 * X[:, 1]

    u = np.zeros((D, g.n))
    u[0] = ux_f(x)
    g.init_equilibrium(rho_f(x), u)
    g.collide()
    g.write_vtu(str(tmp_path), 7)
    assert os.path.exists(tmp_path / "t_0.7.pvtu")
    root = ET.parse(tmp_path / "t_0.7.vtu").getroot()
    piece = root.find("UnstructuredGrid/Piece")
    sub = 1 if p == 1 else p + 1
    ncell = int(np.prod(N))
    npts = int(piece.get("NumberOfPoints"))
    assert npts == ncell * (sub + 1) ** D
    assert int(piece.get("NumberOfCells")) == ncell * sub ** D
    assert root.get("compressor") == "vtkZLibDataCompressor"
    pts = _vtk_array(piece.find("Points/DataArray")).reshape(-1, 3)
    arrays = {a.get("Name"): _vtk_array(a) for a in piece.findall("PointData/DataArray")}
    conn = _vtk_array(piece.find("Cells/DataArray[@Name='connectivity']"))
    offs = _vtk_array(piece.find("Cells/DataArray[@Name='offsets']"))
    types = _vtk_array(piece.find("Cells/DataArray[@Name='types']"))
    nv = 4 if D == 2 else 8
    assert conn.size == ncell * sub ** D * nv and conn.min() == 0 and conn.max() == npts - 1
    np.testing.assert_array_equal(offs, nv * np.arange(1, ncell * sub ** D + 1))
    assert (types == (9 if D == 2 else 12)).all()
    assert sorted(arrays) == sorted(["rho", "ux", "uy"] + (["uz"] if D == 3 else []))
    np.testing.assert_allclose(arrays["rho"], rho_f(pts), rtol=0, atol=1e-12)
    np.testing.assert_allclose(arrays["ux"], ux_f(pts), rtol=0, atol=1e-12)
    # every patch point lies in the box
    for a in range(D):
        assert pts[:, a].min() >= -1e-12 and pts[:, a].max() <= L[a] + 1e-12


@pytest.mark.parametrize("precision,scheme", [("fp64", "SL"), ("fp32", "SL"), ("fp32", "SEDG")])
def test_checkpoint_restart_is_bit_identical(tmp_path, precision, scheme):
    cs, s = lbm_parameters(1.0, 0.1)
    L = 2 * math.pi
    v = perturbed_vertices(2, [4, 3], [0, 0], [L, L], amplitude=0.1, seed=7)
    kw = dict(cartesian=False, scaling=s, viscosity=0.05, cfl=0.4, precision=precision, scheme=scheme)
    a = nat.Solver("D2Q9", 3, [4, 3], v, **kw)
    rho, u = tgv2d_fields(a.coords(), 0.05, cs, L)
    a.init_equilibrium(rho, u)
    a.collide()
    a.step(10)
    a.checkpoint_write(str(tmp_path), 10)
    a.step(5)
    a.checkpoint_write(str(tmp_path), 15)
    a.step(5)
    stat = (tmp_path / "checkpoint_15.stat").read_text().split()
    assert int(stat[0]) == 15 and int(stat[3]) == 3 and abs(float(stat[2]) - s) < 1e-12 * s
    b = nat.Solver("D2Q9", 3, [4, 3], v, **kw)
    assert b.checkpoint_read(str(tmp_path)) == 15  # the most recent
    assert b.time == pytest.approx(15 * a.dt, rel=1e-14)
    b.step(5)
    assert np.array_equal(a.download_f(), b.download_f())
    c = nat.Solver("D2Q9", 3, [4, 3], v, **kw)
    assert c.checkpoint_read(str(tmp_path), 10) == 10
    c.step(10)
    assert np.array_equal(a.download_f(), c.download_f())
    d = nat.Solver("D2Q9", 2, [4, 3], v, **kw)
    with pytest.raises(nat.NatError, match="FE order"):
        d.checkpoint_read(str(tmp_path), 10)


def test_checkpoint_compressible_f_and_g(tmp_path):
    N, p = [4, 4], 2
    v = box_vertices(2, N, [0, 0], [1, 1])
    kw = dict(scaling=2.0, viscosity=1e-2, cfl=0.4, compressible=True, gamma=1.4, precision="fp32")
    a = nat.Solver("D2Q25H", p, N, v, **kw)
    x = a.coords()
    rho = 1 + 0.1 * np.cos(math.pi * x[:, 0])
    T = 1 + 0.05 * np.sin(math.pi * x[:, 1])
    u = np.zeros((2, a.n))
    u[0] = 0.1 * np.sin(2 * math.pi * x[:, 1])
    a.init_equilibrium(rho, u, T)
    a.collide()
    a.step(4)
    a.checkpoint_write(str(tmp_path), 4)
    assert (tmp_path / "checkpointG_4.stat").exists()
    a.step(3)
    b = nat.Solver("D2Q25H", p, N, v, **kw)
    assert b.checkpoint_read(str(tmp_path), 4) == 4
    b.step(3)
    assert np.array_equal(a.download_f(), b.download_f())
    assert np.array_equal(a.download_g(), b.download_g())
