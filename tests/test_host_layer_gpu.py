"""The C++ host layer (natrium_amd/host/natrium_gpu.hpp) on the GPU against the oracle.

tests/cpp/host_layer_check (built by build()) drives CFDSolver<2>::run (CFDSolver.cpp:877-892),
SemiLagrangian<2>::stream(f_old, f, t) (SemiLagrangian.h:150-161), selectCollision
(CollisionSelection.h:61-67) and RungeKutta5LowStorage::step (RungeKutta5LowStorage.cpp:56-105)
through the C ABI and writes the populations it obtains; here they are compared with the
oracle's operations on the same 2D Taylor-Green state (D2Q9, 8 x 8 cells, p = 4, FP64).
"""
import math
import os
import subprocess

import numpy as np
import pytest

from natrium_amd.problems import box_vertices, lbm_parameters, tgv2d_fields
from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "host_layer_check")
FP64_TOL = 1e-12


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    out = tmp_path_factory.mktemp("host_layer")
    print(f"running {BIN}", flush=True)
    r = subprocess.run([BIN, str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    return out, r.stdout


def _load(out, name, Q, n):
    return np.fromfile(os.path.join(out, name), dtype=np.float64).reshape(Q, n)


def _oracle(N, p, scheme="SL", cartesian=True):
    cs, s = lbm_parameters(1.0, 0.05)
    L = 2 * math.pi
    nu = L / 10.0
    v = box_vertices(2, [N, N], [0, 0], [L, L])
    o = Oracle("D2Q9", p, [N, N], v, cartesian=cartesian, scaling=s, viscosity=nu, cfl=0.4, scheme=scheme)
    o.assemble()
    rho, u = tgv2d_fields(o.coords(), nu, cs, L)
    o.init(rho, u)
    return o


@pytest.mark.gpu
def test_cfdsolver_run_matches_oracle_and_fused_form_is_bitwise(run):
    out, stdout = run
    o = _oracle(8, 4)
    o.collide()
    o.steps(20)
    ref = o.get_f()
    loop = _load(out, "run_loop.f64", 9, o.n)
    fused = _load(out, "run_fused.f64", 9, o.n)
    assert np.array_equal(loop, fused)
    assert _rel(loop, ref) < FP64_TOL
    assert "run fused=0 iterations=20" in stdout and "run fused=1 iterations=20" in stdout
    t = float(stdout.split("run fused=1 iterations=20 time=")[1].split()[0])
    assert abs(t - 20 * o.dt) < 1e-15


@pytest.mark.gpu
def test_semilagrangian_stream_and_select_collision_match_oracle(run):
    out, stdout = run
    o = _oracle(8, 4)
    f0 = o.get_f()
    o.stream()
    o.collide()
    # f_old = f before the stream: the library's initial equilibrium (the same formula as the
    # oracle's, rounded on the device)
    assert _rel(_load(out, "sl_fold.f64", 9, o.n), f0) < 1e-14
    assert _rel(_load(out, "sl_stream.f64", 9, o.n), o.get_f()) < FP64_TOL
    dt = float(stdout.split("stream dt=")[1].split()[0])
    assert abs(dt - o.dt) < 1e-18
    assert "same-vector stream: SemiLagrangianException" in stdout
    assert "zero density: DensityZeroException" in stdout
    assert "compressible sedg: NotImplementedException" in stdout
    assert "host_layer_check done" in stdout


@pytest.mark.gpu
def test_rk5_step_of_sedg_operator_matches_oracle(run):
    out, stdout = run
    o = _oracle(4, 3, scheme="SEDG", cartesian=False)
    for _ in range(3):
        o.stream()  # one low-storage RK5 step of the SEDG operator, no collision
    assert _rel(_load(out, "sedg_rk5.f64", 9, o.n), o.get_f()) < FP64_TOL
    t = float(stdout.split("rk5 t=")[1].split()[0])
    assert abs(t - 3 * o.dt) < 1e-15


def test_host_layer_check_is_built_against_the_library():
    """build() compiles the check program with the host compiler; it resolves the library
    in-tree through its run path."""
    assert os.path.exists(BIN), "run __graft_entry__.build() first"
    r = subprocess.run(["ldd", BIN], capture_output=True, text=True)
    line = [ln for ln in r.stdout.splitlines() if "libnatrium_amd.so" in ln]
    assert line and "not found" not in line[0], r.stdout
