"""GPU tests of the request server (continuous batching, no reference counterpart).

A request's result must not depend on the other requests in flight: the same
request run alone and inside a mix of concurrently resident requests gives
bit-identical velocity fields (every request owns its context, packed launches run
each member's tiles as its own launch would, and the step kernels are
deterministic). Graph replays of captured step chunks must match plain launches bit
for bit.
"""
import numpy as np
import pytest

import natrium_amd as nat

pytestmark = pytest.mark.gpu

PROBE = dict(problem="TGV2D", stencil="D2Q9", refinement=3, order=4, reynolds=10.0, mach=0.05, nsteps=120)


def _run_alone(req, use_graphs=True, precision="fp64"):
    srv = nat.Server(max_active=4, chunk_steps=20, use_graphs=use_graphs)
    try:
        rid = srv.submit(precision=precision, **req)
        n, D = nat.request_dofs(req["problem"], req["stencil"], req["refinement"], req["order"])
        return srv.wait(rid, want_u=True, n_dofs=n, dim=D)
    finally:
        srv.close()


def test_server_single_request_matches_direct_solver():
    import math

    from natrium_amd.problems import box_vertices, lbm_parameters, tgv2d_fields

    out = _run_alone(PROBE)
    assert out["status"] == 0
    cs, s = lbm_parameters(1.0, PROBE["mach"])
    L = 2 * math.pi
    N = 2 ** PROBE["refinement"]
    sol = nat.Solver("D2Q9", PROBE["order"], [N, N], box_vertices(2, [N, N], [0, 0], [L, L]), scaling=s,
                     viscosity=L / PROBE["reynolds"], cfl=0.4)
    rho, u = tgv2d_fields(sol.coords(), L / PROBE["reynolds"], cs, L)
    sol.init_equilibrium(rho, u)
    sol.collide()
    sol.step(PROBE["nsteps"])
    _, ud, _ = sol.macro()
    # initial fields come from C++ vs numpy trigonometry: agreement at rounding level
    assert np.abs(out["u"] - ud).max() < 1e-12 * np.abs(ud).max()
    m, _, e = sol.sums()
    assert abs(out["mass"] - m) < 1e-12 * m
    assert abs(out["ekin"] - e) < 1e-10 * e


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_server_result_same_alone_and_in_mix(precision):
    alone = _run_alone(PROBE, precision=precision)
    assert alone["status"] == 0, alone["error"]
    mix = [
        dict(problem="TGV3D", stencil="D3Q27", refinement=3, order=2, reynolds=400.0, mach=0.1, nsteps=60),
        dict(problem="SHOCK_TUBE", stencil="D2Q9", refinement=5, order=3, reynolds=100.0, mach=0.1, nsteps=90),
        dict(problem="TGV3D", stencil="D3Q19", refinement=3, order=3, reynolds=100.0, mach=0.05, nsteps=40),
        dict(problem="TGV2D", stencil="D2Q9", refinement=4, order=6, reynolds=1600.0, mach=0.2, nsteps=70),
        dict(problem="TGV2D", stencil="D2Q9", refinement=3, order=2, reynolds=50.0, mach=0.1, nsteps=30),
    ]
    srv = nat.Server(max_active=3, chunk_steps=20, use_graphs=True)
    try:
        ids = [srv.submit(precision=precision, **r) for r in mix[:3]]
        pid = srv.submit(precision=precision, **PROBE)
        ids += [srv.submit(precision=precision, **r) for r in mix[3:]]
        n, D = nat.request_dofs(PROBE["problem"], PROBE["stencil"], PROBE["refinement"], PROBE["order"])
        inmix = srv.wait(pid, want_u=True, n_dofs=n, dim=D)
        others = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    assert inmix["status"] == 0, inmix["error"]
    assert all(o["status"] == 0 for o in others), [(o["status"], o["error"]) for o in others]
    np.testing.assert_array_equal(inmix["u"], alone["u"])
    assert inmix["mass"] == alone["mass"] and inmix["ekin"] == alone["ekin"]  # fixed-order reductions
    assert st["used_bytes"] == 0 and st["queued"] == 0 and st["active"] == 0


def test_server_graph_replay_matches_plain_launches():
    a = _run_alone(PROBE, use_graphs=True)
    b = _run_alone(PROBE, use_graphs=False)
    assert a["status"] == b["status"] == 0
    np.testing.assert_array_equal(a["u"], b["u"])


def test_server_admission_under_small_budget():
    reqs = [dict(problem="TGV2D", stencil="D2Q9", refinement=4, order=4, reynolds=100.0, mach=0.1, nsteps=40)
            for _ in range(6)]
    one = nat.Server(max_active=1, chunk_steps=20)
    try:
        r0 = one.wait(one.submit(**reqs[0]))
    finally:
        one.close()
    # budget for two resident requests: all six still complete, at most two at a time
    srv = nat.Server(hbm_budget_bytes=int(2.5 * r0["footprint_bytes"]), max_active=8, chunk_steps=20)
    try:
        ids = [srv.submit(**r) for r in reqs]
        outs = [srv.wait(i) for i in ids]
    finally:
        srv.close()
    assert all(o["status"] == 0 for o in outs)
    assert all(abs(o["mass"] - r0["mass"]) <= 1e-14 * r0["mass"] for o in outs)
    # later requests queued behind the budget
    assert max(o["queue_ms"] for o in outs) > min(o["queue_ms"] for o in outs)


def _c5_slice():
    """A seeded slice of the C5 serving workload drawn from its full ranges (refinement 3-5,
    p in {2, 3, 4, 6}, D2Q9 / D3Q19 / D3Q27, Re log-uniform [10, 1600], Ma [0.05, 0.2], steps
    log-uniform [100, 5000]): the draws of seed 2024 whose node updates (DoFs x steps) stay
    within 1.5e8 each, 56 of the first 64 (about 1e9 node updates in total)."""
    from natrium_amd.problems import serving_mix

    out = []
    for r in serving_mix(64, 2024, full=True):
        n, D = nat.request_dofs(r["problem"], r["stencil"], r["refinement"], r["order"])
        if n * r["nsteps"] <= 1.5e8:
            out.append((r, n, D))
    return out


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_server_c5_slice_same_alone_and_in_mix(precision):
    sl = _c5_slice()
    reqs = [r for r, _, _ in sl]
    # the slice reaches the ends of the ranges
    assert max(r["refinement"] for r in reqs) == 5 and max(r["nsteps"] for r in reqs) > 2000
    assert {r["order"] for r in reqs} == {2, 3, 4, 6}
    assert {r["problem"] for r in reqs} == {"TGV2D", "TGV3D", "SHOCK_TUBE"}
    assert {r["stencil"] for r in reqs} == {"D2Q9", "D3Q19", "D3Q27"}
    # probes: the first request of each problem kind at refinement 5, and the first p = 6 one
    probes = []
    for kind in ("TGV2D", "TGV3D", "SHOCK_TUBE"):
        k = next((i for i, r in enumerate(reqs) if r["problem"] == kind and r["refinement"] == 5), None)
        if k is None:
            k = next(i for i, r in enumerate(reqs) if r["problem"] == kind)
        probes.append(k)
    probes.append(next(i for i, r in enumerate(reqs) if r["order"] == 6))
    probes = sorted(set(probes))
    srv = nat.Server(max_active=8, chunk_steps=50, use_graphs=True)
    try:
        ids = [srv.submit(precision=precision, **r) for r in reqs]
        outs = []
        for k, i in enumerate(ids):
            _, n, D = sl[k]
            outs.append(srv.wait(i, want_u=k in probes, n_dofs=n, dim=D))
        st = srv.stats()
    finally:
        srv.close()
    bad = [(reqs[k], o["status"], o["error"]) for k, o in enumerate(outs) if o["status"] != 0]
    assert not bad, bad
    assert st["used_bytes"] == 0 and st["queued"] == 0 and st["active"] == 0
    print(f"C5 slice {precision}: {st['packed_chunks']} packed chunks", flush=True)
    assert st["packed_chunks"] > 0  # structured requests of one instantiation advanced together
    for k in probes:
        alone = _run_alone(reqs[k], precision=precision)
        assert alone["status"] == 0
        np.testing.assert_array_equal(outs[k]["u"], alone["u"], err_msg=str(reqs[k]))
        assert outs[k]["mass"] == alone["mass"] and outs[k]["ekin"] == alone["ekin"]
