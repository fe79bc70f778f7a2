#!/usr/bin/env python3
"""Headline benchmark: fused semi-Lagrangian stream + BGK collide, FP32-shifted vs FP64.

Workload (one MI355X): 3D Taylor-Green vortex, [0, 2 pi]^3, 16^3 affine hexes, FE_Q order p = 4,
D3Q27, Re = 400 (nu = 1/400), Ma = 0.1, CFL 0.4. One step = one fused stream + collide over all
nodes, populations resident in HBM. Reports GLUPS (lattice-node updates per second) of the
FP32-shifted path, the FP64 path beside it, the FP32 velocity L2 error vs FP64 after 1000 steps,
the HBM roofline of the fused kernel and the CPU oracle timed on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps 1000] [--warmup 100]
"""
import argparse
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GLUPS FP32-shifted vs FP64 at L2 err≤1e-6 (TGV 3D); p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0


def tgv3d_problem(N=16, p=4, Ma=0.1, Re=400.0):
    from natrium_amd.problems import box_vertices, lbm_parameters

    cs, s = lbm_parameters(1.0, Ma)
    L = 2 * math.pi
    v = box_vertices(3, [N] * 3, [0.0] * 3, [L] * 3)
    return dict(stencil="D3Q27", order=p, ncells=[N] * 3, vertices=v, scaling=s, viscosity=1.0 / Re, cfl=0.4,
                cs=cs)


def make_solver(nat, prob, precision):
    return nat.Solver(prob["stencil"], prob["order"], prob["ncells"], prob["vertices"], cartesian=True,
                      scaling=prob["scaling"], viscosity=prob["viscosity"], cfl=prob["cfl"], precision=precision)


def init_tgv(sol, prob):
    from natrium_amd.problems import tgv3d_fields

    x = sol.coords()
    rho, u = tgv3d_fields(x, prob["cs"])
    sol.init_equilibrium(rho, u)
    sol.collide()


def timed_glups(sol, steps, warmup):
    """warmup steps through the same entry point as the timed ones, then exactly `steps`
    steps between two synchronisations: wall time and the device time of their launches."""
    if warmup:
        sol.step_timed(warmup)
    sol.sync()
    t0 = time.perf_counter()
    ms_dev = sol.step_timed(steps)
    sol.sync()
    wall = time.perf_counter() - t0
    return wall, ms_dev


def read_traffic(workload):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, or None."""
    path = os.path.join(ROOT, "profiles", "traffic_" + workload + ".md")
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def host_cpu():
    """(model name, logical CPUs of the host, threads this process may use)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    n = os.cpu_count() or 1
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = n
    env = os.environ.get("OMP_NUM_THREADS")
    threads = int(env) if env and env.isdigit() else share
    return model, n, max(1, min(threads, share))


def cpu_oracle_run(prob, steps):
    """The oracle (FP64 CPU restatement of the reference) on the same workload and initial
    state: assembly untimed, then `steps` whole stream + collide steps on the host's thread
    share, timed by wall clock. Returns the cpu_baseline record and the oracle's velocity
    after `steps` steps (the FP64 reference of the accuracy check)."""
    from natrium_amd.problems import tgv3d_fields
    from oracle import oracle as O

    model, ncpu, threads = host_cpu()
    O.set_threads(threads)
    o = O.Oracle(prob["stencil"], prob["order"], prob["ncells"], prob["vertices"], cartesian=True,
                 scaling=prob["scaling"], viscosity=prob["viscosity"], cfl=prob["cfl"])
    o.assemble()
    x = o.coords()
    rho, u = tgv3d_fields(x, prob["cs"])
    o.init(rho, u)
    o.collide()
    t0 = time.perf_counter()
    o.steps(steps)
    el = time.perf_counter() - t0
    _, uo, _, _ = o.macro()
    rec = {"value": o.n * steps / el / 1e9, "unit": "GLUPS", "cores": threads, "kind": "port",
           "sample": f"{steps} FP64 steps (CSR SpMV stream + BGK collide) of the same 16^3 p=4 D3Q27 TGV from the "
                     f"same initial state, {o.n} FE_Q DoFs, {el:.1f} s wall, OpenMP {threads} threads "
                     f"(the process's CPU share; host has {ncpu} logical CPUs, {model}); assembly untimed",
           "host_cpu_model": model, "host_logical_cpus": ncpu}
    return rec, uo


def extra_configs(nat, steps_sl=100, steps_err=500, cpu=False):
    """Beside the headline: the other hot-path configurations, FP32-shifted and FP64 GLUPS on
    the device clock (HIP events), and FP32-shifted vs FP64 velocity errors.

    C3: [0,2pi]^3, 24^3 hexes with interior vertices moved by U(-0.1h, 0.1h) (seed 1234), p=3,
        D3Q19, Ma=0.05, nu=0.01 (Re=100 with unit length), seeded divergence-free Fourier
        velocity (8 modes, amplitude 0.05 c_s): general SL backend (389,017 DoFs) and SEDG with
        RK5 (884,736 DG nodes).
    C4: unit square, 128^2 quads, p=2, D2Q25H compressible f/g, Riemann quadrant states
        (seed 7), velocity-neq bounce-back walls at rest on all four sides, errors after 2000 steps.
    Large C2: the same 3D TGV on 48^3 cells (7.08M stored nodes, beyond the MALL), 40 timed steps.
    Reassembly: the C3 mesh moved to a second perturbation (amplitude 0.12 h, seed 77) and the SL
        operator rebuilt (FP64 departure walk through the Q1 map, 389,017 DoFs x 18 directions):
        wall time of nat_set_vertices on the GPU and, with --cpu-assembly, of the oracle's
        assembly of the same operator (about a minute and 7 GB of host memory).
    """
    import numpy as np

    from natrium_amd.problems import (box_vertices, divergence_free_field, lbm_parameters, perturbed_vertices,
                                      riemann_quadrant_fields, riemann_states)

    out = {}
    # the C2 kernel on a working set beyond the 256 MB MALL: 48^3 cells, p = 4 (192^3 = 7.08M nodes,
    # 1.53 GB of FP32 populations per buffer pair, 3.06 GB in FP64); one tile per block as in C2,
    # 27 tiles per CU
    big = tgv3d_problem(N=48)
    res = {}
    for prec in ("fp32", "fp64"):
        sol = make_solver(nat, big, prec)
        init_tgv(sol, big)
        sol.step(4)
        ms = sol.step_timed(40)
        n = sol.n_stored
        res[prec + "_glups"] = round(n * 40 / (ms / 1e3) / 1e9, 4)
        res[prec + "_us_per_step"] = round(ms * 1e3 / 40, 2)
        if prec == "fp32":
            res["hbm_frac_fp32"] = round(n * 2 * 27 * 4 / (ms / 1e3 / 40) / 1e9 / HBM_PEAK_GBS, 4)
        del sol
    res["stored_nodes"] = 192 ** 3
    res["fp32_over_fp64"] = round(res["fp32_glups"] / res["fp64_glups"], 3)
    out["tgv3d_48c_p4_d3q27"] = res

    cs, s = lbm_parameters(1.0, 0.05)
    L = 2 * math.pi
    v3 = perturbed_vertices(3, [24] * 3, [0] * 3, [L] * 3, amplitude=0.1, seed=1234)

    def run(make, init, steps, errsteps):
        res = {}
        sols = {}
        for prec in ("fp32", "fp64"):
            sol = make(prec)
            init(sol)
            sol.step(2)
            ms = sol.step_timed(steps)
            res[prec + "_glups"] = round(sol.n * steps / (ms / 1e3) / 1e9, 4)
            res[prec + "_ms_per_step"] = round(ms / steps, 4)
            sols[prec] = sol
        res["n_dofs"] = sols["fp64"].n
        res["fp32_over_fp64"] = round(res["fp32_glups"] / res["fp64_glups"], 3)
        if errsteps:
            a32, a64 = make("fp32"), make("fp64")
            init(a32)
            init(a64)
            a32.step(errsteps)
            a64.step(errsteps)
            _, u32, T32 = a32.macro()
            _, u64, T64 = a64.macro()
            res["l2_err_u_fp32_vs_fp64"] = float(np.sqrt(((u32 - u64) ** 2).sum() / (u64 ** 2).sum()))
            if T64 is not None:
                res["l2_err_T_fp32_vs_fp64"] = float(np.sqrt(((T32 - T64) ** 2).sum() / (T64 ** 2).sum()))
            res["l2_err_steps"] = errsteps
        return res

    def init3(sol):
        x = sol.coords()
        u = divergence_free_field(x, modes=8, amplitude=0.05 * cs, seed=1234)
        sol.init_equilibrium(np.ones(sol.n), u)
        sol.collide()

    mk_sl = lambda prec: nat.Solver("D3Q19", 3, [24] * 3, v3, cartesian=False, scaling=s, viscosity=0.01, cfl=0.4,
                                    precision=prec)
    out["c3_sl_general"] = run(mk_sl, init3, steps_sl, steps_err)
    v3b = perturbed_vertices(3, [24] * 3, [0] * 3, [L] * 3, amplitude=0.12, seed=77)
    sol = mk_sl("fp32")
    init3(sol)
    sol.sync()
    t0 = time.perf_counter()
    sol.set_vertices(v3b)
    sol.sync()
    out["c3_sl_general"]["reassembly_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    del sol
    if cpu:
        from oracle import oracle as O

        O.set_threads(host_cpu()[2])
        o = O.Oracle("D3Q19", 3, [24] * 3, v3b, cartesian=False, scaling=s, viscosity=0.01, cfl=0.4)
        t0 = time.perf_counter()
        o.assemble()
        out["c3_sl_general"]["oracle_assembly_s"] = round(time.perf_counter() - t0, 2)
        out["c3_sl_general"]["oracle_assembly_threads"] = O.get_threads()
        del o
    mk_dg = lambda prec: nat.Solver("D3Q19", 3, [24] * 3, v3, cartesian=False, scaling=s, viscosity=0.01, cfl=0.4,
                                    precision=prec, scheme="SEDG")
    out["c3_sedg_rk5"] = run(mk_dg, init3, max(steps_sl // 5, 10), steps_err)

    v4 = box_vertices(2, [128, 128], [0, 0], [1, 1])

    def init4(sol):
        rho, u, T = riemann_quadrant_fields(sol.coords(), riemann_states(7))
        sol.init_equilibrium(rho, u, T)
        sol.collide()

    mk_c4 = lambda prec: nat.Solver("D2Q25H", 2, [128, 128], v4, scaling=2.0, viscosity=1e-3, cfl=0.4,
                                    compressible=True, gamma=1.4, precision=prec, periodic=[0, 0])
    out["c4_compressible"] = run(mk_c4, init4, steps_sl, 2000)
    return out


def side_config(nat, name):
    """(make(precision), init(solver), algorithmic FP32 bytes per step) of a side configuration,
    for profiling runs: C3 on the general SL backend or with SEDG + RK5, or C4."""
    from natrium_amd.problems import (box_vertices, divergence_free_field, lbm_parameters, perturbed_vertices,
                                      riemann_quadrant_fields, riemann_states)

    cs, s = lbm_parameters(1.0, 0.05)
    L = 2 * math.pi
    if name in ("c3_sl", "c3_sedg"):
        v3 = perturbed_vertices(3, [24] * 3, [0] * 3, [L] * 3, amplitude=0.1, seed=1234)
        scheme = "SL" if name == "c3_sl" else "SEDG"

        def make(prec):
            return nat.Solver("D3Q19", 3, [24] * 3, v3, cartesian=False, scaling=s, viscosity=0.01, cfl=0.4,
                              precision=prec, scheme=scheme)

        def init(sol):
            u = divergence_free_field(sol.coords(), modes=8, amplitude=0.05 * cs, seed=1234)
            sol.init_equilibrium(np.ones(sol.n), u)
            sol.collide()

        if name == "c3_sl":  # populations read and written once, plus one 16-byte record per (DoF, moving direction)
            return make, init, lambda sol: sol.n * (2 * 19 * 4 + 18 * 16)
        return make, init, lambda sol: sol.n * 5 * 4 * 18 * 4  # five stages: f, df read and written
    v4 = box_vertices(2, [128, 128], [0, 0], [1, 1])

    def make4(prec):
        return nat.Solver("D2Q25H", 2, [128, 128], v4, scaling=2.0, viscosity=1e-3, cfl=0.4, compressible=True,
                          gamma=1.4, precision=prec, periodic=[0, 0])

    def init4(sol):
        rho, u, T = riemann_quadrant_fields(sol.coords(), riemann_states(7))
        sol.init_equilibrium(rho, u, T)
        sol.collide()

    return make4, init4, lambda sol: sol.n * (2 * 2 * 25 * 4 + 8)


def serving_requests(n, seed, full):
    """The seeded serving mix (natrium_amd.problems.serving_mix)."""
    from natrium_amd.problems import serving_mix

    return serving_mix(n, seed, full)


def run_serving(nat, reqs, precision, seed, load=0.8):
    """Closed-loop pass (all submitted at t=0) measures the service capacity; the open-loop pass
    then submits the same requests with Poisson arrivals at `load` x capacity and records each
    request's latency (submit -> retired)."""
    import threading

    import numpy as np

    def submit_all(srv, times):
        ids = [None] * len(reqs)
        res = [None] * len(reqs)

        def waiter():
            for k in range(len(reqs)):
                while ids[k] is None:
                    time.sleep(1e-4)
                res[k] = srv.wait(ids[k])

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for k, r in enumerate(reqs):
            dt = times[k] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids[k] = srv.submit(precision=precision, **r)
        th.join()
        return res, time.perf_counter() - t0

    srv = nat.Server(max_active=8, chunk_steps=50, use_graphs=True)
    try:
        closed, t_closed = submit_all(srv, [0.0] * len(reqs))
        print(f"serving {precision}: closed loop of {len(reqs)} requests in {t_closed:.2f} s", file=sys.stderr,
              flush=True)
        st = srv.stats()
    finally:
        srv.close()
    nodes = sum(r["nsteps"] * o["n_dofs"] for r, o in zip(reqs, closed))
    rate = len(reqs) / t_closed * load
    rng = np.random.default_rng(seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(reqs)))
    srv = nat.Server(max_active=8, chunk_steps=50, use_graphs=True)
    try:
        opened, t_open = submit_all(srv, list(arrivals))
        print(f"serving {precision}: open loop in {t_open:.2f} s", file=sys.stderr, flush=True)
    finally:
        srv.close()
    lat = np.array([o["latency_ms"] for o in opened])
    bad = sum(1 for o in closed + opened if o["status"] != 0)
    return {"requests": len(reqs), "precision": precision, "failed": bad,
            "closed_loop_s": round(t_closed, 3), "closed_loop_glups": round(nodes / t_closed / 1e9, 4),
            "arrival_rate_per_s": round(rate, 3), "open_loop_s": round(t_open, 3),
            "open_loop_glups": round(nodes / t_open / 1e9, 4),
            "latency_p50_ms": round(float(np.percentile(lat, 50)), 3),
            "latency_p99_ms": round(float(np.percentile(lat, 99)), 3),
            "node_updates": int(nodes), "server_busy_ms": round(st["busy_ms"], 1),
            "packed_chunks_closed_loop": int(st["packed_chunks"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--error-steps", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-assembly", action="store_true",
                    help="also time the oracle's assembly of the moved C3 operator on the host")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--serve-full", action="store_true", help="512 requests, full refinement and step ranges")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the C3/C4 side measurements")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="profiling run: only the FP32-shifted fused kernel of --profile-config for this many steps")
    ap.add_argument("--profile-config", default="c2", choices=["c2", "c3_sl", "c3_sedg", "c4"],
                    help="workload of --profile-steps: C2 (headline), C3 general SL or SEDG, C4 compressible")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="profiling run: a plain device copy of 1 GiB, ten times (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the hot path runs on a single MI355X; "
              "multi-GPU sharding is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch

        n = 1 << 28  # 2^28 floats = 1 GiB
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_launch": 4 * n, "copy_bytes_written_per_launch": 4 * n,
                          "launches": 10}))
        return

    import natrium_amd as nat

    prob = tgv3d_problem()
    if args.profile_steps:
        if args.profile_config == "c2":
            s32 = make_solver(nat, prob, "fp32")
            init_tgv(s32, prob)
            s32.step(args.profile_steps)
            s32.sync()
            print(json.dumps({"profile_steps": args.profile_steps, "nodes": s32.n_stored,
                              "algorithmic_bytes_per_launch": s32.n_stored * 2 * 27 * 4}))
        else:
            make, init, nbytes = side_config(nat, args.profile_config)
            sol = make("fp32")
            init(sol)
            sol.step(args.profile_steps)
            sol.sync()
            print(json.dumps({"profile_config": args.profile_config, "profile_steps": args.profile_steps,
                              "dofs": sol.n, "algorithmic_bytes_per_step": nbytes(sol)}))
        return
    Q = 27
    workload = "tgv3d_16c_p4_d3q27"

    # both solves set up first (mesh, operator, initial state), then timed one after the other:
    # FP32-shifted (the headline), FP64 beside it
    s32 = make_solver(nat, prob, "fp32")
    assert s32.backend_name == "structured"
    n_unique = s32.n_stored
    n_dofs = s32.n
    init_tgv(s32, prob)
    s64 = make_solver(nat, prob, "fp64")
    init_tgv(s64, prob)
    wall32, ms32 = timed_glups(s32, args.steps, args.warmup)
    glups32 = n_unique * args.steps / wall32 / 1e9
    wall64, ms64 = timed_glups(s64, args.steps, args.warmup)
    glups64 = n_unique * args.steps / wall64 / 1e9

    # accuracy: FP32-shifted and FP64 GPU runs of error-steps steps from the same initial state,
    # against each other and against the FP64 oracle run of the cpu_baseline leg
    import numpy as np

    a32 = make_solver(nat, prob, "fp32")
    a64 = make_solver(nat, prob, "fp64")
    init_tgv(a32, prob)
    init_tgv(a64, prob)
    _, _, e0 = a64.sums()
    a32.step(args.error_steps)
    a64.step(args.error_steps)
    _, u32, _ = a32.macro()
    _, u64, _ = a64.macro()
    rl2 = lambda a, b: float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
    l2 = rl2(u32, u64)
    _, _, e1 = a64.sums()
    t_end = a64.time
    del a32, a64

    # roofline of the fused kernel: every population read once and written once per launch
    bytes_per_launch = n_unique * 2 * Q * 4
    avg_launch_s = ms32 / 1e3 / args.steps
    achieved = bytes_per_launch / avg_launch_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(workload),
                "kernel": "k_sl_struct<3,27,4,float>", "bytes_per_launch": bytes_per_launch,
                "avg_launch_us": round(avg_launch_s * 1e6, 3)}

    cpu, l2_oracle = None, None
    if not args.no_cpu:
        cpu, u_oracle = cpu_oracle_run(prob, args.error_steps)
        l2_oracle = {"fp32": rl2(u32, u_oracle), "fp64": rl2(u64, u_oracle)}

    extra = None if args.no_extra else extra_configs(nat, cpu=args.cpu_assembly)

    serving = None
    if not args.no_serve:
        nreq = 512 if args.serve_full else args.serve_requests
        reqs = serving_requests(nreq, 2024, args.serve_full)
        serving = {pr: run_serving(nat, reqs, pr, 2024) for pr in ("fp32", "fp64")}

    out = {
        "metric": METRIC,
        "value": round(glups32, 4),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall32 * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "storage_dtype": "f32 (populations shifted by w rho0)",
        "data": "synthetic",
        "config": {"workload": workload, "mesh": "[0,2pi]^3 16^3 affine hexes", "order": 4, "stencil": "D3Q27",
                   "Re": 400, "Ma": 0.1, "cfl": 0.4, "nodes_updated_per_step": n_unique,
                   "reference_fe_q_dofs": n_dofs, "parallelism": "single GPU"},
        "fp64_glups": round(glups64, 4),
        "fp32_over_fp64": round(glups32 / glups64, 3),
        "glups_reference_dof_convention": round(n_dofs * args.steps / wall32 / 1e9, 4),
        "fp64_ms_per_step": round(wall64 * 1e3 / args.steps, 5),
        "l2_err_u_fp32_vs_oracle": l2_oracle["fp32"] if l2_oracle else None,
        "l2_err_u_fp64_vs_oracle": l2_oracle["fp64"] if l2_oracle else None,
        "l2_err_u_fp32_vs_fp64": l2,
        "l2_err_steps": args.error_steps,
        "ekin_ratio_fp64": e1 / e0,
        "ekin_ratio_stokes": math.exp(-6 * prob["viscosity"] * t_end),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "latency_p50_ms": serving["fp32"]["latency_p50_ms"] if serving else None,
        "latency_p99_ms": serving["fp32"]["latency_p99_ms"] if serving else None,
        "serving": serving,
        "other_configs": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
