"""Converged N-1 Newton-Raphson solves per second on the 2,000-bus synthetic grid, one MI355X.

Workload: the 2,000-bus / 2,700-branch synthetic grid (1 slack, 300 PV), base case solved from flat start, then every
non-bridge single-branch outage solved from the base-case voltages in one batched launch: max 10 Newton iterations,
tolerance 1e-8 pu, FP64 mismatch and update, the LU arithmetic (FP32 with FP64 refinement, or FP64) and launch shape
chosen by the autotuner on the sweep itself under the |dV| <= 1e-9 bar. One step = one full N-1 sweep.

Printed as one JSON line: throughput (converged solves/s), the sweep latency, the plain FP64 one-instance-per-launch
path on a sample (the baseline the batched engine is measured against), the roofline of the batched kernel, and the
CPU oracle (a C++ restatement of the reference's CPU sparse NR) timed on this host's cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(plan: dict, n_bus: int, nnz_y: int, iterations: np.ndarray, refine: int, lu_bytes: int) -> int:
    """Bytes one batched launch must move at minimum: per Newton step and instance, the mismatch (Y values and
    pattern, state), the Jacobian fill, the LU refactor (A values in, L/U indices and values), the triangular solves
    and refinement residuals, and the update; plus the final mismatch evaluation of every instance."""
    m = len(plan["q"])
    nj = len(plan["j_colidx"])
    nlu = len(plan["li"]) + len(plan["ui"])
    mis = 20 * nnz_y + 44 * n_bus
    jac = 16 * nnz_y + 8 * nj
    lu = 8 * nj + (4 + lu_bytes) * nlu
    solve = (4 + lu_bytes) * nlu + 12 * m
    refine_b = 12 * nj + 24 * m + solve
    upd = 8 * m + 32 * n_bus
    per_step = mis + jac + lu + solve + refine * refine_b + upd
    steps = int(iterations.sum())
    return per_step * steps + mis * len(iterations)


def measured_traffic(stats: dict, lu: str, flags: int, alg: int, kernel_ms: float):
    """HBM bytes of one launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this same workload, when
    the plan, arithmetic and iteration counts here are the ones profiled and this run's kernel time lies within 10 % of
    the profiled one (a kernel changed since would move it); otherwise None (no stale figure)."""
    path = os.path.join(ROOT, "profiles", "nr_batch_kernel_n1_2000bus_traffic.json")
    if not os.path.exists(path) or flags != 0:
        return None, None
    with open(path) as fh:
        rec = json.load(fh)
    fp = rec["fingerprint"]
    mine = {k: stats.get(k) for k in ("m", "nnz_j", "nnz_l", "nnz_u", "refactor_rows", "forward_steps", "back_steps",
                                      "factor_slots", "jacobian_transposed")}
    mine.update(lu=lu, algorithmic_bytes=int(alg))
    if any(fp.get(k) != v for k, v in mine.items()):
        return None, None
    if not abs(kernel_ms / rec["kernel_ms"] - 1.0) <= 0.10:
        return None, None
    return int(rec["read_bytes"] + rec["write_bytes"]), rec["source"]


def _schedule_name(opts):
    from gridsim_amd import engine
    if opts.schedule == engine.SCHEDULE_LEVELS:
        return "levels"
    if opts.schedule == engine.SCHEDULE_GROUP:
        w = (opts.flags >> 8) & 15
        return f"work-group of {w or 8} waves per instance"
    return "one wavefront per instance" + (f", launches of {opts.schedule}" if opts.schedule > 0 else "")


def run_seam(args):
    """Per-call latency of the sparse-solve seam that replaces the reference's GPU linear solver
    (ISparseLinearSolver.Solve, src/GridSim.Core/Numerics/ISparseLinearSolver.cs:12-16; the backend it replaces
    allocates and frees its device buffers on every call, RocmSparseLinearSolver.cs:120-140): gs_sparse_solve (the
    plan cached per pattern, the buffers resident), gs_sparse_solve_noalloc (the same in a static device array) and
    gs_sparse_solve_batch (64 systems of one pattern per call), on the Jacobian patterns of MATPOWER case300 and the
    2,000-bus grid at their flat starts. The values are synthetic on those patterns: off-diagonal entries N(0, 1),
    each diagonal entry its row's or its column's absolute sum, the larger, plus one, so the reference's pivot search
    keeps the diagonal and a cached plan replays for new values (the first call's time is the fresh factorisation). One
    call = one host-to-device upload of the values and the right-hand side, the factorisation and both solves on the
    device, and the solution back. CPU baseline: the oracle's sparse LU (ordering, factorisation and solves, as the
    reference's CpuSparseLuSolver does per call) on the same systems, one thread."""
    import ctypes as C
    from gridsim_amd import engine, gridgen
    from gridsim_amd.matpower import read_case
    ctx = engine.Context(0)
    info = ctx.device_info()
    rng = np.random.default_rng(gridgen.BASE_SEED + 17)
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "matpower", "case300.m")
    out_rows = []
    lib = engine.lib()
    calls = max(50, 20 * args.steps)
    for name, net in (("case300", read_case(golden)), ("2,000-bus grid", gridgen.config2_grid())):
        plan = engine.HostPlan(net, engine.NrOptions(flat_start=True))
        rp, ci = plan.j_rowptr.astype(np.int32), plan.j_colidx.astype(np.int32)
        n = len(rp) - 1

        rows = np.repeat(np.arange(n), np.diff(rp))
        diag = np.nonzero(rows == ci)[0]

        def values():
            v = rng.standard_normal(len(ci))
            v[diag] = 0.0
            rsum, csum = np.zeros(n), np.zeros(n)
            np.add.at(rsum, rows, np.abs(v))
            np.add.at(csum, ci, np.abs(v))
            v[diag] = np.maximum(rsum, csum) + 1.0   # dominant by rows and by columns: no pivot moves
            return v

        v = values()
        b = rng.standard_normal(n)
        x = np.empty(n)
        solver = engine.HipSparseLinearSolver(ctx)
        t0 = time.perf_counter()
        x0 = solver.solve(rp, ci, v, b)
        first_ms = 1000 * (time.perf_counter() - t0)
        assert x0 is not None
        resid = float(np.abs(_csr_mul(rp, ci, v, x0) - b).max())

        def timed(fn):
            out = []
            for _ in range(calls):
                ts = time.perf_counter()
                fn()
                out.append(1e6 * (time.perf_counter() - ts))
            return out

        args_c = (C.c_int32(n), engine._p(rp), engine._p(ci), engine._p(v), engine._p(b), engine._p(x))
        lat = timed(lambda: lib.gs_sparse_solve(ctx.h, *args_c))
        if lib.gs_sparse_solve_noalloc(ctx.h, *args_c) != engine.GS_OK:
            raise RuntimeError("gs_sparse_solve_noalloc: " + ctx.error())
        lat_na = timed(lambda: lib.gs_sparse_solve_noalloc(ctx.h, *args_c))
        batch = engine.HipBatchedSparseLinearSolver(ctx)
        mats = [(rp, ci, values()) for _ in range(64)]
        rhs = [rng.standard_normal(n) for _ in range(64)]
        errs = [None] * 64
        batch.solve_batch(mats, rhs, errs)
        tb = time.perf_counter()
        reps = max(3, args.steps)
        for _ in range(reps):
            batch.solve_batch(mats, rhs, errs)
        batch_us = 1e6 * (time.perf_counter() - tb) / (reps * 64)
        row = {"system": f"{name} Jacobian pattern", "n": n, "nnz": int(len(ci)), "first_call_ms": round(first_ms, 3),
               "p50_us": round(float(np.percentile(lat, 50)), 1), "p99_us": round(float(np.percentile(lat, 99)), 1),
               "noalloc_p50_us": round(float(np.percentile(lat_na, 50)), 1),
               "noalloc_p99_us": round(float(np.percentile(lat_na, 99)), 1),
               "batch64_us_per_system": round(batch_us, 1), "max_abs_residual": resid, "calls": calls}
        if not args.no_cpu:
            from oracle import oracle as O
            cpu = []
            for _ in range(max(5, calls // 10)):
                ts = time.perf_counter()
                O.sparse_solve(rp, ci, v, b)
                cpu.append(1e6 * (time.perf_counter() - ts))
            row["cpu_oracle_p50_us"] = round(float(np.percentile(cpu, 50)), 1)
        print(f"# seam {row}", file=sys.stderr, flush=True)
        out_rows.append(row)
    head = out_rows[-1]
    line = {"metric": "sparse-solve seam: per-call latency of gs_sparse_solve (ISparseLinearSolver.Solve), 1 MI355X",
            "value": head["p50_us"], "unit": "us per call (p50)", "n_gpus": 1, "steps": calls, "warmup": 1,
            "ms_per_step": round(head["p50_us"] / 1000, 4), "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic values on real Jacobian patterns",
            "config": {"workload": "one linear system per call on the Jacobian pattern of the 2,000-bus grid "
                                   "(case300 beside it)", "systems": out_rows, "device": info["name"] or info["arch"]},
            "cpu_baseline": ({"value": head.get("cpu_oracle_p50_us"), "unit": "us per call (p50)", "cores": 1,
                              "kind": "port", "sample": "the same system, ordering + factorisation + solves per call "
                              "(C++ restatement of the reference's CpuSparseLuSolver)"} if not args.no_cpu else None)}
    print(json.dumps(line), flush=True)
    ctx.close()


def _csr_mul(rp, ci, v, x):
    y = np.zeros(len(rp) - 1)
    np.add.at(y, np.repeat(np.arange(len(rp) - 1), np.diff(rp)), v * x[ci])
    return y


def run_config3(args):
    """512 instances, a third each on 300-, 2,000- and 10,000-bus grids (branches 1.35 n, 15 % PV), each with a load
    multiplier U[0.8, 1.2] and one random non-bridge branch outage, warm-started from its grid's base case; options
    autotuned per grid (a wavefront or a work-group of waves per instance; the level-parallel schedule offered to the
    largest too); all 512 solved by one replay of a captured hipGraph (each grid's launches on its own branch,
    concurrent: one kernel per grid, or the level-parallel schedule's unrolled Newton steps)."""
    from gridsim_amd import engine, gridgen
    t0 = time.time()
    ctx = engine.Context(0)
    info = ctx.device_info()
    seed = gridgen.BASE_SEED + 3
    rng = np.random.default_rng(seed)
    sizes, counts = (300, 2000, 10000), (171, 171, 170)
    groups, nets, insts = [], [], []
    tuned = []
    for n, count in zip(sizes, counts):
        net = gridgen.sized_grid(n, seed)
        base = engine.DeviceNetwork(ctx, net, engine.NrOptions(flat_start=True, max_iterations=60,
                                                               lu_precision=engine.LU_FP64))
        b = base.solve(count=1)
        assert b.status[0] == engine.CONVERGED, f"{n}-bus base case did not converge"
        base.close()
        warm = net.with_start(b.vm[0], np.degrees(b.va[0]))
        outs = rng.choice(gridgen.config2_outages(warm), count, replace=False)
        scales = rng.uniform(0.8, 1.2, count)
        dn = engine.DeviceNetwork(ctx, warm, engine.NrOptions(max_iterations=10))
        # the level-parallel schedule and a work-group of waves per instance are offered to the largest grid only: a
        # graph runs two groups' level-parallel step chains one after the other (measured), and a wavefront group's
        # single kernel overlaps the chain, where work-groups of many waves would hold the compute units it needs
        offer = engine.SCHEDULE_LEVELS if n == max(sizes) else 0
        opts, ms = engine.autotune(dn, outages=outs, load_scale=scales,
                                   options=engine.NrOptions(max_iterations=10, schedule=offer))
        tuned.append({"n_bus": n, "lu": "fp32" if opts.lu_precision == engine.LU_FP32 else "fp64",
                      "refine_steps": opts.refine_steps,
                      "schedule": _schedule_name(opts),
                      "solve_vector": "global" if opts.flags & engine.FLAG_X_GLOBAL else "lds when it fits",
                      "sample_kernel_ms": round(ms, 3)})
        print(f"# {n} buses: m={len(dn.plan()['q'])}, autotuned {tuned[-1]}", file=sys.stderr, flush=True)
        groups.append(engine.Group(dn, outages=outs, load_scale=scales, options=opts))
        nets.append(warm)
        insts.append((outs, scales))
    batch = engine.Batch(ctx, groups, use_graph=True)
    print(f"# setup {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        batch.run()
    run_ms = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        dms = batch.run()
        run_ms.append(1000 * (time.perf_counter() - ts))
        print(f"# step {k}: {run_ms[-1]:.2f} ms (device {dms:.2f} ms)", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    res = [batch.results(g, keep_state=True) for g in range(len(groups))]
    total = sum(counts)
    converged = int(sum((r.status == engine.CONVERGED).sum() for r in res))
    value = converged * args.steps / elapsed
    # parity with the oracle on a sample of each grid
    match = None
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        ok = True
        for net, (outs, scales), r in zip(nets, insts, res):
            for i in range(0, len(outs), max(1, len(outs) // 4))[:4]:
                rec, vm, va = O.solve(net.with_load_multiplier(float(scales[i])), O.options(max_iterations=10),
                                      skip=int(outs[i]))
                ok &= int(rec["status"]) == int(r.status[i])
                if rec["status"] == 0:
                    ok &= bool(np.abs(r.vm[i] - vm).max() <= 1e-9 and np.abs(r.va[i] - va).max() <= 1e-9)
        match = bool(ok)
        threads = min(os.cpu_count() or 1, 16)
        tc = time.perf_counter()
        cpu_conv = 0
        cpu_n = 0
        for net, (outs, scales) in zip(nets, insts):
            k = 24
            for s in np.unique(np.round(scales[:k], 6)):
                sel = np.nonzero(np.round(scales[:k], 6) == s)[0]
                cres, _, _ = O.screen(net.with_load_multiplier(float(s)), outs[:k][sel], O.options(max_iterations=10),
                                      threads=threads, keep_state=False)
                cpu_conv += int((cres["status"] == 0).sum())
                cpu_n += len(sel)
        csec = time.perf_counter() - tc
        cpu = {"value": round(cpu_conv / csec, 1), "unit": "converged solves/s", "cores": threads, "kind": "port",
               "sample": f"the first 24 instances of each grid ({cpu_n} solves), {threads} threads (C++ restatement "
                         f"of the reference's CPU sparse NR)"}
    line = {
        "metric": "NR solves/s, 512 uneven instances (300 / 2,000 / 10,000 buses) in one hipGraph, 1 MI355X",
        "value": round(value, 1), "unit": "converged solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "autotuned per grid (f32-lu+f64 or f64)", "data": "synthetic",
        "config": {"workload": "512 instances, a third each on 300-, 2,000- and 10,000-bus synthetic grids "
                               "(branches 1.35 n, 15 % PV), load multiplier U[0.8, 1.2], one random non-bridge outage "
                               "each, warm start from each grid's base case, max 10 iterations, 1e-8 pu",
                   "instances": total, "seed": int(seed), "graph": True, "autotuned": tuned,
                   "device": info["name"] or info["arch"]},
        "latency_ms": {"p50": round(float(np.percentile(run_ms, 50)), 3),
                       "p99": round(float(np.percentile(run_ms, 99)), 3),
                       "note": "one replay of the graph solves all 512 instances: per-solve latency is the replay"},
        "verdicts": {"converged": converged, "of": total, "match_cpu_oracle_on_sample": match},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    batch.close()
    ctx.close()


def run_config4(args):
    """One large instance: the 70,000-bus grid from flat start (max 100 iterations, 1e-8 pu), FP64 LU against FP32
    LU with one and two FP64 refinement steps: iterations to 1e-8, device time and wall time of each, and |dV|
    against the oracle's converged state."""
    from gridsim_amd import engine, gridgen
    t0 = time.time()
    ctx = engine.Context(0)
    info = ctx.device_info()
    net = gridgen.config4_grid()
    base = engine.NrOptions(flat_start=True, max_iterations=100, tolerance=1e-8)
    dn = engine.DeviceNetwork(ctx, net, base)
    st = dn.plan_stats()
    print(f"# setup {time.time() - t0:.1f}s: {net.n_bus} buses, {net.n_branch} branches, m={st['m']}, "
          f"nnz(J)={st['nnz_j']}, nnz(L+U)={st['nnz_l'] + st['nnz_u']}, levels {st['levels']}",
          file=sys.stderr, flush=True)
    ref = None
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        tc = time.perf_counter()
        rec, rvm, rva = O.solve(net, O.options(flat_start=True, max_iterations=100))
        csec = time.perf_counter() - tc
        ref = (rec, rvm, rva)
        cpu = {"value": round(1000 * csec, 1), "unit": "ms per solve", "cores": 1, "kind": "port",
               "sample": f"the same solve, single thread (C++ restatement of the reference's CPU sparse NR): "
                         f"{int(rec['iterations'])} iterations, status {int(rec['status'])}"}
        print(f"# oracle: {csec:.2f} s, {int(rec['iterations'])} iterations", file=sys.stderr, flush=True)
    variants = [("fp64", engine.LU_FP64, 0, 0), ("fp64", engine.LU_FP64, 0, engine.SCHEDULE_LEVELS),
                ("fp32+1", engine.LU_FP32, 1, engine.SCHEDULE_LEVELS),
                ("fp32+2", engine.LU_FP32, 2, engine.SCHEDULE_LEVELS)]
    if args.c4_variants:
        keep = set(args.c4_variants.split(","))
        variants = [v for v in variants if f"{v[0]}-{'levels' if v[3] else 'wave'}" in keep]
    rows = []
    for name, lu, refine, sched in variants:
        o = engine.NrOptions(flat_start=True, max_iterations=100, tolerance=1e-8, lu_precision=lu,
                             refine_steps=refine, schedule=sched)
        for _ in range(args.warmup):
            dn.solve(count=1, options=o, keep_state=False)
        walls, kms = [], []
        for _ in range(args.steps):
            ts = time.perf_counter()
            r = dn.solve(count=1, options=o, keep_state=True)
            walls.append(1000 * (time.perf_counter() - ts))
            kms.append(r.kernel_ms)
        row = {"lu": name, "schedule": "level-parallel" if sched == engine.SCHEDULE_LEVELS else "one wavefront",
               "status": engine.STATUS_NAMES[int(r.status[0])], "graph_replays": int(r.launches),
               "iterations": int(r.results["iterations"][0]), "max_mismatch": float(r.results["max_mismatch"][0]),
               "wall_ms": round(float(np.median(walls)), 2), "kernel_ms": round(float(np.median(kms)), 2),
               "phase_us": {n: round(float(v), 1) for n, v in zip(engine.PHASES, r.results["phase_us"][0])}}
        if ref is not None and ref[0]["status"] == 0 and r.status[0] == 0:
            row["max_dv_vs_oracle"] = float(max(np.abs(r.vm[0] - ref[1]).max(), np.abs(r.va[0] - ref[2]).max()))
            row["iterations_oracle"] = int(ref[0]["iterations"])
        rows.append(row)
        print(f"# {json.dumps(row)}", file=sys.stderr, flush=True)
    conv = [r for r in rows if r["status"] == "Converged"]
    best = min(conv or rows, key=lambda r: r["wall_ms"])
    line = {
        "metric": "large single instance: wall time of one 70,000-bus NR solve from flat start to 1e-8 pu, FP64 LU vs "
                  "FP32 LU + FP64 refinement, 1 MI355X",
        "value": best["wall_ms"], "unit": "ms per solve", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": best["wall_ms"], "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64" if best["lu"] == "fp64" else "f32-lu+f64", "data": "synthetic",
        "config": {"workload": "70,000 buses (8,000 PV, 35 regions of 2,000 buses with their own slack), 88,000 "
                               "branches, 12 % transformers, lognormal loads, flat start, max 100 iterations, 1e-8 pu",
                   "n_bus": net.n_bus, "n_branch": net.n_branch, "seed": int(net.seed), "m": st["m"],
                   "nnz_lu": st["nnz_l"] + st["nnz_u"], "levels": st["levels"],
                   "best": f"{best['lu']}, {best['schedule']}",
                   "device": info["name"] or info["arch"]},
        "variants": rows,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    dn.close()
    ctx.close()


def _serve_stream(submit, wait, reqs, arrivals):
    """Open loop: submit request i at arrivals[i] (seconds from the start) from this thread while a collector thread
    waits for the tickets in submission order. Returns (results, wall seconds from the first arrival to the last
    result)."""
    import queue
    import threading
    q = queue.Queue()
    out = [None] * len(reqs)

    def collect():
        for _ in range(len(reqs)):
            i, t = q.get()
            out[i] = wait(t, reqs[i][0])

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    for i, r in enumerate(reqs):
        while True:
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt <= 0:
                break
            time.sleep(min(dt, 0.0005) if dt > 0.0002 else 0)
        q.put((i, submit(*r)))
    th.join()
    return out, time.perf_counter() - t0


def _plain_loop(dns, opts, reqs, arrivals):
    """The one-solve-per-launch loop the server is measured against: requests served one at a time in arrival order
    (gs_nr_solve_batch of one instance each, the network's own options), latency = completion - arrival."""
    lat, conv = [], 0
    t0 = time.perf_counter()
    for (k, o, sc), a in zip(reqs, arrivals):
        now = time.perf_counter() - t0
        if now < a:
            time.sleep(a - now)
        r = dns[k].solve(outages=[o], load_scale=[sc], options=opts[k], keep_state=True)
        done = time.perf_counter() - t0
        lat.append(1000 * (done - a))
        conv += int(r.status[0] == 0)
    wall = time.perf_counter() - t0
    return lat, conv, wall


def run_config5(args):
    """Serving: an open-loop Poisson stream of power-flow and single-contingency requests over 8 resident networks
    (bus counts log-uniform in [100, 25,000]), load multiplier U[0.8, 1.2], served by the continuous-batching loop
    (gs_server_*) at offered rates from 100 to 5,000 requests/s; the same streams through a plain one-solve-per-launch
    loop; the oracle on a sample of requests."""
    from gridsim_amd import engine, gridgen
    t0 = time.time()
    ctx = engine.Context(0)
    info = ctx.device_info()
    seed = gridgen.BASE_SEED + 5
    rng = np.random.default_rng(seed)
    sizes = gridgen.config5_sizes(seed)
    nets, warm, dns, opts, tuned, outs, seeds = [], [], [], [], [], [], []
    for i, n_bus in enumerate(sizes):
        # a grid whose base case does not converge from flat start is redrawn with the next seed (a resident
        # network is served from its solved base case)
        gseed = seed + 1 + i
        while True:
            net = gridgen.sized_grid(n_bus, gseed)
            base = engine.DeviceNetwork(ctx, net, engine.NrOptions(flat_start=True, max_iterations=100,
                                                                   lu_precision=engine.LU_FP64))
            b = base.solve(count=1)
            base.close()
            if b.status[0] == engine.CONVERGED:
                break
            print(f"# {n_bus}-bus grid, seed {gseed}: base case did not converge, redrawn", file=sys.stderr,
                  flush=True)
            gseed += 1000
        nets.append(net)
        seeds.append(gseed)
        w = net.with_start(b.vm[0], np.degrees(b.va[0]))
        dn = engine.DeviceNetwork(ctx, w, engine.NrOptions(max_iterations=10, lu_precision=engine.LU_FP32,
                                                           refine_steps=1))
        o_all = gridgen.config2_outages(w)
        so = np.where(rng.random(16) < 0.5, rng.choice(o_all, 16), -1)
        # (a work-group of waves per instance offered too: the serving loop runs either)
        o, ms = engine.autotune(dn, outages=so, load_scale=rng.uniform(0.8, 1.2, 16),
                                options=dataclasses.replace(dn.options, schedule=engine.SCHEDULE_GROUP))
        tuned.append({"n_bus": w.n_bus, "lu": "fp32" if o.lu_precision == engine.LU_FP32 else "fp64",
                      "refine_steps": o.refine_steps, "schedule": _schedule_name(o)})
        print(f"# {w.n_bus} buses: {tuned[-1]}", file=sys.stderr, flush=True)
        warm.append(w); dns.append(dn); opts.append(o); outs.append(o_all)
    srv = engine.Server(ctx)
    ids = [srv.add_network(dn, o) for dn, o in zip(dns, opts)]
    print(f"# setup {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    def draw(count):
        reqs = []
        for _ in range(count):
            k = int(rng.integers(len(nets)))
            o = int(rng.choice(outs[k])) if rng.random() < 0.5 else -1
            reqs.append((k, o, float(rng.uniform(0.8, 1.2))))
        return reqs

    def submit(k, o, sc):
        return srv.submit(ids[k], o, sc)

    def wait(t, k):
        return srv.wait(t, ids[k], keep_state=False)

    # warm-up: every network once
    _serve_stream(submit, wait, [(k, -1, 1.0) for k in range(len(nets))], np.zeros(len(nets)))
    sweep = []
    rates = [100, 250, 500, 1000, 2000, 5000]
    for rate in rates:
        n_req = int(min(max(200, rate * args.serve_seconds), 4000))
        reqs = draw(n_req)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
        srv.reset_stats()
        res, wall = _serve_stream(submit, wait, reqs, arrivals)
        st = srv.stats()
        conv = sum(int(r[0]["status"] == 0) for r in res)
        pl_n = min(n_req, args.plain_requests)
        lat, pconv, pwall = _plain_loop(dns, opts, reqs[:pl_n], arrivals[:pl_n])
        row = {"offered_rps": rate, "requests": n_req, "solves_per_s": round(conv / wall, 1),
               "p50_ms": round(st["latency_p50_ms"], 3), "p99_ms": round(st["latency_p99_ms"], 3),
               "max_ms": round(st["latency_max_ms"], 3), "mean_batch": round(st["mean_batch"], 2),
               "launches": st["launches"], "joined_running": st["joined"], "converged": conv,
               "plain": {"requests": pl_n, "solves_per_s": round(pconv / pwall, 1),
                         "p50_ms": round(float(np.percentile(lat, 50)), 3),
                         "p99_ms": round(float(np.percentile(lat, 99)), 3)}}
        sweep.append(row)
        print(f"# {rate} req/s: {json.dumps(row)}", file=sys.stderr, flush=True)

    # the oracle on a sample: two requests per network
    match, max_dv, cpu = None, None, None
    if not args.no_cpu:
        from concurrent.futures import ThreadPoolExecutor
        from oracle import oracle as O
        sample = [(k, o, sc) for k in range(len(nets)) for (o, sc) in ((-1, 0.9), (int(outs[k][len(outs[k]) // 2]), 1.1))]
        tickets = [srv.submit(ids[k], o, sc) for k, o, sc in sample]
        got = [srv.wait(t, ids[k]) for t, (k, _, _) in zip(tickets, sample)]
        threads = min(os.cpu_count() or 1, 16)

        def one(r):
            k, o, sc = r
            ts = time.perf_counter()
            rec, vm, va = O.solve(warm[k].with_load_multiplier(sc), O.options(max_iterations=10), skip=o)
            return rec, vm, va, time.perf_counter() - ts

        tc = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            ref = list(ex.map(one, sample))
        csec = time.perf_counter() - tc
        match, max_dv = True, 0.0
        for (rec, vm, va, _), g in zip(ref, got):
            match &= int(rec["status"]) == int(g[0]["status"])
            if rec["status"] == 0:
                max_dv = max(max_dv, float(np.abs(vm - g[1]).max()), float(np.abs(va - g[2]).max()))
        cpu = {"value": round(sum(int(r[0]["status"] == 0) for r in ref) / csec, 1), "unit": "converged solves/s",
               "cores": threads, "kind": "port",
               "sample": f"{len(sample)} requests (a power flow and a contingency per network), {threads} threads, "
                         f"C++ restatement of the reference's CPU sparse NR; single-thread per-request p50 "
                         f"{1000 * float(np.median([r[3] for r in ref])):.1f} ms"}
    top = sweep[-1]
    line = {
        "metric": "serving: converged solves/s and p50/p99 request latency, Poisson stream 100-5,000 req/s, "
                  "8 resident networks (100-25,000 buses), 1 MI355X",
        "value": top["solves_per_s"], "unit": "converged solves/s", "n_gpus": 1, "steps": len(rates), "warmup": 1,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "autotuned per network (f32-lu+f64 or f64)", "data": "synthetic",
        "config": {"workload": "open-loop Poisson stream, half power flows and half single contingencies, network "
                               "uniform over 8 resident grids with bus counts log-uniform in [100, 25,000], load "
                               "multiplier U[0.8, 1.2], warm start from each grid's resident base case, max 10 "
                               "iterations, 1e-8 pu; continuous batching on 4 streams",
                   "networks": [w.n_bus for w in warm], "seed": int(seed), "grid_seeds": seeds, "autotuned": tuned,
                   "device": info["name"] or info["arch"]},
        "latency_ms": {"p50": top["p50_ms"], "p99": top["p99_ms"],
                       "note": f"at the highest offered rate ({top['offered_rps']} req/s); every rate in sweep"},
        "sweep": sweep,
        "verdicts": {"match_cpu_oracle_on_sample": match, "max_dv_vs_oracle_on_sample": max_dv},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    srv.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-sample", type=int, default=48)
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="outages the CPU oracle solves (0: every outage of the sweep)")
    ap.add_argument("--variant-steps", type=int, default=5,
                    help="timed sweeps of the other LU arithmetic (FP64 beside FP32 + 1, or the reverse)")
    ap.add_argument("--serve-rate", type=float, default=2000.0,
                    help="headline: offered rate of the single-contingency request stream (req/s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lu", default="auto", choices=("auto", "fp32", "fp64"),
                    help="LU arithmetic of the batched sweep: auto = the autotuner's choice on the sweep itself")
    ap.add_argument("--config", type=int, default=2, choices=(2, 3, 4, 5),
                    help="2: N-1 sweep of the 2,000-bus grid (the headline); 3: 512 uneven instances, hipGraph; "
                         "4: one 70,000-bus solve, FP64 vs FP32 + refinement; 5: serving, Poisson request stream "
                         "over 8 resident networks")
    ap.add_argument("--seam", action="store_true",
                    help="instead: per-call latency of the sparse-solve seam (gs_sparse_solve and its no-allocation "
                         "and batched forms) on the case300 and 2,000-bus Jacobian patterns")
    ap.add_argument("--c4-variants", default="",
                    help="config 4: comma-separated subset of fp64-wave, fp64-levels, fp32+1-levels, fp32+2-levels")
    ap.add_argument("--serve-seconds", type=float, default=2.0,
                    help="config 5: length of each offered-rate step (requests = rate x seconds, 200..4000)")
    ap.add_argument("--plain-requests", type=int, default=200,
                    help="config 5: requests of each step replayed through the plain one-solve-per-launch loop")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (a kernel of known traffic, for scaling the "
                         "HBM counters of a profiler run)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (run with --gpus 1).", file=sys.stderr)
        sys.exit(2)

    from gridsim_amd import engine, gridgen

    if args.copy_probe_mib:
        import torch
        src = torch.ones(args.copy_probe_mib * 2**18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"# copy probe: {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written", file=sys.stderr)
        del src, dst

    if args.seam:
        return run_seam(args)
    if args.config == 3:
        return run_config3(args)
    if args.config == 4:
        return run_config4(args)
    if args.config == 5:
        return run_config5(args)

    t0 = time.time()
    net = gridgen.config2_grid()
    ctx = engine.Context(0)
    info = ctx.device_info()
    # base case from flat start on the device (FP64), then the warm-start grid the outages run from
    base = engine.DeviceNetwork(ctx, net, engine.NrOptions(flat_start=True, max_iterations=60,
                                                           lu_precision=engine.LU_FP64))
    b = base.solve(count=1)
    assert b.status[0] == engine.CONVERGED, "base case did not converge"
    base.close()
    warm = net.with_start(b.vm[0], np.degrees(b.va[0]))
    outages = gridgen.config2_outages(warm)
    fp32 = engine.NrOptions(max_iterations=10, tolerance=1e-8, lu_precision=engine.LU_FP32, refine_steps=1)
    fp64 = engine.NrOptions(max_iterations=10, tolerance=1e-8, lu_precision=engine.LU_FP64, refine_steps=0)
    dn = engine.DeviceNetwork(ctx, warm, fp32)
    plan = dn.plan()
    tuned_ms = None
    if args.lu == "auto":
        opts, tuned_ms = engine.autotune(dn, outages=outages, options=fp32)
    else:
        opts = fp32 if args.lu == "fp32" else fp64
    lu_name = _lu_name(opts)
    x_place = "global" if opts.flags & engine.FLAG_X_GLOBAL else "lds when it fits"
    print(f"# arithmetic: {lu_name}, schedule {_schedule_name(opts)}, solve vector {x_place}" +
          (" (autotuned)" if tuned_ms else ""), file=sys.stderr, flush=True)
    print(f"# setup {time.time() - t0:.1f}s: {net.n_bus} buses, {net.n_branch} branches, {len(outages)} outages, "
          f"m={len(plan['q'])}, nnz(J)={len(plan['j_colidx'])}, nnz(L+U)={len(plan['li']) + len(plan['ui'])}",
          file=sys.stderr, flush=True)

    for w in range(args.warmup):
        r = dn.solve(outages=outages, options=opts, keep_state=False)
    sweep_ms, kernel_ms = [], []
    t_start = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        r = dn.solve(outages=outages, options=opts, keep_state=False)
        sweep_ms.append(1000 * (time.perf_counter() - ts))
        kernel_ms.append(r.kernel_ms)
        print(f"# step {k}: sweep {sweep_ms[-1]:.2f} ms, kernel {r.kernel_ms:.2f} ms", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    res = r.results
    phase = res["phase_us"].mean(axis=0)
    print("# mean device time per instance (us): " +
          ", ".join(f"{n} {v:.0f}" for n, v in zip(engine.PHASES, phase)), file=sys.stderr, flush=True)
    converged = int((res["status"] == engine.CONVERGED).sum())
    ms_per_step = 1000 * elapsed / args.steps
    value = converged * args.steps / elapsed

    # roofline of the batched kernel: algorithmic bytes of one launch / its average duration (HIP events)
    nnz_y = net.n_bus + 2 * net.n_branch
    alg = algorithmic_bytes(plan, net.n_bus, nnz_y, res["iterations"], opts.refine_steps if opts.lu_precision ==
                            engine.LU_FP32 else 0, 4 if opts.lu_precision == engine.LU_FP32 else 8)
    kavg = float(np.mean(kernel_ms))
    achieved = alg / (kavg / 1000) / 1e9
    stats = dn.plan_stats()
    shape = dn.launch_shape(opts)
    traffic, traffic_src = measured_traffic(stats, "fp32" if opts.lu_precision == engine.LU_FP32 else "fp64",
                                            opts.flags, alg, kavg)

    # the other arithmetic on the same sweep: FP32 LU + one FP64 refinement step beside FP64 LU
    other = fp64 if opts.lu_precision == engine.LU_FP32 else fp32
    dn.solve(outages=outages, options=other, keep_state=False)
    t_o = time.perf_counter()
    o_kms = []
    for _ in range(args.variant_steps):
        ro = dn.solve(outages=outages, options=other, keep_state=False)
        o_kms.append(ro.kernel_ms)
    o_el = (time.perf_counter() - t_o) / args.variant_steps
    variant = {"lu": _lu_name(other), "value": round(int((ro.status == engine.CONVERGED).sum()) / o_el, 1),
               "unit": "converged solves/s", "ms_per_step": round(1000 * o_el, 3),
               "kernel_ms": round(float(np.mean(o_kms)), 3), "steps": args.variant_steps,
               "verdicts_equal_headline": bool(np.array_equal(ro.status, res["status"]))}
    print(f"# variant {variant}", file=sys.stderr, flush=True)

    # the plain FP64 one-instance-per-launch path on a sample of the same outages
    sample = outages[:: max(1, len(outages) // args.plain_sample)][: args.plain_sample]
    tp = time.perf_counter()
    pr = dn.solve(outages=sample, plain=True, keep_state=False,
                  options=engine.NrOptions(max_iterations=10, lu_precision=engine.LU_FP64))
    plain_s = time.perf_counter() - tp
    plain_conv = int((pr.status == engine.CONVERGED).sum())
    plain_rate = plain_conv / plain_s
    plain_lat = plain_s * 1000 / len(sample)

    # per-request latency: the same outages as single-contingency requests, a Poisson stream through the serving loop
    serving = _serve_segment(engine, ctx, dn, opts, outages, args)
    print(f"# serving {serving}", file=sys.stderr, flush=True)

    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads, host = _host_threads()
        cs = outages if args.cpu_sample <= 0 else outages[:: max(1, len(outages) // args.cpu_sample)][: args.cpu_sample]
        tc = time.perf_counter()
        cres, cvm, cva = O.screen(warm, cs, O.options(max_iterations=10), threads=threads, keep_state=True)
        csec = time.perf_counter() - tc
        cpu = {"value": round(int((cres["status"] == 0).sum()) / csec, 1), "unit": "converged solves/s",
               "cores": threads, "kind": "port",
               "sample": f"{len(cs)} of the {len(outages)} outages, {threads} threads, one solver per thread "
                         f"(C++ restatement of the reference's CPU sparse NR, ScreenBranchesParallel's shape; "
                         f"per-solve p50 {1000 * np.median(cres['solve_seconds']):.2f} ms, p99 "
                         f"{1000 * np.percentile(cres['solve_seconds'], 99):.2f} ms single-thread)",
               "host": host}
        # parity on every outage of the sweep: the verdicts, and |dV| and the final mismatch of the converged ones
        chk = dn.solve(outages=cs, options=opts, keep_state=True)
        conv = cres["status"] == 0
        dconv = chk.status == 0
        parity = {"instances": int(len(cs)),
                  "verdict_mismatches": int((chk.status != cres["status"]).sum()),
                  "iteration_mismatches": int((chk.results["iterations"] != cres["iterations"]).sum()),
                  "max_dv_vs_oracle": float(max(np.abs(chk.vm[conv] - cvm[conv]).max(),
                                                np.abs(chk.va[conv] - cva[conv]).max())) if conv.any() else None,
                  "max_mismatch_pu": float(chk.results["max_mismatch"][dconv].max()) if dconv.any() else None,
                  "oracle_max_mismatch_pu": float(cres["max_mismatch"][conv].max()) if conv.any() else None,
                  "bar": "same verdict per outage, |dV| <= 1e-9 (pu, rad), mismatch < 1e-8 pu"}
        parity["pass"] = bool(parity["verdict_mismatches"] == 0 and (parity["max_dv_vs_oracle"] or 0.0) <= 1e-9 and
                              (parity["max_mismatch_pu"] or 0.0) < 1e-8)

    line = {
        "metric": "N-1 NR solves/s at 2k buses (1e-8 pu) + p50/p99 solve latency, 1 MI355X",
        "value": round(value, 1),
        "unit": "converged solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-lu+f64" if opts.lu_precision == engine.LU_FP32 else "f64",
        "data": "synthetic",
        "config": {"workload": "batched N-1 screen, 2000-bus synthetic grid (1 slack / 300 PV), every non-bridge "
                               "single-branch outage from the base-case voltages, max 10 iterations, 1e-8 pu",
                   "n_bus": net.n_bus, "n_branch": net.n_branch, "outages": int(len(outages)),
                   "seed": int(net.seed), "lu": lu_name, "schedule": _schedule_name(opts), "solve_vector": x_place,
                   "autotuned": tuned_ms is not None,
                   "device": info["name"] or info["arch"]},
        "latency_ms": {"p50": serving["p50_ms"], "p99": serving["p99_ms"],
                       "note": f"per request through the serving loop: {serving['requests']} single-contingency "
                               f"requests on this grid, Poisson arrivals at {serving['offered_rps']} req/s offered "
                               f"(submit to result ready)",
                       "sweep_p50": round(float(np.percentile(sweep_ms, 50)), 3),
                       "sweep_p99": round(float(np.percentile(sweep_ms, 99)), 3)},
        "serving": serving,
        "verdicts": {"converged": converged, "max_iterations": int((res["status"] == 1).sum()),
                     "diverged": int((res["status"] == 2).sum()), "nonfinite": int((res["status"] == 3).sum()),
                     "symbolic_refreshes": int(res["symbolic_refreshes"].sum())},
        "parity_vs_cpu_oracle": parity,
        "variant": variant,
        "plain_fp64_one_instance_per_launch": {"value": round(plain_rate, 2), "unit": "converged solves/s",
                                               "sample": int(len(sample)), "latency_ms_per_solve": round(plain_lat, 2),
                                               "speedup_of_batched": round(value / plain_rate, 1)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "plan_stream_bytes": stats["stream_bytes"],
                     "kernel": "nr_batch_kernel<%s,%s,%s>" % ("float" if shape["lu_fp32"] else "double",
                                                               str(shape["x_in_lds"]).lower(),
                                                               str(shape["waves_per_group"] == 2).lower()),
                     "kernel_ms": round(kavg, 3),
                     "launch": shape,
                     "algorithmic_bytes": int(alg),
                     "phase_us_per_instance": {n: round(float(v), 1) for n, v in zip(engine.PHASES, phase)}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    dn.close()
    ctx.close()


def _lu_name(o) -> str:
    from gridsim_amd import engine
    return "fp32 + %d fp64 refinement step(s)" % o.refine_steps if o.lu_precision == engine.LU_FP32 else "fp64"


def _host_threads():
    """Worker threads for the CPU baseline: every core this process may run on (Environment.ProcessorCount, the
    reference's Parallel.For default, honours the affinity mask), at most 16 (the host share one GPU gets on the
    benchmark machines); and a description of the host."""
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = nproc
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = max(1, min(avail, 16))
    return threads, {"nproc": nproc, "affinity": avail, "cpu_model": model, "threads_used": threads}


def _serve_segment(engine, ctx, dn, opts, outages, args):
    """Single-contingency requests on the resident 2,000-bus network through the continuous-batching server
    (gs_server_*): an open-loop Poisson stream at args.serve_rate req/s for args.serve_seconds; p50 / p99 of submit
    to result ready (the server's own clock) and the served rate."""
    rng = np.random.default_rng(gridgen_seed())
    srv = engine.Server(ctx, max_batch=512)
    nid = srv.add_network(dn, opts)
    n_req = max(50, int(args.serve_rate * args.serve_seconds))
    reqs = [(0, int(rng.choice(outages)), 1.0) for _ in range(n_req)]
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, n_req))
    # warm-up: a few requests
    _serve_stream(lambda k, o, sc: srv.submit(nid, o, sc), lambda t, k: srv.wait(t, nid, keep_state=False),
                  reqs[:8], np.zeros(8))
    srv.reset_stats()
    out, wall = _serve_stream(lambda k, o, sc: srv.submit(nid, o, sc), lambda t, k: srv.wait(t, nid, keep_state=False),
                              reqs, arrivals)
    st = srv.stats()
    srv.close()
    conv = sum(int(r[0]["status"] == 0) for r in out)
    return {"offered_rps": args.serve_rate, "requests": n_req, "served_rps": round(n_req / wall, 1),
            "converged": conv, "p50_ms": round(st["latency_p50_ms"], 3), "p99_ms": round(st["latency_p99_ms"], 3),
            "max_ms": round(st["latency_max_ms"], 3), "mean_batch": round(st["mean_batch"], 2),
            "launches": st["launches"], "joined_running": st["joined"]}


def gridgen_seed() -> int:
    from gridsim_amd import gridgen
    return gridgen.BASE_SEED + 2


if __name__ == "__main__":
    main()
