"""The serving loop (gs_server_*) on an MI355X, through the C ABI.

Continuous batching must not change answers: each request's result is the same whether it runs alone (submitted and
waited for before the next) or in a mix of requests of several networks that the dispatcher batches and overlaps on
its streams, and the same as the batch API's for that instance. Bars: identical verdicts and iteration counts, |dV|
<= 1e-12 between alone and mix (the arithmetic of one instance does not depend on its launch), and the oracle's
verdicts with |dV| <= 1e-9 pu / rad for the FP64 network.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_SAME = 1e-12
TOL_DV = 1e-9


@pytest.fixture(scope="module")
def ctx():
    from gridsim_amd import engine
    c = engine.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def warm_nets(oracle):
    """Three grids (118, 300 and 1,200 buses), each warm-started from its solved base case."""
    from gridsim_amd import gridgen
    out = []
    for net in (gridgen.config1_grid(), gridgen.sized_grid(300, gridgen.BASE_SEED + 11),
                gridgen.sized_grid(1200, gridgen.BASE_SEED + 12)):
        rec, vm, va = oracle.solve(net, oracle.options(flat_start=True, max_iterations=60))
        assert rec["status"] == 0
        out.append(net.with_start(vm, np.degrees(va)))
    return out


def _requests(nets, count, seed):
    """(network id, outage or -1, load multiplier) drawn like the serving workload: power flows and single
    contingencies, load multiplier U[0.8, 1.2]."""
    from gridsim_amd import gridgen
    rng = np.random.default_rng(seed)
    outs = [gridgen.config2_outages(n) for n in nets]
    reqs = []
    for _ in range(count):
        k = int(rng.integers(len(nets)))
        o = int(rng.choice(outs[k])) if rng.random() < 0.6 else -1
        reqs.append((k, o, float(rng.uniform(0.8, 1.2))))
    return reqs


def test_server_result_is_the_same_alone_and_in_the_mix(ctx, warm_nets, oracle):
    from gridsim_amd import engine
    opts = [engine.NrOptions(max_iterations=10, lu_precision=engine.LU_FP64, refine_steps=0),
            engine.NrOptions(max_iterations=10, lu_precision=engine.LU_FP32, refine_steps=1),
            engine.NrOptions(max_iterations=10, lu_precision=engine.LU_FP32, refine_steps=2)]
    dns = [engine.DeviceNetwork(ctx, n, o) for n, o in zip(warm_nets, opts)]
    srv = engine.Server(ctx, max_batch=16)
    ids = [srv.add_network(dn, o) for dn, o in zip(dns, opts)]
    reqs = _requests(warm_nets, 60, seed=5)

    print("alone: one request at a time", flush=True)
    alone = []
    for k, o, s in reqs:
        alone.append(srv.wait(srv.submit(ids[k], o, s), ids[k]))
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["mean_batch"] == 1.0

    print("mix: every request submitted at once", flush=True)
    srv.reset_stats()
    tickets = [srv.submit(ids[k], o, s) for k, o, s in reqs]
    mix = [srv.wait(t, ids[k]) for t, (k, _, _) in zip(tickets, reqs)]
    st = srv.stats()
    print("mix stats", st, flush=True)
    assert st["completed"] == len(reqs) and st["mean_batch"] > 1.0, "the mix was not batched"
    assert st["latency_p99_ms"] >= st["latency_p50_ms"] > 0

    for (k, o, s), a, m in zip(reqs, alone, mix):
        assert a[0]["status"] == m[0]["status"] and a[0]["iterations"] == m[0]["iterations"], (k, o, s)
        assert np.abs(a[1] - m[1]).max() <= TOL_SAME and np.abs(a[2] - m[2]).max() <= TOL_SAME, (k, o, s)

    # the same instances through the batch API, network by network
    for k, dn in enumerate(dns):
        sel = [i for i, r in enumerate(reqs) if r[0] == k]
        if not sel:
            continue
        r = dn.solve(outages=[reqs[i][1] for i in sel], load_scale=[reqs[i][2] for i in sel])
        for j, i in enumerate(sel):
            assert r.status[j] == mix[i][0]["status"]
            assert np.abs(r.vm[j] - mix[i][1]).max() <= TOL_SAME and np.abs(r.va[j] - mix[i][2]).max() <= TOL_SAME

    # the oracle's verdicts and converged voltages (to 1e-9) on every request, the FP32 networks' included
    for (k, o, s), m in zip(reqs, mix):
        net = warm_nets[k].with_load_multiplier(s)
        rec, vm, va = oracle.solve(net, oracle.options(max_iterations=10), skip=o)
        assert rec["status"] == m[0]["status"], (k, o, s)
        if rec["status"] == 0:
            assert m[0]["max_mismatch"] < 1e-8
            assert np.abs(vm - m[1]).max() <= TOL_DV and np.abs(va - m[2]).max() <= TOL_DV, (k, o, s)
    srv.close()
    for dn in dns:
        dn.close()


def test_server_queue_limit_and_argument_checks(ctx, warm_nets):
    from gridsim_amd import engine
    dn = engine.DeviceNetwork(ctx, warm_nets[2], engine.NrOptions(max_iterations=10))
    srv = engine.Server(ctx, max_batch=1, streams=1, queue_capacity=4)
    nid = srv.add_network(dn)
    with pytest.raises(engine.GsError) as e:
        srv.submit(nid + 1)
    assert e.value.status == engine.GS_INVALID_ARGUMENT
    with pytest.raises(engine.GsError) as e:
        srv.submit(nid, outage=warm_nets[2].n_branch)
    assert e.value.status == engine.GS_INVALID_ARGUMENT
    with pytest.raises(engine.GsError) as e:
        srv.wait(12345, nid)
    assert e.value.status == engine.GS_INVALID_ARGUMENT
    # one launch slot of one instance: a burst of submissions fills the queue of four
    tickets, busy = [], 0
    for i in range(40):
        try:
            tickets.append(srv.submit(nid, -1, 0.8 + 0.01 * i))
        except engine.GsError as err:
            assert err.status == engine.GS_BUSY
            busy += 1
    print(f"accepted {len(tickets)}, busy {busy}", flush=True)
    assert busy > 0
    for t in tickets:
        r = srv.wait(t, nid)
        assert r[0]["status"] == engine.CONVERGED
    st = srv.stats()
    assert st["rejected"] == busy and st["completed"] == len(tickets)
    srv.close()
    dn.close()


def test_rpc_session_over_the_gpu_server(ctx, oracle):
    """A gridsim-rpc/1 session (gridsim_amd/rpc.py) over the GPU server: two cases loaded, a burst of pipelined solves
    (power flows and single contingencies) answered in request order, each equal to the oracle's solve of the same
    request (verdict, |dV| <= 1e-9, losses), with the solves batched on the device."""
    import io
    import json
    from conftest import GOLDEN
    from gridsim_amd import gridgen, rpc
    backend = rpc.GpuBackend(ctx, max_batch=64)
    host = rpc.RpcHost(backend, base_dir=GOLDEN)
    syn = gridgen.sized_grid(300, gridgen.BASE_SEED + 21)
    outs9 = [1, 2, 4, 5, 7, 8]                       # case9's non-bridge branches
    outs300 = list(gridgen.config2_outages(syn)[:10])
    lines = [{"op": "load", "id": "l1", "case": "matpower/case9.m"},
             {"op": "load", "id": "l2", "case": f"synthetic:300:{gridgen.BASE_SEED + 21}"}]
    reqs = []
    for i in range(40):
        if i % 2 == 0:
            reqs.append(("case-1", outs9[i % len(outs9)] if i % 4 == 0 else -1, 1.0 + 0.01 * (i % 5)))
        else:
            reqs.append(("case-2", int(outs300[i % len(outs300)]) if i % 3 else -1, 0.9 + 0.02 * (i % 7)))
    for i, (h, o, s) in enumerate(reqs):
        lines.append({"op": "solve", "id": i, "handle": h, "outage": o, "load_scale": s})
    lines += [{"op": "stats", "id": "s"}, {"op": "shutdown", "id": "x"}]
    out = io.StringIO()
    host.serve(io.StringIO("\n".join(json.dumps(l) for l in lines) + "\n"), out)
    resp = [json.loads(x) for x in out.getvalue().splitlines()]
    assert [r["id"] for r in resp] == ["l1", "l2"] + list(range(40)) + ["s", "x"]
    from conftest import matpower_case
    nets = {"case-1": matpower_case("case9"), "case-2": syn}
    for (h, o, s), r in zip(reqs, resp[2:42]):
        net = nets[h].with_load_multiplier(s) if s != 1.0 else nets[h]
        rec, vm, va = oracle.solve(net, oracle.options(max_iterations=100), skip=o)
        res = r["result"]
        assert res["converged"] == (rec["status"] == 0), (h, o, s)
        if rec["status"] == 0:
            assert np.abs(np.array(res["vm_pu"]) - vm).max() <= TOL_DV
            assert np.abs(np.radians(res["va_deg"]) - va).max() <= 1e-9
            p, q, _ = oracle.mismatch(net, vm, va, skip=o)
            assert abs(res["loss_mw"] - (p * net.base_mva).sum()) <= 1e-6
    st = resp[42]["result"]
    print("rpc stats", st, flush=True)
    # (stats is a snapshot taken when its line is read, while the pipelined solves may still be on the device)
    assert st["solves"] == 40 and st["loads"] == 2 and st["handles"] == 2
    done = backend.stats()
    print("server after the session", done, flush=True)
    assert done["completed"] == 40 and done["mean_batch"] > 1.0, "the pipelined solves were not batched"
    backend.close()


def test_server_n2_requests_match_the_oracle(ctx, warm_nets, oracle):
    """Requests that take two branches out (an outage list) through the serving loop, mixed with single outages:
    the oracle's verdicts and converged voltages."""
    from gridsim_amd import engine, gridgen
    net = warm_nets[1]
    dn = engine.DeviceNetwork(ctx, net, engine.NrOptions(max_iterations=10, lu_precision=engine.LU_FP64,
                                                         refine_steps=0))
    srv = engine.Server(ctx, max_batch=8)
    nid = srv.add_network(dn)
    rng = np.random.default_rng(41)
    cand = gridgen.config2_outages(net)
    reqs = [[int(b) for b in rng.choice(cand, 2, replace=False)] if j % 2 else int(rng.choice(cand))
            for j in range(24)]
    tickets = [srv.submit(nid, r, 1.0) for r in reqs]
    for r, t in zip(reqs, tickets):
        res, vm, va, _ = srv.wait(t, nid)
        rec, rvm, rva = oracle.solve(net, oracle.options(max_iterations=10), skip=r)
        assert rec["status"] == res["status"], r
        if rec["status"] == 0:
            assert np.abs(vm - rvm).max() <= TOL_DV and np.abs(va - rva).max() <= TOL_DV, r
    srv.close()
    dn.close()


def test_server_runs_a_work_group_per_instance_network(ctx, warm_nets, oracle):
    """A network served on the work-group-per-instance schedule (8 waves per instance): every request's answer is the
    wavefront batch API's bit for bit, alone and mixed with the other networks' requests, and the oracle's."""
    from gridsim_amd import engine
    group = engine.NrOptions(max_iterations=10, lu_precision=engine.LU_FP64, refine_steps=0,
                             schedule=engine.SCHEDULE_GROUP, flags=engine.group_waves(8))
    plain = engine.NrOptions(max_iterations=10, lu_precision=engine.LU_FP32, refine_steps=1)
    dns = [engine.DeviceNetwork(ctx, warm_nets[2], group), engine.DeviceNetwork(ctx, warm_nets[0], plain)]
    srv = engine.Server(ctx, max_batch=16)
    ids = [srv.add_network(dns[0], group), srv.add_network(dns[1], plain)]
    rng = np.random.default_rng(9)
    from gridsim_amd import gridgen
    outs = gridgen.config2_outages(warm_nets[2])
    reqs = [(int(rng.integers(2)), int(rng.choice(outs)) if rng.random() < 0.7 else -1, float(rng.uniform(0.8, 1.2)))
            for _ in range(40)]
    reqs = [(k, o if k == 0 else -1, s) for k, o, s in reqs]
    tickets = [srv.submit(ids[k], o, s) for k, o, s in reqs]
    mix = [srv.wait(t, ids[k]) for t, (k, _, _) in zip(tickets, reqs)]
    sel = [i for i, r in enumerate(reqs) if r[0] == 0]
    wave = dns[0].solve(outages=[reqs[i][1] for i in sel], load_scale=[reqs[i][2] for i in sel],
                        options=engine.NrOptions(max_iterations=10, lu_precision=engine.LU_FP64, refine_steps=0))
    for j, i in enumerate(sel):
        assert wave.status[j] == mix[i][0]["status"] and wave.results["iterations"][j] == mix[i][0]["iterations"]
        assert np.array_equal(wave.vm[j], mix[i][1]) and np.array_equal(wave.va[j], mix[i][2]), reqs[i]
    for i in sel[:6]:
        k, o, s = reqs[i]
        rec, vm, va = oracle.solve(warm_nets[0 if k else 2].with_load_multiplier(s), oracle.options(max_iterations=10),
                                   skip=o)
        assert rec["status"] == mix[i][0]["status"]
        if rec["status"] == 0:
            assert np.abs(vm - mix[i][1]).max() <= TOL_DV and np.abs(va - mix[i][2]).max() <= TOL_DV
    srv.close()
    for dn in dns:
        dn.close()


def test_server_requests_join_a_running_batch_between_newton_steps(ctx, warm_nets, oracle):
    """More requests of one network than its slot has positions: the first ones start the batch, the rest wait and take
    positions as earlier instances finish (outages and load multipliers make them finish at different Newton steps), so
    they join a batch whose other instances are between steps. Each answer is the batch API's for the same instance bit
    for bit, and the oracle's verdict and voltages on a sample."""
    from gridsim_amd import engine, gridgen
    net = warm_nets[2]
    o = engine.NrOptions(max_iterations=10, lu_precision=engine.LU_FP64, refine_steps=0)
    dn = engine.DeviceNetwork(ctx, net, o)
    srv = engine.Server(ctx, max_batch=16, streams=1)
    nid = srv.add_network(dn, o)
    rng = np.random.default_rng(17)
    outs = gridgen.config2_outages(net)
    reqs = [(int(rng.choice(outs)) if j % 3 else -1, float(rng.uniform(0.8, 1.2))) for j in range(64)]
    tickets = [srv.submit(nid, out, s) for out, s in reqs]
    res = [srv.wait(t, nid) for t in tickets]
    st = srv.stats()
    print("join stats", st, flush=True)
    assert st["completed"] == len(reqs)
    assert st["joined"] > 0, "no request joined a running batch"
    assert st["mean_batch"] > 1.0
    r = dn.solve(outages=[x[0] for x in reqs], load_scale=[x[1] for x in reqs])
    for j, (out, s) in enumerate(reqs):
        assert r.status[j] == res[j][0]["status"] and r.results["iterations"][j] == res[j][0]["iterations"], (out, s)
        assert np.array_equal(r.vm[j], res[j][1]) and np.array_equal(r.va[j], res[j][2]), (out, s)
    for j in range(0, len(reqs), 8):
        out, s = reqs[j]
        rec, vm, va = oracle.solve(net.with_load_multiplier(s), oracle.options(max_iterations=10), skip=out)
        assert rec["status"] == res[j][0]["status"], (out, s)
        if rec["status"] == 0:
            assert np.abs(vm - res[j][1]).max() <= TOL_DV and np.abs(va - res[j][2]).max() <= TOL_DV, (out, s)
    srv.close()
    dn.close()
