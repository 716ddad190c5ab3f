"""The workloads of the uneven-size batch and the large single instance against the oracle on an MI355X.

* Uneven batch: 300-, 2,000- and 10,000-bus grids (branches 1.35 n, 15 % PV), each warm-started from its base case,
  instances with a load multiplier U[0.8, 1.2] and one random non-bridge outage, all solved by one captured hipGraph
  replay (the 10,000-bus group on the level-parallel schedule with its Newton steps unrolled, or on a work-group of
  wavefronts per instance), FP64 LU and FP32 LU with one FP64 refinement step.
* Large instance: the 70,000-bus grid from flat start (max 100 iterations), level-parallel, FP64 and FP32 + 1.

Bar: the oracle's verdict and iteration count per instance, final mismatch < 1e-8 pu, |dV| <= 1e-9 (pu, rad).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_DV = 1e-9


@pytest.fixture(scope="module")
def ctx():
    from gridsim_amd import engine
    c = engine.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def uneven(oracle):
    from gridsim_amd import gridgen
    seed = gridgen.BASE_SEED + 3
    rng = np.random.default_rng(seed)
    groups = []
    for n in (300, 2000, 10000):
        net = gridgen.sized_grid(n, seed)
        rec, vm, va = oracle.solve(net, oracle.options(flat_start=True, max_iterations=60))
        assert rec["status"] == 0, f"{n}-bus base case"
        warm = net.with_start(vm, np.degrees(va))
        outs = rng.choice(gridgen.config2_outages(warm), 6, replace=False)
        scales = rng.uniform(0.8, 1.2, 6)
        refs = [oracle.solve(warm.with_load_multiplier(float(s)), oracle.options(max_iterations=10), skip=int(o))
                for o, s in zip(outs, scales)]
        groups.append((warm, outs, scales, refs))
    return groups


@pytest.mark.parametrize("large_schedule", ["levels", "group"])
@pytest.mark.parametrize("lu", ["fp64", "fp32+1"])
def test_uneven_batch_in_one_graph_matches_the_oracle(ctx, uneven, lu, large_schedule):
    from gridsim_amd import engine as eng
    prec, refine = (eng.LU_FP64, 0) if lu == "fp64" else (eng.LU_FP32, 1)
    dns, groups = [], []
    for k, (warm, outs, scales, _) in enumerate(uneven):
        sched = 0 if k < 2 else eng.SCHEDULE_LEVELS if large_schedule == "levels" else eng.SCHEDULE_GROUP
        o = eng.NrOptions(max_iterations=10, lu_precision=prec, refine_steps=refine, schedule=sched)
        dn = eng.DeviceNetwork(ctx, warm, eng.NrOptions(max_iterations=10))
        dns.append(dn)
        groups.append(eng.Group(dn, outages=outs, load_scale=scales, options=o))
    b = eng.Batch(ctx, groups, use_graph=True)
    for rep in range(2):
        ms = b.run()
        print(f"uneven batch ({lu}, 10,000-bus group {large_schedule}) replay {rep}: {ms:.2f} ms", flush=True)
        for g, (warm, outs, scales, refs) in enumerate(uneven):
            r = b.results(g, keep_state=True)
            for i, (rec, vm, va) in enumerate(refs):
                where = (warm.n_bus, int(outs[i]), float(scales[i]))
                assert r.status[i] == rec["status"], where
                assert r.results["iterations"][i] == rec["iterations"], where
                if rec["status"] == 0:
                    assert r.results["max_mismatch"][i] < 1e-8, where
                    assert np.abs(r.vm[i] - vm).max() <= TOL_DV and np.abs(r.va[i] - va).max() <= TOL_DV, where
    b.close()
    for dn in dns:
        dn.close()


@pytest.fixture(scope="module")
def large(oracle):
    from gridsim_amd import gridgen
    net = gridgen.config4_grid()
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True, max_iterations=100))
    assert rec["status"] == 0
    return net, rec, vm, va


@pytest.mark.parametrize("lu", ["fp64", "fp32+1"])
def test_large_instance_level_parallel_matches_the_oracle(ctx, large, lu):
    from gridsim_amd import engine as eng
    net, rec, vm, va = large
    prec, refine = (eng.LU_FP64, 0) if lu == "fp64" else (eng.LU_FP32, 1)
    o = eng.NrOptions(flat_start=True, max_iterations=100, lu_precision=prec, refine_steps=refine,
                      schedule=eng.SCHEDULE_LEVELS)
    dn = eng.DeviceNetwork(ctx, net, eng.NrOptions(flat_start=True, max_iterations=100))
    r = dn.solve(count=1, options=o)
    dv = max(np.abs(r.vm[0] - vm).max(), np.abs(r.va[0] - va).max())
    print(f"70,000 buses ({lu}): {r.results['iterations'][0]} iterations (oracle {rec['iterations']}), "
          f"{r.kernel_ms:.1f} ms device, |dV| {dv:.2e}", flush=True)
    assert r.status[0] == eng.CONVERGED and r.results["iterations"][0] == rec["iterations"]
    assert r.results["max_mismatch"][0] < 1e-8 and dv <= TOL_DV
    dn.close()
