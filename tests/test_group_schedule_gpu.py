"""The work-group-per-instance schedule (several wavefronts of one work-group share an instance's Newton loop) on an
MI355X: the same bits as one wavefront per instance, and the oracle's verdicts, iterations and voltages.

Grids: the 300- and 10,000-bus grids of the uneven batch (warm-started from their base case, load multipliers
U[0.8, 1.2], one random non-bridge outage each) and the 2,000-bus screening grid (single outages from the base-case
voltages). FP64 LU and FP32 LU with one FP64 refinement step; 2, 4 and 8 waves per work-group.
Bar against the oracle: its verdict and iteration count per instance, final mismatch < 1e-8 pu, |dV| <= 1e-9 (pu, rad).
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
def cases(oracle):
    from gridsim_amd import gridgen
    out = {}
    rng = np.random.default_rng(gridgen.BASE_SEED + 11)
    for n in (300, 10000):
        net = gridgen.sized_grid(n, gridgen.BASE_SEED + 3)
        rec, vm, va = oracle.solve(net, oracle.options(flat_start=True, max_iterations=60))
        assert rec["status"] == 0, f"{n}-bus base case"
        warm = net.with_start(vm, np.degrees(va))
        outs = rng.choice(gridgen.config2_outages(warm), 6, replace=False)
        scales = rng.uniform(0.8, 1.2, 6)
        out[n] = (warm, outs, scales)
    net = gridgen.config2_grid()
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True, max_iterations=60))
    warm = net.with_start(vm, np.degrees(va))
    outs = rng.choice(gridgen.config2_outages(warm), 24, replace=False)
    out[2000] = (warm, outs, np.ones(len(outs)))
    return out


@pytest.fixture(scope="module")
def oracle_refs(oracle, cases):
    refs = {}
    for n, (warm, outs, scales) in cases.items():
        k = min(len(outs), 6)
        refs[n] = [oracle.solve(warm.with_load_multiplier(float(s)), oracle.options(max_iterations=10), skip=int(o))
                   for o, s in zip(outs[:k], scales[:k])]
    return refs


@pytest.mark.parametrize("n", [300, 2000, 10000])
@pytest.mark.parametrize("lu", ["fp64", "fp32+1"])
def test_group_schedule_is_the_wavefront_schedule_bit_for_bit_and_the_oracle(ctx, cases, oracle_refs, n, lu):
    from gridsim_amd import engine as eng
    warm, outs, scales = cases[n]
    prec, refine = (eng.LU_FP64, 0) if lu == "fp64" else (eng.LU_FP32, 1)
    dn = eng.DeviceNetwork(ctx, warm, eng.NrOptions(max_iterations=10))
    base = eng.NrOptions(max_iterations=10, lu_precision=prec, refine_steps=refine)
    wave = dn.solve(outs, scales, base)
    print(f"{n} buses {lu}: wavefront {wave.kernel_ms:.2f} ms", flush=True)
    for w in (2, 4, 8):
        o = eng.NrOptions(max_iterations=10, lu_precision=prec, refine_steps=refine, schedule=eng.SCHEDULE_GROUP,
                          flags=eng.group_waves(w))
        shape = dn.launch_shape(o)
        assert 1 <= shape["waves_per_group"] <= w
        r = dn.solve(outs, scales, o)
        print(f"{n} buses {lu}: {shape['waves_per_group']} waves (x in LDS {shape['x_in_lds']}) "
              f"{r.kernel_ms:.2f} ms", flush=True)
        for f in ("status", "iterations", "max_mismatch", "symbolic_refreshes", "step_failures", "min_pivot_ratio"):
            np.testing.assert_array_equal(r.results[f], wave.results[f], err_msg=f"{f}, {w} waves")
        np.testing.assert_array_equal(r.vm, wave.vm)
        np.testing.assert_array_equal(r.va, wave.va)
    for i, (rec, vm, va) in enumerate(oracle_refs[n]):
        where = (n, int(outs[i]), float(scales[i]))
        assert r.status[i] == rec["status"], where
        assert r.results["iterations"][i] == rec["iterations"], where
        if rec["status"] == 0:
            assert r.results["max_mismatch"][i] < 1e-8, where
            assert np.abs(r.vm[i] - vm).max() <= TOL_DV and np.abs(r.va[i] - va).max() <= TOL_DV, where
    dn.close()


def test_group_schedule_with_the_solve_vector_in_global_memory(ctx, cases):
    from gridsim_amd import engine as eng
    warm, outs, scales = cases[300]
    dn = eng.DeviceNetwork(ctx, warm, eng.NrOptions(max_iterations=10))
    base = eng.NrOptions(max_iterations=10, lu_precision=eng.LU_FP64, refine_steps=0)
    wave = dn.solve(outs, scales, base)
    o = eng.NrOptions(max_iterations=10, lu_precision=eng.LU_FP64, refine_steps=0, schedule=eng.SCHEDULE_GROUP,
                      flags=eng.group_waves(4) | eng.FLAG_X_GLOBAL)
    assert not dn.launch_shape(o)["x_in_lds"]
    r = dn.solve(outs, scales, o)
    np.testing.assert_array_equal(r.vm, wave.vm)
    np.testing.assert_array_equal(r.va, wave.va)
    np.testing.assert_array_equal(r.results["iterations"], wave.results["iterations"])
    dn.close()
