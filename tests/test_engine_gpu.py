"""Parity of the HIP engine with the oracle, through the C ABI, on an MI355X.

Bars: bit-exact for FDLIBM and for the FP64 sparse-solver seam; for Newton solves the reference's correctness
contract: the same verdict per instance, max mismatch < 1e-8 pu at convergence, and |dV| <= 1e-9 pu (magnitude in
pu, angle in rad) against the oracle's converged state.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, matpower_case
import strictmath

pytestmark = pytest.mark.gpu

TOL_DV = 1e-9


@pytest.fixture(scope="module")
def ctx():
    from gridsim_amd import engine
    c = engine.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def eng():
    from gridsim_amd import engine
    return engine


def test_device_probe(ctx):
    info = ctx.device_info()
    print("device", info, flush=True)
    assert info["fp64_ok"] and info["fdlibm_ok"]
    assert "gfx950" in info["arch"]


def test_device_fdlibm_is_strictmath_bit_for_bit(ctx):
    vec, dig = strictmath.load(os.path.join(GOLDEN, "strictmath_sincos.txt"))
    xs = np.array([strictmath.from_bits(v[1]) for v in vec])
    s, c = ctx.device_sincos(xs)
    got = np.where([v[0] == "sin" for v in vec], s, c)
    want = np.array([strictmath.from_bits(v[2]) for v in vec])
    ok = ~np.isnan(want)
    assert np.array_equal(got[ok].view(np.uint64), want[ok].view(np.uint64))
    for fn, cls, seed, count, d in dig:
        s, c = ctx.device_sincos(strictmath.draw(cls, seed, count))
        assert strictmath.digest(s if fn == "sin" else c) == d, (fn, cls)


def _compare(oracle, res, vm, va, ref_res, ref_vm, ref_va, label):
    assert np.array_equal(res["status"], ref_res["status"]), f"{label}: verdicts differ"
    conv = res["status"] == 0
    assert np.all(res["max_mismatch"][conv] < 1e-8), label
    if conv.any():
        dvm = np.abs(vm[conv] - ref_vm[conv]).max()
        dva = np.abs(va[conv] - ref_va[conv]).max()
        print(f"{label}: {conv.sum()} converged, max |dVm| {dvm:.2e}, max |dVa| {dva:.2e} rad", flush=True)
        assert dvm <= TOL_DV and dva <= TOL_DV, label


@pytest.mark.parametrize("case", ["case9", "case14", "case118", "case300"])
@pytest.mark.parametrize("lu", ["fp64", "fp32"])
def test_single_solve_matches_oracle(ctx, eng, oracle, case, lu):
    net = matpower_case(case)
    opts = eng.NrOptions(lu_precision=eng.LU_FP64 if lu == "fp64" else eng.LU_FP32, refine_steps=1)
    dn = eng.DeviceNetwork(ctx, net, opts)
    r = dn.solve(count=1)
    rec, vm, va = oracle.solve(net)
    _compare(oracle, r.results, r.vm, r.va, np.array([rec]), vm[None], va[None], f"{case}/{lu}")
    assert r.results["iterations"][0] == rec["iterations"]


def test_fp64_solve_agrees_with_oracle_to_1e13_case118(ctx, eng, oracle):
    """FP64 LU on the reference's own plan, no FMA: every Newton step is the reference's arithmetic. The LU uses the
    fixed-pattern plan (zeros kept, diagonal preference 0.1) replayed, where the reference factors afresh each step, so
    the factors round alike only while the pivots and reach order agree; the states agree to 1e-13."""
    net = matpower_case("case118")
    dn = eng.DeviceNetwork(ctx, net, eng.NrOptions(lu_precision=eng.LU_FP64))
    r = dn.solve(count=1)
    rec, vm, va = oracle.solve(net)
    assert np.abs(r.vm[0] - vm).max() < 1e-13 and np.abs(r.va[0] - va).max() < 1e-13


@pytest.fixture(scope="module")
def n1_118(oracle):
    from gridsim_amd import gridgen
    net = gridgen.config1_grid()
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True))
    warm = net.with_start(vm, np.degrees(va))
    outs = gridgen.config2_outages(warm)
    opts = oracle.options(max_iterations=10)
    ref = oracle.screen(warm, outs, opts, threads=8)
    return warm, outs, ref


@pytest.mark.parametrize("lu", ["fp64", "fp32"])
def test_n1_batch_118_matches_oracle(ctx, eng, oracle, n1_118, lu):
    warm, outs, (rres, rvm, rva) = n1_118
    opts = eng.NrOptions(max_iterations=10, lu_precision=eng.LU_FP64 if lu == "fp64" else eng.LU_FP32)
    dn = eng.DeviceNetwork(ctx, warm, opts)
    r = dn.solve(outages=outs)
    print(f"118 N-1 {lu}: {len(outs)} outages, kernel {r.kernel_ms:.2f} ms, statuses {np.bincount(r.status)}, "
          f"refreshes {r.results['symbolic_refreshes'].sum()}", flush=True)
    _compare(oracle, r.results, r.vm, r.va, rres, rvm, rva, f"N-1 118 {lu}")


@pytest.mark.parametrize("lu", ["fp64", "fp32"])
def test_n1_hub_grid_matches_oracle(ctx, eng, oracle, lu):
    """A grid with one Y row of 107 cells (a hub bus tied to 100 others): the mismatch pass's blocks of buses hold
    that row alone or with few others (its cells more than a lane block), in FP64 and FP32 workspaces; every outage's
    verdict and voltages against the oracle."""
    from gridsim_amd import gridgen
    net = gridgen.hub_grid()
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True, max_iterations=60))
    warm = net.with_start(vm, np.degrees(va))
    outs = gridgen.config2_outages(warm)
    rres, rvm, rva = oracle.screen(warm, outs, oracle.options(max_iterations=10), threads=8)
    opts = eng.NrOptions(max_iterations=10, lu_precision=eng.LU_FP64 if lu == "fp64" else eng.LU_FP32,
                         refine_steps=0 if lu == "fp64" else 1)
    dn = eng.DeviceNetwork(ctx, warm, opts)
    r = dn.solve(outages=outs)
    print(f"hub N-1 {lu}: {len(outs)} outages, kernel {r.kernel_ms:.2f} ms", flush=True)
    _compare(oracle, r.results, r.vm, r.va, rres, rvm, rva, f"N-1 hub {lu}")


def test_plain_path_equals_batched_fp64(ctx, eng, n1_118):
    warm, outs, _ = n1_118
    opts = eng.NrOptions(max_iterations=10, lu_precision=eng.LU_FP64)
    dn = eng.DeviceNetwork(ctx, warm, opts)
    a = dn.solve(outages=outs[:12], options=opts)
    b = dn.solve(outages=outs[:12], options=opts, plain=True)
    assert np.array_equal(a.status, b.status)
    assert np.array_equal(a.vm, b.vm) and np.array_equal(a.va, b.va)
    assert b.launches > 12


def test_n1_batch_2000_subset_matches_oracle(ctx, eng, oracle):
    from gridsim_amd import gridgen
    net = gridgen.config2_grid()
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True, max_iterations=60))
    warm = net.with_start(vm, np.degrees(va))
    outs = gridgen.config2_outages(warm)[::10]
    rres, rvm, rva = oracle.screen(warm, outs, oracle.options(max_iterations=10), threads=8)
    dn = eng.DeviceNetwork(ctx, warm, eng.NrOptions(max_iterations=10))
    r = dn.solve(outages=outs)
    print(f"2000 N-1 subset: {len(outs)} outages, kernel {r.kernel_ms:.2f} ms", flush=True)
    _compare(oracle, r.results, r.vm, r.va, rres, rvm, rva, "N-1 2000 subset")


@pytest.mark.parametrize("case", ["case118", "case300"])
@pytest.mark.parametrize("lu", ["fp64", "fp32"])
def test_n1_loop_transformer_outages_match_oracle(ctx, eng, oracle, case, lu):
    """Every non-bridge outage of a MATPOWER grid whose off-nominal transformers sit inside loops (9 of case118's 177
    outages, 40 of case300's 322), warm-started from the base solution: each outaged transformer's tap-scaled stamp
    leaves the instance's Y exactly as the network rebuilt without it, and the verdicts and states match the oracle."""
    from gridsim_amd import gridgen
    net = matpower_case(case)
    rec, vm, va = oracle.solve(net)
    warm = net.with_start(vm, np.degrees(va))
    outs = gridgen.config2_outages(warm)
    tap = warm.br_tap[outs]
    n_xf = int(((tap != 0) & (tap != 1)).sum())
    assert n_xf >= 9
    rres, rvm, rva = oracle.screen(warm, outs, oracle.options(max_iterations=10), threads=8)
    opts = eng.NrOptions(max_iterations=10, lu_precision=eng.LU_FP64 if lu == "fp64" else eng.LU_FP32)
    dn = eng.DeviceNetwork(ctx, warm, opts)
    r = dn.solve(outages=outs)
    print(f"{case} N-1 {lu}: {len(outs)} outages ({n_xf} transformers), statuses {np.bincount(r.status)}",
          flush=True)
    _compare(oracle, r.results, r.vm, r.va, rres, rvm, rva, f"N-1 {case} {lu}")
    if lu == "fp64":
        assert np.array_equal(r.results["iterations"], rres["iterations"])
    dn.close()


def test_two_instances_per_work_group_equal_one(ctx, eng, oracle):
    """Where the solve vector sits in LDS and two instances per work-group keep more of them resident, the batched
    kernel pairs them on one shared LDS solve vector (gs_network_launch_shape says so). The pairing changes no bit:
    an odd batch (the last work-group has one idle wave) gives the same statuses, iterations and voltages as one
    instance per work-group (GS_FLAG_SINGLE_WAVE), and both match the oracle."""
    from gridsim_amd import gridgen
    net = gridgen.config2_grid()
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True, max_iterations=60))
    warm = net.with_start(vm, np.degrees(va))
    outs = gridgen.config2_outages(warm)[::25][:41]
    pair = eng.NrOptions(max_iterations=10, lu_precision=eng.LU_FP64)
    single = eng.NrOptions(max_iterations=10, lu_precision=eng.LU_FP64, flags=eng.FLAG_SINGLE_WAVE)
    dn = eng.DeviceNetwork(ctx, warm, pair)
    sp, ss = dn.launch_shape(pair), dn.launch_shape(single)
    print("launch shapes", sp, ss, flush=True)
    assert sp["x_in_lds"] and sp["waves_per_group"] == 2 and ss["waves_per_group"] == 1
    assert sp["lds_bytes"] < 2 * ss["lds_bytes"]
    fp32 = dn.launch_shape(eng.NrOptions(lu_precision=eng.LU_FP32, refine_steps=1))
    glob = dn.launch_shape(eng.NrOptions(flags=eng.FLAG_X_GLOBAL))
    assert fp32["lu_fp32"] and not glob["x_in_lds"] and glob["waves_per_group"] == 1
    a = dn.solve(outages=outs, options=pair)
    b = dn.solve(outages=outs, options=single)
    assert len(outs) % 2 == 1
    assert np.array_equal(a.status, b.status)
    assert np.array_equal(a.results["iterations"], b.results["iterations"])
    assert np.array_equal(a.vm, b.vm) and np.array_equal(a.va, b.va)
    rres, rvm, rva = oracle.screen(warm, outs, oracle.options(max_iterations=10), threads=8)
    _compare(oracle, a.results, a.vm, a.va, rres, rvm, rva, "paired N-1 2000")
    dn.close()


def test_solve_passes_in_segments_change_no_bit(ctx, eng, oracle, monkeypatch):
    """Plans whose triangular-solve passes run in segments of one step of products (GRIDSIM_SOLVE_SCRATCH_CAP=64)
    give bit-identical batched results to the uncapped plan (the products and each row's subtraction order are the
    same; only where the running values wait between CHAIN steps differs), on the wavefront and level-parallel
    schedules of the 2,000-bus grid."""
    from gridsim_amd import gridgen
    net = gridgen.config2_grid()
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True, max_iterations=60))
    warm = net.with_start(vm, np.degrees(va))
    outs = gridgen.config2_outages(warm)
    for sched, sel in ((0, outs[::80]), (eng.SCHEDULE_LEVELS, outs[:3])):
        opts = eng.NrOptions(max_iterations=10, lu_precision=eng.LU_FP64, schedule=sched)
        monkeypatch.delenv("GRIDSIM_SOLVE_SCRATCH_CAP", raising=False)
        full = eng.DeviceNetwork(ctx, warm, opts)
        monkeypatch.setenv("GRIDSIM_SOLVE_SCRATCH_CAP", "64")
        seg = eng.DeviceNetwork(ctx, warm, opts)
        monkeypatch.delenv("GRIDSIM_SOLVE_SCRATCH_CAP")
        fs, ff = seg.plan_stats(), full.plan_stats()
        assert fs["forward_steps"] + fs["back_steps"] > ff["forward_steps"] + ff["back_steps"]
        a = full.solve(outages=sel, options=opts)
        b = seg.solve(outages=sel, options=opts)
        print(f"schedule {sched}: statuses {np.bincount(a.status)}", flush=True)
        assert np.array_equal(a.status, b.status)
        assert np.array_equal(a.results["iterations"], b.results["iterations"])
        assert np.array_equal(a.vm, b.vm) and np.array_equal(a.va, b.va)
        full.close()
        seg.close()


def test_load_scaled_instances_match_oracle(ctx, eng, oracle, n1_118):
    warm, outs, _ = n1_118
    scales = np.array([0.8, 0.95, 1.1, 1.2])
    dn = eng.DeviceNetwork(ctx, warm, eng.NrOptions(max_iterations=10))
    r = dn.solve(load_scale=scales, outages=outs[:4])
    for i, s in enumerate(scales):
        rec, vm, va = oracle.solve(warm.with_load_multiplier(s), oracle.options(max_iterations=10),
                                   skip=int(outs[i]))
        _compare(oracle, r.results[i:i + 1], r.vm[i:i + 1], r.va[i:i + 1], np.array([rec]), vm[None], va[None],
                 f"scale {s}")


def test_sparse_solve_seam_is_bit_identical(ctx, eng, oracle):
    net = matpower_case("case300")
    t, vm, va, _, _ = oracle.bus_spec(net)
    rp, ci, val = oracle.jacobian(net, vm, va)
    b = np.linspace(-1, 1, len(rp) - 1)
    solver = eng.HipSparseLinearSolver(ctx)
    x = solver.solve(rp, ci, val, b)
    rc, xr = oracle.sparse_solve(rp, ci, val, b)
    assert rc == 0 and np.array_equal(x, xr)
    # second call on the same pattern replays the cached plan: still the reference's bits
    val2 = val * (1 + 1e-6 * np.sin(np.arange(len(val))))
    x2 = solver.solve(rp, ci, val2, b)
    rc, xr2 = oracle.sparse_solve(rp, ci, val2, b)
    assert np.array_equal(x2, xr2)


def test_sparse_solve_seam_errors(ctx, eng):
    solver = eng.HipSparseLinearSolver(ctx)
    rp = np.array([0, 2, 4], np.int32)
    ci = np.array([0, 1, 0, 1], np.int32)
    assert solver.solve(rp, ci, np.array([1.0, 2.0, 2.0, 4.0]), np.ones(2)) is None
    with pytest.raises(ValueError):
        solver.solve(rp, ci, np.array([1.0, np.nan, 2.0, 4.0]), np.ones(2))
    with pytest.raises(ValueError):
        solver.solve(rp, ci, np.array([1.0, 2.0, 3.0, 4.0]), np.array([1.0, np.inf]))
    batch = eng.HipBatchedSparseLinearSolver(ctx)
    errs = [None] * 3
    out = batch.solve_batch([(rp, ci, np.array([1.0, 2.0, 2.0, 4.0])), (rp, ci, np.array([1.0, 2.0, 3.0, 4.0])),
                             (rp, ci, np.array([np.nan, 2.0, 3.0, 4.0]))], [np.ones(2)] * 3, errs)
    assert out[0] is None and errs[0] is None
    np.testing.assert_allclose(out[1], np.linalg.solve([[1, 2], [3, 4]], [1, 1]))
    assert out[2] is None and isinstance(errs[2], ValueError)


@pytest.fixture(scope="module")
def two_networks(oracle):
    """The 118-bus grid and a 300-bus sized grid, each warm-started from its solved base case."""
    from gridsim_amd import gridgen
    out = []
    for net in (gridgen.config1_grid(), gridgen.sized_grid(300, gridgen.BASE_SEED + 3)):
        rec, vm, va = oracle.solve(net, oracle.options(flat_start=True, max_iterations=60))
        assert rec["status"] == 0
        out.append(net.with_start(vm, np.degrees(va)))
    return out


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_of_two_networks_equals_separate_solves(ctx, eng, two_networks, use_graph):
    """A multi-network batch (one stream per group, or one hipGraph replayed twice) gives every instance exactly the
    result of solving its network's instances on their own."""
    from gridsim_amd import gridgen
    rng = np.random.default_rng(7)
    groups, singles = [], []
    for net in two_networks:
        dn = eng.DeviceNetwork(ctx, net, eng.NrOptions(max_iterations=10))
        outs = rng.choice(gridgen.config2_outages(net), 12, replace=False)
        scales = rng.uniform(0.8, 1.2, 12)
        groups.append(eng.Group(dn, outages=outs, load_scale=scales))
        singles.append(dn.solve(outages=outs, load_scale=scales))
    b = eng.Batch(ctx, groups, use_graph=use_graph)
    for rep in range(2):
        ms = b.run()
        print(f"batch (graph={use_graph}) run {rep}: {ms:.3f} ms", flush=True)
        for g, single in enumerate(singles):
            r = b.results(g, keep_state=True)
            assert np.array_equal(r.status, single.status)
            assert np.array_equal(r.results["iterations"], single.results["iterations"])
            assert np.array_equal(r.vm, single.vm) and np.array_equal(r.va, single.va)
    b.close()


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_with_a_level_parallel_group(ctx, eng, two_networks, use_graph):
    """A batch whose second group runs the level-parallel schedule (its Newton steps unrolled into the batch's launches
    or graph) beside a wavefront group: every instance, on two runs, equals the wavefront solve of the same instances
    bit for bit (verdicts, iteration counts, voltages)."""
    from gridsim_amd import gridgen
    rng = np.random.default_rng(13)
    groups, singles = [], []
    for k, net in enumerate(two_networks):
        dn = eng.DeviceNetwork(ctx, net, eng.NrOptions(max_iterations=10))
        outs = rng.choice(gridgen.config2_outages(net), 12, replace=False)
        scales = rng.uniform(0.8, 1.2, 12)
        opts = eng.NrOptions(max_iterations=10, schedule=eng.SCHEDULE_LEVELS if k == 1 else 0)
        groups.append(eng.Group(dn, outages=outs, load_scale=scales, options=opts))
        singles.append(dn.solve(outages=outs, load_scale=scales, options=eng.NrOptions(max_iterations=10)))
    b = eng.Batch(ctx, groups, use_graph=use_graph)
    for rep in range(2):
        ms = b.run()
        print(f"batch with a level-parallel group (graph={use_graph}) run {rep}: {ms:.3f} ms", flush=True)
        for g, single in enumerate(singles):
            r = b.results(g, keep_state=True)
            assert np.array_equal(r.status, single.status)
            assert np.array_equal(r.results["iterations"], single.results["iterations"])
            assert np.array_equal(r.vm, single.vm) and np.array_equal(r.va, single.va)
    b.close()


def test_batch_voltages_of_instances_the_host_finished(ctx, eng, oracle, two_networks):
    """With a growth threshold of 0.9 static pivots fail their check, so instances pause and the host finishes them
    with fresh symbolic passes; those instances meet the oracle's verdicts and converged voltages, and a batch's
    voltages (read from HBM when asked for, the host-finished rows replaced) equal gs_nr_solve_batch's for the same
    instances bit for bit."""
    from gridsim_amd import gridgen
    net = two_networks[1]
    rng = np.random.default_rng(17)
    outs = rng.choice(gridgen.config2_outages(net), 12, replace=False)
    opts = eng.NrOptions(max_iterations=10, lu_precision=eng.LU_FP64, refine_steps=0, growth_tol=0.9)
    dn = eng.DeviceNetwork(ctx, net, opts)
    single = dn.solve(outages=outs, options=opts)
    refreshed = int((single.results["symbolic_refreshes"] > 0).sum())
    print(f"growth 0.9: {refreshed} of {len(outs)} instances finished by the host", flush=True)
    assert refreshed > 0
    ref, rvm, rva = oracle.screen(net, outs, oracle.options(max_iterations=10), threads=4)
    assert np.array_equal(single.status, ref["status"])
    conv = ref["status"] == 0
    assert (single.results["max_mismatch"][conv] < 1e-8).all()
    assert np.abs(single.vm[conv] - rvm[conv]).max() <= TOL_DV and np.abs(single.va[conv] - rva[conv]).max() <= TOL_DV
    b = eng.Batch(ctx, [eng.Group(dn, outages=outs, options=opts)], use_graph=True)
    for rep in range(2):
        b.run()
        r = b.results(0, keep_state=True)
        assert np.array_equal(r.status, single.status)
        assert np.array_equal(r.results["symbolic_refreshes"], single.results["symbolic_refreshes"])
        assert np.array_equal(r.vm, single.vm) and np.array_equal(r.va, single.va)
    b.close()
    dn.close()


def test_autotune_with_the_level_parallel_schedule_offered(ctx, eng, oracle, two_networks):
    """Options asking for GS_SCHEDULE_LEVELS let the autotuner weigh the level-parallel schedule and a work-group of
    waves per instance too; its choice keeps the oracle's verdicts and voltages."""
    from gridsim_amd import gridgen
    net = two_networks[1]
    outs = gridgen.config2_outages(net)[::40][:8]
    rres, rvm, rva = oracle.screen(net, outs, oracle.options(max_iterations=10), threads=8)
    dn = eng.DeviceNetwork(ctx, net, eng.NrOptions(max_iterations=10))
    chosen, ms = eng.autotune(dn, outages=outs, options=eng.NrOptions(max_iterations=10, schedule=eng.SCHEDULE_LEVELS))
    print(f"autotune with levels offered: lu {chosen.lu_precision} refine {chosen.refine_steps} "
          f"schedule {chosen.schedule}, {ms:.3f} ms", flush=True)
    assert chosen.schedule in (0, 256, 512, eng.SCHEDULE_LEVELS, eng.SCHEDULE_GROUP)
    r = dn.solve(outages=outs, options=chosen)
    _compare(oracle, r.results, r.vm, r.va, rres, rvm, rva, "300-bus N-1 sample, autotuned with levels offered")
    dn.close()


def test_autotune_choice_keeps_the_oracle_verdicts(ctx, eng, oracle, n1_118):
    warm, outs, (rres, rvm, rva) = n1_118
    dn = eng.DeviceNetwork(ctx, warm, eng.NrOptions(max_iterations=10))
    chosen, ms = eng.autotune(dn, outages=outs)
    print(f"autotune: lu {chosen.lu_precision} refine {chosen.refine_steps} schedule {chosen.schedule}, "
          f"{ms:.3f} ms", flush=True)
    assert chosen.lu_precision in (eng.LU_FP32, eng.LU_FP64) and 0 <= chosen.refine_steps <= 2
    r = dn.solve(outages=outs, options=chosen)
    _compare(oracle, r.results, r.vm, r.va, rres, rvm, rva, "N-1 118 autotuned")


def test_10k_bus_grid_solve_vector_in_global_memory(ctx, eng, oracle):
    """A network whose triangular-solve vector does not fit in LDS (the kernel keeps it in global memory)."""
    from gridsim_amd import gridgen
    net = gridgen.sized_grid(10000, gridgen.BASE_SEED + 3)
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True, max_iterations=60))
    assert rec["status"] == 0
    warm = net.with_start(vm, np.degrees(va))
    outs = gridgen.config2_outages(warm)[::1500][:6]
    rres, rvm, rva = oracle.screen(warm, outs, oracle.options(max_iterations=10), threads=6)
    dn = eng.DeviceNetwork(ctx, warm, eng.NrOptions(max_iterations=10))
    m = len(dn.plan()["q"])
    assert 4 * m > 64 * 1024, "the solve vector should not fit in LDS for this test"
    r = dn.solve(outages=outs)
    print(f"10k N-1 sample: m={m}, {len(outs)} outages, kernel {r.kernel_ms:.2f} ms", flush=True)
    _compare(oracle, r.results, r.vm, r.va, rres, rvm, rva, "N-1 10k sample")


@pytest.mark.parametrize("lu", ["fp64", "fp32"])
def test_level_parallel_schedule_equals_wavefront_per_instance(ctx, eng, oracle, two_networks, lu):
    """GS_SCHEDULE_LEVELS (one kernel per phase and per elimination level, a work-group per chunk and instance, the
    Newton step replayed as a hipGraph) runs the same arithmetic as one wavefront per instance: identical verdicts,
    iteration counts and voltages, bit for bit; and the oracle's verdicts and voltages to 1e-9."""
    from gridsim_amd import gridgen
    rng = np.random.default_rng(11)
    prec = eng.LU_FP64 if lu == "fp64" else eng.LU_FP32
    for net in two_networks:
        outs = rng.choice(gridgen.config2_outages(net), 10, replace=False)
        scales = rng.uniform(0.8, 1.2, 10)
        wave = eng.NrOptions(max_iterations=10, lu_precision=prec, refine_steps=0 if lu == "fp64" else 2)
        lev = eng.NrOptions(max_iterations=10, lu_precision=prec, refine_steps=wave.refine_steps,
                            schedule=eng.SCHEDULE_LEVELS)
        dn = eng.DeviceNetwork(ctx, net, wave)
        a = dn.solve(outages=outs, load_scale=scales, options=wave)
        b = dn.solve(outages=outs, load_scale=scales, options=lev)
        print(f"{net.n_bus} buses {lu}: wavefront {a.kernel_ms:.2f} ms, levels {b.kernel_ms:.2f} ms "
              f"({b.launches} step replays)", flush=True)
        assert np.array_equal(a.status, b.status)
        assert np.array_equal(a.results["iterations"], b.results["iterations"])
        assert np.array_equal(a.vm, b.vm) and np.array_equal(a.va, b.va)
        for i in range(len(outs)):
            rec, vm, va = oracle.solve(net.with_load_multiplier(float(scales[i])), oracle.options(max_iterations=10),
                                       skip=int(outs[i]))
            assert rec["status"] == b.status[i]
            if rec["status"] == 0:
                assert np.abs(vm - b.vm[i]).max() <= TOL_DV and np.abs(va - b.va[i]).max() <= TOL_DV
        dn.close()


def test_level_parallel_flat_start(ctx, eng, oracle):
    """A flat start on the level-parallel schedule (the 118-bus grid): the oracle's iterations and voltages."""
    from gridsim_amd import gridgen
    net = gridgen.config1_grid()
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True))
    o = eng.NrOptions(flat_start=True, lu_precision=eng.LU_FP64, refine_steps=0, schedule=eng.SCHEDULE_LEVELS)
    dn = eng.DeviceNetwork(ctx, net, o)
    r = dn.solve(count=1, options=o)
    assert r.status[0] == rec["status"] == 0
    assert r.results["iterations"][0] == rec["iterations"]
    assert np.abs(r.vm[0] - vm).max() <= TOL_DV and np.abs(r.va[0] - va).max() <= TOL_DV
    dn.close()


def _check_screen_against_oracle(oracle, net, outs, res, summ, ld, fl, skip_label):
    for j, k in enumerate(outs):
        if res.status[j] != 0:
            continue
        ref = oracle.screen_summary(net, res.vm[j], res.va[j], skip=int(k))
        pf, qf, pt, qt, load = oracle.line_flows(net, res.vm[j], res.va[j], skip=int(k))
        s = summ[j]
        assert abs(s["max_loading_pct"] - ref["max_loading_pct"]) <= 1e-9 * max(1.0, ref["max_loading_pct"]), skip_label
        assert s["overloads"] == ref["overloads"] and s["undervoltage"] == ref["undervoltage"], skip_label
        assert s["overvoltage"] == ref["overvoltage"], skip_label
        assert s["vmin"] == res.vm[j].min() and s["vmax"] == res.vm[j].max()
        assert abs(s["loss_mw"] - ref["loss_mw"]) <= 1e-9 * max(1.0, abs(ref["loss_mw"]))
        assert abs(s["loss_mvar"] - ref["loss_mvar"]) <= 1e-9 * max(1.0, abs(ref["loss_mvar"]))
        if ref["max_loading_pct"] > 0:
            assert s["max_loading_branch"] == int(np.argmax(load))
        scale = max(1.0, np.abs(fl[j]).max())
        np.testing.assert_allclose(fl[j], np.stack([pf, qf, pt, qt], axis=1), rtol=0, atol=1e-12 * scale)
        np.testing.assert_allclose(ld[j], load, rtol=0, atol=1e-12 * max(1.0, load.max()))


def test_screen_case9_matches_oracle_and_the_reference_cli_fixture(ctx, eng, oracle):
    """gs_nr_screen_batch on every case9 branch outage (the bridges island and do not converge): the device's line
    flows, loadings and classification numbers equal the oracle's (computed from the same voltages), and printed as
    the reference's `gridsim nminus1` prints them they reproduce tests/golden/cli/nminus1-case9.txt."""
    from test_oracle import _parse_nminus1
    net = matpower_case("case9")
    o = eng.NrOptions(max_iterations=100, lu_precision=eng.LU_FP64, refine_steps=0)
    dn = eng.DeviceNetwork(ctx, net, o)
    outs = np.arange(net.n_branch)
    res, summ, ld, fl = dn.screen(outages=outs, loading=True, flows=True, keep_state=True)
    _check_screen_against_oracle(oracle, net, outs, res, summ, ld, fl, "case9")
    ids = list(net.bus_ids)
    for label, status, maxload, vmin, vmax in _parse_nminus1(os.path.join(GOLDEN, "cli", "nminus1-case9.txt")):
        a, b = (int(x) for x in label.split("-"))
        k = [i for i in range(net.n_branch) if (ids[net.br_from[i]], ids[net.br_to[i]]) == (a, b)][0]
        if status == "ISLANDED":
            assert res.status[k] != eng.CONVERGED, label
            continue
        s = summ[k]
        assert res.status[k] == eng.CONVERGED
        assert (f"{s['max_loading_pct']:.0f}%", f"{s['vmin']:.3f}", f"{s['vmax']:.3f}") == (maxload, vmin, vmax), label
    dn.close()


def test_screen_synthetic_n1_matches_oracle(ctx, eng, oracle, n1_118):
    """The 118-bus N-1 screen with ratings at 1.05 x each branch's base-case MVA (so outages overload some branches):
    device summaries, loadings and flows against the oracle's on the same voltages."""
    warm, outs, _ = n1_118
    rec, vm, va = oracle.solve(warm)
    pf, qf, pt, qt, _ = oracle.line_flows(warm, vm, va)
    rated = warm.copy(br_rate_a=1.05 * np.maximum(np.hypot(pf, qf), np.hypot(pt, qt)) + 1e-3)
    dn = eng.DeviceNetwork(ctx, rated, eng.NrOptions(max_iterations=10))
    res, summ, ld, fl = dn.screen(outages=outs, loading=True, flows=True, keep_state=True)
    conv = res.status == 0
    assert conv.sum() > 0 and summ["overloads"][conv].sum() > 0
    _check_screen_against_oracle(oracle, rated, outs, res, summ, ld, fl, "118")
    dn.close()


def test_assembled_result_and_contingency_records_case9(ctx, eng, oracle):
    """HipNewtonPowerFlow assembles the reference's PowerFlowResult (injections, losses, line flows) from the device,
    and screen_branches_batched's ContingencyResult records give the reference CLI's case9 verdicts."""
    from test_oracle import _parse_nminus1
    net = matpower_case("case9")
    o = eng.NrOptions(lu_precision=eng.LU_FP64, refine_steps=0)
    r = eng.HipNewtonPowerFlow(ctx, o).solve(net)
    assert r.converged and abs(r.p_inj_mw[0] - 71.64) < 0.05 and abs(r.total_loss_mw - 4.641) < 0.005
    assert abs(r.total_loss_mw - r.p_inj_mw.sum()) < 1e-9
    pf, qf, pt, qt, load = oracle.line_flows(net, r.vm, np.radians(r.va_deg))
    got = np.array([[f.p_from_mw, f.q_from_mvar, f.p_to_mw, f.q_to_mvar, f.loading_pct] for f in r.line_flows])
    np.testing.assert_allclose(got, np.stack([pf, qf, pt, qt, load], axis=1), rtol=0, atol=1e-9)
    res, recs = eng.screen_branches_batched(ctx, net, options=eng.NrOptions(max_iterations=100,
                                                                            lu_precision=eng.LU_FP64, refine_steps=0))
    by_label = {c.outage: c for c in recs}
    for label, status, maxload, vmin, vmax in _parse_nminus1(os.path.join(GOLDEN, "cli", "nminus1-case9.txt")):
        c = by_label[label]
        if status == "ISLANDED":
            assert not c.converged and c.islanded, label
        elif status == "secure":
            assert c.secure, (label, c)
        else:
            assert len(c.violations) == int(status[len("VIOLATIONx"):]), (label, c)
            assert f"{c.max_branch_loading_pct:.0f}%" == maxload


def test_solve_export_gpu_against_cpu_the_reference_comparer_way(ctx, eng, oracle):
    """The reference's CPU-vs-GPU check (tools/run-gpu.sh:27-50) on case118 and case300: the device's solve export
    against the oracle's, same bus set, max |dVm| <= 1e-6 pu and max |dVa| <= 1e-4 deg (they agree far closer)."""
    from gridsim_amd import export
    from test_export import _oracle_result
    for name in ("case118", "case300"):
        net = matpower_case(name)
        gpu = eng.HipNewtonPowerFlow(ctx, eng.NrOptions(lu_precision=eng.LU_FP64, refine_steps=0)).solve(net)
        a = export.solve_export(net, _oracle_result(oracle, net), case=name, backend="cpu (oracle)")
        b = export.solve_export(net, gpu, case=name, backend="gpu (gfx950)")
        ok, dvm, dva = export.compare_exports(a, b)
        print(f"COMPARE {name}: max|dVm| = {dvm:.3g} pu, max|dVa| = {dva:.3g} deg", flush=True)
        assert ok and dvm <= 1e-9 and dva <= 1e-7
        assert abs(a["result"]["totalLossMw"] - b["result"]["totalLossMw"]) <= 1e-6


def test_reactive_limits_case14_n1_matches_oracle_and_the_reference_cli_fixture(ctx, eng, oracle):
    """GS_FLAG_Q_LIMITS on case14's N-1 screen (where generators hit their limits): every outage's verdict, total
    iterations and voltages equal the oracle's reactive-limit loop (|dV| <= 1e-9), and the screen prints the
    reference's nminus1-case14 rows (1-2 NO SOLUTION; Vmin / Vmax to 3 decimals)."""
    from test_oracle import _parse_nminus1
    net = matpower_case("case14")
    o = eng.NrOptions(max_iterations=100, lu_precision=eng.LU_FP64, refine_steps=0, flags=eng.FLAG_Q_LIMITS)
    dn = eng.DeviceNetwork(ctx, net, o)
    outs = np.arange(net.n_branch)
    res, summ, _, _ = dn.screen(outages=outs, keep_state=True)
    for k in outs:
        rec, vm, va = oracle.solve(net, oracle.options(max_iterations=100, q_limits=True), skip=int(k))
        assert rec["status"] == res.status[k], k
        assert rec["iterations"] == res.results["iterations"][k], k
        if rec["status"] == 0:
            assert np.abs(vm - res.vm[k]).max() <= TOL_DV and np.abs(va - res.va[k]).max() <= TOL_DV, k
            s = oracle.screen_summary(net, vm, va, skip=int(k))
            assert abs(summ[k]["vmin"] - s["vmin"]) <= TOL_DV and abs(summ[k]["loss_mw"] - s["loss_mw"]) <= 1e-6
    ids = list(net.bus_ids)
    for label, status, maxload, vmin, vmax in _parse_nminus1(os.path.join(GOLDEN, "cli", "nminus1-case14.txt")):
        a, b = (int(x) for x in label.split("-"))
        k = [i for i in range(net.n_branch) if (ids[net.br_from[i]], ids[net.br_to[i]]) == (a, b)][0]
        if status == "NO SOLUTION":
            assert res.status[k] != eng.CONVERGED, label
            continue
        assert (f"{summ[k]['vmin']:.3f}", f"{summ[k]['vmax']:.3f}") == (vmin, vmax), label
    dn.close()


def test_iterative_fallback_when_the_lu_refuses_the_step(ctx, eng, oracle):
    """case14's outage of 7-8 cuts the synchronous condenser at bus 8 off: the Jacobian is singular, the LU refuses
    the step and NewtonShell.SolveStep's ILU(0) + BiCGSTAB fallback takes it (on the device). Every outage's verdict,
    iterations, step failures and voltages (|dV| <= 1e-9) equal the oracle's, and 7-8 does take the fallback."""
    net = matpower_case("case14")
    o = eng.NrOptions(max_iterations=100, lu_precision=eng.LU_FP64, refine_steps=0)
    dn = eng.DeviceNetwork(ctx, net, o)
    outs = np.arange(net.n_branch)
    r = dn.solve(outages=outs)
    fallback = 0
    for k in outs:
        rec, vm, va = oracle.solve(net, oracle.options(max_iterations=100), skip=int(k))
        print(f"outage {k}: oracle {rec['status']} / {rec['iterations']} / {rec['step_failures']}, gpu "
              f"{r.status[k]} / {r.results['iterations'][k]} / {r.results['step_failures'][k]}", flush=True)
        assert rec["status"] == r.status[k], k
        assert rec["iterations"] == r.results["iterations"][k], k
        assert rec["step_failures"] == r.results["step_failures"][k], k
        if rec["status"] == 0:
            assert np.abs(vm - r.vm[k]).max() <= TOL_DV and np.abs(va - r.va[k]).max() <= TOL_DV, k
        fallback += int(rec["step_failures"] > 0)
    assert fallback > 0
    dn.close()


def test_reactive_limits_synthetic_batch_matches_oracle(ctx, eng, oracle, n1_118):
    """The 118-bus N-1 batch with every generator's reactive range cut to +-60 % of its base-case output (so outages
    switch buses to PQ, in different patterns per outage): verdicts, iterations and voltages equal the oracle's."""
    warm, outs, _ = n1_118
    rec, vm, va = oracle.solve(warm)
    p, q, _ = oracle.mismatch(warm, vm, va)
    qgen = q[warm.gen_bus] * warm.base_mva + warm.qd[warm.gen_bus]
    lim = 0.6 * np.abs(qgen) + 5.0
    tight = warm.copy(gen_qmin=-lim, gen_qmax=lim)
    o = eng.NrOptions(max_iterations=30, lu_precision=eng.LU_FP64, refine_steps=0, flags=eng.FLAG_Q_LIMITS)
    dn = eng.DeviceNetwork(ctx, tight, o)
    r = dn.solve(outages=outs, options=o)
    switched = 0
    for j, k in enumerate(outs):
        rec, vm, va = oracle.solve(tight, oracle.options(max_iterations=30, q_limits=True), skip=int(k))
        assert rec["status"] == r.status[j], k
        assert rec["iterations"] == r.results["iterations"][j], k
        if rec["status"] == 0:
            assert np.abs(vm - r.vm[j]).max() <= TOL_DV and np.abs(va - r.va[j]).max() <= TOL_DV, k
        plain, _, _ = oracle.solve(tight, oracle.options(max_iterations=30), skip=int(k))
        switched += int(plain["iterations"] != rec["iterations"])
    print(f"{switched} of {len(outs)} outages ran limit passes", flush=True)
    assert switched > 0
    dn.close()


def test_sparse_solve_batch_launches_shared_structures_together(ctx, eng, oracle):
    """gs_sparse_solve_batch: systems of one structure with a cached plan go to the device in one launch; a new
    structure, a singular system and a non-finite right-hand side take the single-system path. Every answer equals
    gs_sparse_solve's and the oracle's, bit for bit, and the statuses are the single calls' statuses."""
    net = matpower_case("case118")
    t, vm, va, _, _ = oracle.bus_spec(net)
    rp, ci, val = oracle.jacobian(net, vm, va)
    rng = np.random.default_rng(5)
    single = eng.HipSparseLinearSolver(ctx)
    single.solve(rp, ci, val, np.ones(len(rp) - 1))   # caches the plan of this structure
    mats, rhs = [], []
    for k in range(12):
        mats.append((rp, ci, val * (1 + 1e-7 * rng.standard_normal(len(val)))))
        rhs.append(rng.standard_normal(len(rp) - 1))
    rp9, ci9, val9 = oracle.jacobian(matpower_case("case9"), *oracle.bus_spec(matpower_case("case9"))[1:3])
    mats.append((rp9, ci9, val9))                                # a structure with no cached plan yet
    rhs.append(rng.standard_normal(len(rp9) - 1))
    r2 = np.array([0, 2, 4], np.int32)
    c2 = np.array([0, 1, 0, 1], np.int32)
    mats.append((r2, c2, np.array([1.0, 2.0, 2.0, 4.0])))        # singular
    rhs.append(np.ones(2))
    mats.append((rp, ci, val))                                   # non-finite right-hand side
    bad = np.ones(len(rp) - 1)
    bad[3] = np.inf
    rhs.append(bad)
    errs = [None] * len(mats)
    out = eng.HipBatchedSparseLinearSolver(ctx).solve_batch(mats, rhs, errs)
    for k, ((r, c, v), b) in enumerate(zip(mats, rhs)):
        if k == len(mats) - 1:
            assert out[k] is None and isinstance(errs[k], ValueError)
            continue
        rc, xr = oracle.sparse_solve(r, c, v, b)
        if rc != 0:
            assert out[k] is None and errs[k] is None, k
            continue
        assert out[k] is not None and np.array_equal(out[k], xr), k
        assert np.array_equal(single.solve(r, c, v, b), xr), k
