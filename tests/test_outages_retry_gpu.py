"""Outage sets (N-2 / N-k), the N-2 screen, the DC-angle start retry and the sparse-solver seam's plan cache on the
device, against the oracle (and the reference's recorded N-2 CLI output), through the C ABI."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, matpower_case
from test_nk_screen import parse_nk_fixture
from test_oracle_retry import TWO_ISLANDS_OVERLOADED, chain

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


@pytest.fixture(scope="module")
def warm118(oracle):
    from gridsim_amd import gridgen
    net = gridgen.config1_grid()
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True))
    return net.with_start(vm, np.degrees(va))


def _sets(net, rng):
    """40 random pairs of non-bridge branches, 10 triples and 4 sets of six (more Y cells than an instance patches in
    place: they run on Y values of their own)."""
    from gridsim_amd import gridgen
    cand = gridgen.config2_outages(net)
    sets = [list(rng.choice(cand, 2, replace=False)) for _ in range(40)]
    sets += [list(rng.choice(cand, 3, replace=False)) for _ in range(10)]
    sets += [list(rng.choice(cand, 6, replace=False)) for _ in range(4)]
    return [[int(b) for b in s] for s in sets]


@pytest.mark.parametrize("lu", ["fp64", "fp32"])
def test_outage_sets_match_the_oracle(ctx, eng, oracle, warm118, lu):
    rng = np.random.default_rng(29)
    sets = _sets(warm118, rng)
    o = eng.NrOptions(max_iterations=10, lu_precision=eng.LU_FP64 if lu == "fp64" else eng.LU_FP32,
                      refine_steps=0 if lu == "fp64" else 1)
    dn = eng.DeviceNetwork(ctx, warm118, o)
    r = dn.solve(outages=sets)
    ref, rvm, rva = oracle.screen(warm118, sets, oracle.options(max_iterations=10), threads=8)
    print(f"N-k sets ({lu}): {int((ref['status'] == 0).sum())} of {len(sets)} converge", flush=True)
    assert np.array_equal(r.status, ref["status"])
    assert np.array_equal(r.results["iterations"], ref["iterations"])
    conv = ref["status"] == 0
    assert (r.results["max_mismatch"][conv] < 1e-8).all()
    assert np.abs(r.vm[conv] - rvm[conv]).max() <= TOL_DV and np.abs(r.va[conv] - rva[conv]).max() <= TOL_DV
    dn.close()


def test_outage_set_loadings_match_the_oracle(ctx, eng, oracle, warm118):
    rng = np.random.default_rng(31)
    sets = _sets(warm118, rng)[:30]
    dn = eng.DeviceNetwork(ctx, warm118, eng.NrOptions(max_iterations=10, lu_precision=eng.LU_FP64, refine_steps=0))
    res, summ, ld, fl = dn.screen(outages=sets, loading=True, flows=True, keep_state=True)
    for j, st in enumerate(sets):
        if res.status[j] != 0:
            continue
        pf, qf, pt, qt, load = oracle.line_flows(warm118, res.vm[j], res.va[j], skip=st)
        scale = max(1.0, np.abs(fl[j]).max())
        np.testing.assert_allclose(fl[j], np.stack([pf, qf, pt, qt], axis=1), rtol=0, atol=1e-12 * scale)
        np.testing.assert_allclose(ld[j], load, rtol=0, atol=1e-12 * max(1.0, load.max()))
        assert (fl[j][st] == 0).all()
        s = oracle.screen_summary(warm118, res.vm[j], res.va[j], skip=st)
        assert abs(summ[j]["vmin"] - s["vmin"]) == 0 and abs(summ[j]["vmax"] - s["vmax"]) == 0
    dn.close()


def test_nminus2_case9_matches_the_reference_cli_fixture(ctx, eng):
    """`gridsim nminus2 case9` (nminus2-case9.txt): N-1 first, the six single outages that solve ranked by stress, the
    15 pairs among them screened on the device (each an outage list of one batched solve): all islanded, shown worst
    first, and the three single outages left out of the pairing."""
    net = matpower_case("case9")
    rep = eng.screen_n_minus_k(ctx, net, 2, eng.NrOptions(max_iterations=100, lu_precision=eng.LU_FP64,
                                                          refine_steps=0))
    rows, counts, excluded = parse_nk_fixture(os.path.join(GOLDEN, "cli", "nminus2-case9.txt"))
    assert len(rep.results) == counts["screened"]
    assert sum(c.islanded for c in rep.results) == counts["islanded"]
    assert sum(c.secure for c in rep.results) == counts["secure"]
    for (label, status), c in zip(rows, rep.results):
        assert c.outage == label and status == "ISLANDED" and c.islanded
    assert [(c.outage, c.violations[0]) for c in rep.excluded] == excluded


def _flat(eng, cap, retry=True):
    return eng.NrOptions(flat_start=True, max_iterations=cap, lu_precision=eng.LU_FP64, refine_steps=0,
                         flags=eng.FLAG_DC_RETRY if retry else 0)


def test_dc_start_retry_converges_the_stalled_chain(ctx, eng, oracle):
    """The reference's 8-bus chain at 60 degrees a section: 21 Newton steps from flat start; at a cap of 10 the
    ordinary solve stalls and the retry from the DC angles converges, as the oracle's (and the reference's) does."""
    net = chain(8, 60.0)
    dn = eng.DeviceNetwork(ctx, net, _flat(eng, 50))
    r50 = dn.solve(count=1, options=_flat(eng, 50))
    assert r50.status[0] == eng.CONVERGED and r50.results["iterations"][0] == 21
    assert r50.results["dc_retry"][0] == eng.RETRY_NOT_TRIED
    off = dn.solve(count=1, options=_flat(eng, 10, retry=False))
    assert off.status[0] != eng.CONVERGED and off.results["dc_retry"][0] == eng.RETRY_NOT_TRIED
    r = dn.solve(count=1, options=_flat(eng, 10))
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True, max_iterations=10, retry_dc_start=True))
    print(f"chain retry: device {r.results['iterations'][0]} steps, oracle {rec['iterations']}", flush=True)
    assert r.status[0] == eng.CONVERGED and r.results["dc_retry"][0] == eng.RETRY_CONVERGED
    assert rec["dc_retry"] == oracle.RETRY_CONVERGED and r.results["iterations"][0] == rec["iterations"]
    assert r.results["max_mismatch"][0] < 1e-8
    assert np.abs(r.vm[0] - vm).max() <= TOL_DV and np.abs(r.va[0] - va).max() <= TOL_DV
    dn.close()


@pytest.mark.parametrize("case", ["case9", "case14", "case118"])
def test_dc_start_retry_skips_a_run_stopped_one_step_short(ctx, eng, oracle, case):
    net = matpower_case(case)
    dn = eng.DeviceNetwork(ctx, net, _flat(eng, 50))
    k = int(dn.solve(count=1, options=_flat(eng, 50)).results["iterations"][0])
    off = dn.solve(count=1, options=_flat(eng, k - 1, retry=False))
    r = dn.solve(count=1, options=_flat(eng, k - 1))
    assert r.status[0] == eng.MAX_ITERATIONS and r.results["iterations"][0] == k - 1
    assert r.results["dc_retry"][0] == eng.RETRY_NOT_TRIED
    assert np.array_equal(r.vm, off.vm) and np.array_equal(r.va, off.va)
    one = dn.solve(count=1, options=_flat(eng, 1))   # a cap of 1 is never retried
    assert one.results["dc_retry"][0] == eng.RETRY_NOT_TRIED
    dn.close()


def test_dc_start_retry_unavailable_without_a_single_slack_island(ctx, eng, oracle):
    from gridsim_amd.matpower import read_text
    net = read_text(TWO_ISLANDS_OVERLOADED)
    dn = eng.DeviceNetwork(ctx, net, _flat(eng, 50))
    r = dn.solve(count=1, options=_flat(eng, 50))
    off = dn.solve(count=1, options=_flat(eng, 50, retry=False))
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True, max_iterations=50, retry_dc_start=True))
    assert r.status[0] == rec["status"] != eng.CONVERGED
    assert r.results["dc_retry"][0] == eng.RETRY_UNAVAILABLE == rec["dc_retry"]
    assert np.array_equal(r.vm, off.vm) and np.array_equal(r.va, off.va)
    dn.close()


def test_dc_start_retry_in_a_screen_batch(ctx, eng, oracle):
    """The retry inside gs_nr_screen_batch beside instances that converge ordinarily: the retried instance's state is
    what the screen classifies (its flows equal the oracle's on the oracle's retried state)."""
    net = chain(8, 60.0)
    dn = eng.DeviceNetwork(ctx, net, _flat(eng, 10))
    res, summ, ld, fl = dn.screen(count=3, load_scale=[1.0, 0.2, 1.0], options=_flat(eng, 10), loading=True,
                                  flows=True, keep_state=True)
    for j, sc in enumerate([1.0, 0.2, 1.0]):
        rec, vm, va = oracle.solve(net.with_load_multiplier(sc), oracle.options(flat_start=True, max_iterations=10,
                                                                                retry_dc_start=True))
        assert res.status[j] == rec["status"] == 0 and res.results["dc_retry"][j] == rec["dc_retry"]
        assert np.abs(res.vm[j] - vm).max() <= TOL_DV and np.abs(res.va[j] - va).max() <= TOL_DV
        pf, qf, pt, qt, _ = oracle.line_flows(net, vm, va)
        np.testing.assert_allclose(fl[j][:, 0], pf, rtol=0, atol=1e-6)
    dn.close()


def test_seam_plans_are_keyed_by_the_whole_structure(ctx, eng, oracle):
    """Systems with the same number of stored entries but different patterns (and sizes), solved alternately through
    gs_sparse_solve, each equal the oracle's CpuSparseLuSolver bit for bit: no plan is reused across structures."""
    from gridsim_amd import gridgen
    solver = eng.HipSparseLinearSolver(ctx)
    systems = [
        (np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32), np.array([4.0, 1.0, 2.0, 3.0])),
        (np.array([0, 1, 2, 3, 4], np.int32), np.array([0, 1, 2, 3], np.int32), np.array([2.0, 3.0, 5.0, 7.0])),
    ]
    net = matpower_case("case14")
    rec, vm, va = oracle.solve(net)
    outs = gridgen.config2_outages(net)
    shapes = {}
    for k in outs:
        rp, ci, val = oracle.jacobian(net, vm, va, skip=int(k))[:3]
        shapes.setdefault(len(val), []).append((rp, ci, val))
    same = next(v for v in shapes.values() if len(v) >= 2 and not np.array_equal(v[0][1], v[1][1]))
    systems += same[:2]
    rng = np.random.default_rng(3)
    for rep in range(2):
        for rp, ci, val in systems:
            n = len(rp) - 1
            b = rng.standard_normal(n)
            x = solver.solve(rp, ci, val, b)
            rc, xr = oracle.sparse_solve(rp, ci, val, b)
            assert rc == 0 and x is not None and np.array_equal(x, xr)
