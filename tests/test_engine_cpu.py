"""CPU-side checks of libgridsim_hip.so: the C ABI is complete, the host half of the library (FDLIBM, the symbolic
plan) matches the oracle bit for bit. No GPU needed."""
import ctypes as C
import os
import re
from dataclasses import replace

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, matpower_case
import strictmath


@pytest.fixture(scope="module")
def eng():
    from gridsim_amd import engine
    engine.lib()
    return engine


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "gridsim_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"GS_API\s+[\w\s\*]+?\b(gs_\w+)\s*\(", text)


def test_library_exports_every_declared_symbol(eng):
    names = _declared_symbols()
    assert len(names) >= 20
    lib = C.CDLL(eng.LIB_PATH)
    for name in names:
        assert hasattr(lib, name), name


def test_struct_sizes_match_the_header(eng):
    """The ctypes mirrors of the C structs have the C layout (the P/Invoke shape test of RocmInteropShapeTests)."""
    assert C.sizeof(eng.GsNrOptions) == 48
    assert C.sizeof(eng.GsInstance) == 40
    assert eng.RESULT_DTYPE.itemsize == 72
    assert C.sizeof(eng.GsCsr) == 32
    assert C.sizeof(eng.GsNetwork) == 4 * 3 + 4 + 8 + 24 * 8
    assert eng.SUMMARY_DTYPE.itemsize == 56
    assert C.sizeof(eng.GsServerConfig) == 32 and C.sizeof(eng.GsServerStats) == 6 * 8 + 5 * 8 + 2 * 8
    assert C.sizeof(eng.GsPlanStats) == 10 * 4 + 8 + 9 * 4 + 4


def test_no_device_status_without_a_gpu(eng):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert eng.lib().gs_init(0, 0, C.byref(h)) == eng.GS_NO_DEVICE


def test_host_fdlibm_matches_strictmath(eng):
    vec, dig = strictmath.load(os.path.join(GOLDEN, "strictmath_sincos.txt"))
    xs = np.array([strictmath.from_bits(v[1]) for v in vec])
    s, c = eng.host_sincos(xs)
    got = np.where([v[0] == "sin" for v in vec], s, c)
    want = np.array([strictmath.from_bits(v[2]) for v in vec])
    assert np.array_equal(got.view(np.uint64)[~np.isnan(want)], want.view(np.uint64)[~np.isnan(want)])
    assert np.all(np.isnan(got[np.isnan(want)]))
    for fn, cls, seed, count, d in dig:
        x = strictmath.draw(cls, seed, count)
        s, c = eng.host_sincos(x)
        assert strictmath.digest(s if fn == "sin" else c) == d, (fn, cls)


def _oracle_plan(oracle, net, flat):
    t, vm, va, _, _ = oracle.bus_spec(net, flat_start=flat)
    rp, ci, val = oracle.jacobian(net, vm, va)
    q = oracle.md_order(rp, ci)
    lu = oracle.lu_factor(rp, ci, val, q, keep_zeros=True, diagonal_preference=0.1)
    return rp, ci, lu


@pytest.mark.parametrize("case,flat", [("case9", True), ("case14", False), ("case118", True), ("case300", True),
                                       ("synthetic-118", True), ("synthetic-2000", True)])
def test_product_plan_equals_oracle_plan(eng, oracle, case, flat):
    """The engine's symbolic pass (minimum degree + Gilbert-Peierls pivot sequence, zeros kept, diagonal preference
    0.1) reproduces the oracle's integer arrays exactly."""
    from gridsim_amd import gridgen
    if case == "synthetic-118":
        net = gridgen.config1_grid()
    elif case == "synthetic-2000":
        net = gridgen.config2_grid()
    else:
        net = matpower_case(case)
    plan = eng.HostPlan(net, eng.NrOptions(flat_start=flat))
    rp, ci, lu = _oracle_plan(oracle, net, flat)
    np.testing.assert_array_equal(plan.j_rowptr, rp)
    np.testing.assert_array_equal(plan.j_colidx, ci)
    np.testing.assert_array_equal(plan.q, lu.q)
    np.testing.assert_array_equal(plan.pivot_row, lu.pivot_row)
    np.testing.assert_array_equal(plan.lp, lu.lp)
    np.testing.assert_array_equal(plan.li, lu.li)
    np.testing.assert_array_equal(plan.up, lu.up)
    np.testing.assert_array_equal(plan.ui, lu.ui)


@pytest.mark.parametrize("case", ["case9", "case118", "case300", "synthetic-2000", "synthetic-10000", "hub"])
def test_device_lu_indexing_emulated_on_host(eng, case):
    """The device refactor (level chunks, 64-record rows) and level-scheduled solves, emulated on the host with every
    index checked against its buffer, solve J dx = f of the start point to FP64 accuracy."""
    from gridsim_amd import gridgen
    if case == "synthetic-2000":
        net = gridgen.config2_grid()
    elif case == "synthetic-10000":
        net = gridgen.sized_grid(10000, gridgen.BASE_SEED + 3)
    elif case == "hub":
        net = gridgen.hub_grid()
    else:
        net = matpower_case(case)
    plan = eng.HostPlan(net, eng.NrOptions(flat_start=True))
    res = C.c_double()
    why = C.create_string_buffer(128)
    st = eng.lib().gs_plan_selfcheck(plan.h, C.byref(res), why)
    assert st == eng.GS_OK, why.value.decode()
    assert res.value < 1e-12


@pytest.mark.parametrize("case", ["case118", "synthetic-2000"])
def test_device_plan_covers_every_update_and_solve_entry_once(eng, case):
    """Closed forms of the compiled streams' work: the refactor's UPDATE lanes are exactly the Gilbert-Peierls updates
    (for each column k and each off-diagonal U(j, k), one per off-diagonal L(:, j) entry), the forward and back
    PRODUCT lanes exactly the off-diagonal entries of L and U, and the row kinds add up to the stream."""
    from gridsim_amd import gridgen
    net = gridgen.config2_grid() if case == "synthetic-2000" else matpower_case(case)
    plan = eng.HostPlan(net, eng.NrOptions(flat_start=True))
    st = plan.stats()
    lp, up, ui, m = plan.lp, plan.up, plan.ui, plan.m
    updates = sum(int(lp[j + 1] - lp[j] - 1) for k in range(m) for j in ui[up[k]:up[k + 1] - 1])
    assert st["update_lanes"] == updates
    assert st["forward_lanes"] == plan.nnz_l - m and st["back_lanes"] == plan.nnz_u - m
    kinds = st["rows_by_kind"]
    assert sum(kinds) <= st["refactor_rows"] < sum(kinds) + 16   # padding: at most one ring of no-op rows
    assert kinds[3] >= plan.n_levels                               # at least one PIVOT row (chunk) per level
    print(case, st, flush=True)


@pytest.mark.parametrize("case", ["case300", "synthetic-2000"])
def test_solve_passes_in_segments_emulated_on_host(eng, case, monkeypatch):
    """With the solves' scratch area capped at one step of products (GRIDSIM_SOLVE_SCRATCH_CAP=64), passes with more
    entries run in several segments, each lane's running value carried from CHAIN to CHAIN: the host emulation still
    checks every index and solves J dx = f to FP64 accuracy, and the streams hold more steps than uncapped."""
    from gridsim_amd import gridgen
    net = gridgen.config2_grid() if case == "synthetic-2000" else matpower_case(case)
    full = eng.HostPlan(net, eng.NrOptions(flat_start=True)).stats()
    monkeypatch.setenv("GRIDSIM_SOLVE_SCRATCH_CAP", "64")
    plan = eng.HostPlan(net, eng.NrOptions(flat_start=True))
    st = plan.stats()
    assert st["forward_lanes"] == full["forward_lanes"] and st["back_lanes"] == full["back_lanes"]
    assert st["forward_steps"] + st["back_steps"] > full["forward_steps"] + full["back_steps"]
    res = C.c_double()
    why = C.create_string_buffer(128)
    assert eng.lib().gs_plan_selfcheck(plan.h, C.byref(res), why) == eng.GS_OK, why.value.decode()
    assert res.value < 1e-12


def test_sine_is_odd_bit_for_bit(eng):
    """The mismatch pass forms a transposed cell's Jacobian values with sin(va_j - va_i) taken as 0.0 - sin(va_i - va_j)
    (and, for a zero difference, as that difference itself): FDLIBM's sine is odd and its cosine even bit for bit, and
    IEEE subtraction is antisymmetric, including at huge arguments (Payne-Hanek reduction) and subnormals."""
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-4, 4, 20000), rng.uniform(-1e6, 1e6, 5000), 10.0 ** rng.uniform(-320, 300, 5000),
                        [5e-324, 1e-300, np.pi, np.pi / 2, np.pi / 4, 1e22, 1.7e308]])
    s1, c1 = eng.host_sincos(x)
    s2, c2 = eng.host_sincos(-x)
    assert np.array_equal((0.0 - s1).view(np.uint64), s2.view(np.uint64))
    assert np.array_equal(c1.view(np.uint64), c2.view(np.uint64))
    a, b = rng.uniform(-1, 1, 20000), rng.uniform(-1, 1, 20000)
    assert np.array_equal((a - b).view(np.uint64), (0.0 - (b - a)).view(np.uint64))
    z, _ = eng.host_sincos(np.array([0.0, -0.0]))
    assert z.view(np.uint64).tolist() == np.array([0.0, -0.0]).view(np.uint64).tolist()


def test_jacobian_transposed_only_for_a_bitwise_symmetric_ybus(eng):
    """Cell lanes form their transposes' values only when every off-diagonal Y cell's stamps equal its transpose's bit
    for bit: the synthetic grids (r > 0, shift 0) are; one phase shifter, or a reactance-only tapped branch (whose
    conductance stamps are +0 one way and -0 the other), is not."""
    from gridsim_amd import gridgen
    net = gridgen.config1_grid()
    assert eng.HostPlan(net).stats()["jacobian_transposed"] == 1
    shifted = net.br_shift_deg.copy()
    shifted[len(shifted) // 2] = 5.0
    assert eng.HostPlan(replace(net, br_shift_deg=shifted)).stats()["jacobian_transposed"] == 0
    r, tap = net.br_r.copy(), net.br_tap.copy()
    r[3], tap[3] = 0.0, 0.98
    assert eng.HostPlan(replace(net, br_r=r, br_tap=tap)).stats()["jacobian_transposed"] == 0
    assert eng.HostPlan(matpower_case("case9")).stats()["jacobian_transposed"] == 0


@pytest.mark.parametrize("transposed", [True, False])
def test_transposed_fill_and_lu_of_a_warm_grid_emulated_on_host(eng, oracle, transposed):
    """At a warm start point (nonzero angle differences) the self-check forms every lane's Jacobian values as the
    mismatch kernel does and compares them bit for bit with each cell's own (transposed layout), then runs the device
    LU emulation on them; with a phase shifter the grid keeps the untransposed layout and passes the same way."""
    from gridsim_amd import gridgen
    net = gridgen.config1_grid()
    if not transposed:
        shifted = net.br_shift_deg.copy()
        shifted[7] = -3.0
        net = replace(net, br_shift_deg=shifted)
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True))
    warm = net.with_start(vm, np.degrees(va))
    plan = eng.HostPlan(warm, eng.NrOptions(flat_start=False))
    assert plan.stats()["jacobian_transposed"] == int(transposed)
    res = C.c_double()
    why = C.create_string_buffer(128)
    assert eng.lib().gs_plan_selfcheck(plan.h, C.byref(res), why) == eng.GS_OK, why.value.decode()
    assert res.value < 1e-12


def test_plan_of_warm_started_outage_grid(eng, oracle):
    """The N-1 grid's plan is taken at the base solution (warm start), as the screening run uses it."""
    from gridsim_amd import gridgen
    net = gridgen.config1_grid()
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True))
    warm = net.with_start(vm, np.degrees(va))
    plan = eng.HostPlan(warm, eng.NrOptions(flat_start=False))
    rp, ci, lu = _oracle_plan(oracle, warm, False)
    np.testing.assert_array_equal(plan.pivot_row, lu.pivot_row)
    np.testing.assert_array_equal(plan.li, lu.li)


def test_instance_records_filled_as_columns(eng):
    """The gs_instance array the wrapper hands to the library: outage, load scale and start pointers per record, with
    the numpy view's field layout equal to the C struct's."""
    assert eng.INSTANCE_DTYPE.itemsize == C.sizeof(eng.GsInstance)
    for name, _ in eng.GsInstance._fields_:
        assert eng.INSTANCE_DTYPE.fields[name][1] == getattr(eng.GsInstance, name).offset, name

    class Unbound(eng.DeviceNetwork):
        def __init__(self):
            pass

    dn = Unbound()
    arr, count, keep = dn._instances(outages=np.arange(1995))
    assert count == 1995 and arr[1994].outage_branch == 1994 and arr[7].load_scale == 1.0 and arr[7].vm0 is None
    rng = np.random.default_rng(5)
    vm, va, ls = rng.random((7, 11)), rng.random((7, 11)), np.linspace(0.8, 1.2, 7)
    arr, count, keep = dn._instances(load_scale=ls, starts=(vm, va))
    for i in range(count):
        assert arr[i].outage_branch == -1 and arr[i].load_scale == ls[i]
        assert np.array_equal(np.ctypeslib.as_array((C.c_double * 11).from_address(arr[i].vm0)), vm[i])
        assert np.array_equal(np.ctypeslib.as_array((C.c_double * 11).from_address(arr[i].va0)), va[i])
    arr, count, keep = dn._instances(outages=[3, 4], starts=([vm[0], vm[1]], [va[0], va[1]]))
    assert arr[1].outage_branch == 4
    assert np.array_equal(np.ctypeslib.as_array((C.c_double * 11).from_address(arr[1].va0)), va[1])
    assert dn._instances(outages=[], starts=([], []))[1] == 0


def test_fresh_symbolic_passes_are_thread_independent_and_share_orders(eng, oracle):
    """The host's fresh symbolic passes of paused instances (gs_plan_fresh_batch) give the same plans on 1 and on 8
    threads, take each instance's order from the pattern cache when its outage set repeats (the order depends on the
    pattern alone), and equal the oracle's factorisation of the same reduced network at the same state (ToCsc keeping
    zeros, diagonal preference 0.1): same order and pivot rows."""
    import ctypes as C
    from gridsim_amd import gridgen
    net = gridgen.config1_grid()
    rec, vm, va = oracle.solve(net, oracle.options(flat_start=True))
    warm = net.with_start(vm, np.degrees(va))
    outs = list(gridgen.config2_outages(warm)[:12])
    sets = outs + outs[:6]                 # six outage sets repeat: their patterns are met twice
    scales = [1.0] * 12 + [0.9] * 6
    one, hits1 = eng.fresh_plans(warm, sets, scales, threads=1)
    many, hits8 = eng.fresh_plans(warm, sets, scales, threads=8)
    assert hits1 == 6 and hits8 >= 0
    for a, b in zip(one, many):
        for k in ("q", "pivot_row", "lp", "li", "up", "ui", "j_rowptr", "j_colidx"):
            assert np.array_equal(getattr(a, k), getattr(b, k)), k
    # against the oracle at the instance's own start state
    for j in (0, 5, 14):
        o, sc = sets[j], scales[j]
        red = warm.with_load_multiplier(sc)
        rp, ci, val = oracle.jacobian(red, warm.vm, np.radians(warm.va_deg), skip=int(o))
        n = len(rp) - 1
        q = np.empty(n, np.int32)
        oracle.lib().orc_md_order(C.c_int32(n), oracle._p(rp), oracle._p(ci), oracle._p(q))
        h = oracle.lib().orc_lu_factor(C.c_int32(n), oracle._p(rp), oracle._p(ci), oracle._p(val), oracle._p(q),
                                       C.c_int32(1), C.c_double(0.1))
        assert h
        lnz, unz, rn = C.c_int32(), C.c_int32(), C.c_int32()
        oracle.lib().orc_lu_sizes(C.c_void_p(h), C.byref(lnz), C.byref(unz), C.byref(rn))
        arrs = [np.empty(k, t) for k, t in ((n + 1, np.int32), (lnz.value, np.int32), (lnz.value, np.float64),
                                            (n + 1, np.int32), (unz.value, np.int32), (unz.value, np.float64),
                                            (n, np.int32), (n, np.int32), (n, np.int32), (n + 1, np.int32),
                                            (rn.value, np.int32))]
        oracle.lib().orc_lu_get(C.c_void_p(h), *(oracle._p(a) for a in arrs))
        oracle.lib().orc_lu_free(C.c_void_p(h))
        assert np.array_equal(one[j].q, arrs[7]) and np.array_equal(one[j].pivot_row, arrs[8]), j


def test_schedule_and_group_wave_options_are_checked(eng):
    """gs_nr_options.schedule and the work-group size in flags are validated where the options enter the library (the
    host-only plan builder runs the same check as every device call)."""
    net = matpower_case("case9")
    eng.HostPlan(net, eng.NrOptions(schedule=eng.SCHEDULE_GROUP, flags=eng.group_waves(8)))
    eng.HostPlan(net, eng.NrOptions(schedule=eng.SCHEDULE_GROUP))
    for bad in (eng.NrOptions(schedule=eng.SCHEDULE_GROUP, flags=eng.group_waves(9)),
                eng.NrOptions(schedule=-3), eng.NrOptions(flags=1 << 12 << 4)):
        with pytest.raises(eng.GsError):
            eng.HostPlan(net, bad)
