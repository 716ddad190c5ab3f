"""gs_sparse_solve_noalloc: the sparse-solver seam with its plan and buffers in a fixed device scratch array (no
device allocation). Called through the C ABI and compared with gs_sparse_solve, bit for bit, on two calls in a row,
the second a smaller system; a system beyond the scratch array is refused with GS_INVALID_ARGUMENT."""
import ctypes as C

import numpy as np
import pytest

from conftest import matpower_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from gridsim_amd import engine
    c = engine.Context(0)
    yield c
    c.close()


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _solve(fn, ctx, rp, ci, val, b):
    from gridsim_amd import engine
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.full(len(rp) - 1, np.nan)
    st = getattr(engine.lib(), fn)(ctx.h, C.c_int32(len(rp) - 1), _p(rp), _p(ci), _p(val), _p(b), _p(x))
    return st, x


def _jacobian(oracle, case, skip=-1):
    net = matpower_case(case)
    rec, vm, va = oracle.solve(net)
    return oracle.jacobian(net, vm, va, skip=skip)


def test_no_alloc_seam_equals_the_allocating_seam(ctx, oracle):
    rng = np.random.default_rng(11)
    systems = [_jacobian(oracle, "case300"), _jacobian(oracle, "case14", skip=3)]   # the second is smaller
    for rep in range(2):
        for rp, ci, val in systems:
            b = rng.standard_normal(len(rp) - 1)
            st0, x0 = _solve("gs_sparse_solve", ctx, rp, ci, val, b)
            st1, x1 = _solve("gs_sparse_solve_noalloc", ctx, rp, ci, val, b)
            print(f"no-alloc seam: n={len(rp) - 1}, statuses {st0} {st1}", flush=True)
            assert st0 == st1 == 0
            assert np.array_equal(x0, x1)
            rc, xr = oracle.sparse_solve(rp, ci, val, b)
            assert rc == 0 and np.array_equal(x1, xr)


def test_no_alloc_seam_replays_the_plan_its_array_holds(ctx, oracle):
    """Calls on one structure in a row: the second and third replay the plan the scratch array holds (new values, the
    same pivots), and still give gs_sparse_solve's and the oracle's bits."""
    rng = np.random.default_rng(12)
    rp, ci, val = _jacobian(oracle, "case118")
    for rep in range(3):
        v = val * (1 + 1e-6 * rng.standard_normal(len(val)))
        b = rng.standard_normal(len(rp) - 1)
        st0, x0 = _solve("gs_sparse_solve", ctx, rp, ci, v, b)
        st1, x1 = _solve("gs_sparse_solve_noalloc", ctx, rp, ci, v, b)
        assert st0 == st1 == 0 and np.array_equal(x0, x1), rep
        rc, xr = oracle.sparse_solve(rp, ci, v, b)
        assert rc == 0 and np.array_equal(x1, xr), rep


def test_no_alloc_seam_keeps_the_seam_contract(ctx):
    from gridsim_amd import engine
    rp = np.array([0, 2, 4], np.int32)
    ci = np.array([0, 1, 0, 1], np.int32)
    st, _ = _solve("gs_sparse_solve_noalloc", ctx, rp, ci, np.array([1.0, 2.0, 2.0, 4.0]), np.ones(2))
    assert st == engine.GS_SINGULAR
    st, _ = _solve("gs_sparse_solve_noalloc", ctx, rp, ci, np.array([1.0, np.nan, 2.0, 4.0]), np.ones(2))
    assert st == engine.GS_NONFINITE_INPUT
    st, x = _solve("gs_sparse_solve_noalloc", ctx, rp, ci, np.array([1.0, 2.0, 3.0, 4.0]), np.ones(2))
    assert st == engine.GS_OK and np.allclose(x, np.linalg.solve([[1, 2], [3, 4]], [1, 1]))


def test_no_alloc_seam_refuses_a_system_beyond_its_scratch(ctx):
    """A tridiagonal system of 20,000 rows (one elimination level per column: a plan of about 143 MB) exceeds the
    64 MiB scratch array: GS_INVALID_ARGUMENT, naming the bytes it needs, and nothing written to x."""
    from gridsim_amd import engine
    n = 20000
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                rows.append(i); cols.append(j); vals.append(4.0 if i == j else -1.0)
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, np.array(rows) + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    b = np.ones(n)
    st, x = _solve("gs_sparse_solve_noalloc", ctx, rp, np.array(cols, np.int32), np.array(vals), b)
    assert st == engine.GS_INVALID_ARGUMENT and np.isnan(x).all()
    assert "bytes of device scratch" in ctx.error()
