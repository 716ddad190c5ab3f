"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so). TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this module, as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# GRIDSIM_ORACLE_LIB: another build of the same library (the host-sanitizer build of the CPU tests)
LIB_PATH = os.environ.get("GRIDSIM_ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

CONVERGED, MAX_ITERATIONS, DIVERGED, NONFINITE = 0, 1, 2, 3
STATUS_NAMES = {0: "Converged", 1: "MaxIterations", 2: "Diverged", 3: "NonFinite"}


class OrcNetwork(C.Structure):
    _fields_ = [("n_bus", C.c_int32), ("n_branch", C.c_int32), ("n_gen", C.c_int32), ("base_mva", C.c_double),
                ("bus_type", C.c_void_p), ("pd", C.c_void_p), ("qd", C.c_void_p), ("gs", C.c_void_p),
                ("bs", C.c_void_p), ("vm", C.c_void_p), ("va_deg", C.c_void_p),
                ("br_from", C.c_void_p), ("br_to", C.c_void_p), ("br_r", C.c_void_p), ("br_x", C.c_void_p),
                ("br_b", C.c_void_p), ("br_tap", C.c_void_p), ("br_shift_deg", C.c_void_p),
                ("br_in_service", C.c_void_p), ("gen_bus", C.c_void_p), ("gen_pg", C.c_void_p),
                ("gen_qg", C.c_void_p), ("gen_vg", C.c_void_p), ("gen_qmin", C.c_void_p), ("gen_qmax", C.c_void_p)]


class OrcOptions(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iterations", C.c_int32), ("diverge_limit", C.c_double),
                ("flat_start", C.c_int32), ("enforce_q_limits", C.c_int32), ("max_q_passes", C.c_int32),
                ("retry_dc_start", C.c_int32)]


class OrcResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("max_mismatch", C.c_double),
                ("solve_seconds", C.c_double), ("step_failures", C.c_int32), ("dc_retry", C.c_int32)]


RESULT_DTYPE = np.dtype([("status", np.int32), ("iterations", np.int32), ("max_mismatch", np.float64),
                         ("solve_seconds", np.float64), ("step_failures", np.int32), ("dc_retry", np.int32)])
# NewtonRetryOutcome
RETRY_NOT_TRIED, RETRY_CONVERGED, RETRY_DID_NOT_CONVERGE, RETRY_REJECTED, RETRY_UNAVAILABLE = 0, 1, 2, 3, 4

_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_lu_factor.restype = C.c_void_p
        _lib.orc_lu_refactor.restype = C.c_void_p
        _lib.orc_ybus.restype = C.c_int32
        _lib.orc_jacobian.restype = C.c_int32
        _lib.orc_mismatch.restype = C.c_int32
        _lib.orc_sparse_solve.restype = C.c_int32
        for name in ("orc_lu_sizes", "orc_lu_get", "orc_lu_solve", "orc_lu_free"):
            getattr(_lib, name).argtypes = None
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class _NetView:
    """Keeps the numpy arrays alive while the C struct points at them."""

    def __init__(self, net):
        self.net = net
        self.s = OrcNetwork(net.n_bus, net.n_branch, net.n_gen, float(net.base_mva),
                            _p(net.bus_type), _p(net.pd), _p(net.qd), _p(net.gs), _p(net.bs), _p(net.vm),
                            _p(net.va_deg), _p(net.br_from), _p(net.br_to), _p(net.br_r), _p(net.br_x),
                            _p(net.br_b), _p(net.br_tap), _p(net.br_shift_deg), _p(net.br_in_service),
                            _p(net.gen_bus), _p(net.gen_pg), _p(net.gen_qg), _p(net.gen_vg), _p(net.gen_qmin),
                            _p(net.gen_qmax))

    @property
    def ref(self):
        return C.byref(self.s)


def options(tolerance=1e-8, max_iterations=30, diverge_limit=10.0, flat_start=False, q_limits=False,
            max_q_passes=100, retry_dc_start=False):
    return OrcOptions(tolerance, max_iterations, diverge_limit, 1 if flat_start else 0, 1 if q_limits else 0,
                      max_q_passes, 1 if retry_dc_start else 0)


def without(net, branches):
    """The network with the given branches out of service (WithoutMany: every routine here skips an out-of-service
    branch exactly as it skips a removed one)."""
    mask = np.array(net.br_in_service, dtype=np.uint8, copy=True)
    for b in np.atleast_1d(np.asarray(branches, np.int64)):
        if b >= 0:
            mask[int(b)] = 0
    return net.copy(br_in_service=mask)


def sincos(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib().orc_sincos(_p(x), C.c_int64(x.size), _p(s), _p(c))
    return s, c


def complex_div(a, b, c, d):
    re, im = C.c_double(), C.c_double()
    lib().orc_complex_div(C.c_double(a), C.c_double(b), C.c_double(c), C.c_double(d), C.byref(re), C.byref(im))
    return re.value, im.value


def branch_stamp(r, x, b, tap, shift_deg):
    out = np.empty(8)
    lib().orc_branch_stamp(C.c_double(r), C.c_double(x), C.c_double(b), C.c_double(tap), C.c_double(shift_deg),
                           _p(out))
    return out[0::2] + 1j * out[1::2]


def ybus(net, skip=-1):
    v = _NetView(net)
    nnz = lib().orc_ybus(v.ref, skip, None, None, None, None)
    rp = np.empty(net.n_bus + 1, np.int32)
    ci = np.empty(nnz, np.int32)
    g = np.empty(nnz)
    b = np.empty(nnz)
    lib().orc_ybus(v.ref, skip, _p(rp), _p(ci), _p(g), _p(b))
    return rp, ci, g, b


def bus_spec(net, flat_start=False):
    n = net.n_bus
    t = np.empty(n, np.int32)
    vm, va, ps, qs = (np.empty(n) for _ in range(4))
    lib().orc_bus_spec(_NetView(net).ref, 1 if flat_start else 0, _p(t), _p(vm), _p(va), _p(ps), _p(qs))
    return t, vm, va, ps, qs


def mismatch(net, vm, va, skip=-1):
    n = net.n_bus
    vm = np.ascontiguousarray(vm, np.float64)
    va = np.ascontiguousarray(va, np.float64)
    p, q, f = np.empty(n), np.empty(n), np.empty(2 * n)
    m = lib().orc_mismatch(_NetView(net).ref, skip, _p(vm), _p(va), _p(p), _p(q), _p(f))
    return p, q, f[:m].copy()


def jacobian(net, vm, va, skip=-1):
    v = _NetView(net)
    vm = np.ascontiguousarray(vm, np.float64)
    va = np.ascontiguousarray(va, np.float64)
    m = C.c_int32()
    nnz = lib().orc_jacobian(v.ref, skip, _p(vm), _p(va), C.byref(m), None, None, None)
    rp = np.empty(m.value + 1, np.int32)
    ci = np.empty(nnz, np.int32)
    val = np.empty(nnz)
    lib().orc_jacobian(v.ref, skip, _p(vm), _p(va), None, _p(rp), _p(ci), _p(val))
    return rp, ci, val


def md_order(rowptr, colidx):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colidx = np.ascontiguousarray(colidx, np.int32)
    n = len(rowptr) - 1
    q = np.empty(n, np.int32)
    lib().orc_md_order(C.c_int32(n), _p(rowptr), _p(colidx), _p(q))
    return q


class LU:
    """A Gilbert-Peierls factorisation with its plan (the arrays SparseLuFactorization and SparseLuPlan hold)."""

    def __init__(self, handle, n):
        self.h = C.c_void_p(handle)
        self.n = n
        L = lib()
        lnz, unz, rn = C.c_int32(), C.c_int32(), C.c_int32()
        L.orc_lu_sizes(self.h, C.byref(lnz), C.byref(unz), C.byref(rn))
        self.lp = np.empty(n + 1, np.int32); self.li = np.empty(lnz.value, np.int32); self.lx = np.empty(lnz.value)
        self.up = np.empty(n + 1, np.int32); self.ui = np.empty(unz.value, np.int32); self.ux = np.empty(unz.value)
        self.pinv = np.empty(n, np.int32); self.q = np.empty(n, np.int32); self.pivot_row = np.empty(n, np.int32)
        self.reach_ptr = np.empty(n + 1, np.int32); self.reach = np.empty(rn.value, np.int32)
        L.orc_lu_get(self.h, _p(self.lp), _p(self.li), _p(self.lx), _p(self.up), _p(self.ui), _p(self.ux),
                     _p(self.pinv), _p(self.q), _p(self.pivot_row), _p(self.reach_ptr), _p(self.reach))

    def solve(self, b):
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty(self.n)
        lib().orc_lu_solve(self.h, _p(b), _p(x))
        return x

    def __del__(self):
        try:
            lib().orc_lu_free(self.h)
        except Exception:
            pass


def lu_factor(rowptr, colidx, val, q, keep_zeros=False, diagonal_preference=0.0):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colidx = np.ascontiguousarray(colidx, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    q = np.ascontiguousarray(q, np.int32)
    n = len(rowptr) - 1
    h = lib().orc_lu_factor(C.c_int32(n), _p(rowptr), _p(colidx), _p(val), _p(q), C.c_int32(int(keep_zeros)),
                            C.c_double(diagonal_preference))
    return LU(h, n) if h else None


def lu_refactor(plan: LU, rowptr, colidx, val):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colidx = np.ascontiguousarray(colidx, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    h = lib().orc_lu_refactor(plan.h, C.c_int32(plan.n), _p(rowptr), _p(colidx), _p(val))
    return LU(h, plan.n) if h else None


def sparse_solve(rowptr, colidx, val, b):
    """CpuSparseLuSolver.Solve: (rc, x) with rc 0 ok / 1 singular / 2 non-finite input."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colidx = np.ascontiguousarray(colidx, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.empty(len(rowptr) - 1)
    rc = lib().orc_sparse_solve(C.c_int32(len(rowptr) - 1), _p(rowptr), _p(colidx), _p(val), _p(b), _p(x))
    return rc, x


def ilu0(rowptr, colidx, val):
    """ILU(0) factors in the pattern's positions, or None when a row has no diagonal entry."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colidx = np.ascontiguousarray(colidx, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    out = np.zeros_like(val)
    ok = lib().orc_ilu0(C.c_int32(len(rowptr) - 1), _p(rowptr), _p(colidx), _p(val), _p(out))
    return out if ok else None


def bicgstab(rowptr, colidx, val, b, tol=1e-10, max_iter=None):
    """(solved, x) of ILU(0)-preconditioned BiCGSTAB; max_iter defaults to SolveStep's 4 n + 200."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colidx = np.ascontiguousarray(colidx, np.int32)
    n = len(rowptr) - 1
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n)
    ok = lib().orc_bicgstab(C.c_int32(n), _p(rowptr), _p(colidx), _p(val), _p(b), C.c_double(tol),
                            C.c_int32(4 * n + 200 if max_iter is None else max_iter), _p(x))
    return bool(ok), x


def solve(net, opts=None, skip=-1):
    """One solve: (result record, vm, va in radians). skip: one outaged branch, or a sequence of them (N-k)."""
    opts = opts or options()
    if np.ndim(skip) > 0:
        net, skip = without(net, skip), -1
    vm = np.empty(net.n_bus)
    va = np.empty(net.n_bus)
    res = OrcResult()
    lib().orc_solve(_NetView(net).ref, skip, C.byref(opts), _p(vm), _p(va), C.byref(res))
    rec = np.zeros((), RESULT_DTYPE)
    for k in RESULT_DTYPE.names:
        rec[k] = getattr(res, k)
    return rec, vm, va


def screen(net, outages, opts=None, threads=1, keep_state=True):
    """N-1 screen: (results [count] structured array, vm [count, n] or None, va [count, n] or None). outages: one
    branch per solve, or per solve a sequence of branches (N-k, WithoutMany's index list)."""
    opts = opts or options()
    if any(np.ndim(o) > 0 for o in list(outages)):
        sets = [np.atleast_1d(np.asarray(o, np.int32)) for o in outages]
        offsets = np.zeros(len(sets) + 1, np.int32)
        offsets[1:] = np.cumsum([len(x) for x in sets])
        flat = np.ascontiguousarray(np.concatenate(sets) if sets else np.zeros(0, np.int32), np.int32)
        count = len(sets)
        res = np.zeros(count, RESULT_DTYPE)
        vm = np.empty((count, net.n_bus)) if keep_state else None
        va = np.empty((count, net.n_bus)) if keep_state else None
        lib().orc_screen_sets(_NetView(net).ref, _p(offsets), _p(flat), C.c_int32(count), C.byref(opts),
                              C.c_int32(threads), _p(res), _p(vm), _p(va))
        return res, vm, va
    outages = np.ascontiguousarray(outages, np.int32)
    count = len(outages)
    res = np.zeros(count, RESULT_DTYPE)
    vm = np.empty((count, net.n_bus)) if keep_state else None
    va = np.empty((count, net.n_bus)) if keep_state else None
    lib().orc_screen(_NetView(net).ref, _p(outages), C.c_int32(count), C.byref(opts), C.c_int32(threads),
                     _p(res), _p(vm), _p(va))
    return res, vm, va


def dc_seed(net, flat_start=False):
    """NewtonShell.TrySeedDcAngles on the ordinary start: the seeded angles (rad), or None when unavailable."""
    va = np.empty(net.n_bus)
    ok = lib().orc_dc_seed(_NetView(net).ref, C.c_int32(1 if flat_start else 0), _p(va))
    return va if ok else None


def line_flows(net, vm, va, skip=-1):
    """NewtonShell.ComputeLineFlows: (pf, qf, pt, qt in MW / MVAr, loading percent of rate A), per branch."""
    if np.ndim(skip) > 0:
        net, skip = without(net, skip), -1
    nb = net.n_branch
    out = [np.zeros(nb) for _ in range(5)]
    rate = np.ascontiguousarray(net.br_rate_a, np.float64)
    lib().orc_line_flows(_NetView(net).ref, int(skip), _p(rate), _p(np.ascontiguousarray(vm, np.float64)),
                         _p(np.ascontiguousarray(va, np.float64)), *(_p(o) for o in out))
    return tuple(out)


SUMMARY_FIELDS = ("max_loading_pct", "overloads", "vmin", "vmax", "undervoltage", "overvoltage", "loss_mw",
                  "loss_mvar")


def screen_summary(net, vm, va, skip=-1):
    """The classification numbers of a converged state (ScreenStateCore, each bus against its own limits where the
    network gives them, as the CLI's default) and AssembleResult's losses, as a dict."""
    if np.ndim(skip) > 0:
        net, skip = without(net, skip), -1
    vm = np.ascontiguousarray(vm, np.float64)
    va = np.ascontiguousarray(va, np.float64)
    p, q, _ = mismatch(net, vm, va, skip=skip)
    out = np.zeros(8)
    rate = np.ascontiguousarray(net.br_rate_a, np.float64)
    lo = np.ascontiguousarray(net.bus_vmin, np.float64)
    hi = np.ascontiguousarray(net.bus_vmax, np.float64)
    lib().orc_screen_summary(_NetView(net).ref, int(skip), _p(rate), _p(lo), _p(hi), _p(vm), _p(va),
                             _p(np.ascontiguousarray(p)), _p(np.ascontiguousarray(q)), _p(out))
    return dict(zip(SUMMARY_FIELDS, out.tolist()))
