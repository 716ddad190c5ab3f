"""Python bindings of libgridsim_hip.so and a mirror of the reference's solver seams on top of it.

The classes follow the reference interfaces the library replaces, with the same names, argument meaning and error
behaviour:

* ``HipSparseLinearSolver``        - ISparseLinearSolver (src/GridSim.Core/Abstractions/ISparseLinearSolver.cs:12-16):
  ``solve(a, b)`` returns the solution or ``None`` when the matrix is singular, raises ``ValueError`` (the
  reference's ArgumentException) for a NaN/Inf matrix or right-hand side.
* ``HipBatchedSparseLinearSolver`` - IBatchedSparseLinearSolver (IBatchedSparseLinearSolver.cs:20-37): one entry
  per system, a throwing system stores its exception in ``errors[s]`` and does not abort the others.
* ``HipNewtonPowerFlow``           - IPowerFlowSolver (IPowerFlowSolver.cs:10-12) on a device-resident network.
* ``screen_branches_batched``      - ContingencyAnalysis.ScreenBranchesBatched (ContingencyAnalysis.Batched.cs:55).

The library must be built (``python -m gridsim_amd.build``); nothing here falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from .network import SLACK, Network, bridges, splits

PKG = os.path.dirname(os.path.abspath(__file__))
# GRIDSIM_HIP_LIB: another build of the same library (the host-sanitizer build of the CPU tests)
LIB_PATH = os.environ.get("GRIDSIM_HIP_LIB") or os.path.join(PKG, "libgridsim_hip.so")

GS_OK, GS_SINGULAR, GS_NONFINITE_INPUT, GS_OUT_OF_MEMORY, GS_INVALID_ARGUMENT, GS_DEVICE_ERROR, GS_NO_DEVICE, \
    GS_BUSY, GS_PENDING = range(9)
CONVERGED, MAX_ITERATIONS, DIVERGED, NONFINITE = 0, 1, 2, 3
LU_FP64, LU_FP32 = 0, 1
FLAG_X_GLOBAL = 1
FLAG_SINGLE_WAVE = 4
FLAG_Q_LIMITS = 2      # enforce generator reactive limits (the outer loop of PV -> PQ switching)
FLAG_DC_RETRY = 8      # retry a stalled or diverged solve from the DC power-flow angle start (RetryFromDcStart)
SCHEDULE_LEVELS = -1   # every instance spread over the whole GPU, level by level (GS_SCHEDULE_LEVELS)
SCHEDULE_GROUP = -2    # a work-group of several wavefronts per instance (GS_SCHEDULE_GROUP)


def group_waves(w: int) -> int:
    """The flags bits that set the waves of a work-group per instance (GS_GROUP_WAVES)."""
    return int(w) << 8
STATUS_NAMES = {0: "Converged", 1: "MaxIterations", 2: "Diverged", 3: "NonFinite"}


class GsError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"gridsim-hip status {status}: {message}")
        self.status = status


class GsNetwork(C.Structure):
    _fields_ = [("n_bus", C.c_int32), ("n_branch", C.c_int32), ("n_gen", C.c_int32), ("base_mva", C.c_double),
                ("bus_type", C.c_void_p), ("pd", C.c_void_p), ("qd", C.c_void_p), ("gs", C.c_void_p),
                ("bs", C.c_void_p), ("vm", C.c_void_p), ("va_deg", C.c_void_p),
                ("br_from", C.c_void_p), ("br_to", C.c_void_p), ("br_r", C.c_void_p), ("br_x", C.c_void_p),
                ("br_b", C.c_void_p), ("br_tap", C.c_void_p), ("br_shift_deg", C.c_void_p),
                ("br_in_service", C.c_void_p), ("gen_bus", C.c_void_p), ("gen_pg", C.c_void_p),
                ("gen_qg", C.c_void_p), ("gen_vg", C.c_void_p), ("br_rate_a", C.c_void_p),
                ("bus_vmin", C.c_void_p), ("bus_vmax", C.c_void_p), ("gen_qmin", C.c_void_p),
                ("gen_qmax", C.c_void_p)]


class GsNrOptions(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iterations", C.c_int32), ("flat_start", C.c_int32),
                ("diverge_limit", C.c_double), ("lu_precision", C.c_int32), ("refine_steps", C.c_int32),
                ("growth_tol", C.c_double), ("schedule", C.c_int32), ("flags", C.c_int32)]


class GsInstance(C.Structure):
    _fields_ = [("outage_branch", C.c_int32), ("n_outages", C.c_int32), ("load_scale", C.c_double),
                ("vm0", C.c_void_p), ("va0", C.c_void_p), ("outages", C.c_void_p)]


MAX_OUTAGES = 8   # GS_MAX_OUTAGES
INSTANCE_DTYPE = np.dtype([("outage_branch", np.int32), ("n_outages", np.int32), ("load_scale", np.float64),
                           ("vm0", np.uint64), ("va0", np.uint64), ("outages", np.uint64)])   # the layout of GsInstance
RESULT_DTYPE = np.dtype([("status", np.int32), ("iterations", np.int32), ("max_mismatch", np.float64),
                         ("symbolic_refreshes", np.int32), ("step_failures", np.int32),
                         ("min_pivot_ratio", np.float64), ("phase_us", np.float64, (4,)),
                         ("engine_stop", np.int32), ("dc_retry", np.int32)])
# gs_nr_result.engine_stop and .dc_retry (GS_STOP_*, GS_RETRY_* = PowerFlowResult.NewtonRetry)
STOP_NONE, STOP_LOCK_TIMEOUT, STOP_REFRESH_CAP, STOP_FALLBACK_CAP = 0, 1, 2, 3
RETRY_NOT_TRIED, RETRY_CONVERGED, RETRY_DID_NOT_CONVERGE, RETRY_REJECTED, RETRY_UNAVAILABLE = 0, 1, 2, 3, 4
PHASES = ("mismatch+jacobian", "refactor", "solve+refine", "update")
SUMMARY_DTYPE = np.dtype([("max_loading_pct", np.float64), ("max_loading_branch", np.int32), ("overloads", np.int32),
                          ("vmin", np.float64), ("vmax", np.float64), ("undervoltage", np.int32),
                          ("overvoltage", np.int32), ("loss_mw", np.float64), ("loss_mvar", np.float64)])


class GsCsr(C.Structure):
    _fields_ = [("n", C.c_int32), ("reserved", C.c_int32), ("rowptr", C.c_void_p), ("colidx", C.c_void_p),
                ("val", C.c_void_p)]


class GsPlanStats(C.Structure):
    _fields_ = [("m", C.c_int32), ("nnz_j", C.c_int32), ("nnz_l", C.c_int32), ("nnz_u", C.c_int32),
                ("levels", C.c_int32), ("refactor_rows", C.c_int32), ("forward_steps", C.c_int32),
                ("back_steps", C.c_int32), ("workspace_stride", C.c_int32), ("factor_slots", C.c_int32),
                ("stream_bytes", C.c_int64), ("rows_by_kind", C.c_int32 * 5), ("update_lanes", C.c_int32),
                ("forward_lanes", C.c_int32), ("back_lanes", C.c_int32), ("jacobian_transposed", C.c_int32)]


def _plan_stats_dict(st: "GsPlanStats") -> dict:
    out = {}
    for k, _ in GsPlanStats._fields_:
        if k == "reserved":
            continue
        v = getattr(st, k)
        out[k] = list(v) if k == "rows_by_kind" else v
    return out


class GsServerConfig(C.Structure):
    _fields_ = [("max_batch", C.c_int32), ("streams", C.c_int32), ("queue_capacity", C.c_int32),
                ("reserved", C.c_int32), ("state_bytes", C.c_size_t), ("pinned_bytes", C.c_size_t)]


class GsServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
                ("launches", C.c_int64), ("queued", C.c_int64), ("in_flight", C.c_int64),
                ("mean_batch", C.c_double), ("latency_p50_ms", C.c_double), ("latency_p99_ms", C.c_double),
                ("latency_max_ms", C.c_double), ("busy_ms", C.c_double), ("state_bytes_peak", C.c_int64),
                ("joined", C.c_int64)]


class GsDeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int32),
                ("fp64_ok", C.c_int32), ("hbm_bytes", C.c_size_t), ("fdlibm_ok", C.c_int32),
                ("reserved", C.c_int32)]


_lib = None


def lib():
    """The loaded library; raises when it has not been built (the engine has no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gridsim_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.gs_last_error.restype = C.c_char_p
        L.gs_version.restype = C.c_char_p
        for name in ("gs_init", "gs_device_info_get", "gs_sparse_solve", "gs_sparse_solve_batch",
                     "gs_network_upload", "gs_nr_solve_batch", "gs_nr_solve_plain", "gs_network_plan_sizes",
                     "gs_network_plan_get", "gs_last_timing", "gs_device_sincos", "gs_plan_build",
                     "gs_plan_sizes", "gs_plan_get", "gs_batch_create", "gs_batch_run", "gs_batch_results",
                     "gs_nr_autotune", "gs_network_plan_stats", "gs_plan_shape", "gs_server_create",
                     "gs_server_add_network", "gs_server_submit", "gs_server_wait", "gs_server_stats_get",
                     "gs_server_stats_reset", "gs_nr_screen_batch", "gs_network_launch_shape"):
            getattr(L, name).restype = C.c_int32
        L.gs_host_sincos.restype = None
        L.gs_server_free.restype = None
        L.gs_server_wait.argtypes = [C.c_void_p, C.c_uint64, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p]
        L.gs_server_submit.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class NetView:
    """A ``gs_network`` pointing at the numpy arrays of a Network (kept alive with it)."""

    def __init__(self, net: Network):
        self.net = net
        self.s = GsNetwork(net.n_bus, net.n_branch, net.n_gen, float(net.base_mva), _p(net.bus_type), _p(net.pd),
                           _p(net.qd), _p(net.gs), _p(net.bs), _p(net.vm), _p(net.va_deg), _p(net.br_from),
                           _p(net.br_to), _p(net.br_r), _p(net.br_x), _p(net.br_b), _p(net.br_tap),
                           _p(net.br_shift_deg), _p(net.br_in_service), _p(net.gen_bus), _p(net.gen_pg),
                           _p(net.gen_qg), _p(net.gen_vg), _p(net.br_rate_a), _p(net.bus_vmin), _p(net.bus_vmax),
                           _p(net.gen_qmin), _p(net.gen_qmax))

    @property
    def ref(self):
        return C.byref(self.s)


@dataclass
class NrOptions:
    """PowerFlowOptions for the engine (PowerFlowOptions.cs): Q limits and the DC-start retry are off."""
    tolerance: float = 1e-8
    max_iterations: int = 30
    flat_start: bool = False
    diverge_limit: float = 10.0
    lu_precision: int = LU_FP32
    refine_steps: int = 1
    growth_tol: float = 1e-3
    schedule: int = 0
    flags: int = 0                 # FLAG_X_GLOBAL: the solve vector in global memory

    def c(self) -> GsNrOptions:
        return GsNrOptions(self.tolerance, self.max_iterations, int(self.flat_start), self.diverge_limit,
                           self.lu_precision, self.refine_steps, self.growth_tol, self.schedule, self.flags)


def host_sincos(x):
    """The product's FDLIBM compiled for the host (no device needed)."""
    x = np.ascontiguousarray(x, np.float64)
    s, c = np.empty_like(x), np.empty_like(x)
    lib().gs_host_sincos(C.c_int64(x.size), _p(x), _p(s), _p(c))
    return s, c


class HostPlan:
    """The symbolic plan the engine would upload for a network (gs_plan_build), computed on the host."""

    def __init__(self, net: Network, options: NrOptions | None = None, handle=None):
        if handle is None:
            opts = (options or NrOptions()).c()
            h = C.c_void_p()
            st = lib().gs_plan_build(NetView(net).ref, C.byref(opts), C.byref(h))
            if st != GS_OK:
                raise GsError(st, "gs_plan_build")
        else:
            h = handle
        self.h = h
        m, nj, nl, nu, lv = (C.c_int32() for _ in range(5))
        lib().gs_plan_sizes(h, C.byref(m), C.byref(nj), C.byref(nl), C.byref(nu), C.byref(lv))
        self.m, self.nnz_j, self.nnz_l, self.nnz_u, self.n_levels = m.value, nj.value, nl.value, nu.value, lv.value
        self.q = np.empty(self.m, np.int32); self.pivot_row = np.empty(self.m, np.int32)
        self.lp = np.empty(self.m + 1, np.int32); self.li = np.empty(self.nnz_l, np.int32)
        self.up = np.empty(self.m + 1, np.int32); self.ui = np.empty(self.nnz_u, np.int32)
        self.j_rowptr = np.empty(self.m + 1, np.int32); self.j_colidx = np.empty(self.nnz_j, np.int32)
        lib().gs_plan_get(h, _p(self.q), _p(self.pivot_row), _p(self.lp), _p(self.li), _p(self.up), _p(self.ui),
                          _p(self.j_rowptr), _p(self.j_colidx))

    def stats(self) -> dict:
        """The compiled device plan's shape (gs_plan_shape): record streams, levels, shared bytes."""
        st = GsPlanStats()
        if lib().gs_plan_shape(self.h, C.byref(st)) != GS_OK:
            raise GsError(GS_INVALID_ARGUMENT, "gs_plan_shape")
        return _plan_stats_dict(st)

    def __del__(self):
        try:
            lib().gs_plan_free(self.h)
        except Exception:
            pass


def fresh_plans(net: Network, outages, load_scale=None, options: NrOptions | None = None, states=None,
                threads: int = 8):
    """The engine's fresh symbolic passes for instances of `net` (gs_plan_fresh_batch) on `threads` host threads:
    ([HostPlan or None per instance], minimum-degree orders the pattern cache supplied). states: (vm, va) per instance
    ([count, n] each), or None for each instance's start point."""
    probe = DeviceNetwork.__new__(DeviceNetwork)
    probe.net = net
    arr, count, keep = probe._instances(outages, load_scale)
    handles = (C.c_void_p * max(count, 1))()
    hits = C.c_int64()
    opts = (options or NrOptions()).c()
    vm = va = None
    if states is not None:
        vm = np.ascontiguousarray(states[0], np.float64)
        va = np.ascontiguousarray(states[1], np.float64)
    st = lib().gs_plan_fresh_batch(NetView(net).ref, C.byref(opts), C.c_int32(count), arr, _p(vm), _p(va),
                                   C.c_int32(threads), handles, C.byref(hits))
    if st not in (GS_OK, GS_SINGULAR):
        raise GsError(st, "gs_plan_fresh_batch")
    plans = [HostPlan(net, handle=C.c_void_p(handles[i])) if handles[i] else None for i in range(count)]
    return plans, hits.value


class Context:
    """A device context (gs_init). One per thread or shared: calls are serialised inside the library."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = C.c_void_p()
        st = lib().gs_init(C.c_int32(device), C.c_size_t(hbm_budget), C.byref(h))
        if st != GS_OK:
            raise GsError(st, "gs_init: no usable GPU" if st == GS_NO_DEVICE else "gs_init")
        self.h = h

    def error(self) -> str:
        return lib().gs_last_error(self.h).decode()

    def check(self, st, what):
        if st != GS_OK:
            raise GsError(st, f"{what}: {self.error()}")

    def device_info(self) -> dict:
        info = GsDeviceInfo()
        self.check(lib().gs_device_info_get(self.h, C.byref(info)), "gs_device_info_get")
        return {"name": info.name.decode(), "arch": info.arch.decode(), "compute_units": info.compute_units,
                "fp64_ok": bool(info.fp64_ok), "fdlibm_ok": bool(info.fdlibm_ok), "hbm_bytes": info.hbm_bytes}

    def device_sincos(self, x):
        x = np.ascontiguousarray(x, np.float64)
        s, c = np.empty_like(x), np.empty_like(x)
        self.check(lib().gs_device_sincos(self.h, C.c_int64(x.size), _p(x), _p(s), _p(c)), "gs_device_sincos")
        return s, c

    def close(self):
        if getattr(self, "h", None):
            lib().gs_shutdown(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- the sparse linear-solver seams

class HipSparseLinearSolver:
    """ISparseLinearSolver on the GPU: CpuSparseLuSolver's contract (CpuSparseLuSolver.cs:30-48)."""

    name = "sparse-lu (hip, gfx950)"

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def solve(self, rowptr, colidx, val, b):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        colidx = np.ascontiguousarray(colidx, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        n = len(rowptr) - 1
        if len(b) != n:
            raise ValueError(f"The right-hand side has {len(b)} entries; the system has {n}.")
        x = np.empty(n)
        st = lib().gs_sparse_solve(self.ctx.h, C.c_int32(n), _p(rowptr), _p(colidx), _p(val), _p(b), _p(x))
        if st == GS_SINGULAR:
            return None
        if st == GS_NONFINITE_INPUT:
            raise ValueError(self.ctx.error())
        self.ctx.check(st, "gs_sparse_solve")
        return x


class HipBatchedSparseLinearSolver:
    """IBatchedSparseLinearSolver on the GPU (IBatchedSparseLinearSolver.cs:20-37)."""

    name = "sparse-lu batched (hip, gfx950)"

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def solve_batch(self, matrices, rhs, errors):
        if len(matrices) != len(rhs) or len(errors) != len(matrices):
            raise ValueError("matrices, right-hand sides and errors must have one entry per system")
        count = len(matrices)
        keep = []
        systems = (GsCsr * max(count, 1))()
        bptr = (C.c_void_p * max(count, 1))()
        xptr = (C.c_void_p * max(count, 1))()
        xs = []
        for s, ((rp, ci, v), b) in enumerate(zip(matrices, rhs)):
            rp = np.ascontiguousarray(rp, np.int32); ci = np.ascontiguousarray(ci, np.int32)
            v = np.ascontiguousarray(v, np.float64); b = np.ascontiguousarray(b, np.float64)
            if len(b) != len(rp) - 1:
                raise ValueError(f"system {s}: right-hand side length {len(b)} != {len(rp) - 1}")
            x = np.empty(len(rp) - 1)
            keep += [rp, ci, v, b]
            xs.append(x)
            systems[s] = GsCsr(len(rp) - 1, 0, _p(rp), _p(ci), _p(v))
            bptr[s] = b.ctypes.data
            xptr[s] = x.ctypes.data
        status = np.zeros(max(count, 1), np.int32)
        self.ctx.check(lib().gs_sparse_solve_batch(self.ctx.h, C.c_int32(count), systems, bptr, xptr, _p(status)),
                       "gs_sparse_solve_batch")
        out = []
        for s in range(count):
            if status[s] == GS_OK:
                out.append(xs[s])
            else:
                if status[s] == GS_NONFINITE_INPUT:
                    errors[s] = ValueError(f"system {s}: non-finite input")
                elif status[s] != GS_SINGULAR:
                    errors[s] = GsError(int(status[s]), f"system {s}")
                out.append(None)
        return out


# ---------------------------------------------------------------- whole Newton-Raphson

@dataclass
class BatchResult:
    results: np.ndarray            # RESULT_DTYPE per instance
    vm: np.ndarray | None
    va: np.ndarray | None          # radians
    kernel_ms: float
    total_ms: float
    launches: int

    @property
    def status(self):
        return self.results["status"]


class DeviceNetwork:
    """A network uploaded to HBM with its symbolic plan (gs_network_upload)."""

    def __init__(self, ctx: Context, net: Network, options: NrOptions | None = None):
        self.ctx = ctx
        self.net = net
        self.options = options or NrOptions()
        self._view = NetView(net)
        h = C.c_void_p()
        ctx.check(lib().gs_network_upload(ctx.h, self._view.ref, C.byref(self.options.c()), C.byref(h)),
                  "gs_network_upload")
        self.h = h

    def plan(self) -> dict:
        m, nj, nl, nu = (C.c_int32() for _ in range(4))
        lib().gs_network_plan_sizes(self.h, C.byref(m), C.byref(nj), C.byref(nl), C.byref(nu))
        out = {k: np.empty(s, np.int32) for k, s in (("q", m.value), ("pivot_row", m.value), ("lp", m.value + 1),
                                                       ("li", nl.value), ("up", m.value + 1), ("ui", nu.value),
                                                       ("j_rowptr", m.value + 1), ("j_colidx", nj.value))}
        lib().gs_network_plan_get(self.h, *(_p(out[k]) for k in ("q", "pivot_row", "lp", "li", "up", "ui",
                                                                 "j_rowptr", "j_colidx")))
        return out

    def launch_shape(self, options: NrOptions | None = None) -> dict:
        """How the batched kernel would be launched for these options (gs_network_launch_shape)."""
        fp32, xl, waves, lds = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int64()
        o = (options or self.options).c()
        self.ctx.check(lib().gs_network_launch_shape(self.h, C.byref(o), C.byref(fp32), C.byref(xl), C.byref(waves),
                                                     C.byref(lds)), "gs_network_launch_shape")
        return {"lu_fp32": bool(fp32.value), "x_in_lds": bool(xl.value), "waves_per_group": waves.value,
                "lds_bytes": lds.value}

    def plan_stats(self) -> dict:
        """Shape of the compiled device plan (record streams, levels, shared bytes)."""
        st = GsPlanStats()
        self.ctx.check(lib().gs_network_plan_stats(self.h, C.byref(st)), "gs_network_plan_stats")
        return _plan_stats_dict(st)

    def _instances(self, outages=None, load_scale=None, count=None, starts=None):
        """GsInstance records. outages: one branch index per instance (-1: none), or per instance a sequence of branch
        indices (N-2 / N-k: WithoutMany's index list)."""
        count = count if count is not None else (len(outages) if outages is not None else len(load_scale))
        arr = (GsInstance * max(count, 1))()
        view = np.frombuffer(arr, INSTANCE_DTYPE)[:count]    # filled column by column, not record by record
        keep = []
        flat = None
        if outages is not None and not isinstance(outages, (list, tuple)):
            a = np.asarray(outages)
            flat = a if a.ndim == 1 and a.dtype != object else None   # one branch per instance: no per-item scan
        if outages is not None and flat is None and any(np.ndim(o) > 0 for o in list(outages)[:count]):
            lists = [np.atleast_1d(np.asarray(o, np.int32)) for o in list(outages)[:count]]
            view["outage_branch"] = [int(o[0]) if len(o) else -1 for o in lists]
            rest = [np.ascontiguousarray(o[1:]) for o in lists]
            keep += rest
            view["n_outages"] = [len(r) for r in rest]
            view["outages"] = [r.ctypes.data if len(r) else 0 for r in rest]
        else:
            view["outage_branch"] = -1 if outages is None else np.asarray(outages, np.int64)[:count]
        view["load_scale"] = 1.0 if load_scale is None else np.asarray(load_scale, np.float64)[:count]
        if starts is not None and count > 0:
            vm0 = np.ascontiguousarray(np.asarray(starts[0], np.float64)[:count].reshape(count, -1))
            va0 = np.ascontiguousarray(np.asarray(starts[1], np.float64)[:count].reshape(count, -1))
            keep += [vm0, va0]
            view["vm0"] = vm0.ctypes.data + vm0.strides[0] * np.arange(count, dtype=np.uint64)
            view["va0"] = va0.ctypes.data + va0.strides[0] * np.arange(count, dtype=np.uint64)
        return arr, count, keep

    def solve(self, outages=None, load_scale=None, options: NrOptions | None = None, keep_state=True,
              plain=False, starts=None, count=None) -> BatchResult:
        opts = options or self.options
        arr, count, keep = self._instances(outages, load_scale, count, starts)
        res = np.zeros(count, RESULT_DTYPE)
        n = self.net.n_bus
        vm = np.empty((count, n)) if keep_state else None
        va = np.empty((count, n)) if keep_state else None
        fn = lib().gs_nr_solve_plain if plain else lib().gs_nr_solve_batch
        self.ctx.check(fn(self.h, C.c_int32(count), arr, C.byref(opts.c()), _p(res), _p(vm), _p(va)),
                       "gs_nr_solve_plain" if plain else "gs_nr_solve_batch")
        tot, ker, nl = C.c_double(), C.c_double(), C.c_int32()
        lib().gs_last_timing(self.h, C.byref(tot), C.byref(ker), C.byref(nl))
        return BatchResult(res, vm, va, ker.value, tot.value, nl.value)

    def screen(self, outages=None, load_scale=None, options: NrOptions | None = None, loading=False, flows=False,
               keep_state=False, starts=None, count=None, injections=False):
        """gs_nr_screen_batch: the solve and, on the device, line flows and the classification numbers of every
        instance. Returns (BatchResult, summaries [SUMMARY_DTYPE], loading [count, nb] or None, flows [count, nb, 4]
        or None); with injections, also (P, Q) [count, n] in MW / MVAr."""
        opts = options or self.options
        arr, count, keep = self._instances(outages, load_scale, count, starts)
        res = np.zeros(count, RESULT_DTYPE)
        summ = np.zeros(count, SUMMARY_DTYPE)
        nb, n = self.net.n_branch, self.net.n_bus
        ld = np.empty((count, nb)) if loading else None
        fl = np.empty((count, nb, 4)) if flows else None
        vm = np.empty((count, n)) if keep_state else None
        va = np.empty((count, n)) if keep_state else None
        pi = np.empty((count, n)) if injections else None
        qi = np.empty((count, n)) if injections else None
        self.ctx.check(lib().gs_nr_screen_batch(self.h, C.c_int32(count), arr, C.byref(opts.c()), _p(res), _p(summ),
                                                _p(ld), _p(fl), _p(vm), _p(va), _p(pi), _p(qi)),
                       "gs_nr_screen_batch")
        tot, ker, nl = C.c_double(), C.c_double(), C.c_int32()
        lib().gs_last_timing(self.h, C.byref(tot), C.byref(ker), C.byref(nl))
        out = (BatchResult(res, vm, va, ker.value, tot.value, nl.value), summ, ld, fl)
        return out + ((pi, qi),) if injections else out

    def close(self):
        if getattr(self, "h", None):
            lib().gs_network_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class LineFlow:
    """One branch's flows (LineFlow, NewtonShell.ComputeLineFlows :699-734): MW / MVAr at each end, the branch losses
    and the loading against rate A (percent, 0 when unrated)."""
    from_bus: int
    to_bus: int
    p_from_mw: float
    q_from_mvar: float
    p_to_mw: float
    q_to_mvar: float
    loss_mw: float
    loss_mvar: float
    rate_a_mva: float
    loading_pct: float


@dataclass
class PowerFlowResult:
    """The fields of the reference's PowerFlowResult (PowerFlowResult.cs:5-21, NewtonShell.AssembleResult :649-690)."""
    status: int
    converged: bool
    iterations: int
    max_mismatch: float
    vm: np.ndarray
    va_deg: np.ndarray
    symbolic_refreshes: int
    p_inj_mw: np.ndarray | None = None
    q_inj_mvar: np.ndarray | None = None
    total_loss_mw: float = 0.0
    total_loss_mvar: float = 0.0
    line_flows: list | None = None


class HipNewtonPowerFlow:
    """IPowerFlowSolver on the GPU: SparseNewtonRaphsonPowerFlow.Solve with Q limits and the DC retry off, and the
    result assembled as NewtonShell.AssembleResult does (line flows and losses computed on the device)."""

    name = "newton-raphson (hip, gfx950)"

    def __init__(self, ctx: Context, options: NrOptions | None = None):
        self.ctx = ctx
        self.options = options or NrOptions()

    def solve(self, net: Network, options: NrOptions | None = None) -> PowerFlowResult:
        opts = options or self.options
        dn = DeviceNetwork(self.ctx, net, opts)
        try:
            r, summ, ld, fl, (p_inj, q_inj) = dn.screen(count=1, options=opts, loading=True, flows=True,
                                                        keep_state=True, injections=True)
        finally:
            dn.close()
        rec = r.results[0]
        ids = net.bus_ids if net.bus_ids is not None else np.arange(net.n_bus)
        flows = [LineFlow(int(ids[net.br_from[k]]), int(ids[net.br_to[k]]), *(float(v) for v in fl[0, k]),
                          float(fl[0, k, 0] + fl[0, k, 2]), float(fl[0, k, 1] + fl[0, k, 3]),
                          float(net.br_rate_a[k]), float(ld[0, k])) for k in range(net.n_branch)]
        return PowerFlowResult(int(rec["status"]), int(rec["status"]) == CONVERGED, int(rec["iterations"]),
                               float(rec["max_mismatch"]), r.vm[0], np.degrees(r.va[0]),
                               int(rec["symbolic_refreshes"]), p_inj[0], q_inj[0], float(summ[0]["loss_mw"]),
                               float(summ[0]["loss_mvar"]), flows)


OVERLOAD_THRESHOLD_PCT = 100.0   # ContingencyAnalysis.OverloadThresholdPercent


@dataclass
class ContingencyResult:
    """ContingencyAnalysis's ContingencyResult (ContingencyAnalysis.cs:9-135) for a branch outage: the verdict, the
    maximum branch loading, Vmin / Vmax and the violation lines ScreenStateCore writes (:1150-1236)."""
    outage: str
    converged: bool
    max_branch_loading_pct: float
    min_vm: float
    max_vm: float
    violations: list
    islanded: bool = False

    @property
    def secure(self) -> bool:
        return self.converged and not self.violations


def branch_label(net: Network, k: int) -> str:
    """A branch named by its end buses, "from-to" (ScreenStateCore's label for an unnamed branch)."""
    ids = net.bus_ids if net.bus_ids is not None else np.arange(net.n_bus)
    return f"{int(ids[net.br_from[k]])}-{int(ids[net.br_to[k]])}"


def set_label(net: Network, outages) -> str:
    """An outage set named as ContingencyAnalysis.LabelMany names it: the branch labels joined by " + "."""
    return " + ".join(branch_label(net, int(k)) for k in np.atleast_1d(outages))


def classify(net: Network, outage, converged: bool, summary, loading_row) -> ContingencyResult:
    """One screened outage (a branch, or a set of branches) as ScreenStateCore classifies it: each branch above 100 %
    of rate A, then one line for the buses under their limits (the lowest voltage) and one for those over them (the
    highest). A non-converged outage reports whether it split the network (ContingencyAnalysis.NonConverged /
    Splits)."""
    label = set_label(net, outage)
    if not converged:
        if np.ndim(outage) == 0:
            split = bool(bridges(net)[outage]) if net.br_in_service[outage] else False
        else:
            split = splits(net, outage)
        return ContingencyResult(label, False, 0.0, 0.0, 0.0, ["islanded" if split else "no solution"], split)
    violations = []
    if summary["overloads"] > 0:
        for k in np.nonzero(loading_row > OVERLOAD_THRESHOLD_PCT)[0]:
            violations.append(f"branch {branch_label(net, int(k))} {loading_row[k]:.0f}%")
    if summary["undervoltage"] > 0:
        violations.append(f"undervoltage {summary['vmin']:.3f} p.u.")
    if summary["overvoltage"] > 0:
        violations.append(f"overvoltage {summary['vmax']:.3f} p.u.")
    return ContingencyResult(label, True, float(summary["max_loading_pct"]), float(summary["vmin"]),
                             float(summary["vmax"]), violations)


def screen_branches_batched(ctx: Context, net: Network, branch_indices=None, options: NrOptions | None = None,
                            keep_state=False):
    """N-1 screen of every in-service branch (or the given ones) in one batched launch with the classification on the
    device (ContingencyAnalysis.ScreenBranchesBatched, ContingencyAnalysis.Batched.cs:55-136). Returns
    (BatchResult, [ContingencyResult])."""
    idx = np.nonzero(net.br_in_service)[0] if branch_indices is None else np.asarray(branch_indices)
    for i in idx:
        if not net.br_in_service[int(i)]:
            raise ValueError(f"branch {int(i)} is out of service: removing it is not an outage.")
    dn = DeviceNetwork(ctx, net, options)
    try:
        res, summ, ld, _ = dn.screen(outages=idx, options=options, loading=True, keep_state=keep_state)
    finally:
        dn.close()
    out = [classify(net, int(k), bool(res.status[j] == CONVERGED), summ[j], ld[j]) for j, k in enumerate(idx)]
    return res, out


# ---------------------------------------------------------------- N-k screening

STRESS_VOLTAGE_WEIGHT_PER_PU = 1000.0     # ContingencyAnalysis.StressVoltageWeightPerPu
DEFAULT_CANDIDATE_COUNT = 20              # ContingencyAnalysis.DefaultCandidateCount
MAX_COMBINATIONS = 20_000_000             # ContingencyAnalysis.MaxCombinations
OPERATIONAL_LOW_PU, OPERATIONAL_HIGH_PU = 0.94, 1.06   # VoltageLimits.OperationalLowPu / HighPu
UNSOLVABLE_MARGIN_PU = 1e-6               # ContingencyAnalysis.UnsolvableMarginPu


def stress_score(c: ContingencyResult) -> float:
    """ContingencyAnalysis.StressScore (ContingencyAnalysis.cs:1073-1080): +inf for a non-converged outage, else the
    largest branch loading plus 1000 x the voltage excursion beyond the 0.94 - 1.06 pu band (no boundaries here)."""
    if not c.converged:
        return float("inf")
    s = c.max_branch_loading_pct
    s += max(0.0, OPERATIONAL_LOW_PU - c.min_vm) * STRESS_VOLTAGE_WEIGHT_PER_PU
    s += max(0.0, c.max_vm - OPERATIONAL_HIGH_PU) * STRESS_VOLTAGE_WEIGHT_PER_PU
    return s


def rank_candidates(single, candidate_count: int, minimum: int):
    """ContingencyAnalysis.RankCandidates (:474-486): the single outages that solve, most stressed first (a stable
    sort, ties in branch order), the top max(candidate_count, minimum) of them; the ones that do not solve are
    returned apart, in branch order. Returns (positions in `single`, excluded results)."""
    solvable = [i for i, c in enumerate(single) if c.converged]
    excluded = [c for c in single if not c.converged]
    take = min(max(candidate_count, minimum), len(solvable))
    ranked = sorted(solvable, key=lambda i: -stress_score(single[i]))   # sorted() is stable
    return ranked[:take], excluded


def combinations(items, k: int):
    """All k-subsets of items in the order ContingencyAnalysis.Combinations (:1851-1866) appends them: by the
    positions of their members, lexicographically."""
    out, pick = [], [0] * k

    def rec(start, depth):
        if depth == k:
            out.append(list(pick))
            return
        for i in range(start, len(items) - (k - depth) + 1):
            pick[depth] = int(items[i])
            rec(i + 1, depth + 1)

    if k <= len(items):
        rec(0, 0)
    return out


def provably_unsolvable(net: Network, outages, tolerance: float) -> bool:
    """ContingencyAnalysis.ProvablyUnsolvablePart (:1750-1829) for constant-power loads: after the outage, a connected
    part with no slack bus and no negative loss (no branch R < 0, no Gs < 0) whose specified net injection is below
    -(buses x tolerance + 1e-6) pu (demand beyond generation: losses cannot be negative), or, when the part is lossless,
    whose |net injection| reaches that margin. Such an outage is reported not converged without a solve."""
    if not (tolerance > 0 and np.isfinite(tolerance)) or not (net.base_mva > 0 and np.isfinite(net.base_mva)):
        return False
    n = net.n_bus
    parent = np.arange(n)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    gone = set(int(b) for b in np.atleast_1d(np.asarray(outages, np.int64)) if b >= 0)
    live = [k for k in range(net.n_branch) if net.br_in_service[k] and k not in gone]
    for k in live:
        a, c = find(int(net.br_from[k])), find(int(net.br_to[k]))
        if a != c:
            parent[a] = c
    pg = np.zeros(n)
    np.add.at(pg, net.gen_bus, net.gen_pg)
    root = np.array([find(i) for i in range(n)])
    has_slack = np.zeros(n, bool); neg_loss = np.zeros(n, bool); loss = np.zeros(n, bool)
    count = np.zeros(n, np.int64); inj = np.zeros(n)
    for i in range(n):
        r = root[i]
        count[r] += 1
        has_slack[r] |= net.bus_type[i] == SLACK
        neg_loss[r] |= net.gs[i] < 0.0
        loss[r] |= net.gs[i] != 0.0
        inj[r] += (pg[i] - net.pd[i]) / net.base_mva
    for k in live:
        r = root[int(net.br_from[k])]
        neg_loss[r] |= net.br_r[k] < 0.0
        loss[r] |= net.br_r[k] != 0.0
    for r in dict.fromkeys(root.tolist()):   # parts in the order of their first bus
        if has_slack[r] or neg_loss[r]:
            continue
        margin = count[r] * tolerance + UNSOLVABLE_MARGIN_PU
        if inj[r] <= -margin or (not loss[r] and abs(inj[r]) >= margin):
            return True
    return False


@dataclass
class NkScreenReport:
    """ContingencyAnalysis.NkScreenReport: the screened combinations worst first, and the single outages the pruned
    screen left out of the pairing because they already fail at N-1."""
    results: list
    excluded: list
    batch: object = None


def screen_outage_sets(ctx: Context, net: Network, sets, options: NrOptions | None = None, chunk: int = 4096):
    """ScreenOutageCore over outage sets (each a branch list, WithoutMany): the provably unsolvable ones reported not
    converged at once, the others solved and classified on the device in batches of `chunk`. Returns a
    ContingencyResult per set, in the order given."""
    options = options or NrOptions()
    out = [None] * len(sets)
    todo = []
    for j, st in enumerate(sets):
        if provably_unsolvable(net, st, options.tolerance):
            out[j] = classify(net, list(st), False, None, None)
        else:
            todo.append(j)
    if todo:
        dn = DeviceNetwork(ctx, net, options)
        try:
            for a in range(0, len(todo), chunk):
                idx = todo[a:a + chunk]
                res, summ, ld, _ = dn.screen(outages=[list(sets[j]) for j in idx], options=options, loading=True)
                for t, j in enumerate(idx):
                    out[j] = classify(net, list(sets[j]), bool(res.status[t] == CONVERGED), summ[t], ld[t])
        finally:
            dn.close()
    return out


def screen_n_minus_k(ctx: Context, net: Network, k: int, options: NrOptions | None = None,
                     candidate_count: int = DEFAULT_CANDIDATE_COUNT, exhaustive: bool = False) -> NkScreenReport:
    """ContingencyAnalysis.ScreenNMinusKReport (ContingencyAnalysis.cs:378-441): unless exhaustive, screen N-1 first,
    rank the single outages that solve by stress and keep the top candidate_count, then screen every k-subset of that
    set (the outage lists of one batched device solve); results worst first (a stable sort on StressScore)."""
    if k < 1:
        raise ValueError("contingency depth k must be >= 1")
    live = np.nonzero(net.br_in_service)[0]
    if k > len(live):
        return NkScreenReport([], [])
    if k == 1:
        single = screen_outage_sets(ctx, net, [[int(b)] for b in live], options)
        return NkScreenReport(sorted(single, key=lambda c: -stress_score(c)), [])
    excluded = []
    if exhaustive:
        candidates = [int(b) for b in live]
    else:
        single = screen_outage_sets(ctx, net, [[int(b)] for b in live], options)
        pos, excluded = rank_candidates(single, candidate_count, k)
        candidates = [int(live[i]) for i in pos]
    count = 1.0
    for c in range(k):
        count = count * (len(candidates) - c) / (c + 1)
    if count > MAX_COMBINATIONS:
        raise ValueError(f"N-{k}: {len(candidates)} candidate branches give about {count:.4g} outage combinations, "
                         f"more than the {MAX_COMBINATIONS} one screen takes")
    combos = combinations(candidates, k)
    results = screen_outage_sets(ctx, net, combos, options)
    return NkScreenReport(sorted(results, key=lambda c: -stress_score(c)), excluded)


# ---------------------------------------------------------------- multi-network batches and the autotuner

@dataclass
class Group:
    """One group of a Batch: instances of one uploaded network."""
    net: "DeviceNetwork"
    outages: object = None
    load_scale: object = None
    options: NrOptions | None = None
    starts: object = None          # (vm[count, n], va[count, n]) or None
    count: int | None = None


class Batch:
    """Instances of several networks solved together (gs_batch_*): each run launches every group's kernel on its own
    stream, or replays one captured hipGraph of those launches (use_graph)."""

    def __init__(self, ctx: Context, groups: list[Group], use_graph: bool = True):
        self.ctx = ctx
        self.groups = groups
        k = len(groups)
        nets = (C.c_void_p * k)(*[g.net.h.value for g in groups])
        counts = (C.c_int32 * k)()
        insts = (C.c_void_p * k)()
        opts = (GsNrOptions * k)()
        self._keep = []
        for i, g in enumerate(groups):
            arr, count, keep = g.net._instances(g.outages, g.load_scale, g.count, g.starts)
            self._keep += [arr, keep]
            counts[i] = count
            insts[i] = C.cast(arr, C.c_void_p)
            opts[i] = (g.options or g.net.options).c()
        self.counts = [int(c) for c in counts]
        h = C.c_void_p()
        ctx.check(lib().gs_batch_create(ctx.h, C.c_int32(k), nets, counts, insts, opts, C.c_int32(int(use_graph)),
                                        C.byref(h)), "gs_batch_create")
        self.h = h
        self._keep = []   # start points are in HBM now

    def run(self) -> float:
        """One launch of the whole batch; returns the device time in ms (HIP events)."""
        ms = C.c_double()
        self.ctx.check(lib().gs_batch_run(self.h, C.byref(ms)), "gs_batch_run")
        return ms.value

    def results(self, group: int, keep_state: bool = False) -> BatchResult:
        count = self.counts[group]
        n = self.groups[group].net.net.n_bus
        res = np.zeros(count, RESULT_DTYPE)
        vm = np.empty((count, n)) if keep_state else None
        va = np.empty((count, n)) if keep_state else None
        self.ctx.check(lib().gs_batch_results(self.h, C.c_int32(group), _p(res), _p(vm), _p(va)), "gs_batch_results")
        return BatchResult(res, vm, va, 0.0, 0.0, 1)

    def close(self):
        if getattr(self, "h", None):
            lib().gs_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(dn: "DeviceNetwork", outages=None, load_scale=None, options: NrOptions | None = None, starts=None,
             count=None) -> tuple[NrOptions, float]:
    """gs_nr_autotune on a sample of instances: the fastest arithmetic / launch slicing whose verdicts equal the FP64
    LU's. Returns the chosen options and their kernel time on the sample (ms)."""
    base = options or dn.options
    arr, count, keep = dn._instances(outages, load_scale, count, starts)
    o = base.c()
    ms = C.c_double()
    dn.ctx.check(lib().gs_nr_autotune(dn.h, C.c_int32(count), arr, C.byref(o), C.byref(ms)), "gs_nr_autotune")
    chosen = NrOptions(o.tolerance, o.max_iterations, bool(o.flat_start), o.diverge_limit, o.lu_precision,
                       o.refine_steps, o.growth_tol, o.schedule, o.flags)
    return chosen, ms.value


class Server:
    """Continuous-batching request loop over resident networks (gs_server_*): submit() queues a power-flow or
    single-contingency request and returns a ticket at once; a dispatcher thread in the library runs up to four slots,
    each a batch of one network stepped a few Newton iterations per launch, and between launches finished requests
    leave and waiting ones of the same network take their positions; wait() returns the result of one ticket."""

    def __init__(self, ctx: Context, max_batch: int = 0, streams: int = 0, queue_capacity: int = 0,
                 state_bytes: int = 0, pinned_bytes: int = 0):
        self.ctx = ctx
        cfg = GsServerConfig(max_batch, streams, queue_capacity, 0, state_bytes, pinned_bytes)
        h = C.c_void_p()
        ctx.check(lib().gs_server_create(ctx.h, C.byref(cfg), C.byref(h)), "gs_server_create")
        self.h = h
        self.nets: list[DeviceNetwork] = []

    def add_network(self, dn: "DeviceNetwork", options: NrOptions | None = None) -> int:
        nid = C.c_int32()
        opts = (options or dn.options).c()
        self.ctx.check(lib().gs_server_add_network(self.h, dn.h, C.byref(opts), C.byref(nid)), "gs_server_add_network")
        self.nets.append(dn)
        return nid.value

    def submit(self, network_id: int, outage=-1, load_scale: float = 1.0, start=None) -> int:
        """Queues one request (outage: a branch index, -1, or a list of branches for an N-2 request); raises
        GsError(GS_BUSY) when the queue is full. The library copies the start point and the outage list."""
        vm0 = va0 = None
        if start is not None:
            vm0 = np.ascontiguousarray(start[0], np.float64)
            va0 = np.ascontiguousarray(start[1], np.float64)
        more = None
        if np.ndim(outage) > 0:
            lst = np.atleast_1d(np.asarray(outage, np.int32))
            outage, more = (int(lst[0]), np.ascontiguousarray(lst[1:])) if len(lst) else (-1, None)
        inst = GsInstance(int(outage), 0 if more is None else len(more), float(load_scale), _p(vm0), _p(va0),
                          _p(more) if more is not None and len(more) else None)
        t = C.c_uint64()
        st = lib().gs_server_submit(self.h, C.c_int32(network_id), C.byref(inst), C.byref(t))
        if st != GS_OK:
            raise GsError(st, "gs_server_submit")
        return t.value

    def wait(self, ticket: int, network_id: int, timeout_ms: int = -1, keep_state: bool = True,
             injections: bool = False):
        """(result record, vm, va (radians), latency ms) of a finished request, with (P, Q) in MW / MVAr appended when
        injections; None when timeout_ms ran out."""
        n = self.nets[network_id].net.n_bus
        res = np.zeros(1, RESULT_DTYPE)
        vm = np.empty(n) if keep_state else None
        va = np.empty(n) if keep_state else None
        pi = np.empty(n) if injections else None
        qi = np.empty(n) if injections else None
        lat = C.c_double()
        st = lib().gs_server_wait(self.h, C.c_uint64(ticket), C.c_int32(timeout_ms), _p(res), _p(vm), _p(va),
                                  _p(pi), _p(qi), C.byref(lat))
        if st == GS_PENDING:
            return None
        if st != GS_OK:
            raise GsError(st, f"gs_server_wait (ticket {ticket})")
        out = (res[0], vm, va, lat.value)
        return out + (pi, qi) if injections else out

    def stats(self) -> dict:
        s = GsServerStats()
        self.ctx.check(lib().gs_server_stats_get(self.h, C.byref(s)), "gs_server_stats_get")
        return {k: getattr(s, k) for k, _ in GsServerStats._fields_}

    def reset_stats(self):
        self.ctx.check(lib().gs_server_stats_reset(self.h), "gs_server_stats_reset")

    def close(self):
        if getattr(self, "h", None):
            lib().gs_server_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
