"""The gridsim-rpc/1 line-JSON front door in front of the continuous-batching server.

The reference's host (src/GridSim.Cli/RpcCommand.cs) reads one JSON request per line and writes one JSON response per
line: ``hello``, ``load`` (a case held by handle), ``release``, ``solve`` (the power flow of a held case), ``stats``
and ``shutdown`` (:46-51, :460-497); it solves each request to completion before it reads the next (:668-700). This
host keeps the same request and response shapes (:880-913: ``protocol``, ``id``, ``op``, ``status``, ``exitCode``,
``case_sha256``, then ``result`` or ``error``) but does not wait: a ``solve`` is submitted to the GPU server as soon
as its line is read, so solves that arrive while others are on the device are batched with them, and the responses
are written in request order as their results come back.

Differences, all additive or stated: ``solve`` also takes ``outage`` (a branch index: the case with that branch out)
and ``load_scale`` (every load and generator P scaled); ``stats`` also reports the server's latency figures; a
``load`` takes a MATPOWER file or ``synthetic:<buses>:<seed>`` (this project's generator; the reference's built-in
case names are not shipped); reactive limits are not enforced (the engine's scope). ``run-verb`` and
``line-constants`` run other CLI verbs and are not served. Run as ``python -m gridsim_amd.rpc`` (stdin / stdout).
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import queue
import sys
import threading

import numpy as np

PROTOCOL = "gridsim-rpc/1"
TOOL = "gridsim-hip 0.1 (gfx950)"
OPERATIONS = ["hello", "load", "release", "solve", "stats", "shutdown"]
MAX_LINE_CHARS = 16 * 1024 * 1024

EXIT_OK, EXIT_INPUT_ERROR, EXIT_USAGE, EXIT_VERDICT = 0, 1, 2, 3   # CliExit


class RpcError(Exception):
    def __init__(self, kind, message, exit_code, sha=None):
        super().__init__(message)
        self.kind, self.message, self.exit_code, self.sha = kind, message, exit_code, sha


def number(d: float):
    """A double as the reference writes it (CliResultWriter.Number): round-trip digits; NaN / Infinity as strings."""
    d = float(d)
    if math.isnan(d):
        return "NaN"
    if math.isinf(d):
        return "Infinity" if d > 0 else "-Infinity"
    return _Raw(repr(d)[:-2] if repr(d).endswith(".0") else repr(d).replace("e", "E"))


class _Raw(str):
    """A number already spelled; written into the JSON text as it stands."""


def _encode(obj) -> str:
    """Compact JSON (no spaces, as the reference's writer), numbers spelled by number()."""
    if isinstance(obj, _Raw):
        return str(obj)
    if isinstance(obj, dict):
        return "{" + ",".join(json.dumps(k, ensure_ascii=False) + ":" + _encode(v) for k, v in obj.items()) + "}"
    if isinstance(obj, (list, tuple)):
        return "[" + ",".join(_encode(v) for v in obj) + "]"
    if isinstance(obj, float):
        n = number(obj)
        return n if isinstance(n, _Raw) else json.dumps(n)
    return json.dumps(obj, ensure_ascii=False)


def respond(id_, op, status, exit_code, sha, result=None, error=None) -> str:
    body = {"protocol": PROTOCOL, "id": id_, "op": op, "status": status, "exitCode": exit_code, "case_sha256": sha}
    if error is None:
        body["result"] = result or {}
    else:
        body["error"] = {"kind": error.kind, "message": error.message}
    return _encode(body)


def _allow(req, op, *keys):
    for k in req:
        if k not in ("id", "op") and k not in keys:
            more = f" (it takes: {', '.join(keys)})." if keys else "."
            raise RpcError("bad-argument", f"operation '{op}' does not take '{k}'" + more, EXIT_USAGE)


def _require_string(req, key):
    v = req.get(key)
    if not isinstance(v, str) or not v:
        raise RpcError("bad-argument", f'"{key}" is required and must be a non-empty string.', EXIT_USAGE)
    return v


class GpuBackend:
    """The GPU side: each loaded case uploaded once (Y-bus, symbolic plan, voltages resident in HBM) and registered
    with one continuous-batching server."""

    def __init__(self, ctx=None, **server_config):
        from . import engine
        self.engine = engine
        self.ctx = ctx or engine.Context(0)
        self.server = engine.Server(self.ctx, **server_config)
        self.options = engine.NrOptions(max_iterations=100, lu_precision=engine.LU_FP64, refine_steps=0)

    def load(self, net):
        dn = self.engine.DeviceNetwork(self.ctx, net, self.options)
        return self.server.add_network(dn, self.options)

    def submit(self, nid, outage, load_scale):
        return self.server.submit(nid, outage, load_scale)

    def wait(self, ticket, nid):
        rec, vm, va, _lat, p, q = self.server.wait(ticket, nid, injections=True)
        return {"status": int(rec["status"]), "iterations": int(rec["iterations"]),
                "max_mismatch": float(rec["max_mismatch"]), "vm": vm, "va": va, "p_mw": p, "q_mvar": q}

    def stats(self):
        return self.server.stats()

    def close(self):
        self.server.close()


class _Held:
    def __init__(self, arg, net, sha, identity, nid):
        self.arg, self.net, self.sha, self.identity, self.nid = arg, net, sha, identity, nid


class RpcHost:
    """One gridsim-rpc/1 session over a backend (GpuBackend in production)."""

    def __init__(self, backend, base_dir="."):
        self.backend = backend
        self.base_dir = base_dir
        self.cases: dict[str, _Held] = {}
        self.next_handle = 1
        self.requests = self.loads = self.solves = 0

    # ---- the case loader
    def _load_case(self, arg):
        from . import gridgen, matpower
        if arg.startswith("synthetic:"):
            try:
                _, n, seed = arg.split(":")
                net = gridgen.sized_grid(int(n), int(seed))
            except (ValueError, TypeError) as ex:
                raise RpcError("case-load-failed", f"error loading case '{arg}': {ex}", EXIT_USAGE)
            sha = hashlib.sha256(f"built-in:{arg}:{TOOL}".encode()).hexdigest()
            return net, sha, "built-in-name"
        path = os.path.join(self.base_dir, arg)
        if not os.path.isfile(path):
            raise RpcError("case-load-failed", f"error loading case '{arg}': no such file.", EXIT_USAGE)
        try:
            net = matpower.read_case(path)
        except Exception as ex:  # the reader's own errors, reported as the reference reports a failed load
            raise RpcError("case-load-failed", f"error loading case '{arg}': {ex}", EXIT_INPUT_ERROR)
        with open(path, "rb") as fh:
            sha = hashlib.sha256(fh.read()).hexdigest()
        return net, sha, "file"

    def _held(self, req):
        handle = _require_string(req, "handle")
        if handle not in self.cases:
            raise RpcError("unknown-handle",
                           f"no case is held under handle '{handle}' (it was never issued or has been released).",
                           EXIT_USAGE)
        return handle, self.cases[handle]

    # ---- one request line: a response string, or a pending solve the writer completes
    def handle(self, line):
        self.requests += 1
        id_, op = None, ""
        try:
            if len(line) > MAX_LINE_CHARS:
                raise RpcError("malformed-request", f"the request line exceeds {MAX_LINE_CHARS} characters.",
                               EXIT_USAGE)
            try:
                req = json.loads(line, object_pairs_hook=_no_duplicates)
            except ValueError as ex:
                raise RpcError("malformed-request", "the line is not valid JSON: " + str(ex), EXIT_USAGE)
            if not isinstance(req, dict):
                raise RpcError("malformed-request", "a request must be a JSON object.", EXIT_USAGE)
            if "id" in req:
                v = req["id"]
                if isinstance(v, bool) or not (v is None or isinstance(v, (str, int, float))):
                    raise RpcError("malformed-request", "id must be a string, a number or null.", EXIT_USAGE)
                id_ = v
            if not isinstance(req.get("op"), str):
                raise RpcError("malformed-request", 'a request needs a string "op".', EXIT_USAGE)
            op = req["op"]
            return self.dispatch(req, id_, op)
        except RpcError as e:
            return respond(id_, op, "error", e.exit_code, e.sha, error=e), False

    def dispatch(self, req, id_, op):
        if op == "hello":
            _allow(req, op)
            return respond(id_, op, "ok", 0, None, {"protocol": PROTOCOL, "tool": TOOL, "build_status": "source",
                                                    "operations": OPERATIONS}), False
        if op == "stats":
            _allow(req, op)
            st = self.backend.stats()
            return respond(id_, op, "ok", 0, None, {
                "requests": self.requests, "loads": self.loads, "solves": self.solves, "verb_runs": 0,
                "handles": len(self.cases),
                "server": {"completed": int(st["completed"]), "launches": int(st["launches"]),
                           "mean_batch": float(st["mean_batch"]), "latency_p50_ms": float(st["latency_p50_ms"]),
                           "latency_p99_ms": float(st["latency_p99_ms"])}}), False
        if op == "shutdown":
            _allow(req, op)
            return respond(id_, op, "ok", 0, None, {"stopping": True}), True
        if op == "load":
            _allow(req, op, "case")
            arg = _require_string(req, "case")
            net, sha, identity = self._load_case(arg)
            nid = self.backend.load(net)
            handle = f"case-{self.next_handle}"
            self.next_handle += 1
            self.cases[handle] = _Held(arg, net, sha, identity, nid)
            self.loads += 1
            return respond(id_, op, "ok", 0, sha, {"handle": handle, "case": net.name, "buses": net.n_bus,
                                                   "branches": net.n_branch, "case_identity": identity}), False
        if op == "release":
            _allow(req, op, "handle")
            handle, held = self._held(req)
            del self.cases[handle]
            return respond(id_, op, "ok", 0, held.sha, {"released": True}), False
        if op == "solve":
            _allow(req, op, "handle", "outage", "load_scale")
            _, held = self._held(req)
            outage = req.get("outage", -1)
            scale = req.get("load_scale", 1.0)
            if not isinstance(outage, int) or isinstance(outage, bool) or not -1 <= outage < held.net.n_branch:
                raise RpcError("bad-argument", f'"outage" must be a branch index below {held.net.n_branch}.',
                               EXIT_USAGE, held.sha)
            if not isinstance(scale, (int, float)) or isinstance(scale, bool) or not (scale > 0 and
                                                                                   math.isfinite(scale)):
                raise RpcError("bad-argument", '"load_scale" must be a number above 0.', EXIT_USAGE, held.sha)
            if outage >= 0 and not held.net.br_in_service[outage]:
                raise RpcError("bad-argument", f"branch {outage} is out of service: removing it is not an outage.",
                               EXIT_USAGE, held.sha)
            ticket = self.backend.submit(held.nid, outage, float(scale))
            self.solves += 1
            return _PendingSolve(id_, held, ticket), False
        raise RpcError("unknown-operation",
                       f"unknown operation '{op}'; the operations are: {', '.join(OPERATIONS)}.", EXIT_USAGE)

    def solve_response(self, pending) -> str:
        held = pending.held
        try:
            r = self.backend.wait(pending.ticket, held.nid)
        except Exception as ex:  # a failed launch is this request's error, not the session's
            return respond(pending.id, "solve", "error", EXIT_INPUT_ERROR, held.sha,
                           error=RpcError("solve-failed", str(ex), EXIT_INPUT_ERROR))
        converged = r["status"] == 0
        vm, va_deg = r["vm"], np.degrees(r["va"])
        ids = held.net.bus_ids if held.net.bus_ids is not None else np.arange(held.net.n_bus)
        loss_mw, loss_mvar = 0.0, 0.0
        for v in r["p_mw"]:
            loss_mw += float(v)
        for v in r["q_mvar"]:
            loss_mvar += float(v)
        return respond(pending.id, "solve", "ok", EXIT_OK if converged else EXIT_VERDICT, held.sha, {
            "case": held.net.name, "converged": bool(converged), "iterations": r["iterations"],
            "max_mismatch_pu": float(r["max_mismatch"]), "loss_mw": loss_mw, "loss_mvar": loss_mvar,
            "bus": [int(b) for b in ids], "vm_pu": [float(v) for v in vm], "va_deg": [float(v) for v in va_deg],
            "p_inj_mw": [float(v) for v in r["p_mw"]], "q_inj_mvar": [float(v) for v in r["q_mvar"]]})

    # ---- the session loop
    def serve(self, inp, out):
        """Reads request lines until end of input or shutdown; responses are written in request order by a writer
        thread, which completes each pending solve when its result is back."""
        pending = queue.Queue()

        def writer():
            while True:
                item = pending.get()
                if item is None:
                    return
                text = self.solve_response(item) if isinstance(item, _PendingSolve) else item
                out.write(text + "\n")
                out.flush()

        th = threading.Thread(target=writer, daemon=True)
        th.start()
        try:
            for line in inp:
                line = line.rstrip("\r\n")
                if not line.strip():
                    continue
                item, stop = self.handle(line)
                pending.put(item)
                if stop:
                    break
        finally:
            pending.put(None)
            th.join()


class _PendingSolve:
    def __init__(self, id_, held, ticket):
        self.id, self.held, self.ticket = id_, held, ticket


def _no_duplicates(pairs):
    seen = {}
    for k, v in pairs:
        if k in seen:
            raise ValueError(f"duplicate property '{k}'")
        seen[k] = v
    return seen


def main(argv=None):
    backend = GpuBackend()
    try:
        RpcHost(backend).serve(sys.stdin, sys.stdout)
    finally:
        backend.close()


if __name__ == "__main__":
    main()
