"""The gridsim-rpc/1 front door (gridsim_amd/rpc.py): request and response shapes of the reference's host
(src/GridSim.Cli/RpcCommand.cs:46-51, :460-497, :590-700, :880-913), run here over a stand-in backend that solves with
the oracle (no GPU); the same session over the GPU server is in test_server_gpu.py."""
import io
import json

import numpy as np
import pytest

from conftest import GOLDEN


class OracleBackend:
    """Stand-in for rpc.GpuBackend in CPU tests: each submitted request solved by the oracle."""

    def __init__(self, oracle):
        self.o = oracle
        self.nets = []
        self.jobs = {}
        self.next = 1

    def load(self, net):
        self.nets.append(net)
        return len(self.nets) - 1

    def submit(self, nid, outage, scale):
        t = self.next
        self.next += 1
        self.jobs[t] = (nid, outage, scale)
        return t

    def wait(self, ticket, nid):
        nid, outage, scale = self.jobs.pop(ticket)
        net = self.nets[nid].with_load_multiplier(scale) if scale != 1.0 else self.nets[nid]
        rec, vm, va = self.o.solve(net, self.o.options(max_iterations=100), skip=outage)
        p, q, _ = self.o.mismatch(net, vm, va, skip=outage)
        return {"status": int(rec["status"]), "iterations": int(rec["iterations"]),
                "max_mismatch": float(rec["max_mismatch"]), "vm": vm, "va": va, "p_mw": p * net.base_mva,
                "q_mvar": q * net.base_mva}

    def stats(self):
        return {"completed": 0, "launches": 0, "mean_batch": 0.0, "latency_p50_ms": 0.0, "latency_p99_ms": 0.0}


def _session(oracle, lines):
    from gridsim_amd import rpc
    host = rpc.RpcHost(OracleBackend(oracle), base_dir=GOLDEN)
    out = io.StringIO()
    host.serve(io.StringIO("\n".join(json.dumps(l) if isinstance(l, dict) else l for l in lines) + "\n"), out)
    return [json.loads(x) for x in out.getvalue().splitlines()]


def test_session_shapes_and_case9_solve(oracle):
    r = _session(oracle, [
        {"op": "hello", "id": 1},
        {"op": "load", "id": 2, "case": "matpower/case9.m"},
        {"op": "solve", "id": 3, "handle": "case-1"},
        {"op": "solve", "id": 4, "handle": "case-1", "outage": 1, "load_scale": 1.1},
        {"op": "stats", "id": 5},
        {"op": "release", "id": 6, "handle": "case-1"},
        {"op": "solve", "id": 7, "handle": "case-1"},
        {"op": "shutdown", "id": 8},
        {"op": "hello", "id": 9},
    ])
    assert [x["id"] for x in r] == [1, 2, 3, 4, 5, 6, 7, 8]   # nothing is answered after shutdown
    for x in r:
        assert x["protocol"] == "gridsim-rpc/1" and set(x) >= {"id", "op", "status", "exitCode", "case_sha256"}
    assert r[0]["result"]["operations"] == ["hello", "load", "release", "solve", "stats", "shutdown"]
    load = r[1]["result"]
    assert load["handle"] == "case-1" and load["buses"] == 9 and load["branches"] == 9
    assert len(r[1]["case_sha256"]) == 64 and load["case_identity"] == "file"
    s = r[2]["result"]
    assert r[2]["status"] == "ok" and r[2]["exitCode"] == 0 and s["converged"] is True
    assert s["bus"] == list(range(1, 10)) and abs(s["vm_pu"][0] - 1.04) < 1e-12
    assert abs(s["loss_mw"] - 4.641) < 0.005 and abs(s["p_inj_mw"][0] - 71.64) < 0.05
    assert abs(s["loss_mw"] - sum(s["p_inj_mw"])) < 1e-9
    assert r[3]["result"]["converged"] is True and r[3]["result"]["loss_mw"] != s["loss_mw"]
    assert r[4]["result"]["solves"] == 2 and r[4]["result"]["loads"] == 1 and r[4]["result"]["handles"] == 1
    assert r[5]["result"]["released"] is True
    assert r[6]["status"] == "error" and r[6]["error"]["kind"] == "unknown-handle" and r[6]["exitCode"] == 2
    assert r[7]["result"]["stopping"] is True


def test_session_errors(oracle):
    r = _session(oracle, [
        "this is not json",
        {"op": "frobnicate", "id": "a"},
        {"op": "hello", "id": "b", "extra": 1},
        {"op": "load", "id": "c"},
        {"op": "load", "id": "d", "case": "no-such-case.m"},
        {"op": "load", "id": "e", "case": "matpower/case9.m"},
        {"op": "solve", "id": "f", "handle": "case-1", "outage": 99},
        '{"op": "hello", "op": "stats"}',
        {"id": [1], "op": "hello"},
    ])
    kinds = [x.get("error", {}).get("kind") for x in r]
    assert kinds == ["malformed-request", "unknown-operation", "bad-argument", "bad-argument", "case-load-failed",
                     None, "bad-argument", "malformed-request", "malformed-request"]
    assert r[1]["id"] == "a" and r[1]["op"] == "frobnicate"


def test_numbers_are_written_in_round_trip_form():
    from gridsim_amd import rpc
    assert rpc._encode({"a": 1.0, "b": 0.1, "c": 1e-10, "d": float("nan"), "e": -float("inf"), "f": 3}) == \
        '{"a":1,"b":0.1,"c":1E-10,"d":"NaN","e":"-Infinity","f":3}'
