"""The solve export document (``gridsim-solve-export/1``) for a solved power flow.

The reference writes one JSON document per traced solve (src/GridSim.Core/IO/SolveExport.cs:109-210): ``schema``,
``meta`` (case, backend, commit, time, elapsed, machine), ``model``, ``solve`` (solver, size, converged, status,
iterations, max mismatch, and its per-iteration trace), ``result`` (total losses, per bus ``id``, ``vm``, ``vaDeg``,
``pInjMw``, ``qInjMvar``, per branch the line flow rows) and ``physics``. Its CPU-versus-GPU comparer
(tools/run-gpu.sh:27-50) reads ``result.buses[].{id, vm, vaDeg}``. This writer emits the same ``schema``, ``meta``,
``solve`` summary and ``result`` members with the reference's names and units; the ``model`` block carries the
network's name, base and sizes only, and the per-iteration trace and the physics audit are not written (the fields
are additive in the schema, so a reader of the reference's documents reads these).
"""
from __future__ import annotations

import json
import platform

from .engine import STATUS_NAMES, PowerFlowResult
from .network import Network


def solve_export(net: Network, result: PowerFlowResult, case: str, backend: str, elapsed_ms: float = 0.0,
                 git_commit: str = "", timestamp_utc: str = "") -> dict:
    ids = net.bus_ids if net.bus_ids is not None else range(net.n_bus)
    flows = result.line_flows or []
    return {
        "schema": "gridsim-solve-export/1",
        "meta": {"Case": case, "Backend": backend, "GitCommit": git_commit, "TimestampUtc": timestamp_utc,
                 "ElapsedMs": float(elapsed_ms), "Machine": platform.node()},
        "model": {"name": net.name, "baseMva": float(net.base_mva), "busCount": int(net.n_bus),
                  "branchCount": int(net.n_branch), "generatorCount": int(net.n_gen)},
        "solve": {"solver": "newton-raphson (hip, gfx950)", "n": int(net.n_bus), "converged": bool(result.converged),
                  "status": STATUS_NAMES.get(result.status, str(result.status)),
                  "iterations": int(result.iterations), "maxMismatch": float(result.max_mismatch)},
        "result": {
            "totalLossMw": float(result.total_loss_mw),
            "totalLossMvar": float(result.total_loss_mvar),
            "buses": [{"id": int(b), "vm": float(result.vm[i]), "vaDeg": float(result.va_deg[i]),
                       "pInjMw": float(result.p_inj_mw[i]) if result.p_inj_mw is not None else 0.0,
                       "qInjMvar": float(result.q_inj_mvar[i]) if result.q_inj_mvar is not None else 0.0}
                      for i, b in enumerate(ids)],
            "lineFlows": [{"from": f.from_bus, "to": f.to_bus, "pFromMw": f.p_from_mw, "qFromMvar": f.q_from_mvar,
                           "pToMw": f.p_to_mw, "qToMvar": f.q_to_mvar, "pLossMw": f.loss_mw,
                           "qLossMvar": f.loss_mvar, "name": "", "ratingMva": f.rate_a_mva,
                           "loadingPercent": f.loading_pct} for f in flows],
        },
    }


def write_solve_export(path: str, doc: dict) -> None:
    with open(path, "w", encoding="utf-8") as fh:
        json.dump(doc, fh, indent=2)


def compare_exports(cpu: dict, gpu: dict, vm_tol: float = 1e-6, va_tol: float = 1e-4):
    """The reference comparer's rule (tools/run-gpu.sh:27-50): the same bus set, max |dVm| <= vm_tol pu and max |dVa|
    <= va_tol degrees. Returns (ok, max |dVm|, max |dVa|)."""
    a = {b["id"]: b for b in cpu["result"]["buses"]}
    b = {x["id"]: x for x in gpu["result"]["buses"]}
    if a.keys() != b.keys():
        return False, float("inf"), float("inf")
    dvm = max(abs(a[i]["vm"] - b[i]["vm"]) for i in a)
    dva = max(abs(a[i]["vaDeg"] - b[i]["vaDeg"]) for i in a)
    return dvm <= vm_tol and dva <= va_tol, dvm, dva
