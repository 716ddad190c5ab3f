"""The solve export document (gridsim_amd/export.py) and the reference comparer's rule (tools/run-gpu.sh:27-50)."""
import json

import numpy as np

from conftest import matpower_case


def _oracle_result(oracle, net):
    from gridsim_amd import engine
    rec, vm, va = oracle.solve(net)
    p, q, _ = oracle.mismatch(net, vm, va)
    pf, qf, pt, qt, load = oracle.line_flows(net, vm, va)
    ids = net.bus_ids
    flows = [engine.LineFlow(int(ids[net.br_from[k]]), int(ids[net.br_to[k]]), pf[k], qf[k], pt[k], qt[k],
                             pf[k] + pt[k], qf[k] + qt[k], net.br_rate_a[k], load[k]) for k in range(net.n_branch)]
    return engine.PowerFlowResult(int(rec["status"]), rec["status"] == 0, int(rec["iterations"]),
                                  float(rec["max_mismatch"]), vm, np.degrees(va), 0, p * net.base_mva,
                                  q * net.base_mva, float((p * net.base_mva).sum()), float((q * net.base_mva).sum()),
                                  flows)


def test_export_shape_and_comparer(oracle, tmp_path):
    from gridsim_amd import export
    net = matpower_case("case9")
    r = _oracle_result(oracle, net)
    doc = export.solve_export(net, r, case="case9", backend="cpu (oracle)")
    path = tmp_path / "case9.json"
    export.write_solve_export(str(path), doc)
    back = json.load(open(path))
    assert back["schema"] == "gridsim-solve-export/1"
    buses = back["result"]["buses"]
    assert [b["id"] for b in buses] == list(range(1, 10)) and buses[0]["vm"] == 1.04
    assert abs(back["result"]["totalLossMw"] - 4.641) < 0.005
    assert len(back["result"]["lineFlows"]) == 9 and back["solve"]["status"] == "Converged"
    ok, dvm, dva = export.compare_exports(back, doc)
    assert ok and dvm == 0 and dva == 0
    worse = json.loads(json.dumps(doc))
    worse["result"]["buses"][4]["vaDeg"] += 2e-4
    assert not export.compare_exports(doc, worse)[0]
