"""A MATPOWER version-2 case reader (``mpc.baseMVA``, ``mpc.bus``, ``mpc.gen``, ``mpc.branch``).

Follows the reference's reading rules (src/GridSim.Core/IO/Matpower/MatpowerParser.cs):
* BUS_TYPE 3 -> Slack, 2 -> PV, 1 -> PQ; 4 (isolated) buses are dropped with every branch touching them (:188-199);
* status-0 branches are dropped (:359-360); status-0 generators are kept out of the solve (:248-309);
* a PV bus with no in-service generator is read as PQ (:342-347);
* TAP 0 means a line (:368-379).
Buses are renumbered to positions 0..n-1 in file order; ``bus_ids`` keeps the file's ids.
"""
from __future__ import annotations

import re

import numpy as np

from .network import Network, PQ, PV, SLACK


def _matrix(text: str, name: str) -> np.ndarray:
    m = re.search(r"mpc\." + name + r"\s*=\s*\[(.*?)\];", text, re.S)
    if not m:
        raise ValueError(f"mpc.{name} not found")
    rows = []
    for line in m.group(1).splitlines():
        line = line.split("%", 1)[0].strip().rstrip(";").strip()
        if not line:
            continue
        for part in line.split(";"):
            part = part.strip()
            if part:
                rows.append([float(v) for v in part.replace(",", " ").split()])
    width = max(len(r) for r in rows)
    return np.array([r + [0.0] * (width - len(r)) for r in rows], dtype=np.float64)


def read_case(path: str) -> Network:
    with open(path, encoding="utf-8") as fh:
        return read_text(fh.read())


def read_text(text: str) -> Network:
    """A case from the text of a MATPOWER file."""
    base = float(re.search(r"mpc\.baseMVA\s*=\s*([0-9.eE+-]+)", text).group(1))
    bus = _matrix(text, "bus")
    gen = _matrix(text, "gen")
    br = _matrix(text, "branch")
    keep_bus = bus[:, 1] != 4
    bus = bus[keep_bus]
    ids = bus[:, 0].astype(np.int64)
    index = {int(b): i for i, b in enumerate(ids)}
    gen = gen[(gen[:, 7] > 0) & np.isin(gen[:, 0].astype(np.int64), ids)]
    br = br[(br[:, 10] > 0) & np.isin(br[:, 0].astype(np.int64), ids) & np.isin(br[:, 1].astype(np.int64), ids)]
    t = bus[:, 1].astype(np.int64)
    bus_type = np.where(t == 3, SLACK, np.where(t == 2, PV, PQ)).astype(np.int32)
    gen_bus = np.array([index[int(b)] for b in gen[:, 0]], dtype=np.int32)
    has_gen = np.zeros(len(ids), dtype=bool)
    has_gen[gen_bus] = True
    bus_type[(bus_type == PV) & ~has_gen] = PQ
    ratio = br[:, 8]
    net = Network(
        base_mva=base, bus_type=bus_type, pd=bus[:, 2], qd=bus[:, 3], gs=bus[:, 4], bs=bus[:, 5],
        vm=bus[:, 7], va_deg=bus[:, 8], bus_vmax=bus[:, 11], bus_vmin=bus[:, 12],
        br_from=np.array([index[int(b)] for b in br[:, 0]]), br_to=np.array([index[int(b)] for b in br[:, 1]]),
        br_r=br[:, 2], br_x=br[:, 3], br_b=br[:, 4], br_tap=np.where(ratio != 0.0, ratio, 1.0),
        br_shift_deg=br[:, 9], br_in_service=np.ones(len(br), dtype=np.uint8), br_rate_a=br[:, 5],
        gen_bus=gen_bus, gen_pg=gen[:, 1], gen_qg=gen[:, 2], gen_vg=gen[:, 5], gen_qmin=gen[:, 4],
        gen_qmax=gen[:, 3], name=re.search(r"function\s+mpc\s*=\s*(\w+)", text).group(1), bus_ids=ids)
    return net
