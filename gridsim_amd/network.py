"""Network data held as flat numpy arrays, the shape the C ABI takes.

A network mirrors the reference's ``GridModel`` records (src/GridSim.Core/Model/Model.cs:20-35 ``Bus``,
:91-115 ``Generator``, :253-276 ``Branch``) with buses numbered by position: bus ``i`` has id ``i``, so the
reference's id->index map (SparseNewtonRaphsonPowerFlow.cs:75-76) is the identity.

Bus types use the reference's ``BusType`` enum order (Model.cs:4-12): 0 = Slack, 1 = PV, 2 = PQ.
Loads are in MW / MVAr, shunts in MW / MVAr at 1 pu, angles in degrees, as the reference stores them.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

SLACK, PV, PQ = 0, 1, 2


@dataclass
class Network:
    base_mva: float
    bus_type: np.ndarray      # int32 [n]
    pd: np.ndarray            # MW
    qd: np.ndarray            # MVAr
    gs: np.ndarray            # MW at 1 pu
    bs: np.ndarray            # MVAr at 1 pu
    vm: np.ndarray            # pu (start / stored)
    va_deg: np.ndarray        # degrees (start / stored)
    br_from: np.ndarray       # int32 [nb]
    br_to: np.ndarray         # int32 [nb]
    br_r: np.ndarray
    br_x: np.ndarray
    br_b: np.ndarray
    br_tap: np.ndarray        # 0 or 1 = a line
    br_shift_deg: np.ndarray
    br_in_service: np.ndarray  # uint8
    gen_bus: np.ndarray       # int32 [ng]
    gen_pg: np.ndarray        # MW
    gen_qg: np.ndarray        # MVAr
    gen_vg: np.ndarray        # pu
    br_rate_a: np.ndarray = field(default=None)   # MVA, 0 = unrated (Branch.RateAMva)
    bus_vmin: np.ndarray = field(default=None)    # pu, NaN = no limit of its own (Bus.VminPu)
    bus_vmax: np.ndarray = field(default=None)    # pu, NaN = no limit of its own (Bus.VmaxPu)
    gen_qmin: np.ndarray = field(default=None)
    gen_qmax: np.ndarray = field(default=None)
    name: str = "network"
    seed: int | None = None
    bus_ids: np.ndarray | None = None

    def __post_init__(self):
        i32 = ("bus_type", "br_from", "br_to", "gen_bus")
        for k in i32:
            setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=np.int32))
        f64 = ("pd", "qd", "gs", "bs", "vm", "va_deg", "br_r", "br_x", "br_b", "br_tap", "br_shift_deg",
               "gen_pg", "gen_qg", "gen_vg")
        for k in f64:
            setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=np.float64))
        self.br_in_service = np.ascontiguousarray(self.br_in_service, dtype=np.uint8)
        if self.br_rate_a is None:
            self.br_rate_a = np.zeros(len(self.br_from))
        self.br_rate_a = np.ascontiguousarray(self.br_rate_a, dtype=np.float64)
        nbus = len(self.bus_type)
        self.bus_vmin = np.ascontiguousarray(np.full(nbus, np.nan) if self.bus_vmin is None else self.bus_vmin,
                                             dtype=np.float64)
        self.bus_vmax = np.ascontiguousarray(np.full(nbus, np.nan) if self.bus_vmax is None else self.bus_vmax,
                                             dtype=np.float64)
        ng = len(self.gen_bus)
        if self.gen_qmin is None:
            self.gen_qmin = np.full(ng, -np.inf)
        if self.gen_qmax is None:
            self.gen_qmax = np.full(ng, np.inf)
        self.gen_qmin = np.ascontiguousarray(self.gen_qmin, dtype=np.float64)
        self.gen_qmax = np.ascontiguousarray(self.gen_qmax, dtype=np.float64)

    @property
    def n_bus(self) -> int:
        return int(self.bus_type.shape[0])

    @property
    def n_branch(self) -> int:
        return int(self.br_from.shape[0])

    @property
    def n_gen(self) -> int:
        return int(self.gen_bus.shape[0])

    def copy(self, **changes) -> "Network":
        fields = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()}
        fields.update(changes)
        return Network(**fields)

    def without_branches(self, idx) -> "Network":
        """The network minus the given branches, remaining branches in model order.

        Mirrors ``ContingencyAnalysis.WithoutMany`` (ContingencyAnalysis.cs:1839-1845): an outage removes the
        branch record, so the Y-bus is rebuilt from the remaining branches in their original order.
        """
        keep = np.ones(self.n_branch, dtype=bool)
        keep[np.asarray(idx, dtype=np.int64)] = False
        return self.copy(br_from=self.br_from[keep], br_to=self.br_to[keep], br_r=self.br_r[keep],
                         br_x=self.br_x[keep], br_b=self.br_b[keep], br_tap=self.br_tap[keep],
                         br_shift_deg=self.br_shift_deg[keep], br_in_service=self.br_in_service[keep],
                         br_rate_a=self.br_rate_a[keep])

    def with_load_multiplier(self, mult: float) -> "Network":
        """Scale every load and every generator's P by ``mult`` (a load-variation instance)."""
        return self.copy(pd=self.pd * mult, qd=self.qd * mult, gen_pg=self.gen_pg * mult)

    def with_start(self, vm: np.ndarray, va_deg: np.ndarray) -> "Network":
        """Stored voltages set to a solved state: the warm start ``BuildBusSpec`` reads when FlatStart is off
        (NewtonShell.cs:464-469)."""
        return self.copy(vm=np.asarray(vm, dtype=np.float64), va_deg=np.asarray(va_deg, dtype=np.float64))


def connected_parts(net: Network, removed=()) -> int:
    """Connected parts of the network with the `removed` branches taken out: buses joined by in-service branches
    (union-find; a bus with no branch is a part of its own), as ContingencyAnalysis.ConnectedParts counts them
    (ContingencyAnalysis.cs:1672-1687)."""
    parent = np.arange(net.n_bus)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    gone = set(int(b) for b in np.atleast_1d(np.asarray(removed, np.int64)) if b >= 0)
    parts = net.n_bus
    for k in range(net.n_branch):
        if not net.br_in_service[k] or k in gone:
            continue
        a, c = find(int(net.br_from[k])), find(int(net.br_to[k]))
        if a != c:
            parent[a] = c
            parts -= 1
    return parts


def splits(net: Network, removed) -> bool:
    """Whether taking the branches out splits the network (ContingencyAnalysis.Splits, :1665)."""
    return connected_parts(net, removed) > connected_parts(net)


def bridges(net: Network) -> np.ndarray:
    """Boolean mask of in-service branches whose removal splits the network (graph bridges).

    Iterative Tarjan low-link over the multigraph of in-service branches; a branch with a parallel twin is never a
    bridge.
    """
    n = net.n_bus
    live = np.nonzero(net.br_in_service)[0]
    f = net.br_from[live].astype(np.int64)
    t = net.br_to[live].astype(np.int64)
    deg = np.bincount(np.concatenate([f, t]), minlength=n)
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(deg, out=ptr[1:])
    adj_node = np.empty(ptr[-1], dtype=np.int64)
    adj_edge = np.empty(ptr[-1], dtype=np.int64)
    fill = ptr[:-1].copy()
    for e in range(len(live)):
        a, b = f[e], t[e]
        adj_node[fill[a]] = b; adj_edge[fill[a]] = e; fill[a] += 1
        adj_node[fill[b]] = a; adj_edge[fill[b]] = e; fill[b] += 1
    disc = np.full(n, -1, dtype=np.int64)
    low = np.zeros(n, dtype=np.int64)
    is_bridge = np.zeros(len(live), dtype=bool)
    timer = 0
    for root in range(n):
        if disc[root] >= 0:
            continue
        stack = [(root, -1, ptr[root])]
        disc[root] = low[root] = timer; timer += 1
        while stack:
            v, pe, it = stack[-1]
            if it < ptr[v + 1]:
                stack[-1] = (v, pe, it + 1)
                u, e = adj_node[it], adj_edge[it]
                if e == pe:
                    continue
                if disc[u] < 0:
                    disc[u] = low[u] = timer; timer += 1
                    stack.append((u, e, ptr[u]))
                else:
                    low[v] = min(low[v], disc[u])
            else:
                stack.pop()
                if stack:
                    p = stack[-1][0]
                    low[p] = min(low[p], low[v])
                    if low[v] > disc[p]:
                        is_bridge[pe] = True
    mask = np.zeros(net.n_branch, dtype=bool)
    mask[live[is_bridge]] = True
    return mask
