"""Seeded synthetic meshed grids for the N-1 / batching / serving workloads.

The reference has no such generator (its synthetic network, src/GridSim.Core/Synthetic/SyntheticGb.cs, is GB-shaped
and unrelated). This one builds the unit-square grids the workloads are specified on:

* buses uniform in [0, 1]^2 from the seed;
* a random spanning tree over the 3-nearest-neighbour graph (Kruskal on random edge weights; components the 3-NN
  graph leaves apart are joined by their shortest connecting edge), then the shortest remaining nearest-neighbour
  edges up to the branch count;
* r ~ U[a, b] pu, x = r * U[3, 10], line charging b ~ U[0, 0.05] pu; a fixed fraction of branches are transformers
  with tap ~ U[0.95, 1.05] (shift 0), placed on radial (bridge) branches first, the way step-up transformers connect
  units, so that no off-nominal tap sits inside a loop;
* every branch's series impedance (r and x together, so x/r is kept) is multiplied by one per-grid factor
  kappa = min(1, theta_max / theta_dc), where theta_dc is the largest bus angle of the DC power flow of the drawn
  grid and theta_max = 1 rad;
* one slack (the bus nearest the centre), ``n_pv`` PV buses drawn at random, the rest PQ;
* PQ loads P ~ U[0, 1] pu at power factor 0.95 (or lognormal(mu, sigma) pu at pf ~ U[pf_lo, pf_hi]);
* PV generation totals 1.03 x the load, split evenly; voltage set points ~ U[1.00, 1.05]; Q limits +-inf.

Everything is drawn from one ``numpy.random.default_rng(seed)`` stream in a fixed order, so a (sizes, seed) pair
names one network on every machine.

Why the two adjustments: drawn literally, these grids have no AC solution. A random spanning tree plus short
nearest-neighbour chords is electrically long (about 200 hops across 2,000 buses, a mean effective reactance of
12 pu from the slack with impedances drawn independently of length), so the DC angles span about 194 rad at the
drawn loading; and a 5 % off-nominal tap on a 0.003 pu branch inside a loop drives about 17 pu of circulating
reactive power. Scaling impedance (not load) keeps every MW and MVAr figure as drawn.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.spatial import cKDTree

from .network import Network, PQ, PV, SLACK, bridges

BASE_SEED = 20261203


class _Dsu:
    def __init__(self, n):
        self.p = np.arange(n)

    def find(self, a):
        p = self.p
        root = a
        while p[root] != root:
            root = p[root]
        while p[a] != root:
            p[a], a = root, p[a]
        return root

    def union(self, a, b):
        ra, rb = self.find(a), self.find(b)
        if ra == rb:
            return False
        if ra < rb:
            self.p[rb] = ra
        else:
            self.p[ra] = rb
        return True


def _knn_edges(tree: cKDTree, pts: np.ndarray, k: int):
    n = len(pts)
    kk = min(k + 1, n)
    d, idx = tree.query(pts, k=kk)
    i = np.repeat(np.arange(n), kk - 1)
    j = idx[:, 1:].reshape(-1)
    a = np.minimum(i, j)
    b = np.maximum(i, j)
    key = np.unique(a.astype(np.int64) * n + b)
    a = (key // n).astype(np.int64)
    b = (key % n).astype(np.int64)
    length = np.hypot(pts[a, 0] - pts[b, 0], pts[a, 1] - pts[b, 1])
    return a, b, length


def _topology(n: int, n_branch: int, rng: np.random.Generator):
    if n_branch < n - 1:
        raise ValueError(f"{n_branch} branches cannot connect {n} buses")
    pts = rng.random((n, 2))
    tree = cKDTree(pts)
    a, b, length = _knn_edges(tree, pts, 3)
    w = rng.random(len(a))
    dsu = _Dsu(n)
    chosen = []
    in_tree = np.zeros(len(a), dtype=bool)
    for e in np.argsort(w, kind="stable"):
        if dsu.union(int(a[e]), int(b[e])):
            chosen.append((int(a[e]), int(b[e])))
            in_tree[e] = True
    # join components the 3-NN graph left apart: shortest edge out of the component holding the lowest bus index
    while len(chosen) < n - 1:
        roots = np.array([dsu.find(i) for i in range(n)])
        r0 = roots[0]
        inside = np.nonzero(roots == r0)[0]
        outside_mask = roots != r0
        out_idx = np.nonzero(outside_mask)[0]
        sub = cKDTree(pts[out_idx])
        d, j = sub.query(pts[inside], k=1)
        best = int(np.argmin(d))
        u, v = int(inside[best]), int(out_idx[j[best]])
        dsu.union(u, v)
        chosen.append((min(u, v), max(u, v)))
    have = set(chosen)
    k = 3
    extra_needed = n_branch - len(chosen)
    extras = []
    while extra_needed > 0:
        a, b, length = _knn_edges(tree, pts, k)
        order = np.lexsort((b, a, length))
        for e in order:
            key = (int(a[e]), int(b[e]))
            if key in have:
                continue
            have.add(key)
            extras.append(key)
            extra_needed -= 1
            if extra_needed == 0:
                break
        k += 2
        if k > n + 2:
            raise ValueError(f"cannot place {n_branch} distinct branches on {n} buses")
    edges = np.array(chosen + extras, dtype=np.int64)
    return pts, edges


def _bridge_mask(n: int, edges: np.ndarray) -> np.ndarray:
    from .network import bridges as _bridges
    nb = len(edges)
    probe = Network(base_mva=100.0, bus_type=np.zeros(n), pd=np.zeros(n), qd=np.zeros(n), gs=np.zeros(n),
                    bs=np.zeros(n), vm=np.ones(n), va_deg=np.zeros(n), br_from=edges[:, 0], br_to=edges[:, 1],
                    br_r=np.zeros(nb), br_x=np.ones(nb), br_b=np.zeros(nb), br_tap=np.zeros(nb),
                    br_shift_deg=np.zeros(nb), br_in_service=np.ones(nb), gen_bus=np.zeros(0), gen_pg=np.zeros(0),
                    gen_qg=np.zeros(0), gen_vg=np.zeros(0))
    return _bridges(probe)


def _dc_angle_range(n: int, edges: np.ndarray, x: np.ndarray, pinj: np.ndarray, slack: int) -> float:
    """Largest |bus angle| (rad) of the DC power flow B' theta = P with the slack angle at 0."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    f, t = edges[:, 0], edges[:, 1]
    bx = 1.0 / x
    bmat = sp.coo_matrix((np.concatenate([bx, bx, -bx, -bx]),
                          (np.concatenate([f, t, f, t]), np.concatenate([f, t, t, f]))), shape=(n, n)).tocsc()
    keep = np.arange(n) != slack
    theta = spl.spsolve(bmat[keep][:, keep].tocsc(), pinj[keep])
    return float(np.abs(theta).max()) if len(theta) else 0.0


def synthetic_grid(n_bus: int, n_branch: int, n_pv: int, seed: int, *, xfmr_frac: float = 0.10,
                   r_range=(0.001, 0.05), load_model: str = "uniform", pf: float = 0.95,
                   pf_range=(0.9, 0.98), lognormal=(-1.0, 1.0), base_mva: float = 100.0,
                   theta_max: float = 1.0, name: str | None = None, return_points: bool = False):
    """One seeded synthetic grid (see the module docstring for the recipe); with return_points also the bus
    coordinates in the unit square."""
    if n_pv > n_bus - 1:
        raise ValueError("too many PV buses")
    rng = np.random.default_rng(seed)
    pts, edges = _topology(n_bus, n_branch, rng)
    nb = len(edges)
    r = rng.uniform(r_range[0], r_range[1], nb)
    x = r * rng.uniform(3.0, 10.0, nb)
    bch = rng.uniform(0.0, 0.05, nb)
    n_x = int(round(xfmr_frac * nb))
    tap = np.zeros(nb)
    radial = _bridge_mask(n_bus, edges)
    cand = np.nonzero(radial)[0]
    if len(cand) >= n_x:
        xf = rng.choice(cand, size=n_x, replace=False)
    else:
        rest = np.nonzero(~radial)[0]
        xf = np.concatenate([cand, rng.choice(rest, size=n_x - len(cand), replace=False)])
    tap[xf] = rng.uniform(0.95, 1.05, n_x)

    slack = int(np.argmin(np.hypot(pts[:, 0] - 0.5, pts[:, 1] - 0.5)))
    others = np.array([i for i in range(n_bus) if i != slack], dtype=np.int64)
    pv = np.sort(rng.choice(others, size=n_pv, replace=False))
    bus_type = np.full(n_bus, PQ, dtype=np.int32)
    bus_type[slack] = SLACK
    bus_type[pv] = PV
    pq = np.nonzero(bus_type == PQ)[0]

    pd = np.zeros(n_bus)
    qd = np.zeros(n_bus)
    if load_model == "uniform":
        p = rng.uniform(0.0, 1.0, len(pq))
        pfs = np.full(len(pq), pf)
    elif load_model == "lognormal":
        p = rng.lognormal(lognormal[0], lognormal[1], len(pq))
        pfs = rng.uniform(pf_range[0], pf_range[1], len(pq))
    else:
        raise ValueError(load_model)
    pd[pq] = p * base_mva
    qd[pq] = pd[pq] * np.tan(np.arccos(pfs))

    total = float(pd.sum())
    gen_bus = np.concatenate([[slack], pv]).astype(np.int32)
    gen_pg = np.concatenate([[0.0], np.full(n_pv, 1.03 * total / max(n_pv, 1))])
    gen_vg = rng.uniform(1.00, 1.05, len(gen_bus))
    gen_qg = np.zeros(len(gen_bus))
    pinj = -pd / base_mva
    np.add.at(pinj, gen_bus, gen_pg / base_mva)
    theta = _dc_angle_range(n_bus, edges, x, pinj, slack)
    kappa = min(1.0, theta_max / theta) if theta > 0 else 1.0
    r = r * kappa
    x = x * kappa
    net = Network(
        base_mva=base_mva, bus_type=bus_type, pd=pd, qd=qd, gs=np.zeros(n_bus), bs=np.zeros(n_bus),
        vm=np.ones(n_bus), va_deg=np.zeros(n_bus),
        br_from=edges[:, 0], br_to=edges[:, 1], br_r=r, br_x=x, br_b=bch, br_tap=tap,
        br_shift_deg=np.zeros(nb), br_in_service=np.ones(nb, dtype=np.uint8),
        gen_bus=gen_bus, gen_pg=gen_pg, gen_qg=gen_qg, gen_vg=gen_vg,
        name=name or f"synthetic-{n_bus}b-{nb}br-s{seed}", seed=seed)
    return (net, pts) if return_points else net


# ---- the five workloads -------------------------------------------------------------------------------------------

def hub_grid(n_spokes: int = 100, seed: int = BASE_SEED + 11):
    """The 118-bus synthetic grid with one bus tied to n_spokes others by long lines (x 2 pu): its Y row then holds
    n_spokes + (its own lines) + 1 cells, far longer than any other row, for testing the paths sized by the longest row
    (the mismatch pass's blocks and scratch)."""
    from dataclasses import replace
    net = config1_grid()
    rng = np.random.default_rng(seed)
    hub = int(np.argmax(np.bincount(np.concatenate([net.br_from, net.br_to]), minlength=len(net.bus_type))))
    others = np.array([i for i in range(len(net.bus_type)) if i != hub])
    to = np.sort(rng.choice(others, size=n_spokes, replace=False))
    k = len(to)
    cat = lambda a, b: np.concatenate([a, b])
    return replace(net, br_from=cat(net.br_from, np.full(k, hub)), br_to=cat(net.br_to, to),
                   br_r=cat(net.br_r, np.full(k, 0.2)), br_x=cat(net.br_x, np.full(k, 2.0)),
                   br_b=cat(net.br_b, np.zeros(k)), br_tap=cat(net.br_tap, np.zeros(k)),
                   br_shift_deg=cat(net.br_shift_deg, np.zeros(k)),
                   br_in_service=cat(net.br_in_service, np.ones(k, dtype=np.uint8)),
                   br_rate_a=cat(net.br_rate_a, np.zeros(k)), name=f"hub{n_spokes}")


def config1_grid(seed: int = BASE_SEED + 1) -> Network:
    """118 buses (1 slack / 53 PV / 64 PQ), 186 branches, flat start."""
    return synthetic_grid(118, 186, 53, seed, name="synthetic-118")


def config2_grid(seed: int = BASE_SEED + 2) -> Network:
    """2,000 buses (1 slack / 300 PV / 1,699 PQ), 2,700 branches: the N-1 screening grid."""
    return synthetic_grid(2000, 2700, 300, seed, name="synthetic-2000")


def config2_outages(net: Network) -> np.ndarray:
    """Every non-bridge in-service branch, in branch order."""
    return np.nonzero((net.br_in_service != 0) & ~bridges(net))[0].astype(np.int32)


def sized_grid(n_bus: int, seed: int, pv_frac: float = 0.15, branch_ratio: float = 1.35) -> Network:
    """Uneven-size workloads: branches 1.35 n, PV fraction 15 %."""
    return synthetic_grid(n_bus, int(round(branch_ratio * n_bus)), max(1, int(round(pv_frac * n_bus))), seed,
                          name=f"synthetic-{n_bus}")


def config5_sizes(seed: int = BASE_SEED + 5, count: int = 8) -> list[int]:
    """Bus counts of the serving workload's resident networks: log-uniform in [100, 25,000], sorted."""
    rng = np.random.default_rng(seed)
    return [int(n) for n in np.sort(np.round(np.exp(rng.uniform(np.log(100), np.log(25000), count))).astype(int))]


def config5_grids(seed: int = BASE_SEED + 5, count: int = 8) -> list[Network]:
    """The serving workload's resident networks: one sized grid per config5_sizes entry, grid i drawn with seed
    seed + 1 + i."""
    return [sized_grid(n, seed + 1 + i) for i, n in enumerate(config5_sizes(seed, count))]


def config4_grid(seed: int = BASE_SEED + 4, cols: int = 7, rows: int = 5) -> Network:
    """The large single instance: 70,000 buses (8,000 PV) and 88,000 branches, 12 % transformers, lognormal loads
    (mu = -1, sigma = 1) pu at pf ~ U[0.9, 0.98], built as cols x rows regions of 2,000 buses joined by tie lines.

    Each region is one synthetic grid (2,509 branches, its share of the 8,000 PV buses, its own slack bus and the
    impedance scaling of the recipe) placed in its own cell of a cols x rows layout; every pair of side-by-side regions
    is joined by tie lines between the bus pairs nearest across their common border (88,000 branches in all), with
    r ~ U[0.01, 0.05] pu, x = r * U[3, 10] and charging b ~ U[0, 0.05] pu, unscaled. Drawn as one 70,000-bus recipe
    grid instead, the impedance scaling that keeps its DC angles within 1 rad shrinks every impedance by about 1e-4
    (median x about 1e-5 pu), and the set-point spread and transformer taps across such branches leave Newton from
    flat start without a reachable solution; a region keeps the 2,000-bus grids' scale. Each region balances its own
    generation at its slack (35 slack buses at angle 0)."""
    regions = cols * rows
    n_region, br_region = 2000, 2509
    pv_split = np.full(regions, 8000 // regions)
    pv_split[: 8000 - pv_split.sum()] += 1
    rng = np.random.default_rng(seed)
    parts = []
    for i in range(regions):
        net, pts = synthetic_grid(n_region, br_region, int(pv_split[i]), seed + 1 + i, xfmr_frac=0.12,
                                  load_model="lognormal", return_points=True)
        parts.append((net, pts + np.array([i % cols, i // cols], dtype=np.float64)))
    # tie lines: side-by-side regions, nearest bus pairs across the border, spread evenly over the pairs
    pairs = [(i, i + 1) for i in range(regions) if i % cols != cols - 1] + \
            [(i, i + cols) for i in range(regions - cols)]
    n_ties = 88000 - regions * br_region
    per_pair = np.full(len(pairs), n_ties // len(pairs))
    per_pair[: n_ties - per_pair.sum()] += 1
    ties = []
    for (a, b), k in zip(pairs, per_pair):
        pa, pb = parts[a][1], parts[b][1]
        tree = cKDTree(pb)
        d, j = tree.query(pa, k=1)
        used_b = set()
        for ia in np.argsort(d, kind="stable"):
            if int(j[ia]) in used_b:
                continue
            used_b.add(int(j[ia]))
            ties.append((a * n_region + int(ia), b * n_region + int(j[ia])))
            if len(used_b) == k:
                break
    nt = len(ties)
    tr = rng.uniform(0.01, 0.05, nt)
    tx = tr * rng.uniform(3.0, 10.0, nt)
    tb = rng.uniform(0.0, 0.05, nt)
    cat = lambda key: np.concatenate([getattr(p[0], key) for p in parts])
    off = lambda key: np.concatenate([getattr(p[0], key) + n_region * i for i, p in enumerate(parts)])
    tie = np.array(ties, dtype=np.int64)
    return Network(
        base_mva=parts[0][0].base_mva, bus_type=cat("bus_type"), pd=cat("pd"), qd=cat("qd"), gs=cat("gs"),
        bs=cat("bs"), vm=cat("vm"), va_deg=cat("va_deg"),
        br_from=np.concatenate([off("br_from"), tie[:, 0]]), br_to=np.concatenate([off("br_to"), tie[:, 1]]),
        br_r=np.concatenate([cat("br_r"), tr]), br_x=np.concatenate([cat("br_x"), tx]),
        br_b=np.concatenate([cat("br_b"), tb]), br_tap=np.concatenate([cat("br_tap"), np.zeros(nt)]),
        br_shift_deg=np.zeros(regions * br_region + nt),
        br_in_service=np.ones(regions * br_region + nt, dtype=np.uint8),
        gen_bus=off("gen_bus"), gen_pg=cat("gen_pg"), gen_qg=cat("gen_qg"), gen_vg=cat("gen_vg"),
        name="synthetic-70000", seed=seed)