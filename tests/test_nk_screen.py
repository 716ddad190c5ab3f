"""N-k screening host logic (ContingencyAnalysis.ScreenNMinusKReport, RankCandidates, Combinations, StressScore,
ProvablyUnsolvablePart; src/GridSim.Core/Analysis/ContingencyAnalysis.cs:378-486,1073-1080,1750-1829,1851-1866) driven
by the oracle on the CPU, against the reference's recorded `gridsim nminus2 case9` output."""
import os

import numpy as np

from conftest import GOLDEN, matpower_case
from gridsim_amd import engine as eng


def parse_nk_fixture(path):
    """The table rows (label, status), the summary counts and the excluded single outages of a `gridsim nminus2`
    output."""
    rows, excluded, counts, dashes = [], [], None, 0
    lines = open(path, encoding="utf-8").read().splitlines()
    for j, line in enumerate(lines):
        if line.startswith("-----"):
            dashes += 1
            continue
        if dashes == 1 and line.strip():
            rows.append((line[:32].strip(), line[32:46].strip()))
        if "outages screened" in line:
            w = line.split()
            counts = {"screened": int(w[0]), "secure": int(w[3]), "violations": int(w[5]), "islanded": int(w[8])}
        if "left out of the pairing" in line:
            for rest in lines[j + 1:]:
                if not rest.startswith("  "):
                    break
                excluded.append((rest[:34].strip(), rest[34:].strip()))
    return rows, counts, excluded


def oracle_screen(oracle, net, sets, opt):
    out = []
    for st in sets:
        st = list(st)
        if eng.provably_unsolvable(net, st, opt.tolerance):
            out.append(eng.classify(net, st, False, None, None))
            continue
        rec, vm, va = oracle.solve(net, opt, skip=st)
        if rec["status"] != oracle.CONVERGED:
            out.append(eng.classify(net, st, False, None, None))
            continue
        summ = oracle.screen_summary(net, vm, va, skip=st)
        loading = oracle.line_flows(net, vm, va, skip=st)[4]
        out.append(eng.classify(net, st, True, summ, loading))
    return out


def test_combinations_and_ranking_follow_the_reference_order():
    assert eng.combinations([7, 3, 5], 2) == [[7, 3], [7, 5], [3, 5]]
    assert eng.combinations([1, 2, 3, 4], 3) == [[1, 2, 3], [1, 2, 4], [1, 3, 4], [2, 3, 4]]
    assert eng.combinations([1, 2], 3) == []
    mk = lambda conv, load, vmin=1.0, vmax=1.0: eng.ContingencyResult("x", conv, load, vmin, vmax, [])
    single = [mk(True, 50.0), mk(False, 0.0), mk(True, 80.0), mk(True, 50.0, vmin=0.93), mk(True, 80.0)]
    pos, excluded = eng.rank_candidates(single, 2, 2)
    # 0.93 pu scores 50 + 10; equal scores keep branch order
    assert pos == [2, 4] and len(excluded) == 1
    pos, _ = eng.rank_candidates(single, 20, 2)
    assert pos == [2, 4, 3, 0]
    assert eng.stress_score(single[1]) == float("inf")


def test_provably_unsolvable_parts():
    net = matpower_case("case9")
    ids = list(net.bus_ids)
    br = {(int(ids[f]), int(ids[t])): k for k, (f, t) in enumerate(zip(net.br_from, net.br_to))}
    # 9-4 and 8-9 out: bus 9 (a load, no slack) stands alone with 125 MW of demand it cannot meet
    assert eng.provably_unsolvable(net, [br[(9, 4)], br[(8, 9)]], 1e-8)
    # a single ring branch out keeps everything connected to the slack
    assert not eng.provably_unsolvable(net, [br[(9, 4)]], 1e-8)
    # 1-4 out leaves the slack (bus 1) alone: the rest has 163 + 85 MW of generation against 315 MW of load
    assert eng.provably_unsolvable(net, [br[(1, 4)]], 1e-8)


def test_nminus2_case9_matches_the_reference_cli_fixture_on_the_oracle(oracle):
    net = matpower_case("case9")
    opt = oracle.options(max_iterations=100)
    live = [int(b) for b in np.nonzero(net.br_in_service)[0]]
    single = oracle_screen(oracle, net, [[b] for b in live], opt)
    pos, excluded = eng.rank_candidates(single, eng.DEFAULT_CANDIDATE_COUNT, 2)
    combos = eng.combinations([live[i] for i in pos], 2)
    results = sorted(oracle_screen(oracle, net, combos, opt), key=lambda c: -eng.stress_score(c))
    rows, counts, fixture_excluded = parse_nk_fixture(os.path.join(GOLDEN, "cli", "nminus2-case9.txt"))
    assert len(results) == counts["screened"] == 15
    assert sum(c.secure for c in results) == counts["secure"]
    assert sum(c.converged and not c.secure for c in results) == counts["violations"]
    assert sum(c.islanded for c in results) == counts["islanded"]
    for (label, status), c in zip(rows, results):
        assert c.outage == label and status == "ISLANDED" and c.islanded
    assert [(c.outage, c.violations[0]) for c in excluded] == fixture_excluded
