"""The effectiveness estimate and the label statistics it reads, on the CPU.

The oracle (oracle/effectiveness.py) is pinned by known answers worked out
by hand for a cycle, a path and a TDS walk, by closed forms of its
probability helpers, by a loop restatement of the statistics, and by the
reference's own per-thread tally, which equals the per-vertex statistic
when every thread gets one vertex.  The host library's estimate and choice
(pjh_constraint_effectiveness, pjh_next_constraint) must then agree with the
oracle bit for bit on the constraints of generated templates over generated
statistics (zero counts included, where the arithmetic gives inf or NaN)."""
import math

import numpy as np
import pytest

from oracle import effectiveness as E
from pattenmatching_sharedmemory_amd.native import STEP_CC, STEP_ENUMERATION, STEP_PC, STEP_TDS

# Triangle statistics: labels 5, 6, 7, four active vertices each, every vertex
# with two neighbours of each other label (8 slots per label pair), so every
# vertex has at least one neighbour of each other label.
TRIANGLE_STAT = {
    "vertex": {5: 4, 6: 4, 7: 4}, "vertex_total": 12, "edge_total": 48,
    "edge": {a: {b: 8 for b in (5, 6, 7) if b != a} for a in (5, 6, 7)},
    "outbound": {5: 16, 6: 16, 7: 16},
    "at_least": {a: {b: {1: 4} for b in (5, 6, 7) if b != a} for a in (5, 6, 7)},
}


def _codes(c):
    """Distinct template labels in template vertex order (the device's codes)."""
    d = c.desc.contents
    out = []
    for p in range(d.vertex_count):
        if d.vertex_label[p] not in out:
            out.append(int(d.vertex_label[p]))
    return out


def _acc_max(c):
    """accumulation_max of the template behind constraints `c`."""
    p = c.pattern
    return E.accumulation_max(p.offsets(), [int(x) for x in p.edges()], [int(x) for x in p.labels()])


def to_product(pj, st, codes, acc):
    """The oracle's statistics as a pj_label_stat over `codes`."""
    K = len(codes)
    mx = [[acc.get(codes[f], {}).get(codes[t], 0) for t in range(K)] for f in range(K)]
    al = {(f, t): [st["at_least"].get(codes[f], {}).get(codes[t], {}).get(k, 0) for k in range(1, mx[f][t] + 1)]
          for f in range(K) for t in range(K) if mx[f][t]}
    return pj.LabelStat.from_arrays(codes, [st["vertex"].get(l, 0) for l in codes],
                                    [[st["edge"].get(codes[f], {}).get(codes[t], 0) for t in range(K)]
                                     for f in range(K)], mx, al)


def oracle_constraints(c):
    """Cycles and paths as label lists, TDS constraints as walks of
    (move, label, move_back), in list order."""
    lab = c.desc.contents.vertex_label
    cycles = [[int(lab[v]) for v in cyc] for cyc in c.cycles()]
    paths = [[int(lab[v]) for v in p] for p in c.paths()]
    tds = [[[(m, int(lab[v]), b) for m, v, b in moves] for _, moves in walks] for walks in c.tds()]
    return cycles, paths, tds


def same(a, b):
    return (math.isnan(a) and math.isnan(b)) or a == b


# --------------------------------------------------------------------------
# the oracle, pinned

@pytest.mark.filterwarnings("ignore::RuntimeWarning")
def test_probability_helpers_closed_forms():
    s = E._Stat(TRIANGLE_STAT)
    # 1 - ((n - 1) / n)^k with n = 4: k = 2 -> 7/16, k = 3 -> 37/64
    assert E.proba_find_same_vertex(s, 2, 5) == 7 / 16
    assert E.proba_find_same_vertex(s, 3, 6) == 37 / 64
    # 1 - (1 / n)^k: k = 2 -> 15/16
    assert E.proba_find_different_vertex(s, 2, 7) == 15 / 16
    # at least one neighbour of the label for every vertex: probability 1
    assert E.proba_find_vertex(s, 5, 6) == 1.0
    assert E.proba_find_one_path(s, 3, 5, 6) == 1.0
    # 8 slots from 4 vertices: 2 per vertex; capped at the 4 vertices of the label
    assert E.average_vertex_knowing(s, 3, 5, 6) == 6.0
    assert E.average_unique_vertex(s, 1, 5, 6) == 2.0
    assert E.average_unique_vertex(s, 3, 5, 6) == 4.0
    assert E.average_vertex(s, 1, 5) == 4.0
    # half the vertices with such a neighbour: 1 - (1/2)^k
    half = dict(TRIANGLE_STAT, at_least={5: {6: {1: 2}}})
    assert E.proba_find_one_path(E._Stat(half), 3, 5, 6) == 1 - 0.125
    # no vertex of the label: 0/0 (the reference's map lookup would throw)
    assert math.isnan(E.proba_find_vertex(s, 9, 6))


def test_cycle_known_answer(pj):
    """Triangle 5-6-7 over TRIANGLE_STAT, by hand.  From every start: each hop
    is found with probability 1 (every vertex has such a neighbour); the
    unique vertices reached go 1 -> min(2, 4) = 2 -> 4 -> min(8, 4) = 4; the
    closure finds the start again with 1 - (3/4)^4 = 175/256.  Matches
    eliminated: 3 x 4 x 81/256 = 3.796875.  Cost per start: 1 + 4 (1 vertex x
    16/4 slots) + 8 (2 vertices) + 16 (closure from 4) + 4 = 33, times
    4 x 81/256, over 1 + (3 - 1)/2 x 175/256: 41.765625 / 1.68359375; three
    starts."""
    m, cst, e = E.effectiveness_cycle(TRIANGLE_STAT, [5, 6, 7])
    assert m == 3.796875
    assert cst == 3 * (41.765625 / 1.68359375)
    assert e == m / cst
    c = pj.Constraints(pj.Graph.pattern(3, [(0, 1), (1, 2), (2, 0)], [5, 6, 7]))
    st = to_product(pj, TRIANGLE_STAT, _codes(c), _acc_max(c))
    assert c.effectiveness(st, STEP_CC, 0) == (m, cst, e)
    assert c.next_constraint(st) == (STEP_CC, 0, e)
    assert c.next_constraint(st, done_cycles=[0])[0] == STEP_ENUMERATION


# Cycle statistics with every count different: V5 = 2, V6 = 4, V7 = 8; slots
# 5->6: 12, 5->7: 8, 6->5: 4, 6->7: 8, 7->5: 8, 7->6: 8 (outbound 20, 12, 16);
# vertices with at least one neighbour: 5->6 all 2, 5->7 all 2, 6->5 all 4,
# 6->7 half (2 of 4), 7->5 half (4 of 8), 7->6 all 8.
SKEWED_STAT = {
    "vertex": {5: 2, 6: 4, 7: 8}, "vertex_total": 14, "edge_total": 48,
    "edge": {5: {6: 12, 7: 8}, 6: {5: 4, 7: 8}, 7: {5: 8, 6: 8}},
    "outbound": {5: 20, 6: 12, 7: 16},
    "at_least": {5: {6: {1: 2}, 7: {1: 2}}, 6: {5: {1: 4}, 7: {1: 2}}, 7: {5: {1: 4}, 6: {1: 8}}},
}


def test_cycle_known_answer_skewed(pj):
    """Triangle 5-6-7 over SKEWED_STAT, by hand.  Finding, from 5: hop 5->6
    certain, min(12/2, 4) = 4 reached; 6->7 with 1 - 0.5^4, min(4 x 8/4, 8)
    = 8; 7->5 with 1 - 0.5^8, min(8, 2) = 2; the start again with
    1 - (1/2)^2: F0 = 15/16 x 255/256 x 3/4.  From 6: 0.5, 2; 0.75, 2; 1, 4;
    1 - (3/4)^4: F1 = 0.375 x 175/256.  From 7: 0.5, 1; 1, 4; 0.9375, 8;
    1 - (7/8)^8: F2 = 0.46875 x (1 - (7/8)^8).  Costs: from 5, 1 + 20/2 +
    6 x 12/4 (6 = 12/2 neighbours, the uncapped first hop) + 8 x 16/8 + 8 =
    53; from 6, 1 + 3 + 4 + 2 x 20/2 + 2 = 30; from 7, 1 + 2 + 10 + 4 x 12/4
    + 4 = 29; each times its label's vertices x (1 - F), over 1 + F."""
    F0 = 15 / 16 * 255 / 256 * 3 / 4
    F1 = 0.375 * 175 / 256
    F2 = 0.46875 * (1 - 5764801 / 16777216)
    m, cst, e = E.effectiveness_cycle(SKEWED_STAT, [5, 6, 7])
    assert m == 2 * (1 - F0) + 4 * (1 - F1) + 8 * (1 - F2)
    want = 0.0
    for n, cost, f in ((2, 53, F0), (4, 30, F1), (8, 29, F2)):
        want += n * cost * (1 - f) / 1 / (1 + 2 / 2 * f)
    assert cst == want
    c = pj.Constraints(pj.Graph.pattern(3, [(0, 1), (1, 2), (2, 0)], [5, 6, 7]))
    (cyc,) = oracle_constraints(c)[0]
    st = to_product(pj, SKEWED_STAT, _codes(c), _acc_max(c))
    assert c.effectiveness(st, STEP_CC, 0) == E.effectiveness_cycle(SKEWED_STAT, cyc)


def test_path_known_answer_four_vertices():
    """Path 5-6-7-5 (positions 0..3): V5 = 4, V6 = 2, V7 = 8; slots 5->6: 4,
    5->7: 16, 6->5: 4, 6->7: 8, 7->5: 16, 7->6: 8 (outbound 20, 12, 24); half
    the label-5 vertices have a label-6 neighbour, half the label-7 ones a
    label-5 neighbour, every other pair all.  Forward: 5->6 with 0.5,
    min(1, 2) = 1; 6->7 certain, min(4, 8) = 4; the end differs from the
    start (label 5) with 1 - (1/4)^4: F0 = 0.5 x 255/256.  Reverse, from
    position 3: 5->7 certain, min(4, 8) = 4; 7->6 certain, min(4, 2) = 2;
    end label read at position 1 (6): 1 - (1/2)^2: F1 = 0.75.  Matches:
    V5 (1 - F0) + V6 (1 - F1).  Costs: forward 1 + 5 + 6 + 4 x 20/4 + 4 = 36;
    reverse 1 + 5 + 4 x 24/8 + 2 x 20/4 + 2 = 30; over 1 + 3/2 F."""
    st = {"vertex": {5: 4, 6: 2, 7: 8}, "vertex_total": 14, "edge_total": 56,
          "edge": {5: {6: 4, 7: 16}, 6: {5: 4, 7: 8}, 7: {5: 16, 6: 8}},
          "outbound": {5: 20, 6: 12, 7: 24},
          "at_least": {5: {6: {1: 2}, 7: {1: 4}}, 6: {5: {1: 2}, 7: {1: 2}}, 7: {5: {1: 4}, 6: {1: 8}}}}
    F0, F1 = 0.5 * 255 / 256, 0.75
    m, cst, e = E.effectiveness_path(st, [5, 6, 7, 5])
    assert m == 4 * (1 - F0) + 2 * (1 - F1)
    assert cst == 4 * 36 * (1 - F0) / 1 / (1 + 3 / 2 * F0) + 2 * 30 * (1 - F1) / 1 / (1 + 3 / 2 * F1)
    assert e == m / cst


def test_path_known_answer(pj):
    """Path 5-6-5 (template 0-1-2, labels 5, 6, 5): four label-5 vertices with
    one label-6 neighbour each, two label-6 vertices with two label-5
    neighbours each.  Forward: the hop 5->6 is certain, min(4/4, 2) = 1
    vertex reached, the end differs from the start with 1 - 1/4 = 0.75.
    Reverse (positions 2 -> 1, end label at position 1 as the reference
    reads it): the same hop, then 1 - 1/2 = 0.5.  Matches: 4 x 0.25 +
    2 x 0.5 = 2.  Costs: (1 + 1 + 1 + 1) x 4 x 0.25 / 1.75 = 16/7 and
    (1 + 1 + 1 + 1) x 2 x 0.5 / 1.5 = 8/3."""
    st = {"vertex": {5: 4, 6: 2}, "vertex_total": 6, "edge_total": 8,
          "edge": {5: {6: 4}, 6: {5: 4}}, "outbound": {5: 4, 6: 4},
          "at_least": {5: {6: {1: 4}}, 6: {5: {1: 2, 2: 2}}}}
    m, cst, e = E.effectiveness_path(st, [5, 6, 5])
    assert m == 2.0
    assert cst == 4 / 1.75 + 4 / 1.5
    assert abs(cst - 104 / 21) < 1e-15
    c = pj.Constraints(pj.Graph.pattern(3, [(0, 1), (1, 2)], [5, 6, 5]))
    assert len(c.paths()) == 1 and not c.cycles()
    prod = to_product(pj, st, _codes(c), _acc_max(c))
    assert c.effectiveness(prod, STEP_PC, 0) == (m, cst, e)


def test_tds_walk_known_answer():
    """Walk [VERTEX 5, VERTEX 6, CHECK 5 (history 0), MOVE_BACK (history 1),
    VERTEX 7]; the reference reads moves 1..length-2, so the last VERTEX is
    not scored.  Labels 5, 6, 7 with four vertices each and 8 slots per
    pair; half of the label-5 vertices have a label-6 neighbour.  Finding:
    the hop 5->6 with 1 - 0.5 = 0.5, reaching min(2, 4) = 2 vertices; the
    CHECK finds a label-5 neighbour surely and the start again with
    1 - (3/4)^2 = 0.4375: 0.21875.  Matches: 4 x 0.78125 = 3.125.  Cost:
    1 + 4 (slots of the root) + 8 (slots of 2 vertices at the CHECK) + 2
    (vertices after the MOVE_BACK) = 15, times 4 x 0.78125, over 1 + (3 -
    1)/2 x 0.21875."""
    st = {"vertex": {5: 4, 6: 4, 7: 4}, "vertex_total": 12, "edge_total": 48,
          "edge": {a: {b: 8 for b in (5, 6, 7) if b != a} for a in (5, 6, 7)},
          "outbound": {5: 16, 6: 16, 7: 16},
          "at_least": {5: {6: {1: 2}, 7: {1: 4}}, 6: {5: {1: 4}, 7: {1: 2}}, 7: {5: {1: 4}, 6: {1: 4}}}}
    walk = [(E.MOVE_VERTEX_NO_STORE, 5, 0), (E.MOVE_VERTEX_NO_STORE, 6, 0), (E.MOVE_CHECK, 5, 0),
            (E.MOVE_BACK, 6, 1), (E.MOVE_VERTEX_NO_STORE, 7, 0)]
    m, cst, e = E.effectiveness_tds(st, [walk])
    assert m == 3.125
    assert cst == 4 * 15 * 0.78125 / (1 + 1 * 0.21875)
    assert e == m / cst
    # the last move is not read: another last move gives the same estimate,
    # while reading it (6->7 is found with 1 - 0.5^2) would not
    assert E.effectiveness_tds(st, [walk[:-1] + [(E.MOVE_VERTEX_NO_STORE, 6, 0)]]) == (m, cst, e)


def test_choice_rule_ties_and_nan():
    """Greatest effectiveness wins, ties go to the earlier constraint and to
    cycles, then paths; NaN never wins; nothing left: enumeration."""
    nan_stat = {"vertex": {}, "edge": {}, "outbound": {}, "at_least": {}, "vertex_total": 0, "edge_total": 0}
    assert E.choose_next(nan_stat, [(0, [1, 2, 3])], [], []) == (STEP_CC, 0, 0.0)
    assert E.choose_next(TRIANGLE_STAT, [], [], []) == (STEP_ENUMERATION, 0, 0.0)
    e = E.effectiveness_cycle(TRIANGLE_STAT, [5, 6, 7])[2]
    assert E.choose_next(TRIANGLE_STAT, [(3, [5, 6, 7]), (4, [6, 7, 5])], [], [])[:2] == (STEP_CC, 3)
    assert E.choose_next(TRIANGLE_STAT, [], [(2, [5, 6, 7])], [(1, [[(3, 5, 0), (3, 6, 0), (3, 7, 0)]])])[0] in (
        STEP_PC, STEP_TDS)
    assert e > 0


def test_accumulation_max_by_hand(pj):
    """Triangle labelled 5, 6, 5: vertex 0 (5) sees 6 and 5, vertex 1 (6)
    sees two 5s, vertex 2 (5) sees 5 and 6."""
    p = pj.Graph.pattern(3, [(0, 1), (1, 2), (2, 0)], [5, 6, 5])
    acc = E.accumulation_max(p.offsets(), [int(x) for x in p.edges()], [int(x) for x in p.labels()])
    assert acc == {5: {6: 1, 5: 1}, 6: {5: 2}}


def _random_state(g, rng, p_active=0.7, p_edge=0.8):
    V, E_ = g.vertex_count, g.edge_count
    cand = np.where(rng.random(V) < p_active, rng.integers(1, 1 << 8, V), 0).astype(np.uint64)
    bits = (rng.random(((E_ + 31) // 32) * 32) < p_edge).astype(np.uint8)
    ea = np.packbits(bits, bitorder="little").view(np.uint32)
    return cand, ea


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_statistics_vectorised_equal_loops_and_thread_tallies(pj, seed):
    """graph_stat == graph_stat_loops on random graphs and states; the
    reference's tally as written equals them with one vertex per thread,
    and with fewer threads its at-least counts are never smaller (a tally
    that is never cleared only grows)."""
    rng = np.random.default_rng(seed)
    g = pj.Graph.rmat(8, 8, 0.57, 0.19, 0.19, seed=seed).labels_uniform(4, seed=seed)
    p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 3), (3, 0), (0, 2)], [0, 1, 0, 2])
    acc = E.accumulation_max(p.offsets(), [int(x) for x in p.edges()], [int(x) for x in p.labels()])
    off, edges, labels = g.offsets(), [int(x) for x in g.edges()], [int(x) for x in g.labels()]
    cand, ea = _random_state(g, rng)
    a = E.graph_stat(off, edges, labels, cand, ea, acc)
    b = E.graph_stat_loops(off, edges, labels, cand, ea, acc)
    assert a == b
    assert a["edge_total"] == sum(a["outbound"].values()) == sum(sum(d.values()) for d in a["edge"].values())
    assert E.graph_stat_thread_tallies(off, edges, labels, cand, ea, acc, g.vertex_count) == a
    one = E.graph_stat_thread_tallies(off, edges, labels, cand, ea, acc, 1)
    assert one["vertex"] == a["vertex"] and one["edge"] == a["edge"]
    grew = False
    for f, d in one["at_least"].items():
        for t, ks in d.items():
            for k, n in ks.items():
                mine = a["at_least"].get(f, {}).get(t, {}).get(k, 0)
                assert n >= mine
                grew = grew or n > mine
    assert grew


# --------------------------------------------------------------------------
# the host library against the oracle

def _random_stat(rng, codes, acc, zero_p=0.1):
    st = {"vertex": {}, "edge": {}, "outbound": {}, "at_least": {}, "vertex_total": 0, "edge_total": 0}
    for l in codes:
        if rng.random() >= zero_p:
            st["vertex"][l] = int(rng.integers(1, 200))
    for a in codes:
        for b in codes:
            if rng.random() < 0.8:
                n = int(rng.integers(0, 800))
                if n:
                    st["edge"].setdefault(a, {})[b] = n
                    st["outbound"][a] = st["outbound"].get(a, 0) + n
            m = acc.get(a, {}).get(b, 0)
            have = st["vertex"].get(a, 0)
            run = int(rng.integers(0, have + 1)) if have else 0
            for k in range(1, m + 1):
                if run:
                    st["at_least"].setdefault(a, {}).setdefault(b, {})[k] = run
                run = int(rng.integers(0, run + 1)) if run else 0
    return st


@pytest.mark.parametrize("seed", [11, 12, 13, 14])
def test_host_estimate_equals_oracle(pj, seed):
    """Every cycle, path and TDS constraint of generated templates (repeated
    labels included) over generated statistics: (matches, cost,
    effectiveness) bit for bit, and the same choice with random constraints
    already run."""
    from pattenmatching_sharedmemory_amd.templates import random_templates

    rng = np.random.default_rng(seed)
    n_checked = {STEP_CC: 0, STEP_PC: 0, STEP_TDS: 0}
    for k, e, l in random_templates(12, seed, 3, 9, 4, labels=6, zipf=0.8):
        c = pj.Constraints(pj.Graph.pattern(k, e, l))
        codes, acc = _codes(c), _acc_max(c)
        cycles, paths, tds = oracle_constraints(c)
        for _ in range(3):
            st = _random_stat(rng, codes, acc)
            prod = to_product(pj, st, codes, acc)
            for step, items, fn in ((STEP_CC, cycles, E.effectiveness_cycle), (STEP_PC, paths, E.effectiveness_path),
                                    (STEP_TDS, tds, E.effectiveness_tds)):
                for i, x in enumerate(items):
                    want = fn(st, x, 3)
                    got = c.effectiveness(prod, step, i, threads=3)
                    assert all(same(a, b) for a, b in zip(got, want)), (step, i, got, want)
                    n_checked[step] += 1
            done = [[i for i in range(len(items)) if rng.random() < 0.4] for items in (cycles, paths, tds)]
            want = E.choose_next(st, [(i, x) for i, x in enumerate(cycles) if i not in done[0]],
                                 [(i, x) for i, x in enumerate(paths) if i not in done[1]],
                                 [(i, x) for i, x in enumerate(tds) if i not in done[2]])
            got = c.next_constraint(prod, *done)
            assert got[:2] == want[:2] and same(got[2], want[2]), (got, want)
    assert all(n > 0 for n in n_checked.values()), n_checked


def test_statistics_round_trip(pj):
    """pj_label_stat built from the oracle's statistics reads back the same
    counts (layout of the pair tables and the at-least runs)."""
    p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 3), (3, 0), (0, 2)], [0, 1, 0, 2])
    c = pj.Constraints(p)
    codes, acc = _codes(c), _acc_max(c)
    st = _random_stat(np.random.default_rng(5), codes, acc, zero_p=0.0)
    prod = to_product(pj, st, codes, acc)
    assert prod.labels == codes
    for f, a in enumerate(codes):
        assert prod.vertex[f] == st["vertex"].get(a, 0)
        for t, b in enumerate(codes):
            assert prod.edge[f, t] == st["edge"].get(a, {}).get(b, 0)
            assert prod.at_least_max[f, t] == acc.get(a, {}).get(b, 0)
            if prod.at_least_max[f, t]:
                assert list(prod.at_least[(f, t)]) == [st["at_least"].get(a, {}).get(b, {}).get(k, 0)
                                                       for k in range(1, prod.at_least_max[f, t] + 1)]
