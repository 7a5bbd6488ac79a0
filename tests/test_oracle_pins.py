"""Pins the oracle's template-driven search (TDS), its memos and the phase-2
scheduling rule by what the mathematics and hand traces fix:

* complete-mode TDS equals brute-force existence of a candidate-respecting
  homomorphism of the merged template (tds_cpu.tpp:13-127 without memos);
* a memo-inertness theorem: under its hypothesis the reference's memoised
  search (edge memo tds_cpu.tpp:43-49, hypothesis memo :51-55) decides the
  same as the complete search, walk by walk;
* two known answers traced by hand through tds_cpu.tpp:13-127 and :153-188,
  where the edge memo and the hypothesis memo each block a walk the complete
  search finds;
* phase 2 (lcc_cpu.tpp:107-142, tds_cpu.tpp:200-237): an unmodified active
  vertex ends scheduled iff one of its active slots leads to a modified
  vertex, a modified one always ends scheduled;
* the TDS constraint order and walk order of the reference's fixtures
  test_3 and test_4, derived by hand from
  multiple_label_generate_constraint_template.cc:13-145 and
  multiple_label_constraint_template.cc:121-169."""
import random

import numpy as np
import pytest

from bruteforce import homomorphism_exists, random_symmetric_graph, walk_template_edges
from conftest import fixture_paths
from pattenmatching_sharedmemory_amd.native import (MOVE_BACK, MOVE_CHECK, MOVE_VERTEX_NO_STORE, MOVE_VERTEX_STORE,
                                                     STEP_CC, STEP_LCC, STEP_LCC0, STEP_PC, STEP_TDS)

DIAMOND = [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3)]  # the test_3 template
TWO_CYCLES_CHORD = [(0, 1), (1, 2), (2, 3), (3, 0), (1, 4), (4, 5), (5, 2), (0, 2), (5, 6), (6, 7)]
HOUSE = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (1, 4)]
K4 = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]

TEMPLATES = {
    "diamond_test3_labels": (4, DIAMOND, [0, 1, 2, 1]),
    "diamond_distinct": (4, DIAMOND, [0, 1, 2, 3]),
    "two_cycles_chord": (8, TWO_CYCLES_CHORD, list(range(8))),
    "house": (5, HOUSE, [0, 1, 2, 3, 4]),
    "k4": (4, K4, [0, 1, 2, 3]),
}

VERTEX_MOVES = (MOVE_VERTEX_STORE, MOVE_VERTEX_NO_STORE)


def _graph(pj, n, p, seed):
    off, ed = random_symmetric_graph(n, p, random.Random(seed))
    return pj.Graph.from_csr(off, ed, [0] * n)


def _symmetric_ea(g, keep, rng):
    """Edge-active words with each undirected edge kept with probability
    `keep`, both of its slots alike; also the active adjacency sets."""
    off, ed = g.offsets(), g.edges()
    V = g.vertex_count
    bits = np.zeros(g.edge_count, dtype=np.uint8)
    alive = {}
    for v in range(V):
        for e in range(int(off[v]), int(off[v + 1])):
            u = int(ed[e])
            k = (min(u, v), max(u, v))
            if k not in alive:
                alive[k] = rng.random() < keep
            bits[e] = alive[k]
    words = np.packbits(bits, bitorder="little")
    words = np.concatenate([words, np.zeros((-len(words)) % 4, dtype=np.uint8)]).view(np.uint32)
    adj = [set() for _ in range(V)]
    for (a, b), on in alive.items():
        if on:
            adj[a].add(b)
            adj[b].add(a)
    return words[: (g.edge_count + 31) // 32].copy(), adj


def _load(orc, g, c, cand, ea, threads=1):
    """Oracle holding `cand` (active iff non-empty), `ea`, nothing modified,
    everything scheduled."""
    cand = np.asarray(cand, dtype=np.uint64)
    vf = np.where(cand != 0, 5, 4).astype(np.uint8)
    bits = np.unpackbits(ea.view(np.uint8), bitorder="little")[: g.edge_count]
    o = orc.Oracle(g, c, threads=threads)
    o.load_state(cand, vf, ea, int(np.count_nonzero(cand)), int(bits.sum()))
    return o


def _tds_failures(orc, g, c, cand, ea, index, mode, threads=1):
    o = _load(orc, g, c, cand, ea, threads)
    o.step(STEP_TDS, index, mode)
    after = o.state()[0]
    return np.asarray(cand, dtype=np.uint64) & ~after, o.last_marks()


# ---------------------------------------------------------------------------
# Complete search = homomorphism existence


@pytest.mark.parametrize("name", sorted(TEMPLATES))
def test_tds_complete_search_is_homomorphism_existence(pj, orc, name):
    """Without memos and shortcut, the walk of root x from v succeeds iff
    some map f of the merged template's vertices, f(x) = v, puts every
    template vertex y into C(f(y)) and every template edge on an active edge
    (CHECK moves verify the non-tree edges, VERTEX moves follow the tree).
    Random candidate sets and symmetric edge activity on random graphs of
    three densities."""
    n, edges, labels = TEMPLATES[name]
    c = pj.Constraints(pj.Graph.pattern(n, edges, labels))
    assert c.tds(), name
    tree_only_differs = checked = fails = 0
    for seed, density in ((1, 0.4), (2, 0.55), (3, 0.7), (4, 0.55)):
        rng = random.Random(seed * 101 + n)
        g = _graph(pj, 11, density, seed * 7 + n)
        ea, adj = _symmetric_ea(g, 0.9, rng)
        cand = [sum(1 << y for y in range(n) if rng.random() < 0.7) if rng.random() < 0.9 else 0
                for _ in range(g.vertex_count)]
        for k, walks in enumerate(c.tds()):
            failed, _ = _tds_failures(orc, g, c, cand, ea, k, orc.COMPLETE)
            for root, moves in walks:
                t_edges = walk_template_edges(root, moves)
                tree = set()
                cur, hist = root, [root]
                for mv, x, mb in moves[1:]:
                    if mv in VERTEX_MOVES:
                        tree.add(frozenset((cur, x)))
                        hist.append(x)
                        cur = x
                    elif mv == MOVE_BACK:
                        cur = hist[mb]
                for v in range(g.vertex_count):
                    if not (cand[v] >> root) & 1:
                        continue
                    exists = homomorphism_exists(adj, cand, t_edges, root, v)
                    assert bool((int(failed[v]) >> root) & 1) == (not exists), (name, seed, k, root, v)
                    checked += 1
                    fails += not exists
                    tree_only_differs += (not exists) and homomorphism_exists(adj, cand, tree, root, v)
    # the cases decide something, and some hinge on a CHECK move
    assert 0 < fails < checked
    assert tree_only_differs > 0
    print(f"{name}: {checked} (source, root) pairs, {fails} without a homomorphism, "
          f"{tree_only_differs} of them with one of the spanning tree")


# ---------------------------------------------------------------------------
# Memo-inertness theorem


def walk_is_memo_inert(root, moves):
    """Template side of the inertness theorem: every MOVE_BACK returns to the
    root or to a vertex whose VERTEX move stores it (VERTEX_STORE)."""
    hist_move = [MOVE_VERTEX_STORE]
    for mv, x, mb in moves[1:]:
        if mv in VERTEX_MOVES:
            hist_move.append(mv)
        elif mv == MOVE_BACK and mb != 0 and hist_move[mb] != MOVE_VERTEX_STORE:
            return False
    return True


def _one_template_vertex_each(cand, walks):
    members = 0
    for r, _ in walks:
        members |= 1 << r
    return all(bin(int(x) & members).count("1") <= 1 for x in cand)


def _planted_graph(pj, n_graph, t_edges, n_t, copies, density, rng):
    """Random graph on n_graph vertices (pairs joined with probability
    `density`) with `copies` copies of the template planted on distinct
    vertices; returns the graph and, per planted vertex, its template vertex."""
    pairs = {(a, b) for a in range(n_graph) for b in range(a + 1, n_graph) if rng.random() < density}
    role = {}
    for _ in range(copies):
        free = [v for v in range(n_graph) if v not in role]
        img = rng.sample(free, n_t)
        for x, v in enumerate(img):
            role[v] = x
        for a, b in t_edges:
            pairs.add((min(img[a], img[b]), max(img[a], img[b])))
    adj = [[] for _ in range(n_graph)]
    for a, b in pairs:
        adj[a].append(b)
        adj[b].append(a)
    off, ed = [0], []
    for v in range(n_graph):
        ed += sorted(adj[v])
        off.append(len(ed))
    return pj.Graph.from_csr(off, ed, [0] * n_graph), role


@pytest.mark.parametrize("name", ["diamond_distinct", "two_cycles_chord", "house", "k4"])
def test_tds_memo_inert_walks_decide_like_the_complete_search(pj, orc, name):
    """Theorem: if no graph vertex holds two template vertices of the TDS
    constraint (distinct labels after LCC0 give this) and a walk returns by
    MOVE_BACK only to the root or to stored vertices, the reference's
    memoised search of that walk decides exactly as the complete search.

    Edge memo: a VERTEX move p->q towards y inserts y for {p, q}.  Each
    template edge is followed by exactly one VERTEX move, in one direction,
    and every graph vertex stands for one template vertex, so the entry of
    {cur, u} can only hold t(u), the target of the move being tested.  The
    test uses x = the template vertex of the previous move (tds_cpu.tpp:32):
    t(cur) after a VERTEX or MOVE_BACK move, a CHECK target (already in the
    history) after a CHECK; neither is t(u), which is new.  So it never fires.
    Hypothesis memo: a revisit of (u, y) with an equal list of stored
    vertices skips a subtree that failed before.  What follows (u, y) reads
    the history only through CHECK targets (stored by construction,
    walk.cc:43-60), MOVE_BACK targets (the root or stored, by hypothesis) and
    the memos (the edge memo never fires; the hypothesis memo, by induction
    over the search order), so the skipped subtree would fail again.
    The alreadyMatched shortcut is off here (every walk runs).  A non-inert
    walk can only fail more often than the complete search."""
    n, edges, labels = TEMPLATES[name]
    c = pj.Constraints(pj.Graph.pattern(n, edges, labels))
    inert_walks = inert_ok = inert_fail = blocked = 0
    for seed, density in ((4, 0.25), (5, 0.35), (6, 0.45), (7, 0.35)):
        rng = random.Random(seed * 13 + n)
        g, role = _planted_graph(pj, 2 * n + 8, edges, n, 2, density, rng)
        ea, _ = _symmetric_ea(g, 0.92, rng)
        cand = [(1 << role[v]) if v in role else (1 << rng.randrange(n)) if rng.random() < 0.9 else 0
                for v in range(g.vertex_count)]
        for k, walks in enumerate(c.tds()):
            assert _one_template_vertex_each(cand, walks)
            f_memo, _ = _tds_failures(orc, g, c, cand, ea, k, orc.FAITHFUL_NO_SHORTCUT)
            f_full, _ = _tds_failures(orc, g, c, cand, ea, k, orc.COMPLETE)
            assert np.all((f_full & ~f_memo) == 0)  # memos only prune
            for root, moves in walks:
                bit = np.uint64(1 << root)
                if walk_is_memo_inert(root, moves):
                    inert_walks += 1
                    assert np.array_equal(f_memo & bit, f_full & bit), (name, seed, k, root)
                    held = sum(1 for x in cand if (int(x) >> root) & 1)
                    failed = int(np.count_nonzero(f_full & bit))
                    inert_ok += held - failed
                    inert_fail += failed
                else:
                    blocked += int(np.count_nonzero((f_memo & ~f_full) & bit))
    assert inert_walks and inert_ok and inert_fail
    print(f"{name}: inert walks {inert_walks}: {inert_ok} sources succeed, {inert_fail} fail; "
          f"non-inert walks blocked {blocked} sources the complete search keeps")


def test_tds_memo_inertness_hypotheses_on_known_walks(pj):
    """The template side of the theorem on walks traced by hand: the test_3
    diamond's walk from 2 returns to the unstored vertex 0 (not inert), the
    walks from 0, 1 and 3 never move back (inert); the chorded two-cycle
    template's walk from 3 is the only one of its six without a MOVE_BACK to
    an unstored vertex."""
    d = pj.Constraints(pj.Graph.pattern(4, DIAMOND, [0, 1, 2, 3]))
    got = {r: walk_is_memo_inert(r, mv) for r, mv in d.tds()[0]}
    assert got == {0: True, 1: True, 2: False, 3: True}
    t = pj.Constraints(pj.Graph.pattern(8, TWO_CYCLES_CHORD, list(range(8))))
    got = {r: walk_is_memo_inert(r, mv) for r, mv in t.tds()[0]}
    assert got == {0: False, 1: False, 2: False, 3: True, 4: False, 5: False}


# ---------------------------------------------------------------------------
# Known answers traced by hand


def _test3_constraints(pj):
    f = fixture_paths("test_3")
    return pj.Constraints(pj.Graph.load(f["pattern"], f["pattern_vertex"]))


def test_edge_memo_known_answer(pj, orc):
    """Triangle s=0, a=1, b=2 with C(s) = {1, 3}, C(a) = {0}, C(b) = {2}
    under the test_3 template (labels 0, 1, 2, 1: template vertices 1 and 3
    share a label).  TDS constraint 0 walks roots 0, 1, 2, 3.

    Root 3 from s: VS3 VS0 VN1 VN2 CHECK0/1 CHECK3/0.  s->a as 0 puts 0 into
    the edge memo of {s, a}.  At a the move towards 1 tests x = 0 (the
    template vertex of the previous move) on {a, s}: 0 is there, so the only
    candidate, s itself, is skipped (f(1) = f(3) = s is the only
    homomorphism).  b lacks 1; the walk fails.  Testing y = 1 instead would
    pass: {a, s} holds only 0.
    Root 1 from s: s->a (0), a->b (2), CHECK s, then b->s as 3 tests x = 1
    (the CHECK's vertex) on {b, s}, empty; CHECK a; success, marking a with 0,
    b with 2 and s with 3.
    So with every walk run (mode 2) s fails 3 and nothing else fails; the
    complete search fails nothing; with the alreadyMatched shortcut on one
    thread s runs root 1 first, is marked with 3 and skips root 3."""
    c = _test3_constraints(pj)
    assert [r for r, _ in c.tds()[0]] == [0, 1, 2, 3]
    g = pj.Graph.from_csr([0, 2, 4, 6], [1, 2, 0, 2, 0, 1], [1, 0, 2])
    ea = np.array([0x3F], dtype=np.uint32)
    cand = [0b1010, 0b0001, 0b0100]
    f2, m2 = _tds_failures(orc, g, c, cand, ea, 0, orc.FAITHFUL_NO_SHORTCUT)
    assert [int(x) for x in f2] == [0b1000, 0, 0]
    assert [int(x) for x in m2] == [0b1010, 0b0001, 0b0100]
    f1, _ = _tds_failures(orc, g, c, cand, ea, 0, orc.COMPLETE)
    assert not f1.any()
    f0, _ = _tds_failures(orc, g, c, cand, ea, 0, orc.FAITHFUL, threads=1)
    assert not f0.any()


def test_hypothesis_memo_known_answer(pj, orc):
    """s=0 with C = {2}; a1=1, a2=2 with C = {0}; b=3 with C = {1}; c=4 with
    C = {3}.  Edges s-a1, s-a2, s-b, s-c, a1-b, a2-b, a2-c.  Template test_3,
    TDS constraint 0.

    Root 2 from s: VS2 VN0 VN1 CHECK2/0 BACK0/1 VN3 CHECK2/0; nothing after
    the root is stored, so the hypothesis stays empty.  s->a1 (0), a1->b (1)
    records (b, 1, []); CHECK s holds; back to a1, which has no neighbour
    with 3: fail.  s->a2 (0): the move a2->b towards 1 finds (b, 1, [])
    recorded and skips b, although a2 has the neighbour c; no other
    neighbour of a2 holds 1, and no other neighbour of s holds 0: s fails 2,
    which the complete search (s, a2, b, c) satisfies.
    a1 fails 0 in both searches (no neighbour of a1 holds 3); a2 (root 0), b
    (root 1, through a2 after a1 fails) and c (root 3) succeed.  With the
    shortcut on one thread s runs first, before any mark exists."""
    c = _test3_constraints(pj)
    off = [0, 4, 6, 9, 12, 14]
    ed = [1, 2, 3, 4, 0, 3, 0, 3, 4, 0, 1, 2, 0, 2]
    g = pj.Graph.from_csr(off, ed, [2, 0, 0, 1, 1])
    ea = np.array([(1 << 14) - 1], dtype=np.uint32)
    cand = [0b0100, 0b0001, 0b0001, 0b0010, 0b1000]
    expect = [0b0100, 0b0001, 0, 0, 0]
    for mode, threads in ((orc.FAITHFUL_NO_SHORTCUT, 1), (orc.FAITHFUL, 1)):
        f, _ = _tds_failures(orc, g, c, cand, ea, 0, mode, threads)
        assert [int(x) for x in f] == expect, mode
    f1, _ = _tds_failures(orc, g, c, cand, ea, 0, orc.COMPLETE)
    assert [int(x) for x in f1] == [0, 0b0001, 0, 0, 0]


# ---------------------------------------------------------------------------
# Phase-2 scheduling rule


def _check_schedule(g, pre_vf, post):
    cand, vf, ea, *_ = post
    off, ed = g.offsets(), g.edges()
    bits = np.unpackbits(ea.view(np.uint8), bitorder="little")
    modified = (vf >> 1) & 1
    scheduled = (vf >> 2) & 1
    n_unmod = n_hit = n_blocked = 0
    for v in range(g.vertex_count):
        if not pre_vf[v] & 1:
            assert scheduled[v] == (pre_vf[v] >> 2) & 1  # untouched
            continue
        if modified[v]:
            assert scheduled[v] == 1, v
            continue
        hits = [bool(bits[e]) and bool(modified[int(ed[e])]) for e in range(int(off[v]), int(off[v + 1]))]
        assert scheduled[v] == any(hits), v
        n_unmod += 1
        n_hit += any(hits)
        n_blocked += (not any(hits)) and any(modified[int(ed[e])] for e in range(int(off[v]), int(off[v + 1])))
    return n_unmod, n_hit, n_blocked


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_phase2_schedules_exactly_the_vertices_next_to_modified_ones(pj, orc, seed):
    """From states with random modified and scheduled flags and random edge
    activity, every LCC step and every NLCC step leaves: modified vertices
    scheduled; an unmodified active vertex scheduled iff one of its active
    slots leads to a modified vertex (the vertex need not be active); other
    vertices' flags as they were."""
    rng = np.random.default_rng(seed)
    g = pj.Graph.rmat(9, 8, seed=seed).labels_uniform(5, seed=seed)
    p = pj.Graph.pattern(6, [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5)], [0, 1, 2, 3, 4, 3])
    c = pj.Constraints(p)
    o = orc.Oracle(g, c, threads=2)
    o.step(STEP_LCC0)
    base = o.state()
    totals = np.zeros(3, dtype=np.int64)
    for step, index in ((STEP_LCC, 0), (STEP_CC, 0), (STEP_PC, 0), (STEP_LCC, 0)):
        cand, vf, ea, av, ae = base
        vf = (vf & 1) | (rng.random(vf.size) < 0.3).astype(np.uint8) << 1 | (rng.random(vf.size) < 0.5).astype(
            np.uint8) << 2
        if step != STEP_LCC:
            vf &= ~np.uint8(2)  # NLCC steps start after resetState
        bits = np.unpackbits(ea.view(np.uint8), bitorder="little")
        bits[: g.edge_count] &= (rng.random(g.edge_count) < 0.8).astype(np.uint8)
        ea = np.packbits(bits, bitorder="little").view(np.uint32)[: ea.size].copy()
        x = orc.Oracle(g, c, threads=2)
        x.load_state(cand, vf, ea, av, ae)
        x.step(step, index, orc.COMPLETE)
        totals += _check_schedule(g, vf, x.state())
    n_unmod, n_hit, n_blocked = totals
    # non-vacuous: some unmodified vertices are scheduled, some are not, and
    # some reach a modified vertex only through inactive slots
    assert 0 < n_hit < n_unmod and n_blocked > 0


# ---------------------------------------------------------------------------
# TDS constraint and walk order of the fixtures


def _roots(c):
    return [[r for r, _ in walks] for walks in c.tds()]


def _fmt(moves):
    name = {MOVE_BACK: "BACK", MOVE_CHECK: "CHECK", MOVE_VERTEX_STORE: "VS", MOVE_VERTEX_NO_STORE: "VN"}
    return " ".join(f"{name[m]}{x}" + (f"/{mb}" if m in (MOVE_BACK, MOVE_CHECK) else "") for m, x, mb in moves)


def test_tds_order_test3_by_hand(pj):
    """test_3: cycles [0,1,2], [0,2,3], [0,1,2,3].  The cycle pass merges
    the first two (common edge 0-2) into M1, erases both from the front of
    the deque, then merges [0,1,2,3] with M1: one E_TDS_CC template M2.  Its
    vertex graph is a copy of [0,1,2,3]'s (keys first inserted 0, 1, 2, 3).
    The four paths all start at label 1 and merge pairwise into one E_TDS_PC
    template whose vertex graph is a copy of path [3,2,0,1]'s (keys 3, 2, 0,
    1).  The subtemplate pass merges both into a graph equal to M2's, already
    in the list, so nothing is added.  libstdc++'s unordered_map with small
    integer keys (13 buckets from the first insertion) iterates in reverse
    order of first insertion, and generateWalkMap inserts walks in the vertex
    graph's order, so walk maps iterate in the order the keys were first
    inserted into the vertex graph: [0,1,2,3] and [3,2,0,1].  Walks: DFS over
    ascending neighbour sets; a CHECK per non-tree edge; MOVE_BACK before
    leaving a vertex again after a subtree; VERTEX_STORE for vertices a
    later CHECK targets."""
    c = _test3_constraints(pj)
    assert _roots(c) == [[0, 1, 2, 3], [3, 2, 0, 1]]
    w = dict(c.tds()[0])
    assert _fmt(w[0]) == "VS0 VN1 VN2 CHECK0/0 VN3 CHECK0/0"
    assert _fmt(w[1]) == "VS1 VS0 VN2 CHECK1/0 VN3 CHECK0/1"
    assert _fmt(w[2]) == "VS2 VN0 VN1 CHECK2/0 BACK0/1 VN3 CHECK2/0"
    assert _fmt(w[3]) == "VS3 VS0 VN1 VN2 CHECK0/1 CHECK3/0"
    assert dict(c.tds()[1]) == w


def test_tds_order_test4_by_hand(pj):
    """test_4 (10 vertices, labels 1,2,3,4,2,6,5,5,6,2).  Cycles [2,3,4,6],
    [2,3,4,7,8], [2,6,4,7,8] merge into one E_TDS_CC template: the first two
    (common edge 2-3), then [2,6,4,7,8] with that (vertex graph: keys
    2,6,4,7,8 from the copy, then 3 from the merged pair).  Paths group by
    start label into an unordered_map first filled with labels 2, 6, 5, so
    the groups are visited 5, 6, 2; each group merges down to one template
    (group 5: vertices 7,4,3,2,6 then 8; group 6: 8,7,4,5 then 6,3,2; group 2:
    9,8,7,4 then 6,1,2,3).  The subtemplate pass merges all four into a
    fifth template over vertices 1..9 (keys 2,6,4,7,8,3 then 5 then 1, 9).
    Walk maps iterate in first-insertion order of each vertex graph."""
    f = fixture_paths("test_4")
    c = pj.Constraints(pj.Graph.load(f["pattern"], f["pattern_vertex"]))
    assert _roots(c) == [
        [2, 6, 4, 7, 8, 3],
        [7, 4, 3, 2, 6, 8],
        [8, 7, 4, 5, 6, 3, 2],
        [9, 8, 7, 4, 6, 1, 2, 3],
        [2, 6, 4, 7, 8, 3, 5, 1, 9],
    ]
    w = dict(c.tds()[0])
    # root 2: 2 -> 3 -> 4 -> 6, CHECK 2, back to 4 -> 7 -> 8, CHECK 2
    assert _fmt(w[2]) == "VS2 VN3 VN4 VN6 CHECK2/0 BACK4/2 VN7 VN8 CHECK2/0"
    assert _fmt(dict(c.tds()[2])[5]) == "VN5 VS4 VN3 VN2 VN6 CHECK4/1 BACK2/3 VN8 VN7 CHECK4/1"


# ---------------------------------------------------------------------------
# Cycle and path memos: known answers traced by hand

# 4-cycle 0-1-3-2-0 with labels 0,1,0,1: normalised cycle [0, 1, 3, 2] (so
# position 2 holds template vertex 3 and position 3 holds 2), and the paths
# [2, 3, 1, 0] and [3, 2, 0, 1] from label 0 (positions differ from ids).
SQUARE = (4, [(0, 1), (1, 3), (3, 2), (2, 0)], [0, 1, 0, 1])


def _nlcc_failures(orc, g, c, cand, step, index, mode):
    ea = np.full((g.edge_count + 31) // 32, 0xFFFFFFFF, dtype=np.uint32)
    o = _load(orc, g, c, cand, ea)
    o.step(step, index, mode)
    return [int(x) for x in np.asarray(cand, dtype=np.uint64) & ~o.state()[0]]


def test_cycle_memo_known_answer(pj, orc):
    """One edge 0-1, C(0) = {0, 2, 3}, C(1) = {1, 2}, cycle [0, 1, 3, 2]:
    every closed walk goes back and forth over the one edge.  The
    reference's memo inserts template ids and tests positions
    (cc_cpu.tpp:53-61).  Start 0 from vertex 0: 0->1 towards id 1 stores 1
    for {0, 1}; at 1 (position 1) the move towards id 3 tests 1: blocked,
    fail.  Start 2 from 0: stores 2, then 0 (at position 3, 3 is absent),
    then at position 0 tests 0: blocked.  Start 3 from 0 needs id 0 at 1:
    absent.  Vertex 1, start 1: stores 3, then 2 (position 2 tested against
    {3}), then at position 3 tests 3: blocked; start 3: stores 0, then at
    position 0 tests 0: blocked.  So the memoised search fails 0, 2, 3 at
    vertex 0 and 1, 2 at vertex 1; the complete search fails only 2 at 0
    (testing the id instead would find every other walk)."""
    c = pj.Constraints(pj.Graph.pattern(*SQUARE))
    assert c.cycles() == [[0, 1, 3, 2]]
    g = pj.Graph.from_csr([0, 1, 2], [1, 0], [0, 1])
    cand = [0b1101, 0b0110]
    assert _nlcc_failures(orc, g, c, cand, STEP_CC, 0, orc.FAITHFUL_NO_SHORTCUT) == [0b1101, 0b0110]
    assert _nlcc_failures(orc, g, c, cand, STEP_CC, 0, orc.COMPLETE) == [0b0100, 0]


def test_path_memo_known_answer(pj, orc):
    """Triangle 0, 1, 2 (all label 0), C(0) = {1, 3}, C(1) = {0, 1},
    C(2) = {0, 1, 2, 3}, path [2, 3, 1, 0] (forward end id 2, reverse end id
    0; pc_cpu.tpp:104-142).  Vertex 1, reverse: 1->0 (id 1) finds no id 3
    neighbour leading on, stores 3 for {0, 2} on the way; 1->2 (id 1), then
    2->0 towards id 3 tests position 2 against {3}: not blocked; 0->2 ends
    at id 2 (2 differs from the source): success.  Vertex 2, forward:
    2->0 (3), 0->1 (1), 1 has only the source with id 0; 0->2 (1, position 1
    tested against {3}), 2->1 ends at id 0: success.  Vertex 2, reverse: the
    only vertex with id 2 is the source itself: fail.  So only vertex 2
    loses id 0, in both searches (testing ids instead of positions would
    block 2->0 and fail vertex 1 as well)."""
    c = pj.Constraints(pj.Graph.pattern(*SQUARE))
    assert c.paths()[0] == [2, 3, 1, 0]
    g = pj.Graph.from_csr([0, 2, 4, 6], [1, 2, 0, 2, 0, 1], [0, 0, 0])
    cand = [0b1010, 0b0011, 0b1111]
    for mode in (orc.FAITHFUL_NO_SHORTCUT, orc.COMPLETE):
        assert _nlcc_failures(orc, g, c, cand, STEP_PC, 0, mode) == [0, 0, 0b0001], mode


def test_path_end_differs_from_the_source_known_answer(pj, orc):
    """Triangle 0, 1, 2, C(0) = {1, 3}, C(1) = {0, 1, 2}, C(2) = {2, 3},
    path [2, 3, 1, 0].  Vertex 1 is the only vertex holding id 0, so its
    forward walks (1->0->1, 1->2->0 then back to 1) can only end at the
    source, which the last hop rejects (pc_cpu.tpp:29-45); its reverse walk
    1->0->2 can only end at id 2 on the source again.  Vertex 2's forward
    walk 2->0 (id 3), 0->1 (id 1) then needs a neighbour of 1 other than 2
    holding id 0: none; 2->1 needs id 3 at 1: absent.  So vertex 1 loses 0
    and 2, vertex 2 loses 2; accepting the source as the end would keep
    vertex 1's ids."""
    c = pj.Constraints(pj.Graph.pattern(*SQUARE))
    g = pj.Graph.from_csr([0, 2, 4, 6], [1, 2, 0, 2, 0, 1], [0, 0, 0])
    cand = [0b1010, 0b0111, 0b1100]
    for mode in (orc.FAITHFUL_NO_SHORTCUT, orc.COMPLETE):
        assert _nlcc_failures(orc, g, c, cand, STEP_PC, 0, mode) == [0, 0b0101, 0b0100], mode
