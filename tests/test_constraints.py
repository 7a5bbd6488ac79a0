"""Host constraint generation against brute force and hand-checked answers."""
import pytest

from bruteforce import csr_rows, label_paths, simple_cycles
from conftest import FIXTURES, fixture_paths
from pattenmatching_sharedmemory_amd.native import (MOVE_BACK, MOVE_CHECK, MOVE_VERTEX_NO_STORE, MOVE_VERTEX_STORE,
                                                     STEP_CC, STEP_PC)


def load_pattern(pj, name):
    f = fixture_paths(name)
    return pj.Graph.load(f["pattern"], f["pattern_vertex"])


def pattern_rows(p):
    return csr_rows(p.offsets(), [int(x) for x in p.edges()])


@pytest.mark.parametrize("name", FIXTURES)
def test_local_constraints_from_pattern_rows(pj, name):
    p = load_pattern(pj, name)
    c = pj.Constraints(p)
    rows, lab = pattern_rows(p), [int(x) for x in p.labels()]
    for q, lc in enumerate(c.local()):
        assert lc["label"] == lab[q]
        need = {}
        for n in rows[q]:
            need[lab[n]] = need.get(lab[n], 0) + 1
        assert lc["need"] == need
        assert lc["neighbors"] == sum(1 << n for n in set(rows[q]))


EXTRA_TEMPLATES = [
    (6, [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5)], [0, 1, 2, 3, 4, 3]),
    (8, [(0, 1), (1, 2), (2, 3), (3, 0), (1, 4), (4, 5), (5, 2), (0, 2), (5, 6), (6, 7)], list(range(8))),
    (7, [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5), (5, 2), (1, 6)], [0, 1, 2, 3, 4, 3, 1]),
    (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0), (0, 2)], [1, 1, 1, 1, 1]),
]


def templates(pj):
    out = [(n, load_pattern(pj, n)) for n in FIXTURES]
    out += [(f"t{i}", pj.Graph.pattern(n, e, l)) for i, (n, e, l) in enumerate(EXTRA_TEMPLATES)]
    return out


def test_cycles_are_all_simple_cycles_normalised_and_ordered(pj):
    for name, p in templates(pj):
        c = pj.Constraints(p)
        got = [tuple(x) for x in c.cycles()]
        assert len(got) == len(set(got)), name
        assert set(got) == simple_cycles(pattern_rows(p)), name
        # first found from its smallest vertex, shorter cycles first
        keys = [(cyc[0], len(cyc)) for cyc in got]
        assert keys == sorted(keys), name


def test_paths_are_all_label_paths_normalised(pj):
    for name, p in templates(pj):
        c = pj.Constraints(p)
        lab = [int(x) for x in p.labels()]
        got = [tuple(x) for x in c.paths()]
        assert len(got) == len(set(got)), name
        assert set(got) == label_paths(pattern_rows(p), lab), name
        assert c.path_labels() == [lab[q[0]] for q in got]


def test_test3_constraints_by_hand(pj):
    c = pj.Constraints(load_pattern(pj, "test_3"))
    # diamond 0-1, 0-2, 0-3, 1-2, 2-3 with labels (0, 1, 2, 1)
    assert c.cycles() == [[0, 1, 2], [0, 2, 3], [0, 1, 2, 3]]
    assert sorted(map(tuple, c.paths())) == sorted([(3, 2, 0, 1), (3, 0, 1), (3, 0, 2, 1), (3, 2, 1)])


def check_walk(root, moves, adjacency):
    """A walk visits every template vertex of `adjacency` once by VERTEX
    moves, covers every template edge exactly once (tree edges by VERTEX
    moves, the rest by CHECK moves) and only moves back to earlier history."""
    hist = [root]
    cur = root
    covered = set()
    assert moves[0][0] in (MOVE_VERTEX_NO_STORE, MOVE_VERTEX_STORE) and moves[0][1] == root
    for mv, x, mb in moves[1:]:
        if mv in (MOVE_VERTEX_NO_STORE, MOVE_VERTEX_STORE):
            assert x in adjacency[cur] and x not in hist
            covered.add(frozenset((cur, x)))
            hist.append(x)
            cur = x
        elif mv == MOVE_CHECK:
            assert hist[mb] == x and x in adjacency[cur]
            e = frozenset((cur, x))
            assert e not in covered
            covered.add(e)
        else:
            # back to the branching vertex x, history entry mb
            assert mv == MOVE_BACK and mb < len(hist) and hist[mb] == x
            cur = x
    assert set(hist) == set(adjacency)
    all_edges = {frozenset((a, b)) for a in adjacency for b in adjacency[a]}
    assert covered == all_edges
    # STORE marks exactly the vertices a later CHECK returns to
    checked = {x for mv, x, _ in moves if mv == MOVE_CHECK}
    stored = {x for mv, x, _ in moves if mv == MOVE_VERTEX_STORE}
    assert stored == checked - ({root} if moves[0][0] == MOVE_VERTEX_NO_STORE else set())


def test_template_walks_are_valid(pj):
    for name, p in templates(pj):
        c = pj.Constraints(p)
        rows = pattern_rows(p)
        full = {v: set(rows[v]) for v in range(len(rows)) if rows[v]}
        root, moves = c.enumeration_walk()
        check_walk(root, moves, full)
        for t in c.tds():
            verts = set()
            for r, mv in t:
                verts |= {x for _, x, _ in mv}
            sub = {v: {u for u in full[v] if u in verts} for v in verts}
            roots = [r for r, _ in t]
            assert sorted(roots) == sorted(verts)  # one walk per template vertex
            for r, mv in t:
                edges = set()
                cur, hist = r, [r]
                for m, x, mb in mv[1:]:
                    if m in (MOVE_VERTEX_NO_STORE, MOVE_VERTEX_STORE):
                        edges.add(frozenset((cur, x)))
                        hist.append(x)
                        cur = x
                    elif m == MOVE_CHECK:
                        edges.add(frozenset((cur, x)))
                    else:
                        cur = hist[mb]
                # the walk covers the merged constraint's edges: a subgraph of the template
                assert all(len(e) == 2 and max(e) in full[min(e)] for e in edges), name


def test_single_cycle_yields_no_tds_and_merged_cycles_do(pj):
    four_cycle = pj.Constraints(pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 3), (3, 0)], [0, 1, 2, 3]))
    assert four_cycle.cycles() == [[0, 1, 2, 3]] and four_cycle.paths() == [] and four_cycle.tds() == []
    tail = pj.Constraints(pj.Graph.pattern(*EXTRA_TEMPLATES[0]))
    assert tail.cycles() == [[0, 1, 2]] and tail.paths() == [[5, 4, 3]] and tail.tds() == []
    two = pj.Constraints(pj.Graph.pattern(*EXTRA_TEMPLATES[1]))
    assert len(two.tds()) >= 1


def test_memo_inertness(pj):
    c = pj.Constraints(pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 3), (3, 0)], [0, 1, 2, 3]))
    assert c.memo_inert(STEP_CC, 0) == 1
    t = pj.Constraints(pj.Graph.pattern(*EXTRA_TEMPLATES[0]))
    assert t.memo_inert(STEP_CC, 0) == 1 and t.memo_inert(STEP_PC, 0) == 1
    same = pj.Constraints(pj.Graph.pattern(*EXTRA_TEMPLATES[3]))  # every label equal: the memo can fire
    assert 0 in [same.memo_inert(STEP_CC, i) for i in range(len(same.cycles()))]
    assert c.memo_inert(STEP_PC, 0) == -1  # no such path


def test_pattern_requirements(pj):
    with pytest.raises(ValueError):
        pj.Constraints(pj.Graph.pattern(1, [], [0]))  # no edges
    g = pj.Graph.rmat(4, 1, seed=1)  # unlabelled
    with pytest.raises(ValueError, match="labels"):
        pj.Constraints(g)
