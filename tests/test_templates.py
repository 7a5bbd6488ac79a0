"""Seeded random query templates of the batched and served forms."""
from pattenmatching_sharedmemory_amd.templates import longest_cycle, random_templates


def _connected(k, edges):
    adj = {v: set() for v in range(k)}
    for a, b in edges:
        adj[a].add(b)
        adj[b].add(a)
    seen, todo = {0}, [0]
    while todo:
        for w in adj[todo.pop()]:
            if w not in seen:
                seen.add(w)
                todo.append(w)
    return len(seen) == k


def test_random_templates_are_seeded_connected_and_bounded():
    a = random_templates(50, 11, 3, 10, 3)
    assert a == random_templates(50, 11, 3, 10, 3)
    assert a != random_templates(50, 12, 3, 10, 3)
    for k, edges, labels in a:
        assert 3 <= k <= 10 and len(labels) == k
        assert k - 1 <= len(edges) <= k - 1 + 3
        assert len({tuple(sorted(e)) for e in edges}) == len(edges)
        assert all(a != b for a, b in edges)
        assert _connected(k, edges)
        assert all(0 <= x < 32 for x in labels)


def test_distinct_labels_and_cycle_bound(pj):
    for k, edges, labels in random_templates(40, 3, 3, 12, 4, distinct_labels=True):
        assert len(set(labels)) == k
        c = pj.Constraints(pj.Graph.pattern(k, edges, labels))
        assert not c.paths()  # paths join template vertices of equal labels
        assert longest_cycle(c) <= k


def test_projected_bytes_grow_with_the_graph(pj):
    from pattenmatching_sharedmemory_amd import native

    k, edges, labels = random_templates(1, 1, 5, 5, 0)[0]
    c = pj.Constraints(pj.Graph.pattern(k, edges, labels))
    small = native.host().pjh_projected_bytes(c.handle, 1 << 10, 1 << 14)
    large = native.host().pjh_projected_bytes(c.handle, 1 << 20, 1 << 24)
    assert 0 < small < large


def test_projected_bytes_count_memo_tables_and_view(pj):
    """A 3-vertex cycle whose memo may fire (repeated label at the tested
    position) and a TDS template project the memo tables, alreadyMatched and
    the active view; a memo-inert triangle projects the state alone."""
    from pattenmatching_sharedmemory_amd import native

    h = native.host()
    plain = pj.Constraints(pj.Graph.pattern(3, [(0, 1), (1, 2), (2, 0)], [0, 1, 2]))
    memo = pj.Constraints(pj.Graph.pattern(5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 2)], [0, 3, 1, 2, 1]))
    tds = pj.Constraints(pj.Graph.pattern(4, [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3)], [0, 1, 2, 3]))
    V, E = 1 << 20, 1 << 24
    assert h.pjh_query_uses_memo(plain.handle) == 0
    assert h.pjh_query_uses_memo(memo.handle) == 1 and h.pjh_query_uses_memo(tds.handle) == 1
    base = h.pjh_projected_bytes(plain.handle, V, E)
    assert base == V * 57 + E // 8
    tables = 2048 * (1 << 16) * 20  # 2^20 vertices: 2^16 entries per wave (the floor)
    view = 2 * (V + 1) * 8 + (E + 1) * 4
    extra = V * 8 + 8 + (2 * V + 2) * 4 + tables + view
    assert h.pjh_projected_bytes(memo.handle, V, E) == base + extra
    assert h.pjh_projected_bytes(tds.handle, V, E) == base + extra
