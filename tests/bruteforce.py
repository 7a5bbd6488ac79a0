"""Independent brute-force references for small inputs (pure Python).

These restate nothing of the reference's code: they compute what the
mathematics fixes (embeddings, greatest fixed points, walk existence) by
exhaustive search, so they can pin the oracle."""
from __future__ import annotations

import itertools
from collections import deque


def csr_rows(offsets, edges):
    return [list(edges[offsets[v]:offsets[v + 1]]) for v in range(len(offsets) - 1)]


def embeddings(g_rows, g_lab, p_rows, p_lab):
    """All injective maps f: template -> graph with lab(f(x)) == plab(x) and
    every template edge (x, y) mapped onto a graph slot f(x) -> f(y)."""
    P = len(p_rows)
    gset = [set(r) for r in g_rows]
    by_label = {}
    for v, l in enumerate(g_lab):
        by_label.setdefault(int(l), []).append(v)
    pedges = [(x, y) for x in range(P) for y in p_rows[x]]
    out = []

    def extend(x, f, used):
        if x == P:
            out.append(tuple(f))
            return
        for v in by_label.get(int(p_lab[x]), []):
            if v in used:
                continue
            ok = True
            for (a, b) in pedges:
                if a == x and b < x and f[b] not in gset[v]:
                    ok = False
                    break
                if b == x and a < x and v not in gset[f[a]]:
                    ok = False
                    break
            if ok:
                f.append(v)
                used.add(v)
                extend(x + 1, f, used)
                f.pop()
                used.discard(v)

    extend(0, [], set())
    return out


def ideal_candidates(g_rows, g_lab, p_rows, p_lab):
    cand = {}
    for f in embeddings(g_rows, g_lab, p_rows, p_lab):
        for x, v in enumerate(f):
            cand.setdefault(v, set()).add(x)
    return cand


def lcc_greatest_fixed_point(g_rows, g_lab, p_rows, p_lab, order_seed=0):
    """Greatest candidate assignment closed under the local constraint, by
    chaotic (random-order, in-place) iteration of its definition:
    p survives at v iff lab(v) == plab(p) and, counting each slot v->u whose
    u still has a candidate among the template neighbours of v's candidates,
    v has at least as many neighbours of each label as p has in the template.
    Edge slots are dropped once no candidate pair supports them."""
    import random

    V, P = len(g_rows), len(p_rows)
    nbr = [set(r) for r in p_rows]
    need = []
    for p in range(P):
        d = {}
        for q in p_rows[p]:
            d[int(p_lab[q])] = d.get(int(p_lab[q]), 0) + 1
        need.append(d)
    cand = [{p for p in range(P) if int(p_lab[p]) == int(g_lab[v])} for v in range(V)]
    alive = [[True] * len(r) for r in g_rows]
    rng = random.Random(order_seed)
    changed = True
    while changed:
        changed = False
        order = list(range(V))
        rng.shuffle(order)
        for v in order:
            if not cand[v]:
                continue
            U = set().union(*(nbr[p] for p in cand[v]))
            counts = {}
            for i, u in enumerate(g_rows[v]):
                if not alive[v][i] or not cand[u]:
                    continue
                if not (cand[u] & U):
                    alive[v][i] = False
                    changed = True
                    continue
                counts[int(g_lab[u])] = counts.get(int(g_lab[u]), 0) + 1
            keep = {p for p in cand[v] if all(counts.get(l, 0) >= k for l, k in need[p].items())}
            if keep != cand[v]:
                cand[v] = keep
                changed = True
    return cand


def walk_exists(g_rows, active_slot, cand, start, positions, closing_target=None, end_excl=None):
    """Breadth-first search over (vertex, position index) states: is there a
    walk start=w0 -> w1 -> ... -> wk with active slots, w_i having template
    vertex positions[i] among its candidates for i >= 1, and, when
    closing_target is given, an active slot wk -> closing_target?  With
    end_excl the last vertex must differ from it."""
    k = len(positions) - 1
    frontier = {start}
    for i in range(1, k + 1):
        q = positions[i]
        nxt = set()
        for w in frontier:
            for idx, u in enumerate(g_rows[w]):
                if active_slot(w, idx) and cand[u] and (cand[u] >> q) & 1:
                    nxt.add(u)
        frontier = nxt
        if not frontier:
            return False
    if end_excl is not None:
        frontier = {w for w in frontier if w != end_excl}
    if closing_target is None:
        return bool(frontier)
    for w in frontier:
        for idx, u in enumerate(g_rows[w]):
            if active_slot(w, idx) and u == closing_target:
                return True
    return False


def simple_cycles(p_rows):
    """Every simple cycle of length >= 3 of an undirected template, as the
    tuple starting at its smallest vertex and continuing towards the smaller
    of that vertex's two cycle neighbours."""
    P = len(p_rows)
    adj = [set(r) for r in p_rows]
    out = set()
    for L in range(3, P + 1):
        for combo in itertools.permutations(range(P), L):
            if combo[0] != min(combo):
                continue
            if all(combo[(i + 1) % L] in adj[combo[i]] for i in range(L)):
                c = list(combo)
                if c[-1] < c[1]:
                    c = [c[0]] + c[1:][::-1]
                out.add(tuple(c))
    return out


def label_paths(p_rows, p_lab):
    """Simple paths h0..hk, k >= 2, with label(hk) == label(h0), hk != h0 and
    no vertex at index 2..k-1 carrying label(h0); oriented so that the first
    id is not smaller than the last."""
    P = len(p_rows)
    out = set()

    def dfs(path):
        k = len(path) - 1
        s, cur = path[0], path[-1]
        if k >= 2 and p_lab[cur] == p_lab[s] and cur != s:
            p = path if path[-1] <= path[0] else path[::-1]
            out.add(tuple(p))
            return
        for n in p_rows[cur]:
            if n in path:
                continue
            dfs(path + [n])

    for s in range(P):
        dfs([s])
    return out


def walk_template_edges(root, moves):
    """Undirected template edges a TDS walk covers: each VERTEX move is an
    edge from the walk's current template vertex to the new one, each CHECK
    an edge from the current template vertex to the checked one."""
    hist, cur, edges = [root], root, set()
    for mv, x, mb in moves[1:]:
        if mv in (2, 3):  # VERTEX_STORE, VERTEX_NO_STORE
            edges.add(frozenset((cur, x)))
            hist.append(x)
            cur = x
        elif mv == 1:  # CHECK
            edges.add(frozenset((cur, x)))
        else:  # MOVE_BACK
            cur = hist[mb]
    return edges


def homomorphism_exists(g_adj, cand, t_edges, root, v):
    """Is there a map f from the template vertices of t_edges to graph
    vertices with f(root) = v, every template vertex x among the candidates
    of f(x), and every template edge {x, y} on an active undirected graph
    edge {f(x), f(y)}?  g_adj[w] is the set of active neighbours of w.
    Not necessarily injective.  Plain backtracking in BFS order from root."""
    nbrs = {}
    for e in t_edges:
        a, b = tuple(e)
        nbrs.setdefault(a, set()).add(b)
        nbrs.setdefault(b, set()).add(a)
    order, seen = [root], {root}
    for x in order:
        for y in sorted(nbrs[x]):
            if y not in seen:
                seen.add(y)
                order.append(y)
    if not (cand[v] >> root) & 1:
        return False
    f = {root: v}

    def extend(i):
        if i == len(order):
            return True
        x = order[i]
        placed = [y for y in nbrs[x] if y in f]
        pool = set(g_adj[f[placed[0]]])
        for y in placed[1:]:
            pool &= g_adj[f[y]]
        for w in sorted(pool):
            if (cand[w] >> x) & 1:
                f[x] = w
                if extend(i + 1):
                    return True
                del f[x]
        return False

    return extend(1)


def random_symmetric_graph(n, p, rng):
    """Undirected simple graph on n vertices (each pair with probability p),
    as ascending CSR rows (offsets, edges)."""
    adj = [[] for _ in range(n)]
    for a in range(n):
        for b in range(a + 1, n):
            if rng.random() < p:
                adj[a].append(b)
                adj[b].append(a)
    off, ed = [0], []
    for a in range(n):
        ed += sorted(adj[a])
        off.append(len(ed))
    return off, ed
