"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline).  CPU restatement of the reference's label statistics and
constraint effectiveness estimate, the inputs and the rule of its dynamic
constraint order:

* accumulation_max: GraphStatCpu::preprocessPatern
  (src/patternmatching/algorithms/graph_stat_cpu.tpp:14-38);
* graph_stat (numpy) and graph_stat_loops (plain loops over vertices and
  slots, for small graphs): GraphStatCpu::compute (graph_stat_cpu.tpp:40-82)
  with GraphStat::addVertex / addEdge / addAccumulation
  (src/patternmatching/algorithms/graph_stat.cc:93-113), one tally per
  vertex;
* graph_stat_thread_tallies: the same with the reference's tally as written,
  one per OpenMP thread and never cleared (graph_stat_cpu.tpp:48-70), under
  OpenMP's static schedule; with one vertex per thread it is graph_stat;
* the estimate, effectiveness_cycle / _path / _tds:
  MultipleLabelConstraintEffectiveness::computeBacktrack
  (src/patternmatching/multiple_label/multiple_label_constraint_effectiveness.cc:
  226-340, 570-696, 972-1118) over the probability helpers (:1120-1230,
  multiple_label_constraint_effectiveness.tpp:10-14);
* choose_next: test_article/test_order_multiple_label_cpu.cc:184-274.

Statistics are dicts keyed by label value, as the reference's maps; a key
the reference never inserted reads 0 (the reference's at() would throw).
Arithmetic is IEEE double in the reference's order of operations (numpy
float64 scalars, so 0/0 and x/0 give NaN / inf as in C++).
"""
from __future__ import annotations

import math
from collections import defaultdict

import numpy as np

STEP_CC, STEP_PC, STEP_TDS, STEP_ENUMERATION = 4, 5, 6, 7
MOVE_BACK, MOVE_CHECK, MOVE_VERTEX_STORE, MOVE_VERTEX_NO_STORE = range(4)

F = np.float64


def _pow(x, y) -> np.float64:
    """C's pow (the C library's, through math.pow); the IEEE special cases
    math.pow reports as errors (0 to a negative power, a negative base to a
    fractional power) as C returns them."""
    try:
        return F(math.pow(float(x), float(y)))
    except (ValueError, OverflowError):
        return np.power(F(x), F(y))


# --------------------------------------------------------------------------
# statistics

def accumulation_max(pattern_offsets, pattern_edges, pattern_labels) -> dict:
    """accumulationMaximumMap[lf][lt]: over the template vertices of label lf,
    the most CSR slots to neighbours of label lt (graph_stat_cpu.tpp:14-38)."""
    out: dict = defaultdict(dict)
    for v in range(len(pattern_offsets) - 1):
        lf = int(pattern_labels[v])
        acc: dict = defaultdict(int)
        for e in range(int(pattern_offsets[v]), int(pattern_offsets[v + 1])):
            acc[int(pattern_labels[int(pattern_edges[e])])] += 1
        for lt, n in acc.items():
            if out[lf].get(lt, 0) < n:
                out[lf][lt] = n
    return dict(out)


def _active_slots(offsets, edges, cand, ea_words):
    V = len(offsets) - 1
    E = int(offsets[-1])
    active = np.asarray(cand[:V]) != 0
    deg = np.diff(np.asarray(offsets, dtype=np.int64))
    src = np.repeat(np.arange(V, dtype=np.int64), deg)
    dst = np.asarray(edges[:E], dtype=np.int64)
    bits = np.unpackbits(np.asarray(ea_words, dtype=np.uint32).view(np.uint8), bitorder="little")[:E].astype(bool)
    keep = bits & active[src] & active[dst]
    return active, src[keep], dst[keep]


def _empty_stat():
    return {"vertex_total": 0, "edge_total": 0, "vertex": {}, "edge": {}, "outbound": {}, "at_least": {}}


def graph_stat(offsets, edges, labels, cand, ea_words, acc_max: dict) -> dict:
    """GraphStat of the active graph, vectorised; each vertex's own tally of
    neighbour labels."""
    labels = np.asarray(labels, dtype=np.uint64)
    active, src, dst = _active_slots(offsets, edges, cand, ea_words)
    st = _empty_stat()
    vl, vc = np.unique(labels[active], return_counts=True)
    st["vertex"] = {int(l): int(c) for l, c in zip(vl, vc)}
    st["vertex_total"] = int(active.sum())
    lf, lt = labels[src], labels[dst]
    if len(src):
        pairs, pc = np.unique(np.stack([lf, lt]), axis=1, return_counts=True)
        for (a, b), c in zip(pairs.T, pc):
            st["edge"].setdefault(int(a), {})[int(b)] = int(c)
            st["outbound"][int(a)] = st["outbound"].get(int(a), 0) + int(c)
        st["edge_total"] = int(len(src))
        # per (vertex, neighbour label) tallies
        keys, tally = np.unique(np.stack([src.astype(np.uint64), lt]), axis=1, return_counts=True)
        for (v, b), n in zip(keys.T, tally):
            a = int(labels[int(v)])
            m = acc_max.get(a, {}).get(int(b), 0)
            for i in range(1, m + 1):
                if n >= i:
                    d = st["at_least"].setdefault(a, {}).setdefault(int(b), {})
                    d[i] = d.get(i, 0) + 1
    return st


def graph_stat_thread_tallies(offsets, edges, labels, cand, ea_words, acc_max: dict, threads: int) -> dict:
    """GraphStatCpu::compute as written: vertices split over `threads` by
    OpenMP's static schedule (contiguous ranges, the first V % threads one
    longer), each thread's tally of neighbour labels never cleared."""
    V = len(offsets) - 1
    E = int(offsets[-1])
    bits = np.unpackbits(np.asarray(ea_words, dtype=np.uint32).view(np.uint8), bitorder="little")[:E]
    st = _empty_stat()
    q, r = divmod(V, threads)
    start = 0
    for t in range(threads):
        n = q + (1 if t < r else 0)
        tally: dict = defaultdict(int)
        for v in range(start, start + n):
            if cand[v] == 0:
                continue
            a = int(labels[v])
            st["vertex_total"] += 1
            st["vertex"][a] = st["vertex"].get(a, 0) + 1
            for e in range(int(offsets[v]), int(offsets[v + 1])):
                u = int(edges[e])
                if not bits[e] or cand[u] == 0:
                    continue
                b = int(labels[u])
                st["edge_total"] += 1
                st["edge"].setdefault(a, {})[b] = st["edge"].get(a, {}).get(b, 0) + 1
                st["outbound"][a] = st["outbound"].get(a, 0) + 1
                tally[b] += 1
            _add_accumulation(st, acc_max, a, tally)
        start += n
    return st


def _add_accumulation(st, acc_max, a, tally):
    for b, n in tally.items():
        if b in acc_max.get(a, {}):
            for i in range(1, acc_max[a][b] + 1):
                if n >= i:
                    d = st["at_least"].setdefault(a, {}).setdefault(b, {})
                    d[i] = d.get(i, 0) + 1


def graph_stat_loops(offsets, edges, labels, cand, ea_words, acc_max: dict) -> dict:
    """graph_stat by plain loops, a fresh tally per vertex (small graphs)."""
    V = len(offsets) - 1
    E = int(offsets[-1])
    bits = np.unpackbits(np.asarray(ea_words, dtype=np.uint32).view(np.uint8), bitorder="little")[:E]
    st = _empty_stat()
    for v in range(V):
        if cand[v] == 0:
            continue
        a = int(labels[v])
        st["vertex_total"] += 1
        st["vertex"][a] = st["vertex"].get(a, 0) + 1
        tally: dict = defaultdict(int)
        for e in range(int(offsets[v]), int(offsets[v + 1])):
            u = int(edges[e])
            if not bits[e] or cand[u] == 0:
                continue
            b = int(labels[u])
            st["edge_total"] += 1
            st["edge"].setdefault(a, {})[b] = st["edge"].get(a, {}).get(b, 0) + 1
            st["outbound"][a] = st["outbound"].get(a, 0) + 1
            tally[b] += 1
        _add_accumulation(st, acc_max, a, tally)
    return st


# --------------------------------------------------------------------------
# estimate

class _Stat:
    def __init__(self, st: dict):
        self.st = st

    def vertices(self, l) -> np.float64:
        return F(self.st["vertex"].get(int(l), 0))

    def edges(self, a, b) -> np.float64:
        return F(self.st["edge"].get(int(a), {}).get(int(b), 0))

    def outbound(self, a) -> np.float64:
        return F(self.st["outbound"].get(int(a), 0))

    def at_least(self, a, b, k) -> np.float64:
        return F(self.st["at_least"].get(int(a), {}).get(int(b), {}).get(int(k), 0))


def proba_find_vertex(s: _Stat, a, b):
    """probaFindVertex (:1120-1133), no known edge or vertex."""
    full = s.vertices(a)
    set_size = full
    proba_different = _pow((full - 0 - 0) / (full - 0), 1)
    return s.at_least(a, b, 1) / set_size * proba_different


def proba_find_one_path(s: _Stat, n, a, b):
    """probaFindOnePathKnowing (:1141-1149)."""
    return F(1) - _pow(F(1) - proba_find_vertex(s, a, b), n)


def proba_find_same_vertex(s: _Stat, n, l):
    """probaFindSameVertex (:1151-1156)."""
    set_size = s.vertices(l)
    return F(1) - _pow((set_size - 1) / set_size, n)


def proba_find_different_vertex(s: _Stat, n, l):
    """probaFindDifferentVertex (:1158-1163)."""
    set_size = s.vertices(l)
    return F(1) - _pow(F(1.0) / set_size, n)


def average_vertex(s: _Stat, n, a):
    """averageNumberOfVertex (:1165-1171)."""
    return F(n) * s.outbound(a) / s.vertices(a)


def average_vertex_knowing(s: _Stat, n, a, b):
    """averageNumberOfVertexKnowing (:1180-1200), no known edge or vertex."""
    tries = F(n) * s.edges(a, b) / s.vertices(a)
    return tries * F(1)


def average_unique_vertex(s: _Stat, n, a, b):
    """averageNumberOfUniqueVertexKnowing (:1209-1230) with
    averageDifferentVertexFromTriesApproximateStrong (.tpp:10-14)."""
    tries = F(n) * s.edges(a, b) / s.vertices(a)
    max_to = s.vertices(b)
    return (tries if tries <= max_to else max_to) * F(1)


def _known_edge(s: _Stat, n, a, b):
    return F(1) - s.edges(a, b) / s.vertices(a) + average_unique_vertex(s, n, a, b)


def _cost_increase(possible, cost, fp, threads, length):
    inc = F(possible) * cost * (F(1) - fp)
    inc = inc / F(threads)
    return inc / (F(1) + F(length - 1) / F(2) * fp)


def effectiveness_cycle(st: dict, labels: list, threads: int = 1):
    """computeBacktrack of a cycle constraint with vertex labels `labels`
    (:226-340): (matchEliminated, costEliminated, effectiveness)."""
    with np.errstate(all="ignore"):
        s = _Stat(st)
        L = len(labels)
        finding = []
        for i in range(L):
            fp, unique = F(1), F(1)
            for j in range(L):
                a, b = labels[(i + j) % L], labels[(i + j + 1) % L]
                if j == 0 or labels[(i + j - 1) % L] != b:
                    fp = fp * proba_find_one_path(s, unique, a, b)
                    unique = average_unique_vertex(s, unique, a, b)
                else:
                    unique = _known_edge(s, unique, a, b)
            fp = fp * proba_find_same_vertex(s, unique, labels[(i + L) % L])
            finding.append(fp)
        match = F(0)
        for i in range(L):
            match = match + s.vertices(labels[i]) * (F(1) - finding[i])
        cost_total = F(0)
        for i in range(L):
            t = F(1)
            cost = t
            for j in range(L - 1):
                a, b = labels[(i + j) % L], labels[(i + j + 1) % L]
                cost = cost + average_vertex(s, t, a)
                if j == 0:
                    t = average_vertex_knowing(s, t, a, b)
                elif labels[(i + j - 1) % L] != b:
                    t = average_unique_vertex(s, t, a, b)
                else:
                    t = _known_edge(s, t, a, b)
            cost = cost + average_vertex(s, t, labels[(i + L - 1) % L])
            cost = cost + t
            cost_total = cost_total + _cost_increase(s.vertices(labels[i]), cost, finding[i], threads, L)
        return float(match), float(cost_total), float(match / cost_total)


def effectiveness_path(st: dict, labels: list, threads: int = 1):
    """computeBacktrack of a path constraint with vertex labels `labels`
    (:570-696)."""
    with np.errstate(all="ignore"):
        s = _Stat(st)
        L = len(labels)
        steps = max(L - 2, 0)
        finding = []
        for i in (0, 1):
            rev = i == 1
            fp, unique = F(1), F(1)
            for j in range(steps):
                cur, nxt = (L - j - 1, L - j - 2) if rev else (j, j + 1)
                a, b = labels[cur], labels[nxt]
                if j == 0 or labels[L - j if rev else j - 1] != b:
                    fp = fp * proba_find_one_path(s, unique, a, b)
                    unique = average_unique_vertex(s, unique, a, b)
                else:
                    unique = _known_edge(s, unique, a, b)
            fp = fp * proba_find_different_vertex(s, unique, labels[(i + L) % L])
            finding.append(fp)
        match = F(0)
        for i in (0, 1):
            match = match + s.vertices(labels[i]) * (F(1) - finding[i])
        cost_total = F(0)
        for i in (0, 1):
            rev = i == 1
            t = F(1)
            cost = t
            for j in range(steps):
                cur, nxt = (L - j - 1, L - j - 2) if rev else (j, j + 1)
                a, b = labels[cur], labels[nxt]
                cost = cost + average_vertex(s, t, a)
                if j == 0:
                    t = average_vertex_knowing(s, t, a, b)
                elif labels[L - j if rev else j - 1] != b:
                    t = average_unique_vertex(s, t, a, b)
                else:
                    t = _known_edge(s, t, a, b)
            cost = cost + average_vertex(s, t, labels[(i + L - 1) % L])
            cost = cost + t
            cost_total = cost_total + _cost_increase(s.vertices(labels[i]), cost, finding[i], threads, L)
        return float(match), float(cost_total), float(match / cost_total)


def effectiveness_tds(st: dict, walks: list, threads: int = 1):
    """computeBacktrack of a TDS constraint (:972-1118); `walks` in the
    constraint's walk order, each a list of (move, label, move_back); moves
    1..length-2 are read."""
    with np.errstate(all="ignore"):
        s = _Stat(st)
        finding = []
        for w in walks:
            hist = [lab for m, lab, _ in w if m in (MOVE_VERTEX_STORE, MOVE_VERTEX_NO_STORE)]
            fp, unique = F(1), F(1)
            a = w[0][1]
            seen = [unique]
            for j in range(max(len(w) - 2, 0)):
                m, lab, back = w[j + 1]
                if m in (MOVE_VERTEX_STORE, MOVE_VERTEX_NO_STORE):
                    fp = fp * proba_find_one_path(s, unique, a, lab)
                    unique = average_unique_vertex(s, unique, a, lab)
                    a = lab
                    seen.append(unique)
                elif m == MOVE_BACK:
                    unique = seen[back]
                    a = hist[back]
                elif m == MOVE_CHECK:
                    fp = fp * (proba_find_one_path(s, unique, a, lab) * proba_find_same_vertex(s, unique, lab))
            finding.append(fp)
        match = F(0)
        for w, fp in zip(walks, finding):
            match = match + s.vertices(w[0][1]) * (F(1) - fp)
        cost_total = F(0)
        for w, fp in zip(walks, finding):
            hist = [lab for m, lab, _ in w if m in (MOVE_VERTEX_STORE, MOVE_VERTEX_NO_STORE)]
            t = F(1)
            cost = t
            a = w[0][1]
            seen = [t]
            for j in range(max(len(w) - 2, 0)):
                m, lab, back = w[j + 1]
                if m in (MOVE_VERTEX_STORE, MOVE_VERTEX_NO_STORE):
                    cost = cost + average_vertex(s, t, a)
                    t = average_vertex_knowing(s, t, a, lab) if j == 0 else average_unique_vertex(s, t, a, lab)
                    seen.append(t)
                    a = lab
                elif m == MOVE_BACK:
                    t = seen[back]
                    a = hist[back]
                elif m == MOVE_CHECK:
                    cost = cost + average_vertex(s, t, a)
            cost = cost + t
            vertex_length = len(hist)
            cost_total = cost_total + _cost_increase(s.vertices(w[0][1]), cost, fp, threads, vertex_length)
        return float(match), float(cost_total), float(match / cost_total)


def choose_next(st: dict, cycles: list, paths: list, tds: list, threads: int = 1):
    """The dynamic order's choice (test_order_multiple_label_cpu.cc:184-274).
    cycles / paths: [(index, labels)], tds: [(index, walks)] of the
    constraints not yet run, in list order.  Returns (step, index,
    effectiveness)."""
    best = {}
    for kind, items, fn in ((STEP_CC, cycles, effectiveness_cycle), (STEP_PC, paths, effectiveness_path),
                            (STEP_TDS, tds, effectiveness_tds)):
        b, bid = 0.0, 0
        for k, (_, x) in enumerate(items):
            e = fn(st, x, threads)[2]
            if e > b:
                b, bid = e, k
        best[kind] = (b, bid)
    c, p, t = best[STEP_CC][0], best[STEP_PC][0], best[STEP_TDS][0]
    if c >= p and c >= t and cycles:
        return STEP_CC, cycles[best[STEP_CC][1]][0], c
    if p >= c and p >= t and paths:
        return STEP_PC, paths[best[STEP_PC][1]][0], p
    if t >= c and t >= p and tds:
        return STEP_TDS, tds[best[STEP_TDS][1]][0], t
    return STEP_ENUMERATION, 0, 0.0
