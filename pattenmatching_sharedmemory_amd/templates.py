"""Seeded random query templates for the batched and served forms.

A template is a random connected graph on k vertices: a random spanning tree
(vertex j > 0 attaches to a uniformly chosen earlier vertex) plus a few extra
edges (each closes one more cycle), with vertex labels drawn from the data
graph's label distribution (Zipf over the label ranks).

With distinct_labels the k labels are drawn without replacement.  Repeated
labels make the reference's cycle and path searches memoised and its
template-driven search exponential in the candidate density (on R-MAT scale
18 with Zipf(1.2) labels single queries then run for minutes), so the
throughput forms use distinct labels.
"""
from __future__ import annotations

import numpy as np


def random_templates(n: int, seed: int, min_vertices: int = 3, max_vertices: int = 10, max_extra_edges: int = 3,
                     labels: int = 32, zipf: float = 1.2,
                     distinct_labels: bool = False) -> list[tuple[int, list[tuple[int, int]], list[int]]]:
    """n templates as (vertex count, undirected edges, labels)."""
    rng = np.random.default_rng(seed)
    p = (np.arange(labels) + 1.0) ** -zipf
    p /= p.sum()
    out = []
    for _ in range(n):
        k = int(rng.integers(min_vertices, max_vertices + 1))
        edges = [(int(rng.integers(0, j)), j) for j in range(1, k)]
        have = {tuple(sorted(e)) for e in edges}
        extra = int(rng.integers(0, max_extra_edges + 1))
        free = [(a, b) for a in range(k) for b in range(a + 1, k) if (a, b) not in have]
        if free and extra:
            pick = rng.choice(len(free), size=min(extra, len(free)), replace=False)
            edges += [free[int(i)] for i in pick]
        labs = [int(x) for x in rng.choice(labels, size=k, p=p, replace=not distinct_labels)]
        out.append((k, edges, labs))
    return out


def longest_cycle(constraints) -> int:
    """Vertices of the longest cycle constraint (0 when there is none)."""
    return max([len(c) for c in constraints.cycles()] + [0])
