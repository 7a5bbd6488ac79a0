"""Label statistics on the device (pj_graph_stat) against the oracle's, and
the dynamic constraint order (PJH_ORDER_DYNAMIC) against its replay on the
oracle: the same statistics at every LCC fixed point, the same estimate bit
for bit, the same choice, and the same per-step results.

Statistics are compared on states after LCC0, where every active vertex
carries a template label (the device counts template labels only), at LCC
fixed points and after cycle steps; graphs with rows above 2048 slots
exercise the hub chunks, templates of 6 to 40 vertices every candidate word
width (8 to 64 bits)."""
import numpy as np
import pytest

from oracle import effectiveness as E
from parity import STEP_CC, STEP_LCC, STEP_LCC0, STEP_PC, STEP_TDS
from test_effectiveness import _acc_max, oracle_constraints

pytestmark = pytest.mark.gpu

STEP_ENUMERATION = 7
GRAPH_SEED, LABEL_SEED = 20221, 7
TRIANGLE_TAIL = [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5)]
TWO_CYCLES_CHORD = [(0, 1), (1, 2), (2, 3), (3, 0), (1, 4), (4, 5), (5, 2), (0, 2), (5, 6), (6, 7)]


def device_stat(ls) -> dict:
    """pj_label_stat as the oracle's dicts (label values, non-zero entries)."""
    K = len(ls.labels)
    st = {"vertex_total": int(ls.vertex_total), "edge_total": int(ls.edge_total), "vertex": {}, "edge": {},
          "outbound": {}, "at_least": {}}
    for f in range(K):
        a = ls.labels[f]
        if ls.vertex[f]:
            st["vertex"][a] = int(ls.vertex[f])
        if ls.outbound[f]:
            st["outbound"][a] = int(ls.outbound[f])
        for t in range(K):
            b = ls.labels[t]
            if ls.edge[f, t]:
                st["edge"].setdefault(a, {})[b] = int(ls.edge[f, t])
            if ls.at_least_max[f, t]:
                for k, n in enumerate(ls.at_least[(f, t)], start=1):
                    if n:
                        st["at_least"].setdefault(a, {}).setdefault(b, {})[k] = int(n)
    return st


def oracle_stat(g, c, cand, vflags, ea):
    live = np.where(vflags & 1, cand, 0).astype(np.uint64)
    return E.graph_stat(g.offsets(), g.edges(), g.labels(), live, ea, _acc_max(c))


def _rmat(pj, scale, labels, zipf):
    return pj.Graph.rmat(scale, 16, 0.57, 0.19, 0.19, seed=GRAPH_SEED).labels_zipf(labels, zipf, seed=LABEL_SEED)


def _check_stat(pj, g, c, pr, label):
    st = pr.state()
    want = oracle_stat(g, c, st.cand, st.vflags, st.ea_words)
    got = device_stat(pr.graph_stat())
    assert got == want, label
    return want


def _hub_active(g, st) -> bool:
    deg = np.diff(g.offsets().astype(np.int64))
    return bool(np.any((deg > 2048) & (st.vflags & 1).astype(bool)))


@pytest.mark.parametrize("scale,labels,zipf,k,edges,plabels", [
    (12, 8, 0.0, 8, TWO_CYCLES_CHORD, [0, 1, 0, 2, 1, 3, 0, 1]),
    (16, 32, 1.2, 6, TRIANGLE_TAIL, [0, 1, 2, 3, 0, 1]),
    (17, 16, 1.0, 8, TWO_CYCLES_CHORD, [0, 1, 2, 0, 1, 2, 3, 0]),
])
def test_label_stat_matches_oracle(pj, orc, scale, labels, zipf, k, edges, plabels):
    """After LCC0, at every LCC step to the fixed point and after each cycle
    step: vertex, label-pair and at-least counts equal the oracle's."""
    g = _rmat(pj, scale, labels, zipf) if zipf else pj.Graph.rmat(scale, 16, 0.57, 0.19, 0.19,
                                                                  seed=GRAPH_SEED).labels_uniform(labels, LABEL_SEED)
    c = pj.Constraints(pj.Graph.pattern(k, edges, plabels))
    pr = pj.Pruner(g, c)
    label = f"s{scale} k{k}"
    print(f"[label stat {label}] V={g.vertex_count} E={g.edge_count}", flush=True)
    pr.step(STEP_LCC0)
    st0 = _check_stat(pj, g, c, pr, label + " after LCC0")
    assert any(m > 1 for d in _acc_max(c).values() for m in d.values()) or k == 6
    if scale >= 16:
        assert _hub_active(g, pr.state()), "no active row above 2048 slots"
    assert st0["vertex_total"] > 0
    while not pr.step(STEP_LCC).is_empty():
        _check_stat(pj, g, c, pr, label + " LCC")
    _check_stat(pj, g, c, pr, label + " LCC fixed point")
    for i in range(len(c.cycles())):
        pr.reset_modified()
        pr.step(STEP_CC, i)
        _check_stat(pj, g, c, pr, label + f" after CC {i}")
    pr.close()


@pytest.mark.parametrize("k,seed", [(6, 1), (12, 2), (24, 3), (40, 4)])
def test_label_stat_word_widths(pj, orc, k, seed):
    """Templates of 6, 12, 24 and 40 vertices (8-, 16-, 32- and 64-bit
    candidate words; up to 40 label codes) on R-MAT scale 13 with 48 labels."""
    from pattenmatching_sharedmemory_amd.templates import random_templates

    g = _rmat(pj, 13, 48, 0.6)
    (kk, e, l), = random_templates(1, seed, k, k, 6, labels=48, zipf=0.6)
    c = pj.Constraints(pj.Graph.pattern(kk, e, l))
    pr = pj.Pruner(g, c)
    pr.step(STEP_LCC0)
    _check_stat(pj, g, c, pr, f"k{k} after LCC0")
    pr.step(STEP_LCC)
    _check_stat(pj, g, c, pr, f"k{k} after LCC")
    pr.close()


def replay_dynamic(orc, g, c, mode, threads=8):
    """The dynamic order run on the oracle: per-step keys, the choices
    (step, index, effectiveness) made at the LCC fixed points, and the
    iteration of the first memoised step whose result depends on the
    reference's schedule (None: every step determined; see parity.py)."""
    cycles, paths, tds = oracle_constraints(c)
    left = {STEP_CC: list(range(len(cycles))), STEP_PC: list(range(len(paths))), STEP_TDS: list(range(len(tds)))}
    o = orc.Oracle(g, c, threads=threads)
    keys, choices, undetermined = [], [], None
    step, idx, it = STEP_LCC0, 0, 0
    while step != STEP_ENUMERATION:
        if step in (STEP_CC, STEP_PC, STEP_TDS):
            o.reset_modified()
            if mode == orc.FAITHFUL and undetermined is None:
                before = o.state()
                live = np.where(before[1] & 1, before[0], 0).astype(np.uint64)
                x = orc.Oracle(g, c, threads=threads)
                x.load_state(*before)
                x.step(step, idx, orc.FAITHFUL_NO_SHORTCUT)
                if np.any(live & ~x.state()[0] & x.last_marks()):
                    undetermined = it
        r = o.step(step, idx, mode)
        cand, vf, ea, av, ae = o.state()
        keys.append((it, step) + tuple(r) + (av, ae))
        if step == STEP_LCC and not any(r):
            stat = oracle_stat(g, c, cand, vf, ea)
            ch = E.choose_next(stat, [(i, cycles[i]) for i in left[STEP_CC]], [(i, paths[i]) for i in left[STEP_PC]],
                               [(i, tds[i]) for i in left[STEP_TDS]])
            choices.append((it, ch))
            step, idx = ch[0], ch[1]
            if step != STEP_ENUMERATION:
                left[step].remove(idx)
        else:
            step = STEP_LCC
        it += 1
    return keys, choices, undetermined


@pytest.mark.parametrize("scale,k,edges,plabels,labels", [
    (16, 6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3], 32),
    (18, 6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3], 32),
    (15, 6, [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5), (5, 2)], [0, 1, 2, 3, 4, 5], 16),
])
def test_dynamic_order_matches_oracle_replay(pj, orc, scale, k, edges, plabels, labels):
    """The whole dynamic order on the device against its replay on the
    oracle (complete search; every cycle and path constraint here is
    memo-inert, where the reference's memoised search decides the same):
    per-step results, the chosen constraints and their estimates."""
    g = _rmat(pj, scale, labels, 1.2)
    c = pj.Constraints(pj.Graph.pattern(k, edges, plabels))
    assert all(c.memo_inert(STEP_CC, i) == 1 for i in range(len(c.cycles())))
    assert all(c.memo_inert(STEP_PC, i) == 1 for i in range(len(c.paths())))
    assert not c.tds()
    pr = pj.Pruner(g, c)
    rows, _ = pr.run(enumerate_matches=False, order="dynamic")
    pr.close()
    keys, choices, _ = replay_dynamic(orc, g, c, orc.COMPLETE)
    print(f"[dynamic order s{scale} k{k}] {len(rows)} steps, choices {[ch[:2] for _, ch in choices]}", flush=True)
    assert [r.key() for r in rows] == keys
    made = [(r.step, r.constraint, r.effectiveness) for r in rows if r.step in (STEP_CC, STEP_PC, STEP_TDS)]
    assert made == [ch for _, ch in choices if ch[0] != STEP_ENUMERATION]
    assert len(made) == len(c.cycles()) + len(c.paths())
    assert sum(r.stat_ms > 0 for r in rows) == len(choices)


def test_dynamic_order_with_tds_against_oracle(pj, orc):
    """The 8-vertex two-cycle + chord template (cycles, paths and a TDS
    constraint) on R-MAT scale 14 with 64 Zipf(1.0) labels: the dynamic
    order on the device against its replay on the oracle in the reference's
    memoised mode, step for step up to the first memoised step whose result
    depends on the reference's schedule (the whole run when there is none),
    and every choice made before it."""
    g = _rmat(pj, 14, 64, 1.0)
    c = pj.Constraints(pj.Graph.pattern(8, TWO_CYCLES_CHORD, list(range(8, 16))))
    assert len(c.tds()) == 1
    pr = pj.Pruner(g, c)
    rows, _ = pr.run(enumerate_matches=False, order="dynamic")
    pr.close()
    keys, choices, undetermined = replay_dynamic(orc, g, c, orc.FAITHFUL)
    n = len(keys) if undetermined is None else undetermined
    print(f"[dynamic order TDS] {len(rows)} steps, choices {[ch[:2] for _, ch in choices]}, "
          f"first undetermined step {undetermined}", flush=True)
    assert [r.key() for r in rows][:n] == keys[:n]
    made = {r.iteration: (r.step, r.constraint, r.effectiveness) for r in rows}
    for it, ch in choices:
        if it + 1 < n and ch[0] != STEP_ENUMERATION:
            assert made[it + 1] == ch, it
    if undetermined is None:
        assert len(rows) == len(keys)
        assert STEP_TDS in [r.step for r in rows]
