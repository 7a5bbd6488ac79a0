"""Parity of the HIP pipeline with the oracle's faithful mode (the
reference's loops, memos included) on the same seeded inputs: per-step
counter trace, final candidate sets, vertex flags and the raw edge-active
bitset, all bit-exact."""
import numpy as np
import pytest

from conftest import FIXTURES, fixture_paths

pytestmark = pytest.mark.gpu


def _fixture(pj, name):
    f = fixture_paths(name)
    g = pj.Graph.load(f["graph"], f["vertex"])
    p = pj.Graph.load(f["pattern"], f["pattern_vertex"])
    return g, p


def _compare_pipeline(pj, orc, g, p, label, enumerate_matches=True):
    c = pj.Constraints(p)
    pr = pj.Pruner(g, c)
    rows, _ = pr.run(enumerate_matches=enumerate_matches)
    st = pr.state()
    o = orc.Oracle(g, c, threads=8)
    orows, _ = o.run(orc.FAITHFUL, enumerate_matches=enumerate_matches)
    cand, vf, ea, av, ae = o.state()
    print(f"[{label}] V={g.vertex_count} E={g.edge_count} steps={len(rows)}", flush=True)
    assert [r.key() for r in rows] == [r.key() for r in orows], label
    assert np.array_equal(st.cand, cand), label
    assert np.array_equal(st.vflags, vf), label
    assert np.array_equal(st.ea_words, ea), label
    assert (st.active_vertices, st.active_edges) == (av, ae)
    pr.close()
    return rows


@pytest.mark.parametrize("name", FIXTURES)
def test_reference_fixtures(pj, orc, name):
    g, p = _fixture(pj, name)
    rows = _compare_pipeline(pj, orc, g, p, name)
    expected_enum = {"test_1": 1, "test_2": 2, "test_3": 2, "test_4": 1}[name]
    assert rows[-1].enumeration == expected_enum


def test_config1_scale14_four_cycle(pj, orc):
    g = pj.Graph.rmat(14, 16, seed=1).labels_uniform(8, seed=2)
    p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 3), (3, 0)], [0, 1, 2, 3])
    _compare_pipeline(pj, orc, g, p, "C1 s14")


TRIANGLE_TAIL = [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5)]


def test_triangle_tail_zipf_scale16_with_enumeration(pj, orc):
    g = pj.Graph.rmat(16, 16, seed=20221).labels_zipf(32, 1.2, seed=7)
    p = pj.Graph.pattern(6, TRIANGLE_TAIL, [7, 8, 9, 10, 11, 10])
    _compare_pipeline(pj, orc, g, p, "C2-shape s16 enum")


def test_triangle_tail_frequent_labels_pruning_scale16(pj, orc):
    # the bench template (labels of Zipf ranks 0..4): pruning fixed point only
    g = pj.Graph.rmat(16, 16, seed=20221).labels_zipf(32, 1.2, seed=7)
    p = pj.Graph.pattern(6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3])
    _compare_pipeline(pj, orc, g, p, "C2-shape s16 prune", enumerate_matches=False)


TDS_EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (1, 4), (4, 5), (5, 2), (0, 2), (5, 6), (6, 7)]


# Graphs on which the reference's TDS result does not depend on thread
# scheduling (the oracle's faithful mode gives one outcome at 1, 2, 3, 4, 8,
# 16 and 32 threads).
@pytest.mark.parametrize("scale,seed", [(10, 5), (12, 7), (11, 3), (11, 4), (10, 9)])
def test_tds_template(pj, orc, scale, seed):
    # two 4-cycles sharing the edge 1-2 plus the chord 0-2: merged TDS templates,
    # searched with the reference's hypothesis and edge memos
    g = pj.Graph.rmat(scale, 8, seed=seed).labels_zipf(8, 1.0, seed=seed + 1)
    p = pj.Graph.pattern(8, TDS_EDGES, [0, 1, 2, 3, 4, 5, 6, 7])
    _compare_pipeline(pj, orc, g, p, f"TDS s{scale} seed {seed}", enumerate_matches=scale <= 10)


@pytest.mark.parametrize("n", [12, 20, 36])
def test_template_sizes_for_each_word_width(pj, orc, n):
    # candidate words are 16-bit (P=12), 32-bit (P=20) and 64-bit (P=36) on the device
    edges = [(i, i + 1) for i in range(n - 1)] + [(0, 2)]
    g = pj.Graph.rmat(13, 16, seed=9).labels_uniform(40, seed=10)
    p = pj.Graph.pattern(n, edges, [i % 40 for i in range(n)])
    _compare_pipeline(pj, orc, g, p, f"P={n}", enumerate_matches=False)


def test_hub_rows(pj, orc):
    # a star of 20000 leaves over a sparse random graph: hub rows take the chunked path
    rng = np.random.default_rng(12)
    V = 30000
    pairs = {(0, v) for v in range(1, 20001)}
    for _ in range(60000):
        a, b = (int(x) for x in rng.integers(1, V, 2))
        if a != b:
            pairs.add((min(a, b), max(a, b)))
    adj = [[] for _ in range(V)]
    for a, b in pairs:
        adj[a].append(b)
        adj[b].append(a)
    off = np.zeros(V + 1, dtype=np.uint64)
    off[1:] = np.cumsum([len(r) for r in adj])
    ed = np.concatenate([np.sort(np.array(r, dtype=np.uint32)) for r in adj if r])
    lab = rng.integers(0, 4, V).astype(np.uint64)
    g = pj.Graph.from_csr(off, ed, lab)
    p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 0), (2, 3)], [0, 1, 2, 3])
    _compare_pipeline(pj, orc, g, p, "hub")


def _random_state(rng, V, E, P):
    cand = rng.integers(0, 1 << min(P, 62), V, dtype=np.uint64)
    cand[rng.random(V) < 0.2] = 0
    vf = (cand != 0).astype(np.uint8)
    vf |= (rng.random(V) < 0.3).astype(np.uint8) << 1
    vf |= (rng.random(V) < 0.6).astype(np.uint8) << 2
    ea = rng.integers(0, 1 << 32, (E + 31) // 32, dtype=np.uint64).astype(np.uint32)
    if E % 32:
        ea[-1] &= np.uint32((1 << (E % 32)) - 1)  # no bits past the last slot
    return cand, vf, ea


@pytest.mark.parametrize("seed", [21, 22])
def test_single_steps_from_random_states(pj, orc, seed):
    from pattenmatching_sharedmemory_amd import STEP_CC, STEP_LCC, STEP_LCC0, STEP_PC, STEP_TDS

    rng = np.random.default_rng(seed)
    g = pj.Graph.rmat(12, 16, seed=seed).labels_uniform(6, seed=seed + 1)
    edges = [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5), (5, 2), (1, 6)]
    p = pj.Graph.pattern(7, edges, [0, 1, 2, 3, 4, 3, 1])
    c = pj.Constraints(p)
    pr = pj.Pruner(g, c)
    o = orc.Oracle(g, c, threads=4)
    V, E = g.vertex_count, g.edge_count
    # (enumeration from a random candidate state is combinatorially huge; it
    # is covered by the pipeline cases above)
    steps = [(STEP_LCC0, 0), (STEP_LCC, 0)]
    steps += [(STEP_CC, i) for i in range(len(c.cycles()))]
    steps += [(STEP_PC, i) for i in range(len(c.paths()))]
    steps += [(STEP_TDS, i) for i in range(len(c.tds()))]
    for step, idx in steps:
        cand, vf, ea = _random_state(rng, V, E, 7)
        av = int((vf & 1).sum())
        pr.load_state(pj.State(cand, vf, ea, av, E))
        o.load_state(cand, vf, ea, av, E)
        r = pr.step(step, idx).as_tuple()
        label = f"step {step} index {idx} seed {seed}"
        print(label, r, flush=True)
        st = pr.state()
        if step in (STEP_CC, STEP_PC, STEP_TDS):
            # the alreadyMatched shortcut can make the reference's result
            # depend on scheduling: check the bounds every schedule obeys,
            # which pin the result exactly when no failing root is marked
            o.step(step, idx, orc.FAITHFUL_NO_SHORTCUT)
            f0 = np.where(vf & 1, cand, 0) & ~o.state()[0]
            m0 = o.last_marks()
            f = np.where(vf & 1, cand, 0) & ~st.cand
            assert np.all((f & ~f0) == 0) and np.all((f0 & ~m0 & ~f) == 0), label
            if np.any(f0 & m0):
                continue
        ro = o.step(step, idx, orc.FAITHFUL) if step not in (STEP_CC, STEP_PC, STEP_TDS) else None
        if ro is None:  # exact case: compare with the faithful run from the same state
            o.load_state(cand, vf, ea, av, E)
            ro = o.step(step, idx, orc.FAITHFUL)
        oc, ovf, oea, oav, oae = o.state()
        assert r == ro, label
        assert np.array_equal(st.cand, oc), label
        assert np.array_equal(st.vflags, ovf), label
        assert np.array_equal(st.ea_words, oea), label
    pr.close()


@pytest.mark.parametrize("case", ["C2-shape s16 prune", "TDS s10", "hub"])
def test_clean_vertex_skipping_changes_nothing(pj, orc, case, monkeypatch):
    """Skipping clean scheduled vertices in LCC phase 1 leaves the step trace
    (including the reference-order slot counts), the candidate sets, flags and
    edge-active bits identical to evaluating every scheduled vertex."""
    if case == "TDS s10":
        edges = [(0, 1), (1, 2), (2, 3), (3, 0), (1, 4), (4, 5), (5, 2), (0, 2), (5, 6), (6, 7)]
        g = pj.Graph.rmat(10, 8, seed=5).labels_zipf(8, 1.0, seed=6)
        p = pj.Graph.pattern(8, edges, [0, 1, 2, 3, 4, 5, 6, 7])
    elif case == "hub":
        g = pj.Graph.rmat(14, 32, seed=3).labels_uniform(4, seed=4)  # R-MAT rows above 2048 slots
        p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 0), (2, 3)], [0, 1, 2, 3])
    else:
        g = pj.Graph.rmat(16, 16, seed=20221).labels_zipf(32, 1.2, seed=7)
        p = pj.Graph.pattern(6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3])
    c = pj.Constraints(p)
    runs = []
    for skip in ("0", "1"):
        monkeypatch.setenv("PJ_SKIP_CLEAN", skip)
        pr = pj.Pruner(g, c)
        pr.kernel_timing(True)
        rows, _ = pr.run(enumerate_matches=False)
        examined = sum(k["slots"] for k in pr.kernel_stats())
        runs.append((rows, pr.state(), examined))
        pr.close()
    (r0, s0, x0), (r1, s1, x1) = runs
    print(f"[{case}] steps={len(r0)} slots examined: every vertex {x0}, skipping unchanged ones {x1}", flush=True)
    assert [r.key() for r in r0] == [r.key() for r in r1]
    # slot counts of the LCC steps are reference-order counts and must agree;
    # memoised walks may skip a source another wave has already matched, so
    # their counts depend on timing
    lcc = [i for i, r in enumerate(r0) if r.step == 3]
    assert [r0[i].slots for i in lcc] == [r1[i].slots for i in lcc]
    assert np.array_equal(s0.cand, s1.cand) and np.array_equal(s0.vflags, s1.vflags)
    assert np.array_equal(s0.ea_words, s1.ea_words)
    o = orc.Oracle(g, c, threads=8)
    orows, _ = o.run(orc.FAITHFUL, enumerate_matches=False)
    assert [r.key() for r in r1] == [r.key() for r in orows]


@pytest.mark.parametrize("case", ["plain s16", "hub", "many labels"])
def test_first_lcc_step_from_slot_labels_changes_nothing(pj, orc, case, monkeypatch):
    """The first LCC step after LCC0 reads each slot's neighbour label
    (built at upload for graphs of at most 255 labels) instead of the
    neighbour's candidates; the step trace, candidate sets, flags and
    edge-active bits equal those with the gathers and the oracle's.  A graph
    of more than 255 labels has no slot labels and takes the gathers."""
    if case == "hub":
        g = pj.Graph.rmat(14, 32, seed=3).labels_uniform(4, seed=4)  # R-MAT rows above 2048 slots
        p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 0), (2, 3)], [0, 1, 2, 3])
    elif case == "many labels":
        g = pj.Graph.rmat(13, 16, seed=9).labels_uniform(300, seed=2)
        p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 3), (3, 0)], [0, 1, 2, 1])
    else:
        g = pj.Graph.rmat(16, 16, seed=20221).labels_zipf(32, 1.2, seed=7)
        p = pj.Graph.pattern(6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3])
    c = pj.Constraints(p)
    runs = []
    for on in ("0", "1"):
        monkeypatch.setenv("PJ_SLOT_LABELS", on)
        pr = pj.Pruner(g, c)
        rows, _ = pr.run(enumerate_matches=False)
        runs.append((rows, pr.state()))
        rows2, _ = pr.run(enumerate_matches=False)  # a second solve on the same context
        assert [r.key() for r in rows2] == [r.key() for r in rows]
        pr.close()
    (r0, s0), (r1, s1) = runs
    print(f"[{case}] steps={len(r0)}", flush=True)
    assert [r.key() for r in r0] == [r.key() for r in r1]
    assert [r.slots for r in r0 if r.step == 3] == [r.slots for r in r1 if r.step == 3]
    assert np.array_equal(s0.cand, s1.cand) and np.array_equal(s0.vflags, s1.vflags)
    assert np.array_equal(s0.ea_words, s1.ea_words)
    o = orc.Oracle(g, c, threads=8)
    orows, _ = o.run(orc.FAITHFUL, enumerate_matches=False)
    assert [r.key() for r in r1] == [r.key() for r in orows]


# Templates on which the reference's memos change the result (the oracle's
# faithful and complete modes differ; the faithful one is the same at 1 and 8
# threads): repeated labels on a cycle, which also yield paths and TDS steps.
MEMO_TEMPLATES = {
    "cycle 0-1-0-1": (4, [(0, 1), (1, 2), (2, 3), (3, 0)], [0, 1, 0, 1]),
    "cycle 0-1-0-1-2": (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)], [0, 1, 0, 1, 2]),
    "cycle 0-1-0-2": (4, [(0, 1), (1, 2), (2, 3), (3, 0)], [0, 1, 0, 2]),
}


@pytest.mark.parametrize("name,seed", [(n, s) for n in MEMO_TEMPLATES for s in (1, 2, 3)
                                       if (n, s) != ("cycle 0-1-0-1-2", 3)])  # that one: see below
def test_memoised_walks_follow_the_reference(pj, orc, name, seed):
    n, edges, labels = MEMO_TEMPLATES[name]
    g = pj.Graph.rmat(11, 8, seed=seed).labels_uniform(3, seed=seed + 5)
    p = pj.Graph.pattern(n, edges, labels)
    c = pj.Constraints(p)
    complete, _ = orc.Oracle(g, c, threads=8).run(orc.COMPLETE, enumerate_matches=False)
    rows = _compare_pipeline(pj, orc, g, p, f"{name} seed {seed}", enumerate_matches=False)
    assert [r.key() for r in rows] != [r.key() for r in complete]  # the memos matter here


@pytest.mark.parametrize("case", ["C1 s14", "C2-shape s16 prune"])
def test_memoised_cycle_and_path_walks_on_memo_inert_templates(pj, orc, case, monkeypatch):
    """PJ_MEMO_WALKS=1 runs every cycle and path step as the memoised search;
    where the memo cannot fire it must agree with the complete search."""
    monkeypatch.setenv("PJ_MEMO_WALKS", "1")
    if case == "C1 s14":
        g = pj.Graph.rmat(14, 16, seed=1).labels_uniform(8, seed=2)
        p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 3), (3, 0)], [0, 1, 2, 3])
        _compare_pipeline(pj, orc, g, p, case)
    else:
        g = pj.Graph.rmat(16, 16, seed=20221).labels_zipf(32, 1.2, seed=7)
        p = pj.Graph.pattern(6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3])
        _compare_pipeline(pj, orc, g, p, case, enumerate_matches=False)


def _nlcc_steps(pj, orc, g, c):
    """Runs the step machine on the oracle (faithful, 8 threads) and yields,
    before every cycle / path / TDS step, (step, index, state) with the
    modified flags already reset."""
    from pattenmatching_sharedmemory_amd import STEP_CC, STEP_LCC, STEP_LCC0, STEP_PC, STEP_TDS

    o = orc.Oracle(g, c, threads=8)

    def lcc_fixed_point():
        while any(o.step(STEP_LCC, 0, orc.FAITHFUL)):
            pass

    o.step(STEP_LCC0, 0, orc.FAITHFUL)
    lcc_fixed_point()
    plan = [(STEP_CC, i) for i in range(len(c.cycles()))] + [(STEP_PC, i) for i in range(len(c.paths()))]
    plan += [(STEP_TDS, 0)] * len(c.tds())  # the reference never advances the TDS iterator
    for step, idx in plan:
        o.reset_modified()
        yield step, idx, o.state()
        o.step(step, idx, orc.FAITHFUL)
        lcc_fixed_point()


# Graphs on which the reference's result depends on thread scheduling: a
# source skips a walk whose root another source's successful walk has marked
# (alreadyMatched).  Every schedule's failures F satisfy F0 - M0 <= F <= F0,
# F0 and M0 being the failures and marks with every walk run; the device must
# land in that range for every memoised step, from the reference's states.
@pytest.mark.parametrize("case", ["TDS s12 seed 5", "TDS s13 seed 5", "cycle 0-1-0-1-2 seed 3"])
def test_schedule_dependent_steps_stay_within_the_reference_outcomes(pj, orc, case):
    if case.startswith("TDS"):
        scale, seed = (12, 5) if "s12" in case else (13, 5)
        g = pj.Graph.rmat(scale, 8, seed=seed).labels_zipf(8, 1.0, seed=seed + 1)
        p = pj.Graph.pattern(8, TDS_EDGES, [0, 1, 2, 3, 4, 5, 6, 7])
    else:
        n, edges, labels = MEMO_TEMPLATES["cycle 0-1-0-1-2"]
        g = pj.Graph.rmat(11, 8, seed=3).labels_uniform(3, seed=8)
        p = pj.Graph.pattern(n, edges, labels)
    c = pj.Constraints(p)
    pr = pj.Pruner(g, c)
    checked = 0
    for step, idx, (cand, vf, ea, av, ae) in _nlcc_steps(pj, orc, g, c):
        pr.load_state(pj.State(cand, vf, ea, av, ae))
        pr.step(step, idx)
        dev = pr.state().cand
        o = orc.Oracle(g, c, threads=8)
        o.load_state(cand, vf, ea, av, ae)
        o.step(step, idx, orc.FAITHFUL_NO_SHORTCUT)
        f0 = cand & ~o.state()[0]
        m0 = o.last_marks()
        f = cand & ~dev
        print(f"[{case}] step {step} index {idx}: failures {int(np.count_nonzero(f))}, "
              f"bounds {int(np.count_nonzero(f0 & ~m0))}..{int(np.count_nonzero(f0))}", flush=True)
        assert np.all((f & ~f0) == 0), (case, step, idx)
        assert np.all((f0 & ~m0 & ~f) == 0), (case, step, idx)
        checked += 1
    assert checked
    pr.close()


@pytest.mark.parametrize("scale,seed", [(12, 7), (11, 4)])
def test_memo_retry_passes(pj, orc, scale, seed, monkeypatch):
    """With 16 live entries in the LDS tables and 64-entry HBM tables most TDS
    sources outgrow both and are rerun by the retry passes; the results stay
    the reference's."""
    monkeypatch.setenv("PJ_MEMO_LDS_LIMIT", "16")
    monkeypatch.setenv("PJ_MEMO_LOG2", "6")
    g = pj.Graph.rmat(scale, 8, seed=seed).labels_zipf(8, 1.0, seed=seed + 1)
    p = pj.Graph.pattern(8, TDS_EDGES, [0, 1, 2, 3, 4, 5, 6, 7])
    _compare_pipeline(pj, orc, g, p, f"TDS retry s{scale} seed {seed}", enumerate_matches=False)


@pytest.mark.parametrize("case", ["path of 5", "cycle of 6", "cycle of 5 + chord"])
def test_long_walks_with_failed_state_memo(pj, orc, case):
    """Complete cycle walks of 5+ vertices and path walks of 4+ vertices skip
    (vertex, hop) states already known to fail; the results are unchanged."""
    g = pj.Graph.rmat(14, 16, seed=11).labels_uniform(6, seed=12)
    if case == "path of 5":
        p = pj.Graph.pattern(5, [(0, 1), (1, 2), (2, 3), (3, 4)], [5, 1, 2, 3, 5])
    elif case == "cycle of 6":
        p = pj.Graph.pattern(6, [(i, (i + 1) % 6) for i in range(6)], [0, 1, 2, 3, 4, 5])
    else:
        p = pj.Graph.pattern(5, [(i, (i + 1) % 5) for i in range(5)] + [(0, 2)], [0, 1, 2, 3, 4])
    c = pj.Constraints(p)
    assert max([len(x) for x in c.cycles()] + [len(x) for x in c.paths()]) >= 5
    _compare_pipeline(pj, orc, g, p, case, enumerate_matches=False)


# Triangle 2-3-4 whose memo can block walks (template id 2 is the position
# tested at the second hop of start 1; vertices 4 and 2 share a label).
MEMO_TRIANGLE = (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 2)], [0, 3, 1, 2, 1])


@pytest.mark.parametrize("kernel", ["triangle", "walk"])
@pytest.mark.parametrize("seed,perturb", [(5, False), (6, True), (7, True)])
def test_memoised_triangle_steps_stay_within_the_reference_outcomes(pj, orc, seed, perturb, kernel, monkeypatch):
    """The memoised 3-vertex cycle step (k_triangles_memo, or the memoised
    walk with PJ_TRIANGLE_MEMO=0) from the reference's state, and from states
    with a fifth of the edge-active bits cleared, lands within the outcomes of
    the reference's schedules: F0 - M0 <= F <= F0 (failures F0 and marks M0
    with every walk run)."""
    from pattenmatching_sharedmemory_amd import STEP_CC, STEP_LCC, STEP_LCC0

    if kernel == "walk":
        monkeypatch.setenv("PJ_TRIANGLE_MEMO", "0")
    n, edges, labels = MEMO_TRIANGLE
    g = pj.Graph.rmat(12, 8, seed=seed).labels_uniform(4, seed=seed + 10)
    c = pj.Constraints(pj.Graph.pattern(n, edges, labels))
    k = [list(x) for x in c.cycles()].index([2, 3, 4])
    o = orc.Oracle(g, c, threads=8)
    o.step(STEP_LCC0, 0, orc.FAITHFUL)
    while any(o.step(STEP_LCC, 0, orc.FAITHFUL)):
        pass
    o.reset_modified()
    cand, vf, ea, av, ae = o.state()
    if perturb:
        rng = np.random.default_rng(seed)
        bits = np.unpackbits(ea.view(np.uint8), bitorder="little")
        bits &= (rng.random(bits.size) >= 0.2).astype(np.uint8)
        ea = np.packbits(bits, bitorder="little").view(np.uint32)[: ea.size].copy()
    cand = np.where(vf & 1, cand, 0).astype(np.uint64)
    x = orc.Oracle(g, c, threads=8)
    x.load_state(cand, vf, ea, av, ae)
    x.step(STEP_CC, k, orc.FAITHFUL_NO_SHORTCUT)
    f0 = cand & ~x.state()[0]
    m0 = x.last_marks()
    pr = pj.Pruner(g, c)
    pr.load_state(pj.State(cand, vf, ea, av, ae))
    pr.step(STEP_CC, k)
    f = cand & ~pr.state().cand
    print(f"[{kernel} seed {seed}] failures {int(np.count_nonzero(f))}, "
          f"bounds {int(np.count_nonzero(f0 & ~m0))}..{int(np.count_nonzero(f0))}", flush=True)
    assert np.any(f0)
    assert np.all((f & ~f0) == 0)
    assert np.all((f0 & ~m0 & ~f) == 0)
    pr.close()


@pytest.mark.parametrize("case", ["C2-shape s16", "hub"])
def test_reverse_positions_change_nothing(pj, orc, case, monkeypatch):
    """Cycle steps closing through the reverse slot positions and through
    binary searches (PJ_REVERSE_SLOTS=0) give identical traces and states,
    and the reference's."""
    if case == "hub":
        g = pj.Graph.rmat(14, 32, seed=3).labels_uniform(4, seed=4)
        p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 0), (2, 3)], [0, 1, 2, 3])
    else:
        g = pj.Graph.rmat(16, 16, seed=20221).labels_zipf(32, 1.2, seed=7)
        p = pj.Graph.pattern(6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3])
    c = pj.Constraints(p)
    runs = []
    for rev in ("0", "1"):
        monkeypatch.setenv("PJ_REVERSE_SLOTS", rev)
        pr = pj.Pruner(g, c)
        rows, _ = pr.run(enumerate_matches=False)
        runs.append((rows, pr.state()))
        pr.close()
    (r0, s0), (r1, s1) = runs
    assert [r.key() for r in r0] == [r.key() for r in r1]
    assert np.array_equal(s0.cand, s1.cand) and np.array_equal(s0.vflags, s1.vflags)
    assert np.array_equal(s0.ea_words, s1.ea_words)
    o = orc.Oracle(g, c, threads=8)
    orows, _ = o.run(orc.FAITHFUL, enumerate_matches=False)
    assert [r.key() for r in r1] == [r.key() for r in orows]


@pytest.mark.parametrize("case", ["C2-shape s16", "hub", "four-cycle", "bound by the budget"])
def test_pair_index_changes_nothing(pj, orc, case, monkeypatch):
    """Cycle steps with the pair index of heavy rows (rows above 16 or 32
    slots, so that most searches between a source's neighbours read it; with
    a 1 MB bound the threshold doubles from 4 slots until the index fits) and
    with binary searches only (PJ_PAIR_INDEX_MB=0) give identical traces and
    states, and the reference's; the index replaces searches, so the cycle
    steps read fewer slots with it."""
    from pattenmatching_sharedmemory_amd import STEP_CC

    budget_on = "1024"
    if case == "bound by the budget":
        g = pj.Graph.rmat(14, 32, seed=3).labels_uniform(4, seed=4)
        p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 0), (2, 3)], [0, 1, 2, 3])
        degree, budget_on = "4", "1"
    elif case == "hub":
        g = pj.Graph.rmat(14, 32, seed=3).labels_uniform(4, seed=4)
        p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 0), (2, 3)], [0, 1, 2, 3])
        degree = "16"
    elif case == "four-cycle":
        g = pj.Graph.rmat(14, 16, seed=1).labels_uniform(8, seed=2)
        p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 3), (3, 0)], [0, 1, 2, 3])
        degree = "16"
    else:
        g = pj.Graph.rmat(16, 16, seed=20221).labels_zipf(32, 1.2, seed=7)
        p = pj.Graph.pattern(6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3])
        degree = "32"
    c = pj.Constraints(p)
    monkeypatch.setenv("PJ_PAIR_DEGREE", degree)
    runs = []
    for budget in ("0", budget_on):
        monkeypatch.setenv("PJ_PAIR_INDEX_MB", budget)
        pr = pj.Pruner(g, c)
        rows, _ = pr.run(enumerate_matches=False)
        runs.append((rows, pr.state()))
        pr.close()
    (r0, s0), (r1, s1) = runs
    assert [r.key() for r in r0] == [r.key() for r in r1], case
    assert np.array_equal(s0.cand, s1.cand) and np.array_equal(s0.vflags, s1.vflags)
    assert np.array_equal(s0.ea_words, s1.ea_words)
    o = orc.Oracle(g, c, threads=8)
    orows, _ = o.run(orc.FAITHFUL, enumerate_matches=False)
    assert [r.key() for r in r1] == [r.key() for r in orows], case
    cc0 = sum(r.slots for r in r0 if r.step == STEP_CC)
    cc1 = sum(r.slots for r in r1 if r.step == STEP_CC)
    print(f"[{case}] cycle-step slots: searches {cc0}, with the pair index {cc1}", flush=True)
    assert cc1 < cc0, case


@pytest.mark.parametrize("case", ["C2-shape s16", "hub"])
def test_wide_triangle_pairs_change_nothing(pj, orc, case, monkeypatch):
    """3-vertex cycle sources of 257-2048 slots decided by pair tests against
    the LDS list of closing vertices (k_triangles_wide, on the side stream
    beside the other bins' walks, or after them with PJ_SIDE_WALKS=0) and by
    the walk (PJ_WIDE_TRIANGLES=0) give identical traces and states, and the
    reference's."""
    if case == "hub":
        g = pj.Graph.rmat(14, 32, seed=3).labels_uniform(4, seed=4)
        p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 0), (2, 3)], [0, 1, 2, 3])
    else:
        g = pj.Graph.rmat(16, 16, seed=20221).labels_zipf(32, 1.2, seed=7)
        p = pj.Graph.pattern(6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3])
    deg = np.diff(g.offsets().astype(np.int64))
    assert np.any((deg > 256) & (deg <= 2048))
    c = pj.Constraints(p)
    runs = []
    for wide, side in (("0", "1"), ("1", "0"), ("1", "1")):
        monkeypatch.setenv("PJ_WIDE_TRIANGLES", wide)
        monkeypatch.setenv("PJ_SIDE_WALKS", side)
        pr = pj.Pruner(g, c)
        rows, _ = pr.run(enumerate_matches=False)
        runs.append((rows, pr.state()))
        pr.close()
    (r0, s0) = runs[0]
    for r1, s1 in runs[1:]:
        assert [r.key() for r in r0] == [r.key() for r in r1], case
        assert np.array_equal(s0.cand, s1.cand) and np.array_equal(s0.vflags, s1.vflags)
        assert np.array_equal(s0.ea_words, s1.ea_words)
    r1 = runs[-1][0]
    o = orc.Oracle(g, c, threads=8)
    orows, _ = o.run(orc.FAITHFUL, enumerate_matches=False)
    assert [r.key() for r in r1] == [r.key() for r in orows], case


@pytest.mark.parametrize("case", ["C2-shape s16", "P=20, 40 labels"])
def test_phase1_hub_chunks_in_the_bins_launch_change_nothing(pj, orc, case, monkeypatch):
    """LCC phase 1 with the hub chunks in the launch of the degree bins
    (default) and in a launch of their own (PJ_PHASE1_FUSED=0) gives
    identical traces and states, and the reference's."""
    if case == "C2-shape s16":
        g = pj.Graph.rmat(16, 16, seed=20221).labels_zipf(32, 1.2, seed=7)
        p = pj.Graph.pattern(6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3])
    else:  # 32-bit candidate words, label counts over five words
        n = 20
        g = pj.Graph.rmat(15, 16, seed=9).labels_uniform(40, seed=10)
        p = pj.Graph.pattern(n, [(i, i + 1) for i in range(n - 1)] + [(0, 2)], [i % 40 for i in range(n)])
    deg = np.diff(g.offsets().astype(np.int64))
    assert np.any(deg > 2048), case  # hub rows exist
    c = pj.Constraints(p)
    runs = []
    for fused in ("0", "1"):
        monkeypatch.setenv("PJ_PHASE1_FUSED", fused)
        pr = pj.Pruner(g, c)
        rows, _ = pr.run(enumerate_matches=False)
        runs.append((rows, pr.state()))
        pr.close()
    (r0, s0), (r1, s1) = runs
    assert [r.key() for r in r0] == [r.key() for r in r1], case
    assert np.array_equal(s0.cand, s1.cand) and np.array_equal(s0.vflags, s1.vflags)
    assert np.array_equal(s0.ea_words, s1.ea_words)
    o = orc.Oracle(g, c, threads=8)
    orows, _ = o.run(orc.FAITHFUL, enumerate_matches=False)
    assert [r.key() for r in r1] == [r.key() for r in orows], case


@pytest.mark.parametrize("pair_degree", ["512", "16"])
@pytest.mark.parametrize("seed", [31, 32])
def test_triangle_steps_from_random_states_in_every_degree_bin(pj, orc, seed, pair_degree, monkeypatch):
    """3-vertex cycle steps from random states (candidates, flags and
    edge-active bits) on a graph with sources in every degree bin, so that
    k_triangles (1-8 / 9-32 slots), k_triangles_mid (33-64 slots in one
    chunk, 65-256 in up to four), k_triangles_wide (257-2048: the first hop
    closed as the walk does, the rest by pairs) and the hub walk all decide
    sources, many of them failing through cleared slots.  Rows above
    `pair_degree` slots are heavy: searches between two of them read the
    pair index."""
    from pattenmatching_sharedmemory_amd import STEP_CC

    monkeypatch.setenv("PJ_PAIR_DEGREE", pair_degree)
    rng = np.random.default_rng(seed)
    g = pj.Graph.rmat(14, 32, seed=3).labels_uniform(3, seed=4)
    deg = np.diff(g.offsets().astype(np.int64))
    # walk task list bin 0 holds rows of up to 8 slots (k_triangles<Word, 8>:
    # groups of 8 lanes, so rows of 1-4 and of 5-8 slots both occur), bin 1
    # rows of 9-32 (k_triangles<Word, 32>)
    for lo, hi in [(1, 4), (5, 8), (9, 32), (33, 64), (65, 256), (257, 2048), (2049, 1 << 40)]:
        assert np.any((deg >= lo) & (deg <= hi)), (lo, hi)
    p = pj.Graph.pattern(3, [(0, 1), (1, 2), (2, 0)], [0, 1, 2])
    c = pj.Constraints(p)
    assert len(c.cycles()) == 1
    pr = pj.Pruner(g, c)
    o = orc.Oracle(g, c, threads=8)
    V, E = g.vertex_count, g.edge_count
    exact = 0
    for rep in range(3):
        cand, vf, ea = _random_state(rng, V, E, 3)
        av = int((vf & 1).sum())
        pr.load_state(pj.State(cand, vf, ea, av, E))
        o.load_state(cand, vf, ea, av, E)
        r = pr.step(STEP_CC, 0).as_tuple()
        label = f"seed {seed} state {rep}"
        print(label, r, flush=True)
        st = pr.state()
        # bounds every schedule of the reference obeys; exact comparison when
        # no failing root is marked by another source's walk
        o.step(STEP_CC, 0, orc.FAITHFUL_NO_SHORTCUT)
        f0 = np.where(vf & 1, cand, 0) & ~o.state()[0]
        m0 = o.last_marks()
        f = np.where(vf & 1, cand, 0) & ~st.cand
        assert np.any(f), label
        assert np.all((f & ~f0) == 0) and np.all((f0 & ~m0 & ~f) == 0), label
        if np.any(f0 & m0):
            continue
        o.load_state(cand, vf, ea, av, E)
        ro = o.step(STEP_CC, 0, orc.FAITHFUL)
        oc, ovf, oea, oav, oae = o.state()
        assert r == ro, label
        assert np.array_equal(st.cand, oc), label
        assert np.array_equal(st.vflags, ovf), label
        assert np.array_equal(st.ea_words, oea), label
        exact += 1
    print(f"seed {seed}: {exact} of 3 states compared exactly", flush=True)
    pr.close()


def test_live_list_compaction_follows_the_reference(pj, orc, monkeypatch):
    """With the live-vertex list compacted at every halving of the active
    vertices (PJ_LIVE_MIN=16), the pipeline still matches the reference."""
    monkeypatch.setenv("PJ_LIVE_MIN", "16")
    g = pj.Graph.rmat(14, 16, seed=1).labels_uniform(8, seed=2)
    p = pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 3), (3, 0)], [0, 1, 2, 3])
    _compare_pipeline(pj, orc, g, p, "C1 s14 compacted")
    g = pj.Graph.rmat(16, 16, seed=20221).labels_zipf(32, 1.2, seed=7)
    p = pj.Graph.pattern(6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3])
    _compare_pipeline(pj, orc, g, p, "C2-shape s16 compacted", enumerate_matches=False)


def test_gathered_and_derived_label_codes_agree(pj):
    """LCC phase 1 takes neighbour label codes from candidate words after
    LCC0 and gathers them after a state upload; both give the same sweep."""
    from pattenmatching_sharedmemory_amd import STEP_LCC, STEP_LCC0

    g = pj.Graph.rmat(15, 16, seed=11).labels_zipf(16, 1.1, seed=12)
    c = pj.Constraints(pj.Graph.pattern(6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3]))
    a, b = pj.Pruner(g, c), pj.Pruner(g, c)
    a.step(STEP_LCC0)
    b.load_state(a.state())  # uploaded: codes gathered until LCC0 runs again
    while True:
        ra, rb = a.step(STEP_LCC), b.step(STEP_LCC)
        assert (ra.vertex_eliminated, ra.edge_eliminated, ra.match_eliminated) == \
            (rb.vertex_eliminated, rb.edge_eliminated, rb.match_eliminated)
        if not (ra.vertex_eliminated or ra.edge_eliminated or ra.match_eliminated):
            break
    sa, sb = a.state(), b.state()
    assert np.array_equal(sa.cand, sb.cand) and np.array_equal(sa.vflags, sb.vflags)
    assert np.array_equal(sa.ea_words, sb.ea_words)
    a.close()
    b.close()


def test_full_scale_plain_workload_invariants(pj):
    """The bench's plain configuration at full size (R-MAT scale 22, 128 M
    slots), checked through properties that do not need the oracle:
    LCC0's closed form, the counters against the state (every eliminated
    edge is one cleared edge-active bit, every active vertex has
    candidates), symmetric edge-active bits between active vertices at the
    LCC fixed point (the template adjacency is symmetric), and the fixed
    point itself (one more LCC step removes nothing)."""
    from pattenmatching_sharedmemory_amd import STEP_LCC, STEP_LCC0

    g = pj.Graph.rmat(22, 16, 0.57, 0.19, 0.19, seed=20221).labels_zipf(32, 1.2, seed=7)
    plab = [0, 1, 2, 3, 4, 3]
    c = pj.Constraints(pj.Graph.pattern(6, TRIANGLE_TAIL, plab))
    pr = pj.Pruner(g, c)
    V, E = g.vertex_count, g.edge_count
    # LCC0: a vertex keeps the template vertices of its label
    r0 = pr.step(STEP_LCC0)
    lab = g.labels()
    per_label = np.bincount(np.array(plab), minlength=int(lab.max()) + 1)
    keep = per_label[np.minimum(lab, len(per_label) - 1)] * (lab < len(per_label))
    assert r0.match_eliminated == int(np.sum(len(plab) - keep))
    assert r0.vertex_eliminated == int(np.count_nonzero(keep == 0))
    pr.reset()
    rows, _ = pr.run(enumerate_matches=False)
    st = pr.state()
    active = (st.vflags & 1).astype(bool)
    assert np.array_equal(active, st.cand != 0)
    assert rows[-1].active_vertices == int(np.count_nonzero(active))
    bits = np.unpackbits(st.ea_words.view(np.uint8), bitorder="little")[:E].astype(bool)
    eliminated = sum(r.edge_eliminated for r in rows)
    assert int(np.count_nonzero(bits)) == E - eliminated
    assert rows[-1].active_edges == E - eliminated
    # edge-active bits agree in both directions between active vertices
    off, edges = g.offsets(), g.edges()
    src = np.repeat(np.arange(V, dtype=np.int64), np.diff(off).astype(np.int64))
    both = active[src] & active[edges]
    key = src * V + edges.astype(np.int64)  # ascending: rows in order, each row ascending
    rkey = edges[both].astype(np.int64) * V + src[both]
    pos = np.searchsorted(key, rkey)
    assert np.all(key[pos] == rkey)  # symmetric graph
    assert np.array_equal(bits[both], bits[pos])
    # fixed point
    r = pr.step(STEP_LCC)
    assert (r.vertex_eliminated, r.edge_eliminated, r.match_eliminated) == (0, 0, 0)
    print(f"scale 22: {len(rows)} steps, {int(np.count_nonzero(active))} active vertices, "
          f"{E - eliminated} active slots", flush=True)
    pr.close()


@pytest.mark.parametrize("off,ed,msg", [
    ([0, 2, 1, 3], [1, 2, 0], "decrease"),
    ([0, 1, 2, 2], [1, 3], "not a vertex"),
    ([0, 1, 2, 5], [1, 0], "edge_count"),
])
def test_upload_refuses_malformed_csr(pj, off, ed, msg):
    """pj_upload_graph checks offsets on the host and neighbour ids in its
    row-order pass (which reads each slot once, in bounds) before any kernel
    indexes with them; the context stays usable for a good graph."""
    import ctypes as C

    from pattenmatching_sharedmemory_amd import native

    lib = native.hip()
    ctx = C.c_void_p()
    assert lib.pj_create(C.byref(ctx), 0) == 0
    o = np.array(off, dtype=np.uint64)
    e = np.array(ed, dtype=np.uint32)
    lab = np.zeros(3, dtype=np.uint64)
    u64, u32 = C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)
    rc = lib.pj_upload_graph(ctx, 3, len(ed), o.ctypes.data_as(u64), e.ctypes.data_as(u32), lab.ctypes.data_as(u64))
    assert rc == -1 and msg in lib.pj_last_error(ctx).decode()
    good_o, good_e = np.array([0, 1, 2, 2], dtype=np.uint64), np.array([1, 0], dtype=np.uint32)
    assert lib.pj_upload_graph(ctx, 3, 2, good_o.ctypes.data_as(u64), good_e.ctypes.data_as(u32),
                               lab.ctypes.data_as(u64)) == 0
    lib.pj_destroy(ctx)


@pytest.mark.parametrize("eid_bytes,file_values", [(4, True), (8, True), (8, False)])
def test_streamed_binary_file_equals_the_host_upload(pj, tmp_path, eid_bytes, file_values):
    """pj_upload_totem_file (edge list streamed through pinned buffers in
    64 MiB chunks: a scale-23 graph has several) gives the same resident
    graph as uploading the host CSR: equal step traces and state checksums.
    A truncated file is refused."""
    g = pj.Graph.rmat(23, 8, seed=9).labels_zipf(16, 1.1, seed=10)
    assert g.edge_count * 4 > 2 * (64 << 20)
    path = tmp_path / "g.bin"
    g.store_binary(str(path), eid_bytes)
    c = pj.Constraints(pj.Graph.pattern(6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3]))
    a = pj.Pruner(g, c)
    b = pj.Pruner(None, c, path=path) if file_values else pj.Pruner(None, c, path=path, labels=g.labels())
    assert (b.vertex_count, b.edge_count) == (g.vertex_count, g.edge_count)
    ra, _ = a.run(enumerate_matches=False)
    rb, _ = b.run(enumerate_matches=False)
    assert [r.key() for r in ra] == [r.key() for r in rb]
    assert a.checksum() == b.checksum()
    a.close()
    b.close()
    data = path.read_bytes()
    path.write_bytes(data[: len(data) - 12345])
    with pytest.raises(RuntimeError, match="truncated"):
        pj.Pruner(None, c, path=path)
