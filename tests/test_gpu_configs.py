"""The measured configurations against the oracle on the GPU:

* plain pipeline at full size: R-MAT scale 22, 32 Zipf(1.2) labels, the
  triangle + tail template [0,1,2,3,4,3] (the bench's workload), every step
  bit-exact;
* memory-scale form at oracle-feasible sizes: 64 Zipf(1.0) labels, the
  8-vertex two-cycle + chord template with its labels (ranks 8..15), the TDS
  step and the enumeration included;
* batched form: templates drawn as configured (3-10 vertices, a random
  spanning tree plus 0-3 extra edges, labels from the graph's Zipf
  distribution with repeats), every template against the oracle, and each
  result the same alone and in the batch;
* served form: templates of 3-12 vertices with 0-4 extra edges arriving as a
  Poisson process under a memory budget, likewise.
"""
import numpy as np
import pytest

from parity import STEP_CC, STEP_PC, compare_pipeline, oracle_checksum, stepwise

pytestmark = pytest.mark.gpu

GRAPH_SEED, LABEL_SEED = 20221, 7
TRIANGLE_TAIL = [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5)]
TWO_CYCLES_CHORD = [(0, 1), (1, 2), (2, 3), (3, 0), (1, 4), (4, 5), (5, 2), (0, 2), (5, 6), (6, 7)]


def _rmat(pj, scale, labels, zipf):
    return pj.Graph.rmat(scale, 16, 0.57, 0.19, 0.19, seed=GRAPH_SEED).labels_zipf(labels, zipf, seed=LABEL_SEED)


def test_plain_configuration_full_scale_matches_the_oracle(pj, orc):
    """R-MAT scale 22 (128 M slots): the device's whole solve against the
    oracle's.  Both constraints are memo-inert (distinct labels along the
    cycle and between the path's ends), where the reference's memoised search
    and the complete search decide alike, so the oracle runs its complete
    mode (cycle closing by binary search) to finish in seconds."""
    g = _rmat(pj, 22, 32, 1.2)
    c = pj.Constraints(pj.Graph.pattern(6, TRIANGLE_TAIL, [0, 1, 2, 3, 4, 3]))
    assert c.memo_inert(STEP_CC, 0) == 1 and c.memo_inert(STEP_PC, 0) == 1
    rows, _ = compare_pipeline(pj, orc, g, c, "plain s22", enumerate_matches=False, mode=orc.COMPLETE, threads=16)
    assert [r.step for r in rows].count(STEP_CC) == 1 and [r.step for r in rows].count(STEP_PC) == 1


@pytest.mark.parametrize("scale,enumerate_matches", [(14, True), (16, True), (18, False)])
def test_memory_scale_configuration_with_tds(pj, orc, scale, enumerate_matches):
    """64 Zipf(1.0) labels, the 8-vertex template with labels of ranks 8..15
    (two 4-cycles sharing t1-t2, chord t0-t2, tail t5-t6-t7): every step of
    the device against the oracle from the oracle's states (LCC bit-exact,
    cycle and TDS steps within the outcomes of the reference's schedules and
    exact where the step is determined), then, when every step is
    determined, the device's own whole trace bit for bit.  At scale 18 the
    TDS step is schedule-dependent in the reference (two failing roots are
    marked by other sources' successful walks)."""
    g = _rmat(pj, scale, 64, 1.0)
    c = pj.Constraints(pj.Graph.pattern(8, TWO_CYCLES_CHORD, list(range(8, 16))))
    assert len(c.tds()) == 1
    pr = pj.Pruner(g, c)
    determined = stepwise(pj, orc, g, c, pr, f"memory-scale form s{scale}", enumerate_matches=enumerate_matches)
    pr.close()
    print(f"[memory-scale form s{scale}] every step determined: {determined}", flush=True)
    if determined:
        rows, _ = compare_pipeline(pj, orc, g, c, f"memory-scale form s{scale}",
                                   enumerate_matches=enumerate_matches)
        assert 6 in [r.step for r in rows]  # the TDS step ran
    else:
        assert scale == 18


def _configured(pj, n, seed, max_vertices, max_extra):
    from pattenmatching_sharedmemory_amd.templates import random_templates

    return [pj.Constraints(pj.Graph.pattern(k, e, l))
            for k, e, l in random_templates(n, seed, 3, max_vertices, max_extra, labels=32, zipf=1.2)]


def _every_template_against_the_oracle(pj, orc, g, cs, label):
    """Step-by-step comparison of each template with the oracle; returns the
    indices of the templates whose result is determined (no memoised step's
    outcome depends on the reference's schedule) and the oracle's final
    state of each."""
    determined, finals = [], []
    for i, c in enumerate(cs):
        pr = pj.Pruner(g, c)
        ok = stepwise(pj, orc, g, c, pr, f"{label} template {i}")
        pr.close()
        o = orc.Oracle(g, c, threads=8)
        o.run(orc.FAITHFUL, enumerate_matches=False)
        finals.append(o.state())
        if ok:
            determined.append(i)
    print(f"[{label}] {len(cs)} templates against the oracle, {len(determined)} determined", flush=True)
    return determined, finals


def test_batched_configuration_templates_against_the_oracle(pj, orc):
    """32 templates drawn as the batched form specifies, on R-MAT scale 12
    with 32 Zipf(1.2) labels; TDS steps included.  Every template's every
    step against the oracle; the batch (4 worker contexts sharing the graph)
    gives each template the result it gets alone, and for determined
    templates the oracle's final state."""
    g = _rmat(pj, 12, 32, 1.2)
    cs = _configured(pj, 32, 4242, 10, 3)
    assert any(c.tds() for c in cs) and any(len(set(c.desc.contents.vertex_label[i] for i in range(c.vertex_count)))
                                            < c.vertex_count for c in cs)
    determined, finals = _every_template_against_the_oracle(pj, orc, g, cs, "batched")
    assert len(determined) >= len(cs) // 2
    owner = pj.Pruner(g)
    alone, _ = owner.run_queries(cs, workers=1, run_tds=True)
    batch, _ = owner.run_queries(cs, workers=4, run_tds=True)
    assert len({r.worker for r in batch}) > 1
    # the LCC0 and first LCC fixed point of every template in shared sweeps
    shared, _ = owner.run_queries(cs, workers=4, run_tds=True, batch_lcc=True)
    assert any(r.batch_lane >= 0 for r in shared)
    for i in determined:
        assert alone[i].result() == batch[i].result(), i
        assert alone[i].result() == shared[i].result(), i
        assert batch[i].state_checksum == oracle_checksum(pj, g, cs[i], finals[i]), i
    owner.close()


def test_served_configuration_templates_against_the_oracle(pj, orc):
    """24 templates drawn as the served form specifies (3-12 vertices, 0-4
    extra edges), on R-MAT scale 13, arriving as a Poisson process and
    admitted under a device-memory budget by 6 workers: every template
    against the oracle step by step; each determined template's result is
    the same alone and in the mix, and the oracle's."""
    g = _rmat(pj, 13, 32, 1.2)
    cs = _configured(pj, 24, 777, 12, 4)
    determined, finals = _every_template_against_the_oracle(pj, orc, g, cs, "served")
    assert len(determined) >= len(cs) // 2
    owner = pj.Pruner(g)
    alone, _ = owner.run_queries(cs, workers=1, run_tds=True)
    rng = np.random.default_rng(3)
    arrivals = np.cumsum(rng.exponential(1.0, len(cs)))
    budget = int(max(r.projected_bytes for r in alone) * 3)
    served, _ = owner.run_queries(cs, workers=6, run_tds=True, arrivals_ms=arrivals, budget_bytes=budget)
    # continuous batching: arrivals join the shared LCC sweeps as lanes free up
    # (every arrival takes a lane: this small a load would otherwise be sent
    # whole to idle workers)
    import os

    os.environ["PJ_SERVED_BATCH_MIN"] = "0"
    try:
        shared, _ = owner.run_queries(cs, workers=6, run_tds=True, arrivals_ms=arrivals, budget_bytes=budget,
                                      batch_lcc=True)
    finally:
        del os.environ["PJ_SERVED_BATCH_MIN"]
    assert any(r.batch_lane >= 0 for r in shared)
    # the default schedule: arrivals that find a worker idle run whole, a
    # batch starts once enough of them wait
    adaptive, _ = owner.run_queries(cs, workers=6, run_tds=True, arrivals_ms=arrivals, budget_bytes=budget,
                                    batch_lcc=True)
    for form in (served, shared, adaptive):
        for r, a in zip(form, arrivals):
            assert r.start_ms + 1e-6 >= a and r.worker_ms + 1e-6 >= r.start_ms
        for r in form:  # bytes of the queries on workers when r reached one
            live = sum(q.projected_bytes for q in form if q.worker_ms <= r.worker_ms < q.end_ms)
            assert live <= budget
    for i in determined:
        assert alone[i].result() == served[i].result(), i
        assert alone[i].result() == shared[i].result(), i
        assert alone[i].result() == adaptive[i].result(), i
        assert served[i].state_checksum == oracle_checksum(pj, g, cs[i], finals[i]), i
    owner.close()
