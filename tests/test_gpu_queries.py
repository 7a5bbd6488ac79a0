"""Batched and served forms (pjh_run_queries): worker contexts sharing one
resident graph solve every template exactly as a context of its own does,
alone or in the mix, and admission keeps the projected bytes in flight
within the budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(pj, n, seed=5):
    from pattenmatching_sharedmemory_amd.templates import longest_cycle, random_templates

    g = pj.Graph.rmat(13, 16, seed=3).labels_zipf(32, 1.2, seed=4)
    cs = []
    for k, e, l in random_templates(4 * n, seed, 3, 10, 3, distinct_labels=True):
        c = pj.Constraints(pj.Graph.pattern(k, e, l))
        if longest_cycle(c) <= 4:
            cs.append(c)
        if len(cs) == n:
            break
    return g, cs


def test_batched_results_equal_separate_solves(pj):
    g, cs = _setup(pj, 16)
    owner = pj.Pruner(g)
    seq, _ = owner.run_queries(cs, workers=1, run_tds=False)
    bat, _ = owner.run_queries(cs, workers=4, run_tds=False)
    assert [r.result() for r in seq] == [r.result() for r in bat]
    assert len({r.worker for r in bat}) > 1
    for c, r in zip(cs, bat):  # a context of its own, through the plain entry points
        pr = pj.Pruner(g, c)
        rows, _ = pr.run(enumerate_matches=False, run_tds=False)
        assert (r.steps, r.active_vertices, r.active_edges) == (len(rows), rows[-1].active_vertices,
                                                                 rows[-1].active_edges)
        assert r.state_checksum == pr.checksum()
        pr.close()
    owner.close()


def test_shared_lcc_sweeps_give_every_template_its_own_result(pj):
    """PJH_BATCH_LCC: templates of at most 16 vertices run LCC0 and their
    first LCC fixed point in a batch, with fewer lanes than templates (lanes
    are taken again as templates leave), a template of 18 vertices goes to
    the workers as it is; every result equals the template's own solve."""
    import os

    g, cs = _setup(pj, 40, seed=7)
    k = 18
    cs.append(pj.Constraints(pj.Graph.pattern(k, [(i, i + 1) for i in range(k - 1)], [i % 6 for i in range(k)])))
    owner = pj.Pruner(g)
    seq, _ = owner.run_queries(cs, workers=1, run_tds=False)
    os.environ["PJ_BATCH_LANES"] = "16"
    try:
        shared, _ = owner.run_queries(cs, workers=3, run_tds=False, batch_lcc=True)
    finally:
        del os.environ["PJ_BATCH_LANES"]
    assert [r.result() for r in seq] == [r.result() for r in shared]
    assert all(r.batch_lane >= 0 for r in shared[:-1]) and shared[-1].batch_lane == -1
    owner.close()


def test_served_results_equal_alone_and_admission_respects_budget(pj):
    g, cs = _setup(pj, 24, seed=9)
    owner = pj.Pruner(g)
    alone, _ = owner.run_queries(cs, workers=1, run_tds=False)
    rng = np.random.default_rng(1)
    arrivals = np.cumsum(rng.exponential(2.0, len(cs)))
    budget = int(max(r.projected_bytes for r in alone) * 2.5)
    srv, _ = owner.run_queries(cs, workers=6, run_tds=False, arrivals_ms=arrivals, budget_bytes=budget)
    assert [r.result() for r in alone] == [r.result() for r in srv]
    for r, a in zip(srv, arrivals):
        assert r.start_ms + 1e-6 >= a and r.end_ms >= r.start_ms
    for r in srv:  # bytes of the queries running when r started
        live = sum(q.projected_bytes for q in srv if q.start_ms <= r.start_ms < q.end_ms)
        assert live <= budget
    owner.close()


def test_checksum_tracks_the_state(pj):
    g, cs = _setup(pj, 1)
    a, b = pj.Pruner(g, cs[0]), pj.Pruner(g, cs[0])
    assert a.checksum() == b.checksum()
    from pattenmatching_sharedmemory_amd import STEP_LCC0

    a.step(STEP_LCC0)
    assert a.checksum() != b.checksum()
    b.step(STEP_LCC0)
    assert a.checksum() == b.checksum()
    a.close()
    b.close()


@pytest.mark.parametrize("kind", ["inert triangle", "memoised triangle", "tds diamond", "long cycle"])
def test_projection_covers_what_a_query_context_allocates(pj, kind):
    """Device memory taken by a worker context (sharing the owner's graph)
    over a whole solve stays within pjh_projected_bytes of its template, up
    to the allocator's rounding (2 MiB per allocation).  The retry tables are
    not the context's: one set per device, allocated when a source first
    outgrows the first-pass tables and shared afterwards; so the first
    context's solve may add up to PJ_RETRY_POOL_BYTES, a second context's
    solve of the same template stays within its projection."""
    import ctypes as C

    import torch

    from pattenmatching_sharedmemory_amd import native

    shapes = {
        "inert triangle": (3, [(0, 1), (1, 2), (2, 0)], [0, 1, 2]),
        "memoised triangle": (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 2)], [0, 3, 1, 2, 1]),
        "tds diamond": (4, [(0, 1), (0, 2), (0, 3), (1, 2), (2, 3)], [0, 1, 2, 3]),
        "long cycle": (5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)], [0, 1, 2, 3, 4]),
    }
    pool = (32 << 22) * 20 + (4 << 26) * 20 + (1 << 29) * 20  # PJ_RETRY_POOL_BYTES
    g = pj.Graph.rmat(16, 16, seed=3).labels_zipf(8, 1.0, seed=4)
    c = pj.Constraints(pj.Graph.pattern(*shapes[kind]))
    owner = pj.Pruner(g)
    hip, host = native.hip(), native.host()
    projected = host.pjh_projected_bytes(c.handle, g.vertex_count, g.edge_count)
    slack = 48 * (2 << 20)
    torch.cuda.init()
    for attempt in range(2):
        ctx = C.c_void_p()
        assert hip.pj_create(C.byref(ctx), 0) == 0
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        assert hip.pj_share_graph(ctx, owner.ctx) == 0
        assert hip.pj_set_template(ctx, c.desc) == 0
        total = C.c_double()
        assert host.pjh_run_pipeline(ctx, c.handle, 1, None, 0, C.byref(total)) > 0
        torch.cuda.synchronize()
        used = free0 - torch.cuda.mem_get_info()[0]
        print(f"[{kind} context {attempt}] used {used} bytes, projected {projected}", flush=True)
        assert used <= projected + slack + (pool if attempt == 0 else 0), (kind, attempt)
        hip.pj_destroy(ctx)
    owner.close()
