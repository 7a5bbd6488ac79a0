"""Batched LCC (pj_batch): LCC0 and the LCC steps to the first fixed point of
many templates in shared sweeps.  Every template's per-step counters and its
parked state (candidate sets, flags, edge-active bits, active counters) must
equal the oracle's LCC0 and LCC steps on the same template, whichever lanes
run beside it and whenever it was admitted."""
import numpy as np
import pytest

from parity import STEP_LCC, STEP_LCC0, assert_same_state

pytestmark = pytest.mark.gpu


def _templates(pj, n, seed, max_vertices=10, max_extra=3):
    from pattenmatching_sharedmemory_amd.templates import random_templates

    return [pj.Constraints(pj.Graph.pattern(k, e, l))
            for k, e, l in random_templates(n, seed, 3, max_vertices, max_extra)]


def _special(pj):
    """Templates at the edges of the batch's range: 16 vertices (the widest
    word), a star needing 6 neighbours of one label (4-plane counters), a
    single edge, and a template whose labels no vertex carries."""
    out = []
    k = 16
    out.append(pj.Constraints(pj.Graph.pattern(k, [(i, i + 1) for i in range(k - 1)] + [(0, 15), (3, 9)],
                                               [i % 5 for i in range(k)])))
    out.append(pj.Constraints(pj.Graph.pattern(7, [(0, i) for i in range(1, 7)], [0, 1, 1, 1, 1, 1, 1])))
    out.append(pj.Constraints(pj.Graph.pattern(2, [(0, 1)], [0, 0])))
    out.append(pj.Constraints(pj.Graph.pattern(3, [(0, 1), (1, 2)], [200, 201, 200])))
    return out


def _oracle_lcc(orc, g, c):
    """Rows (vertices, edges, matches eliminated) and state of the oracle's
    LCC0 and LCC steps up to the first step that eliminates nothing."""
    o = orc.Oracle(g, c, threads=8)
    rows = [o.step(STEP_LCC0, 0, orc.FAITHFUL)]
    while True:
        r = o.step(STEP_LCC, 0, orc.FAITHFUL)
        rows.append(r)
        if not any(r):
            break
    return [r[:3] for r in rows], o.state()


def _check_lane(pj, orc, g, batch, lane, c, pr, label):
    trace = batch.trace(lane)
    rows, state = _oracle_lcc(orc, g, c)
    assert [t[:3] for t in trace] == rows, label
    pr.set_template(c)
    batch.export(lane, pr)
    assert_same_state(pr.state(), state, label)
    # the exported state continues exactly like the oracle's
    o = orc.Oracle(g, c, threads=8)
    o.load_state(*state)
    pr.reset_modified()
    o.reset_modified()
    assert pr.step(STEP_LCC).as_tuple() == o.step(STEP_LCC, 0, orc.FAITHFUL), label


def test_batch_lcc_matches_the_oracle(pj, orc):
    g = pj.Graph.rmat(15, 16, seed=3).labels_zipf(32, 1.2, seed=4)
    deg = np.diff(g.offsets().astype(np.int64))
    assert (deg > 2048).sum() > 0  # hub rows go by chunks
    cs = _special(pj) + _templates(pj, 80, seed=11)
    owner = pj.Pruner(g)
    batch = pj.Batch(owner, lanes=len(cs))
    lanes = [batch.admit(c, strict=False) for c in cs]
    held = [(lane, c) for lane, c in zip(lanes, cs) if lane is not None]
    assert len(held) >= 70 and all(lane is not None for lane in lanes[:4])
    cs = [c for _, c in held]
    lanes = [lane for lane, _ in held]
    assert len(set(lanes)) == len(cs)
    done = batch.run_all()
    assert sorted(done) == sorted(lanes)
    pr = pj.Pruner(g)
    for lane, c in zip(lanes, cs):
        print(f"[batch lane {lane}] P={c.vertex_count}", flush=True)
        _check_lane(pj, orc, g, batch, lane, c, pr, f"lane {lane}")
    batch.close()
    pr.close()
    owner.close()


def test_lanes_admitted_while_others_run(pj, orc):
    """Continuous admission: templates join while others are mid-way, parked
    lanes are released and taken by new templates."""
    g = pj.Graph.rmat(14, 16, seed=5).labels_zipf(32, 1.2, seed=6)
    cs = _templates(pj, 40, seed=23, max_vertices=8)
    owner = pj.Pruner(g)
    batch = pj.Batch(owner, lanes=16)
    pr = pj.Pruner(g)
    pending = list(range(len(cs)))
    lane_of = {}
    checked = 0
    while pending or batch.info()["running"]:
        while pending and batch.info()["running"] + batch.info()["parked"] < 16:
            i = pending.pop(0)
            lane = batch.admit(cs[i], strict=False)
            if lane is None:  # not for the batch
                checked += 1
                continue
            lane_of[lane] = i
            if len(pending) % 3 == 0:
                break  # the rest joins after another step
        for lane in batch.run(max_steps=1):
            i = lane_of.pop(lane)
            print(f"[admission] template {i} lane {lane}", flush=True)
            _check_lane(pj, orc, g, batch, lane, cs[i], pr, f"template {i}")
            batch.release(lane)
            checked += 1
    assert checked == len(cs)
    batch.close()
    pr.close()
    owner.close()


def test_batch_rejects_what_it_cannot_hold(pj):
    g = pj.Graph.rmat(10, 16, seed=1).labels_zipf(8, 1.2, seed=2)
    owner = pj.Pruner(g)
    batch = pj.Batch(owner, lanes=1)
    big = pj.Constraints(pj.Graph.pattern(17, [(i, i + 1) for i in range(16)], [i % 8 for i in range(17)]))
    with pytest.raises(RuntimeError, match="1..16"):
        batch.admit(big)
    small = _templates(pj, 2, seed=3)
    batch.admit(small[0])
    assert batch.info()["lanes"] == 1
    with pytest.raises(RuntimeError, match="no free lane"):
        batch.admit(small[1])
    batch.close()
    owner.close()
