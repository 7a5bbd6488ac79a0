"""Helpers of the GPU parity tests: compare the device's pipeline with the
oracle (CPU restatement of the reference) on the same inputs.

Where the reference's own result depends on thread timing (a source skips a
walk whose root another source's successful walk has marked, alreadyMatched,
multiple_label_cc_cpu.tpp:102-108), every schedule's failures F of a memoised
step satisfy F0 - M0 <= F <= F0, F0 and M0 being the failures and marks with
every walk run (oracle mode 2).  A step with F0 & M0 empty is determined: all
schedules, and the device, give F0."""
from __future__ import annotations

import numpy as np

STEP_LCC0, STEP_LCC, STEP_CC, STEP_PC, STEP_TDS, STEP_ENUMERATION = 2, 3, 4, 5, 6, 7
NLCC = (STEP_CC, STEP_PC, STEP_TDS)


def assert_same_state(st, oracle_state, label):
    cand, vf, ea, av, ae = oracle_state
    assert np.array_equal(st.cand, cand), label
    assert np.array_equal(st.vflags, vf), label
    assert np.array_equal(st.ea_words, ea), label
    assert (st.active_vertices, st.active_edges) == (av, ae), label


def oracle_plan(c):
    """The step machine's NLCC order: cycles, paths, then TDS (always
    constraint 0: the reference never advances the TDS iterator)."""
    plan = [(STEP_CC, i) for i in range(len(c.cycles()))] + [(STEP_PC, i) for i in range(len(c.paths()))]
    return plan + [(STEP_TDS, 0)] * len(c.tds())


def stepwise(pj, orc, g, c, pr, label, threads=8, enumerate_matches=False):
    """Runs the step machine on the oracle (faithful mode) and the device
    step by step from the oracle's states: LCC0 and every LCC step must be
    bit-exact; every NLCC step must land within the bounds of the reference's
    schedules, and equal the oracle when the step is determined.  Returns
    True when every NLCC step was determined (then the pipeline's result is
    unique)."""
    from pattenmatching_sharedmemory_amd import State

    o = orc.Oracle(g, c, threads=threads)
    determined = True

    def lcc_step(step):
        st = o.state()
        pr.load_state(State(*st))
        r = pr.step(step).as_tuple()
        ro = o.step(step, 0, orc.FAITHFUL)
        assert r == ro, (label, step)
        assert_same_state(pr.state(), o.state(), (label, step))
        return any(ro)

    lcc_step(STEP_LCC0)
    while lcc_step(STEP_LCC):
        pass
    for step, idx in oracle_plan(c):
        o.reset_modified()
        before = o.state()
        cand, vf = before[0], before[1]
        live = np.where(vf & 1, cand, 0).astype(np.uint64)
        x = orc.Oracle(g, c, threads=threads)
        x.load_state(*before)
        x.step(step, idx, orc.FAITHFUL_NO_SHORTCUT)
        f0 = live & ~x.state()[0]
        m0 = x.last_marks()
        pr.load_state(State(*before))
        pr.step(step, idx)
        f = live & ~pr.state().cand
        assert np.all((f & ~f0) == 0), (label, step, idx)
        assert np.all((f0 & ~m0 & ~f) == 0), (label, step, idx)
        o.step(step, idx, orc.FAITHFUL)
        if np.any(f0 & m0):
            determined = False
        else:
            assert_same_state(pr.state(), o.state(), (label, step, idx))
        while lcc_step(STEP_LCC):
            pass
    if enumerate_matches:
        st = o.state()
        pr.load_state(State(*st))
        assert pr.step(STEP_ENUMERATION).as_tuple() == o.step(STEP_ENUMERATION, 0, orc.FAITHFUL), label
    return determined


def compare_pipeline(pj, orc, g, c, label, enumerate_matches=True, mode=None, threads=8):
    """Whole step machine on the device and on the oracle: per-step counter
    trace, candidate sets, flags, edge-active bits and counters bit-exact."""
    mode = orc.FAITHFUL if mode is None else mode
    pr = pj.Pruner(g, c)
    rows, _ = pr.run(enumerate_matches=enumerate_matches)
    o = orc.Oracle(g, c, threads=threads)
    orows, _ = o.run(mode, enumerate_matches=enumerate_matches)
    print(f"[{label}] V={g.vertex_count} E={g.edge_count} steps={len(rows)}", flush=True)
    assert [r.key() for r in rows] == [r.key() for r in orows], label
    assert_same_state(pr.state(), o.state(), label)
    pr.close()
    return rows, o


def oracle_checksum(pj, g, c, oracle_state) -> int:
    """pj_state_checksum of the oracle's state, taken by loading it into a
    device context (the checksum covers candidate words, flags and
    edge-active bits)."""
    from pattenmatching_sharedmemory_amd import State

    pr = pj.Pruner(g, c)
    pr.load_state(State(*oracle_state))
    h = pr.checksum()
    pr.close()
    return h
