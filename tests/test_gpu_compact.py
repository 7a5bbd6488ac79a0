"""Active-graph compaction on the device (pj_compact_active /
pj_download_active) against the surviving graph computed on the host from
the downloaded state: vertices with a non-empty candidate set (ascending,
with their sets) and edge-active slots between two of them (CSR order) --
what printActiveVertex / printActiveGraph dump
(multiple_label_cpu.cc:270-297)."""
import ctypes as C

import numpy as np
import pytest

from conftest import FIXTURES, fixture_paths

pytestmark = pytest.mark.gpu


def expected(g, st):
    off, ed = g.offsets().astype(np.int64), g.edges()
    alive = st.cand != 0
    bits = np.unpackbits(st.ea_words.view(np.uint8), bitorder="little")[: g.edge_count].astype(bool)
    src = np.repeat(np.arange(g.vertex_count, dtype=np.int64), np.diff(off))
    keep = bits & alive[src] & alive[ed]
    v = np.nonzero(alive)[0].astype(np.uint32)
    return v, st.cand[alive], src[keep].astype(np.uint32), ed[keep].astype(np.uint32)


def cases(pj):
    for name in FIXTURES:
        f = fixture_paths(name)
        yield name, pj.Graph.load(f["graph"], f["vertex"]), pj.Graph.load(f["pattern"], f["pattern_vertex"]), True
    g = pj.Graph.rmat(16, 16, seed=20221).labels_zipf(32, 1.2, seed=7)
    yield "rmat16 triangle+tail", g, pj.Graph.pattern(6, [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5)],
                                                      [0, 1, 2, 3, 4, 3]), False
    h = pj.Graph.rmat(17, 32, seed=3).labels_uniform(4, seed=4)  # rows crossing the 2^16-slot block boundaries
    yield "rmat17 hubs", h, pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 0), (2, 3)], [0, 1, 2, 3]), False


def test_download_active_matches_the_state(pj):
    for name, g, p, enum in cases(pj):
        pr = pj.Pruner(g, pj.Constraints(p))
        for when in ("initial", "pruned"):
            if when == "pruned":
                pr.run(enumerate_matches=enum)
            got = pr.active_graph()
            want = expected(g, pr.state())
            print(f"[{name} {when}] {len(got[0])} vertices, {len(got[2])} slots survive", flush=True)
            for a, b in zip(got, want):
                assert np.array_equal(a, b), (name, when)
        pr.close()


def test_compact_into_device_buffers_and_capacity(pj):
    import torch

    from pattenmatching_sharedmemory_amd import native

    lib = native.hip()
    g = pj.Graph.rmat(15, 16, seed=5).labels_zipf(16, 1.1, seed=6)
    pr = pj.Pruner(g, pj.Constraints(pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 3), (3, 0)], [0, 1, 2, 3])))
    pr.run(enumerate_matches=False)
    v, c, s, d = expected(g, pr.state())
    nv, ne = C.c_uint64(), C.c_uint64()
    dv = torch.zeros(len(v) + 7, dtype=torch.int32, device="cuda")
    dc = torch.zeros(len(v) + 7, dtype=torch.int64, device="cuda")
    ds = torch.zeros(len(s) + 5, dtype=torch.int32, device="cuda")
    dd = torch.zeros(len(s) + 5, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = lib.pj_compact_active(pr.ctx, dv.data_ptr(), dc.data_ptr(), dv.numel(), ds.data_ptr(), dd.data_ptr(),
                               ds.numel(), C.byref(nv), C.byref(ne))
    assert rc == 0 and (nv.value, ne.value) == (len(v), len(s))
    assert np.array_equal(dv.cpu().numpy()[: len(v)].view(np.uint32), v)
    assert np.array_equal(dc.cpu().numpy()[: len(v)].view(np.uint64), c)
    assert np.array_equal(ds.cpu().numpy()[: len(s)].view(np.uint32), s)
    assert np.array_equal(dd.cpu().numpy()[: len(s)].view(np.uint32), d)
    assert int(ds[len(s):].abs().sum()) == 0  # nothing past the survivors
    # too small: -1, counts still reported, nothing written past the capacity
    ds.zero_()
    torch.cuda.synchronize()
    rc = lib.pj_compact_active(pr.ctx, dv.data_ptr(), None, dv.numel(), ds.data_ptr(), dd.data_ptr(), 10,
                               C.byref(nv), C.byref(ne))
    assert rc == -1 and ne.value == len(s) and b"capacity" in lib.pj_last_error(pr.ctx)
    assert np.array_equal(ds.cpu().numpy()[:10].view(np.uint32), s[:10]) and int(ds[10:].abs().sum()) == 0
    pr.close()
