"""pj_compact_active_noalloc (static scratch, no allocation) against
pj_compact_active on two calls in a row through the C ABI, the second on a
smaller graph: same counts, same vertices, candidate sets and slots."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _compact(lib, fn, pr, cap_v, cap_e):
    import torch

    dv = torch.full((cap_v,), -1, dtype=torch.int32, device="cuda")
    dc = torch.full((cap_v,), -1, dtype=torch.int64, device="cuda")
    ds = torch.full((cap_e,), -1, dtype=torch.int32, device="cuda")
    dd = torch.full((cap_e,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    nv, ne = C.c_uint64(), C.c_uint64()
    rc = getattr(lib, fn)(pr.ctx, dv.data_ptr(), dc.data_ptr(), cap_v, ds.data_ptr(), dd.data_ptr(), cap_e,
                          C.byref(nv), C.byref(ne))
    assert rc == 0, lib.pj_last_error(pr.ctx)
    return nv.value, ne.value, [t.cpu().numpy() for t in (dv, dc, ds, dd)]


def test_no_alloc_compaction_matches_the_allocating_one(pj):
    from pattenmatching_sharedmemory_amd import native

    lib = native.hip()
    big = pj.Graph.rmat(18, 16, seed=20221).labels_zipf(32, 1.2, seed=7)
    small = pj.Graph.rmat(13, 8, seed=3).labels_uniform(4, seed=4)
    tmpl = pj.Constraints(pj.Graph.pattern(6, [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5)], [0, 1, 2, 3, 4, 3]))
    sq = pj.Constraints(pj.Graph.pattern(4, [(0, 1), (1, 2), (2, 3), (3, 0)], [0, 1, 2, 3]))
    for label, g, c in (("scale 18", big, tmpl), ("scale 13", small, sq)):
        pr = pj.Pruner(g, c)
        pr.run(enumerate_matches=False)
        cap_v, cap_e = g.vertex_count + 3, g.edge_count + 3
        a = _compact(lib, "pj_compact_active", pr, cap_v, cap_e)
        b = _compact(lib, "pj_compact_active_noalloc", pr, cap_v, cap_e)
        print(f"[{label}] {a[0]} vertices, {a[1]} slots survive", flush=True)
        assert a[0] > 0 and a[1] > 0
        assert (a[0], a[1]) == (b[0], b[1]), label
        for x, y in zip(a[2], b[2]):
            assert np.array_equal(x, y), label
        pr.close()
