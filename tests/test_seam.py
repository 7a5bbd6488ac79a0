"""The reference's driver seam (multiple_label_cpu.h:34-52) with the typedef
block swapped to this repository's step classes: tests/seam/seam_driver.cc
declares stand-ins of the reference's graph_t and constraint classes with
the reference's field names and drives the step classes the way
MultipleLabelCpu does.  On the CPU: the template the step classes gather from
the reference-shaped constraint objects equals the generator's.  On the GPU:
the whole step loop through the seam, with 32- and 64-bit graph_t offsets,
equals the oracle (step trace, active vertices, active edges, candidate
sets)."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from conftest import FIXTURES, fixture_paths

LIB = Path(__file__).resolve().parent / "seam" / "libpj_seam.so"

TEMPLATES = [
    (6, [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5)], [0, 1, 2, 3, 4, 3]),
    (8, [(0, 1), (1, 2), (2, 3), (3, 0), (1, 4), (4, 5), (5, 2), (0, 2), (5, 6), (6, 7)], list(range(8))),
    (4, [(0, 1), (1, 2), (2, 3), (3, 0)], [0, 1, 0, 1]),
]


def _lib():
    if not LIB.exists():
        raise RuntimeError(f"seam check not built: {LIB} (python tests/seam/build.py)")
    from pattenmatching_sharedmemory_amd import native

    native.host()  # device and host libraries first
    lib = C.CDLL(str(LIB))
    u32p, u64p, u8p = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint8)
    lib.pjs_assembly_matches.restype = C.c_int
    lib.pjs_assembly_matches.argtypes = [C.c_uint32, C.c_uint32, u32p, u32p, u64p]
    lib.pjs_run.restype = C.c_int
    lib.pjs_run.argtypes = [C.c_uint32, C.c_uint64, u64p, u32p, u64p, C.c_uint32, C.c_uint32, u32p, u32p, u64p, C.c_int,
                            C.c_void_p, C.c_int, u8p, u8p, u64p, u64p]
    return lib


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _pattern_arrays(p):
    off, ed = p.offsets(), p.edges()
    src, dst = [], []
    for v in range(p.vertex_count):
        for e in range(int(off[v]), int(off[v + 1])):
            if v < int(ed[e]):
                src.append(v)
                dst.append(int(ed[e]))
    return (p.vertex_count, np.array(src, dtype=np.uint32), np.array(dst, dtype=np.uint32),
            np.ascontiguousarray(p.labels(), dtype=np.uint64))


def _patterns(pj):
    out = []
    for name in FIXTURES:
        f = fixture_paths(name)
        out.append((name, pj.Graph.load(f["pattern"], f["pattern_vertex"])))
    for i, (n, e, l) in enumerate(TEMPLATES):
        out.append((f"t{i}", pj.Graph.pattern(n, e, l)))
    return out


def test_step_classes_gather_the_generators_template(pj):
    lib = _lib()
    for name, p in _patterns(pj):
        P, src, dst, lab = _pattern_arrays(p)
        assert lib.pjs_assembly_matches(P, len(src), _ptr(src, C.c_uint32), _ptr(dst, C.c_uint32),
                                        _ptr(lab, C.c_uint64)) == 1, name


@pytest.mark.gpu
@pytest.mark.parametrize("eid64", [0, 1])
def test_step_loop_through_the_seam_matches_the_oracle(pj, orc, eid64):
    from pattenmatching_sharedmemory_amd.native import Row

    lib = _lib()
    cases = []
    for name in FIXTURES:
        f = fixture_paths(name)
        cases.append((name, pj.Graph.load(f["graph"], f["vertex"]), pj.Graph.load(f["pattern"], f["pattern_vertex"])))
    # every memoised step determined (no failing root marked by another
    # source's walk), enumeration of at most ~10^5 matches
    g = pj.Graph.rmat(11, 16, seed=7).labels_uniform(16, seed=11)
    for i, (n, e, l) in enumerate(TEMPLATES):
        cases.append((f"rmat11 t{i}", g, pj.Graph.pattern(n, e, l)))
    for name, g, p in cases:
        c = pj.Constraints(p)
        P, src, dst, lab = _pattern_arrays(p)
        V, E = g.vertex_count, g.edge_count
        off = np.ascontiguousarray(g.offsets(), dtype=np.uint64)
        ed = np.ascontiguousarray(g.edges(), dtype=np.uint32)
        glab = np.ascontiguousarray(g.labels(), dtype=np.uint64)
        rows = (Row * 1000)()
        va, ea, cand = np.zeros(V, np.uint8), np.zeros(E, np.uint8), np.zeros(V, np.uint64)
        undirected = C.c_uint64()
        n = lib.pjs_run(V, E, _ptr(off, C.c_uint64), _ptr(ed, C.c_uint32), _ptr(glab, C.c_uint64), P, len(src),
                        _ptr(src, C.c_uint32), _ptr(dst, C.c_uint32), _ptr(lab, C.c_uint64), eid64, rows, 1000,
                        _ptr(va, C.c_uint8), _ptr(ea, C.c_uint8), _ptr(cand, C.c_uint64), C.byref(undirected))
        assert n > 0, name
        o = orc.Oracle(g, c, threads=1)
        orows, _ = o.run(orc.FAITHFUL)
        key = lambda r: (r.iteration, r.step, r.vertex_eliminated, r.edge_eliminated, r.match_eliminated,
                         r.enumeration, r.active_vertices, r.active_edges)
        print(f"[{name} eid64={eid64}] {n} steps through the seam", flush=True)
        assert [key(r) for r in rows[:n]] == [r.key() for r in orows], name
        ocand, ovf, oea, *_ = o.state()
        assert np.array_equal(va, ovf & 1), name
        assert np.array_equal(ea, np.unpackbits(oea.view(np.uint8), bitorder="little")[:E]), name
        assert np.array_equal(cand, np.where(ovf & 1, ocand, 0)), name
        assert undirected.value == 0, name  # symmetric edge activity between active vertices
