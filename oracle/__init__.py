"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke(), bench.py's
cpu_baseline).  ctypes wrapper of oracle/libpj_oracle.so, the CPU restatement
of the reference's pruning steps (see pj_oracle.cc)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "libpj_oracle.so"

FAITHFUL = 0   # reference DFS with memo + alreadyMatched shortcut
COMPLETE = 1   # complete search (no memo, no shortcut)
FAITHFUL_NO_SHORTCUT = 2  # memos on, every walk runs (bounds every schedule of FAITHFUL)


class OracleRow(C.Structure):
    _fields_ = [("iteration", C.c_int32), ("step", C.c_int32), ("vertex_eliminated", C.c_uint64),
                ("edge_eliminated", C.c_uint64), ("match_eliminated", C.c_uint64), ("enumeration", C.c_uint64),
                ("active_vertices", C.c_uint64), ("active_edges", C.c_uint64), ("slots", C.c_uint64),
                ("ms", C.c_double)]

    def key(self) -> tuple:
        return (self.iteration, self.step, self.vertex_eliminated, self.edge_eliminated, self.match_eliminated,
                self.enumeration, self.active_vertices, self.active_edges)


class _Results(C.Structure):
    _fields_ = [("vertex_eliminated", C.c_uint64), ("edge_eliminated", C.c_uint64),
                ("match_eliminated", C.c_uint64), ("enumeration", C.c_uint64)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"oracle not built: {LIB} (python oracle/build.py)")
        l = C.CDLL(str(LIB))
        P = C.c_void_p
        u8p, u32p, u64p = C.POINTER(C.c_uint8), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
        sigs = {
            "or_create": (P, []), "or_destroy": (None, [P]), "or_set_threads": (None, [P, C.c_int]),
            "or_set_graph": (C.c_int, [P, C.c_uint32, C.c_uint64, u64p, u32p, u64p]),
            "or_set_template": (C.c_int, [P, P]), "or_reset": (None, [P]), "or_reset_modified": (None, [P]),
            "or_run_step": (C.c_int, [P, C.c_int, C.c_uint32, C.c_int, C.POINTER(_Results)]),
            "or_get_state": (None, [P, u64p, u8p, u32p, u64p, u64p]),
            "or_set_state": (None, [P, u64p, u8p, u32p, C.c_uint64, C.c_uint64]),
            "or_run_pipeline": (C.c_int, [P, C.c_int, C.c_int, C.POINTER(OracleRow), C.c_int,
                                          C.POINTER(C.c_double)]),
            "or_last_slots": (C.c_uint64, [P]),
            "or_last_marks": (None, [P, u64p]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


class Oracle:
    """CPU pruning state over a Graph (pattenmatching_sharedmemory_amd.Graph)
    and a Constraints object.  The graph must outlive the oracle."""

    def __init__(self, graph, constraints, threads: int = 1):
        self.graph = graph
        self.constraints = constraints
        self._s = C.c_void_p(lib().or_create())
        off, ed, lab = graph.raw_arrays()
        lib().or_set_graph(self._s, graph.vertex_count, graph.edge_count, off, ed, lab)
        lib().or_set_template(self._s, C.cast(constraints.desc, C.c_void_p))
        lib().or_set_threads(self._s, threads)
        lib().or_reset(self._s)

    def __del__(self):
        s = getattr(self, "_s", None)
        if s is not None and s.value:
            lib().or_destroy(s)
            self._s = C.c_void_p(0)

    def set_threads(self, n: int) -> None:
        lib().or_set_threads(self._s, n)

    def reset(self) -> None:
        lib().or_reset(self._s)

    def reset_modified(self) -> None:
        lib().or_reset_modified(self._s)

    def step(self, step: int, index: int = 0, mode: int = COMPLETE) -> tuple[int, int, int, int]:
        r = _Results()
        if lib().or_run_step(self._s, step, index, mode, C.byref(r)):
            raise ValueError("bad step")
        return (r.vertex_eliminated, r.edge_eliminated, r.match_eliminated, r.enumeration)

    def last_marks(self):
        """alreadyMatched as the last NLCC step's phase 1 left it."""
        out = np.zeros(self.graph.vertex_count, dtype=np.uint64)
        lib().or_last_marks(self._s, out.ctypes.data_as(C.POINTER(C.c_uint64)))
        return out

    def last_slots(self) -> int:
        return lib().or_last_slots(self._s)

    def state(self):
        V, E = self.graph.vertex_count, self.graph.edge_count
        cand = np.zeros(V, dtype=np.uint64)
        vf = np.zeros(V, dtype=np.uint8)
        ea = np.zeros((E + 31) // 32, dtype=np.uint32)
        av, ae = C.c_uint64(), C.c_uint64()
        lib().or_get_state(self._s, cand.ctypes.data_as(C.POINTER(C.c_uint64)), vf.ctypes.data_as(C.POINTER(C.c_uint8)),
                           ea.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(av), C.byref(ae))
        return cand, vf, ea, av.value, ae.value

    def load_state(self, cand, vflags, ea_words, av: int, ae: int) -> None:
        cand = np.ascontiguousarray(cand, dtype=np.uint64)
        vflags = np.ascontiguousarray(vflags, dtype=np.uint8)
        ea_words = np.ascontiguousarray(ea_words, dtype=np.uint32)
        lib().or_set_state(self._s, cand.ctypes.data_as(C.POINTER(C.c_uint64)),
                           vflags.ctypes.data_as(C.POINTER(C.c_uint8)), ea_words.ctypes.data_as(C.POINTER(C.c_uint32)),
                           av, ae)

    def run(self, mode: int = COMPLETE, max_rows: int = 1000, enumerate_matches: bool = True, run_tds: bool = True):
        rows = (OracleRow * max_rows)()
        total = C.c_double()
        flags = (0 if enumerate_matches else 1) | (0 if run_tds else 2)
        n = lib().or_run_pipeline(self._s, mode, flags, rows, max_rows, C.byref(total))
        if n < 0:
            raise ValueError("oracle pipeline failed")
        return list(rows[: min(n, max_rows)]), total.value
