"""Python host API over the native libraries.

Graph        TOTEM graph (CSR) on the host: load / write / R-MAT / labels
Constraints  every constraint of a template (pattern) graph, generated on the host
Pruner       one device context: HBM-resident graph, template and pruning state;
             runs single steps (the reference's step classes) or the whole
             step machine (MultipleLabelCpu::runPatternMatching,
             src/patternmatching/multiple_label_cpu.cc:113-185)
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import native
from .native import STEP_CC, STEP_ENUMERATION, STEP_LCC, STEP_LCC0, STEP_PC, STEP_TDS, AlgoResults, Row


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Graph:
    """A CSR graph owned by the host library (pj::Graph)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            native.host().pjh_graph_free(h)
            self._h = C.c_void_p(0)

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @classmethod
    def load(cls, path: str, vertex_file: str | None = None) -> "Graph":
        out = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = native.host().pjh_graph_load(str(path).encode(), vertex_file.encode() if vertex_file else None,
                                          C.byref(out), err, len(err))
        if rc:
            raise ValueError(err.value.decode())
        return cls(out.value)

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1) -> "Graph":
        out = C.c_void_p()
        if native.host().pjh_graph_rmat(scale, edge_factor, a, b, c, seed, C.byref(out)):
            raise ValueError("bad R-MAT parameters")
        return cls(out.value)

    @classmethod
    def from_csr(cls, offsets, edges, labels=None) -> "Graph":
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        ed = np.ascontiguousarray(edges, dtype=np.uint32)
        lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.uint64)
        out = C.c_void_p()
        rc = native.host().pjh_graph_from_csr(len(off) - 1, len(ed), _ptr(off, C.c_uint64), _ptr(ed, C.c_uint32),
                                              None if lab is None else _ptr(lab, C.c_uint64), C.byref(out))
        if rc:
            raise ValueError("bad CSR")
        return cls(out.value)

    @classmethod
    def pattern(cls, vertex_count: int, edges, labels) -> "Graph":
        """Undirected template: each (a, b) stored both ways, rows ascending."""
        e = np.asarray(edges, dtype=np.uint32).reshape(-1, 2)
        src, dst = np.ascontiguousarray(e[:, 0]), np.ascontiguousarray(e[:, 1])
        lab = np.ascontiguousarray(labels, dtype=np.uint64)
        out = C.c_void_p()
        rc = native.host().pjh_pattern_from_edges(vertex_count, len(src), _ptr(src, C.c_uint32),
                                                  _ptr(dst, C.c_uint32), _ptr(lab, C.c_uint64), C.byref(out))
        if rc:
            raise ValueError("bad pattern")
        return cls(out.value)

    def labels_uniform(self, n_labels: int, seed: int) -> "Graph":
        native.host().pjh_graph_labels_uniform(self._h, n_labels, seed)
        return self

    def labels_zipf(self, n_labels: int, s: float, seed: int) -> "Graph":
        native.host().pjh_graph_labels_zipf(self._h, n_labels, s, seed)
        return self

    @property
    def vertex_count(self) -> int:
        return native.host().pjh_graph_vertex_count(self._h)

    @property
    def edge_count(self) -> int:
        return native.host().pjh_graph_edge_count(self._h)

    def offsets(self) -> np.ndarray:
        p = native.host().pjh_graph_offsets(self._h)
        return np.ctypeslib.as_array(p, shape=(self.vertex_count + 1,)).copy()

    def edges(self) -> np.ndarray:
        if self.edge_count == 0:
            return np.zeros(0, dtype=np.uint32)
        p = native.host().pjh_graph_edges(self._h)
        return np.ctypeslib.as_array(p, shape=(self.edge_count,)).copy()

    def labels(self) -> np.ndarray | None:
        p = native.host().pjh_graph_labels(self._h)
        if not p or self.vertex_count == 0:
            return None
        return np.ctypeslib.as_array(p, shape=(self.vertex_count,)).copy()

    def raw_arrays(self):
        """(offsets, edges, labels) pointers owned by the graph, for C calls."""
        h = native.host()
        return h.pjh_graph_offsets(self._h), h.pjh_graph_edges(self._h), h.pjh_graph_labels(self._h)

    def store_binary(self, path: str, eid_bytes: int = 4) -> None:
        if native.host().pjh_graph_store_binary(self._h, str(path).encode(), eid_bytes):
            raise OSError(f"cannot write {path}")

    def store_text(self, path: str) -> None:
        if native.host().pjh_graph_store_text(self._h, str(path).encode()):
            raise OSError(f"cannot write {path}")

    def store_vertex_file(self, path: str) -> None:
        if native.host().pjh_graph_store_vertex_file(self._h, str(path).encode()):
            raise OSError(f"cannot write {path}")

    def asymmetric_slots(self) -> int:
        return native.host().pjh_graph_asymmetric_slots(self._h)


class Constraints:
    """All constraints of a template, in the reference's generation order."""

    def __init__(self, pattern: Graph):
        out = C.c_void_p()
        err = C.create_string_buffer(512)
        if native.host().pjh_constraints_build(pattern.handle, C.byref(out), err, len(err)):
            raise ValueError(err.value.decode())
        self._h = out
        self.pattern = pattern

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            native.host().pjh_constraints_free(h)
            self._h = C.c_void_p(0)

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def desc(self):
        return native.host().pjh_constraints_desc(self._h)

    def dump(self) -> str:
        n = native.host().pjh_constraints_dump(self._h, None, 0)
        buf = C.create_string_buffer(n + 1)
        native.host().pjh_constraints_dump(self._h, buf, n + 1)
        return buf.value.decode()

    @property
    def vertex_count(self) -> int:
        return self.desc.contents.vertex_count

    def cycles(self) -> list[list[int]]:
        d = self.desc.contents
        return [[d.cycle_vertex[i] for i in range(d.cycle_offset[c], d.cycle_offset[c + 1])]
                for c in range(d.cycle_count)]

    def paths(self) -> list[list[int]]:
        d = self.desc.contents
        return [[d.path_vertex[i] for i in range(d.path_offset[c], d.path_offset[c + 1])]
                for c in range(d.path_count)]

    def path_labels(self) -> list[int]:
        d = self.desc.contents
        return [d.path_initial_label[c] for c in range(d.path_count)]

    def tds(self) -> list[list[tuple[int, list[tuple[int, int, int]]]]]:
        """Per TDS constraint: [(root, [(move, vertex, move_back), ...]), ...] in walk order."""
        d = self.desc.contents
        out = []
        for c in range(d.tds_count):
            out.append([(d.walks[k].root, d.walks[k].moves())
                        for k in range(d.tds_walk_offset[c], d.tds_walk_offset[c + 1])])
        return out

    def enumeration_walk(self) -> tuple[int, list[tuple[int, int, int]]]:
        w = self.desc.contents.enumeration_walk
        return w.root, w.moves()

    def local(self) -> list[dict]:
        d = self.desc.contents
        out = []
        for p in range(d.vertex_count):
            need = {d.need_label[i]: d.need_count[i] for i in range(d.need_offset[p], d.need_offset[p + 1])}
            out.append({"label": d.vertex_label[p], "need": need, "neighbors": d.neighbor_mask[p]})
        return out

    def memo_inert(self, step: int, index: int) -> int:
        return native.host().pjh_constraints_memo_inert(self._h, step, index)

    def effectiveness(self, stat: "LabelStat", step: int, index: int, threads: int = 1) -> tuple[float, float, float]:
        """(approximateMatchEliminated, approximateCostEliminated,
        approximateEffectiveness) of one cycle / path / TDS constraint
        (pjh_constraint_effectiveness)."""
        out = (C.c_double * 3)()
        if native.host().pjh_constraint_effectiveness(C.byref(stat.desc), self._h, step, index, threads, out):
            raise ValueError(f"no constraint {index} of step {step}")
        return out[0], out[1], out[2]

    def next_constraint(self, stat: "LabelStat", done_cycles=None, done_paths=None, done_tds=None,
                        threads: int = 1) -> tuple[int, int, float]:
        """The dynamic order's choice among the constraints not done
        (pjh_next_constraint): (step, index, effectiveness)."""
        def flags(done, n):
            if done is None:
                return None
            a = np.zeros(max(n, 1), dtype=np.uint8)
            for i in done:
                a[i] = 1
            return _ptr(a, C.c_uint8), a

        keep = [flags(done_cycles, len(self.cycles())), flags(done_paths, len(self.paths())),
                flags(done_tds, len(self.tds()))]
        step, index, eff = C.c_int(), C.c_uint32(), C.c_double()
        if native.host().pjh_next_constraint(C.byref(stat.desc), self._h, *[k[0] if k else None for k in keep],
                                             threads, C.byref(step), C.byref(index), C.byref(eff)):
            raise ValueError("pjh_next_constraint failed")
        return step.value, index.value, eff.value


class LabelStat:
    """pj_label_stat as arrays: labels (template label per code), vertex[K],
    edge[K, K], outbound[K], at_least_max[K, K] and at_least[(f, t)] =
    counts for k = 1..at_least_max[f, t] (codes f, t)."""

    def __init__(self, desc: native.LabelStatDesc):
        self.desc = desc
        K = desc.label_count
        self.labels = [desc.label[k] for k in range(K)]
        self.vertex = np.array(desc.vertex[:K], dtype=np.uint64)
        self.outbound = np.array(desc.outbound[:K], dtype=np.uint64)
        e = np.ctypeslib.as_array(desc.edge).reshape(64, 64)
        self.edge = e[:K, :K].astype(np.uint64)
        mx = np.ctypeslib.as_array(desc.at_least_max).reshape(64, 64)
        self.at_least_max = mx[:K, :K].astype(np.uint32)
        off = np.ctypeslib.as_array(desc.at_least_offset).reshape(64, 64)
        al = np.ctypeslib.as_array(desc.at_least)
        self.at_least = {(f, t): al[off[f, t]: off[f, t] + mx[f, t]].astype(np.uint64).copy()
                         for f in range(K) for t in range(K) if mx[f, t]}
        self.vertex_total = desc.vertex_total
        self.edge_total = desc.edge_total

    @classmethod
    def from_arrays(cls, labels, vertex, edge, at_least_max, at_least: dict) -> "LabelStat":
        """A pj_label_stat built from host arrays (codes in `labels` order)."""
        d = native.LabelStatDesc()
        K = len(labels)
        d.label_count = K
        off = 0
        for f in range(K):
            d.label[f] = int(labels[f])
            d.vertex[f] = int(vertex[f])
            for t in range(K):
                d.edge[f * 64 + t] = int(edge[f][t])
                d.outbound[f] += int(edge[f][t])
                m = int(at_least_max[f][t])
                d.at_least_max[f * 64 + t] = m
                d.at_least_offset[f * 64 + t] = off
                for k in range(m):
                    d.at_least[off + k] = int(at_least[(f, t)][k])
                off += m
            d.vertex_total += int(vertex[f])
            d.edge_total += d.outbound[f]
        d.at_least_count = off
        return cls(d)


@dataclass
class State:
    cand: np.ndarray      # uint64 [V]
    vflags: np.ndarray    # uint8 [V]: bit0 active, bit1 modified, bit2 scheduled
    ea_words: np.ndarray  # uint32 [ceil(E/32)]
    active_vertices: int
    active_edges: int

    def edge_active(self, edge_count: int) -> np.ndarray:
        bits = np.unpackbits(self.ea_words.view(np.uint8), bitorder="little")
        return bits[:edge_count].astype(bool)


class Pruner:
    """One device context (pj_ctx) with a resident graph and template."""

    def __init__(self, graph: Graph | None, constraints: Constraints | None = None, device: int = 0,
                 path: str | None = None, labels=None):
        """A context over `graph` (uploaded from host memory), or over the
        binary TOTEM file `path` streamed into HBM (pj_upload_totem_file;
        `labels` replaces the file's vertex values)."""
        lib = native.hip()
        ctx = C.c_void_p()
        if lib.pj_create(C.byref(ctx), device):
            raise RuntimeError("pj_create failed (no HIP device?)")
        self._ctx = ctx
        self.graph = graph
        if path is not None:
            lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.uint64)
            self._check(lib.pj_upload_totem_file(ctx, str(path).encode(),
                                                 None if lab is None else _ptr(lab, C.c_uint64)))
            self.vertex_count, self.edge_count = self.counters()
        else:
            off, ed, lab = graph.raw_arrays()
            if not lab:
                raise ValueError("graph has no labels")
            self._check(lib.pj_upload_graph(ctx, graph.vertex_count, graph.edge_count, off, ed, lab))
            self.vertex_count, self.edge_count = graph.vertex_count, graph.edge_count
        self.constraints = None
        if constraints is not None:
            self.set_template(constraints)

    def _check(self, rc: int) -> None:
        if rc:
            raise RuntimeError(native.hip().pj_last_error(self._ctx).decode())

    def close(self) -> None:
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            native.hip().pj_destroy(self._ctx)
            self._ctx = C.c_void_p(0)

    def __del__(self):
        self.close()

    @property
    def ctx(self) -> C.c_void_p:
        return self._ctx

    def set_template(self, constraints: Constraints) -> None:
        self.constraints = constraints
        self._check(native.hip().pj_set_template(self._ctx, constraints.desc))

    def reset(self) -> None:
        self._check(native.hip().pj_state_reset(self._ctx))

    def reset_modified(self) -> None:
        self._check(native.hip().pj_reset_modified(self._ctx))

    def step(self, step: int, index: int = 0) -> AlgoResults:
        r = AlgoResults()
        self._check(native.hip().pj_run_step(self._ctx, step, index, C.byref(r)))
        return r

    def last_ms(self) -> float:
        ms = C.c_double()
        native.hip().pj_last_step_ms(self._ctx, C.byref(ms))
        return ms.value

    def kernel_timing(self, enable: bool, only: str | None = None) -> None:
        """Per-kernel HIP-event timing of later steps (clears the totals);
        `only` names the one kernel kind to time."""
        if enable and only:
            self._check(native.hip().pj_set_kernel_timing_kind(self._ctx, only.encode()))
        else:
            self._check(native.hip().pj_set_kernel_timing(self._ctx, 1 if enable else 0))

    def kernel_stats(self) -> list[dict]:
        n = native.hip().pj_kernel_stats(self._ctx, None, 0)
        arr = (native.KernelStat * max(n, 1))()
        n = native.hip().pj_kernel_stats(self._ctx, arr, n)
        return [{"name": a.name.decode(), "launches": a.launches, "ms": a.ms, "items": a.items, "slots": a.slots,
                 "edges_eliminated": a.edges_eliminated, "word_bytes": a.word_bytes} for a in arr[:n]]

    def last_slots(self) -> int:
        n = C.c_uint64()
        native.hip().pj_last_step_slots(self._ctx, C.byref(n))
        return n.value

    def counters(self) -> tuple[int, int]:
        av, ae = C.c_uint64(), C.c_uint64()
        native.hip().pj_get_counters(self._ctx, C.byref(av), C.byref(ae))
        return av.value, ae.value

    def state(self) -> State:
        V, E = self.vertex_count, self.edge_count
        cand = np.zeros(V, dtype=np.uint64)
        vf = np.zeros(V, dtype=np.uint8)
        ea = np.zeros((E + 31) // 32, dtype=np.uint32)
        self._check(native.hip().pj_download_state(self._ctx, _ptr(cand, C.c_uint64), _ptr(vf, C.c_uint8),
                                                   _ptr(ea, C.c_uint32)))
        av, ae = self.counters()
        return State(cand, vf, ea, av, ae)

    def active_graph(self):
        """The surviving graph compacted on the device (pj_download_active):
        (vertices, candidate words, slot sources, slot destinations), vertices
        ascending, slots in CSR order."""
        lib = native.hip()
        nv, ne = C.c_uint64(), C.c_uint64()
        self._check(lib.pj_download_active(self._ctx, None, None, 0, None, None, 0, C.byref(nv), C.byref(ne)))
        v = np.zeros(nv.value, dtype=np.uint32)
        c = np.zeros(nv.value, dtype=np.uint64)
        s = np.zeros(ne.value, dtype=np.uint32)
        d = np.zeros(ne.value, dtype=np.uint32)
        self._check(lib.pj_download_active(self._ctx, v.ctypes.data, c.ctypes.data, nv.value, s.ctypes.data,
                                           d.ctypes.data, ne.value, C.byref(nv), C.byref(ne)))
        return v, c, s, d

    def graph_stat(self) -> LabelStat:
        """Label statistics of the active graph (pj_graph_stat)."""
        d = native.LabelStatDesc()
        self._check(native.hip().pj_graph_stat(self._ctx, C.byref(d)))
        return LabelStat(d)

    def checksum(self) -> int:
        """pj_state_checksum of the current state."""
        h = C.c_uint64()
        self._check(native.hip().pj_state_checksum(self._ctx, C.byref(h)))
        return h.value

    def run_queries(self, templates: list, workers: int = 4, enumerate_matches: bool = False,
                    run_tds: bool = True, arrivals_ms=None, budget_bytes: int = 0,
                    batch_lcc: bool = False) -> tuple[list, float]:
        """Batched / served form (pjh_run_queries): every template is solved
        over this context's resident graph by `workers` contexts sharing it;
        with batch_lcc the LCC0 and first LCC fixed point of templates of at
        most 16 vertices run in shared sweeps (PJH_BATCH_LCC).
        Returns (QueryResult per template, total wall ms)."""
        n = len(templates)
        handles = (C.c_void_p * max(n, 1))(*[t.handle.value for t in templates])
        out = (native.QueryResult * max(n, 1))()
        arr = None
        if arrivals_ms is not None:
            arr = (C.c_double * max(n, 1))(*[float(a) for a in arrivals_ms])
        total = C.c_double()
        err = C.create_string_buffer(512)
        flags = (0 if enumerate_matches else 1) | (0 if run_tds else 2) | (8 if batch_lcc else 0)
        rc = native.host().pjh_run_queries(self._ctx, handles, n, workers, flags, arr, budget_bytes, out,
                                           C.byref(total), err, 512)
        if rc:
            raise RuntimeError(err.value.decode())
        return list(out[:n]), total.value

    def load_state(self, st: State) -> None:
        cand = np.ascontiguousarray(st.cand, dtype=np.uint64)
        vf = np.ascontiguousarray(st.vflags, dtype=np.uint8)
        ea = np.ascontiguousarray(st.ea_words, dtype=np.uint32)
        self._check(native.hip().pj_upload_state(self._ctx, _ptr(cand, C.c_uint64), _ptr(vf, C.c_uint8),
                                                 _ptr(ea, C.c_uint32), st.active_vertices, st.active_edges))

    def run(self, max_rows: int = 1000, enumerate_matches: bool = True,
            run_tds: bool = True, order: str = "static") -> tuple[list[Row], float]:
        """Whole step machine; returns (rows, total device ms).  Without
        enumerate_matches it stops at the pruning fixed point; without run_tds
        it stops when the first TDS step is due.  order="dynamic" runs the
        constraints in the order of the effectiveness estimate (a statistics
        pass at every LCC fixed point; rows carry the chosen constraint, its
        estimate and the pass's device time)."""
        if order not in ("static", "dynamic"):
            raise ValueError(f"unknown order {order!r}")
        rows = (Row * max_rows)()
        total = C.c_double()
        flags = (0 if enumerate_matches else 1) | (0 if run_tds else 2) | (4 if order == "dynamic" else 0)
        n = native.host().pjh_run_pipeline(self._ctx, self.constraints.handle, flags, rows, max_rows,
                                           C.byref(total))
        if n < 0:
            self._check(1)
        return list(rows[: min(n, max_rows)]), total.value


class Batch:
    """Batched LCC (pj_batch): LCC0 and the LCC steps to the first fixed
    point of many templates in sweeps they share, over the graph resident in
    `owner`.  A template that reaches that fixed point is parked; its state
    can be exported into a context with the same template installed."""

    def __init__(self, owner: Pruner, lanes: int = 64):
        lib = native.hip()
        b = C.c_void_p()
        rc = lib.pj_batch_create(C.byref(b), owner.ctx, lanes)
        self._b = b
        self.owner = owner
        if rc:
            msg = lib.pj_batch_last_error(b).decode() if b.value else "pj_batch_create failed"
            self.close()
            raise RuntimeError(msg)
        self.constraints: dict[int, Constraints] = {}

    def _check(self, rc: int) -> None:
        if rc:
            raise RuntimeError(native.hip().pj_batch_last_error(self._b).decode())

    def close(self) -> None:
        if getattr(self, "_b", None) is not None and self._b.value:
            native.hip().pj_batch_destroy(self._b)
            self._b = C.c_void_p(0)

    def __del__(self):
        self.close()

    def admit(self, constraints: Constraints, strict: bool = True) -> int | None:
        """The lane the template takes.  A template the batch cannot hold
        (pj_batch_admit returns -2) raises, or gives None when not strict."""
        lane = C.c_uint32()
        rc = native.hip().pj_batch_admit(self._b, constraints.desc, C.byref(lane))
        if rc == -2 and not strict:
            return None
        self._check(rc)
        self.constraints[lane.value] = constraints
        return lane.value

    def run(self, max_steps: int = 1000) -> list[int]:
        """Steps until some lanes reach their fixed point; returns them (parked)."""
        done = (C.c_uint32 * 4096)()
        n = C.c_uint32()
        self._check(native.hip().pj_batch_run(self._b, max_steps, done, 4096, C.byref(n)))
        return list(done[: n.value])

    def run_all(self) -> list[int]:
        """Steps until no lane runs; the parked lanes in the order they finished."""
        out = []
        while self.info()["running"]:
            out += self.run()
        return out

    def trace(self, lane: int) -> list[tuple[int, int, int, int]]:
        n = C.c_uint32()
        self._check(native.hip().pj_batch_trace(self._b, lane, None, 0, C.byref(n)))
        rows = (AlgoResults * max(n.value, 1))()
        self._check(native.hip().pj_batch_trace(self._b, lane, rows, n.value, C.byref(n)))
        return [r.as_tuple() for r in rows[: n.value]]

    def export(self, lane: int, pruner: Pruner) -> None:
        self._check(native.hip().pj_batch_export(self._b, lane, pruner.ctx))

    def release(self, lane: int) -> None:
        self._check(native.hip().pj_batch_release(self._b, lane))
        self.constraints.pop(lane, None)

    def info(self) -> dict:
        lanes, running, parked = C.c_uint32(), C.c_uint32(), C.c_uint32()
        ms, steps = C.c_double(), C.c_uint64()
        self._check(native.hip().pj_batch_info(self._b, C.byref(lanes), C.byref(running), C.byref(parked),
                                               C.byref(ms), C.byref(steps)))
        return {"lanes": lanes.value, "running": running.value, "parked": parked.value, "device_ms": ms.value,
                "steps": steps.value}


def run_cli(args: list[str]) -> int:
    """patternmatching_binary in-process (writes results.totem etc.)."""
    argv = (C.c_char_p * (len(args) + 1))(b"patternmatching_hip", *[a.encode() for a in args])
    return native.host().pjh_main(len(args) + 1, argv)


__all__ = ["Graph", "Constraints", "LabelStat", "Pruner", "State", "run_cli", "STEP_LCC0", "STEP_LCC", "STEP_CC", "STEP_PC",
           "STEP_TDS", "STEP_ENUMERATION"]
