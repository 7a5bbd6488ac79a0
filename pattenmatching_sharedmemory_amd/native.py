"""ctypes bindings of the two native libraries.

``libprunejuice_hip.so`` exports the device C ABI of include/prunejuice.h and
``libpj_host.so`` the host C ABI of include/prunejuice_host.h.  Loading fails
loudly when the built libraries are missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"
HIP_LIB = LIB_DIR / "libprunejuice_hip.so"
HOST_LIB = LIB_DIR / "libpj_host.so"
CLI = LIB_DIR / "patternmatching_hip"

# enum pj_step (include/prunejuice.h)
STEP_START, STEP_END, STEP_LCC0, STEP_LCC, STEP_CC, STEP_PC, STEP_TDS, STEP_ENUMERATION = range(8)
STEP_NAMES = {STEP_START: "START", STEP_END: "END", STEP_LCC0: "LCC0", STEP_LCC: "LCC", STEP_CC: "CC",
              STEP_PC: "PC", STEP_TDS: "TDS", STEP_ENUMERATION: "ENUMERATION"}
# enum pj_move
MOVE_BACK, MOVE_CHECK, MOVE_VERTEX_STORE, MOVE_VERTEX_NO_STORE = range(4)


class AlgoResults(C.Structure):
    _fields_ = [("vertex_eliminated", C.c_uint64), ("edge_eliminated", C.c_uint64),
                ("match_eliminated", C.c_uint64), ("enumeration", C.c_uint64)]

    def as_tuple(self) -> tuple[int, int, int, int]:
        return (self.vertex_eliminated, self.edge_eliminated, self.match_eliminated, self.enumeration)

    def is_empty(self) -> bool:
        return not any(self.as_tuple())


class WalkDesc(C.Structure):
    _fields_ = [("root", C.c_uint32), ("length", C.c_uint32), ("move", C.POINTER(C.c_uint8)),
                ("vertex", C.POINTER(C.c_uint32)), ("move_back", C.POINTER(C.c_uint32))]

    def moves(self) -> list[tuple[int, int, int]]:
        return [(self.move[i], self.vertex[i], self.move_back[i]) for i in range(self.length)]


class TemplateDesc(C.Structure):
    _fields_ = [
        ("vertex_count", C.c_uint32), ("vertex_label", C.POINTER(C.c_uint64)),
        ("neighbor_mask", C.POINTER(C.c_uint64)), ("need_offset", C.POINTER(C.c_uint32)),
        ("need_label", C.POINTER(C.c_uint64)), ("need_count", C.POINTER(C.c_uint32)),
        ("cycle_count", C.c_uint32), ("cycle_offset", C.POINTER(C.c_uint32)), ("cycle_vertex", C.POINTER(C.c_uint32)),
        ("path_count", C.c_uint32), ("path_offset", C.POINTER(C.c_uint32)), ("path_vertex", C.POINTER(C.c_uint32)),
        ("path_initial_label", C.POINTER(C.c_uint64)),
        ("tds_count", C.c_uint32), ("tds_walk_offset", C.POINTER(C.c_uint32)), ("walk_count", C.c_uint32),
        ("walks", C.POINTER(WalkDesc)), ("enumeration_walk", WalkDesc),
    ]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64), ("ms", C.c_double), ("items", C.c_uint64),
                ("slots", C.c_uint64), ("edges_eliminated", C.c_uint64), ("word_bytes", C.c_uint32)]


class Row(C.Structure):
    """pjh_row: one step of the pipeline."""
    _fields_ = [("iteration", C.c_int32), ("step", C.c_int32), ("vertex_eliminated", C.c_uint64),
                ("edge_eliminated", C.c_uint64), ("match_eliminated", C.c_uint64), ("enumeration", C.c_uint64),
                ("active_vertices", C.c_uint64), ("active_edges", C.c_uint64), ("slots", C.c_uint64),
                ("ms", C.c_double), ("constraint", C.c_int32), ("reserved", C.c_int32),
                ("effectiveness", C.c_double), ("stat_ms", C.c_double)]

    def key(self) -> tuple:
        """The step's results (what the reference's results.totem row holds,
        times excluded).  Examined-slot counts are not part of it: they
        measure how much a given algorithm read, not what it decided."""
        return (self.iteration, self.step, self.vertex_eliminated, self.edge_eliminated, self.match_eliminated,
                self.enumeration, self.active_vertices, self.active_edges)


STAT_MAX_AT_LEAST = 4096


class LabelStatDesc(C.Structure):
    """pj_label_stat: label statistics of the active graph (pj_graph_stat)."""
    _fields_ = [("label_count", C.c_uint32), ("at_least_count", C.c_uint32), ("vertex_total", C.c_uint64),
                ("edge_total", C.c_uint64), ("label", C.c_uint64 * 64), ("vertex", C.c_uint64 * 64),
                ("outbound", C.c_uint64 * 64), ("edge", C.c_uint64 * 4096), ("at_least_max", C.c_uint32 * 4096),
                ("at_least_offset", C.c_uint32 * 4096), ("at_least", C.c_uint64 * STAT_MAX_AT_LEAST)]


class QueryResult(C.Structure):
    """pjh_query_result: one template of a batched / served run."""
    _fields_ = [("steps", C.c_int32), ("worker", C.c_int32), ("rows_hash", C.c_uint64),
                ("state_checksum", C.c_uint64), ("active_vertices", C.c_uint64), ("active_edges", C.c_uint64),
                ("projected_bytes", C.c_uint64), ("arrival_ms", C.c_double), ("start_ms", C.c_double),
                ("end_ms", C.c_double), ("device_ms", C.c_double), ("worker_ms", C.c_double),
                ("batch_lane", C.c_int32), ("reserved", C.c_int32)]

    def result(self) -> tuple:
        """What the solve decided (its rows and final state), times excluded."""
        return (self.steps, self.rows_hash, self.state_checksum, self.active_vertices, self.active_edges)


_hip = None
_host = None

P = C.c_void_p
u8p, u32p, u64p = C.POINTER(C.c_uint8), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)

HIP_SIGNATURES = {
    "pj_create": (C.c_int, [C.POINTER(P), C.c_int]),
    "pj_upload_graph": (C.c_int, [P, C.c_uint32, C.c_uint64, u64p, u32p, u64p]),
    "pj_upload_graph_device": (C.c_int, [P, C.c_uint32, C.c_uint64, P, P, P]),
    "pj_set_template": (C.c_int, [P, C.POINTER(TemplateDesc)]),
    "pj_state_reset": (C.c_int, [P]),
    "pj_reset_modified": (C.c_int, [P]),
    "pj_run_step": (C.c_int, [P, C.c_int, C.c_uint32, C.POINTER(AlgoResults)]),
    "pj_last_step_ms": (C.c_int, [P, C.POINTER(C.c_double)]),
    "pj_last_step_slots": (C.c_int, [P, u64p]),
    "pj_set_kernel_timing": (C.c_int, [P, C.c_int]),
    "pj_set_kernel_timing_kind": (C.c_int, [P, C.c_char_p]),
    "pj_kernel_stats": (C.c_int, [P, C.POINTER(KernelStat), C.c_int]),
    "pj_get_counters": (C.c_int, [P, u64p, u64p]),
    "pj_download_state": (C.c_int, [P, u64p, u8p, u32p]),
    "pj_upload_state": (C.c_int, [P, u64p, u8p, u32p, C.c_uint64, C.c_uint64]),
    "pj_state_checksum": (C.c_int, [P, u64p]),
    "pj_share_graph": (C.c_int, [P, P]),
    "pj_get_device": (C.c_int, [P, C.POINTER(C.c_int)]),
    "pj_last_error": (C.c_char_p, [P]),
    "pj_upload_totem_file": (C.c_int, [P, C.c_char_p, u64p]),
    "pj_compact_active": (C.c_int, [P, P, P, C.c_uint64, P, P, C.c_uint64, u64p, u64p]),
    "pj_compact_active_noalloc": (C.c_int, [P, P, P, C.c_uint64, P, P, C.c_uint64, u64p, u64p]),
    "pj_download_active": (C.c_int, [P, P, P, C.c_uint64, P, P, C.c_uint64, u64p, u64p]),
    "pj_graph_stat": (C.c_int, [P, C.POINTER(LabelStatDesc)]),
    "pj_destroy": (None, [P]),
    "pj_batch_create": (C.c_int, [C.POINTER(P), P, C.c_uint32]),
    "pj_batch_admit": (C.c_int, [P, C.POINTER(TemplateDesc), u32p]),
    "pj_batch_run": (C.c_int, [P, C.c_uint32, u32p, C.c_uint32, u32p]),
    "pj_batch_trace": (C.c_int, [P, C.c_uint32, C.POINTER(AlgoResults), C.c_uint32, u32p]),
    "pj_batch_export": (C.c_int, [P, C.c_uint32, P]),
    "pj_batch_release": (C.c_int, [P, C.c_uint32]),
    "pj_batch_info": (C.c_int, [P, u32p, u32p, u32p, C.POINTER(C.c_double), u64p]),
    "pj_batch_last_error": (C.c_char_p, [P]),
    "pj_batch_destroy": (None, [P]),
}

HOST_SIGNATURES = {
    "pjh_graph_load": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(P), C.c_char_p, C.c_size_t]),
    "pjh_graph_rmat": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64, C.POINTER(P)]),
    "pjh_graph_from_csr": (C.c_int, [C.c_uint32, C.c_uint64, u64p, u32p, u64p, C.POINTER(P)]),
    "pjh_pattern_from_edges": (C.c_int, [C.c_uint32, C.c_uint32, u32p, u32p, u64p, C.POINTER(P)]),
    "pjh_graph_labels_uniform": (None, [P, C.c_uint32, C.c_uint64]),
    "pjh_graph_labels_zipf": (None, [P, C.c_uint32, C.c_double, C.c_uint64]),
    "pjh_graph_vertex_count": (C.c_uint32, [P]),
    "pjh_graph_edge_count": (C.c_uint64, [P]),
    "pjh_graph_offsets": (u64p, [P]),
    "pjh_graph_edges": (u32p, [P]),
    "pjh_graph_labels": (u64p, [P]),
    "pjh_graph_store_binary": (C.c_int, [P, C.c_char_p, C.c_int]),
    "pjh_graph_store_text": (C.c_int, [P, C.c_char_p]),
    "pjh_graph_store_vertex_file": (C.c_int, [P, C.c_char_p]),
    "pjh_graph_asymmetric_slots": (C.c_uint64, [P]),
    "pjh_graph_free": (None, [P]),
    "pjh_constraints_build": (C.c_int, [P, C.POINTER(P), C.c_char_p, C.c_size_t]),
    "pjh_constraints_desc": (C.POINTER(TemplateDesc), [P]),
    "pjh_constraints_dump": (C.c_size_t, [P, C.c_char_p, C.c_size_t]),
    "pjh_constraints_memo_inert": (C.c_int, [P, C.c_int, C.c_uint32]),
    "pjh_constraints_free": (None, [P]),
    "pjh_run_pipeline": (C.c_int, [P, P, C.c_int, C.POINTER(Row), C.c_int, C.POINTER(C.c_double)]),
    "pjh_run_queries": (C.c_int, [P, C.POINTER(P), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_uint64,
                                  C.POINTER(QueryResult), C.POINTER(C.c_double), C.c_char_p, C.c_size_t]),
    "pjh_projected_bytes": (C.c_uint64, [P, C.c_uint32, C.c_uint64]),
    "pjh_query_uses_memo": (C.c_int, [P]),
    "pjh_constraint_effectiveness": (C.c_int, [C.POINTER(LabelStatDesc), P, C.c_int, C.c_uint32, C.c_int,
                                               C.POINTER(C.c_double)]),
    "pjh_next_constraint": (C.c_int, [C.POINTER(LabelStatDesc), P, u8p, u8p, u8p, C.c_int, C.POINTER(C.c_int),
                                      u32p, C.POINTER(C.c_double)]),
    "pjh_main": (C.c_int, [C.c_int, C.POINTER(C.c_char_p)]),
}


def _bind(lib: C.CDLL, sigs: dict) -> None:
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def hip() -> C.CDLL:
    """The device library; raises when it has not been built."""
    global _hip
    if _hip is None:
        if not HIP_LIB.exists():
            raise RuntimeError(f"HIP library missing: {HIP_LIB} (run pattenmatching_sharedmemory_amd.build)")
        lib = C.CDLL(str(HIP_LIB), mode=C.RTLD_GLOBAL)
        _bind(lib, HIP_SIGNATURES)
        _hip = lib
    return _hip


def host() -> C.CDLL:
    global _host
    if _host is None:
        hip()
        if not HOST_LIB.exists():
            raise RuntimeError(f"host library missing: {HOST_LIB} (run pattenmatching_sharedmemory_amd.build)")
        lib = C.CDLL(str(HOST_LIB))
        _bind(lib, HOST_SIGNATURES)
        _host = lib
    return _host
