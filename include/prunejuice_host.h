/*
 * prunejuice_host.h — C ABI of the host-side library (libpj_host.so): TOTEM
 * graph loading and writing, synthetic R-MAT graphs, constraint generation
 * from a template, and the pipeline step machine driven over the device
 * C ABI of prunejuice.h.
 *
 * Reference counterparts:
 *   pjh_graph_load         graph_initialize + parseVertexFile
 *                          (src/totem/totem_graph.cu:388-409,
 *                           src/patternmatching/utils/common_utils.cc:9-59)
 *   pjh_graph_store_binary graph_store_binary (totem_graph.cu:883-923)
 *   pjh_constraints_build  the five generators run by MultipleLabelCpu::allocate
 *                          (src/patternmatching/multiple_label_cpu.cc:54-80)
 *   pjh_run_pipeline       MultipleLabelCpu::runPatternMatching
 *                          (multiple_label_cpu.cc:113-185)
 *   pjh_main               main of patternmatching_binary
 *                          (src/patternmatching/patternmatching_binary.cc:13-42)
 * All functions return 0 on success and -1 on failure unless stated.
 */
#ifndef PRUNEJUICE_HOST_H
#define PRUNEJUICE_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "prunejuice.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pjh_graph pjh_graph;
typedef struct pjh_constraints pjh_constraints;

/* Graphs ------------------------------------------------------------------ */
int pjh_graph_load(const char* path, const char* vertex_file /* may be NULL */, pjh_graph** out, char* err,
                   size_t errlen);
int pjh_graph_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, pjh_graph** out);
int pjh_graph_from_csr(uint32_t vertex_count, uint64_t edge_count, const uint64_t* offsets, const uint32_t* edges,
                       const uint64_t* labels /* may be NULL */, pjh_graph** out);
int pjh_pattern_from_edges(uint32_t vertex_count, uint32_t edge_count, const uint32_t* src, const uint32_t* dst,
                           const uint64_t* labels, pjh_graph** out);
void pjh_graph_labels_uniform(pjh_graph* g, uint32_t n_labels, uint64_t seed);
void pjh_graph_labels_zipf(pjh_graph* g, uint32_t n_labels, double s, uint64_t seed);
uint32_t pjh_graph_vertex_count(const pjh_graph* g);
uint64_t pjh_graph_edge_count(const pjh_graph* g);
const uint64_t* pjh_graph_offsets(const pjh_graph* g);
const uint32_t* pjh_graph_edges(const pjh_graph* g);
const uint64_t* pjh_graph_labels(const pjh_graph* g); /* NULL when not valued */
int pjh_graph_store_binary(const pjh_graph* g, const char* path, int eid_bytes);
int pjh_graph_store_text(const pjh_graph* g, const char* path);
int pjh_graph_store_vertex_file(const pjh_graph* g, const char* path);
uint64_t pjh_graph_asymmetric_slots(const pjh_graph* g);
void pjh_graph_free(pjh_graph* g);

/* Constraints ------------------------------------------------------------- */
int pjh_constraints_build(const pjh_graph* pattern, pjh_constraints** out, char* err, size_t errlen);
const pj_template_desc* pjh_constraints_desc(const pjh_constraints* c);
/* Writes the constraint dump (NUL-terminated, truncated to len); returns the
 * full length. */
size_t pjh_constraints_dump(const pjh_constraints* c, char* buf, size_t len);
/* 1 when the reference's per-source edge memo of this cycle / path
 * constraint can never fire (then its memoised search equals the complete
 * search the device runs), 0 when it may fire, -1 for other steps. */
int pjh_constraints_memo_inert(const pjh_constraints* c, int step, uint32_t index);
void pjh_constraints_free(pjh_constraints* c);

/* Pipeline ---------------------------------------------------------------- */
typedef struct pjh_row {
  int32_t iteration;
  int32_t step;
  uint64_t vertex_eliminated, edge_eliminated, match_eliminated, enumeration;
  uint64_t active_vertices, active_edges;
  uint64_t slots; /* CSR slots examined (pj_last_step_slots) */
  double ms;      /* device time of the step (HIP events) */
  int32_t constraint;   /* cycle / path / TDS constraint the step ran, else -1 */
  int32_t reserved;
  double effectiveness; /* its estimate, when the dynamic order chose it */
  double stat_ms;       /* device time of the label statistics taken after the step (dynamic order) */
} pjh_row;

/* Runs the step machine on a context that holds the graph and the
 * constraints' template (pj_set_template done).  Resets the state first.
 * With PJH_STOP_BEFORE_ENUMERATION in `flags` the machine stops when it
 * reaches the enumeration step (the pruning fixed point) without running it.
 * Returns the number of steps (rows written up to max_rows), -1 on error. */
#define PJH_STOP_BEFORE_ENUMERATION 1
/* With PJH_STOP_BEFORE_TDS it stops when the first TDS step is due (used to
 * measure the other phases on graphs where the memoised TDS search of the
 * reference does not finish in reasonable time). */
#define PJH_STOP_BEFORE_TDS 2
/* With PJH_ORDER_DYNAMIC the constraints run in the order of the reference's
 * effectiveness estimate (test_article/test_order_multiple_label_cpu.cc:
 * 170-290): at each LCC fixed point the label statistics of the active graph
 * (pj_graph_stat) score every cycle, path and TDS constraint not yet run and
 * the best runs next, once.  The cost divisor of the estimate (the
 * reference's omp_get_max_threads()) is 1; it does not change the order. */
#define PJH_ORDER_DYNAMIC 4
int pjh_run_pipeline(pj_ctx* ctx, const pjh_constraints* c, int flags, pjh_row* rows, int max_rows,
                     double* total_ms);

/* Batched and served forms ------------------------------------------------
 * Runs n templates over the graph of `owner` with `workers` contexts sharing
 * it (pj_share_graph), one host thread and one stream each.  Query i may
 * start once the clock reaches arrival_ms[i] (NULL: all at 0), a worker is
 * free and, when budget_bytes > 0, the device bytes projected for the
 * queries in flight (pjh_projected_bytes) stay within the budget; queries
 * are admitted in index order.  Each query sets its template, runs the step
 * machine with `flags` (PJH_STOP_*) and takes pj_state_checksum.  Worker
 * contexts are created before the clock starts.  Returns 0, or -1 with the
 * message of the first failing query in err.  No reference counterpart: the
 * reference runs one template per process (multiple_label_cpu.cc:113-185). */
typedef struct pjh_query_result {
  int32_t steps;            /* rows of the solve */
  int32_t worker;           /* worker that ran it */
  uint64_t rows_hash;       /* hash of every row's iteration, step, counters and active counts */
  uint64_t state_checksum;  /* pj_state_checksum after the solve */
  uint64_t active_vertices, active_edges;
  uint64_t projected_bytes;
  double arrival_ms, start_ms, end_ms; /* on the batch clock */
  double device_ms;         /* summed step device time */
  double worker_ms;         /* when a worker context took it (start_ms: admission) */
  int32_t batch_lane;       /* batch lane its LCC0 and first LCC fixed point ran in, or -1 */
  int32_t reserved;
} pjh_query_result;
/* With PJH_BATCH_LCC in `flags` the LCC0 step and the LCC steps to the first
 * fixed point of every template the batch can hold (pj_batch_admit) run in a batch
 * (pj_batch, include/prunejuice.h) whose lanes take queries as they are
 * admitted; a query that reaches that fixed point is handed to a worker,
 * which installs its state and continues the step machine from there, and
 * its lane takes the next query.  Results are those of a context of its own.
 * Lanes bound the batch; the budget (projected bytes of the worker contexts'
 * state and memo tables) then holds from the moment a worker takes a query
 * (worker_ms) to its end. */
#define PJH_BATCH_LCC 8
int pjh_run_queries(pj_ctx* owner, const pjh_constraints* const* templates, int n, int workers, int flags,
                    const double* arrival_ms, uint64_t budget_bytes, pjh_query_result* out, double* total_ms,
                    char* err, size_t errlen);
/* Device bytes a query's context is projected to need beyond the shared
 * graph, mirroring the device library's allocations: the pruning state
 * (candidate, toUnmatch and staging words, flags, marks, scan lengths,
 * frontier bins, live lists, edge-active bits); with memoised walks (TDS,
 * cycle / path constraints whose memo may fire) alreadyMatched, overflow
 * lists, the first-pass memo tables and the active view; with complete walks
 * of 4+ vertices the memo tables of their failed-state memo.  The retry
 * tables are not per query: one set per device, shared by every context and
 * used under a lock, of PJ_RETRY_POOL_BYTES; pjh_run_queries holds them back
 * from the budget once when any query uses memo tables. */
#define PJ_RETRY_POOL_BYTES ((32ull << 22) * 20 + (4ull << 26) * 20 + (1ull << 29) * 20)
uint64_t pjh_projected_bytes(const pjh_constraints* c, uint32_t vertex_count, uint64_t edge_count);
/* 1 when a query of these constraints allocates memo tables (see above). */
int pjh_query_uses_memo(const pjh_constraints* c);

/* Effectiveness estimate ----------------------------------------------------
 * MultipleLabelConstraintEffectiveness::computeBacktrack
 * (src/patternmatching/multiple_label/multiple_label_constraint_effectiveness.cc:
 * 226-340 cycle, 570-696 path, 972-1118 TDS) of constraint `index` of kind
 * `step` (PJ_STEP_CC / PJ_STEP_PC / PJ_STEP_TDS) from label statistics of
 * the active graph (pj_graph_stat): out[0] approximateMatchEliminated,
 * out[1] approximateCostEliminated, out[2] approximateEffectiveness.
 * `threads` is the reference's omp_get_max_threads() cost divisor.  A label
 * or label pair without statistics reads 0 (the reference's map lookups
 * throw there).  Host arithmetic only. */
int pjh_constraint_effectiveness(const pj_label_stat* stat, const pjh_constraints* c, int step, uint32_t index,
                                 int threads, double out[3]);
/* The dynamic order's choice (src/patternmatching/test_article/
 * test_order_multiple_label_cpu.cc:184-274) among the constraints whose
 * done_* flag is 0 (NULL: none done): the greatest effectiveness, ties to
 * the earlier constraint and to cycles, then paths, then TDS; *step is
 * PJ_STEP_ENUMERATION when none is left. */
int pjh_next_constraint(const pj_label_stat* stat, const pjh_constraints* c, const uint8_t* done_cycles,
                        const uint8_t* done_paths, const uint8_t* done_tds, int threads, int* step, uint32_t* index,
                        double* effectiveness);

/* patternmatching_binary: -i graph -b binary-cache -v vertex-data -p pattern-dir
 * -o output-dir [-d device]; writes results.totem, graph_%04d.totem,
 * vertex_%04d.totem and log.totem.  Returns the process exit code. */
int pjh_main(int argc, char** argv);

#ifdef __cplusplus
}
#endif

#endif /* PRUNEJUICE_HOST_H */
