/*
 * prunejuice.h — C ABI of the MI355X (gfx950) PruneJuice pruning library.
 *
 * This is the drop-in boundary between host C++ (the step machine, constraint
 * generation and TOTEM loader) and the HIP kernels.  Plain pointers and sizes
 * only; every call returns 0 on success and -1 on failure (the reference's
 * error_t convention, src/totem/totem_comdef.h:31-32), with the message kept
 * per context and returned by pj_last_error().
 *
 * Reference seam replaced: the step classes chosen by the typedef block at
 * src/patternmatching/multiple_label_cpu.h:34-52, each exposing
 *   AlgoResults compute(const graph_t&, State*)      (e.g. multiple_label_lcc_cpu.h:23)
 *   void resetState(State*)                          (multiple_label_cc_cpu.h:28)
 *   void setConstraintVector(const std::vector<C>&)  (multiple_label_cc_cpu.h:30-31)
 *   void setConstraintIterator(const size_t&)        (multiple_label_cc_cpu.h:32)
 * and the host state object MultipleLabelGlobalState
 * (src/patternmatching/multiple_label/multiple_label_common_cpu.h:23-59).
 *
 * Calls into one context must be serialised by the caller.  A context owns
 * one device CSR, one template and one pruning state on one HIP stream.
 */
#ifndef PRUNEJUICE_H
#define PRUNEJUICE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PJ_MAX_TEMPLATE_VERTICES 64

/* Step identifiers: same numbering as MultipleLabelStep::Step
 * (src/patternmatching/multiple_label_step.h:16-25). */
enum pj_step {
  PJ_STEP_START = 0,
  PJ_STEP_END = 1,
  PJ_STEP_LCC0 = 2,
  PJ_STEP_LCC = 3,
  PJ_STEP_CC = 4,
  PJ_STEP_PC = 5,
  PJ_STEP_TDS = 6,
  PJ_STEP_ENUMERATION = 7
};

/* Walk move types: same numbering as Walk::MoveType
 * (src/patternmatching/utils/walk.h:16-21). */
enum pj_move {
  PJ_MOVE_BACK = 0,
  PJ_MOVE_CHECK = 1,
  PJ_MOVE_VERTEX_STORE = 2,
  PJ_MOVE_VERTEX_NO_STORE = 3
};

/* Per-step counters: AlgoResults (src/patternmatching/algorithms/algo_results.h:11-22). */
typedef struct pj_algo_results {
  uint64_t vertex_eliminated;
  uint64_t edge_eliminated;
  uint64_t match_eliminated;
  uint64_t enumeration;
} pj_algo_results;

/* One template walk (Walk, src/patternmatching/utils/walk.h:14-37): the moves
 * of a DFS over a template rooted at `root`.  move_back[i] indexes the history
 * of VERTEX moves (history[0] is the root). */
typedef struct pj_walk_desc {
  uint32_t root;
  uint32_t length;
  const uint8_t* move;        /* enum pj_move, [length] */
  const uint32_t* vertex;     /* template vertex of each move, [length] */
  const uint32_t* move_back;  /* history index for CHECK / MOVE_BACK, [length] */
} pj_walk_desc;

/* Flattened constraint tables of one template (POD only).
 *  - local: MultipleLabelConstraintLocal per template vertex
 *    (multiple_label_constraint_local.h:18-32): origin label, neighbour-label
 *    multiset (need_label/need_count) and template neighbour set (neighbor_mask).
 *  - cycles: normalised MultipleLabelConstraintCircular::vertexIndexVector,
 *    in generation order (multiple_label_generate_constraint_circular.cc:58-86).
 *  - paths: MultipleLabelConstraintPath::vertexIndexVector + initialLabel
 *    (multiple_label_constraint_path.cc:9-30).
 *  - tds: for each template constraint, its walkMap in iteration order
 *    (multiple_label_constraint_template.h:179-180).
 *  - enumeration: walkMap.cbegin() of the whole-pattern template
 *    (multiple_label_enumeration_cpu.tpp:133-135). */
typedef struct pj_template_desc {
  uint32_t vertex_count;
  const uint64_t* vertex_label;   /* [vertex_count] */
  const uint64_t* neighbor_mask;  /* [vertex_count] */
  const uint32_t* need_offset;    /* [vertex_count + 1] */
  const uint64_t* need_label;     /* [need_offset[vertex_count]] */
  const uint32_t* need_count;     /* [need_offset[vertex_count]] */

  uint32_t cycle_count;
  const uint32_t* cycle_offset;   /* [cycle_count + 1] */
  const uint32_t* cycle_vertex;

  uint32_t path_count;
  const uint32_t* path_offset;    /* [path_count + 1] */
  const uint32_t* path_vertex;
  const uint64_t* path_initial_label; /* [path_count] */

  uint32_t tds_count;
  const uint32_t* tds_walk_offset; /* [tds_count + 1] into walks */
  uint32_t walk_count;
  const pj_walk_desc* walks;

  pj_walk_desc enumeration_walk;
} pj_template_desc;

typedef struct pj_ctx pj_ctx;

/* Creates a context on HIP device `device` with its own stream.
 * Replaces the OpenMP set-up of MultipleLabelCpu::initialiseTotem
 * (src/patternmatching/multiple_label_cpu.cc:12-18). */
int pj_create(pj_ctx** ctx, int device);

/* Copies a CSR graph to HBM: offsets [V+1] (uint64), edges [E] (uint32 ids),
 * labels [V] (uint64 vertex values).  The caller may free its arrays after the
 * call.  Replaces the host graph_t consumed by every compute()
 * (src/totem/totem_graph.h:137-180; loader src/totem/totem_graph.cu:388-409). */
int pj_upload_graph(pj_ctx* ctx, uint32_t vertex_count, uint64_t edge_count,
                    const uint64_t* offsets, const uint32_t* edges,
                    const uint64_t* labels);

/* Same as pj_upload_graph for arrays that already live in device memory of
 * this context's device.  The context copies them (device to device). */
int pj_upload_graph_device(pj_ctx* ctx, uint32_t vertex_count, uint64_t edge_count,
                           const uint64_t* d_offsets, const uint32_t* d_edges,
                           const uint64_t* d_labels);

/* Installs the constraint tables; replaces the setConstraintVector calls of
 * MultipleLabelCpu::allocate (src/patternmatching/multiple_label_cpu.cc:84-95).
 * Also (re)initialises the pruning state as pj_state_reset does. */
int pj_set_template(pj_ctx* ctx, const pj_template_desc* desc);

/* Pruning state to its initial value: every vertex and edge active, candidate
 * set {0..P-1}, nothing modified, everything scheduled
 * (MultipleLabelGlobalState::allocate, multiple_label_common_cpu.tpp:29-66). */
int pj_state_reset(pj_ctx* ctx);

/* resetState of the NLCC steps: modified flags to 0
 * (multiple_label_cc_cpu.tpp:72-75, pc_cpu.tpp:77-80, tds_cpu.tpp:185-188). */
int pj_reset_modified(pj_ctx* ctx);

/* Runs one step.  `constraint_index` selects the cycle / path / TDS
 * constraint for PJ_STEP_CC / PC / TDS and is ignored otherwise.  The
 * results are written to *out (may be NULL) and the context's active-vertex
 * and active-edge counters are updated as the reference's compute() does
 * (e.g. multiple_label_lcc_cpu.tpp:144-151). */
int pj_run_step(pj_ctx* ctx, int step, uint32_t constraint_index,
                pj_algo_results* out);

/* Device time of the last pj_run_step in milliseconds (HIP events on the
 * context stream). */
int pj_last_step_ms(pj_ctx* ctx, double* ms);

/* CSR slots the last step examined: every slot a serial scan in the
 * reference's loop order reads (scans stop at the first hit where the
 * reference's loop does).  Numerator of the pruning TEPS; no reference
 * counterpart (the reference times steps only, multiple_label_cpu.cc:137-164). */
int pj_last_step_slots(pj_ctx* ctx, uint64_t* slots);

/* Per-kernel timing.  pj_set_kernel_timing(ctx, 1) brackets every kernel
 * launch of later steps with HIP events on the context stream and keeps
 * per-launch counters; pj_kernel_stats returns the totals per kernel kind
 * (number of kinds written, at most max; out may be NULL to count).
 * Enabling or disabling clears the totals.  No reference counterpart (the
 * reference times whole steps only). */
typedef struct pj_kernel_stat {
  char name[32];
  uint64_t launches;
  double ms;                 /* summed launch time */
  uint64_t items;            /* vertices (or hub chunks) the launches were given */
  uint64_t slots;            /* CSR slots examined */
  uint64_t edges_eliminated; /* edge-active bits cleared */
  uint32_t word_bytes;       /* candidate word width (1, 2, 4 or 8 bytes) */
} pj_kernel_stat;
int pj_set_kernel_timing(pj_ctx* ctx, int enable);
int pj_kernel_stats(pj_ctx* ctx, pj_kernel_stat* out, int max);
/* As pj_set_kernel_timing(ctx, 1), but only launches of the named kind
 * (a pj_kernel_stat name) are bracketed with events and counted; NULL or ""
 * times every launch.  -1 for an unknown name.  Lets a caller time one
 * kernel without the cost of events around every other launch. */
int pj_set_kernel_timing_kind(pj_ctx* ctx, const char* name);

/* Active vertex / edge counters: graphActiveVertexCount and
 * graphActiveEdgeCount (multiple_label_common_cpu.h:43-46), as 64-bit values. */
int pj_get_counters(pj_ctx* ctx, uint64_t* active_vertices, uint64_t* active_edges);

/* Copies the state back to the host.  Any pointer may be NULL.
 *   cand      [V] candidate set (bit p = template vertex p) (vertexPatternMatch)
 *   vflags    [V] bit0 active, bit1 modified, bit2 scheduled
 *   ea_words  [ceil(E/32)] edge-active bitset, bit e%32 of word e/32 */
int pj_download_state(pj_ctx* ctx, uint64_t* cand, uint8_t* vflags, uint32_t* ea_words);

/* Order-dependent 64-bit checksum of the device state (candidate words,
 * active / modified / scheduled flags, edge-active bits): equal for equal
 * states, so a solve can be compared without downloading the state.  No
 * reference counterpart (the reference dumps the state to files,
 * multiple_label_cpu.cc:252-297). */
int pj_state_checksum(pj_ctx* ctx, uint64_t* checksum);

/* Streams a binary TOTEM graph file (graph_store_binary's format, offsets of
 * 32 or 64 bits) into the context: the neighbour ids go to HBM in 64 MiB
 * chunks through two pinned host buffers, reading one while the other is
 * copied, so no host array holds the edge list; offsets and values are read
 * whole.  `labels` [V] replaces the file's vertex values (NULL: use them; the
 * file must then be valued).  Then the same checks and set-up as
 * pj_upload_graph.  Replaces graph_initialize on a binary file followed by
 * pj_upload_graph (src/totem/totem_graph.cu:253-312, 388-409). */
int pj_upload_totem_file(pj_ctx* ctx, const char* path, const uint64_t* labels);

/* Makes ctx use the graph uploaded into `owner` (same device) without a
 * copy: CSR, labels and hub lists are shared read-only, the pruning state is
 * ctx's own.  `owner` must outlive ctx and keep its graph.  Many contexts
 * over one resident graph serve the batched and served forms (one stream
 * each).  No reference counterpart (the reference holds one graph and one
 * state, multiple_label_cpu.cc:20-57). */
int pj_share_graph(pj_ctx* ctx, const pj_ctx* owner);

/* The device index the context was created on (pj_create). */
int pj_get_device(const pj_ctx* ctx, int* device);

/* Overwrites the pruning state from host arrays of the same layout as
 * pj_download_state (used to start a step from a given state). */
int pj_upload_state(pj_ctx* ctx, const uint64_t* cand, const uint8_t* vflags,
                    const uint32_t* ea_words, uint64_t active_vertices,
                    uint64_t active_edges);

/* The surviving graph, compacted on the device: the vertices whose candidate
 * set is not empty (ascending, with their candidate sets as 64-bit words) and
 * the slots v->u whose edge-active bit is set between two surviving vertices
 * (CSR order, as src / dst pairs).  These are what the reference's final
 * dumps print, printActiveVertex and printActiveGraph
 * (src/patternmatching/multiple_label_cpu.cc:270-297), which read the whole
 * state on the host.  Outputs are device memory of the context's device with
 * room for cap_vertices / cap_slots entries (cand may be NULL); the counts
 * come back in *n_vertices / *n_slots.  When a count exceeds its capacity the
 * call returns -1 after writing what fits, with the counts set.  The
 * library keeps per-block counters for the context (allocated on first use). */
int pj_compact_active(pj_ctx* ctx, uint32_t* vertices, uint64_t* cand, uint64_t cap_vertices, uint32_t* src,
                      uint32_t* dst, uint64_t cap_slots, uint64_t* n_vertices, uint64_t* n_slots);

/* pj_compact_active without allocating: no reference counterpart; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Its per-block
 * counters live in a static device array of the library (256 KiB per device,
 * held until the process ends), so it takes graphs of at most 2^30 vertices
 * and 2^30 slots and returns -1 with an error for larger ones.  Calls on one
 * device share the array: this process serialises them with a lock; calls
 * from another copy of the library in the same process would race.  Same
 * outputs, counts and errors as pj_compact_active. */
int pj_compact_active_noalloc(pj_ctx* ctx, uint32_t* vertices, uint64_t* cand, uint64_t cap_vertices,
                              uint32_t* src, uint32_t* dst, uint64_t cap_slots, uint64_t* n_vertices,
                              uint64_t* n_slots);

/* pj_compact_active into host arrays (the context holds the device copies).
 * With every output pointer NULL only the counts are returned. */
int pj_download_active(pj_ctx* ctx, uint32_t* vertices, uint64_t* cand, uint64_t cap_vertices, uint32_t* src,
                       uint32_t* dst, uint64_t cap_slots, uint64_t* n_vertices, uint64_t* n_slots);

/* Label statistics of the active graph: the reference's GraphStat
 * (src/patternmatching/algorithms/graph_stat.h:14-45) as GraphStatCpu::compute
 * fills it (src/patternmatching/algorithms/graph_stat_cpu.tpp:40-82), for the
 * effectiveness estimate of the dynamic constraint order
 * (src/patternmatching/test_article/test_order_multiple_label_cpu.cc:170-290).
 * Labels are the installed template's distinct labels in template vertex
 * order (codes 0..label_count-1); tables over label pairs are indexed
 * [from * 64 + to].  Counted over active vertices and over active slots whose
 * both ends are active:
 *   vertex[l]     vertexLabelTotalNumberMap
 *   edge[f][t]    edgeLabelTotalNumberMap
 *   outbound[f]   edgeOutboundLabelTotalNumberMap (sum over t of edge[f][t])
 *   vertex_total, edge_total   vertexTotalNumber, edgeTotalNumber
 *   at_least_max[f][t]  accumulationMaximumMap (GraphStatCpu::preprocessPatern,
 *                 graph_stat_cpu.tpp:14-38): the most neighbours of label t
 *                 that a template vertex of label f has (0: no entry)
 *   at_least[at_least_offset[f][t] + k - 1], k = 1..at_least_max[f][t]:
 *                 vertexLabelWithAtLeastNeighborLabelMap[f][t][k], the active
 *                 vertices of label f with at least k such slots to label t.
 * The last counts each vertex's own slots; the reference's per-thread tally
 * is never cleared between vertices (graph_stat_cpu.tpp:51-70), so its
 * counts also depend on the thread schedule.  pj_last_step_ms reports the
 * device time of the pass.  Needs the pruning state of the template
 * (pj_set_template). */
#define PJ_STAT_MAX_AT_LEAST 4096
typedef struct pj_label_stat {
  uint32_t label_count;
  uint32_t at_least_count;
  uint64_t vertex_total;
  uint64_t edge_total;
  uint64_t label[64];
  uint64_t vertex[64];
  uint64_t outbound[64];
  uint64_t edge[64 * 64];
  uint32_t at_least_max[64 * 64];
  uint32_t at_least_offset[64 * 64];
  uint64_t at_least[PJ_STAT_MAX_AT_LEAST];
} pj_label_stat;
int pj_graph_stat(pj_ctx* ctx, pj_label_stat* out);

/* Last error message of the context (never NULL). */
const char* pj_last_error(pj_ctx* ctx);

/* Releases every device buffer and the stream. */
void pj_destroy(pj_ctx* ctx);

/* ---- Batched LCC over one resident graph --------------------------------
 * No counterpart in the reference, which solves one template per run
 * (MultipleLabelCpu::runPatternMatching, multiple_label_cpu.cc:113-185); the
 * batch serves the batched and served forms.  It runs, for many templates at
 * once, the steps the reference runs first for each: LCC0
 * (multiple_label_lcc0_cpu.tpp:13-147) and LCC steps
 * (multiple_label_lcc_cpu.tpp:18-152) until a step eliminates nothing
 * (AlgoResults::isEmpty, algo_results.cc:18-23).  Up to 64 templates share
 * each sweep of the graph (one lane of a wavefront each); every template's
 * per-step results, candidate sets, flags and edge-active bits are those of
 * the same steps run alone.  A template that reaches that fixed point is
 * parked; pj_batch_export installs its state into a context (with the same
 * template installed, sharing the graph) that continues the template's
 * pipeline, and pj_batch_release frees its lane for another template.
 * Calls on one batch may come from several threads. */
typedef struct pj_batch pj_batch;

/* A batch of `lanes` template lanes (groups of 64 share a sweep) over
 * the graph resident in `graph_owner` (which must outlive the batch).  On
 * failure *batch may still be set for pj_batch_last_error; destroy it. */
int pj_batch_create(pj_batch** batch, const pj_ctx* graph_owner, uint32_t lanes);
/* Admits a template into a free lane (its LCC0 runs in the next step).
 * Returns -2 when the batch cannot hold the template (more than 16 vertices,
 * or more than 255 labels over the batch's templates): solve it in a context
 * of its own; -1 on other errors, such as no free lane. */
int pj_batch_admit(pj_batch* batch, const pj_template_desc* desc, uint32_t* lane);
/* Runs up to `max_steps` steps (each: LCC0 of the lanes admitted since the
 * last step and one LCC step of every running lane), returning after the
 * first step in which some lanes reach their fixed point; those lanes are
 * parked and listed in done[0 .. *n_done) (at most max_done are written). */
int pj_batch_run(pj_batch* batch, uint32_t max_steps, uint32_t* done, uint32_t max_done, uint32_t* n_done);
/* Per-step counters of a parked lane: LCC0, then each LCC step (the last one
 * eliminated nothing). */
int pj_batch_trace(pj_batch* batch, uint32_t lane, pj_algo_results* rows, uint32_t max_rows, uint32_t* n_rows);
/* Installs a parked lane's state into `ctx`: candidate sets, modified and
 * scheduled flags, edge-active bits and the active counters, as
 * pj_run_step would have left them after the same steps. */
int pj_batch_export(pj_batch* batch, uint32_t lane, pj_ctx* ctx);
int pj_batch_release(pj_batch* batch, uint32_t lane);
/* Lanes, running and parked lanes, device time and steps run so far. */
int pj_batch_info(pj_batch* batch, uint32_t* lanes, uint32_t* running, uint32_t* parked, double* device_ms,
                  uint64_t* steps);
const char* pj_batch_last_error(pj_batch* batch);
void pj_batch_destroy(pj_batch* batch);

#ifdef __cplusplus
}
#endif

#endif /* PRUNEJUICE_H */
