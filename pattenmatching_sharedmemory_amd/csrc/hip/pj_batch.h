// Internal interface between the batched LCC (pj_batch.hip) and the
// single-template contexts of pj_device.hip: the resident graph a batch
// sweeps, and the import of one template's LCC fixed point into a context
// that continues that template's pipeline.
#ifndef PJ_BATCH_H
#define PJ_BATCH_H

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../../include/prunejuice.h"

namespace pjb {

// The device CSR of a context (owned or shared) and its hub chunk list:
// rows above hub_degree slots are split into chunks of hub_chunk slots,
// chunk i covering [chunk_start[i], min(chunk_start[i] + hub_chunk, end of
// row hubs[chunk_hub[i]])).
struct GraphView {
  int device;
  int cus;
  uint32_t V;
  uint64_t E;
  const uint64_t* off;        // [V + 1]
  const uint32_t* edges;      // [E]
  const uint64_t* labels;     // [V] raw vertex values
  const uint32_t* hubs;       // [nhubs]
  uint32_t nhubs;
  const uint32_t* chunk_hub;  // [nchunks] index into hubs
  const uint64_t* chunk_start;
  uint32_t nchunks;
  uint32_t hub_degree;
  uint32_t hub_chunk;
  // per slot, the index of the neighbour's label in graph_labels (null when
  // the graph has more than 255 labels)
  const uint8_t* slot_lab;
  const uint64_t* graph_labels;  // host memory, owned by the context
  uint32_t ngraph_labels;
};

// Fills `out` from a context holding a graph; -1 when it has none.
int graph_view(const pj_ctx* ctx, GraphView* out);

// Replaces the pruning state of `ctx` (whose template must already be
// installed, with candidate words of `word_bytes` bytes) by a state on the
// device: candidate words, flag bytes (modified 2, scheduled 4) and
// edge-active words, as left by LCC0 and LCC steps to a fixed point.  The
// copies run on the context's stream, which is synchronised before return;
// `ready` (may be null) is waited for first.
int import_lcc_state(pj_ctx* ctx, const void* cand, int word_bytes, const uint8_t* flags, const uint32_t* ea_words,
                     uint64_t active_vertices, uint64_t active_edges, hipEvent_t ready);

}  // namespace pjb

#endif
