// Label statistics of the active graph on the device: the counts that the
// reference's GraphStatCpu::compute gathers (algorithms/graph_stat_cpu.tpp:
// 40-82) for the effectiveness estimate of its dynamic constraint order
// (test_article/test_order_multiple_label_cpu.cc:170-290).  Internal
// interface of pj_stat.hip, called by pj_graph_stat in pj_device.hip.
//
// Labels are the template's label codes 0..K-1.  One pass over the slots of
// every active vertex (the frontier bins of all active vertices, and the
// chunks of the hub rows) counts
//   vertex[lf]          active vertices of label lf,
//   edge[lf][lt]        active slots v->u with u active, label(v) = lf,
//                       label(u) = lt,
//   exact[lf][lt][k]    active vertices of label lf with min(n, max[lf][lt])
//                       = k, n being the number of such slots to label lt,
//                       for the pairs with max[lf][lt] > 0 (the most
//                       neighbours of label lt any template vertex of label
//                       lf has) and n >= 1.
// "At least k" counts are the suffix sums of exact[lf][lt][k..max].
#ifndef PJ_STAT_H
#define PJ_STAT_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace pjs {

constexpr int kLabels = 64;           // label codes
constexpr uint32_t kExactMax = 4096;  // sum of max[lf][lt] over the pairs (<= 64 x 63)

// Statistics buffer on the device, 64-bit words: vertex [64], edge [64 x 64]
// (from * 64 + to), exact [kExactMax]; then the pair table as 32-bit words
// [64 x 64]: (offset of k = 1 in exact) << 8 | max.
constexpr uint64_t kVertexAt = 0;
constexpr uint64_t kEdgeAt = kLabels;
constexpr uint64_t kExactAt = kEdgeAt + kLabels * kLabels;
constexpr uint64_t kCountWords = kExactAt + kExactMax;
constexpr uint64_t kBufferBytes = kCountWords * 8 + kLabels * kLabels * 4;

struct StatArgs {
  const uint64_t* off;         // [V+1]
  const uint32_t* edges;       // [E]
  const uint32_t* ea;          // edge-active bits
  const void* cand;            // candidate words; a vertex is active iff its word is non-zero
  int word_bytes;
  const uint8_t* code;         // label code of each vertex
  uint32_t V;
  const uint32_t* bins;        // 4 bins of V entries: every active vertex by degree (<= 2048 slots)
  const uint32_t* bin_count;   // [4], device
  const uint32_t* hubs;        // rows above 2048 slots
  uint32_t nhubs;
  const uint32_t* chunk_hub;   // hub of each chunk of 2048 slots
  const uint64_t* chunk_start; // first slot of each chunk
  uint32_t nchunks;
  uint32_t* hub_count;         // [nhubs x 64] per-hub label counts: zero on entry, zero again on return
  const uint8_t* pcode;        // [64] label code of each template vertex (device)
  bool derive;                 // every candidate carries its vertex's label: codes from candidate words
  uint32_t K;                  // label codes in use
  uint32_t exact_words;        // sum of max over the pairs
  unsigned long long* buffer;  // kBufferBytes, pair table filled by the caller
  int cus;
  hipStream_t stream;
};

// Clears the counts and runs the pass on a.stream (asynchronous).
int label_stat(const StatArgs& a, std::string* err);

}  // namespace pjs

#endif
