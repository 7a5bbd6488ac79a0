// Active-graph compaction on the device: the surviving vertices (with their
// candidate sets) and the surviving slots (src, dst) in CSR order, which the
// reference's final dumps print (printActiveGraph / printActiveVertex,
// src/patternmatching/multiple_label_cpu.cc:270-297).  Internal interface of
// pj_compact.hip, called by the C ABI in pj_device.hip.
#ifndef PJ_COMPACT_H
#define PJ_COMPACT_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace pjc {

// Slots (and vertices) per block of the count and write passes; a pass has
// one block per 2^16 slots and one per 2^16 vertices, and one 64-bit counter
// of scratch per block.
constexpr uint64_t kSpan = 1ull << 16;
// Scratch of the no-allocation entry point: counters for 2^14 blocks of each
// kind (256 KiB), so graphs of at most 2^30 vertices and 2^30 slots.
constexpr uint64_t kStaticBlocks = 1ull << 14;
constexpr uint64_t kStaticMaxItems = kStaticBlocks * kSpan;

struct CompactArgs {
  const uint64_t* off;     // [V+1]
  const uint32_t* edges;   // [E]
  const uint32_t* ea;      // edge-active bits
  const void* cand;        // candidate words of word_bytes each; active iff non-zero
  int word_bytes;
  uint32_t V;
  uint64_t E;
  uint32_t* vertices;      // out [cap_v]
  uint64_t* cand_out;      // out [cap_v] (may be null)
  uint32_t* src;           // out [cap_e]
  uint32_t* dst;           // out [cap_e]
  uint64_t cap_v, cap_e;
  hipStream_t stream;
};

// Counters the passes need: vertex blocks + edge blocks + 2 totals.
uint64_t scratch_words(uint32_t V, uint64_t E);

// Runs count, scan and write on a.stream with `scratch` (device memory of
// scratch_words() counters); the totals come back through *n_v / *n_e (the
// stream is synchronised).  -1 with *err when an output is too small (the
// totals are still returned; nothing is written past a capacity) or on a HIP
// error.  Allocates nothing.
int compact_active(const CompactArgs& a, unsigned long long* scratch, uint64_t* n_v, uint64_t* n_e, std::string* err);

// Device address of this library's static scratch on `device`, which must be
// the current device (hipGetSymbolAddress, looked up once per device and
// kept: each device holds its own copy of the array).
int static_scratch(int device, unsigned long long** ptr, std::string* err);

}  // namespace pjc

#endif
