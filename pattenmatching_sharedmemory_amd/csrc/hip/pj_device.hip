// MI355X (gfx950) kernels of the PruneJuice pruning path and the C ABI of
// include/prunejuice.h.
//
// HBM layout of one context (V vertices, E directed CSR slots, P <= 64
// template vertices, word = uint32 when P <= 32 else uint64):
//   offsets  u64[V+1]   edges u32[E]   labels u64[V] (raw vertex values)
//   code     u8[V]      template label code of each vertex (0xFF: no template vertex has it)
//   cand     word[V]    candidate set; a vertex is active iff cand != 0
//   unmatch  word[V]    toUnmatch set of the current step
//   flags    u8[V]      bit1 modified, bit2 scheduled
//   ea       u32[ceil(E/32)]  edge-active bitset
// The reference keeps uint32 flags per vertex and per edge and two 64-bit
// bitmaps per vertex (multiple_label_common_cpu.h:23-59); activity is folded
// into the candidate word here because the reference deactivates a vertex
// exactly when its candidate set becomes empty (lcc0_cpu.tpp:42-45,
// lcc_cpu.tpp:114-117, cc_cpu.tpp:141-147).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/prunejuice.h"
#include "pj_batch.h"
#include "pj_compact.h"
#include "pj_stat.h"

namespace {

constexpr uint8_t kModified = 2;
constexpr uint8_t kScheduled = 4;
constexpr uint8_t kNoCode = 0xFF;
constexpr int kBlock = 256;
constexpr int kCounterSlots = 256;  // per-block counters are spread over slots
constexpr uint32_t kHubDegree = 2048;  // rows above this are split into chunks
constexpr uint32_t kHubChunk = 2048;
constexpr int kMaxLabels = 64;
static_assert(kHubChunk % kBlock == 0, "hub chunks are split evenly over a block");

// Counter slots: [slot][counter], summed per launch.  C_VERTEX / C_EDGE /
// C_MATCH: eliminated vertices, edges and matches; C_ENUM: enumerated
// embeddings; C_ITEMS: vertices (or hub chunks) a launch was given;
// C_SLOTS: slots examined here that a serial scan in the reference's loop
// order also reads; C_SKIP: slots such a scan reads that this launch does not
// examine (rows of clean vertices, which are skipped; hub rows, which another
// launch reads); C_XSLOTS: slots examined here that the reference's count
// does not include (dirty pushes, hub chunk reads).
enum { C_VERTEX = 0, C_EDGE = 1, C_MATCH = 2, C_ENUM = 3, C_SLOTS = 4, C_ITEMS = 5, C_SKIP = 6, C_XSLOTS = 7 };
constexpr int kCounterStride = 8;
constexpr size_t kCounterBlock = static_cast<size_t>(kCounterSlots) * kCounterStride;
constexpr int kMaxLaunches = 32;  // timed launches per step
// Grids of the grid-strided sweeps, in blocks per compute unit: more blocks
// than fit at once let blocks that finish early be replaced (measured on the
// scale-22 solve: 16 for LCC phase 1 and the walks, 8 for phase 2, whose
// tail steps are mostly empty launches).
constexpr int kPhase1BlocksPerCu = 16;
constexpr int kPhase2BlocksPerCu = 8;
constexpr int kWalkBlocksPerCu = 16;

// Kernel kinds reported by pj_kernel_stats.
enum KernelId {
  K_LCC0, K_BINS, K_LCC_P1, K_LCC_HUB, K_LCC_HUB_FINISH, K_P2, K_P2_HUB_FINISH, K_NLCC, K_NLCC_HUB, K_TRIANGLES,
  K_TRIANGLES_MEMO, K_MEMO_WALK, K_HUB_PUSH, K_VIEW, K_LABEL_STAT, K_TRIANGLES_WIDE, K_TRIANGLES_MID,
  K_LCC_P1_LABELS, K_LCC_HUB_LABELS, K_COUNT
};
// nlcc_walk: k_nlcc (cycle / path / TDS / enumeration walks over the bins,
// and the memo retry passes); triangles: k_triangles (3-vertex cycles from
// sources of at most 8 / 32 slots); triangles_mid: k_triangles_mid (33-256
// slots); triangles_wide: k_triangles_wide (257-2048 slots); triangles_memo:
// k_triangles_memo; memo_walk: the first pass of the memoised walks
// (k_memo_lds); lcc_phase1_labels and lcc_phase1_hub_chunks_labels: the
// builds of the first LCC step after LCC0, which read slot labels where the
// others read neighbour ids and candidate words.
const char* const kKernelNames[K_COUNT] = {
    "lcc0",          "build_bins",        "lcc_phase1", "lcc_phase1_hub_chunks", "lcc_phase1_hub_finish",
    "phase2",        "phase2_hub_finish", "nlcc_walk",  "nlcc_walk_hubs",        "triangles",
    "triangles_memo", "memo_walk",        "hub_rescan_push", "active_view",    "label_stat",
    "triangles_wide", "triangles_mid",    "lcc_phase1_labels", "lcc_phase1_hub_chunks_labels"};

// Template tables in device memory.
struct DevTemplate {
  uint32_t P;
  uint32_t K;                       // distinct template labels (codes 0..K-1)
  uint64_t label_raw[kMaxLabels];   // label value of each code
  uint64_t code_vertices[kMaxLabels];  // template vertices carrying each code
  uint8_t pcode[64];                // code of each template vertex
  uint64_t nbr[64];                 // template neighbours of each vertex
  uint64_t need[64][8];             // required neighbour count per code, one byte per code
  // NLCC constraint currently installed for the step kernels.
  uint32_t nlcc_len;
  uint32_t nlcc_vertex[64];         // cycle / path vertices, or walk template vertices
  uint8_t nlcc_move[128];
  uint32_t nlcc_mb[128];
  uint32_t nlcc_hidx[128];          // history index written by each VERTEX move
  uint32_t walk_count;              // walks of the TDS constraint (<= 64)
  uint32_t walk_root[64];
  uint32_t walk_off[65];            // into walk_move / walk_vertex / walk_mb / walk_hidx
  uint8_t walk_move[64 * 128];
  uint32_t walk_vertex[64 * 128];
  uint32_t walk_mb[64 * 128];
  uint32_t walk_hidx[64 * 128];
  uint8_t walk_test[64 * 128];      // memoised walks: template id / position tested in the edge memo
  uint8_t walk_mark[64 * 128];      // memoised walks: mark alreadyMatched on success
  uint32_t nlcc_memo;               // run the installed NLCC step as memoised walks (walk_* hold the plans)
  uint32_t walk_hyp;                // memoised walks use the hypothesis memo (TDS)
  uint32_t path_code;               // code of the path's initial label
  uint32_t rows_sorted;             // every CSR row ascending (binary search allowed)
  uint8_t slot_code[256];           // code of each graph label index (slot_lab), kNoCode when absent
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t rl32(uint32_t v, int lane) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), lane));
}
__device__ __forceinline__ uint64_t rl64(uint64_t v, int lane) {
  uint32_t lo = rl32(static_cast<uint32_t>(v), lane);
  uint32_t hi = rl32(static_cast<uint32_t>(v >> 32), lane);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

template <typename Word>
__device__ __forceinline__ int popc(Word w) {
  return sizeof(Word) == 8 ? __popcll(static_cast<unsigned long long>(w)) : __popc(static_cast<unsigned>(w));
}
template <typename Word>
__device__ __forceinline__ int ctz(Word w) {
  return sizeof(Word) == 8 ? __builtin_ctzll(static_cast<unsigned long long>(w)) : __builtin_ctz(static_cast<unsigned>(w));
}

__device__ __forceinline__ bool ea_test(const uint32_t* __restrict__ ea, uint64_t e) {
  return (ea[e >> 5] >> (e & 31)) & 1u;
}

// Slot test for walks that may run on the active view (no bitset: every
// slot of the view is active).
__device__ __forceinline__ bool walk_slot(const uint32_t* __restrict__ ea, uint64_t e) {
  return !ea || ea_test(ea, e);
}

}  // namespace

// Wave reduction of the ROCm device library (ockl), declared by HIP's headers
// only under HIP_ENABLE_EXTRA_WARP_SYNC_TYPES.
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_add_u64(unsigned long long);

namespace {

// Wave sum of a per-lane value, added by lane 0 to a spread counter slot.
// Must be called with every lane of the wave active.
__device__ __forceinline__ void wave_add(unsigned long long* counters, int which, uint64_t v) {
  // A wave whose lanes all hold 0 (most waves of a launch over a small
  // frontier) skips the reduction (DPP steps of the device library).
  if (!__ballot(v != 0)) return;
  v = __ockl_wfred_add_u64(v);
  if (lane_id() == 0 && v) {
    int slot = (blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) & (kCounterSlots - 1);
    atomicAdd(&counters[slot * kCounterStride + which], static_cast<unsigned long long>(v));
  }
}

// Clears the edge-active bits of the eliminated slots among the G
// consecutive slots [sb, sb + G) of a lane group (lane gl holds slot sb + gl;
// `el`: that slot is eliminated) with one atomic per touched word.  Every
// lane of the group must reach it.
template <int G>
__device__ __forceinline__ void clear_slots(uint32_t* __restrict__ ea, uint64_t sb, bool el, int gl, int gbase) {
  const uint64_t rel = (__ballot(el) >> gbase) & (G == 64 ? ~0ull : ((1ull << G) - 1));
  if (!rel) return;
  const int sh = static_cast<int>(sb & 31);
  const uint64_t lo = rel << sh;
  const uint64_t hi = sh ? rel >> (64 - sh) : 0ull;
  const uint32_t m = gl == 0 ? static_cast<uint32_t>(lo)
                   : gl == 1 ? static_cast<uint32_t>(lo >> 32)
                   : gl == 2 ? static_cast<uint32_t>(hi)
                             : 0u;
  if (m) atomicAnd(&ea[(sb >> 5) + gl], ~m);
}

__device__ __forceinline__ void set_flag_atomic(uint8_t* flags, uint32_t v, uint8_t bit) {
  uint32_t* word = reinterpret_cast<uint32_t*>(reinterpret_cast<uintptr_t>(flags + v) & ~uintptr_t(3));
  atomicOr(word, static_cast<uint32_t>(bit) << (8 * (v & 3)));
}

// ---------------------------------------------------------------------------
// Set-up kernels.

template <typename Word>
__global__ void k_state_init(Word* cand, Word* unmatch, uint8_t* flags, uint32_t V, Word full) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
    cand[v] = full;
    unmatch[v] = 0;
    flags[v] = kScheduled;
  }
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t value) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = value;
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long value) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = value;
}

// Row-order check: k_row_starts marks the first slot of every row in a
// zeroed bitset, k_rows_sorted sets bit 0 of *unsorted when a slot that starts
// no row holds a smaller id than the slot before it, bit 1 when it holds the
// same id, bit 2 when a slot holds an id that is not a vertex.
__global__ void k_row_starts(const uint64_t* __restrict__ off, uint32_t V, uint64_t E, uint32_t* __restrict__ bits) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e = off[v];
    if (e < E) atomicOr(&bits[e >> 5], 1u << (e & 31));
  }
}

__global__ void k_rows_sorted(const uint32_t* __restrict__ edges, uint64_t E, uint32_t V,
                              const uint32_t* __restrict__ bits, uint32_t* unsorted) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = edges[e];
    if (u >= V) atomicOr(unsorted, 4u);  // not a vertex: the upload is refused
    if (e && !((bits[e >> 5] >> (e & 31)) & 1u)) {
      if (u < edges[e - 1]) atomicOr(unsorted, 1u);
      else if (u == edges[e - 1]) atomicOr(unsorted, 2u);  // duplicate slot
    }
  }
}

// Symmetry check on rows sorted strictly ascending (k_rows_symmetric and,
// for hub rows, k_rows_symmetric_chunks).  Order vertices by (degree, id).  Every slot u->v with v before u is looked
// up in the shorter row(v) by binary search (flag[0] set when u is missing);
// flag[1] / flag[2] count the slots u->v with v before / after u.  When all
// lookups succeed and the counts agree, the graph is symmetric: reversing the
// "v before u" slots maps them injectively (rows hold no duplicates) onto
// "after" slots, and equal counts make that map onto.  With `rev` given, every
// successful lookup also records the reverse positions of the pair: rev[e] of
// a slot u->v is the position of u in row(v), and the reverse slot is
// off[v] + rev[e] (a self-loop is its own reverse).  rev is complete exactly
// when the graph is found symmetric.

// One slot u->v of the symmetry check (row(u) starts at beg, has du slots).
__device__ __forceinline__ void symmetric_slot(uint64_t e, uint64_t u, uint64_t beg, uint64_t du,
                                               const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges,
                                               uint32_t V, uint32_t* __restrict__ rev, bool& bad, bool& loop,
                                               uint64_t& before, uint64_t& after) {
  const uint32_t v = edges[e];
  if (v >= V) {
    bad = true;
    return;
  }
  if (v == u) {
    if (rev) rev[e] = static_cast<uint32_t>(e - beg);
    loop = true;
    return;
  }
  uint64_t lo = off[v], hi = off[v + 1];
  const uint64_t dv = hi - lo;
  if (dv < du || (dv == du && v < u)) {
    ++before;
    const uint64_t last = hi;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (edges[mid] < u)
        lo = mid + 1;
      else
        hi = mid;
    }
    if (lo == last || edges[lo] != u) {
      bad = true;
    } else if (rev) {
      rev[e] = static_cast<uint32_t>(lo - off[v]);
      rev[lo] = static_cast<uint32_t>(e - beg);
    }
  } else {
    ++after;
  }
}

__device__ __forceinline__ void symmetric_flags(unsigned long long* flag, bool bad, bool loop, uint64_t before,
                                                uint64_t after) {
  const int lane = __lane_id();
  for (int o = 32; o > 0; o >>= 1) {
    before += __shfl_xor(before, o, 64);
    after += __shfl_xor(after, o, 64);
  }
  if (__ballot(bad) && lane == 0) atomicOr(&flag[0], 1ull);
  if (__ballot(loop) && lane == 0) atomicOr(&flag[3], 1ull);  // a self-loop slot
  if (lane == 0) {
    if (before) atomicAdd(&flag[1], before);
    if (after) atomicAdd(&flag[2], after);
  }
}

// Rows up to kHubDegree slots, one wave per vertex; hub rows are split
// into chunks by k_rows_symmetric_chunks (a single wave working through the
// largest row would hold up the whole check).
__global__ void __launch_bounds__(kBlock) k_rows_symmetric(const uint64_t* __restrict__ off,
                                                           const uint32_t* __restrict__ edges, uint32_t V,
                                                           unsigned long long* flag, uint32_t* __restrict__ rev) {
  const int lane = __lane_id();
  const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  bool bad = false, loop = false;
  uint64_t before = 0, after = 0;
  for (uint64_t u = wave; u < V; u += waves) {
    const uint64_t beg = off[u], end = off[u + 1];
    const uint64_t du = end - beg;
    if (du > kHubDegree) continue;
    for (uint64_t e = beg + lane; e < end; e += 64) symmetric_slot(e, u, beg, du, off, edges, V, rev, bad, loop, before, after);
  }
  symmetric_flags(flag, bad, loop, before, after);
}

// Hub rows of the symmetry check: one block per chunk of kHubChunk slots.
__global__ void __launch_bounds__(kBlock) k_rows_symmetric_chunks(
    const uint32_t* __restrict__ chunk_hub, const uint64_t* __restrict__ chunk_start, uint32_t nchunks,
    const uint32_t* __restrict__ hubs, const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges,
    uint32_t V, unsigned long long* flag, uint32_t* __restrict__ rev) {
  bool bad = false, loop = false;
  uint64_t before = 0, after = 0;
  for (uint32_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    const uint64_t u = hubs[chunk_hub[c]];
    const uint64_t beg = off[u], end = off[u + 1];
    const uint64_t cb = chunk_start[c], ce = min(end, cb + kHubChunk);
    for (uint64_t e = cb + threadIdx.x; e < ce; e += blockDim.x)
      symmetric_slot(e, u, beg, end - beg, off, edges, V, rev, bad, loop, before, after);
  }
  symmetric_flags(flag, bad, loop, before, after);
}

// Sums the kCounterSlots slots of counter block blockIdx.x and clears
// them; block 0 also clears the `nring` bin counts of the ring, so the next
// step starts without a launch that clears them.
__global__ void k_sum_counters(unsigned long long* __restrict__ counters, unsigned long long* __restrict__ sums,
                               uint32_t* __restrict__ ring, uint32_t nring) {
  __shared__ unsigned long long part[kBlock / 64][kCounterStride];
  unsigned long long* blk = counters + static_cast<size_t>(blockIdx.x) * kCounterBlock;
  unsigned long long acc[kCounterStride] = {};
  for (int s = threadIdx.x; s < kCounterSlots; s += blockDim.x)
    for (int k = 0; k < kCounterStride; ++k) {
      acc[k] += blk[s * kCounterStride + k];
      blk[s * kCounterStride + k] = 0;
    }
  if (blockIdx.x == 0)
    for (uint32_t i = threadIdx.x; i < nring; i += blockDim.x) ring[i] = 0;
  for (int k = 0; k < kCounterStride; ++k) {
    unsigned long long x = acc[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6][k] = x;
  }
  __syncthreads();
  if (threadIdx.x < kCounterStride) {
    unsigned long long x = 0;
    for (int w = 0; w < kBlock / 64; ++w) x += part[w][threadIdx.x];
    sums[static_cast<size_t>(blockIdx.x) * kCounterStride + threadIdx.x] = x;
  }
}

__global__ void k_clear_modified(uint8_t* flags, uint32_t V) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x)
    flags[v] &= static_cast<uint8_t>(~kModified);
}

__global__ void k_slot_labels(const uint32_t* __restrict__ edges, uint64_t E, const uint8_t* __restrict__ idx,
                              uint8_t* __restrict__ slot_lab) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x)
    slot_lab[e] = idx[edges[e]];
}

__global__ void k_label_codes(const uint64_t* __restrict__ labels, uint8_t* __restrict__ code, uint32_t V,
                              const DevTemplate* __restrict__ t) {
  const uint32_t K = t->K;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t l = labels[v];
    uint8_t c = kNoCode;
    for (uint32_t k = 0; k < K; ++k)
      if (t->label_raw[k] == l) {
        c = static_cast<uint8_t>(k);
        break;
      }
    code[v] = c;
  }
}

// ---------------------------------------------------------------------------
// LCC0 (multiple_label_lcc0_cpu.tpp:25-47): drop every candidate whose label
// differs from the vertex label; deactivate vertices left without candidates.

template <typename Word>
__global__ void k_lcc0(Word* __restrict__ cand, const uint8_t* __restrict__ code, uint32_t V,
                       const DevTemplate* __restrict__ t, unsigned long long* counters) {
  uint64_t velim = 0, melim = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t span = ((V + stride - 1) / stride) * stride;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < span; v += stride) {
    if (v < V) {
      Word c = cand[v];
      if (c) {
        uint8_t k = code[v];
        Word keep = k == kNoCode ? Word(0) : (c & static_cast<Word>(t->code_vertices[k]));
        if (keep != c) {
          melim += popc<Word>(c & ~keep);
          cand[v] = keep;
          if (!keep) ++velim;
        }
      }
    }
  }
  wave_add(counters, C_VERTEX, velim);
  wave_add(counters, C_MATCH, melim);
}

// ---------------------------------------------------------------------------
// Frontier construction.  Vertices are appended to degree bins with one
// atomic per wave and bin; hubs (degree > kHubDegree) go to no bin, they are
// served by the static chunk list built at upload.
// mode 0: active and scheduled (LCC phase 1); mode 1: active (NLCC task
// list); mode 2: active with label code == t->path_code (PC); mode 3: the
// phase 2 frontier (see bin_of).

constexpr int kBins = 4;
constexpr int kBinRing = 16;  // sets of bin counts a step takes before it clears them again
__constant__ uint32_t c_bin_limit[kBins] = {4, 32, 256, kHubDegree};
// Bin 0 of the walk task lists holds rows of up to 8 slots (k_triangles with
// 8 lanes per source).
constexpr uint32_t kWalkBin0 = 8;

constexpr int kBinIters = 4;                  // vertices per thread
constexpr int kBinVerts = kBlock * kBinIters;  // vertices per block

// mode 0 (LCC phase 1 frontier): active, scheduled and, when `marks` (the
// dirty marks) are given, dirty.  A scheduled vertex whose candidate set and
// whose neighbours' candidate sets are unchanged since its last evaluation
// would remove nothing and eliminate no slot: its previous removals were
// applied, its eliminated slots stay eliminated and the surviving slots and
// label counts are the same.  Skipping it leaves every result, flag and
// counter unchanged.
//
// mode 3 (phase 2 frontier): active, and modified or, when `marks` (the
// rescan marks) are given, marked for a rescan.  An unmodified vertex keeps
// the result of its last scan until a neighbour turns modified (the only
// change that can move its first hit: its own edge-active bits change only
// when it is evaluated, and a vertex that loses a slot turns modified), so
// its scheduled flag stands and last_scan[v] holds the slots the scan reads.
// A modified vertex that is scheduled and has nothing to drop is left out:
// phase 2 would change nothing for it and reads no slot.
//
// mode 4 (phase 2 of an NLCC step whose modified flags were all cleared
// before it, symmetric graph): the only modified vertices are the failing
// sources, and each has marked the vertices of its row (`marks`, the rescan
// marks).  By symmetry an unmarked vertex has no slot to a modified vertex:
// its scan would read the whole row and find nothing (-3: the caller records
// that result).  Marked and modified vertices are binned as in mode 3.
//
// mode 5: active and modified (the sources that mark their neighbours).
//
// The walk task lists (modes 1 and 2) use the same bins: one wave per source
// above 32 slots, bin 2 (33-256 slots) taken by k_triangles_mid in 3-vertex
// cycle steps.
//
// Returns the bin, -1 (not in the frontier), -2 (skipped; its slots are
// counted as examined) or -3 (mode 4 default).  Every load is issued before
// the tests.
template <typename Word>
__device__ __forceinline__ int bin_of(uint64_t v, uint32_t V, const Word* __restrict__ cand,
                                      const uint8_t* __restrict__ flags, const uint8_t* __restrict__ code,
                                      const uint8_t* __restrict__ marks, const Word* __restrict__ unmatch,
                                      const uint64_t* __restrict__ off, int mode, uint32_t path_code, uint64_t& deg) {
  deg = 0;
  if (v >= V) return -1;
  const Word c = cand[v];
  const uint8_t f = flags[v];
  const uint8_t mk = marks ? marks[v] : uint8_t(1);
  const uint8_t cd = mode == 2 ? code[v] : uint8_t(0);
  const Word tu = mode == 3 || mode == 4 || mode == 6 ? unmatch[v] : Word(0);
  deg = off[v + 1] - off[v];
  if (!c) return -1;
  if (mode == 0) {
    if (!(f & kScheduled)) return -1;
    if (!mk) return -2;  // clean: skipped
  } else if (mode == 3 || mode == 4 || mode == 6) {
    if (f & kModified) {
      if (!tu && (f & kScheduled)) return -1;  // nothing to apply
    } else if (mode == 6 || !mk) {
      return mode == 3 ? -2 : -3;  // scan result stands / no modified neighbour / pushed first hit
    }
  } else if (mode == 5) {
    if (!(f & kModified)) return -1;
  } else if (mode == 2 && cd != path_code) {
    return -1;
  }
  if (deg <= (mode == 1 || mode == 2 ? kWalkBin0 : c_bin_limit[0])) return 0;
  if (deg <= c_bin_limit[1]) return 1;
  if (deg <= c_bin_limit[2]) return 2;
  if (deg <= c_bin_limit[3]) return 3;
  return -1;
}

// Each block compacts kBinVerts consecutive entries of the live list (the
// vertices that may still be active; `live` null: every vertex): pass 1
// classifies them (kept in registers) and counts per bin with ballots, one
// global atomic per block and bin reserves space, pass 2 writes the vertex
// ids (each wave's run stays contiguous in its bin).
// Phase 1 of a scheduled vertex without slots (mode 0): every count is 0, so
// candidate p fails iff its label differs or it requires any neighbour
// (multiple_label_lcc_cpu.tpp:79-101 over an empty row); the vertex turns
// modified when one fails.  No slot, no edge elimination, no neighbour to
// mark for a rescan.
template <typename Word>
__device__ __forceinline__ void phase1_isolated(uint32_t v, const Word* __restrict__ cand,
                                                const uint8_t* __restrict__ code, uint8_t* __restrict__ flags,
                                                uint8_t* __restrict__ marks, Word* __restrict__ unmatch,
                                                const DevTemplate* __restrict__ t) {
  const Word cv = cand[v];
  const uint8_t vcode = code[v];
  Word fail = 0;
  for (Word m = cv; m; m &= m - 1) {
    const int p = ctz<Word>(m);
    bool bad = t->pcode[p] != vcode;
#pragma unroll
    for (int w = 0; w < 8; ++w) bad |= t->need[p][w] != 0;
    if (bad) fail |= Word(1) << p;
  }
  if (fail) {
    unmatch[v] = fail;
    flags[v] = static_cast<uint8_t>(flags[v] | kModified);
  }
  if (marks) marks[v] = 0;
}

template <typename Word>
__global__ void __launch_bounds__(kBlock) k_build_bins(
    const uint32_t* __restrict__ live, uint32_t n_live, const Word* __restrict__ cand, uint8_t* __restrict__ flags, const uint8_t* __restrict__ code, uint8_t* __restrict__ marks,
    Word* __restrict__ unmatch, const uint64_t* __restrict__ off, uint32_t V, int mode,
    const DevTemplate* __restrict__ t, uint32_t* __restrict__ bins, uint32_t* __restrict__ bin_count,
    uint32_t* __restrict__ last_scan, const uint32_t* __restrict__ first, unsigned long long* counters) {
  __shared__ uint32_t s_wave[kBlock / 64][kBins];
  __shared__ uint32_t s_base[kBlock / 64][kBins];
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  const uint32_t path_code = mode == 2 ? t->path_code : 0u;
  const uint64_t i0 = static_cast<uint64_t>(blockIdx.x) * kBinVerts + threadIdx.x;
  uint32_t vid[kBinIters];
#pragma unroll
  for (int i = 0; i < kBinIters; ++i) {
    const uint64_t idx = i0 + static_cast<uint64_t>(i) * kBlock;
    vid[i] = idx >= n_live ? V : live ? live[idx] : static_cast<uint32_t>(idx);
  }
  int bin[kBinIters];
  uint64_t deg[kBinIters];
#pragma unroll
  for (int i = 0; i < kBinIters; ++i)
    bin[i] = bin_of<Word>(vid[i], V, cand, flags, code, marks, unmatch, off, mode, path_code, deg[i]);
  if (mode == 0) {  // LCC phase 1 of vertices without slots, here instead of in a bin
#pragma unroll
    for (int i = 0; i < kBinIters; ++i)
      if (bin[i] == 0 && deg[i] == 0) {
        phase1_isolated<Word>(vid[i], cand, code, flags, marks, unmatch, t);
        bin[i] = -1;
      }
  }
  uint32_t cnt[kBins] = {};
  uint64_t skipped = 0;
#pragma unroll
  for (int i = 0; i < kBinIters; ++i) {
    if (bin[i] == -2) skipped += mode == 0 ? deg[i] : last_scan[vid[i]];
    if (bin[i] == -3) {  // the scan finds no modified neighbour, or (first given) stops at first[v]
      const uint32_t v = vid[i];
      const uint8_t f = flags[v];
      const uint32_t h = first ? first[v] : 0xFFFFFFFFu;
      const uint64_t scanned = h != 0xFFFFFFFFu ? h + 1ull : deg[i];
      const uint8_t nf = h != 0xFFFFFFFFu ? static_cast<uint8_t>(f | kScheduled) : static_cast<uint8_t>(f & ~kScheduled);
      if (nf != f) flags[v] = nf;
      last_scan[v] = static_cast<uint32_t>(scanned);
      skipped += scanned;
    }
#pragma unroll
    for (int k = 0; k < kBins; ++k) cnt[k] += __popcll(__ballot(bin[i] == k));
  }
  if (marks || first) wave_add(counters, C_SKIP, skipped);
  if (lane == 0)
    for (int k = 0; k < kBins; ++k) s_wave[wave][k] = cnt[k];
  __syncthreads();
  if (threadIdx.x < kBins) {
    const int k = threadIdx.x;
    uint32_t tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) tot += s_wave[w][k];
    uint32_t base = tot ? atomicAdd(&bin_count[k], tot) : 0u;
    for (int w = 0; w < kBlock / 64; ++w) {
      s_base[w][k] = base;
      base += s_wave[w][k];
    }
  }
  __syncthreads();
  uint32_t pos[kBins];
#pragma unroll
  for (int k = 0; k < kBins; ++k) pos[k] = s_base[wave][k];
  const uint64_t below = (1ull << lane) - 1;
#pragma unroll
  for (int i = 0; i < kBinIters; ++i) {
#pragma unroll
    for (int k = 0; k < kBins; ++k) {
      const uint64_t m = __ballot(bin[i] == k);
      if (bin[i] == k) bins[static_cast<uint64_t>(k) * V + pos[k] + __popcll(m & below)] = vid[i];
      pos[k] += __popcll(m);
    }
  }
}

// Live list compaction: the entries of `in` (null: every vertex) that are
// still active, in blocks of kBinVerts entries whose runs keep their order.
template <typename Word>
__global__ void __launch_bounds__(kBlock) k_live_compact(const uint32_t* __restrict__ in, uint32_t n_in,
                                                        const Word* __restrict__ cand, uint32_t* __restrict__ out,
                                                        uint32_t* __restrict__ n_out) {
  __shared__ uint32_t s_wave[kBlock / 64];
  __shared__ uint32_t s_base[kBlock / 64];
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  const uint64_t i0 = static_cast<uint64_t>(blockIdx.x) * kBinVerts + threadIdx.x;
  uint32_t vid[kBinIters];
  bool keep[kBinIters];
  uint32_t cnt = 0;
#pragma unroll
  for (int i = 0; i < kBinIters; ++i) {
    const uint64_t idx = i0 + static_cast<uint64_t>(i) * kBlock;
    vid[i] = idx < n_in ? (in ? in[idx] : static_cast<uint32_t>(idx)) : 0u;
    keep[i] = idx < n_in && cand[vid[i]] != Word(0);
    cnt += __popcll(__ballot(keep[i]));
  }
  if (lane == 0) s_wave[wave] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) tot += s_wave[w];
    uint32_t base = tot ? atomicAdd(n_out, tot) : 0u;
    for (int w = 0; w < kBlock / 64; ++w) {
      s_base[w] = base;
      base += s_wave[w];
    }
  }
  __syncthreads();
  uint32_t pos = s_base[wave];
  const uint64_t below = (1ull << lane) - 1;
#pragma unroll
  for (int i = 0; i < kBinIters; ++i) {
    const uint64_t m = __ballot(keep[i]);
    if (keep[i]) out[pos + __popcll(m & below)] = vid[i];
    pos += __popcll(m);
  }
}

// ---------------------------------------------------------------------------
// LCC phase 1 (multiple_label_lcc_cpu.tpp:26-102) for one vertex per group of
// G lanes.  Slots are read G at a time; a slot whose neighbour (active, over
// an active slot) has no candidate among the template neighbours of v's
// candidates loses its edge; the others are counted per neighbour label with
// one ballot per label code into saturating byte counters (cap 127).  Each
// candidate p of v is then kept only if v's label is p's label and every
// required neighbour-label count of p is met.

// The tables are the template's pcode[p] and need[p][w] (row stride NS):
// the DevTemplate itself, or a block's LDS copy of them.
template <typename Word, int KW, int NS>
__device__ __forceinline__ Word lcc_requirements(Word cv, uint8_t vcode, const uint64_t (&cnt)[KW],
                                                 const uint8_t* pcode, const uint64_t* need) {
  const uint64_t H = 0x8080808080808080ull;
  Word fail = 0;
  for (Word m = cv; m; m &= m - 1) {
    int p = ctz<Word>(m);
    bool bad = pcode[p] != vcode;
#pragma unroll
    for (int w = 0; w < KW; ++w) {
      uint64_t d = (cnt[w] | H) - need[p * NS + w];
      bad |= ((~d) & H) != 0;
    }
    if (bad) fail |= Word(1) << p;
  }
  return fail;
}

template <typename Word, int KW>
__device__ __forceinline__ Word lcc_requirements(Word cv, uint8_t vcode, const uint64_t (&cnt)[KW],
                                                 const DevTemplate* __restrict__ t) {
  return lcc_requirements<Word, KW, 8>(cv, vcode, cnt, t->pcode, &t->need[0][0]);
}

// A block's LDS copy of the template tables LCC phase 1 reads per vertex
// (codes, template neighbours, required counts of the first KW words).
template <int KW>
struct LccTables {
  uint8_t pcode[64];
  uint8_t slot_code[256];  // graph label index -> code (first LCC step, slot_lab)
  uint64_t nbr[64];
  uint64_t cv[kMaxLabels];  // template vertices of each code
  uint64_t need[64 * KW];
};

template <int KW>
__device__ __forceinline__ void load_lcc_tables(LccTables<KW>& s, const DevTemplate* __restrict__ t) {
  for (int i = threadIdx.x; i < 64; i += blockDim.x) {
    s.pcode[i] = t->pcode[i];
    s.nbr[i] = t->nbr[i];
    s.cv[i] = t->code_vertices[i];
  }
  for (int i = threadIdx.x; i < 256; i += blockDim.x) s.slot_code[i] = t->slot_code[i];
  for (int i = threadIdx.x; i < 64 * KW; i += blockDim.x) s.need[i] = t->need[i / KW][i % KW];
}

// Saturating sum of packed byte counters that are each at most 127: the
// byte sums stay below 256, so no carry crosses a byte; bytes that reach
// 128 become 127.
__device__ __forceinline__ uint64_t sat_add_bytes(uint64_t a, uint64_t b) {
  const uint64_t H = 0x8080808080808080ull;
  const uint64_t x = a + b;
  const uint64_t hi = x & H;
  return (x & ~H) | (hi - (hi >> 7));
}

// LCC phase 1 for one vertex per group of G lanes (G = 1, 4, 32 or 64), each
// lane owning S slots of a pass: slot base + k*G + gl, k < S, so that slice
// k of the pass is G consecutive slots.  The ids and edge-active words of a
// whole pass are requested before any neighbour gather, and the gathers
// before any use, so a pass waits on two rounds of memory latency and rows
// of at most G*S slots take one pass.  The vertex's own loads (candidates,
// row bounds, code, flags) are issued together before the row.  Each lane
// counts neighbour label codes in packed bytes (at most 2048 / G slots per
// lane, below 128); the group sums them with shuffles and saturating byte
// adds (cap 127, above every requirement).  With `derive` set, every active
// vertex's candidates carry its own label (true once LCC0 has run: later
// steps only remove candidates), so a neighbour's code is that of any of its
// candidates, read from the LDS table instead of gathering code[u].
template <typename Word, int G, int S, int KW, bool kSlot>
__device__ __forceinline__ void lcc_phase1_lanes(uint32_t block, uint32_t nblocks, const LccTables<KW>& tab,
                                 const uint32_t* __restrict__ list, const uint32_t* __restrict__ list_count,
                                 const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges,
                                 const uint8_t* __restrict__ code, bool derive,
                                 const uint8_t* __restrict__ slot_lab, const Word* __restrict__ cand,
                                 Word* __restrict__ unmatch, uint8_t* __restrict__ flags, uint8_t* __restrict__ dirty,
                                 uint8_t* __restrict__ rescan, uint32_t* __restrict__ ea,
                                 const DevTemplate* __restrict__ t, unsigned long long* counters) {
  const int lane = lane_id();
  const int gl = lane & (G - 1);
  const int gbase = lane & ~(G - 1);
  const uint64_t gmask = G == 64 ? ~0ull : (((1ull << G) - 1) << gbase);
  const uint32_t n = *list_count;
  const uint64_t groups_per_grid = (uint64_t)nblocks * (blockDim.x / G);
  // every lane runs the same number of iterations (shuffles and ballots)
  const uint64_t span = ((n + groups_per_grid - 1) / groups_per_grid) * groups_per_grid;
  uint64_t elim_total = 0, slots_total = 0, xslots = 0;
  for (uint64_t g = (block * (uint64_t)blockDim.x + threadIdx.x) / G; g < span; g += groups_per_grid) {
    if (g < n) {
      const uint32_t v = list[g];
      const Word cv = cand[v];
      const uint64_t beg = off[v], end = off[v + 1];
      const uint8_t vcode = code[v];
      const uint8_t f = flags[v];
      Word U = 0;
      for (Word m = cv; m; m &= m - 1) U |= static_cast<Word>(tab.nbr[ctz<Word>(m)]);
      uint64_t cnt[KW];
#pragma unroll
      for (int w = 0; w < KW; ++w) cnt[w] = 0;
      uint32_t elim = 0;
      for (uint64_t base = beg; base < end; base += static_cast<uint64_t>(G) * S) {
        uint32_t u[S];
        uint32_t act = 0;  // bit k: slot k of this lane exists and is active
        Word cu[S];
        if (kSlot) {
          // First LCC step after LCC0 on a fresh state: a neighbour's
          // candidates are exactly the template vertices of its label, so
          // the slot's label index replaces the gather of cand[u].
          uint8_t sl[S];
#pragma unroll
          for (int k = 0; k < S; ++k) {
            const uint64_t e = base + k * G + gl;
            sl[k] = e < end ? slot_lab[e] : uint8_t(0xFF);
            const uint32_t word = e < end ? ea[e >> 5] : 0u;
            act |= ((word >> (e & 31)) & 1u) << k;
          }
#pragma unroll
          for (int k = 0; k < S; ++k) {
            const uint8_t c = (act >> k) & 1u ? tab.slot_code[sl[k]] : kNoCode;
            cu[k] = c == kNoCode ? Word(0) : static_cast<Word>(tab.cv[c]);
            u[k] = 0;
          }
        } else {
#pragma unroll
          for (int k = 0; k < S; ++k) {
            const uint64_t e = base + k * G + gl;
            u[k] = e < end ? edges[e] : 0u;
            const uint32_t word = e < end ? ea[e >> 5] : 0u;
            act |= ((word >> (e & 31)) & 1u) << k;
          }
#pragma unroll
          for (int k = 0; k < S; ++k) cu[k] = (act >> k) & 1u ? cand[u[k]] : Word(0);
        }
#pragma unroll
        for (int k = 0; k < S; ++k) {
          bool el = false;
          if (cu[k]) {
            if ((cu[k] & U) == 0) {
              el = true;
              ++elim;
            } else {
              const uint32_t lab = derive || kSlot ? tab.pcode[ctz<Word>(cu[k])] : code[u[k]];
              if (lab != kNoCode) {
#pragma unroll
                for (int w = 0; w < KW; ++w)
                  if (static_cast<int>(lab >> 3) == w) cnt[w] += 1ull << (8 * (lab & 7));
              }
            }
          }
          if (G == 1) {
            const uint64_t e = base + k;
            if (el) atomicAnd(&ea[e >> 5], ~(1u << (e & 31)));
          } else {
            clear_slots<G>(ea, base + k * G, el, gl, gbase);
          }
        }
      }
#pragma unroll
      for (int o = 1; o < G; o <<= 1)
#pragma unroll
        for (int w = 0; w < KW; ++w) cnt[w] = sat_add_bytes(cnt[w], __shfl_xor(cnt[w], o, 64));
      const bool any_elim = G == 1 ? elim != 0 : (__ballot(elim != 0) & gmask) != 0;
      elim_total += elim;
      // every lane of the group holds the same counts, so each can evaluate
      const Word fail = lcc_requirements<Word, KW, KW>(cv, vcode, cnt, tab.pcode, tab.need);
      if (gl == 0) {
        slots_total += end - beg;
        if (fail) unmatch[v] = fail;
        if (fail || any_elim) flags[v] = f | kModified;
        if (dirty) dirty[v] = 0;
      }
      if (rescan && (fail || any_elim) && !(f & kModified)) {
        // v turns modified: the phase 2 scans of its neighbours can change.
        // Eight ids per lane are requested before their marks are stored.
        for (uint64_t e0 = beg + gl; e0 < end; e0 += static_cast<uint64_t>(G) * 8) {
          uint32_t x[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const uint64_t e = e0 + static_cast<uint64_t>(k) * G;
            x[k] = e < end ? edges[e] : 0u;
          }
#pragma unroll
          for (int k = 0; k < 8; ++k)
            if (e0 + static_cast<uint64_t>(k) * G < end) rescan[x[k]] = 1;
        }
        if (gl == 0) xslots += end - beg;
      }
    }
  }
  wave_add(counters, C_EDGE, elim_total);
  wave_add(counters, C_SLOTS, slots_total);
  wave_add(counters, C_XSLOTS, xslots);
}

// Hub rows: the static chunk list holds (vertex, first slot) per chunk of
// kHubChunk slots; chunks of a vertex outside the current frontier exit.
// Label counts and eliminations are summed in hub_count[h][code] and
// hub_elim[h]; k_lcc_hub_finish evaluates the requirements and clears them.

// `derive`: neighbour codes from candidates, as in lcc_phase1_lanes.
template <typename Word, int KW, bool kSlot>
__device__ __forceinline__ void lcc_hub_chunks(uint32_t block, uint32_t nblocks, const uint8_t* s_pcode, const uint8_t* s_scode,
                               const uint64_t* s_cv,
                               const uint32_t* __restrict__ chunk_hub, const uint64_t* __restrict__ chunk_start,
                               uint32_t nchunks, const uint32_t* __restrict__ hubs, const uint64_t* __restrict__ off,
                               const uint32_t* __restrict__ edges, const uint8_t* __restrict__ code, bool derive,
                               const uint8_t* __restrict__ slot_lab,
                               const Word* __restrict__ cand, const uint8_t* __restrict__ flags,
                               const uint8_t* __restrict__ dirty, uint32_t* __restrict__ ea,
                               uint32_t* __restrict__ hub_count, uint32_t* __restrict__ hub_elim,
                               const DevTemplate* __restrict__ t, unsigned long long* counters) {
  // Each thread owns slots threadIdx.x + k * kBlock of the chunk (k < kPer);
  // ids and edge-active words of all of them are loaded before the gathers.
  constexpr int kPer = kHubChunk / kBlock;
  const int lane = lane_id();
  uint64_t elim_total = 0, slots_total = 0, items = 0;
  // The activity of up to kBlock of this block's chunks (block + k * nblocks)
  // is tested at once, one thread each; that thread also leaves the chunk's
  // hub, slot range and U (the template neighbours of the hub's candidates)
  // in LDS, so the block then works through the active chunks in order
  // without the chain of loads that leads to them.  The hubs' candidates,
  // flags and dirty marks do not change during the launch.
  __shared__ uint64_t s_act[kBlock / 64];
  __shared__ uint32_t s_h[kBlock];
  __shared__ uint64_t s_beg[kBlock], s_end[kBlock];
  __shared__ Word s_U[kBlock];
  const uint32_t per = block < nchunks ? (nchunks - 1 - block) / nblocks + 1 : 0;
  for (uint32_t k0 = 0; k0 < per; k0 += kBlock) {
    {
      const uint32_t kk = k0 + threadIdx.x;
      bool act = false;
      if (kk < per) {
        const uint32_t c = block + kk * nblocks;
        const uint32_t h = chunk_hub[c];
        const uint32_t v = hubs[h];
        const Word cv = cand[v];
        act = cv && (flags[v] & kScheduled) && (!dirty || dirty[v]);
        if (act) {
          Word U = 0;
          for (Word m = cv; m; m &= m - 1) U |= static_cast<Word>(t->nbr[ctz<Word>(m)]);
          const uint64_t beg = chunk_start[c];
          s_h[threadIdx.x] = h;
          s_beg[threadIdx.x] = beg;
          s_end[threadIdx.x] = min(off[v + 1], beg + kHubChunk);
          s_U[threadIdx.x] = U;
        }
      }
      const uint64_t am = __ballot(act);
      if (lane_id() == 0) s_act[threadIdx.x >> 6] = am;
      __syncthreads();
    }
    for (int aw = 0; aw < kBlock / 64; ++aw) {
      for (uint64_t mm = s_act[aw]; mm; mm &= mm - 1) {
        const int slot = aw * 64 + __builtin_ctzll(mm);
        const uint32_t h = s_h[slot];
        const Word U = s_U[slot];
        const uint64_t beg = s_beg[slot];
        const uint64_t end = s_end[slot];
        uint32_t word[kPer], u[kPer];
        Word cu[kPer];
        uint8_t lb[kPer];
        if (kSlot) {  // first LCC step after LCC0: labels of the slots, as in lcc_phase1_lanes
#pragma unroll
          for (int k = 0; k < kPer; ++k) {
            const uint64_t e = beg + threadIdx.x + k * kBlock;
            word[k] = e < end ? ea[e >> 5] : 0u;
            lb[k] = e < end ? slot_lab[e] : uint8_t(0xFF);
          }
#pragma unroll
          for (int k = 0; k < kPer; ++k) {
            const uint64_t e = beg + threadIdx.x + k * kBlock;
            const bool act = (word[k] >> (e & 31)) & 1u;
            const uint8_t c = act ? s_scode[lb[k]] : kNoCode;
            cu[k] = c == kNoCode ? Word(0) : static_cast<Word>(s_cv[c]);
            lb[k] = c;
          }
        } else {
#pragma unroll
          for (int k = 0; k < kPer; ++k) {
            const uint64_t e = beg + threadIdx.x + k * kBlock;
            word[k] = e < end ? ea[e >> 5] : 0u;
            u[k] = e < end ? edges[e] : 0u;
          }
#pragma unroll
          for (int k = 0; k < kPer; ++k) {
            const uint64_t e = beg + threadIdx.x + k * kBlock;
            const bool act = (word[k] >> (e & 31)) & 1u;
            cu[k] = act ? cand[u[k]] : Word(0);
            lb[k] = act && !derive ? code[u[k]] : kNoCode;
          }
        }
        if (derive && !kSlot) {
#pragma unroll
          for (int k = 0; k < kPer; ++k)
            if (cu[k]) lb[k] = s_pcode[ctz<Word>(cu[k])];
        }
        // label counts in packed bytes (at most kPer per lane), summed over
        // the wave with saturating byte adds (capped partials keep the
        // finishing kernel's min(count, 127) exact); lane L then holds the
        // count of label code L
        uint64_t cnt[KW];
#pragma unroll
        for (int w = 0; w < KW; ++w) cnt[w] = 0;
        uint32_t elim = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
          bool el = false;
          if (cu[k]) {
            if ((cu[k] & U) == 0) {
              el = true;
              ++elim;
            } else if (lb[k] != kNoCode) {
#pragma unroll
              for (int w = 0; w < KW; ++w)
                if (static_cast<int>(lb[k] >> 3) == w) cnt[w] += 1ull << (8 * (lb[k] & 7));
            }
          }
          clear_slots<64>(ea, beg + k * kBlock + (threadIdx.x & ~63u), el, lane, 0);
        }
#pragma unroll
        for (int o = 1; o < 64; o <<= 1)
#pragma unroll
          for (int w = 0; w < KW; ++w) cnt[w] = sat_add_bytes(cnt[w], __shfl_xor(cnt[w], o, 64));
        uint32_t mine = 0;
#pragma unroll
        for (int w = 0; w < KW; ++w)
          if ((lane >> 3) == w) mine = static_cast<uint32_t>(cnt[w] >> (8 * (lane & 7))) & 0xFFu;
        if (mine) atomicAdd(&hub_count[h * kMaxLabels + lane], mine);
        if (__ballot(elim != 0)) {
          elim = __ockl_wfred_add_u32(elim);
          if (lane == 0) atomicAdd(&hub_elim[h], elim);
        }
        elim_total += lane == 0 ? elim : 0u;
        if (threadIdx.x == 0) {
          slots_total += end - beg;
          ++items;
        }
      }
    }
    __syncthreads();
  }
  wave_add(counters, C_EDGE, elim_total);
  wave_add(counters, C_SLOTS, slots_total);
  wave_add(counters, C_ITEMS, items);
}

// Hub rows of LCC phase 1 in a launch of their own (PJ_PHASE1_FUSED=0;
// by default k_lcc_phase1_hubs runs them in the launch of the bins).  Both
// kernels inline the chunk and bin bodies: called from two kernels, the
// compiler otherwise keeps them as functions with a stack in scratch.
template <typename Word, int KW, bool kSlot>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(KW == 1 ? 7 : KW == 8 && sizeof(Word) == 8 ? 5 : 6))) k_lcc_hub_chunks(
    const uint32_t* __restrict__ chunk_hub, const uint64_t* __restrict__ chunk_start, uint32_t nchunks,
    const uint32_t* __restrict__ hubs, const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges,
    const uint8_t* __restrict__ code, bool derive, const uint8_t* __restrict__ slot_lab,
    const Word* __restrict__ cand, const uint8_t* __restrict__ flags,
    const uint8_t* __restrict__ dirty, uint32_t* __restrict__ ea, uint32_t* __restrict__ hub_count,
    uint32_t* __restrict__ hub_elim, const DevTemplate* __restrict__ t, unsigned long long* counters) {
  __shared__ uint8_t s_pcode[64];
  __shared__ uint8_t s_scode[256];
  __shared__ uint64_t s_cv[kMaxLabels];
  if (threadIdx.x < 64) {
    s_pcode[threadIdx.x] = t->pcode[threadIdx.x];
    s_cv[threadIdx.x] = t->code_vertices[threadIdx.x];
  }
  for (int i = threadIdx.x; i < 256; i += blockDim.x) s_scode[i] = t->slot_code[i];
  __syncthreads();
  lcc_hub_chunks<Word, KW, kSlot>(blockIdx.x, gridDim.x, s_pcode, s_scode, s_cv, chunk_hub, chunk_start, nchunks, hubs, off, edges,
                       code, derive, slot_lab, cand, flags, dirty, ea, hub_count, hub_elim, t, counters);
}

// LCC phase 1 of the degree bins in one launch, each bin grid-strided over
// the whole grid.
template <typename Word, int KW, bool kSlot>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(Word) == 8 && !kSlot && KW <= 2 ? 6 : KW == 1 && !kSlot ? 8 : KW <= 2 ? 7 : 1))) k_lcc_phase1_all(
    const uint32_t* __restrict__ bins, const uint32_t* __restrict__ bin_count, uint32_t V,
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges, const uint8_t* __restrict__ code,
    bool derive, const uint8_t* __restrict__ slot_lab, const Word* __restrict__ cand, Word* __restrict__ unmatch,
    uint8_t* __restrict__ flags, uint8_t* __restrict__ dirty, uint8_t* __restrict__ rescan, uint32_t* __restrict__ ea,
    const DevTemplate* __restrict__ t, unsigned long long* counters) {
  __shared__ LccTables<KW> tab;
  load_lcc_tables<KW>(tab, t);
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0)
    atomicAdd(&counters[C_ITEMS],
              static_cast<unsigned long long>(bin_count[0]) + bin_count[1] + bin_count[2] + bin_count[3]);
  const uint64_t V64 = V;
  // bins: 257-2048 slots (64 lanes x 4 slots per pass), 33-256 (32 x 4),
  // 5-32 (4 x 4), 1-4 (a lane per vertex).  Four slots per lane keep the
  // one-word build within 64 VGPRs (8 waves per SIMD) without spills; eight
  // per lane needed 82 (5 waves) and measured slower (1.05 against 1.01 ms
  // per scale-22 solve).  The slot-label and two-word builds need up to 72
  // (7 waves): held to 64 they spilled 40 bytes per lane to scratch, written
  // and read by every lane of every launch.
  lcc_phase1_lanes<Word, 64, 4, KW, kSlot>(blockIdx.x, gridDim.x, tab, bins + 3 * V64, bin_count + 3, off, edges, code,
                                    derive, slot_lab, cand, unmatch, flags, dirty, rescan, ea, t, counters);
  lcc_phase1_lanes<Word, 32, 4, KW, kSlot>(blockIdx.x, gridDim.x, tab, bins + 2 * V64, bin_count + 2, off, edges, code,
                                    derive, slot_lab, cand, unmatch, flags, dirty, rescan, ea, t, counters);
  lcc_phase1_lanes<Word, 4, 4, KW, kSlot>(blockIdx.x, gridDim.x, tab, bins + V64, bin_count + 1, off, edges, code, derive,
                                   slot_lab, cand, unmatch, flags, dirty, rescan, ea, t, counters);
  lcc_phase1_lanes<Word, 1, 4, KW, kSlot>(blockIdx.x, gridDim.x, tab, bins, bin_count + 0, off, edges, code, derive,
                                   slot_lab, cand, unmatch, flags, dirty, rescan, ea, t, counters);
}

// LCC phase 1 of the hub chunks and the degree bins in one launch: each
// block first takes its share of the hub chunks, then its share of every
// bin, so the blocks still on hub rows run beside the blocks already on the
// bins (both parts wait on chains of dependent loads, neither saturates the
// memory system alone) and no launch boundary separates the two.  Hubs are
// never in a bin, so the two parts touch disjoint rows and vertices.
template <typename Word, int KW, bool kSlot>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(Word) == 8 && !kSlot && KW <= 2 ? 6 : KW == 1 && !kSlot ? 8 : KW <= 2 ? 7 : 1))) k_lcc_phase1_hubs(
    const uint32_t* __restrict__ bins, const uint32_t* __restrict__ bin_count, uint32_t V,
    const uint32_t* __restrict__ chunk_hub, const uint64_t* __restrict__ chunk_start, uint32_t nchunks,
    const uint32_t* __restrict__ hubs, uint32_t* __restrict__ hub_count, uint32_t* __restrict__ hub_elim,
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges, const uint8_t* __restrict__ code,
    bool derive, const uint8_t* __restrict__ slot_lab, const Word* __restrict__ cand, Word* __restrict__ unmatch,
    uint8_t* __restrict__ flags, uint8_t* __restrict__ dirty, uint8_t* __restrict__ rescan, uint32_t* __restrict__ ea,
    const DevTemplate* __restrict__ t, unsigned long long* counters) {
  __shared__ LccTables<KW> tab;
  load_lcc_tables<KW>(tab, t);
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0)
    atomicAdd(&counters[C_ITEMS],
              static_cast<unsigned long long>(bin_count[0]) + bin_count[1] + bin_count[2] + bin_count[3]);
  lcc_hub_chunks<Word, KW, kSlot>(blockIdx.x, gridDim.x, tab.pcode, tab.slot_code, tab.cv, chunk_hub, chunk_start,
                                  nchunks, hubs, off, edges, code, derive, slot_lab, cand, flags, dirty, ea,
                                  hub_count, hub_elim, t, counters);
  const uint64_t V64 = V;
  lcc_phase1_lanes<Word, 64, 4, KW, kSlot>(blockIdx.x, gridDim.x, tab, bins + 3 * V64, bin_count + 3, off, edges, code,
                                    derive, slot_lab, cand, unmatch, flags, dirty, rescan, ea, t, counters);
  lcc_phase1_lanes<Word, 32, 4, KW, kSlot>(blockIdx.x, gridDim.x, tab, bins + 2 * V64, bin_count + 2, off, edges, code,
                                    derive, slot_lab, cand, unmatch, flags, dirty, rescan, ea, t, counters);
  lcc_phase1_lanes<Word, 4, 4, KW, kSlot>(blockIdx.x, gridDim.x, tab, bins + V64, bin_count + 1, off, edges, code, derive,
                                   slot_lab, cand, unmatch, flags, dirty, rescan, ea, t, counters);
  lcc_phase1_lanes<Word, 1, 4, KW, kSlot>(blockIdx.x, gridDim.x, tab, bins, bin_count + 0, off, edges, code, derive,
                                   slot_lab, cand, unmatch, flags, dirty, rescan, ea, t, counters);
}

template <typename Word, int KW>
__global__ void k_lcc_hub_finish(const uint32_t* __restrict__ hubs, uint32_t nhubs, const uint8_t* __restrict__ code,
                                 const Word* __restrict__ cand, Word* __restrict__ unmatch, uint8_t* __restrict__ flags,
                                 uint8_t* __restrict__ dirty, uint32_t* __restrict__ hub_count,
                                 uint32_t* __restrict__ hub_elim, uint32_t* __restrict__ hub_new,
                                 const DevTemplate* __restrict__ t) {
  for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < nhubs; h += gridDim.x * blockDim.x) {
    const uint32_t v = hubs[h];
    const Word cv = cand[v];
    uint64_t cnt[KW];
#pragma unroll
    for (int w = 0; w < KW; ++w) cnt[w] = 0;
    for (uint32_t k = 0; k < t->K; ++k) {
      uint32_t c = min(hub_count[h * kMaxLabels + k], 127u);
      hub_count[h * kMaxLabels + k] = 0;
      cnt[k >> 3] |= static_cast<uint64_t>(c) << (8 * (k & 7));
    }
    const uint32_t elim = hub_elim[h];
    hub_elim[h] = 0;
    hub_new[h] = 0;
    const uint8_t f = flags[v];
    if (!cv || !(f & kScheduled) || (dirty && !dirty[v])) continue;
    const Word fail = lcc_requirements<Word, KW>(cv, code[v], cnt, t);
    if (fail) unmatch[v] = fail;
    if (fail || elim) {
      flags[v] = f | kModified;
      hub_new[h] = !(f & kModified);  // k_hub_rescan_push marks the neighbours
    }
    if (dirty) dirty[v] = 0;
  }
}

// The push of slot e = v->u of a modified vertex v: with reverse positions
// (`first` given), u's phase 2 scan would stop at the slot u->v, position
// rev[e] of row(u), if that slot is active: first[u] keeps the smallest such
// position over all modified neighbours.  Otherwise u is marked for a rescan.
__device__ __forceinline__ void push_one(uint64_t e, const uint32_t* __restrict__ edges,
                                         const uint64_t* __restrict__ off, uint8_t* __restrict__ rescan,
                                         const uint32_t* __restrict__ ea, const uint32_t* __restrict__ rev,
                                         uint32_t* __restrict__ first) {
  const uint32_t u = edges[e];
  if (!first) {
    rescan[u] = 1;
    return;
  }
  const uint32_t p = rev[e];
  // A plain load of first[u] beside off[u]: most pushes to a vertex of many
  // modified neighbours cannot lower its minimum (a stale value is only
  // larger) and skip the slot test and the atomic.
  const uint64_t ou = off[u];
  const uint32_t cur = first[u];
  if (p < cur && ea_test(ea, ou + p)) atomicMin(&first[u], p);
}

// Hubs that turned modified in the last phase 1 (hub_new), or every modified
// hub (flags given), push to the vertices of their rows (push_one), one block
// per chunk of their rows.
__global__ void __launch_bounds__(kBlock) k_hub_rescan_push(
    const uint32_t* __restrict__ chunk_hub, const uint64_t* __restrict__ chunk_start, uint32_t nchunks,
    const uint32_t* __restrict__ hubs, const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges,
    const uint32_t* __restrict__ hub_new, const uint8_t* __restrict__ flags, uint8_t* __restrict__ rescan,
    const uint32_t* __restrict__ ea, const uint32_t* __restrict__ rev, uint32_t* __restrict__ first,
    unsigned long long* counters) {
  uint64_t pushed = 0;
  // The activity of up to kBlock of this block's chunks (blockIdx.x + k * gridDim.x)
  // is tested at once, one thread each; the block then works through the
  // active ones in order.
  __shared__ uint64_t s_act[kBlock / 64];
  const uint32_t per = blockIdx.x < nchunks ? (nchunks - 1 - blockIdx.x) / gridDim.x + 1 : 0;
  for (uint32_t k0 = 0; k0 < per; k0 += kBlock) {
    {
      const uint32_t kk = k0 + threadIdx.x;
      bool act = false;
      if (kk < per) {
        const uint32_t c = blockIdx.x + kk * gridDim.x;
        const uint32_t h = chunk_hub[c];
        act = flags ? (flags[hubs[h]] & kModified) != 0 : hub_new[h] != 0;
      }
      const uint64_t am = __ballot(act);
      if (lane_id() == 0) s_act[threadIdx.x >> 6] = am;
      __syncthreads();
    }
    for (int aw = 0; aw < kBlock / 64; ++aw) {
      for (uint64_t mm = s_act[aw]; mm; mm &= mm - 1) {
        const uint32_t c = blockIdx.x + (k0 + aw * 64 + static_cast<uint32_t>(__builtin_ctzll(mm))) * gridDim.x;
        const uint32_t h = chunk_hub[c];
        if (flags ? !(flags[hubs[h]] & kModified) : !hub_new[h]) continue;  // flags given: every modified hub
        const uint32_t v = hubs[h];
        const uint64_t beg = chunk_start[c];
        const uint64_t end = min(off[v + 1], beg + kHubChunk);
        for (uint64_t e = beg + threadIdx.x; e < end; e += blockDim.x) push_one(e, edges, off, rescan, ea, rev, first);
        if (threadIdx.x == 0) pushed += end - beg;
      }
    }
    __syncthreads();
  }
  wave_add(counters, C_XSLOTS, pushed);
}

// Vertices of a bin (mode 5: modified) push to the vertices of their rows
// (push_one), one group of G lanes per vertex.
template <int G>
__global__ void __launch_bounds__(kBlock) k_push_rescan(const uint32_t* __restrict__ list,
                                                        const uint32_t* __restrict__ list_count,
                                                        const uint64_t* __restrict__ off,
                                                        const uint32_t* __restrict__ edges,
                                                        uint8_t* __restrict__ rescan, const uint32_t* __restrict__ ea,
                                                        const uint32_t* __restrict__ rev, uint32_t* __restrict__ first,
                                                        unsigned long long* counters) {
  const int gl = lane_id() & (G - 1);
  const uint32_t n = *list_count;
  const uint64_t groups_per_grid = (uint64_t)gridDim.x * (blockDim.x / G);
  uint64_t pushed = 0;
  for (uint64_t g = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / G; g < n; g += groups_per_grid) {
    const uint32_t v = list[g];
    const uint64_t beg = off[v], end = off[v + 1];
    for (uint64_t e = beg + gl; e < end; e += G) push_one(e, edges, off, rescan, ea, rev, first);
    if (gl == 0) pushed += end - beg;
  }
  wave_add(counters, C_XSLOTS, pushed);
}

// ---------------------------------------------------------------------------
// Phase 2, shared by LCC and every NLCC step (lcc_cpu.tpp:107-142,
// cc_cpu.tpp:134-174): a modified vertex drops its toUnmatch set, is
// deactivated when nothing is left and is scheduled; any other active vertex
// is scheduled iff one of its active slots leads to a modified vertex
// (active or not).  Modified flags are left set.

template <typename Word, int G>
__device__ void phase2_bin(uint32_t block, uint32_t nblocks, const uint32_t* __restrict__ list,
                           const uint32_t* __restrict__ list_count, const uint64_t* __restrict__ off,
                           const uint32_t* __restrict__ edges, Word* __restrict__ cand, Word* __restrict__ unmatch,
                           uint8_t* __restrict__ flags, uint8_t* __restrict__ dirty, uint8_t* __restrict__ rescan,
                           uint32_t* __restrict__ last_scan, const uint32_t* __restrict__ ea,
                           unsigned long long* counters) {
  const int lane = lane_id();
  const int gl = lane & (G - 1);
  const int gbase = lane & ~(G - 1);
  const uint64_t gmask = G == 64 ? ~0ull : (((1ull << G) - 1) << gbase);
  const uint32_t n = *list_count;
  const uint64_t groups_per_grid = (uint64_t)nblocks * (blockDim.x / G);
  const uint64_t span = ((n + groups_per_grid - 1) / groups_per_grid) * groups_per_grid;
  uint64_t velim = 0, melim = 0, slots = 0, xslots = 0;
  for (uint64_t g = (block * (uint64_t)blockDim.x + threadIdx.x) / G; g < span; g += groups_per_grid) {
    if (g < n) {
      const uint32_t v = list[g];
      const uint8_t f = flags[v];
      if (f & kModified) {
        const Word tu = unmatch[v];  // group-uniform
        if (gl == 0) {
          if (tu) {
            const Word c = cand[v];
            const Word nc = c & ~tu;
            melim += popc<Word>(tu);
            cand[v] = nc;
            if (!nc) ++velim;
            unmatch[v] = 0;
          }
          flags[v] = f | kScheduled;
        }
        if (tu && dirty) {  // v's candidates change: v and every vertex reading them re-evaluate
          const uint64_t beg = off[v], end = off[v + 1];
          if (gl == 0) {
            dirty[v] = 1;
            xslots += end - beg;
          }
          for (uint64_t e = beg + gl; e < end; e += G) dirty[edges[e]] = 1;
        }
      } else {
        const uint64_t beg = off[v], end = off[v + 1];
        bool hit = false;
        uint64_t scanned = end - beg;
        // kUnroll slices of G slots per pass, loads batched as in phase 1;
        // the first hit in slot order ends the scan.
        constexpr int kUnroll = G >= 32 ? 4 : G == 16 ? 2 : 1;
        for (uint64_t base = beg; base < end && !hit; base += static_cast<uint64_t>(G) * kUnroll) {
          uint32_t word[kUnroll], u[kUnroll];
#pragma unroll
          for (int k = 0; k < kUnroll; ++k) {
            const uint64_t e = base + k * G + gl;
            word[k] = e < end ? ea[e >> 5] : 0u;
            u[k] = e < end ? edges[e] : 0u;
          }
          uint8_t fl[kUnroll];
#pragma unroll
          for (int k = 0; k < kUnroll; ++k) {
            const uint64_t e = base + k * G + gl;
            fl[k] = ((word[k] >> (e & 31)) & 1u) ? flags[u[k]] : uint8_t(0);
          }
#pragma unroll
          for (int k = 0; k < kUnroll; ++k) {
            const uint64_t m = __ballot((fl[k] & kModified) != 0) & gmask;
            if (m && !hit) {
              hit = true;
              scanned = base + k * G + (__builtin_ctzll(m) - gbase) + 1 - beg;
            }
          }
        }
        if (gl == 0) {
          flags[v] = hit ? (f | kScheduled) : (f & ~kScheduled);
          slots += scanned;
          if (rescan) {
            last_scan[v] = static_cast<uint32_t>(scanned);
            rescan[v] = 0;
          }
        }
      }
    }
  }
  wave_add(counters, C_VERTEX, velim);
  wave_add(counters, C_MATCH, melim);
  wave_add(counters, C_SLOTS, slots);
  wave_add(counters, C_XSLOTS, xslots);
}

// Hub phase 2: a hub that is not modified is scheduled when any chunk finds
// an active slot to a modified vertex.  Chunks take the minimum offset of
// such a slot in the row (hub_hit[h], 0xFFFFFFFF when none) and
// k_phase2_hub_finish applies the result and resets hub_hit.
template <typename Word>
__device__ void phase2_hub_chunks(uint32_t block, uint32_t nblocks, const uint32_t* __restrict__ chunk_hub,
                                  const uint64_t* __restrict__ chunk_start, uint32_t nchunks,
                                  const uint32_t* __restrict__ hubs, const uint64_t* __restrict__ off,
                                  const uint32_t* __restrict__ edges, const Word* __restrict__ cand,
                                  const Word* __restrict__ unmatch, const uint8_t* __restrict__ flags,
                                  uint8_t* __restrict__ dirty, const uint8_t* __restrict__ rescan,
                                  const uint32_t* __restrict__ ea, uint32_t* __restrict__ hub_hit,
                                  unsigned long long* counters) {
  uint64_t examined = 0;  // slots read here; the reference-order count is taken in k_phase2_hub_finish
  // The activity of up to kBlock of this block's chunks (block + k * nblocks)
  // is tested at once, one thread each; that thread also leaves the chunk's
  // hub, row start, slot range and kind (a modified hub dirtying its
  // neighbours, or a scan) in LDS, and the block then works through the
  // active chunks in order.  Hub state does not change during the launch
  // (phase 2 of the hubs is applied by k_phase2_hub_finish).
  __shared__ uint64_t s_act[kBlock / 64];
  __shared__ uint32_t s_h[kBlock];
  __shared__ uint64_t s_row[kBlock], s_beg[kBlock], s_end[kBlock];
  __shared__ uint8_t s_mod[kBlock];
  const uint32_t per = block < nchunks ? (nchunks - 1 - block) / nblocks + 1 : 0;
  for (uint32_t k0 = 0; k0 < per; k0 += kBlock) {
    {
      const uint32_t kk = k0 + threadIdx.x;
      bool act = false;
      if (kk < per) {
        const uint32_t c = block + kk * nblocks;
        const uint32_t h = chunk_hub[c];
        const uint32_t v = hubs[h];
        const bool mod = (flags[v] & kModified) != 0;
        if (cand[v]) act = mod ? (dirty && unmatch[v]) : (!rescan || rescan[v]);
        if (act) {
          const uint64_t row = off[v], beg = chunk_start[c];
          s_h[threadIdx.x] = h;
          s_row[threadIdx.x] = row;
          s_beg[threadIdx.x] = beg;
          s_end[threadIdx.x] = min(off[v + 1], beg + kHubChunk);
          s_mod[threadIdx.x] = mod;
        }
      }
      const uint64_t am = __ballot(act);
      if (lane_id() == 0) s_act[threadIdx.x >> 6] = am;
      __syncthreads();
    }
    for (int aw = 0; aw < kBlock / 64; ++aw) {
      for (uint64_t mm = s_act[aw]; mm; mm &= mm - 1) {
        const int slot = aw * 64 + __builtin_ctzll(mm);
        const uint32_t h = s_h[slot];
        const uint64_t beg = s_beg[slot];
        const uint64_t end = s_end[slot];
        if (s_mod[slot]) {  // the hub's candidates change: dirty its neighbours
          for (uint64_t e = beg + threadIdx.x; e < end; e += blockDim.x) dirty[edges[e]] = 1;
          if (threadIdx.x == 0) examined += end - beg;
          continue;
        }
        const uint64_t row = s_row[slot];
        // A hit already found in an earlier chunk of the row decides the scan:
        // this chunk cannot lower the minimum (one load per wave).
        if (*reinterpret_cast<const volatile uint32_t*>(&hub_hit[h]) < beg - row) continue;
        constexpr int kPer = kHubChunk / kBlock;
        uint32_t word[kPer], u[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
          const uint64_t e = beg + threadIdx.x + k * kBlock;
          word[k] = e < end ? ea[e >> 5] : 0u;
          u[k] = e < end ? edges[e] : 0u;
        }
        uint8_t fl[kPer];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
          const uint64_t e = beg + threadIdx.x + k * kBlock;
          fl[k] = ((word[k] >> (e & 31)) & 1u) ? flags[u[k]] : uint8_t(0);
        }
        uint32_t first = 0xFFFFFFFFu;  // this thread's lowest hit (row offset)
#pragma unroll
        for (int k = kPer - 1; k >= 0; --k)
          if (fl[k] & kModified) first = static_cast<uint32_t>(beg + threadIdx.x + k * kBlock - row);
        if (__ballot(first != 0xFFFFFFFFu)) {
          first = __ockl_wfred_min_u32(first);
          if (lane_id() == 0) atomicMin(&hub_hit[h], first);
        }
        if (threadIdx.x == 0) examined += end - beg;
      }
    }
    __syncthreads();
  }
  wave_add(counters, C_XSLOTS, examined);
}

// Phase 2 in one launch: hub chunks first, then the degree bins of G = 64,
// 16 and 4 lanes per vertex.  The hub chunks touch only hub rows and the
// bins only their own vertices; every vertex reads of its neighbours only
// the modified bit, which phase 2 leaves as it is.
template <typename Word>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) k_phase2_all(
    const uint32_t* __restrict__ bins, const uint32_t* __restrict__ bin_count, uint32_t V,
    const uint32_t* __restrict__ chunk_hub, const uint64_t* __restrict__ chunk_start, uint32_t nchunks,
    const uint32_t* __restrict__ hubs, const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges,
    Word* __restrict__ cand, Word* __restrict__ unmatch, uint8_t* __restrict__ flags, uint8_t* __restrict__ dirty,
    uint8_t* __restrict__ rescan, uint32_t* __restrict__ last_scan, const uint32_t* __restrict__ ea,
    uint32_t* __restrict__ hub_hit, unsigned long long* counters) {
  if (blockIdx.x == 0 && threadIdx.x == 0)
    atomicAdd(&counters[C_ITEMS],
              static_cast<unsigned long long>(bin_count[0]) + bin_count[1] + bin_count[2] + bin_count[3]);
  const uint64_t V64 = V;
  if (nchunks)
    phase2_hub_chunks<Word>(blockIdx.x, gridDim.x, chunk_hub, chunk_start, nchunks, hubs, off, edges, cand, unmatch,
                            flags, dirty, rescan, ea, hub_hit, counters);
  phase2_bin<Word, 64>(blockIdx.x, gridDim.x, bins + 3 * V64, bin_count + 3, off, edges, cand, unmatch, flags, dirty,
                       rescan, last_scan, ea, counters);
  phase2_bin<Word, 32>(blockIdx.x, gridDim.x, bins + 2 * V64, bin_count + 2, off, edges, cand, unmatch, flags, dirty,
                       rescan, last_scan, ea, counters);
  phase2_bin<Word, 16>(blockIdx.x, gridDim.x, bins + V64, bin_count + 1, off, edges, cand, unmatch, flags, dirty,
                       rescan, last_scan, ea, counters);
  phase2_bin<Word, 4>(blockIdx.x, gridDim.x, bins, bin_count + 0, off, edges, cand, unmatch, flags, dirty, rescan,
                      last_scan, ea, counters);
}

template <typename Word>
__global__ void k_phase2_hub_finish(const uint32_t* __restrict__ hubs, uint32_t nhubs,
                                    const uint64_t* __restrict__ off, Word* __restrict__ cand,
                                    Word* __restrict__ unmatch, uint8_t* __restrict__ flags,
                                    uint8_t* __restrict__ dirty, uint8_t* __restrict__ rescan,
                                    uint32_t* __restrict__ last_scan, uint32_t* __restrict__ hub_hit,
                                    unsigned long long* counters) {
  uint64_t velim = 0, melim = 0, slots = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t span = ((nhubs + stride - 1) / stride) * stride;
  for (uint64_t h = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; h < span; h += stride) {
    if (h >= nhubs) continue;
    const uint32_t v = hubs[h];
    const uint32_t hit = hub_hit[h];
    hub_hit[h] = 0xFFFFFFFFu;
    const Word c = cand[v];
    if (!c) continue;
    const uint8_t f = flags[v];
    if (f & kModified) {
      const Word tu = unmatch[v];
      if (tu) {
        const Word nc = c & ~tu;
        melim += popc<Word>(tu);
        cand[v] = nc;
        if (!nc) ++velim;
        unmatch[v] = 0;
        if (dirty) dirty[v] = 1;  // phase2_hub_chunks dirtied the neighbours
      }
      flags[v] = f | kScheduled;
    } else if (!rescan || rescan[v]) {
      const bool any = hit != 0xFFFFFFFFu;
      const uint64_t scanned = any ? hit + 1ull : off[v + 1] - off[v];
      flags[v] = any ? (f | kScheduled) : (f & ~kScheduled);
      slots += scanned;
      if (rescan) {
        last_scan[v] = static_cast<uint32_t>(scanned);
        rescan[v] = 0;
      }
    }
  }
  wave_add(counters, C_VERTEX, velim);
  wave_add(counters, C_MATCH, melim);
  wave_add(counters, C_SKIP, slots);  // reference-order slots of the hub scans (read by phase2_hub_chunks)
}

// ---------------------------------------------------------------------------
// Memoised walks: the reference's own search (cc_cpu.tpp:12-70,
// pc_cpu.tpp:11-75, tds_cpu.tpp:13-127).  Its per-source memos can prune
// branches a complete search would take, so when they can fire (always for
// TDS; for cycles and paths when the host analysis says so) the device runs
// the same search.  Each wave owns a region of an open-addressing table in
// global memory (64-bit key, 64-bit value, 32-bit epoch per entry); an entry
// is live while its epoch equals the wave's epoch, so a new walk (epoch + 1)
// starts from an empty table.  Keys carry a 2-bit kind in bits 62-63:
//   1: undirected edge {a, b}          -> template-id mask (edge memo)
//   2: (hypothesis node, vertex)       -> child node (hash-consed trie)
//   3: (vertex, template id, node)     -> present (hypothesis memo)
// A hypothesis (the STOREd vertices so far, in order) is a node of a trie
// whose nodes are hash-consed, so equal hypotheses have equal node ids.
// Vertex ids are < 2^30 and nodes < 2^26 (checked on the host / capacity).

constexpr uint8_t kMoveClose = 4;  // last hop of a cycle: an active slot back to the source
constexpr uint8_t kMoveLast = 5;   // last hop of a path: a candidate other than the source
constexpr int kMemoWaves = 2048;            // waves of a memoised launch (first pass)
constexpr int kMemoLog2 = 16;               // entries per wave, first pass
// Later passes, for sources whose walks outgrew the previous tables: fewer
// waves with larger tables (the last one holds 2^29 entries, 10.7 GB).
constexpr int kRetryPasses = 3;
constexpr int kRetryWaves[kRetryPasses] = {32, 4, 1};
constexpr int kRetryLog2[kRetryPasses] = {22, 26, 29};

struct MemoArgs {
  uint64_t* keys;
  uint64_t* vals;
  uint32_t* epochs;
  uint32_t* wave_epoch;  // per wave, persists across launches
  uint32_t log2;         // entries per wave = 1 << log2
  void* already;         // alreadyMatched (Word[V])
  uint32_t* ovf_list;    // sources whose table overflowed
  uint32_t* ovf_count;
  uint32_t lds_limit;    // live entries allowed in the LDS tables of k_memo_lds (0: 3/4 of them)
};

struct MemoTable {
  uint64_t* keys;
  uint64_t* vals;
  uint32_t* epochs;
  uint32_t mask;
  uint32_t limit;
  uint32_t epoch;
  uint32_t used;
  uint32_t next_node;
  bool overflow;
};

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

template <typename T>
__device__ __forceinline__ T wg_load(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename T>
__device__ __forceinline__ void wg_store(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Wave-cooperative lookup (64 consecutive probe slots per round).  Returns
// the slot holding `key`, or the free slot where it belongs; *found says
// which, and *val (when given) receives the value of a found key, read in
// the same round as the keys.  Every lane gets the same answer.
__device__ uint32_t memo_probe(const MemoTable& m, uint64_t key, bool* found, uint64_t* val = nullptr) {
  const int lane = lane_id();
  const uint32_t h = static_cast<uint32_t>(mix64(key)) & m.mask;
  for (uint32_t i = 0; i <= m.mask; i += 64) {
    const uint32_t slot = (h + i + lane) & m.mask;
    const uint32_t ep = wg_load(&m.epochs[slot]);
    const uint64_t k = wg_load(&m.keys[slot]);
    const uint64_t v = val ? wg_load(&m.vals[slot]) : 0ull;
    const bool live = ep == m.epoch;
    const bool hit = live && k == key;
    const uint64_t stop = __ballot(hit || !live);
    if (stop) {
      const int l = __builtin_ctzll(stop);
      *found = (__ballot(hit) >> l) & 1;
      if (val) *val = rl64(v, l);
      return rl32(slot, l);
    }
  }
  *found = false;
  return 0;  // not reached: the table never fills past its limit
}

// Whether one lane's candidate is blocked: edge memo of {a, u} holding `bit`,
// or (hyp) the hypothesis memo holding hkey.  One lane probes alone, by
// linear probing from the key's start, so it stops at the slot the
// wave-cooperative probe returns; lanes of a wave probe their own keys at
// once.  Both keys are probed in the same rounds (their loads, the edge
// value included, issued together).
__device__ bool memo_blocked_lane(const MemoTable& m, uint64_t ekey, uint32_t bit, bool hyp, uint64_t hkey) {
  uint32_t he = static_cast<uint32_t>(mix64(ekey)) & m.mask;
  uint32_t hh = static_cast<uint32_t>(mix64(hkey)) & m.mask;
  bool edone = false, hdone = !hyp;
  for (uint32_t i = 0; i <= m.mask && !(edone && hdone); ++i) {
    const uint32_t se = (he + i) & m.mask, sh = (hh + i) & m.mask;
    uint32_t epe = 0, eph = 0;
    uint64_t ke = 0, ve = 0, kh = 0;
    if (!edone) {
      epe = wg_load(&m.epochs[se]);
      ke = wg_load(&m.keys[se]);
      ve = wg_load(&m.vals[se]);
    }
    if (!hdone) {
      eph = wg_load(&m.epochs[sh]);
      kh = wg_load(&m.keys[sh]);
    }
    if (!edone) {
      if (epe != m.epoch) {
        edone = true;
      } else if (ke == ekey) {
        if ((ve >> bit) & 1) return true;
        edone = true;
      }
    }
    if (!hdone) {
      if (eph != m.epoch) {
        hdone = true;
      } else if (kh == hkey) {
        return true;
      }
    }
  }
  return false;
}

__device__ void memo_put(MemoTable& m, uint32_t slot, uint64_t key, uint64_t val) {
  if (lane_id() == 0) {
    wg_store(&m.vals[slot], val);
    wg_store(&m.keys[slot], key);
    wg_store(&m.epochs[slot], m.epoch);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  if (++m.used > m.limit) m.overflow = true;
}

__device__ __forceinline__ uint64_t edge_key(uint32_t a, uint32_t b) {
  const uint64_t lo = min(a, b), hi = max(a, b);
  return (1ull << 62) | (lo << 30) | hi;
}

// Edge memo test: was template id / position `bit` inserted for edge {a, b}?
__device__ bool memo_edge_has(const MemoTable& m, uint32_t a, uint32_t b, uint32_t bit) {
  bool found;
  uint64_t v = 0;
  memo_probe(m, edge_key(a, b), &found, &v);
  return found && ((v >> bit) & 1);
}

__device__ void memo_edge_add(MemoTable& m, uint32_t a, uint32_t b, uint32_t bit) {
  bool found;
  const uint64_t key = edge_key(a, b);
  uint64_t old = 0;
  const uint32_t slot = memo_probe(m, key, &found, &old);
  if (found) {
    const uint64_t v = old | (1ull << bit);
    if (lane_id() == 0) wg_store(&m.vals[slot], v);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  } else {
    memo_put(m, slot, key, 1ull << bit);
  }
}

__device__ __forceinline__ uint64_t hyp_key(uint32_t u, uint32_t y, uint32_t node) {
  return (3ull << 62) | (static_cast<uint64_t>(u) << 32) | (static_cast<uint64_t>(y) << 26) | node;
}

__device__ bool memo_hyp_has(const MemoTable& m, uint32_t u, uint32_t y, uint32_t node) {
  bool found;
  memo_probe(m, hyp_key(u, y, node), &found);
  return found;
}

__device__ void memo_hyp_add(MemoTable& m, uint32_t u, uint32_t y, uint32_t node) {
  bool found;
  const uint64_t key = hyp_key(u, y, node);
  const uint32_t slot = memo_probe(m, key, &found);
  if (!found) memo_put(m, slot, key, 0);
}

// The hypothesis node reached from `node` by storing vertex u.
__device__ uint32_t memo_trie_child(MemoTable& m, uint32_t node, uint32_t u) {
  bool found;
  const uint64_t key = (2ull << 62) | (static_cast<uint64_t>(node) << 30) | u;
  uint64_t v = 0;
  const uint32_t slot = memo_probe(m, key, &found, &v);
  if (found) return static_cast<uint32_t>(v);
  const uint32_t child = ++m.next_node;
  if (child >= (1u << 26) - 1) m.overflow = true;  // node ids are 26 bits in hypothesis keys
  memo_put(m, slot, key, child);
  return child;
}

// One lane's probe for `key` (linear probing from the key's start, so it
// stops where the wave-cooperative probe stops): the slot holding the key,
// with its value, or the first free slot.
__device__ uint32_t memo_probe_one(const MemoTable& m, uint64_t key, bool* found, uint64_t* val) {
  const uint32_t h = static_cast<uint32_t>(mix64(key)) & m.mask;
  for (uint32_t i = 0; i <= m.mask; ++i) {
    const uint32_t sl = (h + i) & m.mask;
    const uint32_t ep = wg_load(&m.epochs[sl]);
    const uint64_t k = wg_load(&m.keys[sl]);
    const uint64_t v = wg_load(&m.vals[sl]);
    if (ep != m.epoch) {
      *found = false;
      return sl;
    }
    if (k == key) {
      *found = true;
      *val = v;
      return sl;
    }
  }
  *found = false;
  return 0;  // not reached: the table never fills past its limit
}

// The memo updates of a walk step that moved from `cur` to u (template
// vertex y): with `store`, the hypothesis trie child of `node` by u; with
// `hyp`, the hypothesis (u, y, child); always the edge memo bit y of {cur, u}
// -- the same inserts memo_trie_child, memo_hyp_add and memo_edge_add make in
// that order, in three rounds instead of up to six: the trie and edge keys
// are probed together, then the hypothesis key, then every store is made
// with one fence.  A new key lands in the first free slot of its probe
// sequence; made one after the other, a later insert would probe past the
// slot an earlier one filled, so the slots agree unless two new keys found
// the same free slot: then nothing is stored and false is returned (the
// caller makes the inserts one by one).  *node_out receives the child.
__device__ bool memo_step_inserts(MemoTable& m, bool store, bool hyp, uint32_t node, uint32_t cur, uint32_t u,
                                  uint32_t y, uint32_t* node_out) {
  const int lane = lane_id();
  const uint64_t tkey = (2ull << 62) | (static_cast<uint64_t>(node) << 30) | u;
  const uint64_t ekey = edge_key(cur, u);
  uint32_t sl = 0;
  bool fnd = false;
  uint64_t val = 0;
  if ((lane == 0 && store) || lane == 1) sl = memo_probe_one(m, lane == 0 ? tkey : ekey, &fnd, &val);
  const uint32_t t_slot = rl32(sl, 0), e_slot = rl32(sl, 1);
  const bool t_found = rl32(fnd ? 1u : 0u, 0) != 0, e_found = rl32(fnd ? 1u : 0u, 1) != 0;
  const uint64_t t_val = rl64(val, 0), e_val = rl64(val, 1);
  const bool t_new = store && !t_found;
  const uint32_t node2 = !store ? node : t_found ? static_cast<uint32_t>(t_val) : m.next_node + 1;
  const uint64_t hkey = hyp_key(u, y, node2);
  uint32_t h_slot = 0;
  bool h_found = true;
  if (hyp) {
    uint32_t s2 = 0;
    bool f2 = false;
    uint64_t v2 = 0;
    if (lane == 0) s2 = memo_probe_one(m, hkey, &f2, &v2);
    h_slot = rl32(s2, 0);
    h_found = rl32(f2 ? 1u : 0u, 0) != 0;
  }
  const bool h_new = hyp && !h_found, e_new = !e_found;
  if ((t_new && e_new && t_slot == e_slot) || (t_new && h_new && t_slot == h_slot) ||
      (e_new && h_new && e_slot == h_slot))
    return false;
  if (t_new) {
    m.next_node = node2;
    if (node2 >= (1u << 26) - 1) m.overflow = true;  // node ids are 26 bits in hypothesis keys
  }
  if (lane == 0) {
    if (t_new) {
      wg_store(&m.vals[t_slot], static_cast<uint64_t>(node2));
      wg_store(&m.keys[t_slot], tkey);
      wg_store(&m.epochs[t_slot], m.epoch);
    }
    if (h_new) {
      wg_store(&m.vals[h_slot], uint64_t{0});
      wg_store(&m.keys[h_slot], hkey);
      wg_store(&m.epochs[h_slot], m.epoch);
    }
    wg_store(&m.vals[e_slot], static_cast<uint64_t>(e_val | (uint64_t{1} << y)));
    if (e_new) {
      wg_store(&m.keys[e_slot], ekey);
      wg_store(&m.epochs[e_slot], m.epoch);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  m.used += (t_new ? 1u : 0u) + (h_new ? 1u : 0u) + (e_new ? 1u : 0u);
  if (m.used > m.limit) m.overflow = true;
  *node_out = node2;
  return true;
}

// Failed-state memo of the complete search: a cycle or path walk has no
// visited set, so whether it can be finished from vertex v after d hops
// depends on (v, d) only (the source and start are fixed for the walk).
// States are remembered while the table has room; the memo only saves work.
__device__ __forceinline__ uint64_t state_key(uint32_t v, int depth) {
  return (1ull << 62) | (static_cast<uint64_t>(depth) << 32) | v;
}

__device__ bool state_failed(const MemoTable& m, uint32_t v, int depth) {
  bool found;
  memo_probe(m, state_key(v, depth), &found);
  return found;
}

__device__ void state_fail(MemoTable& m, uint32_t v, int depth) {
  if (m.used >= m.limit) return;
  bool found;
  const uint64_t key = state_key(v, depth);
  const uint32_t slot = memo_probe(m, key, &found);
  if (!found) memo_put(m, slot, key, 0);
}

template <typename Word>
__device__ __forceinline__ void atomic_or_word(Word* p, Word bits) {
  if (sizeof(Word) == 8) {
    atomicOr(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(bits));
  } else if (sizeof(Word) == 4) {
    atomicOr(reinterpret_cast<unsigned int*>(p), static_cast<unsigned int>(bits));
  } else {
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    atomicOr(reinterpret_cast<unsigned int*>(a & ~uintptr_t(3)), static_cast<unsigned int>(bits) << (8 * (a & 3)));
  }
}

// ---------------------------------------------------------------------------
// NLCC: one wavefront per source vertex runs a depth-first walk whose row
// scans are done 64 slots at a time (ballot, lowest lane first, so the walk
// visits neighbours in CSR order).  Level i of the walk keeps its vertex and
// its resume slot in lane i's registers (walks are at most 64 levels).  The
// search is complete: it decides whether a walk satisfying the constraint
// exists (see the NLCC notes in the oracle for when this equals the
// reference's memoised search).

struct RowScan {
  uint32_t u;
  uint64_t next;  // slot after the hit
  bool found;
};

// First active slot in [from, end) of a row whose neighbour is active and has
// template vertex q (q < 0: neighbour must equal `target`; `excl`: neighbour
// must differ from it).  Adds to `slots` the slots a serial scan reads (up to
// and including the hit).
template <typename Word>
__device__ __forceinline__ RowScan scan_row(uint64_t from, uint64_t end, const uint32_t* __restrict__ edges,
                                            const uint32_t* __restrict__ ea, const Word* __restrict__ cand, int q,
                                            uint32_t target, uint32_t excl, uint64_t& slots) {
  RowScan r{0, end, false};
  const int lane = lane_id();
  for (uint64_t base = from; base < end; base += 64) {
    const uint64_t e = base + lane;
    bool ok = false;
    uint32_t u = 0;
    if (e < end && walk_slot(ea, e)) {
      u = edges[e];
      if (q < 0) {
        ok = u == target;
      } else {
        Word cu = cand[u];
        ok = ((cu >> q) & 1) && u != excl;
      }
    }
    uint64_t m = __ballot(ok);
    if (m) {
      int first = __builtin_ctzll(m);
      r.u = rl32(u, first);
      r.next = base + first + 1;
      r.found = true;
      slots += r.next - from;
      return r;
    }
  }
  if (end > from) slots += end - from;
  return r;
}

// Whether row(a) holds an active slot to b, by binary search (rows sorted
// ascending; ea null: every slot is active); `probes` counts the slots read.
__device__ __forceinline__ bool active_slot(uint32_t a, uint32_t b, const uint64_t* __restrict__ off,
                                            const uint32_t* __restrict__ edges, const uint32_t* __restrict__ ea,
                                            uint64_t& probes) {
  uint64_t lo = off[a], hi = off[a + 1];
  const uint64_t end = hi;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    ++probes;
    if (edges[mid] < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  for (; lo < end; ++lo) {
    ++probes;
    if (edges[lo] != b) break;
    if (walk_slot(ea, lo)) return true;
  }
  return false;
}

// Slot positions between heavy vertices (rows above the index's degree
// threshold), built at upload for symmetric graphs with sorted rows:
// doff[v] = off[v] | (heavy id of v + 1) << kPairShift (no id: 0), and
// pos[ha * H + hb] = first slot of hb in row(ha) minus that row's start, or
// kNoPos.  A search between two heavy rows (each binary-search probe of a
// long row fetches its own cache line for four useful bytes) becomes one
// load.  pos == nullptr: no index (searches only).
constexpr int kPairShift = 40;
constexpr uint64_t kPairOffMask = (1ull << kPairShift) - 1;
constexpr uint32_t kNoPos = 0xFFFFFFFFu;
// The pair rounds of a 3-vertex cycle source start with this many pairs and
// double up to the group width: most sources that succeed do so in their
// first pairs, and every pair tested in the round after a success is read
// for nothing (modelled on the scale-22 plain workload's cycle step: rounds
// of 32, 64 read 76 % of the lines that flat rounds of 64 read, in as many
// rounds; starting at 8 reads 44 % but takes 20 % more rounds, and measured
// slower: the rounds' latency, not the lines, bounds the kernels once the
// pair index decides the searches between heavy rows).
constexpr int kFirstRound = 32;
struct PairIndex {
  const uint64_t* doff;
  const uint32_t* pos;
  uint64_t H;
};

constexpr uint64_t kNoSlot = ~0ull;

// The slot a->b of a symmetric graph with reverse positions (rows hold no
// duplicate slots), or kNoSlot: the shorter of row(a) and row(b) is
// searched, a hit in row(b) being mapped to the slot a->b through rev; two
// heavy rows are decided by the pair index (its position is the slot the
// search finds).
__device__ __forceinline__ uint64_t slot_rev(uint32_t a, uint32_t b, const uint64_t* __restrict__ off,
                                             const uint32_t* __restrict__ edges, const uint32_t* __restrict__ rev,
                                             const PairIndex& pi, uint64_t& probes) {
  uint64_t ab, bb, da, db;
  if (pi.pos) {
    const uint64_t xa = pi.doff[a], xb = pi.doff[b];
    ab = xa & kPairOffMask;
    bb = xb & kPairOffMask;
    const uint64_t ha = xa >> kPairShift, hb = xb >> kPairShift;
    if (ha && hb) {
      ++probes;
      const uint32_t r = pi.pos[(ha - 1) * pi.H + (hb - 1)];
      return r == kNoPos ? kNoSlot : ab + r;
    }
    da = (pi.doff[a + 1] & kPairOffMask) - ab;
    db = (pi.doff[b + 1] & kPairOffMask) - bb;
  } else {
    ab = off[a];
    bb = off[b];
    da = off[a + 1] - ab;
    db = off[b + 1] - bb;
  }
  const bool in_a = da <= db;
  uint64_t lo = in_a ? ab : bb, hi = lo + (in_a ? da : db);
  const uint64_t end = hi;
  const uint32_t key = in_a ? b : a;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    ++probes;
    if (edges[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo == end || edges[lo] != key) return kNoSlot;
  return in_a ? lo : ab + rev[lo];
}

// Whether row(a) holds an active slot to b (slot_rev, then its edge-active bit).
__device__ __forceinline__ bool active_slot_rev(uint32_t a, uint32_t b, const uint64_t* __restrict__ off,
                                                const uint32_t* __restrict__ edges, const uint32_t* __restrict__ ea,
                                                const uint32_t* __restrict__ rev, const PairIndex& pi,
                                                uint64_t& probes) {
  const uint64_t e = slot_rev(a, b, off, edges, rev, pi, probes);
  return e != kNoSlot && ea_test(ea, e);
}

// Pair index positions: one wave per heavy row; each slot to a heavy
// neighbour records its offset in the row (the smallest, which is where the
// binary search's lower bound lands, should a row repeat a neighbour).
__global__ void __launch_bounds__(kBlock) k_pair_positions(const uint32_t* __restrict__ heavy, uint64_t H,
                                                          const uint64_t* __restrict__ doff,
                                                          const uint32_t* __restrict__ edges,
                                                          uint32_t* __restrict__ pos) {
  const int lane = lane_id();
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / 64);
  for (uint64_t h = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6); h < H; h += waves) {
    const uint32_t a = heavy[h];
    const uint64_t beg = doff[a] & kPairOffMask, end = doff[a + 1] & kPairOffMask;
    uint32_t* row = pos + h * H;
    for (uint64_t e = beg + lane; e < end; e += 64) {
      const uint64_t hb = doff[edges[e]] >> kPairShift;
      if (hb) atomicMin(&row[hb - 1], static_cast<uint32_t>(e - beg));
    }
  }
}

// Last two hops of a cycle on sorted rows: is there a u with an active slot
// v->u, template vertex q among u's candidates, and an active slot u->s?
// The shorter of row(v) and row(s) is read 64 slots at a time, each lane
// binary-searching the other side, so a low-degree source next to a hub
// never reads the hub's row.  With reverse positions (symmetric graphs) the
// slot u->s of a u in row(s) is read directly and the remaining lookups
// search the shorter of the two rows.
template <typename Word>
__device__ bool close_cycle(uint32_t v, uint32_t s, int q, const uint64_t* __restrict__ off,
                            const uint32_t* __restrict__ edges, const uint32_t* __restrict__ ea,
                            const uint32_t* __restrict__ rev, const PairIndex& pi, const Word* __restrict__ cand,
                            uint64_t& slots) {
  const int lane = lane_id();
  const uint64_t vb = off[v], ve = off[v + 1], sb = off[s], se = off[s + 1];
  const bool from_v = (ve - vb) <= (se - sb);
  const uint64_t b = from_v ? vb : sb, end = from_v ? ve : se;
  uint64_t probes = 0, read = 0;
  bool found = false;
  for (uint64_t base = b; base < end; base += 64) {
    const uint64_t e = base + lane;
    bool ok = false;
    if (e < end) {
      const uint32_t u = edges[e];
      if ((cand[u] >> q) & 1) {
        if (rev) {
          if (from_v)
            ok = ea_test(ea, e) && active_slot_rev(u, s, off, edges, ea, rev, pi, probes);
          else
            ok = ea_test(ea, off[u] + rev[e]) && active_slot_rev(v, u, off, edges, ea, rev, pi, probes);
        } else if (from_v) {
          ok = ea_test(ea, e) && active_slot(u, s, off, edges, ea, probes);
        } else {
          ok = active_slot(v, u, off, edges, ea, probes) && active_slot(u, s, off, edges, ea, probes);
        }
      }
    }
    read += min<uint64_t>(64, end - base);
    if (__ballot(ok)) {
      found = true;
      break;
    }
  }
  for (int o = 32; o > 0; o >>= 1) probes += __shfl_xor(probes, o, 64);
  slots += read + probes;
  return found;
}

// Triangle existence for a source of at most 64 slots (complete search of a
// 3-vertex cycle, sorted rows without duplicates, symmetric graph with
// reverse positions).  The walk s -> a -> u -> s exists iff some a in
// A = {a : slot s->a active, q1 in C(a)} and some u in
// U = {u in row(s) : q2 in C(u), slot u->s active} have an active slot a->u
// (u reaches s over some slot, so u is in row(s)).  Lane i examines slot i
// of row(s) for both roles (the slot u->s is off[u] + rev[i]); the |A| x |U|
// pairs are then checked 64 at a time, each lane searching the shorter of
// row(a) and row(u).  Existence does not depend on the order, so any pair
// may decide.
template <typename Word>
__device__ bool triangle_small(uint32_t s, int q1, int q2, const uint64_t* __restrict__ off,
                               const uint32_t* __restrict__ edges, const uint32_t* __restrict__ ea,
                               const uint32_t* __restrict__ rev, const PairIndex& pi, const Word* __restrict__ cand,
                               uint64_t& slots) {
  const int lane = lane_id();
  const uint64_t beg = off[s];
  const int deg = static_cast<int>(off[s + 1] - beg);
  uint64_t probes = 0;
  bool is_a = false, is_u = false;
  uint32_t w = 0;
  Word cw = 0;
  const uint64_t e = beg + lane;
  if (lane < deg) {
    w = edges[e];
    cw = cand[w];
    is_a = ea_test(ea, e) && ((cw >> q1) & 1);
  }
  const uint64_t ma = __ballot(is_a);
  if (ma && lane < deg && ((cw >> q2) & 1)) is_u = ea_test(ea, off[w] + rev[e]);
  const uint64_t mu = __ballot(is_u);
  bool found = false;
  if (ma && mu) {
    // lane r learns the slot of the r-th a and of the r-th u
    int a_from = 0, u_from = 0, r = 0;
    for (uint64_t m = ma; m; m &= m - 1, ++r)
      if (lane == r) a_from = __builtin_ctzll(m);
    r = 0;
    for (uint64_t m = mu; m; m &= m - 1, ++r)
      if (lane == r) u_from = __builtin_ctzll(m);
    const uint32_t a_list = __shfl(w, a_from, 64), u_list = __shfl(w, u_from, 64);
    const int na = __popcll(ma), nu = __popcll(mu);
    const int pairs = na * nu;
    for (int p0 = 0, sz = kFirstRound; p0 < pairs && !found; p0 += sz, sz = 2 * sz < 64 ? 2 * sz : 64) {
      const int p = p0 + lane;
      const int ia = p / nu, iu = p - ia * nu;
      const uint32_t a = __shfl(a_list, ia & 63, 64), u = __shfl(u_list, iu & 63, 64);
      const bool ok = lane < sz && p < pairs && active_slot_rev(a, u, off, edges, ea, rev, pi, probes);
      found = __ballot(ok) != 0;
    }
  }
  for (int o = 32; o > 0; o >>= 1) probes += __shfl_xor(probes, o, 64);
  slots += static_cast<uint64_t>(deg) + probes;
  return found;
}

// Cycle constraint idx[0..L-1] from start position j (cc_cpu.tpp:12-70).
// With sorted rows the last two hops are decided by close_cycle.
template <typename Word>
__device__ bool cycle_walk(uint32_t s, uint32_t j, uint32_t L, bool sorted, const uint64_t* __restrict__ off,
                           const uint32_t* __restrict__ edges, const uint32_t* __restrict__ ea,
                           const uint32_t* __restrict__ rev, const PairIndex& pi, const Word* __restrict__ cand,
                           const DevTemplate* __restrict__ t, MemoTable* fm, uint64_t& slots) {
  const int lane = lane_id();
  uint32_t st_v = s;          // lane i: vertex at level i
  uint64_t st_c = off[s];     // lane i: resume slot at level i
  int depth = 0;
  uint32_t v = s;
  uint64_t cursor = off[s];
  if (fm) {  // a fresh failed-state memo for this walk
    fm->epoch = static_cast<uint32_t>(uniform(static_cast<int>(fm->epoch + 1)));
    fm->used = 0;
  }
  while (true) {
    if (sorted && depth == static_cast<int>(L) - 2) {
      const int q = static_cast<int>(t->nlcc_vertex[(j + L - 1) % L]);
      if (close_cycle<Word>(v, s, q, off, edges, ea, rev, pi, cand, slots)) return true;
      if (fm && depth >= 2) state_fail(*fm, v, depth);
      depth = uniform(depth - 1);
      v = rl32(st_v, depth);
      cursor = rl64(st_c, depth);
      continue;
    }
    if (depth == static_cast<int>(L) - 1) {
      RowScan r = scan_row<Word>(off[v], off[v + 1], edges, ea, cand, -1, s, 0xFFFFFFFFu, slots);
      if (r.found) return true;
      depth = uniform(depth - 1);
      v = rl32(st_v, depth);
      cursor = rl64(st_c, depth);
      continue;
    }
    const int q = static_cast<int>(t->nlcc_vertex[(j + depth + 1) % L]);
    RowScan r = scan_row<Word>(cursor, off[v + 1], edges, ea, cand, q, 0, 0xFFFFFFFFu, slots);
    if (r.found) {
      if (fm && depth + 1 >= 2 && depth + 1 <= static_cast<int>(L) - 2 && state_failed(*fm, r.u, depth + 1)) {
        cursor = r.next;  // known dead end: next neighbour
        continue;
      }
      if (lane == depth) st_c = r.next;
      depth = uniform(depth + 1);
      v = r.u;
      cursor = off[v];
      if (lane == depth) {
        st_v = v;
        st_c = cursor;
      }
    } else {
      if (depth == 0) return false;
      if (fm && depth >= 2) state_fail(*fm, v, depth);
      depth = uniform(depth - 1);
      v = rl32(st_v, depth);
      cursor = rl64(st_c, depth);
    }
  }
}

// Path constraint (pc_cpu.tpp:11-75): from position 0 forwards or from the
// last position backwards over L-1 edges, ending on a vertex other than s.
template <typename Word>
__device__ bool path_walk(uint32_t s, bool reverse, uint32_t L, const uint64_t* __restrict__ off,
                          const uint32_t* __restrict__ edges, const uint32_t* __restrict__ ea,
                          const Word* __restrict__ cand, const DevTemplate* __restrict__ t, MemoTable* fm,
                          uint64_t& slots) {
  const int lane = lane_id();
  uint32_t st_v = s;
  uint64_t st_c = off[s];
  int depth = 0;
  uint32_t v = s;
  uint64_t cursor = off[s];
  if (fm) {
    fm->epoch = static_cast<uint32_t>(uniform(static_cast<int>(fm->epoch + 1)));
    fm->used = 0;
  }
  while (true) {
    const uint32_t pos = reverse ? (L - 2 - depth) : (depth + 1);
    const int q = static_cast<int>(t->nlcc_vertex[pos]);
    const bool last = depth == static_cast<int>(L) - 2;
    RowScan r = scan_row<Word>(cursor, off[v + 1], edges, ea, cand, q, 0, last ? s : 0xFFFFFFFFu, slots);
    if (r.found) {
      if (last) return true;
      if (fm && depth + 1 >= 2 && state_failed(*fm, r.u, depth + 1)) {
        cursor = r.next;
        continue;
      }
      if (lane == depth) st_c = r.next;
      depth = uniform(depth + 1);
      v = r.u;
      cursor = off[v];
      if (lane == depth) {
        st_v = v;
        st_c = cursor;
      }
    } else {
      if (depth == 0) return false;
      if (fm && depth >= 2) state_fail(*fm, v, depth);
      depth = uniform(depth - 1);
      v = rl32(st_v, depth);
      cursor = rl64(st_c, depth);
    }
  }
}

// Walk over VERTEX / CHECK / MOVE_BACK moves (tds_cpu.tpp:13-127 and, with
// counting and injectivity, enumeration_cpu.tpp:38-118).  Lane p keeps the
// vertex the walk stands on before move p (st_v) and the resume slot of move
// p (st_c); lane h keeps history entry h.  Walks have at most 64 moves.
// Returns the number of completions when `count_all`, else 1 on the first
// completion.  When a search without counting backtracks through a CHECK,
// the rest of that row is added to `slots`, as a serial scan would read it
// looking for a second slot to the same vertex.
template <typename Word>
__device__ uint64_t template_walk(uint32_t s, uint32_t len, const uint8_t* __restrict__ mv,
                                  const uint32_t* __restrict__ vx, const uint32_t* __restrict__ mb,
                                  const uint32_t* __restrict__ hidx, bool count_all,
                                  const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges,
                                  const uint32_t* __restrict__ ea, const Word* __restrict__ cand, uint64_t& slots) {
  const int lane = lane_id();
  uint32_t st_v = s;
  uint64_t st_c = 0;
  uint32_t hist = s;
  uint64_t total = 0;
  int pos = 1;
  uint32_t cur = s;
  bool resume = false;  // re-entering a VERTEX move after backtracking
  if (len <= 1) return 1;
  auto is_vertex = [&](int p) { return mv[p] == PJ_MOVE_VERTEX_STORE || mv[p] == PJ_MOVE_VERTEX_NO_STORE; };
  // Back to the most recent VERTEX move; false when none is left.
  auto backtrack = [&]() -> bool {
    pos = uniform(pos - 1);
    while (pos >= 1 && !is_vertex(pos)) {
      if (!count_all && mv[pos] == PJ_MOVE_CHECK) {
        const uint32_t at = rl32(st_v, pos);
        slots += off[at + 1] - rl64(st_c, pos);
      }
      pos = uniform(pos - 1);
    }
    if (pos < 1) return false;
    cur = rl32(st_v, pos);
    resume = true;
    return true;
  };
  while (true) {
    if (pos == static_cast<int>(len)) {
      ++total;
      if (!count_all) return total;
      if (!backtrack()) return total;
      continue;
    }
    const uint8_t m = mv[pos];
    if (m == PJ_MOVE_VERTEX_STORE || m == PJ_MOVE_VERTEX_NO_STORE) {
      const uint64_t from = resume ? rl64(st_c, pos) : off[cur];
      resume = false;
      const int q = static_cast<int>(vx[pos]);
      const uint32_t h = hidx[pos];
      const uint64_t end = off[cur + 1];
      RowScan r{0, end, false};
      for (uint64_t base = from; base < end; base += 64) {
        const uint64_t e = base + lane;
        bool ok = false;
        uint32_t u = 0;
        if (e < end && ea_test(ea, e)) {
          u = edges[e];
          ok = (cand[u] >> q) & 1;
        }
        if (count_all)  // injective embeddings; every lane active here
          for (uint32_t k = 0; k < h; ++k) ok = ok && rl32(hist, static_cast<int>(k)) != u;
        uint64_t bm = __ballot(ok);
        if (bm) {
          int first = __builtin_ctzll(bm);
          r.u = rl32(u, first);
          r.next = base + first + 1;
          r.found = true;
          break;
        }
      }
      slots += (r.found ? r.next : end) - from;
      if (r.found) {
        if (lane == pos) st_c = r.next;
        if (lane == static_cast<int>(h)) hist = r.u;
        cur = r.u;
        pos = uniform(pos + 1);
        if (lane == pos) st_v = cur;
        continue;
      }
    } else if (m == PJ_MOVE_CHECK) {
      const uint32_t target = rl32(hist, static_cast<int>(mb[pos]));
      RowScan r = scan_row<Word>(off[cur], off[cur + 1], edges, ea, cand, -1, target, 0xFFFFFFFFu, slots);
      if (r.found) {
        if (lane == pos) st_c = r.next;
        pos = uniform(pos + 1);
        if (lane == pos) st_v = cur;
        continue;
      }
    } else {  // MOVE_BACK
      cur = rl32(hist, static_cast<int>(mb[pos]));
      pos = uniform(pos + 1);
      if (lane == pos) st_v = cur;
      continue;
    }
    if (!backtrack()) return total;
  }
}

// One memoised search (a plan of moves, as template_walk) from source s.
// Move p: VERTEX moves test walk_test[p] and insert mv vertex vx[p] in the
// edge memo, and with `hyp` consult / fill the hypothesis memo; kMoveClose
// and kMoveLast end the walk.  On success every move with mk[p] marks the
// vertex it reached in alreadyMatched with vx[p].  Returns 1 (found), 0 (no
// walk) or -1 (the wave's table overflowed; the caller retries the source
// with a larger table).
template <typename Word>
__device__ int memo_walk(uint32_t s, uint32_t len, const uint8_t* __restrict__ mv, const uint32_t* __restrict__ vx,
                         const uint8_t* __restrict__ tb, const uint8_t* __restrict__ mk,
                         const uint32_t* __restrict__ mb, const uint32_t* __restrict__ hidx, bool hyp,
                         bool sorted, const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges,
                         const uint32_t* __restrict__ ea, const Word* __restrict__ cand, Word* already,
                         MemoTable& m, uint64_t& slots) {
  const int lane = lane_id();
  uint32_t st_v = s;   // lane p: vertex before move p
  uint64_t st_c = 0;   // lane p: resume slot of move p
  uint32_t st_n = 0;   // lane p: hypothesis node before move p
  uint32_t hist = s;   // lane h: history entry h
  int pos = 1;
  uint32_t cur = s, node = 0;
  bool resume = false;
  if (len <= 1) return 1;
  // the plan in registers (lane p: move p), read by readlane at each step
  const bool in_plan = lane < static_cast<int>(len);
  const uint32_t r_mv = in_plan ? mv[lane] : 0u, r_vx = in_plan ? vx[lane] : 0u, r_tb = in_plan ? tb[lane] : 0u;
  const uint32_t r_mb = in_plan ? mb[lane] : 0u, r_hidx = in_plan ? hidx[lane] : 0u;
  // row bounds of cur when known (loaded with the step that moved to cur)
  bool have_bounds = false;
  uint64_t cur_beg = 0, cur_end = 0;
  auto searching = [&](int p) {
    const uint32_t mo = rl32(r_mv, p);
    return mo == PJ_MOVE_VERTEX_STORE || mo == PJ_MOVE_VERTEX_NO_STORE;
  };
  auto backtrack = [&]() -> bool {
    pos = uniform(pos - 1);
    while (pos >= 1 && !searching(pos)) pos = uniform(pos - 1);
    if (pos < 1) return false;
    cur = rl32(st_v, pos);
    node = rl32(st_n, pos);
    resume = true;
    have_bounds = false;
    return true;
  };
  while (true) {
    if (pos == static_cast<int>(len)) {
      const uint32_t after = __shfl(st_v, (lane + 1) & 63, 64);
      if (lane >= 1 && lane < static_cast<int>(len) && mk[lane])
        atomic_or_word<Word>(&already[lane + 1 < static_cast<int>(len) ? after : cur], Word(1) << vx[lane]);
      return 1;
    }
    const uint32_t mo = rl32(r_mv, pos);
    if (mo == PJ_MOVE_VERTEX_STORE || mo == PJ_MOVE_VERTEX_NO_STORE) {
      const uint64_t from = resume ? rl64(st_c, pos) : have_bounds ? cur_beg : off[cur];
      resume = false;
      const uint32_t y = rl32(r_vx, pos), xt = rl32(r_tb, pos);
      const uint64_t end = have_bounds ? cur_end : off[cur + 1];
      bool found = false;
      uint32_t u = 0;
      uint64_t base = from, next = end;
      while (base < end) {
        const uint64_t e = base + lane;
        bool ok = false;
        uint32_t w = 0;
        if (e < end && walk_slot(ea, e)) {
          w = edges[e];
          ok = (cand[w] >> y) & 1;
        }
        // every candidate of the 64 slots is tested against the memos at
        // once (each lane probes on its own); the walk takes the first one in
        // row order that no memo blocks, as testing them one by one does
        const bool blocked = ok && memo_blocked_lane(m, edge_key(cur, w), xt, hyp, hyp_key(w, y, node));
        const uint64_t bm = __ballot(ok && !blocked);
        if (!bm) {
          base += 64;
          continue;
        }
        const int first = __builtin_ctzll(bm);
        u = rl32(w, first);
        next = base + first + 1;
        found = true;
        break;
      }
      slots += (found ? next : end) - from;
      if (found) {
        // u's row bounds, for the step after this one: their loads overlap
        // the memo inserts
        const uint64_t u_beg = off[u], u_end = off[u + 1];
        const bool store = hyp && mo == PJ_MOVE_VERTEX_STORE;
        uint32_t node2 = node;
        if (!memo_step_inserts(m, store, hyp, node, cur, u, y, &node2)) {  // two new keys, one free slot
          node2 = store ? memo_trie_child(m, node, u) : node;
          if (hyp) memo_hyp_add(m, u, y, node2);
          memo_edge_add(m, cur, u, y);
        }
        if (m.overflow) return -1;
        if (lane == pos) st_c = next;
        if (lane == static_cast<int>(rl32(r_hidx, pos))) hist = u;
        cur = u;
        cur_beg = u_beg;
        cur_end = u_end;
        have_bounds = true;
        node = node2;
        pos = uniform(pos + 1);
        if (lane == pos) {
          st_v = cur;
          st_n = node;
        }
        continue;
      }
    } else if (mo == PJ_MOVE_CHECK || mo == kMoveClose) {
      // an existence test (the walk goes on from cur either way): on sorted
      // rows a binary search, else a scan in row order
      const uint32_t target = mo == kMoveClose ? s : rl32(hist, static_cast<int>(rl32(r_mb, pos)));
      bool hit;
      if (sorted) {
        uint64_t probes = 0;
        hit = active_slot(cur, target, off, edges, ea, probes);
        slots += probes;
      } else {
        hit = scan_row<Word>(off[cur], off[cur + 1], edges, ea, cand, -1, target, 0xFFFFFFFFu, slots).found;
      }
      if (hit) {
        pos = uniform(pos + 1);
        if (lane == pos) {
          st_v = cur;
          st_n = node;
        }
        continue;
      }
    } else if (mo == kMoveLast) {
      RowScan r = scan_row<Word>(off[cur], off[cur + 1], edges, ea, cand, static_cast<int>(rl32(r_vx, pos)), 0, s,
                                 slots);
      if (r.found) {
        cur = r.u;
        have_bounds = false;
        pos = uniform(pos + 1);
        continue;
      }
    } else {  // MOVE_BACK
      cur = rl32(hist, static_cast<int>(rl32(r_mb, pos)));
      have_bounds = false;
      pos = uniform(pos + 1);
      if (lane == pos) {
        st_v = cur;
        st_n = node;
      }
      continue;
    }
    if (!backtrack()) return 0;
  }
}

// ---------------------------------------------------------------------------
// Active view for the memoised walks: per active vertex, in row order, the
// slots with the edge-active bit set whose neighbour is active.  Those walks
// only follow such slots and only check slots to vertices they have reached
// (active), so on the view they make the same moves in the same order while
// skipping the dead slots of rows they rescan (at large scales the active
// vertices are a few per cent of the graph, and hub rows are rescanned at
// every backtrack).  The memo keys are vertex pairs, unaffected by the copy.

template <typename Word, int G>
__global__ void __launch_bounds__(kBlock) k_view_count(
    const uint32_t* __restrict__ list, const uint32_t* __restrict__ list_count, const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ edges, const uint32_t* __restrict__ ea, const Word* __restrict__ cand,
    uint64_t* __restrict__ counts) {
  const int lane = lane_id();
  const int gl = lane & (G - 1);
  const uint64_t gmask = G == 64 ? ~0ull : (((1ull << G) - 1) << (lane & ~(G - 1)));
  const uint32_t n = *list_count;
  const uint64_t groups_per_grid = (uint64_t)gridDim.x * (blockDim.x / G);
  const uint64_t span = ((n + groups_per_grid - 1) / groups_per_grid) * groups_per_grid;
  for (uint64_t g = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / G; g < span; g += groups_per_grid) {
    if (g >= n) continue;  // group-uniform
    const uint32_t v = list[g];
    const uint64_t beg = off[v], end = off[v + 1];
    uint64_t kept = 0;
    for (uint64_t base = beg; base < end; base += G) {
      const uint64_t e = base + gl;
      const bool keep = e < end && ea_test(ea, e) && cand[edges[e]] != 0;
      kept += __popcll(__ballot(keep) & gmask);
    }
    if (gl == 0) counts[v] = kept;
  }
}

template <typename Word, int G>
__global__ void __launch_bounds__(kBlock) k_view_fill(
    const uint32_t* __restrict__ list, const uint32_t* __restrict__ list_count, const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ edges, const uint32_t* __restrict__ ea, const Word* __restrict__ cand,
    const uint64_t* __restrict__ voff, uint32_t* __restrict__ vedges) {
  const int lane = lane_id();
  const int gl = lane & (G - 1);
  const uint64_t gmask = G == 64 ? ~0ull : (((1ull << G) - 1) << (lane & ~(G - 1)));
  const uint64_t below = ((1ull << lane) - 1) & gmask;
  const uint32_t n = *list_count;
  const uint64_t groups_per_grid = (uint64_t)gridDim.x * (blockDim.x / G);
  const uint64_t span = ((n + groups_per_grid - 1) / groups_per_grid) * groups_per_grid;
  for (uint64_t g = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / G; g < span; g += groups_per_grid) {
    if (g >= n) continue;
    const uint32_t v = list[g];
    const uint64_t beg = off[v], end = off[v + 1];
    uint64_t out = voff[v];
    for (uint64_t base = beg; base < end; base += G) {
      const uint64_t e = base + gl;
      uint32_t u = 0;
      bool keep = false;
      if (e < end && ea_test(ea, e)) {
        u = edges[e];
        keep = cand[u] != 0;
      }
      const uint64_t m = __ballot(keep) & gmask;
      if (keep) vedges[out + __popcll(m & below)] = u;
      out += __popcll(m);
    }
  }
}

// Hub rows: one wave per hub walks the row in order (count, then fill).
template <typename Word>
__global__ void __launch_bounds__(kBlock) k_view_hubs(const uint32_t* __restrict__ hubs, uint32_t nhubs,
                                                      const uint64_t* __restrict__ off,
                                                      const uint32_t* __restrict__ edges,
                                                      const uint32_t* __restrict__ ea, const Word* __restrict__ cand,
                                                      uint64_t* __restrict__ counts, const uint64_t* __restrict__ voff,
                                                      uint32_t* __restrict__ vedges) {
  const int lane = lane_id();
  const uint64_t below = (1ull << lane) - 1;
  const uint32_t waves = gridDim.x * (blockDim.x / 64);
  for (uint32_t h = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); h < nhubs; h += waves) {
    const uint32_t v = hubs[h];
    if (!cand[v]) continue;  // wave-uniform
    const uint64_t beg = off[v], end = off[v + 1];
    uint64_t out = voff ? voff[v] : 0;
    for (uint64_t base = beg; base < end; base += 64) {
      const uint64_t e = base + lane;
      uint32_t u = 0;
      bool keep = false;
      if (e < end && ea_test(ea, e)) {
        u = edges[e];
        keep = cand[u] != 0;
      }
      const uint64_t m = __ballot(keep);
      if (vedges && keep) vedges[out + __popcll(m & below)] = u;
      out += __popcll(m);
    }
    if (!voff && lane == 0) counts[v] = out;
  }
}

// Exclusive prefix sum of the view's row counts, in three launches: per
// span of kScanSpan counts a block total (k_scan_spans), the exclusive scan
// of those totals in one block (k_scan_totals), then each span rescanned from
// LDS with its total's offset (k_scan_apply).  Each count is read twice and
// each offset written once.
constexpr int kScanPer = 16;
constexpr uint64_t kScanSpan = static_cast<uint64_t>(kBlock) * kScanPer;
constexpr int kTotalsBlock = 1024;

__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, int d) {
  const uint32_t lo = __shfl_up(static_cast<uint32_t>(v), d);
  const uint32_t hi = __shfl_up(static_cast<uint32_t>(v >> 32), d);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Exclusive prefix of x over the block's threads (blockDim.x a multiple of
// 64, at most kTotalsBlock); *total receives the block's sum.  Every thread
// of the block must call it.
__device__ __forceinline__ uint64_t block_exclusive(uint64_t x, uint64_t* total) {
  __shared__ uint64_t s_wave[kTotalsBlock / 64];
  const int lane = lane_id(), w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  uint64_t inc = x;
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t y = shfl_up64(inc, d);
    if (lane >= d) inc += y;
  }
  if (lane == 63) s_wave[w] = inc;
  __syncthreads();
  uint64_t before = 0, all = 0;
  for (int i = 0; i < nw; ++i) {
    const uint64_t t = s_wave[i];
    before += i < w ? t : 0;
    all += t;
  }
  __syncthreads();
  *total = all;
  return before + inc - x;
}

__global__ void __launch_bounds__(kBlock) k_scan_spans(const uint64_t* __restrict__ in, uint64_t n,
                                                      uint64_t* __restrict__ totals) {
  const uint64_t a = blockIdx.x * kScanSpan;
  const uint64_t z = a + kScanSpan < n ? a + kScanSpan : n;
  uint64_t sum = 0;
  for (uint64_t i = a + threadIdx.x; i < z; i += kBlock) sum += in[i];
  uint64_t all;
  block_exclusive(sum, &all);
  if (threadIdx.x == 0) totals[blockIdx.x] = all;
}

__global__ void __launch_bounds__(kTotalsBlock) k_scan_totals(uint64_t* __restrict__ totals, uint64_t n) {
  const uint64_t seg = (n + kTotalsBlock - 1) / kTotalsBlock;
  const uint64_t a = threadIdx.x * seg < n ? threadIdx.x * seg : n;
  const uint64_t z = a + seg < n ? a + seg : n;
  uint64_t sum = 0;
  for (uint64_t i = a; i < z; ++i) sum += totals[i];
  uint64_t all;
  uint64_t run = block_exclusive(sum, &all);
  for (uint64_t i = a; i < z; ++i) {
    const uint64_t x = totals[i];
    totals[i] = run;
    run += x;
  }
}

// One span: counts into LDS (coalesced), kScanPer consecutive counts per
// thread scanned in registers, offsets back through LDS.  The LDS index is
// padded by one word per kScanPer so that a wave's threads, each at the
// start of its own run, fall on distinct banks.
__global__ void __launch_bounds__(kBlock) k_scan_apply(const uint64_t* __restrict__ in, uint64_t n,
                                                      const uint64_t* __restrict__ totals,
                                                      uint64_t* __restrict__ out) {
  __shared__ uint64_t s[kScanSpan + kScanSpan / kScanPer];
  const uint64_t a = blockIdx.x * kScanSpan;
  const uint32_t len = static_cast<uint32_t>(a + kScanSpan < n ? kScanSpan : n - a);
  for (uint32_t i = threadIdx.x; i < kScanSpan; i += kBlock) s[i + i / kScanPer] = i < len ? in[a + i] : 0;
  __syncthreads();
  uint64_t* mine = s + threadIdx.x * (kScanPer + 1);
  uint64_t v[kScanPer];
  uint64_t sum = 0;
#pragma unroll
  for (int j = 0; j < kScanPer; ++j) {
    v[j] = mine[j];
    sum += v[j];
  }
  uint64_t all;
  uint64_t run = totals[blockIdx.x] + block_exclusive(sum, &all);
#pragma unroll
  for (int j = 0; j < kScanPer; ++j) {
    mine[j] = run;
    run += v[j];
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < len; i += kBlock) out[a + i] = s[i + i / kScanPer];
}

// The memoised walks of one source: plans in walk order; a root already
// matched by another source's walk is taken as matched (cc_cpu.tpp:102-108,
// pc_cpu.tpp:104-142, tds_cpu.tpp:160-188).  Returns the failing roots;
// *overflow when the table overflowed (then nothing of s may be written).
template <typename Word>
__device__ Word memo_source(uint32_t s, Word cs, const DevTemplate* __restrict__ t, const uint64_t* __restrict__ off,
                            const uint32_t* __restrict__ edges, const uint32_t* __restrict__ ea,
                            const Word* __restrict__ cand, Word* already, MemoTable& m, uint64_t& slots,
                            bool* overflow) {
  Word fail = 0;
  *overflow = false;
  for (uint32_t k = 0; k < t->walk_count; ++k) {
    const uint32_t x = t->walk_root[k];
    if (!((cs >> x) & 1)) continue;
    const Word done = __hip_atomic_load(&already[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((done >> x) & 1) continue;
    const uint32_t o = t->walk_off[k];
    const uint32_t len = t->walk_off[k + 1] - o;
    m.epoch = static_cast<uint32_t>(uniform(static_cast<int>(m.epoch + 1)));
    m.used = 0;
    m.next_node = 0;
    m.overflow = false;
    const int r = memo_walk<Word>(s, len, t->walk_move + o, t->walk_vertex + o, t->walk_test + o, t->walk_mark + o,
                                  t->walk_mb + o, t->walk_hidx + o, t->walk_hyp != 0, t->rows_sorted != 0, off, edges,
                                  ea, cand, already, m, slots);
    if (r < 0) {
      *overflow = true;
      return 0;
    }
    if (r == 0) fail |= Word(1) << x;
  }
  return fail;
}

// First pass of the memoised walks: one wave per block with its memo table
// in LDS (kLdsMemo entries, 20 KB), so each probe is an LDS round trip.
// Sources whose walks outgrow it go to the overflow list and are rerun with
// tables in HBM (k_nlcc with MemoArgs, launch_nlcc_retry).
constexpr int kLdsMemo = 1024;

template <typename Word>
__global__ void __launch_bounds__(64) k_memo_lds(const uint32_t* __restrict__ list,
                                                 const uint32_t* __restrict__ list_count,
                                                 const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges,
                                                 const Word* __restrict__ cand, Word* __restrict__ unmatch,
                                                 uint8_t* __restrict__ flags, const DevTemplate* __restrict__ t,
                                                 MemoArgs ma, unsigned long long* counters) {
  __shared__ uint64_t s_keys[kLdsMemo];
  __shared__ uint64_t s_vals[kLdsMemo];
  __shared__ uint32_t s_epochs[kLdsMemo];
  const int lane = lane_id();
  for (int i = lane; i < kLdsMemo; i += 64) s_epochs[i] = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  MemoTable m{};
  m.keys = s_keys;
  m.vals = s_vals;
  m.epochs = s_epochs;
  m.mask = kLdsMemo - 1;
  m.limit = ma.lds_limit ? min(ma.lds_limit, 3u * kLdsMemo / 4) : 3u * kLdsMemo / 4;
  m.epoch = 0;
  Word* already = static_cast<Word*>(ma.already);
  const uint32_t n = *list_count;
  if (blockIdx.x == 0 && lane == 0) counters[C_ITEMS] = n;
  uint64_t slots = 0;
  for (uint32_t w = blockIdx.x; w < n; w += gridDim.x) {
    const uint32_t s = list[w];
    const Word cs = cand[s];
    bool overflow = false;
    const Word fail = memo_source<Word>(s, cs, t, off, edges, nullptr, cand, already, m, slots, &overflow);
    if (lane == 0) {
      if (overflow) {
        ma.ovf_list[atomicAdd(ma.ovf_count, 1u)] = s;
      } else if (fail) {  // s is in one task list once: this wave is its only writer
        unmatch[s] = static_cast<Word>(unmatch[s] | fail);
        set_flag_atomic(flags, s, kModified);
      }
    }
  }
  if (lane == 0 && slots) {
    const int slot = blockIdx.x & (kCounterSlots - 1);
    atomicAdd(&counters[slot * kCounterStride + C_SLOTS], static_cast<unsigned long long>(slots));
  }
}

// 3-vertex cycle constraint for sources of at most G slots, one group of G
// lanes per source (64 / G sources per wave, so more independent searches
// are in flight than with one wave per source).  Same test as triangle_small
// with group-local ballots: lane i examines slot i of row(s) for both roles,
// then the |A| x |U| pairs are checked G at a time.  Needs sorted rows
// without duplicates, a symmetric graph and reverse positions.
template <typename Word, int G>
__global__ void __launch_bounds__(kBlock) k_triangles(const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ list_count,
                                                      const uint64_t* __restrict__ off,
                                                      const uint32_t* __restrict__ edges,
                                                      const uint32_t* __restrict__ ea, const uint32_t* __restrict__ rev,
                                                      PairIndex pi, const Word* __restrict__ cand,
                                                      Word* __restrict__ unmatch, uint8_t* __restrict__ flags,
                                                      const DevTemplate* __restrict__ t,
                                                      unsigned long long* counters) {
  const int lane = lane_id();
  const int gl = lane & (G - 1);
  const int gbase = lane & ~(G - 1);
  const uint64_t gmask = G == 64 ? ~0ull : (((1ull << G) - 1) << gbase);
  const uint32_t n = *list_count;
  if (blockIdx.x == 0 && threadIdx.x == 0) counters[C_ITEMS] = n;
  const uint64_t groups_per_grid = (uint64_t)gridDim.x * (blockDim.x / G);
  const uint64_t span = ((n + groups_per_grid - 1) / groups_per_grid) * groups_per_grid;
  const uint32_t qv[3] = {t->nlcc_vertex[0], t->nlcc_vertex[1], t->nlcc_vertex[2]};
  uint64_t slots = 0;
  for (uint64_t g = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / G; g < span; g += groups_per_grid) {
    if (g >= n) continue;
    const uint32_t s = list[g];
    const Word cs = cand[s];
    const uint64_t beg = off[s];
    const int deg = static_cast<int>(off[s + 1] - beg);
    const bool mine = gl < deg;
    const uint64_t e = beg + gl;
    uint32_t w = 0;
    Word cw = 0;
    bool out_ok = false, back_ok = false;
    if (mine) {
      w = edges[e];
      cw = cand[w];
      out_ok = ea_test(ea, e);
      if (cw) back_ok = ea_test(ea, off[w] + rev[e]);  // slot w->s
    }
    uint64_t probes = 0;
    Word fail = 0;
    for (int j = 0; j < 3; ++j) {
      const uint32_t q = qv[j];
      if (!((cs >> q) & 1)) continue;
      const uint32_t q1 = qv[(j + 1) % 3], q2 = qv[(j + 2) % 3];
      const uint64_t ma = (__ballot(out_ok && ((cw >> q1) & 1)) & gmask) >> gbase;
      const uint64_t mu = (__ballot(back_ok && ((cw >> q2) & 1)) & gmask) >> gbase;
      bool found = false;
      if (ma && mu) {
        int a_from = 0, u_from = 0, r = 0;  // lane r of the group: slot of the r-th a / u
        for (uint64_t m = ma; m; m &= m - 1, ++r)
          if (gl == r) a_from = __builtin_ctzll(m);
        r = 0;
        for (uint64_t m = mu; m; m &= m - 1, ++r)
          if (gl == r) u_from = __builtin_ctzll(m);
        const uint32_t a_list = __shfl(w, gbase + a_from, 64), u_list = __shfl(w, gbase + u_from, 64);
        const int na = __popcll(ma), nu = __popcll(mu);
        const int pairs = na * nu;
        for (int p0 = 0, sz = kFirstRound < G ? kFirstRound : G; p0 < pairs && !found;
             p0 += sz, sz = 2 * sz < G ? 2 * sz : G) {
          const int p = p0 + gl;
          const int ia = p / nu, iu = p - ia * nu;
          const uint32_t a = __shfl(a_list, gbase + (ia & (G - 1)), 64);
          const uint32_t u = __shfl(u_list, gbase + (iu & (G - 1)), 64);
          const bool ok = gl < sz && p < pairs && active_slot_rev(a, u, off, edges, ea, rev, pi, probes);
          found = (__ballot(ok) & gmask) != 0;
        }
      }
      if (!found) fail |= Word(1) << q;
    }
    slots += probes + (gl == 0 ? static_cast<uint64_t>(deg) : 0);
    if (fail && gl == 0) {  // s is in one task list once: this group is its only writer
      unmatch[s] = static_cast<Word>(unmatch[s] | fail);
      set_flag_atomic(flags, s, kModified);
    }
  }
  wave_add(counters, C_SLOTS, slots);
}

// Triangle existence for a source of at most 64 * KC slots: the test of
// triangle_small over KC chunks of 64 slots of row(s).  Every chunk is
// classified at once (lane i keeps slot i of each chunk in registers, all
// loads of a level issued together), and each lane learns the slot of the
// r-th a and the r-th u of each chunk; the pairs of each (A chunk, U chunk)
// are then checked 64 at a time, A chunks in row order.  The pairs are
// packed densely (only a in A and u in U), so a source whose candidates fail
// costs |A| x |U| / 64 rounds of searches, not one round per slot of row(s)
// for every a as in the walk.
template <typename Word, int KC>
__device__ bool triangle_chunks(uint32_t s, int q1, int q2, const uint64_t* __restrict__ off,
                                const uint32_t* __restrict__ edges, const uint32_t* __restrict__ ea,
                                const uint32_t* __restrict__ rev, const PairIndex& pi, const Word* __restrict__ cand,
                                uint64_t& slots) {
  const int lane = lane_id();
  const uint64_t beg = off[s];
  const int deg = static_cast<int>(off[s + 1] - beg);
  uint32_t w[KC];
  Word cw[KC];
  uint32_t rv[KC];
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const int i = k * 64 + lane;
    w[k] = i < deg ? edges[beg + i] : 0u;
    rv[k] = i < deg ? rev[beg + i] : 0u;
  }
  uint64_t ma[KC], mu[KC];
  uint64_t wo[KC];
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const int i = k * 64 + lane;
    cw[k] = i < deg ? cand[w[k]] : Word(0);
    wo[k] = (i < deg && ((cw[k] >> q2) & 1)) ? off[w[k]] : 0ull;
  }
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const int i = k * 64 + lane;
    const bool in = i < deg;
    ma[k] = __ballot(in && ((cw[k] >> q1) & 1) && ea_test(ea, beg + i));
    mu[k] = __ballot(in && ((cw[k] >> q2) & 1) && ea_test(ea, wo[k] + rv[k]));  // slot u->s
  }
  uint64_t probes = 0;
  bool found = false;
  uint64_t any_a = 0, any_u = 0;
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    any_a |= ma[k];
    any_u |= mu[k];
  }
  if (any_a && any_u) {
    uint32_t a_list[KC], u_list[KC];
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      int a_from = 0, u_from = 0, r = 0;  // lane r: slot of the r-th a / u of chunk k
      for (uint64_t m = ma[k]; m; m &= m - 1, ++r)
        if (lane == r) a_from = __builtin_ctzll(m);
      r = 0;
      for (uint64_t m = mu[k]; m; m &= m - 1, ++r)
        if (lane == r) u_from = __builtin_ctzll(m);
      a_list[k] = __shfl(w[k], a_from, 64);
      u_list[k] = __shfl(w[k], u_from, 64);
    }
    int sz = kFirstRound;  // grows over the source's rounds, whichever chunk pair they test
#pragma unroll
    for (int ka = 0; ka < KC; ++ka) {
#pragma unroll
      for (int ku = 0; ku < KC; ++ku) {
        if (found || !ma[ka] || !mu[ku]) continue;
        const int na = __popcll(ma[ka]), nu = __popcll(mu[ku]);
        const int pairs = na * nu;
        for (int p0 = 0; p0 < pairs && !found; p0 += sz, sz = 2 * sz < 64 ? 2 * sz : 64) {
          const int p = p0 + lane;
          const int ia = p / nu, iu = p - ia * nu;
          const uint32_t a = __shfl(a_list[ka], ia & 63, 64), u = __shfl(u_list[ku], iu & 63, 64);
          const bool ok = lane < sz && p < pairs && active_slot_rev(a, u, off, edges, ea, rev, pi, probes);
          found = __ballot(ok) != 0;
        }
      }
    }
  }
  probes = __ockl_wfred_add_u64(probes);
  slots += static_cast<uint64_t>(deg) + probes;
  return found;
}

// 3-vertex cycle sources of 33 to 64 * kMidChunks slots (complete search;
// sorted rows without duplicates, symmetric graph, reverse positions): one
// wave per source, triangle_small / triangle_chunks per start position.
constexpr int kMidChunks = 4;
static_assert(kMidChunks * 64 == 256, "bin 2 holds rows of 33 to 256 slots");

template <typename Word>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8)))
k_triangles_mid(const uint32_t* __restrict__ list, const uint32_t* __restrict__ list_count,
                const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges,
                const uint32_t* __restrict__ ea, const uint32_t* __restrict__ rev, PairIndex pi,
                const Word* __restrict__ cand, Word* __restrict__ unmatch, uint8_t* __restrict__ flags,
                const DevTemplate* __restrict__ t, unsigned long long* counters) {
  const uint32_t n = *list_count;
  if (blockIdx.x == 0 && threadIdx.x == 0) counters[C_ITEMS] = n;
  const uint32_t waves = gridDim.x * (blockDim.x / 64);
  const uint32_t wave_id = uniform(blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6));
  const uint32_t qv[3] = {t->nlcc_vertex[0], t->nlcc_vertex[1], t->nlcc_vertex[2]};
  uint64_t slots = 0;
  for (uint32_t w = wave_id; w < n; w += waves) {
    const uint32_t s = list[w];
    const Word cs = cand[s];
    const uint64_t deg = off[s + 1] - off[s];
    Word fail = 0;
    for (int j = 0; j < 3; ++j) {
      const uint32_t q = qv[j];
      if (!((cs >> q) & 1)) continue;
      const int q1 = static_cast<int>(qv[(j + 1) % 3]), q2 = static_cast<int>(qv[(j + 2) % 3]);
      const bool ok = deg <= 64 ? triangle_small<Word>(s, q1, q2, off, edges, ea, rev, pi, cand, slots)
                                : triangle_chunks<Word, kMidChunks>(s, q1, q2, off, edges, ea, rev, pi, cand, slots);
      if (!ok) fail |= Word(1) << q;
    }
    if (fail && lane_id() == 0) {  // s is in one task list once: this wave is its only writer
      unmatch[s] = static_cast<Word>(unmatch[s] | fail);
      set_flag_atomic(flags, s, kModified);
    }
  }
  if (lane_id() == 0 && slots) {
    const int slot = (blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) & (kCounterSlots - 1);
    atomicAdd(&counters[slot * kCounterStride + C_SLOTS], static_cast<unsigned long long>(slots));
  }
}

// 3-vertex cycle sources of 257 to kWideSlots slots (complete search; sorted
// rows without duplicates, symmetric graph, reverse positions): one wave per
// source.  The first first hop a (row order) is closed as the walk closes it
// (close_cycle over the shorter of row(a) and row(s)); if that fails, the
// closing vertices U = {u in row(s) : q2 in C(u), slot u->s
// active} are compacted into the wave's LDS list; the later first hops are classified
// 64 slots at a time (A_k = {a : slot s->a active, q1 in C(a)}) and the
// |A_k| x |U| pairs of each chunk are tested 64 at a time by active_slot_rev.
// The walk it replaces closes the cycle of every first hop a over the
// shorter of row(a) and row(s): a source whose candidate fails read
// |A| x min(deg) slots in |A| x min(deg) / 64 rounds instead of
// |A| x |U| / 64.  Existence does not depend on the order of the pairs.
constexpr uint32_t kWideSlots = 2048;
static_assert(kWideSlots == kHubDegree, "bin 3 holds rows of 257 to kHubDegree slots");

template <typename Word>
__device__ bool triangle_wide(uint32_t s, int q1, int q2, uint64_t beg, uint32_t deg,
                              const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges,
                              const uint32_t* __restrict__ ea, const uint32_t* __restrict__ rev,
                              const PairIndex& pi, const Word* __restrict__ cand, uint32_t* U, uint64_t& slots) {
  const int lane = lane_id();
  const uint64_t below = (1ull << lane) - 1;
  // The first first hop is closed as the walk closes it: most sources that
  // succeed do so there, without the lists.
  const RowScan r1 = scan_row<Word>(beg, beg + deg, edges, ea, cand, q1, 0, 0xFFFFFFFFu, slots);
  if (!r1.found) return false;
  if (close_cycle<Word>(r1.u, s, q2, off, edges, ea, rev, pi, cand, slots)) return true;
  const uint32_t skip = static_cast<uint32_t>(r1.next - beg);  // first hops from here on
  uint64_t probes = 0;
  uint32_t nu = 0;
  for (uint32_t c0 = 0; c0 < deg; c0 += 64) {
    const uint32_t i = c0 + lane;
    bool is_u = false;
    uint32_t w = 0;
    if (i < deg) {
      w = edges[beg + i];
      if ((cand[w] >> q2) & 1) is_u = ea_test(ea, off[w] + rev[beg + i]);  // slot w->s
    }
    const uint64_t m = __ballot(is_u);
    if (is_u) U[nu + __popcll(m & below)] = w;
    nu += __popcll(m);
  }
  // the list's stores by every lane before any lane reads it
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  bool found = false;
  uint64_t read = deg;
  for (uint32_t c0 = skip & ~63u; nu && c0 < deg && !found; c0 += 64) {
    const uint32_t i = c0 + lane;
    bool is_a = false;
    uint32_t w = 0;
    if (i >= skip && i < deg) {
      w = edges[beg + i];
      is_a = ((cand[w] >> q1) & 1) && ea_test(ea, beg + i);
    }
    read += min(64u, deg - c0);
    const uint64_t ma = __ballot(is_a);
    if (!ma) continue;
    int a_from = 0, r = 0;  // lane r: slot of the r-th a of the chunk
    for (uint64_t m = ma; m; m &= m - 1, ++r)
      if (lane == r) a_from = __builtin_ctzll(m);
    const uint32_t a_list = __shfl(w, a_from, 64);
    const uint32_t pairs = static_cast<uint32_t>(__popcll(ma)) * nu;
    for (uint32_t p0 = 0; p0 < pairs && !found; p0 += 64) {
      const uint32_t p = p0 + lane;
      const uint32_t ia = p / nu, iu = p - ia * nu;
      const uint32_t a = __shfl(a_list, static_cast<int>(ia & 63), 64);
      const bool ok = p < pairs && active_slot_rev(a, U[iu], off, edges, ea, rev, pi, probes);
      found = __ballot(ok) != 0;
    }
  }
  // every lane's reads of the list before the wave's next search rewrites it
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  probes = __ockl_wfred_add_u64(probes);
  slots += read + probes;
  return found;
}

template <typename Word>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8)))
k_triangles_wide(const uint32_t* __restrict__ list, const uint32_t* __restrict__ list_count,
                 const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges,
                 const uint32_t* __restrict__ ea, const uint32_t* __restrict__ rev, PairIndex pi,
                 const Word* __restrict__ cand, Word* __restrict__ unmatch, uint8_t* __restrict__ flags,
                 const DevTemplate* __restrict__ t, unsigned long long* counters) {
  __shared__ uint32_t s_u[kBlock / 64][kWideSlots];
  uint32_t* U = s_u[threadIdx.x >> 6];
  const uint32_t n = *list_count;
  if (blockIdx.x == 0 && threadIdx.x == 0) counters[C_ITEMS] = n;
  const uint32_t waves = gridDim.x * (blockDim.x / 64);
  const uint32_t wave_id = uniform(blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6));
  const uint32_t qv[3] = {t->nlcc_vertex[0], t->nlcc_vertex[1], t->nlcc_vertex[2]};
  uint64_t slots = 0;
  for (uint32_t w = wave_id; w < n; w += waves) {
    const uint32_t s = list[w];
    const Word cs = cand[s];
    const uint64_t beg = off[s];
    const uint32_t deg = static_cast<uint32_t>(off[s + 1] - beg);  // at most kWideSlots (bin 3)
    Word fail = 0;
    for (int j = 0; j < 3; ++j) {
      const uint32_t q = qv[j];
      if (!((cs >> q) & 1)) continue;
      const int q1 = static_cast<int>(qv[(j + 1) % 3]), q2 = static_cast<int>(qv[(j + 2) % 3]);
      if (!triangle_wide<Word>(s, q1, q2, beg, deg, off, edges, ea, rev, pi, cand, U, slots)) fail |= Word(1) << q;
    }
    if (fail && lane_id() == 0) {  // s is in one task list once: this wave is its only writer
      unmatch[s] = static_cast<Word>(unmatch[s] | fail);
      set_flag_atomic(flags, s, kModified);
    }
  }
  if (lane_id() == 0 && slots) {
    const int slot = (blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) & (kCounterSlots - 1);
    atomicAdd(&counters[slot * kCounterStride + C_SLOTS], static_cast<unsigned long long>(slots));
  }
}

// Memoised 3-vertex cycle constraint (cc_cpu.tpp:12-70 with L = 3) on a
// graph without duplicate slots or self-loops (sorted rows, reverse
// positions).  For start j the reference walks s -> a -> u -> s with
// a in row(s) order (slot active, q1 in C(a)) and u in row(a) order (slot
// active, q2 in C(u)); its edge memo, cleared per (s, j), stores template ids
// and tests positions.  On such a graph the only test that can succeed is the
// one at the second hop, position (j+1) % 3, on the edge {a, u}, and it holds
// exactly when q2 equals that position and u was itself an earlier first hop
// whose second hops reached a (u in row(s) before a, slot s->u active, q1 in
// C(u); slot u->a active, q2 in C(a)): the first hop's test on {s, a} sees
// only ids inserted for {s, a}, which no earlier step inserts, and a second
// hop to u = s (the only other insertion on {a, s}) cannot close without a
// self-loop.  So the search takes the first (a, u) in that order that is not
// blocked and has an active slot u -> s, and marks a with q1 and u with q2 in
// alreadyMatched, as the walk does.  Rows of a are read 64 slots at a time
// and the first lane in row order decides.  One wave per source.
template <typename Word>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) k_triangles_memo(const uint32_t* __restrict__ list,
                                                           const uint32_t* __restrict__ list_count,
                                                           const uint64_t* __restrict__ off,
                                                           const uint32_t* __restrict__ edges,
                                                           const uint32_t* __restrict__ ea,
                                                           const uint32_t* __restrict__ rev, PairIndex pi,
                                                           const Word* __restrict__ cand, Word* __restrict__ unmatch,
                                                           uint8_t* __restrict__ flags, Word* already,
                                                           const DevTemplate* __restrict__ t,
                                                           unsigned long long* counters) {
  const int lane = lane_id();
  const uint32_t n = *list_count;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&counters[C_ITEMS], static_cast<unsigned long long>(n));
  const uint32_t waves = gridDim.x * (blockDim.x / 64);
  const uint32_t wave_id = uniform(blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6));
  uint64_t slots = 0, probes = 0;
  for (uint32_t wv = wave_id; wv < n; wv += waves) {
    const uint32_t s = list[wv];
    const Word cs = cand[s];
    const uint64_t sb = off[s], se = off[s + 1];
    Word fail = 0;
    for (uint32_t j = 0; j < 3; ++j) {
      const uint32_t q0 = t->nlcc_vertex[j];
      if (!((cs >> q0) & 1)) continue;
      const Word done = __hip_atomic_load(&already[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((done >> q0) & 1) continue;  // alreadyMatched (cc_cpu.tpp:102-108)
      const uint32_t q1 = t->nlcc_vertex[(j + 1) % 3], q2 = t->nlcc_vertex[(j + 2) % 3];
      const bool may_block = q2 == (j + 1) % 3;
      bool found = false;
      uint64_t cursor = sb;
      while (!found) {
        // next first hop a in row(s) order
        const RowScan r = scan_row<Word>(cursor, se, edges, ea, cand, static_cast<int>(q1), 0, 0xFFFFFFFFu, slots);
        if (!r.found) break;
        cursor = r.next;
        const uint32_t a = r.u;
        const uint64_t pa = r.next - 1 - sb;  // position of a in row(s)
        const bool block_a = may_block && ((cand[a] >> q2) & 1);
        const uint64_t ab = off[a], ae = off[a + 1];
        if (se - sb < ae - ab) {
          // Row(s) is the shorter: every u closing the cycle is in both rows,
          // and both are sorted, so walking row(s) meets the u in the order
          // of row(a); each u's slot a->u comes from slot_rev, its slot u->s
          // and its position in row(s) from the walk itself.
          for (uint64_t base = sb; base < se && !found; base += 64) {
            const uint64_t e = base + lane;  // slot s->u
            bool ok = false;
            uint32_t u = 0;
            if (e < se) {
              u = edges[e];
              const Word cu = cand[u];
              if (((cu >> q2) & 1) && ea_test(ea, off[u] + rev[e])) {  // q2 in C(u), slot u->s active
                const uint64_t au = slot_rev(a, u, off, edges, rev, pi, probes);
                ok = au != kNoSlot && ea_test(ea, au);
                // u an earlier first hop (before a in row(s), slot s->u active)
                // whose second hops reached a
                if (ok && block_a && ((cu >> q1) & 1) && e - sb < pa && ea_test(ea, e) &&
                    ea_test(ea, off[u] + rev[au]))
                  ok = false;
              }
            }
            const uint64_t m = __ballot(ok);
            slots += m ? __builtin_ctzll(m) + 1 : min<uint64_t>(64, se - base);
            if (m) {
              found = true;
              u = rl32(u, __builtin_ctzll(m));
              if (lane == 0) {
                atomic_or_word<Word>(&already[a], Word(1) << q1);
                atomic_or_word<Word>(&already[u], Word(1) << q2);
              }
            }
          }
          continue;
        }
        for (uint64_t base = ab; base < ae && !found; base += 64) {
          const uint64_t e = base + lane;
          bool ok = false;
          if (e < ae && ea_test(ea, e)) {
            const uint32_t u = edges[e];
            const Word cu = cand[u];
            ok = (cu >> q2) & 1;
            if (ok && block_a && ((cu >> q1) & 1) && ea_test(ea, off[u] + rev[e])) {
              // u an earlier first hop?  Its position in row(s), by binary search.
              uint64_t lo = sb, hi = se;
              while (lo < hi) {
                const uint64_t mid = (lo + hi) >> 1;
                ++probes;
                if (edges[mid] < u)
                  lo = mid + 1;
                else
                  hi = mid;
              }
              if (lo < se && edges[lo] == u && lo - sb < pa && ea_test(ea, lo)) ok = false;
            }
            if (ok) ok = active_slot_rev(u, s, off, edges, ea, rev, pi, probes);
          }
          const uint64_t m = __ballot(ok);
          slots += m ? __builtin_ctzll(m) + 1 : min<uint64_t>(64, ae - base);
          if (m) {
            found = true;
            const uint32_t u = rl32(e < ae ? edges[e] : 0u, __builtin_ctzll(m));
            if (lane == 0) {
              atomic_or_word<Word>(&already[a], Word(1) << q1);
              atomic_or_word<Word>(&already[u], Word(1) << q2);
            }
          }
        }
      }
      if (!found) fail |= Word(1) << q0;
    }
    if (fail && lane == 0) {  // s is in one task list once: this wave is its only writer
      unmatch[s] = static_cast<Word>(unmatch[s] | fail);
      set_flag_atomic(flags, s, kModified);
    }
  }
  for (int o = 32; o > 0; o >>= 1) probes += __shfl_xor(probes, o, 64);
  if (lane == 0) {
    const int slot = (blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) & (kCounterSlots - 1);
    slots += probes;
    if (slots) atomicAdd(&counters[slot * kCounterStride + C_SLOTS], static_cast<unsigned long long>(slots));
  }
}

// One wave per source in the task list (active vertices); failing
// (source, template vertex) pairs go to toUnmatch and mark the source
// modified (cc_cpu.tpp:94-129, pc_cpu.tpp:95-147, tds_cpu.tpp:153-188).
// kind 0: cycle, 1: path, 2: TDS walks, 3: enumeration (counts only).
// kOcc: registers capped for 8 waves per SIMD (launches of many waves);
// launches with per-wave memo tables have a fixed, smaller wave count and
// are built without the cap.
template <typename Word, bool kOcc>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kOcc ? 8 : 1))) k_nlcc(int kind, const uint32_t* __restrict__ list,
                                                 const uint32_t* __restrict__ list_count,
                                                 const uint32_t* __restrict__ list2,
                                                 const uint32_t* __restrict__ list2_count,
                                                 const uint64_t* __restrict__ off, const uint32_t* __restrict__ edges,
                                                 const uint32_t* __restrict__ ea, const uint32_t* __restrict__ rev,
                                                 PairIndex pi, const Word* __restrict__ cand,
                                                 Word* __restrict__ unmatch, uint8_t* __restrict__ flags,
                                                 const DevTemplate* __restrict__ t, MemoArgs ma,
                                                 unsigned long long* counters) {
  // list2 (may be null) follows list: sources w >= n1 are list2[w - n1]
  const uint32_t n1 = *list_count;
  const uint32_t n = n1 + (list2 ? *list2_count : 0u);
  if (blockIdx.x == 0 && threadIdx.x == 0) counters[C_ITEMS] = n;
  const uint32_t waves = gridDim.x * (blockDim.x / 64);
  const uint32_t wave_id = uniform(blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6));
  const bool memo = t->nlcc_memo != 0 && kind != 3;
  MemoTable m{};
  Word* already = static_cast<Word*>(ma.already);
  const bool tables = ma.keys != nullptr;
  MemoTable* fm = tables && !memo && kind <= 1 ? &m : nullptr;  // failed-state memo of long walks
  if (tables) {
    const uint64_t base = static_cast<uint64_t>(wave_id) << ma.log2;
    m.keys = ma.keys + base;
    m.vals = ma.vals + base;
    m.epochs = ma.epochs + base;
    m.mask = (1u << ma.log2) - 1;
    m.limit = (3u << ma.log2) / 4;
    m.epoch = ma.wave_epoch[wave_id];
  }
  uint64_t enumerated = 0, slots = 0;
  for (uint32_t w = wave_id; w < n; w += waves) {
    const uint32_t s = w < n1 ? list[w] : list2[w - n1];
    const Word cs = cand[s];
    Word fail = 0;
    const uint32_t L = t->nlcc_len;
    if (memo) {
      bool overflow = false;
      fail = memo_source<Word>(s, cs, t, off, edges, ea, cand, already, m, slots, &overflow);
      if (overflow && lane_id() == 0) ma.ovf_list[atomicAdd(ma.ovf_count, 1u)] = s;  // retried with a larger table
    } else if (kind == 0) {
      for (uint32_t j = 0; j < L; ++j) {
        const uint32_t q = t->nlcc_vertex[j];
        if (!((cs >> q) & 1)) continue;
        const bool ok = L == 3 && rev && off[s + 1] - off[s] <= 64
                            ? triangle_small<Word>(s, static_cast<int>(t->nlcc_vertex[(j + 1) % 3]),
                                                   static_cast<int>(t->nlcc_vertex[(j + 2) % 3]), off, edges, ea,
                                                   rev, pi, cand, slots)
                            : cycle_walk<Word>(s, j, L, t->rows_sorted != 0, off, edges, ea, rev, pi, cand, t, fm,
                                               slots);
        if (!ok) fail |= Word(1) << q;
      }
    } else if (kind == 1) {
      const uint32_t q0 = t->nlcc_vertex[0], q1 = t->nlcc_vertex[L - 1];
      if (((cs >> q0) & 1) && !path_walk<Word>(s, false, L, off, edges, ea, cand, t, fm, slots))
        fail |= Word(1) << q0;
      if (((cs >> q1) & 1) && !path_walk<Word>(s, true, L, off, edges, ea, cand, t, fm, slots))
        fail |= Word(1) << q1;
    } else if (kind == 2) {
      for (uint32_t k = 0; k < t->walk_count; ++k) {
        const uint32_t x = t->walk_root[k];
        if (!((cs >> x) & 1)) continue;
        const uint32_t o = t->walk_off[k];
        const uint32_t len = t->walk_off[k + 1] - o;
        if (!template_walk<Word>(s, len, t->walk_move + o, t->walk_vertex + o, t->walk_mb + o, t->walk_hidx + o,
                                 false, off, edges, ea, cand, slots))
          fail |= Word(1) << x;
      }
    } else {
      const uint32_t x = t->walk_root[0];
      if ((cs >> x) & 1) {
        const uint32_t len = t->walk_off[1];
        enumerated += template_walk<Word>(s, len, t->walk_move, t->walk_vertex, t->walk_mb, t->walk_hidx, true, off,
                                          edges, ea, cand, slots);
      }
    }
    if (fail && lane_id() == 0) {  // s is in one task list once: this wave is its only writer
      unmatch[s] = static_cast<Word>(unmatch[s] | fail);
      set_flag_atomic(flags, s, kModified);
    }
  }
  if (tables && lane_id() == 0) ma.wave_epoch[wave_id] = m.epoch;
  if (lane_id() == 0) {
    const int slot = (blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) & (kCounterSlots - 1);
    if (enumerated) atomicAdd(&counters[slot * kCounterStride + C_ENUM], static_cast<unsigned long long>(enumerated));
    if (slots) atomicAdd(&counters[slot * kCounterStride + C_SLOTS], static_cast<unsigned long long>(slots));
  }
}

template <typename Word>
__global__ void k_cand_to_u64(const Word* __restrict__ cand, uint64_t* __restrict__ out, const uint8_t* __restrict__ flags,
                              uint8_t* __restrict__ vflags, uint32_t V) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t c = static_cast<uint64_t>(cand[v]);
    out[v] = c;
    vflags[v] = static_cast<uint8_t>((flags[v] & (kModified | kScheduled)) | (c ? 1 : 0));
  }
}

template <typename Word>
__global__ void k_u64_to_cand(const uint64_t* __restrict__ in, const uint8_t* __restrict__ vflags, Word* __restrict__ cand,
                              Word* __restrict__ unmatch, uint8_t* __restrict__ flags, uint32_t V) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint8_t vf = vflags[v];
    cand[v] = (vf & 1) ? static_cast<Word>(in[v]) : Word(0);
    unmatch[v] = 0;
    flags[v] = vf & (kModified | kScheduled);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Context and C ABI.

struct pj_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t ev_done = nullptr;  // no timing: polled by wait_stream
  // A second stream for walks of one degree bin that run beside the walks
  // of the other bins on `stream` (fork: the side stream waits for ev_fork;
  // join: `stream` waits for ev_join).
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool side_pending = false;
  std::string err;
  int cus = 256;

  uint32_t V = 0;
  uint64_t E = 0;
  uint64_t* d_off = nullptr;
  uint32_t* d_edges = nullptr;
  uint64_t* d_labels = nullptr;
  uint8_t* d_code = nullptr;
  void* d_cand = nullptr;
  void* d_unmatch = nullptr;
  void* d_u64 = nullptr;   // staging for downloads / uploads; first hits of an NLCC phase 2 (push_modified)
  uint8_t* d_flags = nullptr;
  uint8_t* d_vflags = nullptr;
  uint8_t* d_dirty = nullptr;  // per vertex: candidates of v or of a neighbour changed since v's last evaluation
  uint8_t* d_rescan = nullptr;       // per vertex: a neighbour turned modified since v's last phase 2 scan
  uint32_t* d_last_scan = nullptr;   // per vertex: slots its last phase 2 scan read
  uint32_t* d_hub_new = nullptr;     // per hub: turned modified in the last phase 1
  uint32_t* d_ea = nullptr;
  uint32_t* d_rev = nullptr;  // per slot u->v: position of u in row(v) (symmetric graphs only)
  // Per slot: index of the neighbour's label among the graph's distinct
  // labels (graphs of at most 255 labels; null otherwise), and those labels.
  uint8_t* d_slot_lab = nullptr;
  std::vector<uint64_t> graph_labels;
  // Pair index of the heavy rows (PairIndex): decorated offsets [V + 1] and
  // positions [pair_heavy^2]; null when the graph has fewer than two rows
  // above the threshold that fits the budget, or no reverse positions.
  uint64_t* d_pair_doff = nullptr;
  uint32_t* d_pair_pos = nullptr;
  uint64_t pair_heavy = 0;
  uint64_t pair_degree = 0;            // rows of more slots than this are heavy
  uint64_t pair_min_degree = 512;      // PJ_PAIR_DEGREE: the smallest threshold tried
  uint64_t pair_budget = 8ull << 30;   // PJ_PAIR_INDEX_MB: bytes the positions may take (0: no index)
  // 0 after a state reset, 1 after LCC0 on such a state (every active
  // vertex's candidates are then the template vertices of its label), 2 once
  // anything else ran: the first LCC step reads slot labels instead of
  // neighbour candidates only at 1.
  int fresh = 2;
  bool slot_labels_enabled = true;  // PJ_SLOT_LABELS=0: never
  bool phase1_fused = true;          // PJ_PHASE1_FUSED=0: hub chunks of LCC phase 1 in a launch of their own
  int phase1_blocks = kPhase1BlocksPerCu;  // PJ_PHASE1_BLOCKS: blocks per compute unit of the phase-1 launch
  uint32_t side_walks = 0xC;         // PJ_SIDE_WALKS: bins whose 3-vertex cycle walks run on the side stream
  bool wide_triangles = true;        // PJ_WIDE_TRIANGLES=0: 257-2048-slot cycle sources walk
  // Live list: the vertices that may still be active (a superset of the
  // active ones; candidates only shrink until the state is replaced).
  // live == nullptr: every vertex.  Two buffers for the compaction.
  uint32_t* d_live[2] = {nullptr, nullptr};
  const uint32_t* live = nullptr;
  uint32_t n_live = 0;
  uint32_t live_min = 1u << 16;  // lists this short are not compacted further (PJ_LIVE_MIN)
  uint64_t ea_words = 0;
  uint32_t ea_tail = 0;  // host staging of the last edge-active word
  bool rows_sorted = false;
  bool graph_shared = false;  // CSR, labels and hub lists belong to another context (pj_share_graph)
  bool symmetric = false;   // every slot u->v has a reverse slot v->u
  bool skip_clean = false;  // phase 1 skips clean scheduled vertices (rows sorted and symmetric)
  bool dirty_enabled = true;
  bool rescan_all = false;  // every rescan mark is set (phase 1 need not push)
  bool self_loops = true;       // some row holds its own vertex (known only for symmetric graphs)
  bool triangle_memo_enabled = true;  // PJ_TRIANGLE_MEMO=0: memoised 3-cycles as walks
  bool reverse_enabled = true;  // keep d_rev for symmetric graphs (PJ_REVERSE_SLOTS=0: no)
  bool labels_consistent = false;  // every candidate of every vertex carries the vertex's label (after LCC0)
  bool modified_cleared = false;   // no modified flag set since pj_reset_modified
  bool memo_walks = false;  // run every cycle / path step as memoised walks (PJ_MEMO_WALKS=1)
  int memo_log2 = 0;  // first-pass table entries per wave, log2 (PJ_MEMO_LOG2, 6..20; 0: by graph size)
  uint32_t memo_lds_limit = 0;  // PJ_MEMO_LDS_LIMIT: live entries of the LDS tables (0: 768)
  // memoised walks: per-wave tables (two passes), alreadyMatched, overflow list
  MemoArgs memo1{};
  void* d_already = nullptr;  // Word[V]
  // active view for the memoised walks (k_view_*)
  uint64_t* d_view_count = nullptr;  // V + 1
  uint64_t* d_view_off = nullptr;    // V + 1
  uint32_t* d_view_edges = nullptr;
  uint64_t view_cap = 0;
  uint64_t* d_scan_tmp = nullptr;  // per-span totals of the view-count scan
  uint32_t* d_ovf = nullptr;  // [0] first-pass overflows, [1] second-pass overflows, then the list
  uint32_t* d_bins = nullptr;        // kBins * V
  uint32_t* d_bin_count = nullptr;   // kBins: the set of the ring the last frontier build filled
  uint32_t* d_bin_ring = nullptr;    // kBinRing sets of kBins bin counts
  uint32_t bin_slot = 1;             // sets of the ring from here on are zero
  bool counters_clean = false;       // every counter block in use and every ring set are zero
  unsigned long long* d_counters = nullptr;  // kCounterSlots * kCounterStride
  // hubs
  uint32_t nhubs = 0, nchunks = 0;
  uint32_t* d_hubs = nullptr;
  uint32_t* d_chunk_hub = nullptr;
  uint64_t* d_chunk_start = nullptr;
  uint32_t* d_hub_count = nullptr;
  uint32_t* d_hub_elim = nullptr;
  uint32_t* d_nhubs = nullptr;       // device copy of nhubs (NLCC task count)
  uint32_t* d_hub_hit = nullptr;     // phase-2 first-hit offset per hub

  DevTemplate* d_t = nullptr;
  DevTemplate h_t{};
  bool have_template = false;
  int word_bytes = 8;
  int kw = 1;
  // host copy of the constraint tables
  std::vector<std::vector<uint32_t>> cycles, paths;
  std::vector<uint64_t> path_label;
  struct HostWalk {
    uint32_t root;
    std::vector<uint8_t> move;
    std::vector<uint32_t> vertex, mb;
  };
  std::vector<std::vector<HostWalk>> tds;
  HostWalk enumeration;

  uint64_t active_vertices = 0, active_edges = 0;
  // kernel timing (pj_set_kernel_timing)
  bool timing = false, cur_timed = false;
  int nlaunch = 0;
  int launch_kid[kMaxLaunches] = {};
  int timing_kid = -1;  // time only launches of this kind (-1: every launch)
  hipEvent_t lev[2 * kMaxLaunches] = {};
  uint64_t k_launches[K_COUNT] = {}, k_items[K_COUNT] = {}, k_slots[K_COUNT] = {}, k_elim[K_COUNT] = {};
  double k_ms[K_COUNT] = {};
  float last_ms = 0.f;
  uint64_t last_slots = 0;
  unsigned long long* h_sums = nullptr;   // pinned: per counter block, kCounterStride sums
  unsigned long long* h_sums_dev = nullptr;  // the device's address of h_sums (k_sum_counters writes it)
  uint64_t* h_word = nullptr;  // pinned: one value read back between launches (counts, totals)
  unsigned long long* d_sums = nullptr;
  // Active-graph compaction: the counters of pj_compact_active and the
  // device outputs of pj_download_active (grown on demand).
  unsigned long long* d_compact_scratch = nullptr;
  uint32_t *d_out_v = nullptr, *d_out_src = nullptr, *d_out_dst = nullptr;
  uint64_t* d_out_cand = nullptr;
  uint64_t out_cap_v = 0, out_cap_e = 0;
  // Label statistics (pj_graph_stat): counts and pair table, pjs::kBufferBytes.
  unsigned long long* d_stat = nullptr;
};

namespace {

#define PJ_HIP(call)                                                             \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);              \
      return -1;                                                                 \
    }                                                                            \
  } while (0)

// Waits until the work queued on the context's stream is done by polling a
// marker event.  A blocking wait lets the thread sleep until an interrupt
// wakes it; at the end of each pruning step, where the host reads the
// step's counts before it queues the next step, that wake-up leaves the GPU
// idle for tens of microseconds.
hipError_t wait_stream(pj_ctx* ctx) {
  hipError_t e = hipEventRecord(ctx->ev_done, ctx->stream);
  if (e != hipSuccess) return e;
  while ((e = hipEventQuery(ctx->ev_done)) == hipErrorNotReady) __builtin_ia32_pause();
  return e;
}

int grid_for(const pj_ctx* ctx, uint64_t items, int per_block = kBlock) {
  uint64_t g = (items + per_block - 1) / per_block;
  uint64_t cap = static_cast<uint64_t>(ctx->cus) * 16;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return static_cast<int>(g);
}

template <typename T>
int dev_alloc(pj_ctx* ctx, T** p, uint64_t count) {
  if (*p) {
    hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) count = 1;
  PJ_HIP(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
  return 0;
}

void free_graph(pj_ctx* ctx) {
  if (ctx->graph_shared) {  // the owner frees these
    ctx->d_off = nullptr;
    ctx->d_edges = nullptr;
    ctx->d_labels = nullptr;
    ctx->d_rev = nullptr;
    ctx->d_slot_lab = nullptr;
    ctx->d_pair_doff = nullptr;
    ctx->d_pair_pos = nullptr;
    ctx->d_hubs = ctx->d_chunk_hub = nullptr;
    ctx->d_chunk_start = nullptr;
    ctx->graph_shared = false;
  }
  void* bufs[] = {ctx->d_off, ctx->d_edges, ctx->d_labels, ctx->d_code, ctx->d_cand, ctx->d_unmatch, ctx->d_u64,
                  ctx->d_flags, ctx->d_vflags, ctx->d_dirty, ctx->d_rescan, ctx->d_ea, ctx->d_bins, ctx->d_bin_ring,
                  ctx->d_hubs,
                  ctx->d_chunk_hub, ctx->d_chunk_start, ctx->d_hub_count, ctx->d_hub_elim, ctx->d_nhubs, ctx->d_hub_hit,
                  ctx->d_last_scan, ctx->d_hub_new, ctx->memo1.keys, ctx->memo1.vals, ctx->memo1.epochs,
                  ctx->memo1.wave_epoch, ctx->d_already, ctx->d_ovf, ctx->d_view_count, ctx->d_view_off,
                  ctx->d_view_edges, ctx->d_scan_tmp, ctx->d_rev, ctx->d_live[0], ctx->d_live[1],
                  ctx->d_compact_scratch, ctx->d_out_v, ctx->d_out_src, ctx->d_out_dst, ctx->d_out_cand,
                  ctx->d_stat, ctx->d_slot_lab, ctx->d_pair_doff, ctx->d_pair_pos};
  for (void* b : bufs)
    if (b) hipFree(b);
  ctx->d_pair_doff = nullptr;
  ctx->d_pair_pos = nullptr;
  ctx->pair_heavy = ctx->pair_degree = 0;
  ctx->d_off = nullptr;
  ctx->d_edges = nullptr;
  ctx->d_labels = nullptr;
  ctx->d_code = nullptr;
  ctx->d_cand = ctx->d_unmatch = ctx->d_u64 = nullptr;
  ctx->d_flags = ctx->d_vflags = ctx->d_dirty = ctx->d_rescan = nullptr;
  ctx->d_ea = nullptr;
  ctx->d_rev = nullptr;
  ctx->d_slot_lab = nullptr;
  ctx->graph_labels.clear();
  ctx->d_live[0] = ctx->d_live[1] = nullptr;
  ctx->d_compact_scratch = nullptr;
  ctx->d_stat = nullptr;
  ctx->d_out_v = ctx->d_out_src = ctx->d_out_dst = nullptr;
  ctx->d_out_cand = nullptr;
  ctx->out_cap_v = ctx->out_cap_e = 0;
  ctx->live = nullptr;
  ctx->n_live = 0;
  ctx->d_bins = ctx->d_bin_count = ctx->d_bin_ring = nullptr;
  ctx->counters_clean = false;
  ctx->d_hubs = ctx->d_chunk_hub = nullptr;
  ctx->d_chunk_start = nullptr;
  ctx->d_hub_count = ctx->d_hub_elim = nullptr;
  ctx->d_nhubs = nullptr;
  ctx->d_hub_hit = nullptr;
  ctx->d_last_scan = ctx->d_hub_new = nullptr;
  ctx->memo1 = MemoArgs{};
  ctx->d_already = nullptr;
  ctx->d_ovf = nullptr;
  ctx->d_view_count = ctx->d_view_off = nullptr;
  ctx->d_view_edges = nullptr;
  ctx->d_scan_tmp = nullptr;
  ctx->view_cap = 0;
  ctx->V = 0;
  ctx->E = 0;
}

// Builds the static hub chunk list from host offsets.
// The chunk list holds the first chunk of every hub, then the second of every
// hub with two or more, and so on: the grid reaches the start of every row
// first, which lets the phase 2 scan (first hit in slot order) skip the later
// chunks of rows whose hit is already found.
int setup_hubs(pj_ctx* ctx, const uint64_t* h_off) {
  std::vector<uint32_t> hubs, chunk_hub;
  std::vector<uint64_t> chunk_start;
  uint64_t max_chunks = 0;
  for (uint32_t v = 0; v < ctx->V; ++v) {
    uint64_t d = h_off[v + 1] - h_off[v];
    if (d <= kHubDegree) continue;
    hubs.push_back(v);
    max_chunks = std::max<uint64_t>(max_chunks, (d + kHubChunk - 1) / kHubChunk);
  }
  // hubs by chunk count, descending: those with more than r chunks are a prefix
  std::vector<uint32_t> by_count(hubs.size());
  std::vector<uint64_t> count(hubs.size());
  for (uint32_t h = 0; h < hubs.size(); ++h) {
    by_count[h] = h;
    count[h] = (h_off[hubs[h] + 1] - h_off[hubs[h]] + kHubChunk - 1) / kHubChunk;
  }
  std::stable_sort(by_count.begin(), by_count.end(), [&](uint32_t x, uint32_t y) { return count[x] > count[y]; });
  for (uint64_t r = 0; r < max_chunks; ++r)
    for (uint32_t i = 0; i < by_count.size() && count[by_count[i]] > r; ++i) {
      const uint32_t h = by_count[i];
      chunk_hub.push_back(h);
      chunk_start.push_back(h_off[hubs[h]] + r * kHubChunk);
    }
  ctx->nhubs = static_cast<uint32_t>(hubs.size());
  ctx->nchunks = static_cast<uint32_t>(chunk_hub.size());
  if (dev_alloc(ctx, &ctx->d_hubs, hubs.size()) || dev_alloc(ctx, &ctx->d_chunk_hub, chunk_hub.size()) ||
      dev_alloc(ctx, &ctx->d_chunk_start, chunk_start.size()) ||
      dev_alloc(ctx, &ctx->d_hub_count, static_cast<uint64_t>(hubs.size()) * kMaxLabels) ||
      dev_alloc(ctx, &ctx->d_hub_elim, hubs.size()) || dev_alloc(ctx, &ctx->d_nhubs, 1) ||
      dev_alloc(ctx, &ctx->d_hub_hit, hubs.size()) || dev_alloc(ctx, &ctx->d_hub_new, hubs.size()))
    return -1;
  PJ_HIP(hipMemcpyAsync(ctx->d_nhubs, &ctx->nhubs, 4, hipMemcpyHostToDevice, ctx->stream));
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  if (!hubs.empty()) {
    PJ_HIP(hipMemcpyAsync(ctx->d_hubs, hubs.data(), hubs.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    PJ_HIP(hipMemcpyAsync(ctx->d_chunk_hub, chunk_hub.data(), chunk_hub.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    PJ_HIP(hipMemcpyAsync(ctx->d_chunk_start, chunk_start.data(), chunk_start.size() * 8, hipMemcpyHostToDevice,
                          ctx->stream));
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, hubs.size() * kMaxLabels)), dim3(kBlock), 0, ctx->stream,
                       ctx->d_hub_count, static_cast<uint64_t>(hubs.size()) * kMaxLabels, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, hubs.size())), dim3(kBlock), 0, ctx->stream, ctx->d_hub_elim,
                       static_cast<uint64_t>(hubs.size()), 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, hubs.size())), dim3(kBlock), 0, ctx->stream, ctx->d_hub_hit,
                       static_cast<uint64_t>(hubs.size()), 0xFFFFFFFFu);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, hubs.size())), dim3(kBlock), 0, ctx->stream, ctx->d_hub_new,
                       static_cast<uint64_t>(hubs.size()), 0u);
    PJ_HIP(hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

int alloc_state(pj_ctx* ctx) {
  const uint64_t V = ctx->V;
  if (dev_alloc(ctx, &ctx->d_code, V) || dev_alloc(ctx, reinterpret_cast<uint64_t**>(&ctx->d_cand), V) ||
      dev_alloc(ctx, reinterpret_cast<uint64_t**>(&ctx->d_unmatch), V) ||
      dev_alloc(ctx, reinterpret_cast<uint64_t**>(&ctx->d_u64), V) || dev_alloc(ctx, &ctx->d_flags, V + 4) ||
      dev_alloc(ctx, &ctx->d_vflags, V + 4) || dev_alloc(ctx, &ctx->d_dirty, V + 4) ||
      dev_alloc(ctx, &ctx->d_rescan, V + 4) || dev_alloc(ctx, &ctx->d_last_scan, V) ||
      dev_alloc(ctx, &ctx->d_ea, ctx->ea_words) ||
      dev_alloc(ctx, &ctx->d_bins, V * kBins) || dev_alloc(ctx, &ctx->d_bin_ring, kBinRing * kBins) ||
      dev_alloc(ctx, &ctx->d_live[0], V + 1) || dev_alloc(ctx, &ctx->d_live[1], V + 1) ||
      dev_alloc(ctx, &ctx->d_counters, (kMaxLaunches + 1) * kCounterBlock))
    return -1;
  ctx->d_bin_count = ctx->d_bin_ring;  // set 0 serves the upload's checks
  ctx->bin_slot = 1;
  ctx->counters_clean = false;
  return 0;
}

// Marks every vertex for a phase 2 rescan (modified flags were cleared or
// set outside phase 1).
void mark_all_rescan(pj_ctx* ctx) {
  if (!ctx->skip_clean) return;
  ctx->rescan_all = true;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, (ctx->V + 4) / 4)), dim3(kBlock), 0, ctx->stream,
                     reinterpret_cast<uint32_t*>(ctx->d_rescan), static_cast<uint64_t>((ctx->V + 4) / 4), 0x01010101u);
}

// Marks every vertex for re-evaluation (the candidate sets were replaced).
void mark_all_dirty(pj_ctx* ctx) {
  if (!ctx->skip_clean) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, (ctx->V + 4) / 4)), dim3(kBlock), 0, ctx->stream,
                     reinterpret_cast<uint32_t*>(ctx->d_dirty), static_cast<uint64_t>((ctx->V + 4) / 4), 0x01010101u);
  mark_all_rescan(ctx);
}

template <typename Word>
int state_reset_t(pj_ctx* ctx) {
  const Word full = ctx->h_t.P >= 8 * sizeof(Word) ? static_cast<Word>(~Word(0))
                                                   : static_cast<Word>((Word(1) << ctx->h_t.P) - 1);
  hipLaunchKernelGGL(k_state_init<Word>, dim3(grid_for(ctx, ctx->V)), dim3(kBlock), 0, ctx->stream,
                     static_cast<Word*>(ctx->d_cand), static_cast<Word*>(ctx->d_unmatch), ctx->d_flags, ctx->V, full);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, ctx->ea_words)), dim3(kBlock), 0, ctx->stream, ctx->d_ea,
                     ctx->ea_words, 0xFFFFFFFFu);
  if (ctx->E & 31) {  // bits past the last slot stay 0
    ctx->ea_tail = (1u << (ctx->E & 31)) - 1u;
    PJ_HIP(hipMemcpyAsync(ctx->d_ea + ctx->ea_words - 1, &ctx->ea_tail, 4, hipMemcpyHostToDevice, ctx->stream));
  }
  mark_all_dirty(ctx);
  PJ_HIP(hipGetLastError());
  PJ_HIP(wait_stream(ctx));
  ctx->active_vertices = ctx->V;
  ctx->active_edges = ctx->E;
  ctx->labels_consistent = false;  // every template vertex is a candidate again
  ctx->fresh = 0;
  ctx->modified_cleared = false;
  ctx->live = nullptr;              // every vertex is active again
  ctx->n_live = ctx->V;
  return 0;
}

// Counter block of the next launch.  Without kernel timing every launch
// adds into block 0; with it each launch gets its own block and a pair of
// HIP events so that pj_kernel_stats can attribute time and work per kernel.
unsigned long long* begin_launch(pj_ctx* ctx, int kid) {
  if (!ctx->timing || (ctx->timing_kid >= 0 && kid != ctx->timing_kid) || ctx->nlaunch >= kMaxLaunches) {
    ctx->cur_timed = false;
    return ctx->d_counters;
  }
  const int n = ctx->nlaunch;
  ctx->launch_kid[n] = kid;
  hipEventRecord(ctx->lev[2 * n], ctx->stream);
  ctx->cur_timed = true;
  return ctx->d_counters + static_cast<size_t>(n + 1) * kCounterBlock;
}

void end_launch(pj_ctx* ctx) {
  if (!ctx->cur_timed) return;
  hipEventRecord(ctx->lev[2 * ctx->nlaunch + 1], ctx->stream);
  ++ctx->nlaunch;
  ctx->cur_timed = false;
}

// Counters and bin counts start each step at zero.  The sums kernel at the
// end of a step clears them (counters_clean), so a step that follows one
// queues no clearing launch; anything else that may have written them
// leaves counters_clean false and the clearing is queued here.
int zero_counters(pj_ctx* ctx) {
  if (!ctx->counters_clean) {
    const uint64_t n = ctx->timing ? static_cast<uint64_t>(kMaxLaunches + 1) * kCounterBlock : kCounterBlock;
    hipLaunchKernelGGL(k_fill_u64, dim3(ctx->timing ? 64 : 4), dim3(kBlock), 0, ctx->stream, ctx->d_counters, n,
                       0ull);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->d_bin_ring,
                       static_cast<uint64_t>(kBinRing * kBins), 0u);
  }
  ctx->counters_clean = false;
  ctx->bin_slot = 0;
  ctx->nlaunch = 0;
  return 0;
}

// Each frontier build takes the next zero set of bin counts of the ring;
// a step that uses the whole ring clears it (after every launch queued
// before, which may still read it) and starts over.
int zero_bins(pj_ctx* ctx) {
  if (ctx->bin_slot >= kBinRing) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->d_bin_ring,
                       static_cast<uint64_t>(kBinRing * kBins), 0u);
    ctx->bin_slot = 0;
  }
  ctx->d_bin_count = ctx->d_bin_ring + static_cast<size_t>(ctx->bin_slot++) * kBins;
  return 0;
}

template <typename Word>
void launch_bins(pj_ctx* ctx, int mode) {
  zero_bins(ctx);
  unsigned long long* c = begin_launch(ctx, K_BINS);
  const uint32_t blocks = static_cast<uint32_t>((static_cast<uint64_t>(ctx->n_live) + kBinVerts - 1) / kBinVerts);
  uint8_t* marks = !ctx->skip_clean  ? nullptr
                   : mode == 0       ? ctx->d_dirty
                   : mode == 3 || mode == 4 ? ctx->d_rescan
                                            : nullptr;
  hipLaunchKernelGGL(k_build_bins<Word>, dim3(std::max<uint32_t>(blocks, 1)), dim3(kBlock), 0, ctx->stream,
                     ctx->live, ctx->n_live, static_cast<const Word*>(ctx->d_cand), ctx->d_flags, ctx->d_code, marks,
                     static_cast<Word*>(ctx->d_unmatch), ctx->d_off, ctx->V, mode, ctx->d_t, ctx->d_bins,
                     ctx->d_bin_count, ctx->d_last_scan, mode == 6 ? static_cast<const uint32_t*>(ctx->d_u64) : nullptr,
                     c);
  end_launch(ctx);
}

// Phase 2 over every active vertex (bins + hubs).
template <typename Word>
void launch_phase2(pj_ctx* ctx, int mode = 3) {
  launch_bins<Word>(ctx, mode);
  uint8_t* dirty = ctx->skip_clean ? ctx->d_dirty : nullptr;
  uint8_t* rescan = ctx->skip_clean ? ctx->d_rescan : nullptr;
  uint32_t* last = ctx->d_last_scan;
  Word* cand = static_cast<Word*>(ctx->d_cand);
  Word* unm = static_cast<Word*>(ctx->d_unmatch);
  unsigned long long* c = begin_launch(ctx, K_P2);
  hipLaunchKernelGGL(k_phase2_all<Word>, dim3(ctx->cus * kPhase2BlocksPerCu), dim3(kBlock), 0, ctx->stream, ctx->d_bins,
                     ctx->d_bin_count, ctx->V, ctx->d_chunk_hub, ctx->d_chunk_start, ctx->nhubs ? ctx->nchunks : 0u,
                     ctx->d_hubs, ctx->d_off, ctx->d_edges, cand, unm, ctx->d_flags, dirty, rescan, last, ctx->d_ea,
                     ctx->d_hub_hit, c);
  end_launch(ctx);
  if (ctx->nhubs) {
    c = begin_launch(ctx, K_P2_HUB_FINISH);
    hipLaunchKernelGGL(k_phase2_hub_finish<Word>, dim3(grid_for(ctx, ctx->nhubs)), dim3(kBlock), 0, ctx->stream,
                       ctx->d_hubs, ctx->nhubs, ctx->d_off, cand, unm, ctx->d_flags, dirty, rescan, last,
                       ctx->d_hub_hit, c);
    end_launch(ctx);
  }
  ctx->rescan_all = false;  // the scans cleared the marks of the vertices they read
}

// Reverse positions of the full CSR for complete walks and NLCC pushes (null
// where the graph is not known to be sorted and symmetric, or they are
// switched off).
const uint32_t* reverse_positions(const pj_ctx* ctx) {
  return ctx->rows_sorted && ctx->symmetric ? ctx->d_rev : nullptr;
}

PairIndex pair_index(const pj_ctx* ctx) {
  if (!ctx->d_pair_pos || !reverse_positions(ctx)) return PairIndex{nullptr, nullptr, 0};
  return PairIndex{ctx->d_pair_doff, ctx->d_pair_pos, ctx->pair_heavy};
}

// The push of an NLCC phase 2: rescan marks cleared, then every modified
// vertex (bins of mode 5, hubs by chunks) pushes to the vertices of its row.
// With reverse positions the push computes each neighbour's first hit
// (first[u], in the staging array) and the phase 2 that follows (mode 6)
// scans no row of an unmodified vertex; without them it sets rescan marks
// (mode 4).  Returns the phase 2 mode.
template <typename Word>
int push_modified(pj_ctx* ctx) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, (ctx->V + 4) / 4)), dim3(kBlock), 0, ctx->stream,
                     reinterpret_cast<uint32_t*>(ctx->d_rescan), static_cast<uint64_t>((ctx->V + 4) / 4), 0u);
  const uint32_t* rev = reverse_positions(ctx);
  uint32_t* first = rev ? static_cast<uint32_t*>(ctx->d_u64) : nullptr;
  const uint32_t* ea = rev ? ctx->d_ea : nullptr;
  if (first)
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, ctx->V)), dim3(kBlock), 0, ctx->stream, first,
                       static_cast<uint64_t>(ctx->V), 0xFFFFFFFFu);
  launch_bins<Word>(ctx, 5);
  const uint64_t V = ctx->V;
  const int g = ctx->cus * 8;
  unsigned long long* c = begin_launch(ctx, K_HUB_PUSH);
  hipLaunchKernelGGL(k_push_rescan<4>, dim3(g), dim3(kBlock), 0, ctx->stream, ctx->d_bins, ctx->d_bin_count + 0,
                     ctx->d_off, ctx->d_edges, ctx->d_rescan, ea, rev, first, c);
  hipLaunchKernelGGL(k_push_rescan<16>, dim3(g), dim3(kBlock), 0, ctx->stream, ctx->d_bins + V,
                     ctx->d_bin_count + 1, ctx->d_off, ctx->d_edges, ctx->d_rescan, ea, rev, first, c);
  hipLaunchKernelGGL(k_push_rescan<32>, dim3(g), dim3(kBlock), 0, ctx->stream, ctx->d_bins + 2 * V,
                     ctx->d_bin_count + 2, ctx->d_off, ctx->d_edges, ctx->d_rescan, ea, rev, first, c);
  hipLaunchKernelGGL(k_push_rescan<64>, dim3(g), dim3(kBlock), 0, ctx->stream, ctx->d_bins + 3 * V,
                     ctx->d_bin_count + 3, ctx->d_off, ctx->d_edges, ctx->d_rescan, ea, rev, first, c);
  if (ctx->nhubs)
    hipLaunchKernelGGL(k_hub_rescan_push, dim3(std::min<uint32_t>(ctx->nchunks, ctx->cus * 8)), dim3(kBlock), 0,
                       ctx->stream, ctx->d_chunk_hub, ctx->d_chunk_start, ctx->nchunks, ctx->d_hubs, ctx->d_off,
                       ctx->d_edges, ctx->d_hub_new, ctx->d_flags, ctx->d_rescan, ea, rev, first, c);
  end_launch(ctx);
  ctx->rescan_all = false;
  return first ? 6 : 4;
}

template <typename Word, int KW>
void launch_lcc_phase1(pj_ctx* ctx) {
  launch_bins<Word>(ctx, 0);
  const bool derive = ctx->labels_consistent;
  const uint8_t* slot_lab = ctx->fresh == 1 && ctx->slot_labels_enabled ? ctx->d_slot_lab : nullptr;
  uint8_t* dirty = ctx->skip_clean ? ctx->d_dirty : nullptr;
  uint8_t* rescan = ctx->skip_clean && !ctx->rescan_all ? ctx->d_rescan : nullptr;  // push target
  const Word* cand = static_cast<const Word*>(ctx->d_cand);
  Word* unm = static_cast<Word*>(ctx->d_unmatch);
  unsigned long long* c = begin_launch(ctx, slot_lab ? K_LCC_P1_LABELS : K_LCC_P1);
  // the first step after LCC0 (slot labels) has a build of its own: one
  // body holding both ways of reading neighbours spills registers
  const bool fused = ctx->nhubs && ctx->phase1_fused;
  if (fused)
    hipLaunchKernelGGL((slot_lab ? k_lcc_phase1_hubs<Word, KW, true> : k_lcc_phase1_hubs<Word, KW, false>),
                       dim3(ctx->cus * ctx->phase1_blocks), dim3(kBlock), 0, ctx->stream, ctx->d_bins,
                       ctx->d_bin_count, ctx->V, ctx->d_chunk_hub, ctx->d_chunk_start, ctx->nchunks, ctx->d_hubs,
                       ctx->d_hub_count, ctx->d_hub_elim, ctx->d_off, ctx->d_edges, ctx->d_code, derive, slot_lab,
                       cand, unm, ctx->d_flags, dirty, rescan, ctx->d_ea, ctx->d_t, c);
  else
    hipLaunchKernelGGL((slot_lab ? k_lcc_phase1_all<Word, KW, true> : k_lcc_phase1_all<Word, KW, false>),
                       dim3(ctx->cus * ctx->phase1_blocks), dim3(kBlock), 0, ctx->stream, ctx->d_bins,
                       ctx->d_bin_count, ctx->V, ctx->d_off, ctx->d_edges, ctx->d_code, derive, slot_lab, cand, unm,
                       ctx->d_flags, dirty, rescan, ctx->d_ea, ctx->d_t, c);
  end_launch(ctx);
  if (ctx->nhubs) {
    if (!fused) {
    c = begin_launch(ctx, slot_lab ? K_LCC_HUB_LABELS : K_LCC_HUB);
    hipLaunchKernelGGL((slot_lab ? k_lcc_hub_chunks<Word, KW, true> : k_lcc_hub_chunks<Word, KW, false>),
                       dim3(std::min<uint32_t>(ctx->nchunks, ctx->cus * kPhase1BlocksPerCu)),
                       dim3(kBlock), 0,
                       ctx->stream, ctx->d_chunk_hub, ctx->d_chunk_start, ctx->nchunks, ctx->d_hubs, ctx->d_off,
                       ctx->d_edges, ctx->d_code, derive, slot_lab, cand, ctx->d_flags, dirty, ctx->d_ea,
                       ctx->d_hub_count, ctx->d_hub_elim, ctx->d_t, c);
    end_launch(ctx);
    }
    begin_launch(ctx, K_LCC_HUB_FINISH);
    hipLaunchKernelGGL((k_lcc_hub_finish<Word, KW>), dim3(grid_for(ctx, ctx->nhubs)), dim3(kBlock), 0, ctx->stream,
                       ctx->d_hubs, ctx->nhubs, ctx->d_code, cand, unm, ctx->d_flags, dirty, ctx->d_hub_count,
                       ctx->d_hub_elim, ctx->d_hub_new, ctx->d_t);
    end_launch(ctx);
    if (rescan) {
      c = begin_launch(ctx, K_HUB_PUSH);
      hipLaunchKernelGGL(k_hub_rescan_push, dim3(std::min<uint32_t>(ctx->nchunks, ctx->cus * 8)), dim3(kBlock), 0,
                         ctx->stream, ctx->d_chunk_hub, ctx->d_chunk_start, ctx->nchunks, ctx->d_hubs, ctx->d_off,
                         ctx->d_edges, ctx->d_hub_new, nullptr, rescan, nullptr, nullptr, nullptr, c);
      end_launch(ctx);
    }
  }
}

template <typename Word>
void launch_lcc(pj_ctx* ctx) {
  switch (ctx->kw) {
    case 1:
      launch_lcc_phase1<Word, 1>(ctx);
      break;
    case 2:
      launch_lcc_phase1<Word, 2>(ctx);
      break;
    case 4:
      launch_lcc_phase1<Word, 4>(ctx);
      break;
    default:
      launch_lcc_phase1<Word, 8>(ctx);
      break;
  }
  launch_phase2<Word>(ctx);
}

// Memoised walks: tables are allocated on first use (the second pass only
// when a source outgrows the first); alreadyMatched and the overflow
// counters are cleared for every step.
int alloc_memo(pj_ctx* ctx, MemoArgs* m, int waves, int log2) {
  if (m->keys) return 0;
  const uint64_t n = static_cast<uint64_t>(waves) << log2;
  if (dev_alloc(ctx, &m->keys, n) || dev_alloc(ctx, &m->vals, n) || dev_alloc(ctx, &m->epochs, n) ||
      dev_alloc(ctx, &m->wave_epoch, waves))
    return -1;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, n)), dim3(kBlock), 0, ctx->stream, m->epochs, n, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlock), 0, ctx->stream, m->wave_epoch, static_cast<uint64_t>(waves),
                     0u);
  m->log2 = static_cast<uint32_t>(log2);
  m->already = ctx->d_already;
  return 0;
}

int prepare_memo(pj_ctx* ctx, bool tables = true) {
  const uint64_t V = ctx->V;
  if (!ctx->d_already) {
    uint64_t* a = nullptr;
    if (dev_alloc(ctx, &a, V + 1) || dev_alloc(ctx, &ctx->d_ovf, 2 * V + 2)) return -1;
    ctx->d_already = a;
  }
  if (!tables) {  // alreadyMatched only
    const uint64_t words = (V * ctx->word_bytes + 3) / 4;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, words)), dim3(kBlock), 0, ctx->stream,
                       static_cast<uint32_t*>(ctx->d_already), words, 0u);
    PJ_HIP(hipGetLastError());
    return 0;
  }
  // Walks on larger graphs reach more states: 2^16 entries per wave up to
  // 2^22 vertices, doubling per doubling of V up to 2^18 (2048 waves x 2^18
  // x 20 B = 10.7 GB from 2^24 vertices).  A source that outgrows its table
  // reruns from the start with a larger one, so the floor is set by the
  // heavy sources of small graphs: with 2^14 entries those of the scale-18
  // TDS workload ran twice (135 ms per solve against 79 ms with 2^16).
  int log2 = ctx->memo_log2;
  if (!log2) {
    int bits = 0;
    while (bits < 63 && (1ull << bits) < V) ++bits;
    log2 = std::min(18, std::max(kMemoLog2, bits - 6));
  }
  if (alloc_memo(ctx, &ctx->memo1, kMemoWaves, log2)) return -1;
  ctx->memo1.ovf_count = ctx->d_ovf;
  ctx->memo1.ovf_list = ctx->d_ovf + 2;
  ctx->memo1.lds_limit = ctx->memo_lds_limit;
  const uint64_t words = (V * ctx->word_bytes + 3) / 4;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, words)), dim3(kBlock), 0, ctx->stream,
                     static_cast<uint32_t*>(ctx->d_already), words, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->d_ovf, static_cast<uint64_t>(2), 0u);
  PJ_HIP(hipGetLastError());
  return 0;
}

// Builds the active view (k_view_*) from the mode 1 bins (every active
// vertex) and the hub list; the edge array grows to what a view needs.
template <typename Word>
int build_view(pj_ctx* ctx) {
  const uint64_t V = ctx->V;
  if (!ctx->d_view_off) {
    if (dev_alloc(ctx, &ctx->d_view_count, V + 1) || dev_alloc(ctx, &ctx->d_view_off, V + 1)) return -1;
    const uint64_t spans = (V + 1 + kScanSpan - 1) / kScanSpan;
    if (dev_alloc(ctx, &ctx->d_scan_tmp, spans)) return -1;
  }
  const Word* cand = static_cast<const Word*>(ctx->d_cand);
  const int g = ctx->cus * 8;
  const uint32_t hub_blocks = std::max<uint32_t>(1, std::min<uint32_t>((ctx->nhubs + 3) / 4, ctx->cus * 4));
  begin_launch(ctx, K_VIEW);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ctx, V + 1)), dim3(kBlock), 0, ctx->stream,
                     reinterpret_cast<unsigned long long*>(ctx->d_view_count), V + 1, 0ull);
  hipLaunchKernelGGL((k_view_count<Word, 4>), dim3(g), dim3(kBlock), 0, ctx->stream, ctx->d_bins, ctx->d_bin_count + 0,
                     ctx->d_off, ctx->d_edges, ctx->d_ea, cand, ctx->d_view_count);
  hipLaunchKernelGGL((k_view_count<Word, 16>), dim3(g), dim3(kBlock), 0, ctx->stream, ctx->d_bins + V,
                     ctx->d_bin_count + 1, ctx->d_off, ctx->d_edges, ctx->d_ea, cand, ctx->d_view_count);
  hipLaunchKernelGGL((k_view_count<Word, 32>), dim3(g), dim3(kBlock), 0, ctx->stream, ctx->d_bins + 2 * V,
                     ctx->d_bin_count + 2, ctx->d_off, ctx->d_edges, ctx->d_ea, cand, ctx->d_view_count);
  hipLaunchKernelGGL((k_view_count<Word, 64>), dim3(g), dim3(kBlock), 0, ctx->stream, ctx->d_bins + 3 * V,
                     ctx->d_bin_count + 3, ctx->d_off, ctx->d_edges, ctx->d_ea, cand, ctx->d_view_count);
  if (ctx->nhubs)
    hipLaunchKernelGGL(k_view_hubs<Word>, dim3(hub_blocks), dim3(kBlock), 0, ctx->stream, ctx->d_hubs, ctx->nhubs,
                       ctx->d_off, ctx->d_edges, ctx->d_ea, cand, ctx->d_view_count, nullptr, nullptr);
  const uint64_t spans = (V + 1 + kScanSpan - 1) / kScanSpan;
  hipLaunchKernelGGL(k_scan_spans, dim3(spans), dim3(kBlock), 0, ctx->stream, ctx->d_view_count, V + 1,
                     ctx->d_scan_tmp);
  hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(kTotalsBlock), 0, ctx->stream, ctx->d_scan_tmp, spans);
  hipLaunchKernelGGL(k_scan_apply, dim3(spans), dim3(kBlock), 0, ctx->stream, ctx->d_view_count, V + 1,
                     ctx->d_scan_tmp, ctx->d_view_off);
  PJ_HIP(hipGetLastError());
  PJ_HIP(hipMemcpyAsync(ctx->h_word, ctx->d_view_off + V, 8, hipMemcpyDeviceToHost, ctx->stream));
  PJ_HIP(wait_stream(ctx));
  const uint64_t total = *ctx->h_word;
  if (total + 1 > ctx->view_cap) {
    if (ctx->d_view_edges) hipFree(ctx->d_view_edges);
    ctx->d_view_edges = nullptr;
    const uint64_t cap = std::min<uint64_t>(ctx->E + 1, total + total / 4 + 4096);
    if (dev_alloc(ctx, &ctx->d_view_edges, cap)) return -1;
    ctx->view_cap = cap;
  }
  hipLaunchKernelGGL((k_view_fill<Word, 4>), dim3(g), dim3(kBlock), 0, ctx->stream, ctx->d_bins, ctx->d_bin_count + 0,
                     ctx->d_off, ctx->d_edges, ctx->d_ea, cand, ctx->d_view_off, ctx->d_view_edges);
  hipLaunchKernelGGL((k_view_fill<Word, 16>), dim3(g), dim3(kBlock), 0, ctx->stream, ctx->d_bins + V,
                     ctx->d_bin_count + 1, ctx->d_off, ctx->d_edges, ctx->d_ea, cand, ctx->d_view_off,
                     ctx->d_view_edges);
  hipLaunchKernelGGL((k_view_fill<Word, 32>), dim3(g), dim3(kBlock), 0, ctx->stream, ctx->d_bins + 2 * V,
                     ctx->d_bin_count + 2, ctx->d_off, ctx->d_edges, ctx->d_ea, cand, ctx->d_view_off,
                     ctx->d_view_edges);
  hipLaunchKernelGGL((k_view_fill<Word, 64>), dim3(g), dim3(kBlock), 0, ctx->stream, ctx->d_bins + 3 * V,
                     ctx->d_bin_count + 3, ctx->d_off, ctx->d_edges, ctx->d_ea, cand, ctx->d_view_off,
                     ctx->d_view_edges);
  if (ctx->nhubs)
    hipLaunchKernelGGL(k_view_hubs<Word>, dim3(hub_blocks), dim3(kBlock), 0, ctx->stream, ctx->d_hubs, ctx->nhubs,
                       ctx->d_off, ctx->d_edges, ctx->d_ea, cand, ctx->d_view_count, ctx->d_view_off,
                       ctx->d_view_edges);
  end_launch(ctx);
  PJ_HIP(hipGetLastError());
  return 0;
}

// Complete cycle and path walks of 4+ vertices keep a failed-state memo
// (state_fail): several partial walks reach the same (vertex, hop) there.
bool long_walk(const pj_ctx* ctx, int kind) {
  return (kind == 0 || kind == 1) && ctx->h_t.nlcc_len >= 4;
}

// Memoised 3-vertex cycle steps that k_triangles_memo decides exactly:
// graphs with sorted rows, no duplicate slots (implied by the reverse
// positions) and no self-loops.  PJ_TRIANGLE_MEMO=0 runs them as memoised
// walks instead (same results).
bool triangle_memo(const pj_ctx* ctx, int kind) {
  return kind == 0 && ctx->h_t.nlcc_memo && ctx->h_t.nlcc_len == 3 && reverse_positions(ctx) && !ctx->self_loops &&
         ctx->triangle_memo_enabled;
}

template <typename Word>
int launch_nlcc(pj_ctx* ctx, int kind) {
  if (triangle_memo(ctx, kind)) {
    if (prepare_memo(ctx, false)) return -1;
    launch_bins<Word>(ctx, 1);
    for (int b = 0; b < kBins; ++b) {
      unsigned long long* c = begin_launch(ctx, K_TRIANGLES_MEMO);
      hipLaunchKernelGGL(k_triangles_memo<Word>, dim3(ctx->cus * kWalkBlocksPerCu), dim3(kBlock), 0, ctx->stream,
                         ctx->d_bins + b * static_cast<uint64_t>(ctx->V), ctx->d_bin_count + b, ctx->d_off,
                         ctx->d_edges, ctx->d_ea, reverse_positions(ctx), pair_index(ctx),
                         static_cast<const Word*>(ctx->d_cand),
                         static_cast<Word*>(ctx->d_unmatch), ctx->d_flags, static_cast<Word*>(ctx->d_already),
                         ctx->d_t, c);
      end_launch(ctx);
    }
    return 0;
  }
  const bool memo = ctx->h_t.nlcc_memo && kind != 3;
  const bool tables = memo || (!ctx->h_t.nlcc_memo && long_walk(ctx, kind));
  launch_bins<Word>(ctx, kind == 1 && !memo ? 2 : 1);
  if (tables && prepare_memo(ctx)) return -1;
  if (memo) {
    if (build_view<Word>(ctx)) return -1;
    if (kind == 1) launch_bins<Word>(ctx, 2);  // path sources only
  }
  const MemoArgs ma = tables ? ctx->memo1 : MemoArgs{};
  // memoised walks run on the active view (every slot active: no bitset)
  const uint64_t* off = memo ? ctx->d_view_off : ctx->d_off;
  const uint32_t* edges = memo ? ctx->d_view_edges : ctx->d_edges;
  const uint32_t* ea = memo ? nullptr : ctx->d_ea;
  const Word* cand = static_cast<const Word*>(ctx->d_cand);
  Word* unm = static_cast<Word*>(ctx->d_unmatch);
  const uint64_t V = ctx->V;
  // memo tables are indexed by wave: those launches have exactly kMemoWaves
  const int g = tables ? kMemoWaves / (kBlock / 64) : ctx->cus * kWalkBlocksPerCu;
  // 3-vertex cycles from low-degree sources: several sources per wave
  const uint32_t* rev = reverse_positions(ctx);
  const PairIndex pi = memo ? PairIndex{nullptr, nullptr, 0} : pair_index(ctx);  // positions refer to the CSR
  const bool triangles = kind == 0 && !tables && rev && ctx->h_t.nlcc_len == 3;
  // The 3-vertex cycle walks of the bins in side_walks (a bit per bin;
  // default bins 2 and 3, sources of 33-2048 slots) run on the side stream,
  // beside the other bins' walks: each walk writes only its own source's
  // unmatched candidates and reads state that no walk of the step writes,
  // and the walks wait on chains of dependent loads, so the chip has room
  // for both.  A launch timed with events stays on the context stream.
  auto kid_of = [&](int b) {
    return triangles && b <= 1                         ? K_TRIANGLES
           : triangles && b == 2                         ? K_TRIANGLES_MID
           : triangles && b == 3 && ctx->wide_triangles ? K_TRIANGLES_WIDE
           : memo                                       ? K_MEMO_WALK
                                                        : K_NLCC;
  };
  // The side stream is created on first use: streams are spread over the
  // hardware queues in the order they are created, so contexts that never
  // use one (the query forms' workers) must not take a queue's share.
  if (triangles && ctx->side_walks && !ctx->side &&
      (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
       hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
       hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess)) {
    ctx->err = "launch_nlcc: creating the side stream failed";
    return -1;
  }
  auto on_side = [&](int b) {
    return triangles && ctx->side && ((ctx->side_walks >> b) & 1u) && (b != 3 || ctx->wide_triangles) &&
           !(ctx->timing && (ctx->timing_kid < 0 || ctx->timing_kid == kid_of(b)));
  };
  auto launch_bin = [&](int b, hipStream_t st) {
    unsigned long long* c = begin_launch(ctx, kid_of(b));
    if (triangles && b == 0)
      hipLaunchKernelGGL((k_triangles<Word, kWalkBin0>), dim3(g), dim3(kBlock), 0, st, ctx->d_bins + b * V,
                         ctx->d_bin_count + b, off, edges, ea, rev, pi, cand, unm, ctx->d_flags, ctx->d_t, c);
    else if (triangles && b == 1)
      hipLaunchKernelGGL((k_triangles<Word, 32>), dim3(g), dim3(kBlock), 0, st, ctx->d_bins + b * V,
                         ctx->d_bin_count + b, off, edges, ea, rev, pi, cand, unm, ctx->d_flags, ctx->d_t, c);
    else if (triangles && b == 2)
      hipLaunchKernelGGL(k_triangles_mid<Word>, dim3(g), dim3(kBlock), 0, st, ctx->d_bins + b * V,
                         ctx->d_bin_count + b, off, edges, ea, rev, pi, cand, unm, ctx->d_flags, ctx->d_t, c);
    else if (triangles && b == 3 && ctx->wide_triangles)
      hipLaunchKernelGGL(k_triangles_wide<Word>, dim3(g), dim3(kBlock), 0, st, ctx->d_bins + b * V,
                         ctx->d_bin_count + b, off, edges, ea, rev, pi, cand, unm, ctx->d_flags, ctx->d_t, c);
    else if (memo)  // first pass with LDS tables; overflows go to the list of d_ovf[0]
      hipLaunchKernelGGL(k_memo_lds<Word>, dim3(kMemoWaves), dim3(64), 0, st, ctx->d_bins + b * V,
                         ctx->d_bin_count + b, off, edges, cand, unm, ctx->d_flags, ctx->d_t, ma, c);
    else  // bin 3 (with bin 2 unless k_triangles_mid took it)
      hipLaunchKernelGGL((tables ? k_nlcc<Word, false> : k_nlcc<Word, true>), dim3(g), dim3(kBlock), 0, st,
                         kind, ctx->d_bins + b * V,
                         ctx->d_bin_count + b, b == 3 && !triangles ? ctx->d_bins + 2 * V : nullptr,
                         b == 3 && !triangles ? ctx->d_bin_count + 2 : nullptr, off, edges, ea,
                         reverse_positions(ctx), pi, cand, unm, ctx->d_flags, ctx->d_t, ma, c);
    end_launch(ctx);
  };
  bool forked = false;
  for (int b = kBins - 1; b >= 0; --b) {  // the widest sources first
    if (!on_side(b)) continue;
    if (!forked) {  // the side stream starts after the bins were built
      PJ_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
      PJ_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
      forked = true;
    }
    launch_bin(b, ctx->side);
  }
  if (forked) {
    PJ_HIP(hipEventRecord(ctx->ev_join, ctx->side));
    ctx->side_pending = true;
  }
  // Every bin holds active vertices; hubs are run from their own list.
  for (int b = 0; b < kBins; ++b) {
    if (b == 2 && !triangles && !memo) continue;  // taken by the k_nlcc launch of bin 3
    if (on_side(b)) continue;
    launch_bin(b, ctx->stream);
  }
  return 0;
}

// Retry tables (2.7, 5.4 and 10.7 GB) serve the rare sources whose walk
// outgrows the first-pass tables.  One set per device, shared by every
// context of the process and allocated on first use: a context takes the
// lock for its retry passes and leaves them with its stream drained, so the
// passes of different contexts never overlap and each holds the per-wave
// epochs to itself.  The set lives until the process ends.
constexpr int kMaxDevices = 64;
struct RetryPool {
  std::mutex mu;
  MemoArgs tables[kRetryPasses]{};
};
RetryPool g_retry_pool[kMaxDevices];


// Sources whose walks outgrew their tables run again with larger ones
// (kRetryWaves / kRetryLog2), the overflow lists alternating between the two
// halves of d_ovf.  A source's results are written only by the pass that
// completes it.
template <typename Word>
int launch_nlcc_retry(pj_ctx* ctx, int kind) {
  const uint64_t V = ctx->V;
  if (ctx->device < 0 || ctx->device >= kMaxDevices) {
    ctx->err = "device index beyond the retry table pool";
    return -1;
  }
  RetryPool& pool = g_retry_pool[ctx->device];
  std::unique_lock<std::mutex> pool_lock(pool.mu, std::defer_lock);  // held from the first retry pass
  // pass 0: the first-pass HBM tables (memo1, kMemoWaves waves); then the
  // shared retry tables
  for (int r = 0; r <= kRetryPasses; ++r) {
    uint32_t* in_count = ctx->d_ovf + (r & 1);
    uint32_t* out_count = ctx->d_ovf + ((r + 1) & 1);
    uint32_t* in_list = ctx->d_ovf + 2 + (r & 1) * V;
    uint32_t* out_list = ctx->d_ovf + 2 + ((r + 1) & 1) * V;
    PJ_HIP(hipMemcpyAsync(ctx->h_word, in_count, 4, hipMemcpyDeviceToHost, ctx->stream));
    PJ_HIP(wait_stream(ctx));
    const uint32_t n = *reinterpret_cast<const uint32_t*>(ctx->h_word);
    if (!n) return 0;
    MemoArgs m = ctx->memo1;
    int waves = kMemoWaves;
    if (r > 0) {
      if (!pool_lock.owns_lock()) pool_lock.lock();
      MemoArgs& shared = pool.tables[r - 1];
      if (alloc_memo(ctx, &shared, kRetryWaves[r - 1], kRetryLog2[r - 1])) return -1;
      m = shared;
      m.already = ctx->d_already;  // this context's alreadyMatched
      m.lds_limit = ctx->memo_lds_limit;
      waves = kRetryWaves[r - 1];
    }
    m.ovf_count = out_count;
    m.ovf_list = out_list;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, out_count, static_cast<uint64_t>(1), 0u);
    unsigned long long* c = begin_launch(ctx, K_NLCC);
    hipLaunchKernelGGL((k_nlcc<Word, false>), dim3(waves), dim3(64), 0, ctx->stream, kind, in_list, in_count, nullptr,
                       nullptr, ctx->d_view_off, ctx->d_view_edges, nullptr, nullptr, PairIndex{nullptr, nullptr, 0},
                       static_cast<const Word*>(ctx->d_cand),
                       static_cast<Word*>(ctx->d_unmatch), ctx->d_flags, ctx->d_t, m, c);
    end_launch(ctx);
  }
  PJ_HIP(hipMemcpyAsync(ctx->h_word, ctx->d_ovf + ((kRetryPasses + 1) & 1), 4, hipMemcpyDeviceToHost, ctx->stream));
  PJ_HIP(wait_stream(ctx));
  const uint32_t n = *reinterpret_cast<const uint32_t*>(ctx->h_word);
  if (n) {
    ctx->err = "memoised walk of a source exceeds 2^29 memo entries";  // after the last retry table
    return -1;
  }
  return 0;
}

// Hubs are not in the bins; the NLCC task list for them is the hub list with
// a count.  Hubs that are inactive (or of the wrong label for PC) are skipped
// by the kernels' candidate tests, so the whole list is passed.
template <typename Word>
void launch_nlcc_hubs(pj_ctx* ctx, int kind, uint32_t* d_hub_n) {
  if (!ctx->nhubs) return;
  const Word* cand = static_cast<const Word*>(ctx->d_cand);
  Word* unm = static_cast<Word*>(ctx->d_unmatch);
  const bool memo = ctx->h_t.nlcc_memo && kind != 3;
  const bool tables = memo || (!ctx->h_t.nlcc_memo && long_walk(ctx, kind));
  const MemoArgs ma = tables ? ctx->memo1 : MemoArgs{};
  const uint32_t g = std::min<uint32_t>(ctx->nhubs, tables ? kMemoWaves : ctx->cus * 8);
  unsigned long long* c = begin_launch(ctx, K_NLCC_HUB);
  if (triangle_memo(ctx, kind))
    hipLaunchKernelGGL(k_triangles_memo<Word>, dim3(std::min<uint32_t>(ctx->nhubs, ctx->cus * 8)), dim3(64), 0,
                       ctx->stream, ctx->d_hubs, d_hub_n, ctx->d_off, ctx->d_edges, ctx->d_ea, reverse_positions(ctx),
                       pair_index(ctx), cand, unm, ctx->d_flags, static_cast<Word*>(ctx->d_already), ctx->d_t, c);
  else if (memo)
    hipLaunchKernelGGL(k_memo_lds<Word>, dim3(g), dim3(64), 0, ctx->stream, ctx->d_hubs, d_hub_n, ctx->d_view_off,
                       ctx->d_view_edges, cand, unm, ctx->d_flags, ctx->d_t, ma, c);
  else
    hipLaunchKernelGGL((tables ? k_nlcc<Word, false> : k_nlcc<Word, true>), dim3(g), dim3(64), 0, ctx->stream, kind,
                       ctx->d_hubs, d_hub_n, nullptr, nullptr,
                       ctx->d_off,
                       ctx->d_edges, ctx->d_ea, reverse_positions(ctx), pair_index(ctx), cand, unm, ctx->d_flags,
                       ctx->d_t, ma, c);
  end_launch(ctx);
}

// Whether the reference's edge memo can fire on a search whose memo levels
// are `tr` (position -> next position): it inserts template ids and tests
// positions (cc_cpu.tpp:53-61, pc_cpu.tpp:59-67), so a test hits only when a
// template id equal to the tested position was inserted for the same
// undirected edge, which needs matching label codes at both ends.
bool memo_may_fire(const std::vector<uint32_t>& idx, const uint8_t* pcode,
                   const std::vector<std::pair<size_t, size_t>>& tr) {
  for (const auto& cur : tr)
    for (const auto& prev : tr) {
      if (idx[prev.second] != cur.first) continue;
      const uint8_t lc = pcode[idx[cur.first]], lu = pcode[idx[cur.second]];
      const uint8_t pf = pcode[idx[prev.first]], pt = pcode[idx[prev.second]];
      if ((pf == lc && pt == lu) || (pf == lu && pt == lc)) return true;
    }
  return false;
}

// Appends one memoised-walk plan to the walk tables.
struct PlanMove {
  uint8_t move;
  uint32_t vertex, mb, hidx;
  uint8_t test, mark;
};

int add_plan(pj_ctx* ctx, uint32_t root, const std::vector<PlanMove>& mv, uint32_t* o) {
  DevTemplate& t = ctx->h_t;
  if (t.walk_count >= 64 || mv.size() > 64 || *o + mv.size() > 64 * 128) {
    ctx->err = "memoised walk plan exceeds the device tables (64 walks of up to 64 moves)";
    return -1;
  }
  const uint32_t k = t.walk_count++;
  t.walk_root[k] = root;
  t.walk_off[k] = *o;
  for (size_t i = 0; i < mv.size(); ++i) {
    t.walk_move[*o + i] = mv[i].move;
    t.walk_vertex[*o + i] = mv[i].vertex;
    t.walk_mb[*o + i] = mv[i].mb;
    t.walk_hidx[*o + i] = mv[i].hidx;
    t.walk_test[*o + i] = mv[i].test;
    t.walk_mark[*o + i] = mv[i].mark;
  }
  *o += static_cast<uint32_t>(mv.size());
  t.walk_off[k + 1] = *o;
  return 0;
}

int install_nlcc(pj_ctx* ctx, int step, uint32_t index) {
  DevTemplate& t = ctx->h_t;
  t.rows_sorted = ctx->rows_sorted ? 1u : 0u;
  t.nlcc_memo = 0;
  t.walk_hyp = 0;
  const bool force_memo = ctx->memo_walks;
  if (step == PJ_STEP_CC) {
    if (index >= ctx->cycles.size()) {
      ctx->err = "cycle constraint index out of range";
      return -1;
    }
    const auto& c = ctx->cycles[index];
    const size_t L = c.size();
    t.nlcc_len = static_cast<uint32_t>(L);
    for (size_t i = 0; i < L; ++i) t.nlcc_vertex[i] = c[i];
    bool fire = false;
    for (size_t j = 0; j < L && !fire; ++j) {
      std::vector<std::pair<size_t, size_t>> tr;
      for (size_t d = 0; d + 1 < L; ++d) tr.emplace_back((j + d) % L, (j + d + 1) % L);
      fire = memo_may_fire(c, t.pcode, tr);
    }
    if (fire || force_memo) {
      // start j: hops to idx[(j+p)%L] testing position (j+p-1)%L, marking the
      // intermediates (cc_cpu.tpp:24-68), then a slot back to the source.
      t.nlcc_memo = 1;
      t.walk_count = 0;
      uint32_t o = 0;
      for (size_t j = 0; j < L; ++j) {
        std::vector<PlanMove> mv;
        mv.push_back({PJ_MOVE_VERTEX_NO_STORE, c[j], 0, 0, 0, 0});
        for (size_t p = 1; p < L; ++p)
          mv.push_back({PJ_MOVE_VERTEX_NO_STORE, c[(j + p) % L], 0, static_cast<uint32_t>(p),
                        static_cast<uint8_t>((j + p - 1) % L), 1});
        mv.push_back({kMoveClose, c[j], 0, 0, 0, 0});
        if (add_plan(ctx, c[j], mv, &o)) return -1;
      }
    }
  } else if (step == PJ_STEP_PC) {
    if (index >= ctx->paths.size()) {
      ctx->err = "path constraint index out of range";
      return -1;
    }
    const auto& p = ctx->paths[index];
    const size_t L = p.size();
    t.nlcc_len = static_cast<uint32_t>(L);
    for (size_t i = 0; i < L; ++i) t.nlcc_vertex[i] = p[i];
    t.path_code = kNoCode;
    for (uint32_t k = 0; k < t.K; ++k)
      if (t.label_raw[k] == ctx->path_label[index]) t.path_code = k;
    std::vector<std::pair<size_t, size_t>> fw, bw;
    for (size_t d = 0; d + 2 < L; ++d) fw.emplace_back(d, d + 1);
    for (size_t d = 0; d + 2 < L; ++d) bw.emplace_back(L - 1 - d, L - 2 - d);
    if (memo_may_fire(p, t.pcode, fw) || memo_may_fire(p, t.pcode, bw) || force_memo) {
      // forward from idx[0] and backward from idx[L-1]; hop h tests position
      // h-1 (forward) or L-h (backward); the last hop needs a candidate other
      // than the source and marks it (pc_cpu.tpp:11-75)
      t.nlcc_memo = 1;
      t.walk_count = 0;
      uint32_t o = 0;
      for (int r = 0; r < 2; ++r) {
        std::vector<PlanMove> mv;
        mv.push_back({PJ_MOVE_VERTEX_NO_STORE, r ? p[L - 1] : p[0], 0, 0, 0, 0});
        for (size_t h = 1; h + 1 < L; ++h) {
          const size_t np = r ? L - 1 - h : h;
          const size_t cp = r ? L - h : h - 1;
          mv.push_back({PJ_MOVE_VERTEX_NO_STORE, p[np], 0, static_cast<uint32_t>(h), static_cast<uint8_t>(cp), 0});
        }
        mv.push_back({kMoveLast, r ? p[0] : p[L - 1], 0, 0, 0, 1});
        if (add_plan(ctx, r ? p[L - 1] : p[0], mv, &o)) return -1;
      }
    }
  } else {
    const std::vector<pj_ctx::HostWalk>* walks;
    std::vector<pj_ctx::HostWalk> single;
    if (step == PJ_STEP_TDS) {
      if (index >= ctx->tds.size()) {
        ctx->err = "TDS constraint index out of range";
        return -1;
      }
      walks = &ctx->tds[index];
    } else {
      single.push_back(ctx->enumeration);
      walks = &single;
    }
    if (walks->size() > 64) {
      ctx->err = "more than 64 walks in a TDS constraint";
      return -1;
    }
    t.walk_count = static_cast<uint32_t>(walks->size());
    uint32_t o = 0;
    for (size_t k = 0; k < walks->size(); ++k) {
      const auto& w = (*walks)[k];
      if (w.move.size() > 64) {
        ctx->err = "walks of more than 64 moves are not supported on the device";
        return -1;
      }
      t.walk_root[k] = w.root;
      t.walk_off[k] = o;
      uint32_t h = 0;
      for (size_t i = 0; i < w.move.size(); ++i) {
        t.walk_move[o + i] = w.move[i];
        t.walk_vertex[o + i] = w.vertex[i];
        t.walk_mb[o + i] = w.mb[i];
        const bool vtx = w.move[i] == PJ_MOVE_VERTEX_STORE || w.move[i] == PJ_MOVE_VERTEX_NO_STORE;
        t.walk_hidx[o + i] = vtx ? h : 0;
        // TDS memos: the edge memo tests the previous move's template vertex
        // and every VERTEX move marks its vertex (tds_cpu.tpp:31-82)
        t.walk_test[o + i] = static_cast<uint8_t>(i ? w.vertex[i - 1] : 0);
        t.walk_mark[o + i] = vtx && i ? 1 : 0;
        if (vtx) ++h;
      }
      o += static_cast<uint32_t>(w.move.size());
    }
    t.walk_off[walks->size()] = o;
    if (step == PJ_STEP_TDS) {
      t.nlcc_memo = 1;
      t.walk_hyp = 1;
    }
  }
  if (t.nlcc_memo && ctx->V >= (1u << 30)) {
    ctx->err = "memoised walks need vertex ids below 2^30";
    return -1;
  }
  PJ_HIP(hipMemcpyAsync(ctx->d_t, &t, sizeof(DevTemplate), hipMemcpyHostToDevice, ctx->stream));
  return 0;
}

template <typename Word>
int run_step_t(pj_ctx* ctx, int step, uint32_t index) {
  zero_counters(ctx);
  switch (step) {
    case PJ_STEP_LCC0: {
      unsigned long long* c = begin_launch(ctx, K_LCC0);
      hipLaunchKernelGGL(k_lcc0<Word>, dim3(grid_for(ctx, ctx->V)), dim3(kBlock), 0, ctx->stream,
                         static_cast<Word*>(ctx->d_cand), ctx->d_code, ctx->V, ctx->d_t, c);
      end_launch(ctx);
      mark_all_dirty(ctx);  // every candidate set was rewritten
      ctx->labels_consistent = true;
      ctx->fresh = ctx->fresh == 0 ? 1 : 2;
      break;
    }
    case PJ_STEP_LCC:
      launch_lcc<Word>(ctx);
      ctx->fresh = 2;
      break;
    case PJ_STEP_CC:
    case PJ_STEP_PC:
    case PJ_STEP_TDS:
    case PJ_STEP_ENUMERATION: {
      const int kind = step == PJ_STEP_CC ? 0 : step == PJ_STEP_PC ? 1 : step == PJ_STEP_TDS ? 2 : 3;
      ctx->fresh = 2;
      if (launch_nlcc<Word>(ctx, kind)) return -1;
      launch_nlcc_hubs<Word>(ctx, kind, ctx->d_nhubs);
      if (ctx->side_pending) {  // the side stream's walks are done before the step goes on
        PJ_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
        ctx->side_pending = false;
      }
      if (ctx->h_t.nlcc_memo && kind != 3 && !triangle_memo(ctx, kind) && launch_nlcc_retry<Word>(ctx, kind))
        return -1;
      if (step != PJ_STEP_ENUMERATION) {
        if (ctx->skip_clean && ctx->modified_cleared) {
          // Only the failing sources are modified: they mark their rows and
          // every unmarked vertex has no modified neighbour.
          launch_phase2<Word>(ctx, push_modified<Word>(ctx));
        } else {
          mark_all_rescan(ctx);  // sources that failed turned modified
          launch_phase2<Word>(ctx);
        }
      }
      break;
    }
    default:
      ctx->err = "unknown step";
      return -1;
  }
  PJ_HIP(hipGetLastError());
  return 0;
}


template <typename Word>
int compact_live_t(pj_ctx* ctx) {
  uint32_t* out = ctx->live == ctx->d_live[0] ? ctx->d_live[1] : ctx->d_live[0];
  uint32_t* n_out = out + ctx->V;  // count in the spare last entry
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, n_out, static_cast<uint64_t>(1), 0u);
  const uint32_t blocks = static_cast<uint32_t>((static_cast<uint64_t>(ctx->n_live) + kBinVerts - 1) / kBinVerts);
  hipLaunchKernelGGL(k_live_compact<Word>, dim3(std::max<uint32_t>(blocks, 1)), dim3(kBlock), 0, ctx->stream,
                     ctx->live, ctx->n_live, static_cast<const Word*>(ctx->d_cand), out, n_out);
  PJ_HIP(hipGetLastError());
  PJ_HIP(hipMemcpyAsync(ctx->h_word, n_out, 4, hipMemcpyDeviceToHost, ctx->stream));
  PJ_HIP(wait_stream(ctx));
  ctx->live = out;
  ctx->n_live = *reinterpret_cast<const uint32_t*>(ctx->h_word);
  return 0;
}

template <typename Word>
int download_t(pj_ctx* ctx, uint64_t* cand, uint8_t* vflags, uint32_t* ea_words) {
  hipLaunchKernelGGL(k_cand_to_u64<Word>, dim3(grid_for(ctx, ctx->V)), dim3(kBlock), 0, ctx->stream,
                     static_cast<const Word*>(ctx->d_cand), static_cast<uint64_t*>(ctx->d_u64), ctx->d_flags,
                     ctx->d_vflags, ctx->V);
  PJ_HIP(hipGetLastError());
  if (cand) PJ_HIP(hipMemcpyAsync(cand, ctx->d_u64, ctx->V * 8ull, hipMemcpyDeviceToHost, ctx->stream));
  if (vflags) PJ_HIP(hipMemcpyAsync(vflags, ctx->d_vflags, ctx->V, hipMemcpyDeviceToHost, ctx->stream));
  if (ea_words) PJ_HIP(hipMemcpyAsync(ea_words, ctx->d_ea, ctx->ea_words * 4, hipMemcpyDeviceToHost, ctx->stream));
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

template <typename Word>
int upload_state_t(pj_ctx* ctx, const uint64_t* cand, const uint8_t* vflags, const uint32_t* ea_words) {
  PJ_HIP(hipMemcpyAsync(ctx->d_u64, cand, ctx->V * 8ull, hipMemcpyHostToDevice, ctx->stream));
  PJ_HIP(hipMemcpyAsync(ctx->d_vflags, vflags, ctx->V, hipMemcpyHostToDevice, ctx->stream));
  PJ_HIP(hipMemcpyAsync(ctx->d_ea, ea_words, ctx->ea_words * 4, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_u64_to_cand<Word>, dim3(grid_for(ctx, ctx->V)), dim3(kBlock), 0, ctx->stream,
                     static_cast<const uint64_t*>(ctx->d_u64), ctx->d_vflags, static_cast<Word*>(ctx->d_cand),
                     static_cast<Word*>(ctx->d_unmatch), ctx->d_flags, ctx->V);
  mark_all_dirty(ctx);
  ctx->labels_consistent = false;  // an uploaded state may hold candidates of other labels
  ctx->fresh = 2;
  ctx->modified_cleared = false;
  ctx->live = nullptr;              // and any vertex may be active
  ctx->n_live = ctx->V;
  PJ_HIP(hipGetLastError());
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

// Order-dependent checksum of a state: sum over vertices of a hash of (v,
// candidate word, active / modified / scheduled) plus, over edge-active
// words, a hash of (index, word).  Equal states give equal sums; a single
// changed bit changes the sum (up to 64-bit hash collisions).
template <typename Word>
__global__ void k_state_checksum(const Word* __restrict__ cand, const uint8_t* __restrict__ flags,
                                 const uint32_t* __restrict__ ea, uint32_t V, uint64_t words,
                                 unsigned long long* out) {
  uint64_t acc = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += stride) {
    const uint64_t c = static_cast<uint64_t>(cand[v]);
    const uint64_t f = (flags[v] & (kModified | kScheduled)) | (c ? 1u : 0u);
    acc += mix64(mix64(v * 0x9E3779B97F4A7C15ull ^ c) + f);
  }
  for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < words; w += stride)
    acc += mix64((w + 0xD1B54A32D192ED03ull) * 0x94D049BB133111EBull ^ ea[w]);
  // wave sums, then one atomic per block (sums mod 2^64 do not depend on the
  // order; one atomic per wave on a single address serialised ~16 K of them)
  __shared__ uint64_t s_acc[kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane_id() == 0) s_acc[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t b = 0;
    for (int w = 0; w < kBlock / 64; ++w) b += s_acc[w];
    atomicAdd(out, static_cast<unsigned long long>(b));
  }
}

template <typename Word>
int checksum_t(pj_ctx* ctx, uint64_t* out) {
  unsigned long long* d = ctx->d_sums;  // free outside pj_run_step
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, d, static_cast<uint64_t>(1), 0ull);
  hipLaunchKernelGGL(k_state_checksum<Word>, dim3(grid_for(ctx, ctx->V)), dim3(kBlock), 0, ctx->stream,
                     static_cast<const Word*>(ctx->d_cand), ctx->d_flags, ctx->d_ea, ctx->V, ctx->ea_words, d);
  PJ_HIP(hipGetLastError());
  unsigned long long h = 0;
  PJ_HIP(hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, ctx->stream));
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  *out = h;
  return 0;
}

// Candidate words are the narrowest unsigned type holding one bit per
// template vertex (1, 2, 4 or 8 bytes), so the per-slot gathers of
// neighbour candidates touch as few cache lines as possible.
template <typename F>
int with_word(const pj_ctx* ctx, F&& f) {
  switch (ctx->word_bytes) {
    case 1:
      return f(uint8_t{});
    case 2:
      return f(uint16_t{});
    case 4:
      return f(uint32_t{});
    default:
      return f(uint64_t{});
  }
}

}  // namespace

namespace {

// Label statistics (pjs::label_stat): the pair table from the template's
// local constraints, the bins of every active vertex, one pass, and the
// counts turned into the reference's maps.
template <typename Word>
int graph_stat_t(pj_ctx* ctx, pj_label_stat* out) {
  const DevTemplate& t = ctx->h_t;
  const uint32_t K = t.K;
  // accumulationMaximumMap (graph_stat_cpu.tpp:14-38): per template vertex
  // the count of its neighbours of each label, maximised over the vertices
  // of each label; the local constraints hold those counts.
  std::vector<uint32_t> mx(pjs::kLabels * pjs::kLabels, 0), tab(pjs::kLabels * pjs::kLabels, 0);
  for (uint32_t p = 0; p < t.P; ++p)
    for (uint32_t k = 0; k < K; ++k) {
      const uint32_t need = static_cast<uint32_t>((t.need[p][k >> 3] >> (8 * (k & 7))) & 0xFFu);
      uint32_t& m = mx[t.pcode[p] * pjs::kLabels + k];
      m = std::max(m, need);
    }
  uint32_t words = 0;
  for (uint32_t f = 0; f < K; ++f)
    for (uint32_t l = 0; l < K; ++l) {
      const uint32_t m = mx[f * pjs::kLabels + l];
      if (!m) continue;
      tab[f * pjs::kLabels + l] = words << 8 | m;
      words += m;
    }
  if (words > pjs::kExactMax) {
    ctx->err = "pj_graph_stat: template neighbour counts exceed the statistics table";
    return -1;
  }
  if (!ctx->d_stat && dev_alloc(ctx, &ctx->d_stat, pjs::kBufferBytes / 8)) return -1;
  PJ_HIP(hipMemcpyAsync(ctx->d_stat + pjs::kCountWords, tab.data(), tab.size() * 4, hipMemcpyHostToDevice,
                        ctx->stream));
  PJ_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  launch_bins<Word>(ctx, 1);
  pjs::StatArgs a{};
  a.off = ctx->d_off;
  a.edges = ctx->d_edges;
  a.ea = ctx->d_ea;
  a.cand = ctx->d_cand;
  a.word_bytes = static_cast<int>(sizeof(Word));
  a.code = ctx->d_code;
  a.V = ctx->V;
  a.bins = ctx->d_bins;
  a.bin_count = ctx->d_bin_count;
  a.hubs = ctx->d_hubs;
  a.nhubs = ctx->nhubs;
  a.chunk_hub = ctx->d_chunk_hub;
  a.chunk_start = ctx->d_chunk_start;
  a.nchunks = ctx->nhubs ? ctx->nchunks : 0u;
  a.hub_count = ctx->d_hub_count;
  a.pcode = reinterpret_cast<const uint8_t*>(ctx->d_t) + offsetof(DevTemplate, pcode);  // device address
  a.derive = ctx->labels_consistent;
  a.K = K;
  a.exact_words = words;
  a.buffer = ctx->d_stat;
  a.cus = ctx->cus;
  a.stream = ctx->stream;
  begin_launch(ctx, K_LABEL_STAT);
  if (pjs::label_stat(a, &ctx->err)) return -1;
  end_launch(ctx);
  PJ_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  std::vector<unsigned long long> counts(pjs::kCountWords);
  PJ_HIP(hipMemcpyAsync(counts.data(), ctx->d_stat, counts.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  PJ_HIP(hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1));
  ctx->last_slots = 0;
  for (int b = 0; b < ctx->nlaunch; ++b) {  // kernel timing: the bins and the pass
    float ms = 0.f;
    PJ_HIP(hipEventElapsedTime(&ms, ctx->lev[2 * b], ctx->lev[2 * b + 1]));
    ctx->k_launches[ctx->launch_kid[b]] += 1;
    ctx->k_ms[ctx->launch_kid[b]] += ms;
  }
  std::memset(out, 0, sizeof(*out));
  out->label_count = K;
  out->at_least_count = words;
  for (uint32_t f = 0; f < K; ++f) {
    out->label[f] = t.label_raw[f];
    out->vertex[f] = counts[pjs::kVertexAt + f];
    out->vertex_total += out->vertex[f];
    for (uint32_t l = 0; l < K; ++l) {
      const uint32_t i = f * pjs::kLabels + l;
      out->edge[i] = counts[pjs::kEdgeAt + i];
      out->outbound[f] += out->edge[i];
      const uint32_t m = tab[i] & 0xFFu, at = tab[i] >> 8;
      out->at_least_max[i] = m;
      out->at_least_offset[i] = at;
      // at least k = the vertices whose capped tally is k..m
      uint64_t run = 0;
      for (uint32_t k = m; k >= 1; --k) {
        run += counts[pjs::kExactAt + at + k - 1];
        out->at_least[at + k - 1] = run;
      }
    }
    out->edge_total += out->outbound[f];
  }
  return 0;
}

}  // namespace

extern "C" {

int pj_create(pj_ctx** out, int device) {
  if (!out) return -1;
  *out = nullptr;
  auto* ctx = new pj_ctx();
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&ctx->d_t), sizeof(DevTemplate));
  if (e == hipSuccess)  // coherent: the sums kernel's stores are visible once the stream is done
    e = hipHostMalloc(reinterpret_cast<void**>(&ctx->h_sums), (kMaxLaunches + 1) * kCounterStride * 8,
                      hipHostMallocCoherent);
  if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(&ctx->h_sums_dev), ctx->h_sums, 0);
  if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(&ctx->h_word), 64, hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void**>(&ctx->d_sums), (kMaxLaunches + 1) * kCounterStride * 8);
  if (e != hipSuccess) {
    fprintf(stderr, "pj_create: %s\n", hipGetErrorString(e));
    delete ctx;
    return -1;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    ctx->cus = cus;
  // PJ_SKIP_CLEAN=0 evaluates every scheduled vertex in LCC phase 1 (the
  // results are the same either way; the switch exists for comparisons).
  if (const char* sc = getenv("PJ_SKIP_CLEAN")) ctx->dirty_enabled = strcmp(sc, "0") != 0;
  // PJ_SLOT_LABELS=0 keeps the neighbour-candidate gathers in the first LCC
  // step (same results; the switch exists for comparisons).
  if (const char* sl = getenv("PJ_SLOT_LABELS")) ctx->slot_labels_enabled = strcmp(sl, "0") != 0;
  // PJ_WIDE_TRIANGLES=0 walks the 3-vertex cycles of sources of 257-2048
  // slots (first hop by first hop) instead of testing their pairs.
  if (const char* wt = getenv("PJ_WIDE_TRIANGLES")) ctx->wide_triangles = strcmp(wt, "0") != 0;
  // PJ_PHASE1_FUSED=0 runs the hub chunks of LCC phase 1 in a launch after
  // the degree bins instead of in the same launch (same results).
  if (const char* pf = getenv("PJ_PHASE1_FUSED")) ctx->phase1_fused = strcmp(pf, "0") != 0;
  // PJ_PHASE1_BLOCKS: blocks per compute unit of the LCC phase-1 launch (1-64)
  if (const char* pb = getenv("PJ_PHASE1_BLOCKS")) ctx->phase1_blocks = std::min(64, std::max(1, atoi(pb)));
  // PJ_SIDE_WALKS: a bit per degree bin whose 3-vertex cycle walks run on
  // the side stream beside the others (default 12: the bins of 33-256 and
  // 257-2048 slots; 0: every walk on the context stream, one after another).
  if (const char* sw = getenv("PJ_SIDE_WALKS")) ctx->side_walks = static_cast<uint32_t>(strtoul(sw, nullptr, 0)) & 0xFu;
  // PJ_PAIR_INDEX_MB bounds the pair index of heavy rows (0: none, every
  // pair search by binary search); PJ_PAIR_DEGREE is the smallest row length
  // above which a row is heavy (doubled until the index fits).  Same results
  // either way.
  if (const char* pb = getenv("PJ_PAIR_INDEX_MB")) ctx->pair_budget = static_cast<uint64_t>(std::max(0, atoi(pb))) << 20;
  if (const char* pd = getenv("PJ_PAIR_DEGREE")) ctx->pair_min_degree = static_cast<uint64_t>(std::max(1, atoi(pd)));
  // PJ_MEMO_WALKS=1 runs cycle and path steps as the reference's memoised
  // search even where the complete search is known to give the same result.
  if (const char* mw = getenv("PJ_MEMO_WALKS")) ctx->memo_walks = strcmp(mw, "0") != 0;
  // PJ_MEMO_LOG2 sizes the first-pass memo tables (small values send most
  // sources through the retry passes; the results are the same).
  if (const char* ml = getenv("PJ_MEMO_LOG2")) ctx->memo_log2 = std::min(20, std::max(6, atoi(ml)));
  if (const char* ll = getenv("PJ_MEMO_LDS_LIMIT")) ctx->memo_lds_limit = static_cast<uint32_t>(std::max(1, atoi(ll)));
  // PJ_REVERSE_SLOTS=0 keeps no reverse positions (4 bytes per slot) and
  // closes cycles by binary searches instead (same results).
  if (const char* rs = getenv("PJ_REVERSE_SLOTS")) ctx->reverse_enabled = strcmp(rs, "0") != 0;
  if (const char* tm = getenv("PJ_TRIANGLE_MEMO")) ctx->triangle_memo_enabled = strcmp(tm, "0") != 0;
  if (const char* lm = getenv("PJ_LIVE_MIN")) ctx->live_min = static_cast<uint32_t>(std::max(0, atoi(lm)));
  *out = ctx;
  return 0;
}

// A graph upload: upload_begin allocates the CSR and the state of V vertices
// and E slots, the caller fills d_off, d_edges and d_labels on ctx->stream,
// and upload_finish checks the CSR (offsets on the host, neighbour ids in
// the row-order pass), builds the hub chunk list and runs the order and
// symmetry checks.
static int upload_begin(pj_ctx* ctx, uint32_t V, uint64_t E) {
  PJ_HIP(hipSetDevice(ctx->device));
  free_graph(ctx);
  ctx->V = V;
  ctx->E = E;
  ctx->ea_words = (E + 31) / 32;
  if (dev_alloc(ctx, &ctx->d_off, V + 1ull) || dev_alloc(ctx, &ctx->d_edges, E) || dev_alloc(ctx, &ctx->d_labels, V) ||
      alloc_state(ctx))
    return -1;
  return 0;
}

static int upload_finish(pj_ctx* ctx, const uint64_t* host_off);

static int upload_common(pj_ctx* ctx, uint32_t V, uint64_t E, const uint64_t* offsets, const uint32_t* edges,
                         const uint64_t* labels, hipMemcpyKind kind) {
  if (upload_begin(ctx, V, E)) return -1;
  PJ_HIP(hipMemcpyAsync(ctx->d_off, offsets, (V + 1ull) * 8, kind, ctx->stream));
  if (E) PJ_HIP(hipMemcpyAsync(ctx->d_edges, edges, E * 4, kind, ctx->stream));
  if (V) PJ_HIP(hipMemcpyAsync(ctx->d_labels, labels, V * 8ull, kind, ctx->stream));
  std::vector<uint64_t> h_off;
  const uint64_t* host_off = offsets;
  if (kind == hipMemcpyDeviceToDevice) {
    h_off.resize(V + 1ull);
    PJ_HIP(hipMemcpyAsync(h_off.data(), offsets, (V + 1ull) * 8, hipMemcpyDeviceToHost, ctx->stream));
    PJ_HIP(hipStreamSynchronize(ctx->stream));
    host_off = h_off.data();
  }
  return upload_finish(ctx, host_off);
}

// Per-slot label indices (d_slot_lab) for graphs of at most 255 distinct
// labels: the labels are indexed on the host, then every slot takes its
// neighbour's index.  Graphs with more labels go without (the first LCC step
// then gathers neighbour candidates as every other step does).
static int slot_labels(pj_ctx* ctx) {
  const uint32_t V = ctx->V;
  const uint64_t E = ctx->E;
  ctx->graph_labels.clear();
  if (!V || !E) return 0;
  std::vector<uint64_t> lab(V);
  PJ_HIP(hipMemcpyAsync(lab.data(), ctx->d_labels, V * 8ull, hipMemcpyDeviceToHost, ctx->stream));
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<uint8_t> idx(V);
  std::vector<uint64_t> keys(512);
  std::vector<int> slot(512, -1);  // open addressing over the distinct labels
  for (uint32_t v = 0; v < V; ++v) {
    const uint64_t l = lab[v];
    uint64_t h = (l * 0x9E3779B97F4A7C15ull) >> 55;
    while (slot[h] >= 0 && keys[h] != l) h = (h + 1) & 511;
    if (slot[h] < 0) {
      if (ctx->graph_labels.size() == 255) {  // too many labels for a byte per slot
        ctx->graph_labels.clear();
        return 0;
      }
      keys[h] = l;
      slot[h] = static_cast<int>(ctx->graph_labels.size());
      ctx->graph_labels.push_back(l);
    }
    idx[v] = static_cast<uint8_t>(slot[h]);
  }
  uint8_t* d_idx = nullptr;
  if (dev_alloc(ctx, &ctx->d_slot_lab, E) || dev_alloc(ctx, &d_idx, V)) return -1;
  PJ_HIP(hipMemcpyAsync(d_idx, idx.data(), V, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_slot_labels, dim3(grid_for(ctx, E)), dim3(kBlock), 0, ctx->stream, ctx->d_edges, E, d_idx,
                     ctx->d_slot_lab);
  PJ_HIP(hipGetLastError());
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  PJ_HIP(hipFree(d_idx));
  return 0;
}

// The pair index (PairIndex) of a symmetric graph with sorted rows and
// reverse positions: the threshold starts at pair_min_degree and doubles
// until the positions of the rows above it fit pair_budget; heavy ids follow
// vertex order.  Fewer than two heavy rows: no index.
static int pair_index_build(pj_ctx* ctx, const uint64_t* host_off) {
  const uint32_t V = ctx->V;
  const uint64_t E = ctx->E;
  if (!reverse_positions(ctx) || !ctx->pair_budget || E > kPairOffMask) return 0;
  // above[b]: rows of more than T0 << b slots (one pass; b < 40 as E < 2^40)
  const uint64_t T0 = ctx->pair_min_degree;
  uint64_t above[kPairShift + 1] = {};
  for (uint64_t v = 0; v < V; ++v) {
    const uint64_t d = host_off[v + 1] - host_off[v];
    if (d <= T0) continue;
    int b = 0;
    while (b < kPairShift && ((d - 1) >> (b + 1)) >= T0) ++b;  // d > T0 << (b + 1), without overflow
    ++above[b];
  }
  for (int b = kPairShift - 1; b >= 0; --b) above[b] += above[b + 1];
  uint64_t T = 0, H = 0;
  for (int b = 0; b <= kPairShift; ++b) {
    H = above[b];
    if (H < 2) return 0;
    if (H * H * 4 <= ctx->pair_budget && H < (1ull << (64 - kPairShift))) {
      T = T0 << b;
      break;
    }
  }
  if (!T) return 0;
  std::vector<uint64_t> doff(static_cast<uint64_t>(V) + 1);
  std::vector<uint32_t> heavy;
  heavy.reserve(H);
  for (uint64_t v = 0; v < V; ++v) {
    doff[v] = host_off[v];
    if (host_off[v + 1] - host_off[v] > T) {
      heavy.push_back(static_cast<uint32_t>(v));
      doff[v] |= static_cast<uint64_t>(heavy.size()) << kPairShift;
    }
  }
  doff[V] = host_off[V];
  uint32_t* d_heavy = nullptr;
  if (dev_alloc(ctx, &ctx->d_pair_doff, static_cast<uint64_t>(V) + 1) || dev_alloc(ctx, &ctx->d_pair_pos, H * H) ||
      dev_alloc(ctx, &d_heavy, H))
    return -1;
  PJ_HIP(hipMemcpyAsync(ctx->d_pair_doff, doff.data(), doff.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  PJ_HIP(hipMemcpyAsync(d_heavy, heavy.data(), H * 4, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, H * H)), dim3(kBlock), 0, ctx->stream, ctx->d_pair_pos, H * H,
                     kNoPos);
  hipLaunchKernelGGL(k_pair_positions, dim3(std::min<uint64_t>((H + 3) / 4, ctx->cus * 16ull)), dim3(kBlock), 0,
                     ctx->stream, d_heavy, H, ctx->d_pair_doff, ctx->d_edges, ctx->d_pair_pos);
  PJ_HIP(hipGetLastError());
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  PJ_HIP(hipFree(d_heavy));
  ctx->pair_heavy = H;
  ctx->pair_degree = T;
  return 0;
}

static int upload_finish(pj_ctx* ctx, const uint64_t* host_off) {
  const uint32_t V = ctx->V;
  const uint64_t E = ctx->E;
  // Every kernel indexes off[v + 1] and cand[edges[e]] unchecked: refuse a
  // CSR whose offsets decrease or leave [0, E] (here) or whose ids are not
  // vertices (bit 2 of the row-order check below).
  auto refuse = [ctx](const std::string& why) {
    hipStreamSynchronize(ctx->stream);  // copies into the buffers may still run
    free_graph(ctx);
    ctx->V = 0;
    ctx->E = 0;
    ctx->ea_words = 0;
    ctx->err = "malformed CSR: " + why;
    return -1;
  };
  if (host_off[0] != 0 || host_off[V] != E) return refuse("offsets must run from 0 to edge_count");
  for (uint64_t v = 0; v < V; ++v)
    if (host_off[v + 1] < host_off[v]) return refuse("offsets decrease at vertex " + std::to_string(v));
  if (setup_hubs(ctx, host_off)) return -1;
  {
    // The edge-active bitset is free until a template is installed; the
    // check's flags go to a set of bin counts.
    ctx->counters_clean = false;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, ctx->d_bin_count, static_cast<uint64_t>(4), 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, ctx->ea_words)), dim3(kBlock), 0, ctx->stream, ctx->d_ea,
                       ctx->ea_words, 0u);
    hipLaunchKernelGGL(k_row_starts, dim3(grid_for(ctx, V)), dim3(kBlock), 0, ctx->stream, ctx->d_off, V, E,
                       ctx->d_ea);
    hipLaunchKernelGGL(k_rows_sorted, dim3(grid_for(ctx, E)), dim3(kBlock), 0, ctx->stream, ctx->d_edges, E, V,
                       ctx->d_ea, ctx->d_bin_count);
    uint32_t unsorted = 0;
    PJ_HIP(hipMemcpyAsync(&unsorted, ctx->d_bin_count, 4, hipMemcpyDeviceToHost, ctx->stream));
    PJ_HIP(hipStreamSynchronize(ctx->stream));
    if (unsorted & 4u) return refuse("a neighbour id is not a vertex");
    ctx->rows_sorted = (unsorted & 1u) == 0;
    ctx->symmetric = false;
    if (unsorted == 0) {
      unsigned long long* f = ctx->d_sums;  // three counters; d_sums is free outside pj_run_step
      hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, f, static_cast<uint64_t>(4), 0ull);
      if (ctx->reverse_enabled && dev_alloc(ctx, &ctx->d_rev, E)) return -1;
      hipLaunchKernelGGL(k_rows_symmetric, dim3(grid_for(ctx, static_cast<uint64_t>(V) * 64)), dim3(kBlock), 0,
                         ctx->stream, ctx->d_off, ctx->d_edges, V, f, ctx->d_rev);
      if (ctx->nchunks)
        hipLaunchKernelGGL(k_rows_symmetric_chunks, dim3(std::min<uint32_t>(ctx->nchunks, ctx->cus * 16)),
                           dim3(kBlock), 0, ctx->stream, ctx->d_chunk_hub, ctx->d_chunk_start, ctx->nchunks,
                           ctx->d_hubs, ctx->d_off, ctx->d_edges, V, f, ctx->d_rev);
      unsigned long long h[4] = {1, 0, 0, 1};
      PJ_HIP(hipMemcpyAsync(h, f, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
      PJ_HIP(hipStreamSynchronize(ctx->stream));
      ctx->symmetric = h[0] == 0 && h[1] == h[2];
      ctx->self_loops = h[3] != 0;
      if (!ctx->symmetric && ctx->d_rev) {  // incomplete: never read
        PJ_HIP(hipFree(ctx->d_rev));
        ctx->d_rev = nullptr;
      }
    }
    ctx->skip_clean = ctx->rows_sorted && ctx->symmetric && ctx->dirty_enabled;
  }
  if (slot_labels(ctx)) return -1;
  if (pair_index_build(ctx, host_off)) return -1;
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_template = false;
  ctx->active_vertices = V;
  ctx->active_edges = E;
  ctx->live = nullptr;
  ctx->n_live = V;
  return 0;
}

int pj_upload_graph(pj_ctx* ctx, uint32_t V, uint64_t E, const uint64_t* offsets, const uint32_t* edges,
                    const uint64_t* labels) {
  if (!ctx || !offsets || (E && !edges) || (V && !labels)) return -1;
  return upload_common(ctx, V, E, offsets, edges, labels, hipMemcpyHostToDevice);
}

int pj_upload_graph_device(pj_ctx* ctx, uint32_t V, uint64_t E, const uint64_t* offsets, const uint32_t* edges,
                           const uint64_t* labels) {
  if (!ctx || !offsets || (E && !edges) || (V && !labels)) return -1;
  return upload_common(ctx, V, E, offsets, edges, labels, hipMemcpyDeviceToDevice);
}

// Binary TOTEM file (graph_store_binary, src/totem/totem_graph.cu:883-923;
// read back by :253-312): magic, id widths, counts, flags, offsets of 32 or
// 64 bits, vertex values when valued, neighbour ids.  The neighbour ids go
// to HBM in chunks through two pinned host buffers: the read of one chunk
// overlaps the copy of the other, and no host array holds the edge list.
// Offsets (widened to 64 bits) and values are read whole: 8 bytes per vertex.
int pj_upload_totem_file(pj_ctx* ctx, const char* path, const uint64_t* labels) {
  if (!ctx || !path) return -1;
  struct FileCloser {
    FILE* f;
    ~FileCloser() {
      if (f) fclose(f);
    }
  } fc{fopen(path, "rb")};
  FILE* f = fc.f;
  if (!f) {
    ctx->err = std::string("cannot open ") + path;
    return -1;
  }
  uint32_t head[4] = {0, 0, 0, 0};  // magic, vid bytes, eid bytes, vertex count
  uint64_t E = 0;
  uint8_t fl[3] = {0, 0, 0};        // valued, weighted, directed
  bool ok = fread(head, 4, 4, f) == 4 && head[0] == 0x10102048u && head[1] == 4 && (head[2] == 4 || head[2] == 8);
  if (ok) {
    if (head[2] == 4) {
      uint32_t e32 = 0;
      ok = fread(&e32, 4, 1, f) == 1;
      E = e32;
    } else {
      ok = fread(&E, 8, 1, f) == 1;
    }
  }
  ok = ok && fread(fl, 1, 3, f) == 3;
  if (!ok) {
    ctx->err = std::string("not a binary TOTEM graph: ") + path;
    return -1;
  }
  const uint32_t V = head[3];
  if (!labels && !fl[0]) {
    ctx->err = std::string("graph file has no vertex values and no labels were given: ") + path;
    return -1;
  }
  std::vector<uint64_t> off(V + 1ull);
  if (head[2] == 8) {
    ok = fread(off.data(), 8, V + 1ull, f) == V + 1ull;
  } else {
    std::vector<uint32_t> o32(V + 1ull);
    ok = fread(o32.data(), 4, V + 1ull, f) == V + 1ull;
    for (uint64_t i = 0; ok && i <= V; ++i) off[i] = o32[i];
  }
  std::vector<uint64_t> values;
  if (ok && fl[0]) {
    if (labels) {
      ok = fseeko(f, static_cast<off_t>(V) * 8, SEEK_CUR) == 0;
    } else {
      values.resize(V);
      ok = fread(values.data(), 8, V, f) == V;
    }
  }
  if (!ok) {
    ctx->err = std::string("truncated binary graph: ") + path;
    return -1;
  }
  if (upload_begin(ctx, V, E)) return -1;
  PJ_HIP(hipMemcpyAsync(ctx->d_off, off.data(), (V + 1ull) * 8, hipMemcpyHostToDevice, ctx->stream));
  if (V)
    PJ_HIP(hipMemcpyAsync(ctx->d_labels, labels ? labels : values.data(), V * 8ull, hipMemcpyHostToDevice, ctx->stream));
  constexpr uint64_t kChunk = 16ull << 20;  // ids per chunk (64 MiB)
  uint32_t* buf[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  int rc = 0;
  for (int b = 0; b < 2 && !rc; ++b) {
    if (hipHostMalloc(reinterpret_cast<void**>(&buf[b]), kChunk * 4, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&done[b], hipEventDisableTiming) != hipSuccess) {
      ctx->err = "pinned staging buffers for the edge list";
      rc = -1;
    }
  }
  for (uint64_t pos = 0, k = 0; !rc && pos < E; pos += kChunk, ++k) {
    const int b = static_cast<int>(k & 1);
    const uint64_t n = std::min<uint64_t>(kChunk, E - pos);
    if (k >= 2 && hipEventSynchronize(done[b]) != hipSuccess) {  // its previous copy has left the buffer
      ctx->err = "edge list copy";
      rc = -1;
      break;
    }
    if (fread(buf[b], 4, n, f) != n) {
      ctx->err = std::string("truncated binary graph: ") + path;
      rc = -1;
      break;
    }
    if (hipMemcpyAsync(ctx->d_edges + pos, buf[b], n * 4, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipEventRecord(done[b], ctx->stream) != hipSuccess) {
      ctx->err = "edge list copy";
      rc = -1;
    }
  }
  hipStreamSynchronize(ctx->stream);
  for (int b = 0; b < 2; ++b) {
    if (buf[b]) hipHostFree(buf[b]);
    if (done[b]) hipEventDestroy(done[b]);
  }
  if (rc) {
    free_graph(ctx);
    ctx->V = 0;
    ctx->E = 0;
    ctx->ea_words = 0;
    return -1;
  }
  return upload_finish(ctx, off.data());
}

int pj_share_graph(pj_ctx* ctx, const pj_ctx* owner) {
  if (!ctx || !owner || ctx == owner) return -1;
  if (!owner->d_off) {
    ctx->err = "the owner context holds no graph";
    return -1;
  }
  if (owner->device != ctx->device) {
    ctx->err = "contexts on different devices cannot share a graph";
    return -1;
  }
  PJ_HIP(hipSetDevice(ctx->device));
  free_graph(ctx);
  ctx->graph_shared = true;
  // Contexts that share a graph (the workers of the query forms) keep their
  // walks on one stream unless PJ_SIDE_WALKS asks otherwise: several
  // contexts already run side by side, and every extra stream takes a share
  // of the hardware queues that the workers' streams are spread over.
  if (!getenv("PJ_SIDE_WALKS")) ctx->side_walks = 0;
  ctx->V = owner->V;
  ctx->E = owner->E;
  ctx->ea_words = owner->ea_words;
  ctx->d_off = owner->d_off;
  ctx->d_edges = owner->d_edges;
  ctx->d_labels = owner->d_labels;
  ctx->d_rev = owner->d_rev;
  ctx->d_slot_lab = owner->d_slot_lab;
  ctx->graph_labels = owner->graph_labels;
  ctx->d_pair_doff = owner->d_pair_doff;
  ctx->d_pair_pos = owner->d_pair_pos;
  ctx->pair_heavy = owner->pair_heavy;
  ctx->pair_degree = owner->pair_degree;
  ctx->self_loops = owner->self_loops;
  ctx->rows_sorted = owner->rows_sorted;
  ctx->symmetric = owner->symmetric;
  ctx->skip_clean = ctx->rows_sorted && ctx->symmetric && ctx->dirty_enabled;
  ctx->nhubs = owner->nhubs;
  ctx->nchunks = owner->nchunks;
  ctx->d_hubs = owner->d_hubs;
  ctx->d_chunk_hub = owner->d_chunk_hub;
  ctx->d_chunk_start = owner->d_chunk_start;
  const uint64_t nh = ctx->nhubs;
  if (alloc_state(ctx) || dev_alloc(ctx, &ctx->d_hub_count, nh * kMaxLabels) || dev_alloc(ctx, &ctx->d_hub_elim, nh) ||
      dev_alloc(ctx, &ctx->d_nhubs, 1) || dev_alloc(ctx, &ctx->d_hub_hit, nh) || dev_alloc(ctx, &ctx->d_hub_new, nh))
    return -1;
  PJ_HIP(hipMemcpyAsync(ctx->d_nhubs, &ctx->nhubs, 4, hipMemcpyHostToDevice, ctx->stream));
  if (nh) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, nh * kMaxLabels)), dim3(kBlock), 0, ctx->stream,
                       ctx->d_hub_count, nh * kMaxLabels, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, nh)), dim3(kBlock), 0, ctx->stream, ctx->d_hub_elim, nh, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, nh)), dim3(kBlock), 0, ctx->stream, ctx->d_hub_hit, nh,
                       0xFFFFFFFFu);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, nh)), dim3(kBlock), 0, ctx->stream, ctx->d_hub_new, nh, 0u);
  }
  PJ_HIP(hipGetLastError());
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  ctx->have_template = false;
  ctx->active_vertices = ctx->V;
  ctx->active_edges = ctx->E;
  ctx->live = nullptr;
  ctx->n_live = ctx->V;
  return 0;
}

int pj_get_device(const pj_ctx* ctx, int* device) {
  if (!ctx || !device) return -1;
  *device = ctx->device;
  return 0;
}

int pj_state_checksum(pj_ctx* ctx, uint64_t* out) {
  if (!ctx || !out) return -1;
  if (!ctx->have_template) {
    ctx->err = "no template installed";
    return -1;
  }
  PJ_HIP(hipSetDevice(ctx->device));
  return with_word(ctx, [&](auto w) { return checksum_t<decltype(w)>(ctx, out); });
}

int pj_set_template(pj_ctx* ctx, const pj_template_desc* d) {
  if (!ctx) return -1;
  if (!d || d->vertex_count == 0 || d->vertex_count > 64) {
    ctx->err = "template must have 1..64 vertices";
    return -1;
  }
  if (!ctx->d_off) {
    ctx->err = "no graph uploaded";
    return -1;
  }
  PJ_HIP(hipSetDevice(ctx->device));
  DevTemplate& t = ctx->h_t;
  std::memset(&t, 0, sizeof(t));
  const uint32_t P = d->vertex_count;
  t.P = P;
  // Dense codes for the template's distinct labels, in vertex order.
  for (uint32_t p = 0; p < P; ++p) {
    uint32_t k = 0;
    while (k < t.K && t.label_raw[k] != d->vertex_label[p]) ++k;
    if (k == t.K) t.label_raw[t.K++] = d->vertex_label[p];
    t.pcode[p] = static_cast<uint8_t>(k);
    t.code_vertices[k] |= 1ull << p;
    t.nbr[p] = d->neighbor_mask[p];
  }
  for (uint32_t p = 0; p < P; ++p) {
    for (uint32_t i = d->need_offset[p]; i < d->need_offset[p + 1]; ++i) {
      uint32_t k = 0;
      while (k < t.K && t.label_raw[k] != d->need_label[i]) ++k;
      if (k == t.K) {
        ctx->err = "neighbour label not carried by any template vertex";
        return -1;
      }
      uint32_t need = std::min<uint32_t>(d->need_count[i], 127u);
      t.need[p][k >> 3] |= static_cast<uint64_t>(need) << (8 * (k & 7));
    }
  }
  std::memset(t.slot_code, kNoCode, sizeof(t.slot_code));
  for (size_t i = 0; i < ctx->graph_labels.size() && i < 255; ++i)
    for (uint32_t k = 0; k < t.K; ++k)
      if (t.label_raw[k] == ctx->graph_labels[i]) t.slot_code[i] = static_cast<uint8_t>(k);
  ctx->kw = t.K <= 8 ? 1 : t.K <= 16 ? 2 : t.K <= 32 ? 4 : 8;
  ctx->word_bytes = P <= 8 ? 1 : P <= 16 ? 2 : P <= 32 ? 4 : 8;
  ctx->cycles.clear();
  for (uint32_t c = 0; c < d->cycle_count; ++c)
    ctx->cycles.emplace_back(d->cycle_vertex + d->cycle_offset[c], d->cycle_vertex + d->cycle_offset[c + 1]);
  ctx->paths.clear();
  ctx->path_label.clear();
  for (uint32_t c = 0; c < d->path_count; ++c) {
    ctx->paths.emplace_back(d->path_vertex + d->path_offset[c], d->path_vertex + d->path_offset[c + 1]);
    ctx->path_label.push_back(d->path_initial_label[c]);
  }
  auto copy_walk = [](const pj_walk_desc& w) {
    pj_ctx::HostWalk h;
    h.root = w.root;
    h.move.assign(w.move, w.move + w.length);
    h.vertex.assign(w.vertex, w.vertex + w.length);
    h.mb.assign(w.move_back, w.move_back + w.length);
    return h;
  };
  ctx->tds.clear();
  for (uint32_t c = 0; c < d->tds_count; ++c) {
    std::vector<pj_ctx::HostWalk> ws;
    for (uint32_t k = d->tds_walk_offset[c]; k < d->tds_walk_offset[c + 1]; ++k) ws.push_back(copy_walk(d->walks[k]));
    ctx->tds.push_back(std::move(ws));
  }
  ctx->enumeration = copy_walk(d->enumeration_walk);
  PJ_HIP(hipMemcpyAsync(ctx->d_t, &t, sizeof(DevTemplate), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_label_codes, dim3(grid_for(ctx, ctx->V)), dim3(kBlock), 0, ctx->stream, ctx->d_labels,
                     ctx->d_code, ctx->V, ctx->d_t);
  PJ_HIP(hipGetLastError());
  ctx->have_template = true;
  return pj_state_reset(ctx);
}

int pj_state_reset(pj_ctx* ctx) {
  if (!ctx) return -1;
  if (!ctx->have_template) {
    ctx->err = "no template installed";
    return -1;
  }
  PJ_HIP(hipSetDevice(ctx->device));
  return with_word(ctx, [&](auto w) { return state_reset_t<decltype(w)>(ctx); });
}

int pj_reset_modified(pj_ctx* ctx) {
  if (!ctx) return -1;
  if (!ctx->have_template) {
    ctx->err = "no template installed";
    return -1;
  }
  hipLaunchKernelGGL(k_clear_modified, dim3(grid_for(ctx, ctx->V)), dim3(kBlock), 0, ctx->stream, ctx->d_flags,
                     ctx->V);
  mark_all_rescan(ctx);
  PJ_HIP(hipGetLastError());
  ctx->modified_cleared = true;
  ctx->fresh = 2;
  return 0;
}

int pj_run_step(pj_ctx* ctx, int step, uint32_t index, pj_algo_results* out) {
  if (!ctx) return -1;
  if (!ctx->have_template) {
    ctx->err = "no template installed";
    return -1;
  }
  PJ_HIP(hipSetDevice(ctx->device));
  if (step == PJ_STEP_CC || step == PJ_STEP_PC || step == PJ_STEP_TDS || step == PJ_STEP_ENUMERATION)
    if (install_nlcc(ctx, step, index)) return -1;
  PJ_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  int rc = with_word(ctx, [&](auto w) { return run_step_t<decltype(w)>(ctx, step, index); });
  ctx->modified_cleared = false;
  if (rc) return rc;
  PJ_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  const size_t blocks = ctx->timing ? static_cast<size_t>(ctx->nlaunch) + 1 : 1;
  hipLaunchKernelGGL(k_sum_counters, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0, ctx->stream,
                     ctx->d_counters, ctx->h_sums_dev, ctx->d_bin_ring, static_cast<uint32_t>(kBinRing * kBins));
  PJ_HIP(hipGetLastError());
  ctx->counters_clean = true;
  PJ_HIP(wait_stream(ctx));
  PJ_HIP(hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1));
  pj_algo_results r{0, 0, 0, 0};
  ctx->last_slots = 0;
  for (size_t b = 0; b < blocks; ++b) {
    const unsigned long long* sum = ctx->h_sums + b * kCounterStride;
    r.vertex_eliminated += sum[C_VERTEX];
    r.edge_eliminated += sum[C_EDGE];
    r.match_eliminated += sum[C_MATCH];
    r.enumeration += sum[C_ENUM];
    ctx->last_slots += sum[C_SLOTS] + sum[C_SKIP];
    if (b == 0) continue;
    const int kid = ctx->launch_kid[b - 1];
    float ms = 0.f;
    PJ_HIP(hipEventElapsedTime(&ms, ctx->lev[2 * (b - 1)], ctx->lev[2 * (b - 1) + 1]));
    ctx->k_launches[kid] += 1;
    ctx->k_ms[kid] += ms;
    ctx->k_items[kid] += sum[C_ITEMS];
    ctx->k_slots[kid] += sum[C_SLOTS] + sum[C_XSLOTS];
    ctx->k_elim[kid] += sum[C_EDGE];
  }
  ctx->active_vertices -= r.vertex_eliminated;
  ctx->active_edges -= r.edge_eliminated;
  if (out) *out = r;
  // Once fewer than half of the live list can be active, the list is
  // compacted, so frontier construction reads only live vertices.
  if (ctx->n_live > ctx->live_min && ctx->active_vertices < ctx->n_live / 2 &&
      with_word(ctx, [&](auto w) { return compact_live_t<decltype(w)>(ctx); }))
    return -1;
  return 0;
}

int pj_set_kernel_timing(pj_ctx* ctx, int enable) {
  if (!ctx) return -1;
  if (enable && !ctx->lev[0])
    for (int i = 0; i < 2 * kMaxLaunches; ++i) PJ_HIP(hipEventCreate(&ctx->lev[i]));
  ctx->timing = enable != 0;
  ctx->counters_clean = false;  // the steps use a different number of counter blocks
  ctx->timing_kid = -1;
  for (int k = 0; k < K_COUNT; ++k) {
    ctx->k_launches[k] = ctx->k_items[k] = ctx->k_slots[k] = ctx->k_elim[k] = 0;
    ctx->k_ms[k] = 0.0;
  }
  return 0;
}

int pj_set_kernel_timing_kind(pj_ctx* ctx, const char* name) {
  if (!ctx) return -1;
  int kid = -1;
  if (name && *name) {
    for (int k = 0; k < K_COUNT; ++k)
      if (std::strcmp(kKernelNames[k], name) == 0) kid = k;
    if (kid < 0) {
      ctx->err = std::string("unknown kernel kind: ") + name;
      return -1;
    }
  }
  if (pj_set_kernel_timing(ctx, 1)) return -1;
  ctx->timing_kid = kid;
  return 0;
}

int pj_kernel_stats(pj_ctx* ctx, pj_kernel_stat* out, int max) {
  if (!ctx) return -1;
  int n = 0;
  for (int k = 0; k < K_COUNT; ++k) {
    if (!ctx->k_launches[k]) continue;
    if (out && n < max) {
      pj_kernel_stat& st = out[n];
      std::memset(&st, 0, sizeof(st));
      std::strncpy(st.name, kKernelNames[k], sizeof(st.name) - 1);
      st.launches = ctx->k_launches[k];
      st.ms = ctx->k_ms[k];
      st.items = ctx->k_items[k];
      st.slots = ctx->k_slots[k];
      st.edges_eliminated = ctx->k_elim[k];
      st.word_bytes = static_cast<uint32_t>(ctx->word_bytes);
    }
    ++n;
  }
  return n;
}

int pj_last_step_ms(pj_ctx* ctx, double* ms) {
  if (!ctx || !ms) return -1;
  *ms = ctx->last_ms;
  return 0;
}

int pj_last_step_slots(pj_ctx* ctx, uint64_t* slots) {
  if (!ctx || !slots) return -1;
  *slots = ctx->last_slots;
  return 0;
}

int pj_get_counters(pj_ctx* ctx, uint64_t* av, uint64_t* ae) {
  if (!ctx) return -1;
  if (av) *av = ctx->active_vertices;
  if (ae) *ae = ctx->active_edges;
  return 0;
}

int pj_download_state(pj_ctx* ctx, uint64_t* cand, uint8_t* vflags, uint32_t* ea_words) {
  if (!ctx) return -1;
  if (!ctx->have_template) {
    ctx->err = "no template installed";
    return -1;
  }
  PJ_HIP(hipSetDevice(ctx->device));
  return with_word(ctx, [&](auto w) { return download_t<decltype(w)>(ctx, cand, vflags, ea_words); });
}

int pj_upload_state(pj_ctx* ctx, const uint64_t* cand, const uint8_t* vflags, const uint32_t* ea_words,
                    uint64_t active_vertices, uint64_t active_edges) {
  if (!ctx || !cand || !vflags || !ea_words) return -1;
  if (!ctx->have_template) {
    ctx->err = "no template installed";
    return -1;
  }
  PJ_HIP(hipSetDevice(ctx->device));
  int rc = with_word(ctx, [&](auto w) { return upload_state_t<decltype(w)>(ctx, cand, vflags, ea_words); });
  if (rc) return rc;
  ctx->active_vertices = active_vertices;
  ctx->active_edges = active_edges;
  return 0;
}

const char* pj_last_error(pj_ctx* ctx) {
  if (!ctx) return "null context";
  return ctx->err.c_str();
}

// Active-graph compaction (pj_compact.hip).
static int compact_common(pj_ctx* ctx, unsigned long long* scratch, uint32_t* vertices, uint64_t* cand,
                          uint64_t cap_v, uint32_t* src, uint32_t* dst, uint64_t cap_e, uint64_t* n_v,
                          uint64_t* n_e) {
  pjc::CompactArgs a{ctx->d_off, ctx->d_edges, ctx->d_ea, ctx->d_cand, ctx->word_bytes, ctx->V, ctx->E,
                     vertices, cand, src, dst, cap_v, cap_e, ctx->stream};
  uint64_t nv = 0, ne = 0;
  const int rc = pjc::compact_active(a, scratch, &nv, &ne, &ctx->err);
  if (n_v) *n_v = nv;
  if (n_e) *n_e = ne;
  return rc;
}

static int compact_ready(pj_ctx* ctx) {
  if (!ctx) return -1;
  if (!ctx->have_template) {
    ctx->err = "no template installed";
    return -1;
  }
  PJ_HIP(hipSetDevice(ctx->device));
  return 0;
}

int pj_compact_active(pj_ctx* ctx, uint32_t* vertices, uint64_t* cand, uint64_t cap_vertices, uint32_t* src,
                      uint32_t* dst, uint64_t cap_slots, uint64_t* n_vertices, uint64_t* n_slots) {
  if (compact_ready(ctx)) return -1;
  // counters sized for this graph, kept by the context
  if (!ctx->d_compact_scratch && dev_alloc(ctx, &ctx->d_compact_scratch, pjc::scratch_words(ctx->V, ctx->E)))
    return -1;
  return compact_common(ctx, ctx->d_compact_scratch, vertices, cand, cap_vertices, src, dst, cap_slots, n_vertices,
                        n_slots);
}

int pj_compact_active_noalloc(pj_ctx* ctx, uint32_t* vertices, uint64_t* cand, uint64_t cap_vertices, uint32_t* src,
                              uint32_t* dst, uint64_t cap_slots, uint64_t* n_vertices, uint64_t* n_slots) {
  if (compact_ready(ctx)) return -1;
  if (ctx->V > pjc::kStaticMaxItems || ctx->E > pjc::kStaticMaxItems) {
    ctx->err = "pj_compact_active_noalloc: graphs of at most 2^30 vertices and 2^30 slots (its static scratch)";
    return -1;
  }
  // one call at a time in this process: every call on a device shares its scratch
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  unsigned long long* scratch = nullptr;
  if (pjc::static_scratch(ctx->device, &scratch, &ctx->err)) return -1;
  return compact_common(ctx, scratch, vertices, cand, cap_vertices, src, dst, cap_slots, n_vertices, n_slots);
}

int pj_download_active(pj_ctx* ctx, uint32_t* vertices, uint64_t* cand, uint64_t cap_vertices, uint32_t* src,
                       uint32_t* dst, uint64_t cap_slots, uint64_t* n_vertices, uint64_t* n_slots) {
  if (compact_ready(ctx)) return -1;
  uint64_t nv = 0, ne = 0;
  // counts first (outputs of capacity 0), then outputs of exactly that size
  if (!ctx->d_compact_scratch && dev_alloc(ctx, &ctx->d_compact_scratch, pjc::scratch_words(ctx->V, ctx->E)))
    return -1;
  if (compact_common(ctx, ctx->d_compact_scratch, nullptr, nullptr, 0, nullptr, nullptr, 0, &nv, &ne) &&
      (nv == 0 && ne == 0))
    return -1;
  if (n_vertices) *n_vertices = nv;
  if (n_slots) *n_slots = ne;
  if (!vertices && !cand && !src && !dst) return 0;  // counts only
  if (nv > cap_vertices || ne > cap_slots) {
    ctx->err = "pj_download_active: output capacity too small";
    return -1;
  }
  if (nv > ctx->out_cap_v) {
    if (dev_alloc(ctx, &ctx->d_out_v, nv) || dev_alloc(ctx, &ctx->d_out_cand, nv)) return -1;
    ctx->out_cap_v = nv;
  }
  if (ne > ctx->out_cap_e) {
    if (dev_alloc(ctx, &ctx->d_out_src, ne) || dev_alloc(ctx, &ctx->d_out_dst, ne)) return -1;
    ctx->out_cap_e = ne;
  }
  if (compact_common(ctx, ctx->d_compact_scratch, ctx->d_out_v, ctx->d_out_cand, ctx->out_cap_v, ctx->d_out_src,
                     ctx->d_out_dst, ctx->out_cap_e, &nv, &ne))
    return -1;
  if (vertices && nv) PJ_HIP(hipMemcpyAsync(vertices, ctx->d_out_v, nv * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (cand && nv) PJ_HIP(hipMemcpyAsync(cand, ctx->d_out_cand, nv * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (src && ne) PJ_HIP(hipMemcpyAsync(src, ctx->d_out_src, ne * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (dst && ne) PJ_HIP(hipMemcpyAsync(dst, ctx->d_out_dst, ne * 4, hipMemcpyDeviceToHost, ctx->stream));
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int pj_graph_stat(pj_ctx* ctx, pj_label_stat* out) {
  if (!ctx) return -1;
  if (!out) {
    ctx->err = "pj_graph_stat: no output";
    return -1;
  }
  if (!ctx->have_template) {
    ctx->err = "no template installed";
    return -1;
  }
  PJ_HIP(hipSetDevice(ctx->device));
  if (zero_counters(ctx)) return -1;
  return with_word(ctx, [&](auto w) { return graph_stat_t<decltype(w)>(ctx, out); });
}

void pj_destroy(pj_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->side) hipStreamSynchronize(ctx->side);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  free_graph(ctx);
  if (ctx->d_t) hipFree(ctx->d_t);
  if (ctx->d_counters) hipFree(ctx->d_counters);
  if (ctx->d_sums) hipFree(ctx->d_sums);
  if (ctx->h_sums) hipHostFree(ctx->h_sums);
  if (ctx->h_word) hipHostFree(ctx->h_word);
  if (ctx->ev0) hipEventDestroy(ctx->ev0);
  if (ctx->ev1) hipEventDestroy(ctx->ev1);
  if (ctx->ev_done) hipEventDestroy(ctx->ev_done);
  for (int i = 0; i < 2 * kMaxLaunches; ++i)
    if (ctx->lev[i]) hipEventDestroy(ctx->lev[i]);
  if (ctx->side) hipStreamDestroy(ctx->side);
  if (ctx->ev_fork) hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) hipEventDestroy(ctx->ev_join);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

}  // extern "C"

// Internal interface of the batched LCC (pj_batch.h).
namespace pjb {

int graph_view(const pj_ctx* ctx, GraphView* out) {
  if (!ctx || !out || !ctx->d_off) return -1;
  out->device = ctx->device;
  out->cus = ctx->cus;
  out->V = ctx->V;
  out->E = ctx->E;
  out->off = ctx->d_off;
  out->edges = ctx->d_edges;
  out->labels = ctx->d_labels;
  out->hubs = ctx->d_hubs;
  out->nhubs = ctx->nhubs;
  out->chunk_hub = ctx->d_chunk_hub;
  out->chunk_start = ctx->d_chunk_start;
  out->nchunks = ctx->nhubs ? ctx->nchunks : 0u;
  out->hub_degree = kHubDegree;
  out->hub_chunk = kHubChunk;
  out->slot_lab = ctx->d_slot_lab;
  out->graph_labels = ctx->graph_labels.data();
  out->ngraph_labels = static_cast<uint32_t>(ctx->graph_labels.size());
  return 0;
}

int import_lcc_state(pj_ctx* ctx, const void* cand, int word_bytes, const uint8_t* flags, const uint32_t* ea_words,
                     uint64_t active_vertices, uint64_t active_edges, hipEvent_t ready) {
  if (!ctx) return -1;
  if (!ctx->have_template) {
    ctx->err = "no template installed";
    return -1;
  }
  if (word_bytes != ctx->word_bytes) {
    ctx->err = "imported candidate words do not match the installed template";
    return -1;
  }
  PJ_HIP(hipSetDevice(ctx->device));
  if (ready) PJ_HIP(hipStreamWaitEvent(ctx->stream, ready, 0));
  PJ_HIP(hipMemcpyAsync(ctx->d_cand, cand, static_cast<uint64_t>(ctx->V) * word_bytes, hipMemcpyDeviceToDevice,
                        ctx->stream));
  PJ_HIP(hipMemcpyAsync(ctx->d_flags, flags, ctx->V, hipMemcpyDeviceToDevice, ctx->stream));
  PJ_HIP(hipMemcpyAsync(ctx->d_ea, ea_words, ctx->ea_words * 4, hipMemcpyDeviceToDevice, ctx->stream));
  // toUnmatch is empty after every phase 2
  const uint64_t words = (static_cast<uint64_t>(ctx->V) * word_bytes + 3) / 4;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(ctx, words)), dim3(kBlock), 0, ctx->stream,
                     static_cast<uint32_t*>(ctx->d_unmatch), words, 0u);
  mark_all_dirty(ctx);
  ctx->active_vertices = active_vertices;
  ctx->active_edges = active_edges;
  ctx->labels_consistent = true;  // LCC0 ran: every candidate carries its vertex's label
  ctx->modified_cleared = false;
  ctx->live = nullptr;
  ctx->n_live = ctx->V;
  PJ_HIP(hipGetLastError());
  PJ_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

}  // namespace pjb
