// Label statistics of the active graph (see pj_stat.h).
//
// GraphStatCpu::compute (algorithms/graph_stat_cpu.tpp:40-82) visits every
// active vertex, adds its label, and for every active slot to an active
// neighbour adds the label pair and the neighbour's label to a per-thread
// tally, from which GraphStat::addAccumulation (algorithms/graph_stat.cc:
// 101-113) counts the vertex under "at least i neighbours of that label" for
// i up to the template's maximum.  Here each vertex's tally is its own (the
// reference's tally, privateAccumulation, is declared once per thread and
// never cleared, so there a vertex also counts the slots of the vertices its
// thread visited before it; the count of one vertex is what the statistic's
// name and its use in the effectiveness estimate describe).
//
// One launch walks the frontier bins of every active vertex (a group of 8,
// 32 or 64 lanes per row by degree) and then the 2048-slot chunks of the hub
// rows (a wave per chunk).  Counts go to LDS with LDS atomics (the vertex
// and label-pair tables, the "exact" table of capped per-vertex tallies, and
// one tally of K counters per lane group) and leave the block with one
// global atomic per non-zero entry.  A hub's tally is summed over its chunks
// in hub_count and capped by a second, small launch.  The pass is bound by
// the gathers of the neighbours' candidate words and label codes.
#include "pj_stat.h"

namespace pjs {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr uint32_t kChunk = 2048;
constexpr int kRowsPerWave = 64 / 8;      // groups of the narrowest (8-lane) rows in a wave
constexpr int kMaxGroups = kWaves * kRowsPerWave;

__device__ __forceinline__ bool ea_on(const uint32_t* __restrict__ ea, uint64_t e) {
  return (ea[e >> 5] >> (e & 31)) & 1u;
}

struct Lds {
  uint32_t* vtx;    // [K]
  uint32_t* edge;   // [K x K]
  uint32_t* exact;  // [exact_words]
  uint32_t* tab;    // [K x K]: offset << 8 | max
  uint32_t* tally;  // [kMaxGroups x K]
};

// The G lanes of one group count the slots [beg, end) of row v into the
// label-pair table and the group's tally (lane gl of the group).
template <typename Word, int G>
__device__ __forceinline__ void count_slots(const StatArgs& a, const Lds& s, uint32_t K, uint32_t lf, uint64_t beg,
                                            uint64_t end, int gl, uint32_t* tally) {
  const Word* __restrict__ cand = static_cast<const Word*>(a.cand);
  for (uint64_t b = beg; b < end; b += G) {
    const uint64_t e = b + gl;
    if (e < end && ea_on(a.ea, e)) {
      const uint32_t u = a.edges[e];
      const Word cu = cand[u];
      const uint32_t lt = a.code[u];
      if (cu != Word(0) && lt < K) {
        atomicAdd(&s.edge[lf * K + lt], 1u);
        atomicAdd(&tally[lt], 1u);
      }
    }
  }
  // The group's lanes read the tally next: their LDS atomics come first.
  __threadfence_block();
}

// Rows of one bin: a group of G lanes per vertex, grid-strided.
template <typename Word, int G>
__device__ void stat_bin(const StatArgs& a, const Lds& s, uint32_t K, const uint32_t* __restrict__ list,
                         uint32_t n) {
  constexpr uint32_t kGroups = kThreads / G;
  const int gl = threadIdx.x & (G - 1);
  const uint32_t gi = threadIdx.x / G;
  // Tally rows are private to a wave (8 per wave, one per group): a wave may
  // start the next bin while another still counts rows of this one.
  uint32_t* tally = s.tally + ((threadIdx.x >> 6) * kRowsPerWave + (__lane_id() / G)) * K;
  for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * kGroups + gi; g < n;
       g += static_cast<uint64_t>(gridDim.x) * kGroups) {
    const uint32_t v = list[g];
    const uint32_t lf = a.code[v];
    if (lf >= K) continue;  // group-uniform; an active vertex carries a template label
    if (gl == 0) atomicAdd(&s.vtx[lf], 1u);
    count_slots<Word, G>(a, s, K, lf, a.off[v], a.off[v + 1], gl, tally);
    for (uint32_t lt = gl; lt < K; lt += G) {
      const uint32_t c = tally[lt];
      if (!c) continue;
      tally[lt] = 0;
      const uint32_t tb = s.tab[lf * K + lt];
      const uint32_t m = tb & 0xFFu;
      if (m) atomicAdd(&s.exact[(tb >> 8) + min(c, m) - 1], 1u);
    }
    __threadfence_block();
  }
}

// Chunks of the hub rows, a wave each: tallies are summed per hub in
// hub_count (capped by k_stat_hub_finish once every chunk is counted).
template <typename Word>
__device__ void stat_hub_chunks(const StatArgs& a, const Lds& s, uint32_t K) {
  const Word* __restrict__ cand = static_cast<const Word*>(a.cand);
  const int lane = __lane_id();
  const uint32_t w = threadIdx.x >> 6;
  uint32_t* tally = s.tally + w * kRowsPerWave * K;
  for (uint32_t c = blockIdx.x * kWaves + w; c < a.nchunks; c += gridDim.x * kWaves) {
    const uint32_t h = a.chunk_hub[c];
    const uint32_t v = a.hubs[h];
    const uint32_t lf = a.code[v];
    if (cand[v] == Word(0) || lf >= K) continue;  // wave-uniform
    const uint64_t row = a.off[v], row_end = a.off[v + 1];
    const uint64_t beg = a.chunk_start[c];
    const uint64_t end = beg + kChunk < row_end ? beg + kChunk : row_end;
    if (beg == row && lane == 0) atomicAdd(&s.vtx[lf], 1u);
    count_slots<Word, 64>(a, s, K, lf, beg, end, lane, tally);
    for (uint32_t lt = lane; lt < K; lt += 64) {
      const uint32_t n = tally[lt];
      if (!n) continue;
      tally[lt] = 0;
      atomicAdd(&a.hub_count[static_cast<uint64_t>(h) * kLabels + lt], n);
    }
    __threadfence_block();
  }
}

template <typename Word>
__global__ void __launch_bounds__(kThreads) k_label_stat(StatArgs a) {
  extern __shared__ uint32_t lds[];
  const uint32_t K = a.K, X = a.exact_words;
  Lds s;
  s.vtx = lds;
  s.edge = s.vtx + K;
  s.exact = s.edge + K * K;
  s.tab = s.exact + X;
  s.tally = s.tab + K * K;
  const uint32_t zero_words = K + K * K + X;
  for (uint32_t i = threadIdx.x; i < zero_words; i += kThreads) lds[i] = 0;
  for (uint32_t i = threadIdx.x; i < kMaxGroups * K; i += kThreads) s.tally[i] = 0;
  const uint32_t* gtab = reinterpret_cast<const uint32_t*>(a.buffer + kCountWords);
  for (uint32_t i = threadIdx.x; i < K * K; i += kThreads) s.tab[i] = gtab[(i / K) * kLabels + i % K];
  __syncthreads();
  const uint64_t V = a.V;
  // longest rows first
  stat_bin<Word, 64>(a, s, K, a.bins + 3 * V, a.bin_count[3]);
  stat_bin<Word, 64>(a, s, K, a.bins + 2 * V, a.bin_count[2]);
  stat_bin<Word, 32>(a, s, K, a.bins + 1 * V, a.bin_count[1]);
  stat_bin<Word, 8>(a, s, K, a.bins, a.bin_count[0]);
  stat_hub_chunks<Word>(a, s, K);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < K; i += kThreads)
    if (s.vtx[i]) atomicAdd(&a.buffer[kVertexAt + i], static_cast<unsigned long long>(s.vtx[i]));
  for (uint32_t i = threadIdx.x; i < K * K; i += kThreads)
    if (s.edge[i])
      atomicAdd(&a.buffer[kEdgeAt + (i / K) * kLabels + i % K], static_cast<unsigned long long>(s.edge[i]));
  for (uint32_t i = threadIdx.x; i < X; i += kThreads)
    if (s.exact[i]) atomicAdd(&a.buffer[kExactAt + i], static_cast<unsigned long long>(s.exact[i]));
}

// Caps each active hub's summed tally and clears hub_count.
template <typename Word>
__global__ void __launch_bounds__(kThreads) k_stat_hub_finish(StatArgs a) {
  const Word* __restrict__ cand = static_cast<const Word*>(a.cand);
  const uint32_t* gtab = reinterpret_cast<const uint32_t*>(a.buffer + kCountWords);
  for (uint32_t h = blockIdx.x * kThreads + threadIdx.x; h < a.nhubs; h += gridDim.x * kThreads) {
    const uint32_t v = a.hubs[h];
    const uint32_t lf = a.code[v];
    const bool active = cand[v] != Word(0) && lf < a.K;
    uint32_t* cnt = a.hub_count + static_cast<uint64_t>(h) * kLabels;
    for (uint32_t lt = 0; lt < a.K; ++lt) {
      const uint32_t c = cnt[lt];
      if (!c) continue;
      cnt[lt] = 0;
      if (!active) continue;
      const uint32_t tb = gtab[lf * kLabels + lt];
      const uint32_t m = tb & 0xFFu;
      if (m) atomicAdd(&a.buffer[kExactAt + (tb >> 8) + min(c, m) - 1], 1ull);
    }
  }
}

__global__ void k_zero_counts(unsigned long long* p) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < kCountWords;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = 0;
}

template <typename Word>
int run(const StatArgs& a, std::string* err) {
  const size_t lds = 4ull * (a.K + 2ull * a.K * a.K + a.exact_words + static_cast<uint64_t>(kMaxGroups) * a.K);
  hipLaunchKernelGGL(k_zero_counts, dim3(16), dim3(kThreads), 0, a.stream, a.buffer);
  const uint32_t blocks = static_cast<uint32_t>(a.cus > 0 ? a.cus * 4 : 256);
  hipLaunchKernelGGL(k_label_stat<Word>, dim3(blocks), dim3(kThreads), lds, a.stream, a);
  if (a.nhubs)
    hipLaunchKernelGGL(k_stat_hub_finish<Word>, dim3((a.nhubs + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       a.stream, a);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    *err = std::string("label statistics: ") + hipGetErrorString(e);
    return -1;
  }
  return 0;
}

}  // namespace

int label_stat(const StatArgs& a, std::string* err) {
  if (a.K == 0 || a.K > static_cast<uint32_t>(kLabels) || a.exact_words > kExactMax) {
    *err = "label statistics: bad label table";
    return -1;
  }
  switch (a.word_bytes) {
    case 1:
      return run<uint8_t>(a, err);
    case 2:
      return run<uint16_t>(a, err);
    case 4:
      return run<uint32_t>(a, err);
    default:
      return run<uint64_t>(a, err);
  }
}

}  // namespace pjs
