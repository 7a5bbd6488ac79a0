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
// One launch walks the frontier bins of every active vertex (a group of 2,
// 8 or 64 lanes per row by degree, 4 slots per lane per pass) and then the
// 2048-slot chunks of the hub rows (a wave per chunk).  Counts go to LDS (the
// vertex and label-pair tables, the "exact" table of capped per-vertex
// tallies, and one tally of K counters per lane group) and leave the block
// with one global atomic per non-zero entry.  A hub's tally is summed over its chunks
// in hub_count and capped by a second, small launch.  The pass is bound by
// the gathers of the neighbours' candidate words and label codes.
#include "pj_stat.h"

namespace pjs {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr uint32_t kChunk = 2048;
constexpr int kSlots = 4;                   // slots per lane per pass, loaded before any is counted
constexpr int kRowsPerWave = 64 / 2;        // groups of the narrowest (2-lane) rows in a wave
constexpr int kMaxGroups = kWaves * kRowsPerWave;

__device__ __forceinline__ bool ea_on(const uint32_t* __restrict__ ea, uint64_t e) {
  return (ea[e >> 5] >> (e & 31)) & 1u;
}

template <typename Word>
__device__ __forceinline__ int lowest(Word w) {
  return sizeof(Word) == 8 ? __builtin_ctzll(static_cast<uint64_t>(w)) : __builtin_ctz(static_cast<uint32_t>(w));
}

struct Lds {
  uint32_t* vtx;    // [K]
  uint32_t* edge;   // [K x K]
  uint32_t* exact;  // [exact_words]
  uint16_t* tally;  // [kMaxGroups x K]: slots of one row (<= 2048 per row or chunk)
  uint8_t* pcode;   // [64]: label code of each template vertex
};

// One pass of the G lanes of a group over the slots [b, b + G * kSlots) of
// row v (lane gl takes b + gl + j * G): edge ids and edge-active words first,
// then the neighbours' candidate words (and label codes unless they are
// derived from the candidates: once every candidate of every vertex carries
// the vertex's label, the lowest candidate's template vertex gives it), then
// the counts.  Counts are aggregated across the wave by (group, label): one
// LDS update per distinct pair instead of one per slot.
template <typename Word, int G>
__device__ __forceinline__ void count_pass(const StatArgs& a, const Lds& s, uint32_t K, uint32_t lf, uint64_t b,
                                           uint64_t end, int gl, uint16_t* tally) {
  const Word* __restrict__ cand = static_cast<const Word*>(a.cand);
  const int lane = __lane_id();
  uint32_t u[kSlots], on[kSlots];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const uint64_t e = b + gl + static_cast<uint64_t>(j) * G;
    u[j] = e < end ? a.edges[e] : 0u;
    on[j] = e < end ? ea_on(a.ea, e) : 0u;
  }
  uint32_t lt[kSlots];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const Word cu = on[j] ? cand[u[j]] : Word(0);
    lt[j] = cu == Word(0) ? kLabels : a.derive ? s.pcode[lowest<Word>(cu)] : a.code[u[j]];
  }
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    const uint32_t key = static_cast<uint32_t>(lane / G) * kLabels + lt[j];
    uint64_t pending = __ballot(lt[j] < K);
    while (pending) {
      const int leader = __builtin_ctzll(pending);
      const uint32_t k = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(key), leader));
      const uint64_t m = __ballot(lt[j] < K && key == k);
      if (lane == leader) {
        const uint32_t n = static_cast<uint32_t>(__popcll(m));
        atomicAdd(&s.edge[lf * K + lt[j]], n);
        tally[lt[j]] = static_cast<uint16_t>(tally[lt[j]] + n);  // the group's own entry, one writer
      }
      pending &= ~m;
    }
  }
}

// Rows of one bin: a group of G lanes per vertex, grid-strided.
template <typename Word, int G>
__device__ void stat_bin(const StatArgs& a, const Lds& s, uint32_t K, const uint32_t* __restrict__ list,
                         uint32_t n) {
  constexpr uint32_t kGroups = kThreads / G;
  const int gl = threadIdx.x & (G - 1);
  const uint32_t gi = threadIdx.x / G;
  const uint32_t* gtab = reinterpret_cast<const uint32_t*>(a.buffer + kCountWords);
  // Tally rows are private to a wave (one per group): a wave may start the
  // next bin while another still counts rows of this one.
  uint16_t* tally = s.tally + ((threadIdx.x >> 6) * kRowsPerWave + (__lane_id() / G)) * K;
  for (uint64_t g = static_cast<uint64_t>(blockIdx.x) * kGroups + gi; g < n;
       g += static_cast<uint64_t>(gridDim.x) * kGroups) {
    const uint32_t v = list[g];
    const uint32_t lf = a.code[v];
    if (lf >= K) continue;  // group-uniform; an active vertex carries a template label
    if (gl == 0) atomicAdd(&s.vtx[lf], 1u);
    const uint64_t beg = a.off[v], end = a.off[v + 1];
    for (uint64_t b = beg; b < end; b += G * kSlots) count_pass<Word, G>(a, s, K, lf, b, end, gl, tally);
    __threadfence_block();  // the leaders' tally writes before the group reads them
    for (uint32_t lt = gl; lt < K; lt += G) {
      const uint32_t c = tally[lt];
      if (!c) continue;
      tally[lt] = 0;
      const uint32_t tb = gtab[lf * kLabels + lt];
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
  uint16_t* tally = s.tally + w * kRowsPerWave * K;
  for (uint32_t c = blockIdx.x * kWaves + w; c < a.nchunks; c += gridDim.x * kWaves) {
    const uint32_t h = a.chunk_hub[c];
    const uint32_t v = a.hubs[h];
    const uint32_t lf = a.code[v];
    if (cand[v] == Word(0) || lf >= K) continue;  // wave-uniform
    const uint64_t row = a.off[v], row_end = a.off[v + 1];
    const uint64_t beg = a.chunk_start[c];
    const uint64_t end = beg + kChunk < row_end ? beg + kChunk : row_end;
    if (beg == row && lane == 0) atomicAdd(&s.vtx[lf], 1u);
    for (uint64_t b = beg; b < end; b += 64 * kSlots) count_pass<Word, 64>(a, s, K, lf, b, end, lane, tally);
    __threadfence_block();
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
  s.tally = reinterpret_cast<uint16_t*>(s.exact + X);
  s.pcode = reinterpret_cast<uint8_t*>(s.tally + kMaxGroups * K);
  const uint32_t zero_words = K + K * K + X + (kMaxGroups * K + 1) / 2;
  for (uint32_t i = threadIdx.x; i < zero_words; i += kThreads) lds[i] = 0;
  if (threadIdx.x < 64) s.pcode[threadIdx.x] = a.pcode[threadIdx.x];
  __syncthreads();
  const uint64_t V = a.V;
  // longest rows first; 2 / 8 / 64 / 64 lanes of kSlots slots per row
  stat_bin<Word, 64>(a, s, K, a.bins + 3 * V, a.bin_count[3]);
  stat_bin<Word, 64>(a, s, K, a.bins + 2 * V, a.bin_count[2]);
  stat_bin<Word, 8>(a, s, K, a.bins + 1 * V, a.bin_count[1]);
  stat_bin<Word, 2>(a, s, K, a.bins, a.bin_count[0]);
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

// Caps each active hub's summed tally and clears hub_count.  The capped
// counts of the block's hubs are summed in LDS (the exact table has a few
// entries per label pair, which every hub would otherwise hit with its own
// global atomic) and leave the block with one atomic per non-zero entry.
template <typename Word>
__global__ void __launch_bounds__(kThreads) k_stat_hub_finish(StatArgs a) {
  __shared__ uint32_t s_exact[kExactMax];
  const uint32_t X = a.exact_words;
  for (uint32_t i = threadIdx.x; i < X; i += kThreads) s_exact[i] = 0;
  __syncthreads();
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
      if (m) atomicAdd(&s_exact[(tb >> 8) + min(c, m) - 1], 1u);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < X; i += kThreads)
    if (s_exact[i]) atomicAdd(&a.buffer[kExactAt + i], static_cast<unsigned long long>(s_exact[i]));
}

__global__ void k_zero_counts(unsigned long long* p) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < kCountWords;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = 0;
}

template <typename Word>
int run(const StatArgs& a, std::string* err) {
  const size_t lds = 4ull * (a.K + 1ull * a.K * a.K + a.exact_words) + 2ull * kMaxGroups * a.K + 64;
  hipLaunchKernelGGL(k_zero_counts, dim3(16), dim3(kThreads), 0, a.stream, a.buffer);
  const uint32_t blocks = static_cast<uint32_t>(a.cus > 0 ? a.cus * 16 : 1024);
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
