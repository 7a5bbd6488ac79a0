// Batched LCC: the LCC0 step and the LCC steps to the first fixed point of
// many templates over one resident graph, in sweeps that every template
// shares.
//
// The reference solves one template per run: LCC0 (multiple_label_lcc0_cpu.
// tpp:13-147), then LCC phase 1 and phase 2 (multiple_label_lcc_cpu.tpp:
// 26-102, 107-142) until a step eliminates nothing (algo_results.cc:18-23,
// multiple_label_step.cc:52-62).  Here up to 64 templates form a group that
// sits in the 64 lanes of a wavefront: lane t holds template t's candidate
// word of a vertex, so the state of one vertex in one group is one 128-byte
// line (16-bit words, templates of at most 16 vertices), and each row(v) and
// each neighbour's candidate line are read once for all 64 templates.
//
// HBM layout of a batch of G groups (V vertices, E slots):
//   cand, unm  u16[G][V][64]  candidate set / toUnmatch of each template
//   mod, sched, act, pend, dirty  u64[G][V]  per-template bits: modified
//              (sticky until the state leaves the batch), scheduled, active
//              (cand != 0), toUnmatch pending for phase 2, candidates of v or
//              of a neighbour changed since v's last evaluation
//   ea         u64[G][E]      per-template edge-active bits of each slot
//   blab       u8[V]          index of the vertex label in the union of the
//              batch's template labels (0xFF: no template has it)
// Per lane: template tables, the step it is on, and its trace (vertices,
// edges and matches eliminated per step).  A template whose LCC step
// eliminates nothing is parked: its state is written out in the layout of a
// single-template context (candidate words, flag bytes, edge-active words),
// from which pj_batch_export installs it into a context that continues the
// template's pipeline (the NLCC steps and the LCC steps after them).
//
// Each step is the reference's two-phase (Jacobi) sweep per template, so the
// per-template results are those of the single-template kernels bit for
// bit: the same scheduled sets are evaluated, eliminations and removals are
// applied in phase 2 only, and the per-step counters are kept per template.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/prunejuice.h"
#include "pj_batch.h"

namespace {

typedef uint16_t Word;                // candidate word of one template (P <= 16)
constexpr int kBlockB = 256;
constexpr int kWaves = kBlockB / 64;  // waves per block
constexpr uint32_t kMaxP = 16;
constexpr uint32_t kMaxSteps = 1000;  // the reference's step cap (multiple_label_cpu.cc:113)
constexpr int kMaxEnt = 240;          // requirement entries per template (<= sum of template degrees)
constexpr int kS = 16;                // slots per pass of LCC phase 1
constexpr int kEntLds = 32;           // requirement entries per lane kept in LDS by phase 1
constexpr uint8_t kNoLab = 0xFF;
constexpr uint32_t kMaxLabels = 255;  // labels of the union (index 0xFF means none)
constexpr uint8_t kFlagModified = 2, kFlagScheduled = 4;  // flag bytes of a context
// control words per group
constexpr int kRun = 0, kInit = 1, kFirst = 2, kDone = 3, kSecond = 4, kCtl = 8;

// Tables of one template lane.  ent[ent_off[p] .. ent_off[p + 1]) are the
// requirements of template vertex p (multiple_label_generate_constraint_
// local.cc:12-35): for each label of its template neighbours, the template
// vertices carrying that label (low 16 bits) and how many neighbours of that
// label p needs (high bits).
struct LaneTab {
  uint32_t ent[kMaxEnt];
  uint16_t labmask[256];  // template vertices carrying each union label
  uint16_t glabmask[256]; // template vertices carrying each graph label (index into the graph's labels)
  uint16_t nbr[kMaxP];    // template neighbours of each template vertex
  uint8_t ent_off[kMaxP + 1];
  uint8_t P;
  uint16_t full;
};

struct BArgs {
  uint32_t V;
  uint32_t nhubs, nchunks, hub_chunk, hub_degree;
  uint32_t nlab;
  uint32_t nglab;               // graph labels indexed by slot_lab (0: no slot labels)
  uint64_t E;
  const uint8_t* slot_lab;      // [E] index of each slot's neighbour label among the graph's labels
  const uint64_t* off;
  const uint32_t* edges;
  const uint32_t* hubs;
  const uint32_t* chunk_hub;
  const uint64_t* chunk_start;
  Word* cand;
  Word* unm;
  uint64_t* mod;
  uint64_t* sched;
  uint64_t* act;
  uint64_t* pend;
  uint64_t* dirty;              // [G][V] templates whose candidates at v or at a neighbour changed
  uint64_t* ea;
  const uint8_t* blab;
  const LaneTab* tab;
  uint64_t* ctl;                // [G][kCtl]
  uint32_t* lane_step;          // [G * 64]
  unsigned long long* trace;    // [G * 64][kMaxSteps][4]: vertices, edges, matches eliminated
  uint32_t* list1;              // [G][V] phase 1 frontier (rows of at most hub_degree slots)
  uint32_t* list2;              // [G][V] phase 2 frontier
  uint32_t* count;              // [G][4]
  uint32_t* hub_cnt;            // [G][nhubs][64][16] label counts of hub rows (zero between steps)
  unsigned long long* hub_elim; // [G][nhubs] templates that eliminated a hub slot (zero between steps)
};

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t shfl64(uint64_t x, int src) {
  const uint32_t lo = __shfl(static_cast<uint32_t>(x), src, 64);
  const uint32_t hi = __shfl(static_cast<uint32_t>(x >> 32), src, 64);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// A wave-uniform 64-bit value in scalar registers.
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x >> 32)));
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x)));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Lane k's 64-bit value (split in halves hi, lo), wave-uniform.
__device__ __forceinline__ uint64_t lane64(uint32_t hi, uint32_t lo, int k) {
  const uint32_t h = static_cast<uint32_t>(__builtin_amdgcn_readlane(hi, k));
  const uint32_t l = static_cast<uint32_t>(__builtin_amdgcn_readlane(lo, k));
  return (static_cast<uint64_t>(h) << 32) | l;
}

__device__ __forceinline__ uint64_t wave_or64(uint64_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t lo = __shfl_xor(static_cast<uint32_t>(x), o, 64);
    const uint32_t hi = __shfl_xor(static_cast<uint32_t>(x >> 32), o, 64);
    x |= (static_cast<uint64_t>(hi) << 32) | lo;
  }
  return x;
}

// Adds the per-lane counters of every wave of the block (vertices, edges and
// matches eliminated) to the trace row of each lane's current step.  Every
// thread of the block must call it.
__device__ void flush_counters(const BArgs& a, uint32_t g, uint64_t ve, uint64_t ee, uint64_t me) {
  __shared__ unsigned long long s[3 * 64];
  for (int i = threadIdx.x; i < 3 * 64; i += blockDim.x) s[i] = 0;
  __syncthreads();
  const int t = lane_id();
  if (ve) atomicAdd(&s[t], static_cast<unsigned long long>(ve));
  if (ee) atomicAdd(&s[64 + t], static_cast<unsigned long long>(ee));
  if (me) atomicAdd(&s[128 + t], static_cast<unsigned long long>(me));
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * 64; i += blockDim.x) {
    const unsigned long long x = s[i];
    if (!x) continue;
    const uint32_t L = g * 64 + (i & 63);
    atomicAdd(&a.trace[(static_cast<uint64_t>(L) * kMaxSteps + a.lane_step[L]) * 4 + (i >> 6)], x);
  }
}

// Saturating per-template-vertex counters, bit-sliced over NB planes (cap
// 2^NB - 1): x is one bit, the template vertex counted.  NB = 1 keeps
// whether a count is non-zero (every requirement of the running templates
// is one neighbour).
template <int NB>
__device__ __forceinline__ void count_one(Word (&c)[NB], Word x) {
  Word sat = c[0];
#pragma unroll
  for (int k = 1; k < NB; ++k) sat &= c[k];
  Word y = x & ~sat;
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const Word carry = c[k] & y;
    c[k] ^= y;
    y = carry;
  }
}

// Template vertices of `cv` whose requirements fail: for each requirement
// entry (template vertices of one label, count needed) the neighbours counted
// under those template vertices must reach the count.  A neighbour is counted
// under the lowest of its candidates, which carries its label (LCC0 ran), so
// the count of a label is the sum over the template vertices of that label;
// counts saturate at 2^NB - 1, which is at least every count needed.
template <int NB>
__device__ __forceinline__ Word requirements(Word cv, const Word (&c)[NB], const uint32_t* ent, const uint8_t* off) {
  Word fail = 0;
  for (Word m = cv; m; m &= m - 1) {
    const int p = __builtin_ctz(m);
    for (int j = off[p]; j < off[p + 1]; ++j) {
      const uint32_t e = ent[j];
      const Word mask = static_cast<Word>(e & 0xFFFFu);
      uint32_t cnt = 0;
#pragma unroll
      for (int k = 0; k < NB; ++k) cnt += static_cast<uint32_t>(__builtin_popcount(c[k] & mask)) << k;
      if (cnt < (e >> 16)) {
        fail |= Word(1) << p;
        break;
      }
    }
  }
  return fail;
}

// Union label index of every vertex.
__global__ void kb_labels(const uint64_t* __restrict__ labels, const uint64_t* __restrict__ uni, uint32_t nu,
                          uint32_t V, uint8_t* __restrict__ blab) {
  __shared__ uint64_t s[kMaxLabels];
  for (uint32_t i = threadIdx.x; i < nu; i += blockDim.x) s[i] = uni[i];
  __syncthreads();
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t l = labels[v];
    uint8_t k = kNoLab;
    for (uint32_t i = 0; i < nu; ++i)
      if (s[i] == l) {
        k = static_cast<uint8_t>(i);
        break;
      }
    blab[v] = k;
  }
}

// Admission of the lanes `mask` of group g: their trace rows cleared, and
// they are to run LCC0 in the next step.
__global__ void kb_admit(BArgs a, uint32_t g, uint64_t mask) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < 64ull * kMaxSteps * 4;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t t = static_cast<uint32_t>(i / (kMaxSteps * 4));
    if ((mask >> t) & 1) a.trace[(static_cast<uint64_t>(g) * 64) * kMaxSteps * 4 + i] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    uint64_t* c = a.ctl + g * kCtl;
    c[kInit] |= mask;
    c[kRun] &= ~mask;
    c[kFirst] &= ~mask;
    c[kSecond] &= ~mask;
    c[kDone] &= ~mask;
  }
  if (blockIdx.x == 0 && threadIdx.x < 64 && ((mask >> threadIdx.x) & 1)) a.lane_step[g * 64 + threadIdx.x] = 0;
}

// Every slot active again for the lanes `mask` of group g.
__global__ void kb_ea_set(BArgs a, uint32_t g, uint64_t mask) {
  uint64_t* ea = a.ea + static_cast<uint64_t>(g) * a.E;
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < a.E; e += (uint64_t)gridDim.x * blockDim.x)
    ea[e] |= mask;
}

__global__ void kb_fill(uint32_t* p, uint32_t n, uint32_t value) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = value;
}

// LCC0 (multiple_label_lcc0_cpu.tpp:25-47) for the lanes admitted in this
// step: every vertex starts with all template vertices and keeps those
// carrying its label; matches eliminated = P - kept, a vertex left without
// candidates is eliminated.  The lanes' flags start as the reference's
// initial state (scheduled, not modified; common_cpu.tpp:29-66).  One wave
// per vertex, four vertices per pass.
__global__ void __launch_bounds__(kBlockB) kb_lcc0(BArgs a) {
  const uint32_t g = blockIdx.y;
  const uint64_t init = a.ctl[g * kCtl + kInit];
  if (!init) return;
  extern __shared__ uint16_t s_labmask[];  // [64][nlab]
  for (uint32_t i = threadIdx.x; i < 64 * a.nlab; i += blockDim.x)
    s_labmask[i] = a.tab[g * 64 + i / a.nlab].labmask[i % a.nlab];
  __syncthreads();
  const int t = lane_id();
  const bool mine = (init >> t) & 1;
  const LaneTab& tb = a.tab[g * 64 + t];
  const Word full = mine ? tb.full : Word(0);
  const int P = mine ? tb.P : 0;
  uint64_t ve = 0, me = 0;
  const uint64_t gV = static_cast<uint64_t>(g) * a.V;
  const uint32_t wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
  const uint32_t nw = gridDim.x * kWaves;
  for (uint64_t v0 = static_cast<uint64_t>(wave) * 4; v0 < a.V; v0 += static_cast<uint64_t>(nw) * 4) {
    uint8_t lab[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) lab[k] = v0 + k < a.V ? a.blab[v0 + k] : kNoLab;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t v = v0 + k;
      if (v >= a.V) break;
      const Word c = lab[k] == kNoLab ? Word(0) : static_cast<Word>(full & s_labmask[t * a.nlab + lab[k]]);
      if (mine) {
        a.cand[(gV + v) * 64 + t] = c;
        me += static_cast<uint64_t>(P - __builtin_popcount(c));
        ve += c == 0;
      }
      const uint64_t A = __ballot(c != 0);
      const uint64_t gv = gV + v;
      if (t == 0) a.act[gv] = (a.act[gv] & ~init) | A;
      if (t == 1) a.sched[gv] |= init;
      if (t == 2) a.mod[gv] &= ~init;
      if (t == 3) a.pend[gv] &= ~init;
    }
  }
  flush_counters(a, g, ve, 0, me);
}

// Frontiers.  mode 1 (phase 1): vertices scheduled and active in a running
// template that must be evaluated, rows of at most hub_degree slots (hub rows
// go by chunks).  A template in its first or second LCC step evaluates every
// scheduled vertex; later, only those whose candidates or whose neighbours'
// candidates changed since their last evaluation (dirty): a scheduled vertex
// without such a change would remove nothing and eliminate no slot (its
// removals were applied, its eliminated slots stay eliminated, its counts are
// the same), so skipping it changes no result, flag or counter.
// mode 2 (phase 2): vertices with phase 2 work in a running template: a
// modified template with removals pending or not yet scheduled, or an
// unmodified active template whose scheduled bit must be recomputed (every
// one in its first LCC step; later only those not scheduled, see kb_phase2).
// Each block compacts tiles of kTile consecutive vertices with one atomic.
constexpr int kFrontierPer = 8;  // vertices per thread
constexpr uint32_t kTile = kBlockB * kFrontierPer;

__global__ void __launch_bounds__(kBlockB) kb_frontier(BArgs a, int mode) {
  const uint32_t g = blockIdx.y;
  const uint64_t* ctl = a.ctl + g * kCtl;
  const uint64_t run = ctl[kRun];
  if (!run) return;
  const uint64_t first = ctl[kFirst], full = first | ctl[kSecond];
  const uint64_t gV = static_cast<uint64_t>(g) * a.V;
  uint32_t* list = (mode == 1 ? a.list1 : a.list2) + gV;
  uint32_t* cnt = a.count + g * 4 + mode;
  __shared__ uint32_t s_wave[kWaves];
  __shared__ uint32_t s_base;
  const int t = lane_id(), wv = threadIdx.x >> 6;
  for (uint64_t tile = blockIdx.x; tile * kTile < a.V; tile += gridDim.x) {
    const uint64_t v0 = tile * kTile + static_cast<uint64_t>(threadIdx.x) * kFrontierPer;
    uint32_t bits = 0;
#pragma unroll
    for (int k = 0; k < kFrontierPer; ++k) {
      const uint64_t v = v0 + k;
      if (v >= a.V) break;
      const uint64_t gv = gV + v;
      bool in;
      if (mode == 1) {
        in = (a.sched[gv] & a.act[gv] & run & (a.dirty[gv] | full)) != 0 && a.off[v + 1] - a.off[v] <= a.hub_degree;
      } else {
        const uint64_t ac = a.act[gv] & run, m = a.mod[gv], sc = a.sched[gv], p = a.pend[gv];
        in = ((ac & m & (p | ~sc)) | (ac & ~m & (first | ~sc))) != 0;
      }
      bits |= static_cast<uint32_t>(in) << k;
    }
    const uint32_t n = __builtin_popcount(bits);
    uint32_t x = n;  // inclusive prefix over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (t >= o) x += y;
    }
    if (t == 63) s_wave[wv] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t tot = 0;
      for (int w = 0; w < kWaves; ++w) {
        const uint32_t c = s_wave[w];
        s_wave[w] = tot;
        tot += c;
      }
      s_base = tot ? atomicAdd(cnt, tot) : 0u;
    }
    __syncthreads();
    uint32_t pos = s_base + s_wave[wv] + x - n;
    for (int k = 0; k < kFrontierPer; ++k)
      if ((bits >> k) & 1) list[pos++] = static_cast<uint32_t>(v0 + k);
    __syncthreads();
  }
}

// LCC phase 1 (multiple_label_lcc_cpu.tpp:26-102) of every running template
// at the frontier vertices (one wave per vertex, lane t = template t) and at
// the chunks of hub rows.  For template t at vertex v: U = template
// neighbours of v's candidates; a slot whose edge is active and whose
// neighbour u is active for t loses its edge when no candidate of u is in U,
// and otherwise counts u under its lowest candidate.  Slots are read kS at a
// time: lanes 0..kS-1 load the ids and edge masks, then every lane gathers
// its template's word of each neighbour (the 64 lanes read one line).  The
// lane that loaded a slot's mask writes it back when some template
// eliminated the slot (each slot belongs to one wave).  A chunk adds its
// counts to hub_cnt and kb_hub_finish evaluates the hub.
template <int NB>
__global__ void __launch_bounds__(kBlockB) kb_phase1(BArgs a) {
  const uint32_t g = blockIdx.y;
  const uint64_t run = a.ctl[g * kCtl + kRun];
  if (!run) return;
  const uint64_t full = a.ctl[g * kCtl + kFirst] | a.ctl[g * kCtl + kSecond];
  __shared__ Word s_nbr[64 * kMaxP];
  __shared__ uint8_t s_off[64 * (kMaxP + 1)];
  __shared__ uint32_t s_ent[64 * kEntLds];  // the first kEntLds requirement entries of each lane
  for (uint32_t i = threadIdx.x; i < 64 * kEntLds; i += blockDim.x) s_ent[i] = a.tab[g * 64 + i / kEntLds].ent[i % kEntLds];
  // Every running template in its first LCC step (right after LCC0): a
  // neighbour's candidates are the template vertices of its label, read from
  // the slot's label index instead of gathering the neighbour's words.
  const bool fresh = a.slot_lab && a.nglab && (run & ~a.ctl[g * kCtl + kFirst]) == 0;
  // [64][gstride] when fresh: 32-bit entries and an odd row stride, so that
  // the 64 lanes reading one label index hit distinct banks
  extern __shared__ uint32_t s_gmask[];
  const uint32_t gstride = a.nglab | 1u;
  if (fresh)
    for (uint32_t i = threadIdx.x; i < 64 * a.nglab; i += blockDim.x)
      s_gmask[(i / a.nglab) * gstride + i % a.nglab] = a.tab[g * 64 + i / a.nglab].glabmask[i % a.nglab];
  for (uint32_t i = threadIdx.x; i < 64 * kMaxP; i += blockDim.x) s_nbr[i] = a.tab[g * 64 + i / kMaxP].nbr[i % kMaxP];
  for (uint32_t i = threadIdx.x; i < 64 * (kMaxP + 1); i += blockDim.x)
    s_off[i] = a.tab[g * 64 + i / (kMaxP + 1)].ent_off[i % (kMaxP + 1)];
  __syncthreads();
  const int t = lane_id();
  const LaneTab& tb = a.tab[g * 64 + t];
  const uint64_t gV = static_cast<uint64_t>(g) * a.V;
  const Word* cand = a.cand + gV * 64;
  uint64_t* eam = a.ea + static_cast<uint64_t>(g) * a.E;
  const uint32_t n1 = a.count[g * 4 + 1];
  const uint32_t items = n1 + a.nchunks;
  const uint32_t wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
  const uint32_t nw = gridDim.x * kWaves;
  uint64_t ee = 0;
  // Items are taken 64 at a time per wave: lane i loads item i's vertex,
  // slot range, dirty mask and evaluation mask in two rounds of loads; the
  // wave then works through them with their values broadcast from the lanes
  // (one item's chain of dependent loads is then the own words and the row).
  // (item r of a wave's k-th round is wave + (64 k + r) * nw: the rows are
  // spread over the waves as in a grid-strided loop)
  for (uint64_t k0 = 0; wave + k0 * nw < items; k0 += 64) {
    const uint64_t ii64 = wave + (k0 + t) * nw;
    const uint32_t ii = ii64 < items ? static_cast<uint32_t>(ii64) : items;
    uint32_t iv = 0, ih = 0;
    uint64_t ib = 0, ie = 0;
    if (ii < n1) {
      iv = a.list1[gV + ii];
      ib = a.off[iv];
      ie = a.off[iv + 1];
    } else if (ii < items) {
      const uint32_t c = ii - n1;
      ih = a.chunk_hub[c];
      iv = a.hubs[ih];
      ib = a.chunk_start[c];
      const uint64_t row_end = a.off[iv + 1];
      ie = row_end < ib + a.hub_chunk ? row_end : ib + a.hub_chunk;
    }
    uint64_t idv = 0, iw = 0;
    if (ii < items) {
      idv = a.dirty[gV + iv];
      iw = a.sched[gV + iv] & a.act[gV + iv] & run & (idv | full);
    }
    const uint64_t todo = __ballot(iw != 0);
    for (uint64_t tm = todo; tm; tm &= tm - 1) {
    const int j = __builtin_ctzll(tm);
    const uint32_t i = static_cast<uint32_t>(wave + (k0 + j) * nw);
    const bool chunk = i >= n1;
    const uint32_t v = static_cast<uint32_t>(__builtin_amdgcn_readlane(iv, j));
    const uint32_t h = static_cast<uint32_t>(__builtin_amdgcn_readlane(ih, j));
    const uint64_t beg = lane64(static_cast<uint32_t>(ib >> 32), static_cast<uint32_t>(ib), j);
    const uint64_t end = lane64(static_cast<uint32_t>(ie >> 32), static_cast<uint32_t>(ie), j);
    const uint64_t dv = lane64(static_cast<uint32_t>(idv >> 32), static_cast<uint32_t>(idv), j);
    const uint64_t w = lane64(static_cast<uint32_t>(iw >> 32), static_cast<uint32_t>(iw), j);
    // v's own words are read while the first slots' reads are in flight
    Word C = 0, U = 0;
    bool own_read = false;
    Word cp[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) cp[k] = 0;
    uint32_t el = 0;
    for (uint64_t base = beg; base < end; base += kS) {
      const uint64_t e = base + (t & (kS - 1));
      const bool own = t < kS && e < end;
      const uint64_t mk = own ? eam[e] : 0ull;
      const uint32_t src = own ? (fresh ? static_cast<uint32_t>(a.slot_lab[e]) : a.edges[e]) : 0u;
      const uint32_t mlo = static_cast<uint32_t>(mk), mhi = static_cast<uint32_t>(mk >> 32);
      // Slot k's mask and neighbour (or label index) are uniform, in scalar
      // registers; Lk = templates for which slot k is evaluated.  Every
      // lane reads its template's word of each neighbour (one line per
      // slot for the wave), all reads issued before any is used; lanes
      // outside Lk are masked out of the results.
      uint64_t Lk[kS];
      Word cu[kS];
      if (fresh) {
#pragma unroll
        for (int k = 0; k < kS; ++k) {
          Lk[k] = lane64(mhi, mlo, k) & w;
          const uint32_t sk = static_cast<uint32_t>(__builtin_amdgcn_readlane(src, k));
          cu[k] = Lk[k] && sk < a.nglab ? static_cast<Word>(s_gmask[t * gstride + sk]) : Word(0);
        }
      } else {
#pragma unroll
        for (int k = 0; k < kS; ++k) {
          Lk[k] = lane64(mhi, mlo, k) & w;
          const uint32_t sk = static_cast<uint32_t>(__builtin_amdgcn_readlane(src, k));
          cu[k] = Lk[k] ? cand[static_cast<uint64_t>(sk) * 64 + t] : Word(0);
        }
      }
      if (!own_read) {
        C = ((w >> t) & 1) ? cand[static_cast<uint64_t>(v) * 64 + t] : Word(0);
        for (Word m = C; m; m &= m - 1) U |= s_nbr[t * kMaxP + __builtin_ctz(m)];
        own_read = true;
      }
      uint64_t clr = 0;  // lane k: templates that eliminated slot base + k
#pragma unroll
      for (int k = 0; k < kS; ++k) {
        const uint64_t L = Lk[k];
        if (!L) continue;
        const uint64_t ok = __ballot((cu[k] & U) != 0) & L;
        const uint64_t em = __ballot(cu[k] != 0) & L & ~ok;
        el += (em >> t) & 1;
        if (em) clr = t == k ? em : clr;
        count_one<NB>(cp, ((ok >> t) & 1) ? static_cast<Word>(cu[k] & static_cast<Word>(-cu[k])) : Word(0));
      }
      if (own && clr) eam[e] = mk & ~clr;
    }
    if (!own_read) C = ((w >> t) & 1) ? cand[static_cast<uint64_t>(v) * 64 + t] : Word(0);  // a row without slots
    ee += el;
    if (!chunk) {
      const uint8_t* ro = &s_off[t * (kMaxP + 1)];
      const Word fail = !C                        ? Word(0)
                        : ro[kMaxP] <= kEntLds ? requirements<NB>(C, cp, &s_ent[t * kEntLds], ro)
                                               : requirements<NB>(C, cp, tb.ent, ro);
      const uint64_t M = __ballot(fail != 0 || el != 0);
      const uint64_t F = __ballot(fail != 0);
      if (fail) a.unm[(gV + v) * 64 + t] = fail;
      // no other wave writes these words of v in phase 1; atomics without a
      // return value leave no load to wait for
      if (t == 0 && M) atomicOr(reinterpret_cast<unsigned long long*>(&a.mod[gV + v]), static_cast<unsigned long long>(M));
      if (t == 1 && F) atomicOr(reinterpret_cast<unsigned long long*>(&a.pend[gV + v]), static_cast<unsigned long long>(F));
      if (t == 2 && (dv & w)) a.dirty[gV + v] = dv & ~w;
    } else {
      if (C) {
        uint32_t* hc = a.hub_cnt + ((static_cast<uint64_t>(g) * a.nhubs + h) * 64 + t) * 16;
        Word any = cp[0];
#pragma unroll
        for (int k = 1; k < NB; ++k) any |= cp[k];
        for (Word m = any; m; m &= m - 1) {
          const int q = __builtin_ctz(m);
          uint32_t n = 0;
#pragma unroll
          for (int k = 0; k < NB; ++k) n |= static_cast<uint32_t>((cp[k] >> q) & 1) << k;
          atomicAdd(&hc[q], n);
        }
      }
      const uint64_t EL = __ballot(el != 0);
      if (t == 0 && EL) atomicOr(&a.hub_elim[static_cast<uint64_t>(g) * a.nhubs + h], static_cast<unsigned long long>(EL));
    }
    }
  }
  flush_counters(a, g, 0, ee, 0);
}

// Hub rows after their chunks: requirements from the summed counts (exact
// sums of per-chunk counts that saturate at 2^NB - 1, at least every count
// needed), modified when a requirement fails or a chunk eliminated a slot.
// Clears the sums.  One wave per hub.
__global__ void __launch_bounds__(kBlockB) kb_hub_finish(BArgs a) {
  const uint32_t g = blockIdx.y;
  const uint64_t run = a.ctl[g * kCtl + kRun];
  if (!run || !a.nhubs) return;
  const uint64_t full = a.ctl[g * kCtl + kFirst] | a.ctl[g * kCtl + kSecond];
  const int t = lane_id();
  const LaneTab& tb = a.tab[g * 64 + t];
  const uint64_t gV = static_cast<uint64_t>(g) * a.V;
  const uint32_t wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
  const uint32_t nw = gridDim.x * kWaves;
  for (uint32_t h = wave; h < a.nhubs; h += nw) {
    const uint32_t v = a.hubs[h];
    const uint64_t dv = a.dirty[gV + v];
    const uint64_t w = a.sched[gV + v] & a.act[gV + v] & run & (dv | full);
    if (!w) continue;
    Word fail = 0;
    if ((w >> t) & 1) {
      const Word C = a.cand[(gV + v) * 64 + t];
      uint32_t* hc = a.hub_cnt + ((static_cast<uint64_t>(g) * a.nhubs + h) * 64 + t) * 16;
      uint32_t cq[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) cq[q] = hc[q];
#pragma unroll
      for (int q = 0; q < 16; ++q) hc[q] = 0;
      for (Word m = C; m; m &= m - 1) {
        const int p = __builtin_ctz(m);
        for (int j = tb.ent_off[p]; j < tb.ent_off[p + 1]; ++j) {
          const uint32_t e = tb.ent[j];
          uint32_t cnt = 0;
#pragma unroll
          for (int q = 0; q < 16; ++q) cnt += ((e >> q) & 1) ? cq[q] : 0u;
          if (cnt < (e >> 16)) {
            fail |= Word(1) << p;
            break;
          }
        }
      }
    }
    unsigned long long* he = &a.hub_elim[static_cast<uint64_t>(g) * a.nhubs + h];
    const uint64_t EL = *he;
    const uint64_t M = __ballot(fail != 0) | (EL & w);
    const uint64_t F = __ballot(fail != 0);
    if (fail) a.unm[(gV + v) * 64 + t] = fail;
    if (t == 0 && M) a.mod[gV + v] |= M;
    if (t == 1 && F) a.pend[gV + v] |= F;
    if (t == 2 && EL) *he = 0;
    if (t == 3 && (dv & w)) a.dirty[gV + v] = dv & ~w;
  }
}

// LCC phase 2 (multiple_label_lcc_cpu.tpp:107-142) at the phase 2 frontier:
// a modified template drops its toUnmatch set (matches eliminated; a vertex
// left empty is eliminated) and is scheduled; an unmodified active template
// is scheduled iff an active slot of v leads to a vertex modified for it.
// Between two resets of the modified flags the latter can only turn from
// unscheduled to scheduled: the edge bits of an unmodified vertex's row are
// unchanged (a vertex that loses a slot is modified) and modified flags only
// get set.  So after a template's first LCC step only its unscheduled,
// unmodified templates scan; the scan is slot-parallel (lane = slot) over the
// per-template masks and ends once every scanning template found a slot.
__global__ void __launch_bounds__(kBlockB) kb_phase2(BArgs a) {
  const uint32_t g = blockIdx.y;
  const uint64_t run = a.ctl[g * kCtl + kRun];
  if (!run) return;
  const uint64_t first = a.ctl[g * kCtl + kFirst];
  const int t = lane_id();
  const uint64_t gV = static_cast<uint64_t>(g) * a.V;
  const uint64_t* eam = a.ea + static_cast<uint64_t>(g) * a.E;
  const uint32_t n2 = a.count[g * 4 + 2];
  const uint32_t wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
  const uint32_t nw = gridDim.x * kWaves;
  uint64_t ve = 0, me = 0;
  for (uint32_t i = wave; i < n2; i += nw) {
    const uint32_t v = a.list2[gV + i];
    const uint64_t gv = gV + v;
    const uint64_t ac = a.act[gv], m = a.mod[gv], s = a.sched[gv], p = a.pend[gv];
    const uint64_t A = ac & run & m & (p | ~s);
    const uint64_t N = ac & run & ~m & (first | ~s);
    bool died = false, changed = false;
    if (((A & p) >> t) & 1) {
      const uint64_t idx = gv * 64 + t;
      const Word c = a.cand[idx], T = a.unm[idx];
      const Word nc = c & ~T;
      if (T) {
        a.cand[idx] = nc;
        changed = true;
        me += __builtin_popcount(T);
        if (!nc) {
          ++ve;
          died = true;
        }
      }
    }
    const uint64_t D = __ballot(died);
    // Templates past their first LCC step whose candidates changed at v mark
    // v and its neighbours dirty (their next evaluation can change).
    const uint64_t push = __ballot(changed) & ~first;
    if (push) {
      if (t == 0) atomicOr(reinterpret_cast<unsigned long long*>(&a.dirty[gv]), static_cast<unsigned long long>(push));
      const uint64_t beg = a.off[v], end = a.off[v + 1];
      for (uint64_t base = beg; base < end; base += 4 * 64) {
        uint32_t u[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint64_t e = base + k * 64 + t;
          u[k] = e < end ? a.edges[e] : 0xFFFFFFFFu;
        }
        uint64_t d[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) d[k] = u[k] != 0xFFFFFFFFu ? a.dirty[gV + u[k]] : push;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if ((d[k] & push) != push)
            atomicOr(reinterpret_cast<unsigned long long*>(&a.dirty[gV + u[k]]), static_cast<unsigned long long>(push));
      }
    }
    uint64_t H = 0;
    if (N) {
      const uint64_t beg = a.off[v], end = a.off[v + 1];
      for (uint64_t base = beg; base < end && (H & N) != N; base += 4 * 64) {
        uint64_t mk[4];
        uint32_t u[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const uint64_t e = base + k * 64 + t;
          mk[k] = e < end ? eam[e] & N & ~H : 0ull;
          u[k] = e < end ? a.edges[e] : 0u;
        }
        uint64_t hit = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) hit |= mk[k] ? (mk[k] & a.mod[gV + u[k]]) : 0ull;
        H |= wave_or64(hit);
      }
    }
    const uint64_t ns = ((s | A) & ~N) | (H & N);
    if (t == 0 && ns != s) a.sched[gv] = ns;
    if (t == 1 && D) a.act[gv] = ac & ~D;
    if (t == 2 && (p & A)) a.pend[gv] = p & ~A;
  }
  flush_counters(a, g, ve, 0, me);
}

// End of a step, one wave per group: lanes that ran LCC0 move to their first
// LCC step; a lane whose LCC step eliminated nothing has reached the fixed
// point (done); the others move to their next step.
__global__ void kb_update(BArgs a) {
  const uint32_t g = blockIdx.x;
  const int t = threadIdx.x;
  const uint32_t L = g * 64 + t;
  uint64_t* c = a.ctl + g * kCtl;
  const uint64_t run = c[kRun], init = c[kInit], done = c[kDone];
  bool conv = false;
  if ((init >> t) & 1) {
    a.lane_step[L] = 1;
  } else if ((run >> t) & 1) {
    const uint32_t s = a.lane_step[L];
    const unsigned long long* r = a.trace + (static_cast<uint64_t>(L) * kMaxSteps + s) * 4;
    if (r[0] == 0 && r[1] == 0 && r[2] == 0)
      conv = true;
    else
      a.lane_step[L] = s + 1 < kMaxSteps ? s + 1 : s;  // the host reports a lane that reaches the cap
  }
  const uint64_t C = __ballot(conv);
  __syncthreads();
  if (t == 0) {
    c[kRun] = (run & ~C) | init;
    c[kInit] = 0;
    c[kSecond] = c[kFirst] & ~C;
    c[kFirst] = init;
    c[kDone] = done | C;
  }
}

// Parking, vertex part: the candidate words and flag bytes of the lanes
// `mask` of group g in the layout of a context (1-byte words for the lanes
// `narrow`), 64 vertices per block through LDS.  Park arrays hold a multiple
// of 64 vertices.
__global__ void __launch_bounds__(kBlockB) kb_park_vertices(BArgs a, uint32_t g, uint64_t mask, uint64_t narrow,
                                                            uint8_t* __restrict__ park_cand, uint64_t cand_stride,
                                                            uint8_t* __restrict__ park_flags, uint64_t flag_stride) {
  __shared__ Word s_c[64 * 65];
  __shared__ uint64_t s_m[64], s_s[64];
  const uint64_t gV = static_cast<uint64_t>(g) * a.V;
  for (uint64_t v0 = static_cast<uint64_t>(blockIdx.x) * 64; v0 < a.V; v0 += static_cast<uint64_t>(gridDim.x) * 64) {
    for (int i = threadIdx.x; i < 64 * 64; i += blockDim.x) {
      const int vi = i >> 6, t = i & 63;
      s_c[vi * 65 + t] = v0 + vi < a.V ? a.cand[(gV + v0 + vi) * 64 + t] : Word(0);
    }
    if (threadIdx.x < 64) {
      const uint64_t v = v0 + threadIdx.x;
      s_m[threadIdx.x] = v < a.V ? a.mod[gV + v] : 0ull;
      s_s[threadIdx.x] = v < a.V ? a.sched[gV + v] : 0ull;
    }
    __syncthreads();
    const int t = threadIdx.x >> 2, part = threadIdx.x & 3;
    if ((mask >> t) & 1) {
      const uint32_t L = g * 64 + t;
      uint8_t* fl = park_flags + L * flag_stride + v0 + part * 16;
      uint32_t fw[4] = {0, 0, 0, 0};
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int vi = part * 16 + j;
        const uint32_t f = (((s_m[vi] >> t) & 1) ? kFlagModified : 0) | (((s_s[vi] >> t) & 1) ? kFlagScheduled : 0);
        fw[j >> 2] |= f << (8 * (j & 3));
      }
      *reinterpret_cast<uint4*>(fl) = make_uint4(fw[0], fw[1], fw[2], fw[3]);
      if ((narrow >> t) & 1) {
        uint32_t cw[4] = {0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < 16; ++j) cw[j >> 2] |= static_cast<uint32_t>(s_c[(part * 16 + j) * 65 + t] & 0xFF) << (8 * (j & 3));
        *reinterpret_cast<uint4*>(park_cand + L * cand_stride + v0 + part * 16) = make_uint4(cw[0], cw[1], cw[2], cw[3]);
      } else {
        uint32_t cw[8];
#pragma unroll
        for (int j = 0; j < 8; ++j)
          cw[j] = static_cast<uint32_t>(s_c[(part * 16 + 2 * j) * 65 + t]) |
                  (static_cast<uint32_t>(s_c[(part * 16 + 2 * j + 1) * 65 + t]) << 16);
        uint4* dst = reinterpret_cast<uint4*>(park_cand + L * cand_stride + 2 * (v0 + part * 16));
        dst[0] = make_uint4(cw[0], cw[1], cw[2], cw[3]);
        dst[1] = make_uint4(cw[4], cw[5], cw[6], cw[7]);
      }
    }
    __syncthreads();
  }
}

// Parking, edge part: each lane's edge-active words, 64 slots per wave and
// pass (lane i holds slot e0 + i's mask; one ballot per parked template gives
// its two words).  Park arrays hold a multiple of 2 words per 64 slots.
__global__ void __launch_bounds__(kBlockB) kb_park_edges(BArgs a, uint32_t g, uint64_t mask,
                                                         uint32_t* __restrict__ park_ea, uint64_t ea_stride) {
  const uint64_t* eam = a.ea + static_cast<uint64_t>(g) * a.E;
  const int lane = lane_id();
  const uint64_t wave = blockIdx.x * (uint64_t)kWaves + (threadIdx.x >> 6);
  const uint64_t nw = static_cast<uint64_t>(gridDim.x) * kWaves;
  for (uint64_t e0 = wave * 64; e0 < a.E; e0 += nw * 64) {
    const uint64_t e = e0 + lane;
    const uint64_t m = e < a.E ? eam[e] : 0ull;
    uint64_t mine = 0;
    for (uint64_t lm = mask; lm; lm &= lm - 1) {
      const int t = __builtin_ctzll(lm);
      const uint64_t b = __ballot((m >> t) & 1);
      mine = lane == t ? b : mine;
    }
    if ((mask >> lane) & 1) {
      const uint32_t L = g * 64 + lane;
      *reinterpret_cast<uint2*>(park_ea + L * ea_stride + e0 / 32) =
          make_uint2(static_cast<uint32_t>(mine), static_cast<uint32_t>(mine >> 32));
    }
  }
}

enum LaneStatus { kFree = 0, kRunning = 1, kParked = 2 };

}  // namespace

struct pj_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  std::mutex mu;
  pjb::GraphView gv{};
  uint32_t G = 0;
  uint32_t nlanes = 0;  // lanes that take templates (the first nlanes of G * 64)
  BArgs a{};
  // device buffers
  LaneTab* d_tab = nullptr;
  uint64_t* d_union = nullptr;
  uint8_t* d_blab = nullptr;
  uint8_t* park_cand = nullptr;
  uint8_t* park_flags = nullptr;
  uint32_t* park_ea = nullptr;
  uint64_t vpad = 0, ea_stride = 0;
  uint64_t* h_ctl = nullptr;  // pinned copy of the control words
  // host lane state
  std::vector<LaneTab> tab;
  std::vector<int> status;
  std::vector<uint32_t> maxneed;
  std::vector<std::vector<pj_algo_results>> rows;
  std::vector<uint64_t> av, ae;
  std::vector<uint64_t> pending;  // per group: lanes admitted since the last step
  std::vector<uint64_t> uni;      // label union
  std::vector<uint64_t> graph_labels;  // labels indexed by slot_lab
  bool labels_dirty = false;
  double ms = 0.0;
  uint64_t steps = 0;
};

namespace {

#define PJB_HIP(call)                                                \
  do {                                                               \
    hipError_t e_ = (call);                                          \
    if (e_ != hipSuccess) {                                          \
      b->err = std::string(#call) + ": " + hipGetErrorString(e_);    \
      return -1;                                                     \
    }                                                                \
  } while (0)

template <typename T>
int balloc(pj_batch* b, T** p, uint64_t count) {
  if (count == 0) count = 1;
  PJB_HIP(hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)));
  return 0;
}

void bfree(pj_batch* b) {
  void* bufs[] = {b->a.cand, b->a.unm, b->a.mod, b->a.sched, b->a.act, b->a.pend, b->a.dirty, b->a.ea, b->a.ctl, b->a.lane_step,
                  b->a.trace, b->a.list1, b->a.list2, b->a.count, b->a.hub_cnt, b->a.hub_elim, b->d_tab, b->d_union,
                  b->d_blab, b->park_cand, b->park_flags, b->park_ea};
  for (void* p : bufs)
    if (p) hipFree(p);
  if (b->h_ctl) hipHostFree(b->h_ctl);
}

int grid_cap(const pj_batch* b, uint64_t items, int per_cu) {
  uint64_t g = (items + kBlockB - 1) / kBlockB;
  const uint64_t cap = static_cast<uint64_t>(b->gv.cus) * per_cu;
  return static_cast<int>(std::max<uint64_t>(1, std::min(g, cap)));
}

// One step of every group: LCC0 of the lanes admitted since the last step,
// one LCC step of the running lanes, then the lane update.
int launch_step(pj_batch* b, int nb) {
  const BArgs& a = b->a;
  const int cus = b->gv.cus;
  hipLaunchKernelGGL(kb_fill, dim3(1), dim3(64), 0, b->stream, a.count, b->G * 4, 0u);
  hipLaunchKernelGGL(kb_lcc0, dim3(cus * 4, b->G), dim3(kBlockB), 64 * a.nlab * sizeof(uint16_t), b->stream, a);
  const uint32_t tiles = static_cast<uint32_t>(std::min<uint64_t>((a.V + kTile - 1) / kTile, cus * 4));
  hipLaunchKernelGGL(kb_frontier, dim3(tiles, b->G), dim3(kBlockB), 0, b->stream, a, 1);
  const size_t gm = 64 * (a.nglab | 1u) * sizeof(uint32_t);
  if (nb == 4)
    hipLaunchKernelGGL(kb_phase1<4>, dim3(cus * 8, b->G), dim3(kBlockB), gm, b->stream, a);
  else if (nb == 2)
    hipLaunchKernelGGL(kb_phase1<2>, dim3(cus * 8, b->G), dim3(kBlockB), gm, b->stream, a);
  else
    hipLaunchKernelGGL(kb_phase1<1>, dim3(cus * 8, b->G), dim3(kBlockB), gm, b->stream, a);
  if (a.nhubs)
    hipLaunchKernelGGL(kb_hub_finish, dim3(std::max<uint32_t>(1, std::min<uint32_t>((a.nhubs + kWaves - 1) / kWaves,
                                                                                  cus * 4)),
                                           b->G),
                       dim3(kBlockB), 0, b->stream, a);
  hipLaunchKernelGGL(kb_frontier, dim3(tiles, b->G), dim3(kBlockB), 0, b->stream, a, 2);
  hipLaunchKernelGGL(kb_phase2, dim3(cus * 8, b->G), dim3(kBlockB), 0, b->stream, a);
  hipLaunchKernelGGL(kb_update, dim3(b->G), dim3(64), 0, b->stream, a);
  PJB_HIP(hipGetLastError());
  return 0;
}

// Writes the state of the lanes `mask` of group g to their park arrays.
int park(pj_batch* b, uint32_t g, uint64_t mask) {
  uint64_t narrow = 0;
  for (int t = 0; t < 64; ++t)
    if (((mask >> t) & 1) && b->tab[g * 64 + t].P <= 8) narrow |= 1ull << t;
  const int cus = b->gv.cus;
  hipLaunchKernelGGL(kb_park_vertices, dim3(std::max<uint32_t>(1, std::min<uint64_t>((b->gv.V + 63) / 64, cus * 8))),
                     dim3(kBlockB), 0, b->stream, b->a, g, mask, narrow, b->park_cand, 2 * b->vpad, b->park_flags,
                     b->vpad);
  hipLaunchKernelGGL(kb_park_edges, dim3(grid_cap(b, b->gv.E, 8)), dim3(kBlockB), 0, b->stream, b->a, g, mask,
                     b->park_ea, b->ea_stride);
  PJB_HIP(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int pj_batch_create(pj_batch** out, const pj_ctx* owner, uint32_t lanes) {
  if (!out) return -1;
  *out = nullptr;
  if (!owner || lanes == 0) return -1;
  pj_batch* b = new pj_batch();
  if (pjb::graph_view(owner, &b->gv) != 0) {
    delete b;
    return -1;
  }
  b->device = b->gv.device;
  if (!b->gv.labels) {
    b->err = "pj_batch_create: the graph has no vertex labels";
    *out = b;
    return -1;
  }
  auto fail = [&]() {
    *out = b;  // the caller reads the message and destroys it
    return -1;
  };
  if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
    b->err = "pj_batch_create: stream or event creation failed";
    return fail();
  }
  b->G = (lanes + 63) / 64;
  b->nlanes = lanes;
  const uint64_t V = b->gv.V, E = b->gv.E, G = b->G, L = G * 64;
  BArgs& a = b->a;
  a.V = b->gv.V;
  a.E = E;
  a.off = b->gv.off;
  a.edges = b->gv.edges;
  a.hubs = b->gv.hubs;
  a.nhubs = b->gv.nhubs;
  a.chunk_hub = b->gv.chunk_hub;
  a.chunk_start = b->gv.chunk_start;
  a.nchunks = b->gv.nchunks;
  a.hub_chunk = b->gv.hub_chunk;
  a.hub_degree = b->gv.hub_degree;
  a.nlab = 1;
  a.slot_lab = b->gv.slot_lab;
  a.nglab = b->gv.slot_lab ? b->gv.ngraph_labels : 0u;
  b->graph_labels.assign(b->gv.graph_labels, b->gv.graph_labels + b->gv.ngraph_labels);
  b->vpad = (V + 63) / 64 * 64;
  b->ea_stride = (E + 63) / 64 * 2;
  if (balloc(b, &a.cand, G * V * 64) || balloc(b, &a.unm, G * V * 64) || balloc(b, &a.mod, G * V) ||
      balloc(b, &a.sched, G * V) || balloc(b, &a.act, G * V) || balloc(b, &a.pend, G * V) ||
      balloc(b, &a.dirty, G * V) ||
      balloc(b, &a.ea, G * E) || balloc(b, &a.ctl, G * kCtl) || balloc(b, &a.lane_step, L) ||
      balloc(b, &a.trace, L * kMaxSteps * 4) || balloc(b, &a.list1, G * V) || balloc(b, &a.list2, G * V) ||
      balloc(b, &a.count, G * 4) || balloc(b, &a.hub_cnt, G * std::max<uint64_t>(1, a.nhubs) * 64 * 16) ||
      balloc(b, &a.hub_elim, G * std::max<uint64_t>(1, a.nhubs)) || balloc(b, &b->d_tab, L) ||
      balloc(b, &b->d_union, kMaxLabels) || balloc(b, &b->d_blab, V) || balloc(b, &b->park_cand, L * 2 * b->vpad) ||
      balloc(b, &b->park_flags, L * b->vpad) || balloc(b, &b->park_ea, L * b->ea_stride))
    return fail();
  a.blab = b->d_blab;
  a.tab = b->d_tab;
  if (hipHostMalloc(reinterpret_cast<void**>(&b->h_ctl), G * kCtl * 8, 0) != hipSuccess) {
    b->err = "pj_batch_create: pinned control words";
    return fail();
  }
  // Clean state: no lane running, counters and masks zero (every lane is
  // set up on admission).
  if (hipMemsetAsync(a.ctl, 0, G * kCtl * 8, b->stream) != hipSuccess ||
      hipMemsetAsync(a.hub_cnt, 0, G * std::max<uint64_t>(1, a.nhubs) * 64 * 16 * 4, b->stream) != hipSuccess ||
      hipMemsetAsync(a.hub_elim, 0, G * std::max<uint64_t>(1, a.nhubs) * 8, b->stream) != hipSuccess ||
      hipMemsetAsync(a.act, 0, G * V * 8, b->stream) != hipSuccess ||
      hipMemsetAsync(a.mod, 0, G * V * 8, b->stream) != hipSuccess ||
      hipMemsetAsync(a.sched, 0, G * V * 8, b->stream) != hipSuccess ||
      hipMemsetAsync(a.pend, 0, G * V * 8, b->stream) != hipSuccess ||
      hipMemsetAsync(a.dirty, 0, G * V * 8, b->stream) != hipSuccess ||
      hipMemsetAsync(a.ea, 0, G * E * 8, b->stream) != hipSuccess ||
      hipMemsetAsync(b->d_tab, 0, L * sizeof(LaneTab), b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess) {
    b->err = "pj_batch_create: clearing the state failed";
    return fail();
  }
  b->tab.assign(L, LaneTab{});
  b->status.assign(L, kFree);
  b->maxneed.assign(L, 0);
  b->rows.assign(L, {});
  b->av.assign(L, 0);
  b->ae.assign(L, 0);
  b->pending.assign(G, 0);
  *out = b;
  return 0;
}

int pj_batch_admit(pj_batch* b, const pj_template_desc* d, uint32_t* lane) {
  if (!b || !d || !lane) return -1;
  std::lock_guard<std::mutex> lk(b->mu);
  const uint32_t P = d->vertex_count;
  if (P == 0 || P > kMaxP) {
    b->err = "pj_batch_admit: the batch takes templates of 1..16 vertices";
    return -2;
  }
  uint32_t L = 0;
  while (L < b->nlanes && b->status[L] != kFree) ++L;
  if (L == b->nlanes) {
    b->err = "pj_batch_admit: no free lane";
    return -1;
  }
  // label union (append only: indices of labels already in use stay)
  std::vector<uint64_t> uni = b->uni;
  auto index_of = [&](uint64_t l) -> int {
    for (size_t i = 0; i < uni.size(); ++i)
      if (uni[i] == l) return static_cast<int>(i);
    if (uni.size() >= kMaxLabels) return -1;
    uni.push_back(l);
    return static_cast<int>(uni.size() - 1);
  };
  LaneTab tb{};
  tb.P = static_cast<uint8_t>(P);
  tb.full = static_cast<Word>(P >= 16 ? 0xFFFFu : (1u << P) - 1u);
  for (uint32_t p = 0; p < P; ++p) {
    const int li = index_of(d->vertex_label[p]);
    if (li < 0) {
      b->err = "pj_batch_admit: the batch's templates carry more than 255 labels";
      return -2;
    }
    tb.labmask[li] |= static_cast<Word>(1u << p);
    for (size_t gi = 0; gi < b->graph_labels.size(); ++gi)
      if (b->graph_labels[gi] == d->vertex_label[p]) tb.glabmask[gi] |= static_cast<Word>(1u << p);
    tb.nbr[p] = static_cast<Word>(d->neighbor_mask[p] & 0xFFFFu);
  }
  uint32_t nent = 0, maxneed = 0;
  for (uint32_t p = 0; p < P; ++p) {
    tb.ent_off[p] = static_cast<uint8_t>(nent);
    for (uint32_t i = d->need_offset[p]; i < d->need_offset[p + 1]; ++i) {
      Word mask = 0;
      for (uint32_t q = 0; q < P; ++q)
        if (d->vertex_label[q] == d->need_label[i]) mask |= static_cast<Word>(1u << q);
      if (!mask) {
        b->err = "pj_batch_admit: neighbour label not carried by any template vertex";
        return -1;
      }
      if (nent >= static_cast<uint32_t>(kMaxEnt)) {
        b->err = "pj_batch_admit: too many requirement entries";
        return -2;
      }
      const uint32_t need = std::min<uint32_t>(d->need_count[i], 127u);
      maxneed = std::max(maxneed, need);
      tb.ent[nent++] = mask | (need << 16);
    }
  }
  tb.ent_off[P] = static_cast<uint8_t>(nent);
  for (uint32_t p = P + 1; p <= kMaxP; ++p) tb.ent_off[p] = static_cast<uint8_t>(nent);
  if (uni.size() != b->uni.size()) {
    b->uni = uni;
    b->labels_dirty = true;
  }
  b->tab[L] = tb;
  PJB_HIP(hipSetDevice(b->device));
  PJB_HIP(hipMemcpyAsync(b->d_tab + L, &b->tab[L], sizeof(LaneTab), hipMemcpyHostToDevice, b->stream));
  PJB_HIP(hipStreamSynchronize(b->stream));
  b->status[L] = kRunning;
  b->maxneed[L] = maxneed;
  b->rows[L].clear();
  b->pending[L / 64] |= 1ull << (L % 64);
  *lane = L;
  return 0;
}

int pj_batch_run(pj_batch* b, uint32_t max_steps, uint32_t* done, uint32_t max_done, uint32_t* n_done) {
  if (!b) return -1;
  // The lock guards the lane table; it is let go while the device works, so
  // that other threads can export and release parked lanes meanwhile.
  // Admission and runs come from one thread.
  std::unique_lock<std::mutex> lk(b->mu);
  if (n_done) *n_done = 0;
  PJB_HIP(hipSetDevice(b->device));
  BArgs& a = b->a;
  if (b->labels_dirty) {
    PJB_HIP(hipMemcpyAsync(b->d_union, b->uni.data(), b->uni.size() * 8, hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(kb_labels, dim3(grid_cap(b, a.V, 8)), dim3(kBlockB), 0, b->stream, b->gv.labels, b->d_union,
                       static_cast<uint32_t>(b->uni.size()), a.V, b->d_blab);
    a.nlab = static_cast<uint32_t>(std::max<size_t>(1, b->uni.size()));
    b->labels_dirty = false;
  }
  for (uint32_t g = 0; g < b->G; ++g) {
    if (!b->pending[g]) continue;
    hipLaunchKernelGGL(kb_admit, dim3(64), dim3(kBlockB), 0, b->stream, a, g, b->pending[g]);
    hipLaunchKernelGGL(kb_ea_set, dim3(grid_cap(b, a.E, 8)), dim3(kBlockB), 0, b->stream, a, g, b->pending[g]);
    b->pending[g] = 0;
  }
  PJB_HIP(hipGetLastError());
  uint32_t nd = 0;
  for (uint32_t s = 0; s < max_steps; ++s) {
    bool any = false;
    uint32_t maxneed = 0;
    for (size_t L = 0; L < b->status.size(); ++L)
      if (b->status[L] == kRunning) {
        any = true;
        maxneed = std::max(maxneed, b->maxneed[L]);
      }
    if (!any) break;
    lk.unlock();
    PJB_HIP(hipEventRecord(b->ev0, b->stream));
    if (launch_step(b, maxneed > 3 ? 4 : maxneed > 1 ? 2 : 1)) return -1;
    PJB_HIP(hipEventRecord(b->ev1, b->stream));
    PJB_HIP(hipMemcpyAsync(b->h_ctl, a.ctl, b->G * kCtl * 8, hipMemcpyDeviceToHost, b->stream));
    PJB_HIP(hipStreamSynchronize(b->stream));
    lk.lock();
    float ms = 0.f;
    PJB_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->ms += ms;
    ++b->steps;
    // lanes that reached their fixed point in this step
    bool parked_any = false;
    std::vector<uint64_t> parked_now(b->G, 0);
    for (uint32_t g = 0; g < b->G; ++g) {
      uint64_t newly = 0;
      for (int t = 0; t < 64; ++t) {
        const uint32_t L = g * 64 + t;
        if (b->status[L] == kRunning && ((b->h_ctl[g * kCtl + kDone] >> t) & 1)) newly |= 1ull << t;
      }
      if (!newly) continue;
      std::vector<uint32_t> st(64);
      PJB_HIP(hipMemcpy(st.data(), a.lane_step + g * 64, 64 * 4, hipMemcpyDeviceToHost));
      for (int t = 0; t < 64; ++t) {
        if (!((newly >> t) & 1)) continue;
        const uint32_t L = g * 64 + t;
        const uint32_t n = st[t] + 1;  // LCC0 and LCC steps 1..st[t]
        std::vector<unsigned long long> tr(static_cast<size_t>(n) * 4);
        PJB_HIP(hipMemcpy(tr.data(), a.trace + static_cast<uint64_t>(L) * kMaxSteps * 4, tr.size() * 8,
                          hipMemcpyDeviceToHost));
        auto& rows = b->rows[L];
        rows.assign(n, pj_algo_results{});
        uint64_t ve = 0, ee = 0;
        for (uint32_t k = 0; k < n; ++k) {
          rows[k].vertex_eliminated = tr[4 * k];
          rows[k].edge_eliminated = tr[4 * k + 1];
          rows[k].match_eliminated = tr[4 * k + 2];
          ve += tr[4 * k];
          ee += tr[4 * k + 1];
        }
        b->av[L] = b->gv.V - ve;
        b->ae[L] = b->gv.E - ee;
      }
      if (park(b, g, newly)) return -1;
      parked_any = true;
      parked_now[g] = newly;
    }
    if (parked_any) {
      // the parked states are complete before any lane is handed out
      lk.unlock();
      PJB_HIP(hipStreamSynchronize(b->stream));
      lk.lock();
      for (uint32_t g = 0; g < b->G; ++g)
        for (int t = 0; t < 64; ++t)
          if ((parked_now[g] >> t) & 1) {
            const uint32_t L = g * 64 + t;
            b->status[L] = kParked;
            if (done && nd < max_done) done[nd] = L;
            ++nd;
          }
    }
    // a lane at the step cap without a fixed point
    if (b->steps % 64 == 0) {
      std::vector<uint32_t> st(b->G * 64);
      PJB_HIP(hipMemcpy(st.data(), a.lane_step, st.size() * 4, hipMemcpyDeviceToHost));
      for (size_t L = 0; L < st.size(); ++L)
        if (b->status[L] == kRunning && st[L] + 1 >= kMaxSteps) {
          b->err = "pj_batch_run: a template reached the step cap without a fixed point";
          return -1;
        }
    }
    if (parked_any) break;
  }
  if (n_done) *n_done = nd;
  return 0;
}

int pj_batch_trace(pj_batch* b, uint32_t lane, pj_algo_results* rows, uint32_t max_rows, uint32_t* n_rows) {
  if (!b || !n_rows) return -1;
  std::lock_guard<std::mutex> lk(b->mu);
  if (lane >= b->status.size() || b->status[lane] != kParked) {
    b->err = "pj_batch_trace: lane is not parked";
    return -1;
  }
  const auto& r = b->rows[lane];
  *n_rows = static_cast<uint32_t>(r.size());
  for (uint32_t k = 0; k < r.size() && k < max_rows && rows; ++k) rows[k] = r[k];
  return 0;
}

int pj_batch_export(pj_batch* b, uint32_t lane, pj_ctx* ctx) {
  if (!b || !ctx) return -1;
  uint64_t av, ae;
  int wb;
  {
    std::lock_guard<std::mutex> lk(b->mu);
    if (lane >= b->status.size() || b->status[lane] != kParked) {
      b->err = "pj_batch_export: lane is not parked";
      return -1;
    }
    av = b->av[lane];
    ae = b->ae[lane];
    wb = b->tab[lane].P <= 8 ? 1 : 2;
  }
  // The park arrays of a parked lane are not written until it is released.
  if (pjb::import_lcc_state(ctx, b->park_cand + lane * 2 * b->vpad, wb, b->park_flags + lane * b->vpad,
                            b->park_ea + lane * b->ea_stride, av, ae, nullptr) != 0) {
    std::lock_guard<std::mutex> lk(b->mu);
    b->err = std::string("pj_batch_export: ") + pj_last_error(ctx);
    return -1;
  }
  return 0;
}

int pj_batch_release(pj_batch* b, uint32_t lane) {
  if (!b) return -1;
  std::lock_guard<std::mutex> lk(b->mu);
  if (lane >= b->status.size() || b->status[lane] != kParked) {
    b->err = "pj_batch_release: lane is not parked";
    return -1;
  }
  b->status[lane] = kFree;
  b->rows[lane].clear();
  return 0;
}

int pj_batch_info(pj_batch* b, uint32_t* lanes, uint32_t* running, uint32_t* parked, double* device_ms,
                  uint64_t* steps) {
  if (!b) return -1;
  std::lock_guard<std::mutex> lk(b->mu);
  uint32_t r = 0, p = 0;
  for (int s : b->status) {
    r += s == kRunning;
    p += s == kParked;
  }
  if (lanes) *lanes = b->nlanes;
  if (running) *running = r;
  if (parked) *parked = p;
  if (device_ms) *device_ms = b->ms;
  if (steps) *steps = b->steps;
  return 0;
}

const char* pj_batch_last_error(pj_batch* b) {
  if (!b) return "null batch";
  return b->err.c_str();
}

void pj_batch_destroy(pj_batch* b) {
  if (!b) return;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  bfree(b);
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

}  // extern "C"
