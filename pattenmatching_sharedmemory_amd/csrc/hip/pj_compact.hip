// Active-graph compaction (see pj_compact.h).
//
// A vertex survives when its candidate set is not empty (the reference's
// vertexActiveList, multiple_label_common_cpu.h:66-80: a vertex is
// deactivated exactly when its set empties); a slot v->u survives when its
// edge-active bit is set and both v and u survive (printActiveGraph,
// multiple_label_cpu.cc:270-284).  Three passes, bandwidth-bound:
//   count  one block per 2^16 vertices or slots counts its survivors (one
//          ballot per 64 items) into its own scratch counter;
//   scan   one block turns the counters into exclusive offsets and totals;
//   write  each block writes its survivors from its offset, in order
//          (ballot prefix within a wave, wave prefix within the block).
// A slot's source is found by a binary search of the offsets between the
// sources of the block's first and last slot.  Counters are written, never
// accumulated, so the scratch needs no clearing between calls.
#include "pj_compact.h"

#include <mutex>
#include <vector>

namespace pjc {

// The no-allocation entry point's counters: one array per device, used
// through the address hipGetSymbolAddress returns for it (in host code the
// name is only a stand-in); held until the process ends.  Calls that
// overlap on one device would share it: pj_compact_active_noalloc takes a
// lock for the whole call.
__device__ __align__(256) unsigned long long scratch[kStaticBlocks * 2 + 2];

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kScanThreads = 1024;

template <typename Word>
__device__ __forceinline__ bool alive(const Word* cand, uint64_t v) {
  return cand[v] != Word(0);
}

// Largest v in [lo, hi] with off[v] <= e (off non-decreasing, off[lo] <= e).
__device__ __forceinline__ uint32_t source_of(const uint64_t* __restrict__ off, uint32_t lo, uint32_t hi, uint64_t e) {
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo + 1) / 2;
    if (off[mid] <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ uint32_t source_of_all(const uint64_t* __restrict__ off, uint32_t V, uint64_t e) {
  return source_of(off, 0, V - 1, e);
}

struct Ranges {
  uint64_t vblocks, eblocks;
};

// Survivor of item i of the block (vertex or slot), with the slot's source.
template <typename Word>
__device__ __forceinline__ bool survivor(const CompactArgs& a, bool vertex_block, uint64_t i, uint32_t s_lo,
                                         uint32_t s_hi, uint32_t* src_out) {
  if (vertex_block) return alive<Word>(static_cast<const Word*>(a.cand), i);
  const Word* cand = static_cast<const Word*>(a.cand);
  const bool on = (a.ea[i >> 5] >> (i & 31)) & 1u;
  const uint32_t u = a.edges[i];
  const uint32_t s = source_of(a.off, s_lo, s_hi, i);
  *src_out = s;
  return on && alive<Word>(cand, s) && alive<Word>(cand, u);
}

template <typename Word, bool kWrite>
__global__ void __launch_bounds__(kThreads) k_compact(CompactArgs a, Ranges r, unsigned long long* __restrict__ sc) {
  __shared__ uint32_t s_src[2];
  __shared__ uint32_t s_wave[kWaves];
  const uint64_t b = blockIdx.x;
  const bool vb = b < r.vblocks;
  const uint64_t n = vb ? a.V : a.E;
  const uint64_t lo = (vb ? b : b - r.vblocks) * kSpan;
  const uint64_t hi = lo + kSpan < n ? lo + kSpan : n;
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  if (!vb && threadIdx.x < 2 && lo < hi) s_src[threadIdx.x] = source_of_all(a.off, a.V, threadIdx.x ? hi - 1 : lo);
  __syncthreads();
  const uint32_t s_lo = vb ? 0u : s_src[0], s_hi = vb ? 0u : s_src[1];
  unsigned long long pos = kWrite ? sc[b] : 0ull;
  uint64_t count = 0;
  for (uint64_t c0 = lo; c0 < hi; c0 += kThreads) {
    const uint64_t i = c0 + threadIdx.x;
    uint32_t s = 0;
    const bool keep = i < hi && survivor<Word>(a, vb, i, s_lo, s_hi, &s);
    const uint64_t m = __ballot(keep);
    if (!kWrite) {
      count += __popcll(m);  // identical in every lane
      continue;
    }
    if (lane == 0) s_wave[wave] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    uint64_t before = 0, total = 0;
    for (int w = 0; w < kWaves; ++w) {
      before += w < wave ? s_wave[w] : 0u;
      total += s_wave[w];
    }
    if (keep) {
      const uint64_t k = pos + before + __popcll(m & ((1ull << lane) - 1));
      if (vb) {
        if (k < a.cap_v) {
          a.vertices[k] = static_cast<uint32_t>(i);
          if (a.cand_out) a.cand_out[k] = static_cast<uint64_t>(static_cast<const Word*>(a.cand)[i]);
        }
      } else if (k < a.cap_e) {
        a.src[k] = s;
        a.dst[k] = a.edges[i];
      }
    }
    pos += total;
    __syncthreads();
  }
  if (!kWrite) {
    if (lane == 0) s_wave[wave] = static_cast<uint32_t>(count);
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long t = 0;
      for (int w = 0; w < kWaves; ++w) t += s_wave[w];
      sc[b] = t;
    }
  }
}

// Exclusive scan of counters [0, nv) and [nv, nv + ne) in place; totals go to
// sc[nv + ne] and sc[nv + ne + 1].  One block.
__global__ void __launch_bounds__(kScanThreads) k_scan(unsigned long long* __restrict__ sc, uint64_t nv, uint64_t ne) {
  __shared__ unsigned long long s_part[kScanThreads];
  for (int part = 0; part < 2; ++part) {
    unsigned long long* p = sc + (part ? nv : 0);
    const uint64_t n = part ? ne : nv;
    const uint64_t seg = (n + kScanThreads - 1) / kScanThreads;
    const uint64_t a = threadIdx.x * seg, z = a + seg < n ? a + seg : n;
    unsigned long long sum = 0;
    for (uint64_t i = a; i < z; ++i) sum += p[i];
    s_part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long run = 0;
      for (int t = 0; t < kScanThreads; ++t) {
        const unsigned long long x = s_part[t];
        s_part[t] = run;
        run += x;
      }
      sc[nv + ne + part] = run;
    }
    __syncthreads();
    unsigned long long run = s_part[threadIdx.x];
    for (uint64_t i = a; i < z; ++i) {
      const unsigned long long x = p[i];
      p[i] = run;
      run += x;
    }
    __syncthreads();
  }
}

template <typename Word>
int run(const CompactArgs& a, unsigned long long* sc, uint64_t* n_v, uint64_t* n_e, std::string* err) {
  const Ranges r{(static_cast<uint64_t>(a.V) + kSpan - 1) / kSpan, (a.E + kSpan - 1) / kSpan};
  const uint64_t blocks = r.vblocks + r.eblocks;
  unsigned long long tot[2] = {0, 0};
  if (blocks) {
    hipLaunchKernelGGL((k_compact<Word, false>), dim3(blocks), dim3(kThreads), 0, a.stream, a, r, sc);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, a.stream, sc, r.vblocks, r.eblocks);
    hipLaunchKernelGGL((k_compact<Word, true>), dim3(blocks), dim3(kThreads), 0, a.stream, a, r, sc);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(tot, sc + blocks, sizeof(tot), hipMemcpyDeviceToHost, a.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(a.stream);
    if (e != hipSuccess) {
      *err = std::string("active-graph compaction: ") + hipGetErrorString(e);
      return -1;
    }
  }
  *n_v = tot[0];
  *n_e = tot[1];
  if (tot[0] > a.cap_v || tot[1] > a.cap_e) {
    *err = "active-graph compaction: output capacity too small (" + std::to_string(tot[0]) + " vertices, " +
           std::to_string(tot[1]) + " slots survive)";
    return -1;
  }
  return 0;
}

}  // namespace

uint64_t scratch_words(uint32_t V, uint64_t E) {
  return (static_cast<uint64_t>(V) + kSpan - 1) / kSpan + (E + kSpan - 1) / kSpan + 2;
}

int compact_active(const CompactArgs& a, unsigned long long* sc, uint64_t* n_v, uint64_t* n_e, std::string* err) {
  if (a.V == 0) {
    *n_v = *n_e = 0;
    return 0;
  }
  switch (a.word_bytes) {
    case 1:
      return run<uint8_t>(a, sc, n_v, n_e, err);
    case 2:
      return run<uint16_t>(a, sc, n_v, n_e, err);
    case 4:
      return run<uint32_t>(a, sc, n_v, n_e, err);
    default:
      return run<uint64_t>(a, sc, n_v, n_e, err);
  }
}

int static_scratch(int device, unsigned long long** ptr, std::string* err) {
  static std::mutex mu;
  static std::vector<unsigned long long*> by_device;
  std::lock_guard<std::mutex> lk(mu);
  if (device < 0) {
    *err = "bad device";
    return -1;
  }
  if (by_device.size() <= static_cast<size_t>(device)) by_device.resize(device + 1, nullptr);
  if (!by_device[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the current device's copy
    if (e != hipSuccess || !p) {
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return -1;
    }
    by_device[device] = static_cast<unsigned long long*>(p);
  }
  *ptr = by_device[device];
  return 0;
}

}  // namespace pjc
