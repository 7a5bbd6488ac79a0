// ORACLE — test infrastructure only.
//
// CPU restatement of the PruneJuice pruning steps of the reference
// (tahsinreza/pattenmatching_sharedmemory, src/patternmatching/...), used as
// the checker of the HIP path and as the timed CPU baseline.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  It
// is never part of the product path.
//
// State layout follows the reference: uint32 flags per vertex and per edge
// (multiple_label_common_cpu.h:66-80), 64-bit candidate / alreadyMatched /
// toUnmatch sets iterated in ascending order (FixedBitmap<64>,
// utils/bitmap_fixed.h:18-70), per-thread std::unordered_map label counters
// (multiple_label_lcc_cpu.tpp:19-21).
//
// Three modes for the non-local steps (CC, PC, TDS):
//   mode 0 "faithful": the reference's recursive DFS with its per-source
//          edge memo (which stores template ids but tests positions,
//          cc_cpu.tpp:53-61), the TDS hypothesis memo, and the cross-source
//          alreadyMatched shortcut (cc_cpu.tpp:102-108).  With more than one
//          thread the shortcut can make results depend on scheduling,
//          exactly as in the reference.
//   mode 1 "complete": the same walks without memo and without the
//          shortcut, i.e. "does a walk satisfying the constraint exist";
//          on graphs whose rows ascend, a cycle's closing hop is a binary
//          search instead of the reference's scan (same decision, fewer
//          slots read: the slot counts of this mode are not the reference's).
//          The two modes agree whenever the memo never fires on a needed
//          branch (memo-inert constraints: pjh_constraints_memo_inert of the
//          host library decides it for cycle and path constraints) — tests
//          pin that agreement.
//   mode 2 "faithful without the shortcut": memos on, every walk runs.  Its
//          failures F0 and the alreadyMatched marks its successful walks
//          leave (or_last_marks) M0 bound every schedule's failures F of
//          mode 0: F0 minus M0 is contained in F, which is contained in F0
//          (a walk's own outcome does not depend on the marks; the shortcut
//          only skips walks whose root some successful walk marked).
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../include/prunejuice.h"

namespace {

struct PairHash {
  size_t operator()(const std::pair<uint32_t, uint32_t>& p) const {
    return std::hash<uint32_t>()(p.first) ^ std::hash<uint32_t>()(p.second);
  }
};

struct Walk {
  uint32_t root = 0;
  std::vector<uint8_t> move;
  std::vector<uint32_t> vertex, mb;
};

struct Tables {
  uint32_t P = 0;
  std::vector<uint64_t> plabel;
  std::vector<uint64_t> nbr;
  std::vector<std::vector<std::pair<uint64_t, uint32_t>>> need;
  std::vector<std::vector<uint32_t>> cycles, paths;
  std::vector<uint64_t> path_label;
  std::vector<std::vector<Walk>> tds;
  Walk enumeration;
};

inline int ctz64(uint64_t x) { return __builtin_ctzll(x); }

// CSR slots read by the walks of the calling thread: one per iteration of a
// row loop, the TEPS numerator.
thread_local uint64_t tl_slots = 0;

}  // namespace

struct or_state {
  uint32_t V = 0;
  uint64_t E = 0;
  const uint64_t* off = nullptr;
  const uint32_t* edges = nullptr;
  const uint64_t* labels = nullptr;
  Tables t;
  std::vector<uint32_t> active, modified, scheduled, edge_active;
  std::vector<uint64_t> match, already, unmatch;
  std::vector<uint64_t> last_marks;  // alreadyMatched at the end of the last NLCC phase 1
  uint64_t active_v = 0, active_e = 0;
  uint64_t slots = 0;  // slots examined by the last step
  int threads = 1;
  bool sorted_rows = false;  // every row ascending (duplicates allowed)
};

namespace {

// Atomic helpers on the 64-bit sets (FixedBitmap::insertAtomic /
// getBitAtomic, bitmap_fixed.tpp:87-92,156-162).
inline void insert_atomic(uint64_t* w, uint32_t b) { __atomic_fetch_or(w, 1ull << b, __ATOMIC_RELAXED); }
inline bool test_atomic(const uint64_t* w, uint32_t b) { return (__atomic_load_n(w, __ATOMIC_RELAXED) >> b) & 1; }

// ---- LCC0: multiple_label_lcc0_cpu.tpp:13-47 --------------------------------
pj_algo_results lcc0(or_state* s) {
  size_t velim = 0, melim = 0;
#pragma omp parallel for num_threads(s->threads) reduction(+ : velim, melim) schedule(guided)
  for (int64_t v = 0; v < static_cast<int64_t>(s->V); ++v) {
    if (!s->active[v]) continue;
    const uint64_t lab = s->labels[v];
    for (uint64_t m = s->match[v]; m; m &= m - 1) {
      const int p = ctz64(m);
      if (lab != s->t.plabel[p]) {
        ++melim;
        s->match[v] &= ~(1ull << p);
      }
    }
    if (!s->match[v]) {
      s->active[v] = 0;
      ++velim;
    }
  }
  s->active_v -= velim;
  return pj_algo_results{velim, 0, melim, 0};
}

// ---- phase 2 shared by LCC and NLCC -----------------------------------------
// lcc_cpu.tpp:107-142 (scan with break); cc_cpu.tpp:134-174 (same, plus the
// alreadyMatched clear); pc_cpu.tpp:152-189 / tds_cpu.tpp:256-293 (no break,
// same outcome).
void phase2(or_state* s, bool nlcc, size_t* velim_out, size_t* melim_out) {
  size_t velim = 0, melim = 0, slots = 0;
#pragma omp parallel for num_threads(s->threads) reduction(+ : velim, melim, slots) schedule(guided)
  for (int64_t v = 0; v < static_cast<int64_t>(s->V); ++v) {
    if (!s->active[v]) continue;
    if (nlcc) s->already[v] = 0;
    if (s->modified[v]) {
      for (uint64_t m = s->unmatch[v]; m; m &= m - 1) {
        s->match[v] &= ~(1ull << ctz64(m));
        ++melim;
      }
      if (!s->match[v]) {
        s->active[v] = 0;
        ++velim;
      }
      s->scheduled[v] = 1;
      s->unmatch[v] = 0;
    } else {
      bool sched = false;
      for (uint64_t e = s->off[v]; e < s->off[v + 1]; ++e) {
        ++slots;
        if (!s->edge_active[e]) continue;
        if (!s->modified[s->edges[e]]) continue;
        sched = true;
        break;
      }
      s->scheduled[v] = sched;
    }
  }
  s->slots += slots;
  *velim_out = velim;
  *melim_out = melim;
}

// ---- LCC: multiple_label_lcc_cpu.tpp:18-152 ---------------------------------
pj_algo_results lcc(or_state* s) {
  size_t eelim = 0, slots = 0;
  std::unordered_map<uint64_t, size_t> counts;
  counts.reserve(10);
#pragma omp parallel for num_threads(s->threads) reduction(+ : eelim, slots) private(counts) schedule(guided)
  for (int64_t v = 0; v < static_cast<int64_t>(s->V); ++v) {
    if (!s->active[v] || !s->scheduled[v]) continue;
    slots += s->off[v + 1] - s->off[v];
    const uint64_t lab = s->labels[v];
    uint64_t nbr_set = 0;
    for (uint64_t m = s->match[v]; m; m &= m - 1) nbr_set |= s->t.nbr[ctz64(m)];
    for (auto& kv : counts) kv.second = 0;
    for (uint64_t e = s->off[v]; e < s->off[v + 1]; ++e) {
      if (!s->edge_active[e]) continue;
      const uint32_t u = s->edges[e];
      if (!s->active[u]) continue;
      bool possible = false;
      for (uint64_t m = s->match[u]; m; m &= m - 1)
        if (nbr_set >> ctz64(m) & 1) {
          possible = true;
          break;
        }
      if (!possible) {
        ++eelim;
        s->edge_active[e] = 0;
        s->modified[v] = 1;
        continue;
      }
      ++counts[s->labels[u]];
    }
    bool changed = false;
    for (uint64_t m = s->match[v]; m; m &= m - 1) {
      const int p = ctz64(m);
      bool fail = lab != s->t.plabel[p];
      if (!fail)
        for (const auto& nd : s->t.need[p])
          if (counts[nd.first] < nd.second) {
            fail = true;
            break;
          }
      if (fail) {
        s->unmatch[v] |= 1ull << p;
        changed = true;
      }
    }
    if (changed) s->modified[v] = 1;
  }
  s->slots += slots;
  size_t velim = 0, melim = 0;
  phase2(s, false, &velim, &melim);
  s->active_v -= velim;
  s->active_e -= eelim;
  return pj_algo_results{velim, eelim, melim, 0};
}

using EdgeMemo = std::unordered_map<std::pair<uint32_t, uint32_t>, uint64_t, PairHash>;

inline std::pair<uint32_t, uint32_t> ukey(uint32_t a, uint32_t b) {
  return a > b ? std::make_pair(a, b) : std::make_pair(b, a);
}

// Closing hop of the complete search on ascending rows: an active slot
// cur -> src exists (binary search, then every duplicate of src).  Decides
// the same as the reference's linear scan (cc_cpu.tpp:24-34); the faithful
// mode keeps the scan so that its slot counts follow the reference.
bool closes_sorted(const or_state* s, uint32_t cur, uint32_t src) {
  const uint32_t* b = s->edges + s->off[cur];
  const uint32_t* e = s->edges + s->off[cur + 1];
  tl_slots += 1;
  if (!s->active[src]) return false;
  for (const uint32_t* p = std::lower_bound(b, e, src); p != e && *p == src; ++p)
    if (s->edge_active[p - s->edges]) return true;
  return false;
}

// ---- CC walk: multiple_label_cc_cpu.tpp:12-70 -------------------------------
bool cc_walk(or_state* s, const std::vector<uint32_t>& idx, EdgeMemo* memo, bool faithful, uint32_t src, uint32_t cur,
             size_t start, size_t remaining) {
  const size_t L = idx.size();
  if (remaining == 1) {
    if (!faithful && s->sorted_rows) return closes_sorted(s, cur, src);
    for (uint64_t e = s->off[cur]; e < s->off[cur + 1]; ++e) {
      ++tl_slots;
      if (!s->edge_active[e]) continue;
      const uint32_t u = s->edges[e];
      if (!s->active[u]) continue;
      if (u == src) return true;
    }
    return false;
  }
  const size_t cp = (start + L - remaining) % L;
  const size_t np = (start + L - (remaining - 1)) % L;
  const uint32_t q = idx[np];
  for (uint64_t e = s->off[cur]; e < s->off[cur + 1]; ++e) {
    ++tl_slots;
    if (!s->edge_active[e]) continue;
    const uint32_t u = s->edges[e];
    if (!s->active[u]) continue;
    if (!(s->match[u] >> q & 1)) continue;
    if (faithful) {
      const auto k = ukey(cur, u);
      auto it = memo->find(k);
      if (it != memo->end() && (it->second >> cp & 1)) continue;
      (*memo)[k] |= 1ull << q;
    }
    if (cc_walk(s, idx, memo, faithful, src, u, start, remaining - 1)) {
      if (faithful) insert_atomic(&s->already[u], q);
      return true;
    }
  }
  return false;
}

// ---- PC walk: multiple_label_pc_cpu.tpp:11-75 -------------------------------
bool pc_walk(or_state* s, const std::vector<uint32_t>& idx, bool reverse, EdgeMemo* memo, bool faithful, uint32_t src,
             uint32_t cur, size_t remaining) {
  const size_t L = idx.size();
  const size_t cp = reverse ? remaining : (L - remaining - 1);
  const size_t np = reverse ? remaining - 1 : (L - remaining);
  const uint32_t q = idx[np];
  if (remaining == 1) {
    for (uint64_t e = s->off[cur]; e < s->off[cur + 1]; ++e) {
      ++tl_slots;
      if (!s->edge_active[e]) continue;
      const uint32_t u = s->edges[e];
      if (!s->active[u]) continue;
      if (!(s->match[u] >> q & 1)) continue;
      if (u != src) {
        if (faithful) insert_atomic(&s->already[u], q);
        return true;
      }
    }
    return false;
  }
  for (uint64_t e = s->off[cur]; e < s->off[cur + 1]; ++e) {
    ++tl_slots;
    if (!s->edge_active[e]) continue;
    const uint32_t u = s->edges[e];
    if (!s->active[u]) continue;
    if (!(s->match[u] >> q & 1)) continue;
    if (faithful) {
      const auto k = ukey(cur, u);
      auto it = memo->find(k);
      if (it != memo->end() && (it->second >> cp & 1)) continue;
      (*memo)[k] |= 1ull << q;
    }
    if (pc_walk(s, idx, reverse, memo, faithful, src, u, remaining - 1)) return true;
  }
  return false;
}

// ---- TDS walk: multiple_label_tds_cpu.tpp:13-127 ----------------------------
struct TdsMemo {
  std::unordered_map<uint32_t, std::unordered_map<uint32_t, std::vector<std::vector<uint32_t>>>> hyp;
  EdgeMemo edge;
};

bool tds_walk(or_state* s, const Walk& w, TdsMemo* memo, bool faithful, std::vector<uint32_t>* hist,
              std::vector<uint32_t>* hyp, uint32_t cur, size_t pos) {
  if (pos == w.move.size()) return true;
  const uint8_t mv = w.move[pos];
  if (mv == PJ_MOVE_VERTEX_STORE || mv == PJ_MOVE_VERTEX_NO_STORE) {
    const uint32_t x = w.vertex[pos - 1];
    const uint32_t y = w.vertex[pos];
    for (uint64_t e = s->off[cur]; e < s->off[cur + 1]; ++e) {
      ++tl_slots;
      if (!s->edge_active[e]) continue;
      const uint32_t u = s->edges[e];
      if (!s->active[u]) continue;
      if (!(s->match[u] >> y & 1)) continue;
      const auto k = ukey(cur, u);
      if (faithful) {
        auto it = memo->edge.find(k);
        if (it != memo->edge.end() && (it->second >> x & 1)) continue;
        auto hu = memo->hyp.find(u);
        if (hu != memo->hyp.end()) {
          auto hy = hu->second.find(y);
          if (hy != hu->second.end() && std::find(hy->second.begin(), hy->second.end(), *hyp) != hy->second.end())
            continue;
        }
      }
      if (mv == PJ_MOVE_VERTEX_STORE) hyp->push_back(u);
      if (faithful) memo->hyp[u][y].push_back(*hyp);
      hist->push_back(u);
      if (faithful) memo->edge[k] |= 1ull << y;
      if (tds_walk(s, w, memo, faithful, hist, hyp, u, pos + 1)) {
        if (faithful) insert_atomic(&s->already[u], y);
        return true;
      }
      hist->pop_back();
      if (mv == PJ_MOVE_VERTEX_STORE) hyp->pop_back();
    }
  } else if (mv == PJ_MOVE_CHECK) {
    const uint32_t target = (*hist)[w.mb[pos]];
    for (uint64_t e = s->off[cur]; e < s->off[cur + 1]; ++e) {
      ++tl_slots;
      if (!s->edge_active[e]) continue;
      const uint32_t u = s->edges[e];
      if (!s->active[u]) continue;
      if (u != target) continue;
      if (tds_walk(s, w, memo, faithful, hist, hyp, cur, pos + 1)) return true;
    }
  } else {
    const uint32_t next = (*hist)[w.mb[pos]];
    if (tds_walk(s, w, memo, faithful, hist, hyp, next, pos + 1)) return true;
  }
  return false;
}

// Phase 1 of the three NLCC steps (cc_cpu.tpp:94-129, pc_cpu.tpp:95-147,
// tds_cpu.tpp:153-188), then phase 2.
pj_algo_results nlcc(or_state* s, int step, uint32_t index, int mode) {
  const bool faithful = mode == 0 || mode == 2;
  const bool shortcut = mode == 0;
  uint64_t slots = 0;
#pragma omp parallel for num_threads(s->threads) reduction(+ : slots) schedule(guided)
  for (int64_t vi = 0; vi < static_cast<int64_t>(s->V); ++vi) {
    const uint32_t v = static_cast<uint32_t>(vi);
    if (!s->active[v]) continue;
    const uint64_t before = tl_slots;
    bool changed = false;
    if (step == PJ_STEP_CC) {
      const auto& idx = s->t.cycles[index];
      EdgeMemo memo;
      for (size_t j = 0; j < idx.size(); ++j) {
        const uint32_t q = idx[j];
        if (!(s->match[v] >> q & 1)) continue;
        if (shortcut && test_atomic(&s->already[v], q)) continue;
        memo.clear();
        if (!cc_walk(s, idx, &memo, faithful, v, v, j, idx.size())) {
          s->unmatch[v] |= 1ull << q;
          changed = true;
        }
      }
    } else if (step == PJ_STEP_PC) {
      const auto& idx = s->t.paths[index];
      if (s->labels[v] != s->t.path_label[index]) continue;
      EdgeMemo memo;
      const uint32_t ends[2] = {idx.front(), idx.back()};
      for (int r = 0; r < 2; ++r) {
        const uint32_t q = ends[r];
        if (!(s->match[v] >> q & 1)) continue;
        if (shortcut && test_atomic(&s->already[v], q)) continue;
        memo.clear();
        if (!pc_walk(s, idx, r == 1, &memo, faithful, v, v, idx.size() - 1)) {
          s->unmatch[v] |= 1ull << q;
          changed = true;
        }
      }
    } else {
      const auto& walks = s->t.tds[index];
      TdsMemo memo;
      std::vector<uint32_t> hist, hyp;
      for (const auto& w : walks) {
        if (!(s->match[v] >> w.root & 1)) continue;
        if (shortcut && test_atomic(&s->already[v], w.root)) continue;
        memo.hyp.clear();
        memo.edge.clear();
        hist.assign(1, v);
        hyp.clear();
        if (!tds_walk(s, w, &memo, faithful, &hist, &hyp, v, 1)) {
          s->unmatch[v] |= 1ull << w.root;
          changed = true;
        }
      }
    }
    if (changed) s->modified[v] = 1;
    slots += tl_slots - before;
  }
  s->slots += slots;
  s->last_marks = s->already;
  size_t velim = 0, melim = 0;
  phase2(s, true, &velim, &melim);
  s->active_v -= velim;
  return pj_algo_results{velim, 0, melim, 0};
}

// ---- enumeration: multiple_label_enumeration_cpu.tpp:38-174 ------------------
uint64_t enum_walk(or_state* s, const Walk& w, std::vector<uint32_t>* hist, uint32_t cur, size_t pos) {
  if (pos == w.move.size()) return 1;
  const uint8_t mv = w.move[pos];
  uint64_t total = 0;
  if (mv == PJ_MOVE_VERTEX_STORE || mv == PJ_MOVE_VERTEX_NO_STORE) {
    const uint32_t y = w.vertex[pos];
    for (uint64_t e = s->off[cur]; e < s->off[cur + 1]; ++e) {
      ++tl_slots;
      if (!s->edge_active[e]) continue;
      const uint32_t u = s->edges[e];
      if (!s->active[u]) continue;
      if (!(s->match[u] >> y & 1)) continue;
      if (std::find(hist->begin(), hist->end(), u) != hist->end()) continue;
      hist->push_back(u);
      total += enum_walk(s, w, hist, u, pos + 1);
      hist->pop_back();
    }
  } else if (mv == PJ_MOVE_CHECK) {
    const uint32_t target = (*hist)[w.mb[pos]];
    for (uint64_t e = s->off[cur]; e < s->off[cur + 1]; ++e) {
      ++tl_slots;
      if (!s->edge_active[e]) continue;
      const uint32_t u = s->edges[e];
      if (!s->active[u]) continue;
      if (u != target) continue;
      return enum_walk(s, w, hist, cur, pos + 1);
    }
  } else {
    return enum_walk(s, w, hist, (*hist)[w.mb[pos]], pos + 1);
  }
  return total;
}

pj_algo_results enumerate(or_state* s) {
  const Walk& w = s->t.enumeration;
  uint64_t total = 0, slots = 0;
#pragma omp parallel for num_threads(s->threads) reduction(+ : total, slots) schedule(dynamic, 1000)
  for (int64_t v = 0; v < static_cast<int64_t>(s->V); ++v) {
    if (!s->active[v]) continue;
    if (!(s->match[v] >> w.root & 1)) continue;
    const uint64_t before = tl_slots;
    std::vector<uint32_t> hist{static_cast<uint32_t>(v)};
    total += enum_walk(s, w, &hist, static_cast<uint32_t>(v), 1);
    slots += tl_slots - before;
  }
  s->slots += slots;
  return pj_algo_results{0, 0, 0, total};
}

}  // namespace

extern "C" {

typedef struct or_row {
  int32_t iteration;
  int32_t step;
  uint64_t vertex_eliminated, edge_eliminated, match_eliminated, enumeration;
  uint64_t active_vertices, active_edges;
  uint64_t slots;
  double ms;
} or_row;

or_state* or_create(void) { return new or_state(); }
void or_destroy(or_state* s) { delete s; }

void or_set_threads(or_state* s, int threads) { s->threads = threads > 0 ? threads : omp_get_max_threads(); }

// Keeps the caller's graph arrays (caller owns them for the state's lifetime).
int or_set_graph(or_state* s, uint32_t V, uint64_t E, const uint64_t* off, const uint32_t* edges,
                 const uint64_t* labels) {
  s->V = V;
  s->E = E;
  s->off = off;
  s->edges = edges;
  s->labels = labels;
  s->sorted_rows = true;
  for (uint32_t v = 0; v < V && s->sorted_rows; ++v)
    for (uint64_t e = off[v] + 1; e < off[v + 1]; ++e)
      if (edges[e - 1] > edges[e]) {
        s->sorted_rows = false;
        break;
      }
  return 0;
}

int or_set_template(or_state* s, const pj_template_desc* d) {
  Tables t;
  t.P = d->vertex_count;
  t.plabel.assign(d->vertex_label, d->vertex_label + t.P);
  t.nbr.assign(d->neighbor_mask, d->neighbor_mask + t.P);
  t.need.resize(t.P);
  for (uint32_t p = 0; p < t.P; ++p)
    for (uint32_t i = d->need_offset[p]; i < d->need_offset[p + 1]; ++i)
      t.need[p].emplace_back(d->need_label[i], d->need_count[i]);
  for (uint32_t c = 0; c < d->cycle_count; ++c)
    t.cycles.emplace_back(d->cycle_vertex + d->cycle_offset[c], d->cycle_vertex + d->cycle_offset[c + 1]);
  for (uint32_t c = 0; c < d->path_count; ++c) {
    t.paths.emplace_back(d->path_vertex + d->path_offset[c], d->path_vertex + d->path_offset[c + 1]);
    t.path_label.push_back(d->path_initial_label[c]);
  }
  auto copy = [](const pj_walk_desc& w) {
    Walk o;
    o.root = w.root;
    o.move.assign(w.move, w.move + w.length);
    o.vertex.assign(w.vertex, w.vertex + w.length);
    o.mb.assign(w.move_back, w.move_back + w.length);
    return o;
  };
  for (uint32_t c = 0; c < d->tds_count; ++c) {
    std::vector<Walk> ws;
    for (uint32_t k = d->tds_walk_offset[c]; k < d->tds_walk_offset[c + 1]; ++k) ws.push_back(copy(d->walks[k]));
    t.tds.push_back(std::move(ws));
  }
  t.enumeration = copy(d->enumeration_walk);
  s->t = std::move(t);
  return 0;
}

// MultipleLabelGlobalState::allocate (multiple_label_common_cpu.tpp:29-66).
void or_reset(or_state* s) {
  s->active.assign(s->V, 1);
  s->modified.assign(s->V, 0);
  s->scheduled.assign(s->V, 1);
  s->edge_active.assign(s->E, 1);
  const uint64_t full = s->t.P >= 64 ? ~0ull : ((1ull << s->t.P) - 1);
  s->match.assign(s->V, full);
  s->already.assign(s->V, 0);
  s->unmatch.assign(s->V, 0);
  s->active_v = s->V;
  s->active_e = s->E;
}

uint64_t or_last_slots(const or_state* s) { return s->slots; }

// alreadyMatched as the last NLCC step's phase 1 left it (before phase 2
// clears it); zeros before any NLCC step.
void or_last_marks(const or_state* s, uint64_t* out) {
  for (uint32_t v = 0; v < s->V; ++v) out[v] = v < s->last_marks.size() ? s->last_marks[v] : 0;
}

void or_reset_modified(or_state* s) { std::fill(s->modified.begin(), s->modified.end(), 0u); }

int or_run_step(or_state* s, int step, uint32_t index, int mode, pj_algo_results* out) {
  pj_algo_results r{0, 0, 0, 0};
  s->slots = 0;
  switch (step) {
    case PJ_STEP_LCC0:
      r = lcc0(s);
      break;
    case PJ_STEP_LCC:
      r = lcc(s);
      break;
    case PJ_STEP_CC:
      if (index >= s->t.cycles.size()) return -1;
      r = nlcc(s, step, index, mode);
      break;
    case PJ_STEP_PC:
      if (index >= s->t.paths.size()) return -1;
      r = nlcc(s, step, index, mode);
      break;
    case PJ_STEP_TDS:
      if (index >= s->t.tds.size()) return -1;
      r = nlcc(s, step, index, mode);
      break;
    case PJ_STEP_ENUMERATION:
      r = enumerate(s);
      break;
    default:
      return -1;
  }
  if (out) *out = r;
  return 0;
}

void or_get_state(const or_state* s, uint64_t* cand, uint8_t* vflags, uint32_t* ea_words, uint64_t* av,
                  uint64_t* ae) {
  for (uint32_t v = 0; v < s->V; ++v) {
    if (cand) cand[v] = s->match[v];
    if (vflags)
      vflags[v] = static_cast<uint8_t>((s->active[v] ? 1 : 0) | (s->modified[v] ? 2 : 0) | (s->scheduled[v] ? 4 : 0));
  }
  if (ea_words) {
    std::memset(ea_words, 0, ((s->E + 31) / 32) * 4);
    for (uint64_t e = 0; e < s->E; ++e)
      if (s->edge_active[e]) ea_words[e >> 5] |= 1u << (e & 31);
  }
  if (av) *av = s->active_v;
  if (ae) *ae = s->active_e;
}

void or_set_state(or_state* s, const uint64_t* cand, const uint8_t* vflags, const uint32_t* ea_words, uint64_t av,
                  uint64_t ae) {
  for (uint32_t v = 0; v < s->V; ++v) {
    s->match[v] = cand[v];
    s->active[v] = vflags[v] & 1;
    s->modified[v] = (vflags[v] >> 1) & 1;
    s->scheduled[v] = (vflags[v] >> 2) & 1;
    s->unmatch[v] = 0;
    s->already[v] = 0;
  }
  for (uint64_t e = 0; e < s->E; ++e) s->edge_active[e] = (ea_words[e >> 5] >> (e & 31)) & 1;
  s->active_v = av;
  s->active_e = ae;
}

// The step machine of MultipleLabelStep::getNextStep (multiple_label_step.cc:30-82)
// driven as MultipleLabelCpu::runPatternMatching does (multiple_label_cpu.cc:113-185):
// START -> LCC0 -> LCC until a step eliminates nothing -> next CC, else next
// PC, else TDS (always constraint 0: the reference never advances the TDS
// iterator, tds_cpu.tpp:135-244) -> LCC ... -> ENUMERATION -> END, at most
// 1000 steps.  Writes one row per step; returns the number of rows.
// flags: bit 0 stops before the enumeration step, bit 1 before the first TDS
// step (as PJH_STOP_BEFORE_ENUMERATION / PJH_STOP_BEFORE_TDS).
int or_run_pipeline(or_state* s, int mode, int flags, or_row* rows, int max_rows, double* total_ms) {
  or_reset(s);
  int n_cc = static_cast<int>(s->t.cycles.size()), n_pc = static_cast<int>(s->t.paths.size()),
      n_tds = static_cast<int>(s->t.tds.size());
  int cc = 0, pc = 0, tds = 0;
  int step = PJ_STEP_LCC0;
  int n = 0;
  double total = 0.0;
  for (int it = 0; it < 1000; ++it) {
    uint32_t index = 0;
    if ((flags & 1) && step == PJ_STEP_ENUMERATION) break;
    if ((flags & 2) && step == PJ_STEP_TDS) break;
    if (step == PJ_STEP_CC || step == PJ_STEP_PC || step == PJ_STEP_TDS) or_reset_modified(s);
    if (step == PJ_STEP_CC) index = cc;
    if (step == PJ_STEP_PC) index = pc;
    pj_algo_results r{0, 0, 0, 0};
    auto t0 = std::chrono::steady_clock::now();
    if (or_run_step(s, step, index, mode, &r)) return -1;
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    total += ms;
    if (n < max_rows) {
      rows[n] = or_row{it, step, r.vertex_eliminated, r.edge_eliminated, r.match_eliminated, r.enumeration,
                       s->active_v, s->active_e, s->slots, ms};
    }
    ++n;
    const bool empty = !r.vertex_eliminated && !r.edge_eliminated && !r.match_eliminated && !r.enumeration;
    int next = step;
    switch (step) {
      case PJ_STEP_LCC0:
        next = PJ_STEP_LCC;
        break;
      case PJ_STEP_LCC:
        if (empty) {
          if (cc < n_cc)
            next = PJ_STEP_CC;
          else if (pc < n_pc)
            next = PJ_STEP_PC;
          else if (tds < n_tds)
            next = PJ_STEP_TDS;
          else
            next = PJ_STEP_ENUMERATION;
        }
        break;
      case PJ_STEP_CC:
        ++cc;
        next = PJ_STEP_LCC;
        break;
      case PJ_STEP_PC:
        ++pc;
        next = PJ_STEP_LCC;
        break;
      case PJ_STEP_TDS:
        ++tds;
        next = PJ_STEP_LCC;
        break;
      case PJ_STEP_ENUMERATION:
        next = PJ_STEP_END;
        break;
    }
    if (next == PJ_STEP_END) break;
    step = next;
  }
  if (total_ms) *total_ms = total;
  return n;
}

}  // extern "C"
