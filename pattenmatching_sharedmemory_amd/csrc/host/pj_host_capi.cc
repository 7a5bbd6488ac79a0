// extern "C" surface of the host library (include/prunejuice_host.h).
#include <getopt.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <deque>
#include <cstring>
#include <iostream>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/prunejuice_host.h"
#include "pj_constraints.h"
#include "pj_driver.h"
#include "pj_effectiveness.h"
#include "pj_graph.h"

struct pjh_graph {
  pj::Graph g;
};
struct pjh_constraints {
  pj::ConstraintSet c;
};

namespace {

void put_err(const std::string& msg, char* err, size_t len) {
  if (!err || !len) return;
  std::strncpy(err, msg.c_str(), len - 1);
  err[len - 1] = 0;
}

// Whether the reference's per-source edge memo can fire for a constraint.
// The memo maps an undirected edge to template ids entered over it and is
// tested with the current *position* (cc_cpu.tpp:53-61, pc_cpu.tpp:59-67).
// A skip of c@cp -> u@np needs an earlier traversal c'@cp' -> w@np' over the
// same edge with idx[np'] == cp, where {c', w} == {c, u}.  After LCC0 a
// graph vertex can stand at position i only if its label is the label of
// template vertex idx[i], so the skip is possible only if those label
// equalities hold for some pair of memo transitions of one search.
bool memo_may_fire(const std::vector<uint32_t>& idx, const std::vector<uint64_t>& plab,
                   const std::vector<std::pair<size_t, size_t>>& transitions) {
  for (const auto& cur : transitions)
    for (const auto& prev : transitions) {
      const size_t cp = cur.first, np = cur.second, cpp = prev.first, npp = prev.second;
      if (idx[npp] != cp) continue;
      const uint64_t l_c = plab[idx[cp]], l_u = plab[idx[np]];
      const uint64_t l_prev_from = plab[idx[cpp]], l_prev_to = plab[idx[npp]];
      // same orientation: c' = c, w = u
      if (l_prev_from == l_c && l_prev_to == l_u) return true;
      // opposite orientation: c' = u, w = c
      if (l_prev_from == l_u && l_prev_to == l_c) return true;
    }
  return false;
}

}  // namespace

extern "C" {

int pjh_graph_load(const char* path, const char* vertex_file, pjh_graph** out, char* err, size_t errlen) {
  if (!path || !out) return -1;
  auto* h = new pjh_graph();
  std::string e;
  if (pj::load_totem(path, &h->g, &e) || (vertex_file && *vertex_file && pj::parse_vertex_file(vertex_file, &h->g, &e))) {
    put_err(e, err, errlen);
    delete h;
    return -1;
  }
  *out = h;
  return 0;
}

int pjh_graph_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, pjh_graph** out) {
  if (!out || scale < 1 || scale > 31 || edge_factor < 1) return -1;
  auto* h = new pjh_graph();
  pj::RmatParams p;
  p.scale = scale;
  p.edge_factor = edge_factor;
  p.a = a;
  p.b = b;
  p.c = c;
  p.seed = seed;
  pj::rmat_generate(p, &h->g);
  *out = h;
  return 0;
}

int pjh_graph_from_csr(uint32_t V, uint64_t E, const uint64_t* off, const uint32_t* edges, const uint64_t* labels,
                       pjh_graph** out) {
  if (!out || !off || (E && !edges)) return -1;
  auto* h = new pjh_graph();
  h->g.vertex_count = V;
  h->g.edge_count = E;
  h->g.vertices.assign(off, off + V + 1);
  h->g.edges.assign(edges, edges + E);
  if (labels) {
    h->g.valued = true;
    h->g.values.assign(labels, labels + V);
  }
  h->g.directed = false;
  std::string err;
  if (pj::validate_csr(h->g, &err)) {
    fprintf(stderr, "pjh_graph_from_csr: %s\n", err.c_str());
    delete h;
    return -1;
  }
  *out = h;
  return 0;
}

int pjh_pattern_from_edges(uint32_t V, uint32_t n, const uint32_t* src, const uint32_t* dst, const uint64_t* labels,
                           pjh_graph** out) {
  if (!out || !labels || (n && (!src || !dst))) return -1;
  for (uint32_t i = 0; i < n; ++i)
    if (src[i] >= V || dst[i] >= V) return -1;
  auto* h = new pjh_graph();
  pj::pattern_from_edges(V, std::vector<uint32_t>(src, src + n), std::vector<uint32_t>(dst, dst + n),
                         std::vector<uint64_t>(labels, labels + V), &h->g);
  *out = h;
  return 0;
}

void pjh_graph_labels_uniform(pjh_graph* g, uint32_t n, uint64_t seed) { pj::labels_uniform(&g->g, n, seed); }
void pjh_graph_labels_zipf(pjh_graph* g, uint32_t n, double s, uint64_t seed) { pj::labels_zipf(&g->g, n, s, seed); }
uint32_t pjh_graph_vertex_count(const pjh_graph* g) { return g->g.vertex_count; }
uint64_t pjh_graph_edge_count(const pjh_graph* g) { return g->g.edge_count; }
const uint64_t* pjh_graph_offsets(const pjh_graph* g) { return g->g.vertices.data(); }
const uint32_t* pjh_graph_edges(const pjh_graph* g) { return g->g.edges.data(); }
const uint64_t* pjh_graph_labels(const pjh_graph* g) { return g->g.valued ? g->g.values.data() : nullptr; }

int pjh_graph_store_binary(const pjh_graph* g, const char* path, int eid_bytes) {
  std::string e;
  return pj::store_totem_binary(g->g, path, eid_bytes, &e);
}
int pjh_graph_store_text(const pjh_graph* g, const char* path) {
  std::string e;
  return pj::store_totem_text(g->g, path, &e);
}
int pjh_graph_store_vertex_file(const pjh_graph* g, const char* path) {
  std::string e;
  return pj::store_vertex_file(g->g, path, &e);
}
uint64_t pjh_graph_asymmetric_slots(const pjh_graph* g) { return pj::count_asymmetric_slots(g->g); }
void pjh_graph_free(pjh_graph* g) { delete g; }

int pjh_constraints_build(const pjh_graph* pattern, pjh_constraints** out, char* err, size_t errlen) {
  if (!pattern || !out) return -1;
  auto* c = new pjh_constraints();
  std::string e;
  if (c->c.build(pattern->g, &e)) {
    put_err(e, err, errlen);
    delete c;
    return -1;
  }
  *out = c;
  return 0;
}

const pj_template_desc* pjh_constraints_desc(const pjh_constraints* c) { return c->c.desc(); }

size_t pjh_constraints_dump(const pjh_constraints* c, char* buf, size_t len) {
  const std::string s = c->c.dump();
  if (buf && len) {
    std::strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return s.size();
}

int pjh_constraints_memo_inert(const pjh_constraints* c, int step, uint32_t index) {
  const auto& cs = c->c;
  std::vector<uint64_t> plab = cs.vertex_label;
  if (step == PJ_STEP_CC) {
    if (index >= cs.cycles.size()) return -1;
    const auto& idx = cs.cycles[index].vertex;
    const size_t L = idx.size();
    // One search per start position j; memo levels move j+d -> j+d+1, d < L-1.
    for (size_t j = 0; j < L; ++j) {
      std::vector<std::pair<size_t, size_t>> tr;
      for (size_t d = 0; d + 1 < L; ++d) tr.emplace_back((j + d) % L, (j + d + 1) % L);
      if (memo_may_fire(idx, plab, tr)) return 0;
    }
    return 1;
  }
  if (step == PJ_STEP_PC) {
    if (index >= cs.paths.size()) return -1;
    const auto& idx = cs.paths[index].vertex;
    const size_t L = idx.size();
    // Forward: positions 0..L-2 -> 1..L-1, memo on all but the last hop.
    std::vector<std::pair<size_t, size_t>> fw, bw;
    for (size_t d = 0; d + 2 < L; ++d) fw.emplace_back(d, d + 1);
    for (size_t d = 0; d + 2 < L; ++d) bw.emplace_back(L - 1 - d, L - 2 - d);
    return (memo_may_fire(idx, plab, fw) || memo_may_fire(idx, plab, bw)) ? 0 : 1;
  }
  return -1;
}

void pjh_constraints_free(pjh_constraints* c) { delete c; }

int pjh_run_pipeline(pj_ctx* ctx, const pjh_constraints* c, int flags, pjh_row* rows, int max_rows,
                     double* total_ms) {
  if (!ctx || !c) return -1;
  pj::MultipleLabelDeviceState st;
  st.ctx = ctx;
  st.managed = true;
  int n = 0;
  auto on_step = [&](const pj::StepRecord& r) {
    if (rows && n < max_rows)
      rows[n] = pjh_row{r.iteration,
                        static_cast<int32_t>(r.step),
                        r.results.vertexEliminated,
                        r.results.edgeEliminated,
                        r.results.matchEliminated,
                        r.results.enumeration,
                        r.activeVertices,
                        r.activeEdges,
                        r.slots,
                        r.deviceMs,
                        r.constraint,
                        0,
                        r.effectiveness,
                        r.statMs};
    ++n;
  };
  const bool stop_enum = (flags & PJH_STOP_BEFORE_ENUMERATION) != 0;
  const int rc = (flags & PJH_ORDER_DYNAMIC)
                     ? pj::run_dynamic_order(&st, c->c, on_step, total_ms, stop_enum, 1)
                     : pj::run_step_machine(&st, c->c, on_step, total_ms, stop_enum, (flags & PJH_STOP_BEFORE_TDS) != 0);
  return rc < 0 ? -1 : n;
}

int pjh_constraint_effectiveness(const pj_label_stat* stat, const pjh_constraints* c, int step, uint32_t index,
                                 int threads, double out[3]) {
  if (!stat || !c || !out || threads < 1) return -1;
  const auto& cs = c->c;
  const pj::LabelStat st(*stat);
  pj::Effectiveness e;
  if (step == PJ_STEP_CC && index < cs.cycles.size())
    e = pj::effectiveness(st, cs.cycles[index], threads);
  else if (step == PJ_STEP_PC && index < cs.paths.size())
    e = pj::effectiveness(st, cs.paths[index], threads);
  else if (step == PJ_STEP_TDS && index < cs.tds.size())
    e = pj::effectiveness(st, cs.tds[index], threads);
  else
    return -1;
  out[0] = e.match_eliminated;
  out[1] = e.cost_eliminated;
  out[2] = e.effectiveness;
  return 0;
}

int pjh_next_constraint(const pj_label_stat* stat, const pjh_constraints* c, const uint8_t* done_cycles,
                        const uint8_t* done_paths, const uint8_t* done_tds, int threads, int* step, uint32_t* index,
                        double* effectiveness) {
  if (!stat || !c || !step || !index || threads < 1) return -1;
  const auto& cs = c->c;
  std::vector<uint32_t> cl, pl, tl;
  for (uint32_t i = 0; i < cs.cycles.size(); ++i)
    if (!done_cycles || !done_cycles[i]) cl.push_back(i);
  for (uint32_t i = 0; i < cs.paths.size(); ++i)
    if (!done_paths || !done_paths[i]) pl.push_back(i);
  for (uint32_t i = 0; i < cs.tds.size(); ++i)
    if (!done_tds || !done_tds[i]) tl.push_back(i);
  const pj::NextConstraint n = pj::choose_next(pj::LabelStat(*stat), cs, cl, pl, tl, threads);
  *step = n.step;
  *index = n.index;
  if (effectiveness) *effectiveness = n.effectiveness;
  return 0;
}

uint64_t pjh_projected_bytes(const pjh_constraints* c, uint32_t vertex_count, uint64_t edge_count) {
  const auto& cs = c->c;
  const uint64_t V = vertex_count, E = edge_count;
  // Per-context state (pj_device.hip alloc_state): label codes 1, candidate /
  // toUnmatch / staging words 3 x 8, four flag bytes, scan lengths 4,
  // frontier bins 4 x 4, two live lists 2 x 4 bytes per vertex; edge-active
  // bits.
  uint64_t bytes = V * (1 + 24 + 4 + 4 + 16 + 8) + (E + 31) / 32 * 4;
  // Memoised walks: every TDS step, and cycle / path constraints whose memo
  // may fire (pjh_constraints_memo_inert == 0).  Complete cycle and path
  // walks of 4+ vertices keep a failed-state memo in the same tables.
  bool memo = !cs.tds.empty(), tables = memo;
  for (uint32_t i = 0; i < cs.cycles.size(); ++i) {
    const bool fires = pjh_constraints_memo_inert(c, PJ_STEP_CC, i) == 0;
    memo = memo || fires;
    tables = tables || fires || cs.cycles[i].vertex.size() >= 4;
  }
  for (uint32_t i = 0; i < cs.paths.size(); ++i) {
    const bool fires = pjh_constraints_memo_inert(c, PJ_STEP_PC, i) == 0;
    memo = memo || fires;
    tables = tables || fires || cs.paths[i].vertex.size() >= 4;
  }
  if (memo || tables) bytes += V * 8 + 8 + (2 * V + 2) * 4;  // alreadyMatched, overflow lists (prepare_memo)
  if (tables) {  // first-pass tables: 2048 waves x 2^log2 entries x 20 bytes
    int bits = 0;
    while (bits < 63 && (1ull << bits) < V) ++bits;
    const int log2 = std::min(18, std::max(16, bits - 6));
    bytes += 2048ull * (1ull << log2) * 20;
  }
  if (memo) bytes += 2 * (V + 1) * 8 + (E + 1) * 4;  // active view: counts, offsets, at most every slot
  return bytes;
}

int pjh_query_uses_memo(const pjh_constraints* c) {
  const auto& cs = c->c;
  if (!cs.tds.empty()) return 1;
  for (uint32_t i = 0; i < cs.cycles.size(); ++i)
    if (pjh_constraints_memo_inert(c, PJ_STEP_CC, i) == 0 || cs.cycles[i].vertex.size() >= 4) return 1;
  for (uint32_t i = 0; i < cs.paths.size(); ++i)
    if (pjh_constraints_memo_inert(c, PJ_STEP_PC, i) == 0 || cs.paths[i].vertex.size() >= 4) return 1;
  return 0;
}

namespace {

// One query on a worker context: installs the template, imports the batch's
// parked state when `lane` is given, runs the step machine (from the batch's
// rows when given) and takes the state checksum.
void solve_query(pj_ctx* ctx, const pjh_constraints* c, int flags, pj_batch* batch, int64_t lane,
                 pjh_query_result* r, std::string* e) {
  uint64_t h = 1469598103934665603ull;
  auto mixin = [&h](uint64_t x) {
    h ^= x + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    h *= 1099511628211ull;
  };
  int steps = 0;
  double dev = 0.0;
  if (pj_set_template(ctx, c->c.desc()) != 0) {
    *e = pj_last_error(ctx);
    return;
  }
  std::vector<pj::AlgoResults> prefix;
  if (lane >= 0) {
    uint32_t nrows = 0;
    std::vector<pj_algo_results> rows(1000);
    if (pj_batch_trace(batch, static_cast<uint32_t>(lane), rows.data(), static_cast<uint32_t>(rows.size()), &nrows) ||
        pj_batch_export(batch, static_cast<uint32_t>(lane), ctx) ||
        pj_batch_release(batch, static_cast<uint32_t>(lane))) {
      *e = pj_batch_last_error(batch);
      return;
    }
    for (uint32_t k = 0; k < nrows && k < rows.size(); ++k) {
      pj::AlgoResults a;
      a.vertexEliminated = rows[k].vertex_eliminated;
      a.edgeEliminated = rows[k].edge_eliminated;
      a.matchEliminated = rows[k].match_eliminated;
      a.enumeration = rows[k].enumeration;
      prefix.push_back(a);
    }
  }
  pj::MultipleLabelDeviceState st;
  st.ctx = ctx;
  st.managed = true;
  const int rc = pj::run_step_machine(
      &st, c->c,
      [&](const pj::StepRecord& rec) {
        ++steps;
        mixin(static_cast<uint64_t>(rec.iteration));
        mixin(static_cast<uint64_t>(rec.step));
        mixin(rec.results.vertexEliminated);
        mixin(rec.results.edgeEliminated);
        mixin(rec.results.matchEliminated);
        mixin(rec.results.enumeration);
        mixin(rec.activeVertices);
        mixin(rec.activeEdges);
      },
      &dev, (flags & PJH_STOP_BEFORE_ENUMERATION) != 0, (flags & PJH_STOP_BEFORE_TDS) != 0,
      lane >= 0 ? &prefix : nullptr);
  if (rc < 0) *e = st.message.empty() ? std::string(pj_last_error(ctx)) : st.message;
  if (e->empty() && pj_state_checksum(ctx, &r->state_checksum) != 0) *e = pj_last_error(ctx);
  pj_get_counters(ctx, &r->active_vertices, &r->active_edges);
  r->steps = steps;
  r->rows_hash = h;
  r->device_ms = dev;
}

}  // namespace

int pjh_run_queries(pj_ctx* owner, const pjh_constraints* const* templates, int n, int workers, int flags,
                    const double* arrival_ms, uint64_t budget_bytes, pjh_query_result* out, double* total_ms,
                    char* err, size_t errlen) {
  if (!owner || (n > 0 && (!templates || !out)) || workers < 1) return -1;
  const bool batched = (flags & PJH_BATCH_LCC) != 0;
  // every worker owns a context sharing the graph; created before the clock,
  // as is the batch
  std::vector<pj_ctx*> ctxs(workers, nullptr);
  pj_batch* batch = nullptr;
  int device = 0;
  if (pj_get_device(owner, &device) != 0) return -1;
  auto cleanup = [&]() {
    if (batch) pj_batch_destroy(batch);
    for (auto* c : ctxs)
      if (c) pj_destroy(c);
  };
  for (int w = 0; w < workers; ++w) {
    if (pj_create(&ctxs[w], device) || pj_share_graph(ctxs[w], owner)) {
      put_err(ctxs[w] ? pj_last_error(ctxs[w]) : "pj_create failed", err, errlen);
      cleanup();
      return -1;
    }
  }
  uint32_t lanes = 0;
  if (batched) {
    lanes = static_cast<uint32_t>(std::min(256, std::max(64, (n + 63) / 64 * 64)));
    if (const char* env = std::getenv("PJ_BATCH_LANES")) lanes = static_cast<uint32_t>(std::max(1, std::atoi(env)));
    if (pj_batch_create(&batch, owner, lanes) != 0) {
      put_err(batch ? pj_batch_last_error(batch) : "pj_batch_create failed", err, errlen);
      cleanup();
      return -1;
    }
    pj_batch_info(batch, &lanes, nullptr, nullptr, nullptr, nullptr);
  }
  uint64_t vc = 0, ec = 0;
  pj_get_counters(ctxs[0], &vc, &ec);  // a fresh context: V and E
  std::vector<uint64_t> projected(n);
  bool any_memo = false;
  for (int i = 0; i < n; ++i) {
    projected[i] = pjh_projected_bytes(templates[i], static_cast<uint32_t>(vc), ec);
    any_memo = any_memo || pjh_query_uses_memo(templates[i]);
  }
  // The retry tables are one set per device shared by all contexts
  // (PJ_RETRY_POOL_BYTES): held back from the budget once.
  if (budget_bytes && any_memo) budget_bytes = budget_bytes > PJ_RETRY_POOL_BYTES ? budget_bytes - PJ_RETRY_POOL_BYTES : 1;
  std::mutex mu;
  std::condition_variable cv;
  int next = 0;  // next query to admit
  uint64_t reserved = 0;
  std::string first_error;
  // batched: queries ready for a worker (query, parked lane or -1)
  std::deque<std::pair<int, int64_t>> ready;
  bool admission_done = false;
  int idle = 0;  // batched: workers waiting for a ready query
  // served (arrival times given): arrivals waiting before a batch starts
  int batch_min = static_cast<int>(std::min<uint32_t>(lanes ? lanes : 1, 64));
  // PJ_SERVED_BATCH_MIN=0: every arrival takes a batch lane (no direct start)
  if (const char* env = std::getenv("PJ_SERVED_BATCH_MIN")) batch_min = std::max(0, std::atoi(env));
  const auto t0 = std::chrono::steady_clock::now();
  auto now_ms = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); };
  auto fits = [&](int i) { return budget_bytes == 0 || reserved == 0 || reserved + projected[i] <= budget_bytes; };
  auto begin = [&](int i) {
    pjh_query_result& r = out[i];
    r = pjh_query_result{};
    r.projected_bytes = projected[i];
    r.arrival_ms = arrival_ms ? arrival_ms[i] : 0.0;
    r.start_ms = r.worker_ms = now_ms();
    r.batch_lane = -1;
    if (!batched) reserved += projected[i];
  };
  auto finish = [&](int i, int w, const std::string& e) {
    out[i].worker = w;
    out[i].end_ms = now_ms();
    std::lock_guard<std::mutex> lk(mu);
    reserved -= projected[i];
    if (!e.empty() && first_error.empty()) first_error = "query " + std::to_string(i) + ": " + e;
  };

  auto worker = [&](int w) {
    pj_ctx* ctx = ctxs[w];
    while (true) {
      int i;
      int64_t lane = -1;
      {
        std::unique_lock<std::mutex> lk(mu);
        if (batched) {
          // the next ready query, once its bytes fit the budget
          ++idle;
          cv.wait(lk, [&] {
            return !first_error.empty() || (ready.empty() ? admission_done : fits(ready.front().first));
          });
          --idle;
          if (!first_error.empty() || ready.empty()) return;
          i = ready.front().first;
          lane = ready.front().second;
          ready.pop_front();
          reserved += projected[i];
          out[i].worker_ms = now_ms();
        } else {
          while (true) {
            if (next >= n || !first_error.empty()) return;
            const double due = arrival_ms ? arrival_ms[next] : 0.0;
            const double t = now_ms();
            if (t >= due && fits(next)) break;
            if (t < due)
              cv.wait_for(lk, std::chrono::duration<double, std::milli>(due - t));
            else
              cv.wait(lk);
          }
          i = next++;
          begin(i);
        }
      }
      std::string e;
      solve_query(ctx, templates[i], flags, batch, lane, &out[i], &e);
      finish(i, w, e);
      cv.notify_all();
    }
  };
  std::vector<std::thread> threads;
  for (int w = 0; w < workers; ++w) threads.emplace_back(worker, w);

  if (batched) {
    // Admission into the batch's lanes (templates the batch cannot hold go to
    // the workers as they are); every step hands the templates that reached
    // their LCC fixed point to the workers and frees their lanes.
    std::vector<int> lane_query(lanes, -1);
    std::vector<uint32_t> done(lanes);
    while (true) {
      uint32_t running = 0, parked = 0;
      {
        std::unique_lock<std::mutex> lk(mu);
        if (!first_error.empty()) break;
        pj_batch_info(batch, nullptr, &running, &parked, nullptr, nullptr);
        while (next < n) {
          const double due = arrival_ms ? arrival_ms[next] : 0.0;
          if (now_ms() < due) break;
          // With arrival times (served), a query that finds a worker idle and
          // no query waiting for one is solved whole at once: the batch's
          // sweeps cover the whole graph at every step, so at low load a
          // template parks later than a worker finishes it; under load the
          // workers are busy and arrivals take batch lanes.
          if (arrival_ms && batch_min > 0 && static_cast<int>(ready.size()) < idle) {
            begin(next);
            ready.emplace_back(next, -1);
            ++next;
            cv.notify_all();
            continue;
          }
          // A batch step sweeps the whole graph whether few or many lanes
          // take part: arrivals start a batch only once batch_min of them
          // wait (and join one that runs); until then they wait for a worker.
          if (arrival_ms && batch_min > 0 && running == 0) {
            int waiting = 0;
            for (int k = next; k < n && waiting < batch_min && arrival_ms[k] <= now_ms(); ++k) ++waiting;
            if (waiting < batch_min) break;
          }
          if (running + parked >= lanes) break;
          uint32_t lane = 0;
          const int rc = pj_batch_admit(batch, templates[next]->c.desc(), &lane);
          if (rc == -2) {  // the batch cannot hold it: a worker solves it whole
            begin(next);
            ready.emplace_back(next, -1);
            ++next;
            cv.notify_all();
            continue;
          }
          if (rc != 0) {
            first_error = std::string("query ") + std::to_string(next) + ": " + pj_batch_last_error(batch);
            break;
          }
          begin(next);
          out[next].batch_lane = static_cast<int32_t>(lane);
          lane_query[lane] = next;
          ++running;
          ++next;
        }
        if (!first_error.empty()) break;
        if (!running) {
          if (next >= n) break;
          // nothing to sweep: wait for the next arrival, a lane or budget
          const double due = arrival_ms ? arrival_ms[next] : 0.0;
          const double t = now_ms();
          if (t < due)
            cv.wait_for(lk, std::chrono::duration<double, std::milli>(due - t));
          else
            cv.wait_for(lk, std::chrono::milliseconds(1));
          continue;
        }
      }
      uint32_t nd = 0;
      if (pj_batch_run(batch, 1, done.data(), lanes, &nd) != 0) {
        std::lock_guard<std::mutex> lk(mu);
        if (first_error.empty()) first_error = std::string("batch: ") + pj_batch_last_error(batch);
        break;
      }
      if (nd) {
        std::lock_guard<std::mutex> lk(mu);
        for (uint32_t k = 0; k < nd; ++k) ready.emplace_back(lane_query[done[k]], static_cast<int64_t>(done[k]));
      }
      cv.notify_all();
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      admission_done = true;
    }
    cv.notify_all();
  }
  for (auto& t : threads) t.join();
  if (total_ms) *total_ms = now_ms();
  cleanup();
  if (!first_error.empty()) {
    put_err(first_error, err, errlen);
    return -1;
  }
  return 0;
}

// patternmatching_binary.cc:13-42 with the CLI of cmd_line_option.cc:22-78
// (-i -b -v -e -p -o -h) plus -d <device>.
int pjh_main(int argc, char** argv) {
  pj::Options o;
  bool help = false, have_i = false, have_p = false, have_o = false;
  optind = 1;
  int ch;
  while ((ch = getopt(argc, argv, "i:b:v:e:p:o:d:h")) != -1) {
    switch (ch) {
      case 'i':
        o.graph = optarg;
        have_i = true;
        break;
      case 'b':
        o.graphBin = optarg;
        break;
      case 'v':
        o.vertexData = optarg;
        break;
      case 'e':
        o.edgeData = optarg;
        break;
      case 'p':
        o.patternDir = optarg;
        have_p = true;
        break;
      case 'o':
        o.outputDir = optarg;
        have_o = true;
        break;
      case 'd':
        o.device = std::atoi(optarg);
        break;
      default:
        help = true;
    }
  }
  if (help || !have_i || !have_p || !have_o) {
    std::cerr << "Usage: -i <string> -p <string> -o <string>\n"
              << " -i <string>   - input graph base filename (required)\n"
              << " -b <string>   - input binary graph base filename (optional)\n"
              << " -v <string>   - vertex metadata base filename (optional)\n"
              << " -e <string>   - edge metadata base filename (optional, unused)\n"
              << " -p <string>   - pattern base directory (required)\n"
              << " -o <string>   - output base directory (required)\n"
              << " -d <int>      - HIP device (default 0)\n"
              << " -h            - print help and exit\n\n";
    return 255;
  }
  pj::MultipleLabelHip run;
  if (run.allocate(o)) {
    std::cerr << "error: " << run.error() << std::endl;
    run.free();
    return 1;
  }
  int rc = run.runPatternMatching();
  if (rc) std::cerr << "error: " << run.error() << std::endl;
  run.free();
  return rc ? 1 : 0;
}

}  // extern "C"
