// Host pipeline driver (see pj_driver.h).
#include "pj_driver.h"

#include "pj_effectiveness.h"

#include <sys/stat.h>

#include <cstdarg>
#include <cstdlib>
#include <chrono>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iostream>
#include <sstream>

namespace pj {

const char* MultipleLabelStep::name(Step s) {
  switch (s) {
    case E_START:
      return "START";
    case E_END:
      return "END";
    case E_LCC0:
      return "LCC0";
    case E_LCC:
      return "LCC";
    case E_CC:
      return "CC";
    case E_PC:
      return "PC";
    case E_TDS:
      return "TDS";
    case E_ENUMERATION:
      return "ENUMERATION";
  }
  return "?";
}

void MultipleLabelStep::set(Step s, Step* next, std::string* nm) {
  current_ = s;
  *next = s;
  *nm = name(s);
}

bool MultipleLabelStep::getNextStep(const AlgoResults& r, Step* next, std::string* nm) {
  switch (current_) {
    case E_START:
      set(E_LCC0, next, nm);
      return false;
    case E_LCC0:
      set(E_LCC, next, nm);
      return false;
    case E_LCC:
      if (!r.isEmpty()) return false;  // stay on LCC
      if (iCc_ < nCc_)
        set(E_CC, next, nm);
      else if (iPc_ < nPc_)
        set(E_PC, next, nm);
      else if (iTds_ < nTds_)
        set(E_TDS, next, nm);
      else
        set(E_ENUMERATION, next, nm);
      return false;
    case E_CC:
      ++iCc_;
      set(E_LCC, next, nm);
      return false;
    case E_PC:
      ++iPc_;
      set(E_LCC, next, nm);
      return false;
    case E_TDS:
      ++iTds_;
      set(E_LCC, next, nm);
      return false;
    case E_ENUMERATION:
      set(E_END, next, nm);
      return true;
    case E_END:
      return false;
  }
  return false;
}

int MultipleLabelDeviceState::allocate(uint64_t vertexCount, uint64_t edgeCount, uint32_t patternCount) {
  if (ctx) free();
  graphVertexCount = graphActiveVertexCount = vertexCount;
  graphEdgeCount = graphActiveEdgeCount = edgeCount;
  patternVertexCount = patternCount;
  error = 0;
  message.clear();
  uploaded = nullptr;
  template_version = 0;
  host_fresh_ = false;
  const char* dev = std::getenv("PJ_DEVICE");
  if (pj_create(&ctx, dev ? std::atoi(dev) : 0) != 0) {
    ctx = nullptr;
    error = 1;
    message = "pj_create failed";
    return -1;
  }
  return 0;
}

int MultipleLabelDeviceState::free() {
  if (ctx) pj_destroy(ctx);
  ctx = nullptr;
  uploaded = nullptr;
  template_version = 0;
  host_fresh_ = false;
  return 0;
}

void MultipleLabelDeviceState::sync_host() const {
  if (host_fresh_ || !ctx) return;
  h_cand_.assign(graphVertexCount, 0);
  h_vflags_.assign(graphVertexCount, 0);
  h_ea_.assign((graphEdgeCount + 31) / 32, 0);
  if (pj_download_state(ctx, h_cand_.data(), h_vflags_.data(), h_ea_.data()) != 0) {
    std::fprintf(stderr, "pj_download_state: %s\n", pj_last_error(ctx));
    return;
  }
  host_fresh_ = true;
}

uint32_t MultipleLabelDeviceState::ActiveList::operator[](uint64_t i) const {
  s->sync_host();
  if (edges) return i < s->h_ea_.size() * 32 ? (s->h_ea_[i >> 5] >> (i & 31)) & 1u : 0u;
  return i < s->h_vflags_.size() ? s->h_vflags_[i] & 1u : 0u;
}

MultipleLabelDeviceState::CandidateSet MultipleLabelDeviceState::CandidateList::operator[](uint64_t v) const {
  s->sync_host();
  return CandidateSet{v < s->h_cand_.size() ? s->h_cand_[v] : 0};
}

namespace detail {

TemplateAssembly& reference_template() {
  static TemplateAssembly assembly;
  return assembly;
}

int upload_graph(MultipleLabelDeviceState* s, uint32_t V, uint64_t E, const std::vector<uint64_t>& offsets,
                 const uint32_t* edges, const uint64_t* labels, const void* key) {
  if (pj_upload_graph(s->ctx, V, E, offsets.data(), edges, labels) != 0) {
    s->error = 1;
    s->message = pj_last_error(s->ctx);
    return -1;
  }
  s->uploaded = key;
  s->template_version = 0;  // a new graph needs the template again
  s->graphVertexCount = s->graphActiveVertexCount = V;
  s->graphEdgeCount = s->graphActiveEdgeCount = E;
  s->invalidate_host_copy();
  return 0;
}

int install_reference_template(MultipleLabelDeviceState* s) {
  TemplateAssembly& a = reference_template();
  if (a.set.local.empty()) {
    s->error = 1;
    s->message = "no local constraints were given to the step classes";
    return -1;
  }
  if (!a.have_enumeration) {  // setConstraint not called: enumerate along a walk of the first TDS constraint
    a.set.enumeration = a.set.tds.empty() ? TemplateConstraint{} : a.set.tds.front();
  }
  a.set.flatten();
  if (pj_set_template(s->ctx, a.set.desc()) != 0) {
    s->error = 1;
    s->message = pj_last_error(s->ctx);
    return -1;
  }
  s->template_version = a.version;
  pj_get_counters(s->ctx, &s->graphActiveVertexCount, &s->graphActiveEdgeCount);
  s->invalidate_host_copy();
  return 0;
}

AlgoResults run_device_step(MultipleLabelDeviceState* s, int step, uint32_t index) {
  AlgoResults out;
  if (s->error) return out;
  s->invalidate_host_copy();
  pj_algo_results r{};
  if (pj_run_step(s->ctx, step, index, &r) != 0) {
    s->error = 1;
    s->message = pj_last_error(s->ctx);
    return out;
  }
  out.vertexEliminated = r.vertex_eliminated;
  out.edgeEliminated = r.edge_eliminated;
  out.matchEliminated = r.match_eliminated;
  out.enumeration = r.enumeration;
  pj_get_counters(s->ctx, &s->graphActiveVertexCount, &s->graphActiveEdgeCount);
  return out;
}

void reset_modified(MultipleLabelDeviceState* s) {
  if (s->error) return;
  s->invalidate_host_copy();
  if (pj_reset_modified(s->ctx) != 0) {
    s->error = 1;
    s->message = pj_last_error(s->ctx);
  }
}

}  // namespace detail

int run_step_machine(MultipleLabelDeviceState* state, const ConstraintSet& cs,
                     const std::function<void(const StepRecord&)>& on_step, double* total_ms,
                     bool stop_before_enumeration, bool stop_before_tds, const std::vector<AlgoResults>* prefix) {
  MultipleLabelHip::Lcc0Type lcc0;
  MultipleLabelHip::LccType lcc;
  MultipleLabelHip::CcType cc;
  MultipleLabelHip::PcType pc;
  MultipleLabelHip::TdsType tds;
  MultipleLabelHip::EnumerationType en;
  lcc0.setConstraintVector(cs.local);
  lcc.setConstraintVector(cs.local);
  cc.setConstraintVector(cs.cycles);
  pc.setConstraintVector(cs.paths);
  tds.setConstraintVector(cs.tds);
  en.setConstraint(cs.enumeration);
  MultipleLabelStep machine;
  machine.initStepCc(static_cast<int>(cs.cycles.size()));
  machine.initStepPc(static_cast<int>(cs.paths.size()));
  machine.initStepTds(static_cast<int>(cs.tds.size()));

  Graph unused;
  AlgoResults results;
  MultipleLabelStep::Step step;
  std::string name;
  bool finished = false;
  int iteration = 0;
  double total = 0.0;
  const bool trace = std::getenv("PJ_TRACE") != nullptr;  // per-step progress on stderr
  int cc_runs = 0, pc_runs = 0;
  machine.getNextStep(results, &step, &name);
  if (prefix && !prefix->empty()) {
    // The batch ran LCC0 and the LCC steps to the first fixed point; the
    // active counts of each row follow from the eliminations.
    uint64_t V = 0, E = 0;
    pj_get_counters(state->ctx, &state->graphActiveVertexCount, &state->graphActiveEdgeCount);
    for (const AlgoResults& r : *prefix) {
      V += r.vertexEliminated;
      E += r.edgeEliminated;
    }
    V += state->graphActiveVertexCount;
    E += state->graphActiveEdgeCount;
    for (const AlgoResults& r : *prefix) {
      const bool lcc = step == MultipleLabelStep::E_LCC0 || step == MultipleLabelStep::E_LCC;
      if (!lcc) {
        state->error = 1;
        state->message = "batch rows do not follow the step machine";
        return -1;
      }
      V -= r.vertexEliminated;
      E -= r.edgeEliminated;
      if (on_step) on_step(StepRecord{iteration, step, r, V, E, 0, 0.0, 0.0});
      results = r;
      finished = machine.getNextStep(results, &step, &name);
      ++iteration;
    }
    if (step == MultipleLabelStep::E_LCC) {
      state->error = 1;
      state->message = "batch rows end before the LCC fixed point";
      return -1;
    }
  } else if (pj_state_reset(state->ctx) != 0) {
    state->error = 1;
    state->message = pj_last_error(state->ctx);
    return -1;
  }
  pj_get_counters(state->ctx, &state->graphActiveVertexCount, &state->graphActiveEdgeCount);
  while (!finished && iteration < 1000) {
    if (stop_before_enumeration && step == MultipleLabelStep::E_ENUMERATION) break;
    if (stop_before_tds && step == MultipleLabelStep::E_TDS) break;
    switch (step) {
      case MultipleLabelStep::E_CC:
        cc.resetState(state);
        break;
      case MultipleLabelStep::E_PC:
        pc.resetState(state);
        break;
      case MultipleLabelStep::E_TDS:
        tds.resetState(state);
        break;
      default:
        break;
    }
    auto t0 = std::chrono::steady_clock::now();
    switch (step) {
      case MultipleLabelStep::E_LCC0:
        results = lcc0.compute(unused, state);
        break;
      case MultipleLabelStep::E_LCC:
        results = lcc.compute(unused, state);
        break;
      case MultipleLabelStep::E_CC:
        results = cc.compute(unused, state);
        break;
      case MultipleLabelStep::E_PC:
        results = pc.compute(unused, state);
        break;
      case MultipleLabelStep::E_TDS:
        results = tds.compute(unused, state);
        break;
      case MultipleLabelStep::E_ENUMERATION:
        results = en.compute(unused, state);
        break;
      default:
        break;
    }
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (state->error) return -1;
    double dev = 0.0;
    uint64_t slots = 0;
    pj_last_step_ms(state->ctx, &dev);
    pj_last_step_slots(state->ctx, &slots);
    total += dev;
    if (trace)
      std::fprintf(stderr, "[pj] step %d %s %.3f ms device, %llu slots, active %llu\n", iteration,
                   MultipleLabelStep::name(step), dev, static_cast<unsigned long long>(slots),
                   static_cast<unsigned long long>(state->graphActiveVertexCount));
    if (on_step) {
      StepRecord rec{iteration, step, results, state->graphActiveVertexCount, state->graphActiveEdgeCount, slots,
                     wall, dev};
      // the iterators: cycles and paths advance per step, TDS stays at 0
      if (step == MultipleLabelStep::E_CC) rec.constraint = cc_runs;
      if (step == MultipleLabelStep::E_PC) rec.constraint = pc_runs;
      if (step == MultipleLabelStep::E_TDS) rec.constraint = 0;
      on_step(rec);
    }
    if (step == MultipleLabelStep::E_CC) ++cc_runs;
    if (step == MultipleLabelStep::E_PC) ++pc_runs;
    finished = machine.getNextStep(results, &step, &name);
    ++iteration;
  }
  if (total_ms) *total_ms = total;
  return iteration;
}

int run_dynamic_order(MultipleLabelDeviceState* state, const ConstraintSet& cs,
                      const std::function<void(const StepRecord&)>& on_step, double* total_ms,
                      bool stop_before_enumeration, int threads) {
  MultipleLabelHip::Lcc0Type lcc0;
  MultipleLabelHip::LccType lcc;
  MultipleLabelHip::CcType cc;
  MultipleLabelHip::PcType pc;
  MultipleLabelHip::TdsType tds;
  MultipleLabelHip::EnumerationType en;
  lcc0.setConstraintVector(cs.local);
  lcc.setConstraintVector(cs.local);
  cc.setConstraintVector(cs.cycles);
  pc.setConstraintVector(cs.paths);
  tds.setConstraintVector(cs.tds);
  en.setConstraint(cs.enumeration);
  // the constraints not yet run, by original index in list order
  std::vector<uint32_t> cycles_left, paths_left, tds_left;
  for (uint32_t i = 0; i < cs.cycles.size(); ++i) cycles_left.push_back(i);
  for (uint32_t i = 0; i < cs.paths.size(); ++i) paths_left.push_back(i);
  for (uint32_t i = 0; i < cs.tds.size(); ++i) tds_left.push_back(i);
  auto drop = [](std::vector<uint32_t>* v, uint32_t index) {
    for (auto it = v->begin(); it != v->end(); ++it)
      if (*it == index) {
        v->erase(it);
        return;
      }
  };

  if (pj_state_reset(state->ctx) != 0) {
    state->error = 1;
    state->message = pj_last_error(state->ctx);
    return -1;
  }
  pj_get_counters(state->ctx, &state->graphActiveVertexCount, &state->graphActiveEdgeCount);
  Graph unused;
  AlgoResults results;
  MultipleLabelStep::Step step = MultipleLabelStep::E_LCC0;
  NextConstraint chosen;
  pj_label_stat* stat = new pj_label_stat;
  bool finished = false;
  int iteration = 0;
  double total = 0.0;
  int rc = 0;
  while (!finished && iteration < 1000) {
    if (stop_before_enumeration && step == MultipleLabelStep::E_ENUMERATION) break;
    auto t0 = std::chrono::steady_clock::now();
    switch (step) {
      case MultipleLabelStep::E_LCC0:
        results = lcc0.compute(unused, state);
        break;
      case MultipleLabelStep::E_LCC:
        results = lcc.compute(unused, state);
        break;
      case MultipleLabelStep::E_CC:
        cc.resetState(state);
        cc.setConstraintIterator(chosen.index);
        results = cc.compute(unused, state);
        break;
      case MultipleLabelStep::E_PC:
        pc.resetState(state);
        pc.setConstraintIterator(chosen.index);
        results = pc.compute(unused, state);
        break;
      case MultipleLabelStep::E_TDS:
        tds.resetState(state);
        tds.setConstraintIterator(chosen.index);
        results = tds.compute(unused, state);
        break;
      case MultipleLabelStep::E_ENUMERATION:
        results = en.compute(unused, state);
        break;
      default:
        break;
    }
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (state->error) {
      rc = -1;
      break;
    }
    double dev = 0.0;
    uint64_t slots = 0;
    pj_last_step_ms(state->ctx, &dev);
    pj_last_step_slots(state->ctx, &slots);
    total += dev;
    StepRecord rec{iteration, step, results, state->graphActiveVertexCount, state->graphActiveEdgeCount, slots,
                   wall, dev};
    const bool nlcc = step == MultipleLabelStep::E_CC || step == MultipleLabelStep::E_PC ||
                      step == MultipleLabelStep::E_TDS;
    if (nlcc) {
      rec.constraint = static_cast<int>(chosen.index);
      rec.effectiveness = chosen.effectiveness;
    }
    // next step (test_order_multiple_label_cpu.cc:171-287)
    if (step == MultipleLabelStep::E_LCC && results.isEmpty()) {
      if (pj_graph_stat(state->ctx, stat) != 0) {
        state->error = 1;
        state->message = pj_last_error(state->ctx);
        rc = -1;
        break;
      }
      pj_last_step_ms(state->ctx, &rec.statMs);
      total += rec.statMs;
      chosen = choose_next(LabelStat(*stat), cs, cycles_left, paths_left, tds_left, threads);
      if (chosen.step == PJ_STEP_CC) {
        drop(&cycles_left, chosen.index);
        step = MultipleLabelStep::E_CC;
      } else if (chosen.step == PJ_STEP_PC) {
        drop(&paths_left, chosen.index);
        step = MultipleLabelStep::E_PC;
      } else if (chosen.step == PJ_STEP_TDS) {
        drop(&tds_left, chosen.index);
        step = MultipleLabelStep::E_TDS;
      } else {
        step = MultipleLabelStep::E_ENUMERATION;
      }
    } else if (step == MultipleLabelStep::E_ENUMERATION) {
      finished = true;
    } else {
      step = MultipleLabelStep::E_LCC;
    }
    if (on_step) on_step(rec);
    ++iteration;
  }
  delete stat;
  if (rc) return -1;
  if (total_ms) *total_ms = total;
  return iteration;
}

// ---------------------------------------------------------------------------

namespace {

bool is_file(const std::string& p) {
  struct stat st;
  return !p.empty() && stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

void make_dirs(const std::string& dir) {
  std::string cur;
  for (size_t i = 0; i < dir.size(); ++i) {
    cur.push_back(dir[i]);
    if (dir[i] == '/' || i + 1 == dir.size()) mkdir(cur.c_str(), 0755);
  }
}

std::string fmt(const char* f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt(const char* f, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, f);
  vsnprintf(buf, sizeof(buf), f, ap);
  va_end(ap);
  return buf;
}

}  // namespace

void MultipleLabelHip::log(const std::string& line, bool to_cout) const {
  if (to_cout) std::cout << line << std::endl;
  std::ofstream os(options_.outputDir + "log.totem", std::ios::app);
  std::time_t now = std::time(nullptr);
  char ts[64];
  std::strftime(ts, sizeof(ts), "[%F %T] : ", std::localtime(&now));
  os << ts << line << std::endl;
}

// MultipleLabelCpu::allocate (multiple_label_cpu.cc:20-104).
int MultipleLabelHip::allocate(const Options& options) {
  options_ = options;
  if (!options_.outputDir.empty() && options_.outputDir.back() != '/') options_.outputDir += "/";
  if (!options_.patternDir.empty() && options_.patternDir.back() != '/') options_.patternDir += "/";
  if (options_.graphBin.empty()) options_.graphBin = options_.graph + ".bin";
  make_dirs(options_.outputDir);
  // Logger::init truncates the result and log files (logger.cc:52-80).
  std::ofstream(options_.outputDir + "results.totem", std::ios::trunc).close();
  std::ofstream(options_.outputDir + "log.totem", std::ios::trunc).close();
  std::string err;
  log("Loading graph", true);
  // A binary cache streams into HBM (pj_upload_totem_file); the host keeps
  // its counts and vertex values only.
  bool streamed = false;
  int kind = is_file(options_.graphBin) ? load_totem_binary_values(options_.graphBin, &graph_, &err) : 1;
  if (kind < 0) {
    error_ = err;
    return -1;
  }
  if (kind == 0) {
    streamed = true;
  } else if (is_file(options_.graphBin)) {
    if (load_totem(options_.graphBin, &graph_, &err)) {
      error_ = err;
      return -1;
    }
  } else {
    if (load_totem(options_.graph, &graph_, &err)) {
      error_ = err;
      return -1;
    }
    if (!options_.vertexData.empty() && is_file(options_.vertexData)) {
      if (parse_vertex_file(options_.vertexData, &graph_, &err)) {
        error_ = err;
        return -1;
      }
    }
    log("Saved graph as : " + options_.graphBin, true);
    store_totem_binary(graph_, options_.graphBin, graph_.edge_count > 0xFFFFFFFFull ? 8 : 4, &err);
  }
  if (!graph_.valued) {
    error_ = "graph has no vertex labels (give -v or a valued graph file)";
    return -1;
  }
  log("Loading pattern", true);
  const std::string pe = options_.patternDir + "pattern_edge";
  if (load_totem(is_file(pe + ".totem") ? pe + ".totem" : pe, &pattern_, &err)) {
    error_ = err;
    return -1;
  }
  // Pattern labels are read when vertex metadata was requested (multiple_label_cpu.cc:43-47).
  if (!options_.vertexData.empty()) {
    if (parse_vertex_file(options_.patternDir + "pattern_vertex_data", &pattern_, &err)) {
      error_ = err;
      return -1;
    }
  }
  log("Generate pattern multiple_label", true);
  if (constraints_.build(pattern_, &err)) {
    error_ = err;
    return -1;
  }
  log(constraints_.dump(), false);

  if (pj_create(&state_.ctx, options_.device)) {
    error_ = "pj_create failed";
    return -1;
  }
  const int up = streamed ? pj_upload_totem_file(state_.ctx, options_.graphBin.c_str(), graph_.values.data())
                          : pj_upload_graph(state_.ctx, graph_.vertex_count, graph_.edge_count,
                                            graph_.vertices.data(), graph_.edges.data(), graph_.values.data());
  if (up ||
      pj_set_template(state_.ctx, constraints_.desc())) {
    error_ = pj_last_error(state_.ctx);
    return -1;
  }
  state_.managed = true;
  state_.graphVertexCount = graph_.vertex_count;
  state_.graphEdgeCount = graph_.edge_count;
  state_.patternVertexCount = pattern_.vertex_count;
  log("End of initialisation", true);
  return 0;
}

int MultipleLabelHip::free() {
  if (state_.ctx) pj_destroy(state_.ctx);
  state_.ctx = nullptr;
  return 0;
}

// logResults (multiple_label_cpu.cc:187-250) and the final dumps
// printActiveGraph / printActiveVertex (:252-297).
void MultipleLabelHip::logResults(int iteration, bool logGraph) {
  std::ofstream rs(options_.outputDir + "results.totem", std::ios::app);
  if (iteration == 0)
    rs << "Iteration; Step; Step time; Cumulative time; Eliminated vertex; Active vertex; Total vertex; "
          "Eliminated edge; Active edge; Total edge; Eliminated match; Enumeration number"
       << std::endl;
  // graphActiveEdgeCount is a 32-bit vid_t in the reference (multiple_label_common_cpu.h:45).
  rs << fmt("%04d; %s; %.4f ; %.4f ; %zu; %zu; %zu; %zu; %zu; %zu; %zu; %zu", iteration, stepName_.c_str(), stepTime_,
            totalTime_, results_.vertexEliminated, static_cast<size_t>(state_.graphActiveVertexCount),
            static_cast<size_t>(state_.graphVertexCount), results_.edgeEliminated,
            static_cast<size_t>(static_cast<uint32_t>(state_.graphActiveEdgeCount)),
            static_cast<size_t>(state_.graphEdgeCount), results_.matchEliminated, results_.enumeration)
     << std::endl;
  if (!logGraph) return;
  // The surviving vertices and slots, compacted on the device
  // (pj_download_active); the host only formats them.
  uint64_t nv = 0, ne = 0;
  if (pj_download_active(state_.ctx, nullptr, nullptr, 0, nullptr, nullptr, 0, &nv, &ne)) {
    error_ = pj_last_error(state_.ctx);
    return;
  }
  std::vector<uint32_t> vid(nv), src(ne), dst(ne);
  std::vector<uint64_t> cand(nv);
  if (pj_download_active(state_.ctx, vid.data(), cand.data(), nv, src.data(), dst.data(), ne, &nv, &ne)) {
    error_ = pj_last_error(state_.ctx);
    return;
  }
  std::ostringstream g, v;
  for (uint64_t i = 0; i < ne; ++i)
    g << src[i] << " " << dst[i] << " " << graph_.values[src[i]] << " " << graph_.values[dst[i]] << "\n";
  for (uint64_t i = 0; i < nv; ++i) {
    v << vid[i] << " ";
    for (uint64_t m = cand[i]; m; m &= m - 1) v << __builtin_ctzll(m) << " ";
    v << "\n";
  }
  std::ofstream(options_.outputDir + fmt("graph_%04d.totem", iteration)) << g.str() << std::endl;
  std::ofstream(options_.outputDir + fmt("vertex_%04d.totem", iteration)) << v.str() << std::endl;
}

int MultipleLabelHip::runPatternMatching() {
  totalTime_ = 0.0;
  log("Start run", true);
  int last = 0;
  int rc = run_step_machine(
      &state_, constraints_,
      [&](const StepRecord& rec) {
        stepName_ = MultipleLabelStep::name(rec.step);
        std::cout << "Start " << stepName_ << std::endl;
        results_ = rec.results;
        stepTime_ = rec.wallMs;
        totalTime_ += rec.wallMs;
        logResults(rec.iteration, false);
        std::string line;
        if (rec.step == MultipleLabelStep::E_LCC0 || rec.step == MultipleLabelStep::E_LCC)
          line = fmt("Iteration %4d [%s], Running time %.4f/%.4f, Eliminated vertex %zu/%zu/%zu, "
                     "Eliminated edge %zu/%zu/%zu, Eliminated match %zu",
                     rec.iteration, stepName_.c_str(), stepTime_, totalTime_, results_.vertexEliminated,
                     static_cast<size_t>(state_.graphActiveVertexCount), static_cast<size_t>(state_.graphVertexCount),
                     results_.edgeEliminated, static_cast<size_t>(state_.graphActiveEdgeCount),
                     static_cast<size_t>(state_.graphEdgeCount), results_.matchEliminated);
        else if (rec.step == MultipleLabelStep::E_ENUMERATION)
          line = fmt("Iteration %4d [%s], Running time %.4f/%.4f, Enumeration %zu", rec.iteration, stepName_.c_str(),
                     stepTime_, totalTime_, results_.enumeration);
        else
          line = fmt("Iteration %4d [%s], Running time %.4f/%.4f, Eliminated vertex %zu/%zu/%zu, Eliminated match %zu",
                     rec.iteration, stepName_.c_str(), stepTime_, totalTime_, results_.vertexEliminated,
                     static_cast<size_t>(state_.graphActiveVertexCount), static_cast<size_t>(state_.graphVertexCount),
                     results_.matchEliminated);
        log(line, true);
        last = rec.iteration + 1;
      },
      nullptr);
  if (rc < 0) {
    error_ = state_.message;
    return -1;
  }
  log("End run", true);
  log("Saving pruned graph", true);
  stepName_ = "END";
  stepTime_ = 0.0;
  results_ = AlgoResults();
  logResults(last, true);
  return error_.empty() ? 0 : -1;
}

}  // namespace pj
