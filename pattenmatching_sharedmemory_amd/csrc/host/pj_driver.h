// Host pipeline driver: the reference's step classes and step machine,
// re-targeted to the device C ABI.
//
// The reference selects its step implementations with the typedef block of
// MultipleLabelCpu (src/patternmatching/multiple_label_cpu.h:34-52).  The
// classes here keep those names with a Hip suffix and the same method set
// (compute / resetState / setConstraintVector / setConstraintIterator), so a
// driver written against the reference's seam drives the HIP kernels by
// changing only those typedefs.  MultipleLabelHip mirrors MultipleLabelCpu
// (allocate / runPatternMatching / free) including its output files.
#ifndef PJ_DRIVER_H
#define PJ_DRIVER_H

#include <cstdint>
#include <functional>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/prunejuice.h"
#include "pj_constraints.h"
#include "pj_graph.h"

namespace pj {

// algo_results.h:11-22
struct AlgoResults {
  size_t vertexEliminated = 0;
  size_t edgeEliminated = 0;
  size_t matchEliminated = 0;
  size_t enumeration = 0;
  bool isEmpty() const {
    return vertexEliminated == 0 && edgeEliminated == 0 && matchEliminated == 0 && enumeration == 0;
  }
};

// multiple_label_step.h:13-59 / multiple_label_step.cc:30-82
class MultipleLabelStep {
 public:
  enum Step { E_START, E_END, E_LCC0, E_LCC, E_CC, E_PC, E_TDS, E_ENUMERATION };
  void initStepCc(int n) { nCc_ = n, iCc_ = 0; }
  void initStepPc(int n) { nPc_ = n, iPc_ = 0; }
  void initStepTds(int n) { nTds_ = n, iTds_ = 0; }
  bool getNextStep(const AlgoResults& r, Step* next, std::string* name);
  static const char* name(Step s);

 private:
  void set(Step s, Step* next, std::string* name);
  Step current_ = E_START;
  int nCc_ = 0, iCc_ = 0, nPc_ = 0, iPc_ = 0, nTds_ = 0, iTds_ = 0;
};

// Device-resident counterpart of MultipleLabelGlobalState
// (multiple_label_common_cpu.h:23-59).  The arrays live in the context; the
// counters logResults prints are mirrored here (multiple_label_cpu.cc:
// 207-208); the host views vertexActiveList, edgeActiveList and
// vertexPatternMatch, which printActiveGraph / printActiveVertex index
// (multiple_label_cpu.cc:270-297), download the state on first use after a
// step.  allocate() / free() replace the reference's (common_cpu.tpp:29-84):
// they create and release the device context; the graph and the template go
// to the device on the first compute(), from the graph_t it is given and the
// constraint vectors the step classes received.
struct MultipleLabelDeviceState {
  pj_ctx* ctx = nullptr;
  uint64_t graphVertexCount = 0;
  uint64_t graphActiveVertexCount = 0;
  uint64_t graphEdgeCount = 0;
  uint64_t graphActiveEdgeCount = 0;
  uint32_t patternVertexCount = 0;
  int error = 0;  // set when a device call fails (compute() has no error channel)
  std::string message;
  // Set by owners that upload the graph and install the template themselves
  // (MultipleLabelHip, pjh_run_pipeline, pjh_run_queries); otherwise the
  // first compute() does both.
  bool managed = false;
  const void* uploaded = nullptr;    // edge array of the graph_t on the device
  uint64_t template_version = 0;     // of the constraint assembly installed

  MultipleLabelDeviceState() = default;
  MultipleLabelDeviceState(const MultipleLabelDeviceState&) = delete;
  MultipleLabelDeviceState& operator=(const MultipleLabelDeviceState&) = delete;

  // Device from PJ_DEVICE (default 0).  Returns 0 / -1 (error_t).
  int allocate(uint64_t vertexCount, uint64_t edgeCount, uint32_t patternVertexCount);
  int free();

  // Host views (read-only), refreshed from the device when stale.
  struct ActiveList {
    const MultipleLabelDeviceState* s;
    bool edges;
    uint32_t operator[](uint64_t i) const;
  };
  struct CandidateSet {  // template vertices of one vertex, ascending (FixedBitmap iteration)
    uint64_t bits;
    struct iterator {
      uint64_t rest;
      uint32_t operator*() const { return static_cast<uint32_t>(__builtin_ctzll(rest)); }
      iterator& operator++() {
        rest &= rest - 1;
        return *this;
      }
      bool operator!=(const iterator& o) const { return rest != o.rest; }
    };
    iterator begin() const { return iterator{bits}; }
    iterator end() const { return iterator{0}; }
    bool isIn(uint32_t p) const { return (bits >> p) & 1; }
  };
  struct CandidateList {
    const MultipleLabelDeviceState* s;
    CandidateSet operator[](uint64_t v) const;
  };
  ActiveList vertexActiveList{this, false};
  ActiveList edgeActiveList{this, true};
  CandidateList vertexPatternMatch{this};

  void invalidate_host_copy() { host_fresh_ = false; }
  void sync_host() const;

 private:
  mutable bool host_fresh_ = false;
  mutable std::vector<uint64_t> h_cand_;
  mutable std::vector<uint8_t> h_vflags_;
  mutable std::vector<uint32_t> h_ea_;
};

namespace detail {
AlgoResults run_device_step(MultipleLabelDeviceState* s, int step, uint32_t index);
void reset_modified(MultipleLabelDeviceState* s);

// The constraint vectors MultipleLabelCpu::allocate hands to the step
// objects (setConstraintVector / setConstraint, multiple_label_cpu.cc:84-95),
// gathered into one template for the device.  The reference gives them to
// the steps, not to the state, so they meet here: one template per process,
// as in the reference.  Every piece replaces its predecessor and bumps the
// version; the next compute() installs a changed assembly.
struct TemplateAssembly {
  ConstraintSet set;
  uint64_t version = 0;
  bool have_enumeration = false;
};
TemplateAssembly& reference_template();
int upload_graph(MultipleLabelDeviceState* s, uint32_t V, uint64_t E, const std::vector<uint64_t>& offsets,
                 const uint32_t* edges, const uint64_t* labels, const void* key);
int install_reference_template(MultipleLabelDeviceState* s);

// Local constraints (MultipleLabelConstraintLocal, multiple_label_constraint_local.h:18-32):
// originVertex, originLabel, localConstraint (label -> count),
// localConstraintNeighborVertex (iterable template vertex ids).
template <class L>
void assemble_local(const std::vector<L>& v) {
  TemplateAssembly& a = reference_template();
  ConstraintSet& cs = a.set;
  cs.vertex_count = static_cast<uint32_t>(v.size());
  cs.vertex_label.assign(v.size(), 0);
  cs.local.assign(v.size(), LocalConstraint{});
  for (const auto& lc : v) {
    LocalConstraint& o = cs.local.at(lc.originVertex);
    o.origin = static_cast<uint32_t>(lc.originVertex);
    o.origin_label = static_cast<uint64_t>(lc.originLabel);
    for (const auto& kv : lc.localConstraint) o.need.emplace_back(kv.first, static_cast<uint32_t>(kv.second));
    for (const auto& n : lc.localConstraintNeighborVertex) o.neighbor_mask |= 1ull << n;
    cs.vertex_label[lc.originVertex] = o.origin_label;
  }
  ++a.version;
}

// Cycles (MultipleLabelConstraintCircular: vertexIndexVector, vertexLabelVector).
template <class C>
void assemble_cycles(const std::vector<C>& v) {
  TemplateAssembly& a = reference_template();
  a.set.cycles.clear();
  for (const auto& c : v) {
    CycleConstraint o;
    o.vertex.assign(c.vertexIndexVector.begin(), c.vertexIndexVector.end());
    o.label.assign(c.vertexLabelVector.begin(), c.vertexLabelVector.end());
    a.set.cycles.push_back(o);
  }
  ++a.version;
}

// Paths (MultipleLabelConstraintPath: initialLabel, vertexIndexVector, vertexLabelVector).
template <class P>
void assemble_paths(const std::vector<P>& v) {
  TemplateAssembly& a = reference_template();
  a.set.paths.clear();
  for (const auto& p : v) {
    PathConstraint o;
    o.initial_label = static_cast<uint64_t>(p.initialLabel);
    o.vertex.assign(p.vertexIndexVector.begin(), p.vertexIndexVector.end());
    o.label.assign(p.vertexLabelVector.begin(), p.vertexLabelVector.end());
    a.set.paths.push_back(o);
  }
  ++a.version;
}

// A template constraint's walks in its walkMap's iteration order
// (MultipleLabelConstraintTemplate::walkMap, Walk: moveVector,
// vertexIndexVector, moveBackIndexVector; move numbering as pj_move).
template <class T>
TemplateConstraint template_walks(const T& t, bool first_only) {
  TemplateConstraint o;
  for (auto it = t.walkMap.cbegin(); it != t.walkMap.cend(); ++it) {
    Walk w;
    for (const auto& m : it->second.moveVector) w.move.push_back(static_cast<uint8_t>(m));
    w.vertex.assign(it->second.vertexIndexVector.begin(), it->second.vertexIndexVector.end());
    w.move_back.assign(it->second.moveBackIndexVector.begin(), it->second.moveBackIndexVector.end());
    o.walks[static_cast<uint32_t>(it->first)] = w;
    o.order.push_back(static_cast<uint32_t>(it->first));
    if (first_only) break;
  }
  return o;
}

template <class T>
void assemble_tds(const std::vector<T>& v) {
  TemplateAssembly& a = reference_template();
  a.set.tds.clear();
  for (const auto& t : v) a.set.tds.push_back(template_walks(t, false));
  ++a.version;
}

// The enumeration follows walkMap.cbegin() (multiple_label_enumeration_cpu.tpp:133-135).
template <class T>
void assemble_enumeration(const T& t) {
  TemplateAssembly& a = reference_template();
  a.set.enumeration = template_walks(t, true);
  a.have_enumeration = true;
  ++a.version;
}

template <class T>
bool has_values(const std::vector<T>& v) {
  return !v.empty();
}
template <class T>
bool has_values(const T* p) {
  return p != nullptr;
}

// Before a step: the graph_t (TOTEM's graph_t or pj::Graph: vertex_count,
// edge_count, vertices, edges, values) and the assembled template go to the
// device once (graph_t keeps eid_t offsets of 32 or 64 bits: widened here).
template <class G>
void ensure_device(const G& graph, MultipleLabelDeviceState* s) {
  if (s->managed || s->error) return;
  const void* key = graph.edge_count ? static_cast<const void*>(&graph.edges[0]) : static_cast<const void*>(&graph);
  if (s->uploaded != key) {
    const uint64_t V = graph.vertex_count, E = graph.edge_count;
    std::vector<uint64_t> off(V + 1);
    for (uint64_t i = 0; i <= V; ++i) off[i] = static_cast<uint64_t>(graph.vertices[i]);
    std::vector<uint64_t> zeros;
    const uint64_t* labels = nullptr;
    if (V && has_values(graph.values)) {
      labels = reinterpret_cast<const uint64_t*>(&graph.values[0]);
    } else {
      zeros.assign(V, 0);
      labels = zeros.data();
    }
    static_assert(sizeof(graph.edges[0]) == 4, "vertex ids must be 32-bit");
    static_assert(sizeof(graph.values[0]) == 8, "vertex values must be 64-bit");
    if (upload_graph(s, static_cast<uint32_t>(V), E, off, E ? reinterpret_cast<const uint32_t*>(&graph.edges[0]) : nullptr,
                     labels, key))
      return;
  }
  if (s->template_version != reference_template().version) install_reference_template(s);
}
}  // namespace detail

// The step classes (multiple_label_cpu.h:34-52).  compute() takes the graph
// type the driver holds (TOTEM's graph_t at the reference's seam, pj::Graph
// in this repository's driver); setConstraintVector takes the reference's
// constraint classes (gathered for the device) or this repository's (whose
// owner installs the template itself).
template <class State>
class MultipleLabelLcc0Hip {
 public:
  template <class G>
  AlgoResults compute(const G& graph, State* s) const {
    detail::ensure_device(graph, s);
    return detail::run_device_step(s, PJ_STEP_LCC0, 0);
  }
  void setConstraintVector(const std::vector<LocalConstraint>&) {}
  template <class L>
  void setConstraintVector(const std::vector<L>& v) {
    detail::assemble_local(v);
  }
};

template <class State>
class MultipleLabelLccHip {
 public:
  template <class G>
  AlgoResults compute(const G& graph, State* s) const {
    detail::ensure_device(graph, s);
    return detail::run_device_step(s, PJ_STEP_LCC, 0);
  }
  // Never called by the driver, as in the reference (lcc_cpu.tpp:13-16).
  void resetState(State* s) { detail::reset_modified(s); }
  void setConstraintVector(const std::vector<LocalConstraint>&) {}
  template <class L>
  void setConstraintVector(const std::vector<L>& v) {
    detail::assemble_local(v);
  }
};

// Cycle / path steps advance their constraint iterator after each compute
// (cc_cpu.tpp:180, pc_cpu.tpp:190).
template <class State, int StepId>
class MultipleLabelIteratedHip {
 public:
  template <class G>
  AlgoResults compute(const G& graph, State* s) {
    detail::ensure_device(graph, s);
    AlgoResults r = detail::run_device_step(s, StepId, static_cast<uint32_t>(iterator_));
    ++iterator_;
    return r;
  }
  void resetState(State* s) { detail::reset_modified(s); }
  void setConstraintVector(const std::vector<CycleConstraint>&) { iterator_ = 0; }
  void setConstraintVector(const std::vector<PathConstraint>&) { iterator_ = 0; }
  template <class C>
  void setConstraintVector(const std::vector<C>& v) {
    iterator_ = 0;
    assemble(std::integral_constant<int, StepId>(), v);
  }
  void setConstraintIterator(const size_t& index) { iterator_ = index; }

 private:
  template <class C>
  static void assemble(std::integral_constant<int, PJ_STEP_CC>, const std::vector<C>& v) {
    detail::assemble_cycles(v);
  }
  template <class C>
  static void assemble(std::integral_constant<int, PJ_STEP_PC>, const std::vector<C>& v) {
    detail::assemble_paths(v);
  }
  size_t iterator_ = 0;
};
template <class State>
using MultipleLabelCcHip = MultipleLabelIteratedHip<State, PJ_STEP_CC>;
template <class State>
using MultipleLabelPcHip = MultipleLabelIteratedHip<State, PJ_STEP_PC>;

// TDS never advances its iterator in the reference (tds_cpu.tpp:135-244):
// every TDS step runs the constraint the iterator was last set to.
template <class State>
class MultipleLabelTdsHip {
 public:
  template <class G>
  AlgoResults compute(const G& graph, State* s) {
    detail::ensure_device(graph, s);
    return detail::run_device_step(s, PJ_STEP_TDS, static_cast<uint32_t>(iterator_));
  }
  void resetState(State* s) { detail::reset_modified(s); }
  void setConstraintVector(const std::vector<TemplateConstraint>&) { iterator_ = 0; }
  template <class T>
  void setConstraintVector(const std::vector<T>& v) {
    iterator_ = 0;
    detail::assemble_tds(v);
  }
  void setConstraintIterator(const size_t& index) { iterator_ = index; }

 private:
  size_t iterator_ = 0;
};

template <class State>
class MultipleLabelEnumerationHip {
 public:
  template <class G>
  AlgoResults compute(const G& graph, State* s) {
    detail::ensure_device(graph, s);
    return detail::run_device_step(s, PJ_STEP_ENUMERATION, 0);
  }
  void resetState(State*) {}
  void setConstraint(const TemplateConstraint&) {}
  template <class T>
  void setConstraint(const T& t) {
    detail::assemble_enumeration(t);
  }
};

// UndirectedCpu (algorithms/undirected_cpu.tpp:13-51): active slots between
// active vertices whose reverse slot is not active, counted on the host from
// the downloaded state.  The reference only calls it from code switched off
// (multiple_label_cpu.cc:168).
template <class State, class Unused = void>
class UndirectedHip {
 public:
  template <class G>
  size_t compute(const G& graph, State* s) {
    detail::ensure_device(graph, s);
    size_t directed = 0;
    for (uint64_t v = 0; v < graph.vertex_count; ++v) {
      if (!s->vertexActiveList[v]) continue;
      for (uint64_t e = graph.vertices[v]; e < graph.vertices[v + 1]; ++e) {
        if (!s->edgeActiveList[e]) continue;
        const uint64_t u = graph.edges[e];
        if (!s->vertexActiveList[u]) continue;
        bool found = false;
        for (uint64_t f = graph.vertices[u]; f < graph.vertices[u + 1] && !found; ++f)
          found = s->edgeActiveList[f] && graph.edges[f] == v;
        directed += !found;
      }
    }
    return directed;
  }
};

struct StepRecord {
  int iteration;
  MultipleLabelStep::Step step;
  AlgoResults results;
  uint64_t activeVertices, activeEdges;
  uint64_t slots;
  double wallMs, deviceMs;
  int constraint = -1;         // cycle / path / TDS constraint the step ran
  double effectiveness = 0.0;  // its estimate when the dynamic order chose it
  double statMs = 0.0;         // device time of the label statistics taken after the step
};

// Runs the step machine (multiple_label_cpu.cc:113-185) on a state whose
// context holds the graph and template.  Calls `on_step` after every step.
// With stop_before_enumeration the machine ends at the pruning fixed point;
// with stop_before_tds it ends when the first TDS step is due.
// With `prefix` (the LCC0 row and the LCC rows up to the first one that
// eliminated nothing, as a batch ran them: pj_batch_trace) the state already
// holds the result of those steps (pj_batch_export): the machine replays the
// rows without running them and continues from there.
int run_step_machine(MultipleLabelDeviceState* state, const ConstraintSet& cs,
                     const std::function<void(const StepRecord&)>& on_step, double* total_ms,
                     bool stop_before_enumeration = false, bool stop_before_tds = false,
                     const std::vector<AlgoResults>* prefix = nullptr);

// The dynamic constraint order of the reference's article driver
// (test_article/test_order_multiple_label_cpu.cc:127-290, ORDER_USE_DYNAMIC):
// LCC0, then LCC until a step eliminates nothing; there the label statistics
// of the active graph (pj_graph_stat) and the effectiveness estimate
// (choose_next) pick the next cycle, path or TDS constraint, which runs once
// and leaves its list; LCC follows every such step.  With every list empty
// the enumeration runs (USE_ENUMERATE), or the machine stops there with
// stop_before_enumeration.  That driver enters its loop with no step set
// and reaches LCC without LCC0; LCC0 runs first here, as in the plain
// machine: LCC repeats the label test, so the fixed point is the same.
int run_dynamic_order(MultipleLabelDeviceState* state, const ConstraintSet& cs,
                      const std::function<void(const StepRecord&)>& on_step, double* total_ms,
                      bool stop_before_enumeration, int threads);

struct Options {
  std::string graph, graphBin, vertexData, edgeData, patternDir, outputDir;
  int device = 0;
};

// MultipleLabelCpu (multiple_label_cpu.h:32-97) on the device.
class MultipleLabelHip {
 public:
  typedef MultipleLabelDeviceState StateType;
  typedef MultipleLabelLcc0Hip<StateType> Lcc0Type;
  typedef MultipleLabelLccHip<StateType> LccType;
  typedef MultipleLabelCcHip<StateType> CcType;
  typedef MultipleLabelPcHip<StateType> PcType;
  typedef MultipleLabelTdsHip<StateType> TdsType;
  typedef MultipleLabelEnumerationHip<StateType> EnumerationType;

  int allocate(const Options& options);
  int runPatternMatching();
  int free();
  const std::string& error() const { return error_; }

 private:
  void logResults(int iteration, bool logGraph);
  void log(const std::string& line, bool to_cout) const;

  Options options_;
  Graph graph_, pattern_;
  ConstraintSet constraints_;
  StateType state_;
  std::string error_;
  std::string stepName_ = "START";
  AlgoResults results_;
  double stepTime_ = 0.0, totalTime_ = 0.0;
};

}  // namespace pj

#endif
