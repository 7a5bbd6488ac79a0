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
// (multiple_label_common_cpu.h:23-59): the arrays live in the context, the
// counters that logResults prints are mirrored here.
struct MultipleLabelDeviceState {
  pj_ctx* ctx = nullptr;
  uint64_t graphVertexCount = 0;
  uint64_t graphActiveVertexCount = 0;
  uint64_t graphEdgeCount = 0;
  uint64_t graphActiveEdgeCount = 0;
  uint32_t patternVertexCount = 0;
  int error = 0;  // set when a device call fails (compute() has no error channel)
  std::string message;
};

namespace detail {
AlgoResults run_device_step(MultipleLabelDeviceState* s, int step, uint32_t index);
void reset_modified(MultipleLabelDeviceState* s);
}  // namespace detail

template <class State>
class MultipleLabelLcc0Hip {
 public:
  AlgoResults compute(const Graph&, State* s) const { return detail::run_device_step(s, PJ_STEP_LCC0, 0); }
  void setConstraintVector(const std::vector<LocalConstraint>&) {}
};

template <class State>
class MultipleLabelLccHip {
 public:
  AlgoResults compute(const Graph&, State* s) const { return detail::run_device_step(s, PJ_STEP_LCC, 0); }
  // Never called by the driver, as in the reference (lcc_cpu.tpp:13-16).
  void resetState(State* s) { detail::reset_modified(s); }
  void setConstraintVector(const std::vector<LocalConstraint>&) {}
};

// Cycle / path steps advance their constraint iterator after each compute
// (cc_cpu.tpp:180, pc_cpu.tpp:190).
template <class State, int StepId>
class MultipleLabelIteratedHip {
 public:
  AlgoResults compute(const Graph&, State* s) {
    AlgoResults r = detail::run_device_step(s, StepId, static_cast<uint32_t>(iterator_));
    ++iterator_;
    return r;
  }
  void resetState(State* s) { detail::reset_modified(s); }
  template <class C>
  void setConstraintVector(const std::vector<C>&) {
    iterator_ = 0;
  }
  void setConstraintIterator(const size_t& index) { iterator_ = index; }

 private:
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
  AlgoResults compute(const Graph&, State* s) {
    return detail::run_device_step(s, PJ_STEP_TDS, static_cast<uint32_t>(iterator_));
  }
  void resetState(State* s) { detail::reset_modified(s); }
  void setConstraintVector(const std::vector<TemplateConstraint>&) { iterator_ = 0; }
  void setConstraintIterator(const size_t& index) { iterator_ = index; }

 private:
  size_t iterator_ = 0;
};

template <class State>
class MultipleLabelEnumerationHip {
 public:
  AlgoResults compute(const Graph&, State* s) { return detail::run_device_step(s, PJ_STEP_ENUMERATION, 0); }
  void resetState(State*) {}
  void setConstraint(const TemplateConstraint&) {}
};

struct StepRecord {
  int iteration;
  MultipleLabelStep::Step step;
  AlgoResults results;
  uint64_t activeVertices, activeEdges;
  uint64_t slots;
  double wallMs, deviceMs;
};

// Runs the step machine (multiple_label_cpu.cc:113-185) on a state whose
// context holds the graph and template.  Calls `on_step` after every step.
// With stop_before_enumeration the machine ends at the pruning fixed point;
// with stop_before_tds it ends when the first TDS step is due.
int run_step_machine(MultipleLabelDeviceState* state, const ConstraintSet& cs,
                     const std::function<void(const StepRecord&)>& on_step, double* total_ms,
                     bool stop_before_enumeration = false, bool stop_before_tds = false);

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
