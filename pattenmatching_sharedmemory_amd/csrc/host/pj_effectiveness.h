// Effectiveness estimate of the pruning constraints and the dynamic
// constraint order built on it.
//
// The reference's article driver with ORDER_USE_DYNAMIC
// (test_article/test_order_multiple_label_cpu.cc:170-290) takes label
// statistics of the active graph each time LCC reaches a fixed point, scores
// every cycle, path and TDS constraint not yet run with
// MultipleLabelConstraintEffectiveness::computeBacktrack
// (multiple_label/multiple_label_constraint_effectiveness.cc:226-340, 570-696,
// 972-1118): the candidate matches the constraint is expected to eliminate
// over the work it is expected to cost, both from probabilities of finding a
// labelled neighbour (:1120-1230).  The best constraint runs next and is then
// dropped from its list (the step classes that driver uses erase the
// constraint they ran, e.g. multiple_label_test/multiple_label_cc_backtrack_
// cpu.tpp:192).  The statistics come from the device (pj_graph_stat); the
// estimate is host arithmetic in double precision, in the reference's order
// of operations.
//
// The reference looks its maps up with at(), which throws where a label or
// label pair has no entry (no active vertex or slot of it); here such an
// entry reads 0, and the IEEE arithmetic that follows (0/0, x/0) decides:
// a NaN effectiveness never wins the comparison, as in the reference's
// "greater than" test.
#ifndef PJ_EFFECTIVENESS_H
#define PJ_EFFECTIVENESS_H

#include <cstdint>
#include <vector>

#include "../../../include/prunejuice.h"
#include "pj_constraints.h"

namespace pj {

// Map lookups into a pj_label_stat by label value (0 where the reference's
// map has no entry).
class LabelStat {
 public:
  explicit LabelStat(const pj_label_stat& s) : s_(s) {}
  double vertices(uint64_t label) const;                       // vertexLabelTotalNumberMap
  double edges(uint64_t from, uint64_t to) const;              // edgeLabelTotalNumberMap
  double outbound(uint64_t from) const;                        // edgeOutboundLabelTotalNumberMap
  double at_least(uint64_t from, uint64_t to, size_t k) const; // vertexLabelWithAtLeastNeighborLabelMap

 private:
  int code(uint64_t label) const;
  const pj_label_stat& s_;
};

// MultipleLabelConstraintEffectiveness's three results.
struct Effectiveness {
  double match_eliminated = 0.0;  // approximateMatchEliminated
  double cost_eliminated = 0.0;   // approximateCostEliminated
  double effectiveness = 0.0;     // approximateEffectiveness
};

// computeBacktrack of a cycle, a path and a TDS constraint; `threads` is the
// reference's omp_get_max_threads() divisor of the cost (the same for every
// constraint, so it does not change the order).
Effectiveness effectiveness(const LabelStat& st, const CycleConstraint& c, int threads);
Effectiveness effectiveness(const LabelStat& st, const PathConstraint& c, int threads);
Effectiveness effectiveness(const LabelStat& st, const TemplateConstraint& c, int threads);

// The choice at an LCC fixed point (test_order_multiple_label_cpu.cc:
// 184-274) among the constraints not yet run (original indices, in list
// order): the cycle, path or TDS constraint of greatest effectiveness, ties
// to the earlier constraint and to cycles, then paths; PJ_STEP_ENUMERATION
// when every list is empty.
struct NextConstraint {
  int step = PJ_STEP_ENUMERATION;
  uint32_t index = 0;
  double effectiveness = 0.0;
};
NextConstraint choose_next(const LabelStat& st, const ConstraintSet& cs, const std::vector<uint32_t>& cycles_left,
                           const std::vector<uint32_t>& paths_left, const std::vector<uint32_t>& tds_left,
                           int threads);

}  // namespace pj

#endif
