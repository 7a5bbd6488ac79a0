// Test infrastructure: the step classes of pj_driver.h driven through the
// reference's seam.
//
// The reference's MultipleLabelCpu (src/patternmatching/multiple_label_cpu.h:
// 34-97, multiple_label_cpu.cc:20-297) holds a graph_t, a state object and
// one object per step type chosen by typedef, hands the constraint vectors of
// its generators to the step objects (multiple_label_cpu.cc:84-95), calls
// state.allocate(V, E, P) (:52), compute(*graph, &state) per step (:141-161)
// and reads the state's counters and its vertexActiveList / edgeActiveList /
// vertexPatternMatch (:187-297).  This file declares stand-ins of the
// reference's types with the field names the reference declares
// (totem_graph.h:137-180; multiple_label_constraint_local.h:18-32,
// _circular.h:14-26, _path.h:14-26, _template.h:58-59; walk.h:14-37), fills
// them from this repository's generator, and runs a step loop that makes
// those same calls on the pj:: step classes, so that the typedef swap shown in
// INTEGRATION.md is compiled and run here.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <set>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/prunejuice_host.h"
#include "../../pattenmatching_sharedmemory_amd/csrc/host/pj_driver.h"

namespace ref {

typedef uint32_t vid_t;
typedef uint64_t weight_t;
typedef uint32_t pvid_t;

template <class eid_t>
struct graph_t {
  vid_t vertex_count;
  eid_t edge_count;
  eid_t* vertices;
  vid_t* edges;
  weight_t* values;
  bool valued;
};

// FixedBitmap stand-in: iterates its ids in ascending order.
struct IdSet {
  std::set<pvid_t> ids;
  std::set<pvid_t>::const_iterator begin() const { return ids.begin(); }
  std::set<pvid_t>::const_iterator end() const { return ids.end(); }
};

struct MultipleLabelConstraintLocal {
  pvid_t originVertex;
  weight_t originLabel;
  std::unordered_map<weight_t, size_t> localConstraint;
  IdSet localConstraintNeighborVertex;
};

struct MultipleLabelConstraintCircular {
  std::vector<vid_t> vertexIndexVector;
  std::vector<weight_t> vertexLabelVector;
};

struct MultipleLabelConstraintPath {
  weight_t initialLabel;
  std::vector<vid_t> vertexIndexVector;
  std::vector<weight_t> vertexLabelVector;
};

struct Walk {
  enum MoveType { E_MOVE_BACK, E_CHECK, E_VERTEX_STORE, E_VERTEX_NO_STORE };
  std::vector<MoveType> moveVector;
  std::vector<vid_t> vertexIndexVector;
  std::vector<size_t> moveBackIndexVector;
};

// walkMap keeps the reference's iteration order: a sequence of (root, walk)
// pairs offers the cbegin / cend / first / second the step classes use.
struct MultipleLabelConstraintTemplate {
  std::vector<std::pair<vid_t, Walk>> walkMap;
};

struct Constraints {
  std::vector<MultipleLabelConstraintLocal> local;
  std::vector<MultipleLabelConstraintCircular> circular;
  std::vector<MultipleLabelConstraintPath> path;
  std::vector<MultipleLabelConstraintTemplate> tmpl;
  MultipleLabelConstraintTemplate enumeration;
};

MultipleLabelConstraintTemplate from_walks(const pj_walk_desc* w, uint32_t n) {
  MultipleLabelConstraintTemplate t;
  for (uint32_t k = 0; k < n; ++k) {
    Walk x;
    for (uint32_t i = 0; i < w[k].length; ++i) {
      x.moveVector.push_back(static_cast<Walk::MoveType>(w[k].move[i]));
      x.vertexIndexVector.push_back(w[k].vertex[i]);
      x.moveBackIndexVector.push_back(w[k].move_back[i]);
    }
    t.walkMap.emplace_back(w[k].root, x);
  }
  return t;
}

// The reference generators' output, taken from this repository's generator
// (pinned against the reference's fixtures by tests/test_constraints.py and
// tests/test_oracle_pins.py).
Constraints from_desc(const pj_template_desc& d) {
  Constraints c;
  for (uint32_t p = 0; p < d.vertex_count; ++p) {
    MultipleLabelConstraintLocal lc;
    lc.originVertex = p;
    lc.originLabel = d.vertex_label[p];
    for (uint32_t i = d.need_offset[p]; i < d.need_offset[p + 1]; ++i) lc.localConstraint[d.need_label[i]] = d.need_count[i];
    for (uint32_t n = 0; n < 64; ++n)
      if ((d.neighbor_mask[p] >> n) & 1) lc.localConstraintNeighborVertex.ids.insert(n);
    c.local.push_back(lc);
  }
  for (uint32_t k = 0; k < d.cycle_count; ++k) {
    MultipleLabelConstraintCircular cc;
    for (uint32_t i = d.cycle_offset[k]; i < d.cycle_offset[k + 1]; ++i) {
      cc.vertexIndexVector.push_back(d.cycle_vertex[i]);
      cc.vertexLabelVector.push_back(d.vertex_label[d.cycle_vertex[i]]);
    }
    c.circular.push_back(cc);
  }
  for (uint32_t k = 0; k < d.path_count; ++k) {
    MultipleLabelConstraintPath pc;
    pc.initialLabel = d.path_initial_label[k];
    for (uint32_t i = d.path_offset[k]; i < d.path_offset[k + 1]; ++i) {
      pc.vertexIndexVector.push_back(d.path_vertex[i]);
      pc.vertexLabelVector.push_back(d.vertex_label[d.path_vertex[i]]);
    }
    c.path.push_back(pc);
  }
  for (uint32_t k = 0; k < d.tds_count; ++k)
    c.tmpl.push_back(from_walks(d.walks + d.tds_walk_offset[k], d.tds_walk_offset[k + 1] - d.tds_walk_offset[k]));
  c.enumeration = from_walks(&d.enumeration_walk, 1);
  return c;
}

// The typedef block of the seam, swapped to the device step classes.
typedef pj::MultipleLabelDeviceState StateType;
typedef pj::MultipleLabelLcc0Hip<StateType> Lcc0Type;
typedef pj::MultipleLabelLccHip<StateType> LccType;
typedef pj::MultipleLabelCcHip<StateType> CcType;
typedef pj::MultipleLabelPcHip<StateType> PcType;
typedef pj::MultipleLabelTdsHip<StateType> TdsType;
typedef pj::MultipleLabelEnumerationHip<StateType> EnumerationType;
typedef pj::UndirectedHip<StateType> UndirectedType;

struct Steps {
  Lcc0Type lcc0;
  LccType lcc;
  CcType cc;
  PcType pc;
  TdsType tds;
  EnumerationType enumeration;
  UndirectedType undirected;

  void set(const Constraints& c) {
    lcc0.setConstraintVector(c.local);
    lcc.setConstraintVector(c.local);
    cc.setConstraintVector(c.circular);
    pc.setConstraintVector(c.path);
    tds.setConstraintVector(c.tmpl);
    enumeration.setConstraint(c.enumeration);
  }
};

bool same_walk(const pj_walk_desc& a, const pj_walk_desc& b) {
  return a.root == b.root && a.length == b.length && !std::memcmp(a.move, b.move, a.length) &&
         !std::memcmp(a.vertex, b.vertex, 4 * a.length) && !std::memcmp(a.move_back, b.move_back, 4 * a.length);
}

// Flattened template equality (need lists compared as label -> count maps:
// the reference keeps them in an unordered_map).
bool same_desc(const pj_template_desc& a, const pj_template_desc& b) {
  if (a.vertex_count != b.vertex_count || a.cycle_count != b.cycle_count || a.path_count != b.path_count ||
      a.tds_count != b.tds_count || a.walk_count != b.walk_count)
    return false;
  for (uint32_t p = 0; p < a.vertex_count; ++p) {
    if (a.vertex_label[p] != b.vertex_label[p] || a.neighbor_mask[p] != b.neighbor_mask[p]) return false;
    std::unordered_map<uint64_t, uint32_t> na, nb;
    for (uint32_t i = a.need_offset[p]; i < a.need_offset[p + 1]; ++i) na[a.need_label[i]] = a.need_count[i];
    for (uint32_t i = b.need_offset[p]; i < b.need_offset[p + 1]; ++i) nb[b.need_label[i]] = b.need_count[i];
    if (na != nb) return false;
  }
  for (uint32_t k = 0; k <= a.cycle_count; ++k)
    if (a.cycle_offset[k] != b.cycle_offset[k]) return false;
  for (uint32_t i = 0; i < a.cycle_offset[a.cycle_count]; ++i)
    if (a.cycle_vertex[i] != b.cycle_vertex[i]) return false;
  for (uint32_t k = 0; k <= a.path_count; ++k)
    if (a.path_offset[k] != b.path_offset[k]) return false;
  for (uint32_t k = 0; k < a.path_count; ++k)
    if (a.path_initial_label[k] != b.path_initial_label[k]) return false;
  for (uint32_t i = 0; i < a.path_offset[a.path_count]; ++i)
    if (a.path_vertex[i] != b.path_vertex[i]) return false;
  for (uint32_t k = 0; k <= a.tds_count; ++k)
    if (a.tds_walk_offset[k] != b.tds_walk_offset[k]) return false;
  for (uint32_t k = 0; k < a.walk_count; ++k)
    if (!same_walk(a.walks[k], b.walks[k])) return false;
  return same_walk(a.enumeration_walk, b.enumeration_walk);
}

template <class eid_t>
int run(uint32_t V, uint64_t E, const uint64_t* off, const uint32_t* edges, const uint64_t* labels,
        const pj_template_desc& d, pjh_row* rows, int max_rows, uint8_t* vactive, uint8_t* eactive, uint64_t* cand,
        uint64_t* undirected) {
  std::vector<eid_t> vertices(off, off + V + 1);
  std::vector<vid_t> ed(edges, edges + E);
  std::vector<weight_t> values(labels, labels + V);
  graph_t<eid_t> g{V, static_cast<eid_t>(E), vertices.data(), ed.data(), values.data(), true};
  graph_t<eid_t>* graph = &g;
  const Constraints c = from_desc(d);
  Steps steps;
  StateType state;
  state.allocate(graph->vertex_count, graph->edge_count, d.vertex_count);
  steps.set(c);
  pj::MultipleLabelStep machine;
  machine.initStepCc(static_cast<int>(c.circular.size()));
  machine.initStepPc(static_cast<int>(c.path.size()));
  machine.initStepTds(static_cast<int>(c.tmpl.size()));
  pj::AlgoResults results;
  pj::MultipleLabelStep::Step step;
  std::string name;
  bool finished = false;
  int it = 0;
  machine.getNextStep(results, &step, &name);
  while (!finished && it < 1000) {
    switch (step) {
      case pj::MultipleLabelStep::E_CC:
        steps.cc.resetState(&state);
        break;
      case pj::MultipleLabelStep::E_PC:
        steps.pc.resetState(&state);
        break;
      case pj::MultipleLabelStep::E_TDS:
        steps.tds.resetState(&state);
        break;
      default:
        break;
    }
    switch (step) {
      case pj::MultipleLabelStep::E_LCC0:
        results = steps.lcc0.compute(*graph, &state);
        break;
      case pj::MultipleLabelStep::E_LCC:
        results = steps.lcc.compute(*graph, &state);
        break;
      case pj::MultipleLabelStep::E_CC:
        results = steps.cc.compute(*graph, &state);
        break;
      case pj::MultipleLabelStep::E_PC:
        results = steps.pc.compute(*graph, &state);
        break;
      case pj::MultipleLabelStep::E_TDS:
        results = steps.tds.compute(*graph, &state);
        break;
      case pj::MultipleLabelStep::E_ENUMERATION:
        results = steps.enumeration.compute(*graph, &state);
        break;
      default:
        break;
    }
    if (state.error) {
      std::fprintf(stderr, "seam driver: %s\n", state.message.c_str());
      state.free();
      return -1;
    }
    if (it < max_rows)
      rows[it] = pjh_row{it, static_cast<int32_t>(step), results.vertexEliminated, results.edgeEliminated,
                         results.matchEliminated, results.enumeration, state.graphActiveVertexCount,
                         state.graphActiveEdgeCount, 0, 0.0};
    finished = machine.getNextStep(results, &step, &name);
    ++it;
  }
  // what printActiveGraph / printActiveVertex read
  for (vid_t v = 0; v < graph->vertex_count; ++v) {
    vactive[v] = static_cast<uint8_t>(state.vertexActiveList[v]);
    uint64_t bits = 0;
    for (const auto& p : state.vertexPatternMatch[v]) bits |= 1ull << p;
    cand[v] = state.vertexActiveList[v] ? bits : 0;
  }
  for (eid_t e = 0; e < graph->edge_count; ++e) eactive[e] = static_cast<uint8_t>(state.edgeActiveList[e]);
  *undirected = steps.undirected.compute(*graph, &state);
  state.free();
  return it;
}

}  // namespace ref

extern "C" {

// 1 when the step classes, given the reference-shaped constraint objects of
// the pattern (P vertices, n undirected edges src[i]-dst[i], labels), gather
// exactly the template the generator flattens; 0 when they differ, -1 on a
// bad pattern.  Needs no GPU.
int pjs_assembly_matches(uint32_t P, uint32_t n, const uint32_t* src, const uint32_t* dst, const uint64_t* labels) {
  pjh_graph* p = nullptr;
  pjh_constraints* c = nullptr;
  char err[256];
  if (pjh_pattern_from_edges(P, n, src, dst, labels, &p) || pjh_constraints_build(p, &c, err, sizeof(err))) return -1;
  const pj_template_desc& d = *pjh_constraints_desc(c);
  ref::Steps steps;
  steps.set(ref::from_desc(d));
  auto& a = pj::detail::reference_template();
  a.set.flatten();
  const int same = ref::same_desc(*a.set.desc(), d) ? 1 : 0;
  pjh_constraints_free(c);
  pjh_graph_free(p);
  return same;
}

// Runs the step loop on the device through the seam with a graph_t whose
// offsets are 32-bit (eid64 = 0) or 64-bit; returns the number of steps.
int pjs_run(uint32_t V, uint64_t E, const uint64_t* off, const uint32_t* edges, const uint64_t* labels, uint32_t P,
            uint32_t n, const uint32_t* src, const uint32_t* dst, const uint64_t* plabels, int eid64, pjh_row* rows,
            int max_rows, uint8_t* vactive, uint8_t* eactive, uint64_t* cand, uint64_t* undirected) {
  pjh_graph* p = nullptr;
  pjh_constraints* c = nullptr;
  char err[256];
  if (pjh_pattern_from_edges(P, n, src, dst, plabels, &p) || pjh_constraints_build(p, &c, err, sizeof(err))) return -1;
  const pj_template_desc& d = *pjh_constraints_desc(c);
  const int r = eid64 ? ref::run<uint64_t>(V, E, off, edges, labels, d, rows, max_rows, vactive, eactive, cand, undirected)
                      : ref::run<uint32_t>(V, E, off, edges, labels, d, rows, max_rows, vactive, eactive, cand, undirected);
  pjh_constraints_free(c);
  pjh_graph_free(p);
  return r;
}

}  // extern "C"
