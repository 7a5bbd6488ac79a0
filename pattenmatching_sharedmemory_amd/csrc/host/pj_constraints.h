// Host constraint generation from a template (pattern) graph.
//
// Produces, in the reference's order, the local constraints
// (multiple_label_generate_constraint_local.cc:12-35), the normalised cycle
// constraints (multiple_label_generate_constraint_circular.cc:58-86), the
// path constraints (multiple_label_generate_constraint_path.cc:56-82), the
// template-driven-search constraints with their walks
// (multiple_label_generate_constraint_template.cc:13-145,
//  multiple_label_constraint_template.cc:121-169, walk.cc:16-60) and the
// enumeration walk (multiple_label_generate_constraint_enumeration.cc:13-18),
// and flattens them into a pj_template_desc.
//
// Iteration order of the template and walk maps is part of the result (it
// fixes which TDS constraint runs and which walk enumeration follows), so
// those maps are std::unordered_map with the same key types, filled by the
// same sequence of insertions as the reference.
#ifndef PJ_CONSTRAINTS_H
#define PJ_CONSTRAINTS_H

#include <cstdint>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/prunejuice.h"
#include "pj_graph.h"

namespace pj {

struct LocalConstraint {
  uint32_t origin = 0;
  uint64_t origin_label = 0;
  std::vector<std::pair<uint64_t, uint32_t>> need;  // (neighbour label, count), first-seen order
  uint64_t neighbor_mask = 0;
};

struct CycleConstraint {
  std::vector<uint32_t> vertex;
  std::vector<uint64_t> label;
  bool operator==(const CycleConstraint& o) const { return vertex == o.vertex; }
};

struct PathConstraint {
  uint64_t initial_label = 0;
  std::vector<uint32_t> vertex;
  std::vector<uint64_t> label;
  bool operator==(const PathConstraint& o) const {
    return initial_label == o.initial_label && vertex == o.vertex;
  }
};

struct Walk {
  std::vector<uint8_t> move;
  std::vector<uint32_t> vertex;
  std::vector<uint64_t> label;
  std::vector<uint32_t> move_back;
  uint32_t vertex_length = 0;
  void add_vertex(const Graph& pattern, uint32_t v);
  void add_check(const Graph& pattern, uint32_t v, uint32_t history_index);
  void add_move_back(const Graph& pattern, uint32_t v, uint32_t history_index);
  void mark_stored();
  std::string to_string() const;
};

struct TemplateConstraint {
  enum Origin { EMPTY, CC, PC, TDS_CC, TDS_PC, TDS_SUB, TDS_FULL };
  Origin origin = EMPTY;
  size_t length = 0;  // template vertices covered
  std::unordered_map<uint32_t, std::set<uint32_t>> adjacency;
  std::unordered_map<uint32_t, Walk> walks;
  // Root order of `walks` when set from outside (the reference's walkMap
  // iteration order); empty: the order of `walks` itself.
  std::vector<uint32_t> order;

  static TemplateConstraint from_cycle(const CycleConstraint& c);
  static TemplateConstraint from_path(const PathConstraint& p);
  static TemplateConstraint from_pattern(const Graph& pattern);
  static TemplateConstraint merge(const TemplateConstraint& a, const TemplateConstraint& b);
  bool shares_vertex(const TemplateConstraint& o) const;
  bool shares_edge(const TemplateConstraint& o) const;
  bool same_as(const TemplateConstraint& o) const {
    return length == o.length && adjacency == o.adjacency;
  }
  void build_walks(const Graph& pattern);
};

struct ConstraintSet {
  uint32_t vertex_count = 0;
  std::vector<uint64_t> vertex_label;
  std::vector<LocalConstraint> local;
  std::vector<CycleConstraint> cycles;
  std::vector<PathConstraint> paths;
  std::vector<TemplateConstraint> tds;
  TemplateConstraint enumeration;

  // Flattened storage backing desc().
  std::vector<uint64_t> f_nbr, f_need_label;
  std::vector<uint32_t> f_need_off, f_need_count;
  std::vector<uint32_t> f_cycle_off, f_cycle_vtx, f_path_off, f_path_vtx, f_tds_off;
  std::vector<uint64_t> f_path_label;
  std::vector<pj_walk_desc> f_walks;
  pj_template_desc f_desc{};

  // Generates every constraint of `pattern` (labels required).
  int build(const Graph& pattern, std::string* err);
  const pj_template_desc* desc() const { return &f_desc; }
  // Human-readable dump in the spirit of the reference's log.totem prints.
  std::string dump() const;
  // Rebuilds desc() from the vectors above (build() calls it; call it after
  // filling them directly).
  void flatten();
};

}  // namespace pj

#endif
