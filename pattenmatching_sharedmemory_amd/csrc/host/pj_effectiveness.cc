// Effectiveness estimate and dynamic constraint order (see pj_effectiveness.h).
// Each function keeps the reference's expressions and their order of
// evaluation, so that the doubles agree bit for bit with a restatement that
// does the same.
#include "pj_effectiveness.h"

#include <cmath>

namespace pj {

int LabelStat::code(uint64_t label) const {
  for (uint32_t k = 0; k < s_.label_count && k < 64; ++k)
    if (s_.label[k] == label) return static_cast<int>(k);
  return -1;
}

double LabelStat::vertices(uint64_t label) const {
  const int c = code(label);
  return c < 0 ? 0.0 : static_cast<double>(s_.vertex[c]);
}

double LabelStat::edges(uint64_t from, uint64_t to) const {
  const int f = code(from), t = code(to);
  return f < 0 || t < 0 ? 0.0 : static_cast<double>(s_.edge[f * 64 + t]);
}

double LabelStat::outbound(uint64_t from) const {
  const int f = code(from);
  return f < 0 ? 0.0 : static_cast<double>(s_.outbound[f]);
}

double LabelStat::at_least(uint64_t from, uint64_t to, size_t k) const {
  const int f = code(from), t = code(to);
  if (f < 0 || t < 0 || k == 0) return 0.0;
  const uint32_t i = static_cast<uint32_t>(f) * 64 + static_cast<uint32_t>(t);
  if (k > s_.at_least_max[i]) return 0.0;
  return static_cast<double>(s_.at_least[s_.at_least_offset[i] + k - 1]);
}

namespace {

// multiple_label_constraint_effectiveness.cc:1120-1230 and .tpp:10-14, with
// the knowing counts the backtrack estimates use (0 edges, 0 vertices).

// probaFindVertex (:1120-1133): the share of label-`from` vertices with a
// neighbour of label `to`.
double proba_find_vertex(const LabelStat& st, uint64_t from, uint64_t to) {
  const double full = st.vertices(from);
  const double set_size = full;  // knowingEdgeNumber <= 1
  const double proba_different = std::pow((full - 0 - 0) / (full - 0), 1);
  return st.at_least(from, to, 1) / set_size * proba_different;
}

// probaFindOnePathKnowing (:1141-1149): one of `from_number` vertices finds one.
double proba_find_one_path(const LabelStat& st, double from_number, uint64_t from, uint64_t to) {
  return 1 - std::pow(1 - proba_find_vertex(st, from, to), from_number);
}

// probaFindSameVertex (:1151-1156).
double proba_find_same_vertex(const LabelStat& st, double from_number, uint64_t label) {
  const double set_size = st.vertices(label);
  return 1 - std::pow((set_size - 1) / set_size, from_number);
}

// probaFindDifferentVertex (:1158-1163).
double proba_find_different_vertex(const LabelStat& st, double from_number, uint64_t label) {
  const double set_size = st.vertices(label);
  return 1 - std::pow((1.) / set_size, from_number);
}

// averageNumberOfVertex (:1165-1171): slots read from `from_number` vertices.
double average_vertex(const LabelStat& st, double from_number, uint64_t from) {
  return from_number * st.outbound(from) / st.vertices(from);
}

// averageNumberOfVertexKnowing (:1180-1200): neighbours of label `to` reached.
// Its proba_different is an int of value 1 here (no known vertex).
double average_vertex_knowing(const LabelStat& st, double from_number, uint64_t from, uint64_t to) {
  const double tries = from_number * st.edges(from, to) / st.vertices(from);
  return tries * 1;
}

// averageNumberOfUniqueVertexKnowing (:1209-1230): the same, capped at the
// number of label-`to` vertices (averageDifferentVertexFromTriesApproximate
// Strong, .tpp:10-14: tries <= n ? tries : n).
double average_unique_vertex(const LabelStat& st, double from_number, uint64_t from, uint64_t to) {
  const double tries = from_number * st.edges(from, to) / st.vertices(from);
  const double max_to = st.vertices(to);
  return (tries <= max_to ? tries : max_to) * 1;
}

// The "known edge" update (next label equals the one before it).
double known_edge_unique(const LabelStat& st, double n, uint64_t from, uint64_t to) {
  return 1 - st.edges(from, to) / st.vertices(from) + average_unique_vertex(st, n, from, to);
}

}  // namespace

// computeBacktrack, cycle (:226-340).
Effectiveness effectiveness(const LabelStat& st, const CycleConstraint& c, int threads) {
  Effectiveness r;
  const size_t L = c.label.size();
  const auto& lab = c.label;
  std::vector<double> finding;
  for (size_t i = 0; i < L; ++i) {
    double fp = 1.;
    double unique = 1;
    for (size_t j = 0; j < L; ++j) {
      const uint64_t from = lab[(i + j) % L], to = lab[(i + j + 1) % L];
      if (j == 0) {
        fp *= proba_find_one_path(st, unique, from, to);
        unique = average_unique_vertex(st, unique, from, to);
      } else {
        const uint64_t previous = lab[(i + j - 1) % L];
        if (previous != to) {
          fp *= proba_find_one_path(st, unique, from, to);
          unique = average_unique_vertex(st, unique, from, to);
        } else {
          unique = known_edge_unique(st, unique, from, to);
        }
      }
    }
    fp *= proba_find_same_vertex(st, unique, lab[(i + L) % L]);  // closure
    finding.push_back(fp);
  }
  for (size_t i = 0; i < L; ++i) r.match_eliminated += st.vertices(lab[i]) * (1 - finding[i]);
  for (size_t i = 0; i < L; ++i) {
    double traversed = 1;
    double cost = traversed;
    for (size_t j = 0; j + 1 < L; ++j) {
      const uint64_t from = lab[(i + j) % L], to = lab[(i + j + 1) % L];
      cost += average_vertex(st, traversed, from);
      if (j == 0) {
        traversed = average_vertex_knowing(st, traversed, from, to);
      } else {
        const uint64_t previous = lab[(i + j - 1) % L];
        traversed = previous != to ? average_unique_vertex(st, traversed, from, to)
                                   : known_edge_unique(st, traversed, from, to);
      }
    }
    cost += average_vertex(st, traversed, lab[(i + L - 1) % L]);  // closure
    cost += traversed;
    double increase = st.vertices(lab[i]) * cost * (1 - finding[i]);
    increase /= threads;
    increase /= 1 + static_cast<double>(L - 1) / 2 * (finding[i]);
    r.cost_eliminated += increase;
  }
  r.effectiveness = r.match_eliminated / r.cost_eliminated;
  return r;
}

// computeBacktrack, path (:570-696).  Both directions; the end labels read
// are positions 0 and 1 (closure and match count) as in the reference.
Effectiveness effectiveness(const LabelStat& st, const PathConstraint& c, int threads) {
  Effectiveness r;
  const size_t L = c.label.size();
  const auto& lab = c.label;
  const size_t steps = L >= 2 ? L - 2 : 0;
  std::vector<double> finding;
  for (int i = 0; i <= 1; ++i) {
    const bool reverse = i == 1;
    double fp = 1.;
    double unique = 1;
    for (size_t j = 0; j < steps; ++j) {
      const size_t cur = reverse ? L - j - 1 : j, next = reverse ? L - j - 2 : j + 1;
      const uint64_t from = lab[cur], to = lab[next];
      if (j == 0) {
        fp *= proba_find_one_path(st, unique, from, to);
        unique = average_unique_vertex(st, unique, from, to);
      } else {
        const uint64_t previous = lab[reverse ? L - j : j - 1];
        if (previous != to) {
          fp *= proba_find_one_path(st, unique, from, to);
          unique = average_unique_vertex(st, unique, from, to);
        } else {
          unique = known_edge_unique(st, unique, from, to);
        }
      }
    }
    fp *= proba_find_different_vertex(st, unique, lab[(i + L) % L]);  // closure
    finding.push_back(fp);
  }
  for (int i = 0; i <= 1; ++i) r.match_eliminated += st.vertices(lab[i]) * (1 - finding[i]);
  for (int i = 0; i <= 1; ++i) {
    const bool reverse = i == 1;
    double traversed = 1;
    double cost = traversed;
    for (size_t j = 0; j < steps; ++j) {
      const size_t cur = reverse ? L - j - 1 : j, next = reverse ? L - j - 2 : j + 1;
      const uint64_t from = lab[cur], to = lab[next];
      cost += average_vertex(st, traversed, from);
      if (j == 0) {
        traversed = average_vertex_knowing(st, traversed, from, to);
      } else {
        const uint64_t previous = lab[reverse ? L - j : j - 1];
        traversed = previous != to ? average_unique_vertex(st, traversed, from, to)
                                   : known_edge_unique(st, traversed, from, to);
      }
    }
    cost += average_vertex(st, traversed, lab[(i + L - 1) % L]);  // closure
    cost += traversed;
    double increase = st.vertices(lab[i]) * cost * (1 - finding[i]);
    increase /= threads;
    increase /= 1 + static_cast<double>(L - 1) / 2 * (finding[i]);
    r.cost_eliminated += increase;
  }
  r.effectiveness = r.match_eliminated / r.cost_eliminated;
  return r;
}

// computeBacktrack, TDS (:972-1118): every walk of the constraint, moves
// 1..length-2 (the reference's loop stops before the last move).
Effectiveness effectiveness(const LabelStat& st, const TemplateConstraint& c, int threads) {
  Effectiveness r;
  std::vector<const Walk*> walks;
  if (c.order.empty())
    for (const auto& kv : c.walks) walks.push_back(&kv.second);
  else
    for (uint32_t root : c.order) walks.push_back(&c.walks.at(root));
  auto history_labels = [](const Walk& w) {
    std::vector<uint64_t> h;
    for (size_t k = 0; k < w.move.size(); ++k)
      if (w.move[k] == PJ_MOVE_VERTEX_STORE || w.move[k] == PJ_MOVE_VERTEX_NO_STORE) h.push_back(w.label[k]);
    return h;
  };
  std::vector<double> finding;
  for (const Walk* wp : walks) {
    const Walk& w = *wp;
    const std::vector<uint64_t> history = history_labels(w);
    const size_t steps = w.move.size() >= 2 ? w.move.size() - 2 : 0;
    double fp = 1.;
    double unique = 1;
    uint64_t from = w.label[0];
    std::vector<double> unique_at{unique};
    for (size_t j = 0; j < steps; ++j) {
      const uint8_t m = w.move[j + 1];
      if (m == PJ_MOVE_VERTEX_STORE || m == PJ_MOVE_VERTEX_NO_STORE) {
        const uint64_t to = w.label[j + 1];
        fp *= proba_find_one_path(st, unique, from, to);
        unique = average_unique_vertex(st, unique, from, to);
        from = to;
        unique_at.push_back(unique);
      } else if (m == PJ_MOVE_BACK) {
        const uint32_t back = w.move_back[j + 1];
        unique = unique_at[back];
        from = history[back];
      } else if (m == PJ_MOVE_CHECK) {
        const uint64_t to = w.label[j + 1];
        fp *= proba_find_one_path(st, unique, from, to) * proba_find_same_vertex(st, unique, to);
      }
    }
    finding.push_back(fp);
  }
  for (size_t i = 0; i < walks.size(); ++i) r.match_eliminated += st.vertices(walks[i]->label[0]) * (1 - finding[i]);
  for (size_t i = 0; i < walks.size(); ++i) {
    const Walk& w = *walks[i];
    const std::vector<uint64_t> history = history_labels(w);
    const size_t steps = w.move.size() >= 2 ? w.move.size() - 2 : 0;
    double traversed = 1;
    double cost = traversed;
    uint64_t from = w.label[0];
    std::vector<double> traversed_at{traversed};
    for (size_t j = 0; j < steps; ++j) {
      const uint8_t m = w.move[j + 1];
      if (m == PJ_MOVE_VERTEX_STORE || m == PJ_MOVE_VERTEX_NO_STORE) {
        const uint64_t to = w.label[j + 1];
        cost += average_vertex(st, traversed, from);
        traversed = j == 0 ? average_vertex_knowing(st, traversed, from, to)
                           : average_unique_vertex(st, traversed, from, to);
        traversed_at.push_back(traversed);
        from = to;
      } else if (m == PJ_MOVE_BACK) {
        const uint32_t back = w.move_back[j + 1];
        traversed = traversed_at[back];
        from = history[back];
      } else if (m == PJ_MOVE_CHECK) {
        cost += average_vertex(st, traversed, from);
      }
    }
    cost += traversed;
    double increase = st.vertices(w.label[0]) * cost * (1 - finding[i]);
    increase /= threads;
    increase /= 1 + static_cast<double>(w.vertex_length - 1) / 2 * (finding[i]);
    r.cost_eliminated += increase;
  }
  r.effectiveness = r.match_eliminated / r.cost_eliminated;
  return r;
}

NextConstraint choose_next(const LabelStat& st, const ConstraintSet& cs, const std::vector<uint32_t>& cycles_left,
                           const std::vector<uint32_t>& paths_left, const std::vector<uint32_t>& tds_left,
                           int threads) {
  double best_c = 0, best_p = 0, best_t = 0;
  uint32_t id_c = 0, id_p = 0, id_t = 0;
  for (size_t k = 0; k < cycles_left.size(); ++k) {
    const double e = effectiveness(st, cs.cycles[cycles_left[k]], threads).effectiveness;
    if (e > best_c) {
      id_c = static_cast<uint32_t>(k);
      best_c = e;
    }
  }
  for (size_t k = 0; k < paths_left.size(); ++k) {
    const double e = effectiveness(st, cs.paths[paths_left[k]], threads).effectiveness;
    if (e > best_p) {
      id_p = static_cast<uint32_t>(k);
      best_p = e;
    }
  }
  for (size_t k = 0; k < tds_left.size(); ++k) {
    const double e = effectiveness(st, cs.tds[tds_left[k]], threads).effectiveness;
    if (e > best_t) {
      id_t = static_cast<uint32_t>(k);
      best_t = e;
    }
  }
  NextConstraint n;
  if (best_c >= best_p && best_c >= best_t && !cycles_left.empty()) {
    n = NextConstraint{PJ_STEP_CC, cycles_left[id_c], best_c};
  } else if (best_p >= best_c && best_p >= best_t && !paths_left.empty()) {
    n = NextConstraint{PJ_STEP_PC, paths_left[id_p], best_p};
  } else if (best_t >= best_c && best_t >= best_p && !tds_left.empty()) {
    n = NextConstraint{PJ_STEP_TDS, tds_left[id_t], best_t};
  }
  return n;
}

}  // namespace pj
