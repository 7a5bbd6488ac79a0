// Host constraint generation (see pj_constraints.h for the reference map).
#include "pj_constraints.h"

#include <algorithm>
#include <functional>
#include <sstream>

namespace pj {

namespace {

template <class T>
bool contains(const std::vector<T>& v, const T& x) {
  return std::find(v.begin(), v.end(), x) != v.end();
}

// The reference removes two merged entries of a std::deque with
// `list.erase(it1); list.erase(it2);` (generate_constraint_template.cc:30-31
// and :71-72), where the second iterator is used after the first erase.  With
// libstdc++, erasing position i of n elements shifts the front part when
// i < n/2 (the element at j keeps its slot, so the second erase removes it)
// and the back part otherwise (slot j then holds the element that was at j+1,
// or is the end slot, in which case the last element is popped).  This
// reproduces that outcome on a vector so that the template list, and hence
// the TDS constraint order, matches.
template <class T>
void erase_two_like_reference(std::vector<T>* v, size_t i, size_t j) {
  const size_t n = v->size();
  const bool front_shift = i < n / 2;
  v->erase(v->begin() + static_cast<std::ptrdiff_t>(i));
  size_t k = front_shift ? j - 1 : j;
  if (k >= v->size()) k = v->size() - 1;
  v->erase(v->begin() + static_cast<std::ptrdiff_t>(k));
}

// Pairwise merge loop shared by the cycle and path passes
// (generate_constraint_template.cc:24-42 and :64-82).
void merge_until_stable(std::vector<TemplateConstraint>* list,
                        bool (TemplateConstraint::*related)(const TemplateConstraint&) const) {
  size_t i = 0;
  while (i < list->size()) {
    bool merged = false;
    for (size_t j = i + 1; j < list->size(); ++j) {
      if (((*list)[i].*related)((*list)[j])) {
        TemplateConstraint m = TemplateConstraint::merge((*list)[i], (*list)[j]);
        erase_two_like_reference(list, i, j);
        list->push_back(std::move(m));
        i = 0;
        merged = true;
        break;
      }
    }
    if (!merged) ++i;
  }
}

}  // namespace

void Walk::add_vertex(const Graph& pattern, uint32_t v) {
  move.push_back(PJ_MOVE_VERTEX_NO_STORE);
  vertex.push_back(v);
  label.push_back(pattern.values[v]);
  move_back.push_back(0);
  ++vertex_length;
}

void Walk::add_check(const Graph& pattern, uint32_t v, uint32_t history_index) {
  move.push_back(PJ_MOVE_CHECK);
  vertex.push_back(v);
  label.push_back(pattern.values[v]);
  move_back.push_back(history_index);
}

void Walk::add_move_back(const Graph& pattern, uint32_t v, uint32_t history_index) {
  move.push_back(PJ_MOVE_BACK);
  vertex.push_back(v);
  label.push_back(pattern.values[v]);
  move_back.push_back(history_index);
}

// computeStoredVertex (walk.cc:43-60): the VERTEX move of a template vertex
// that a later CHECK revisits becomes a STORE move.
void Walk::mark_stored() {
  std::vector<uint32_t> stored;
  for (size_t i = 0; i < move.size(); ++i) {
    if (move[i] != PJ_MOVE_VERTEX_NO_STORE) continue;
    bool checked_later = false;
    for (size_t j = i + 1; j < move.size(); ++j)
      if (vertex[j] == vertex[i] && move[j] == PJ_MOVE_CHECK) {
        checked_later = true;
        break;
      }
    if (checked_later && !contains(stored, vertex[i])) {
      move[i] = PJ_MOVE_VERTEX_STORE;
      stored.push_back(vertex[i]);
    }
  }
}

std::string Walk::to_string() const {
  static const char* names[] = {"MB", "CK", "VS", "VN"};
  std::ostringstream os;
  os << "Walk : ";
  for (size_t i = 0; i < move.size(); ++i) {
    if (i) os << " -> ";
    os << "[" << names[move[i]] << "," << vertex[i] << "," << label[i] << "," << move_back[i] << "]";
  }
  return os.str();
}

TemplateConstraint TemplateConstraint::from_cycle(const CycleConstraint& c) {
  TemplateConstraint t;
  t.origin = CC;
  t.length = c.vertex.size();
  for (size_t i = 0; i + 1 < c.vertex.size(); ++i) {
    t.adjacency[c.vertex[i]].insert(c.vertex[i + 1]);
    t.adjacency[c.vertex[i + 1]].insert(c.vertex[i]);
  }
  t.adjacency[c.vertex.back()].insert(c.vertex.front());
  t.adjacency[c.vertex.front()].insert(c.vertex.back());
  return t;
}

TemplateConstraint TemplateConstraint::from_path(const PathConstraint& p) {
  TemplateConstraint t;
  t.origin = PC;
  t.length = p.vertex.size();
  for (size_t i = 0; i + 1 < p.vertex.size(); ++i) {
    t.adjacency[p.vertex[i]].insert(p.vertex[i + 1]);
    t.adjacency[p.vertex[i + 1]].insert(p.vertex[i]);
  }
  return t;
}

TemplateConstraint TemplateConstraint::from_pattern(const Graph& pattern) {
  TemplateConstraint t;
  t.origin = TDS_FULL;
  t.length = pattern.vertex_count;
  for (uint32_t v = 0; v < pattern.vertex_count; ++v)
    for (uint64_t e = pattern.vertices[v]; e < pattern.vertices[v + 1]; ++e) {
      uint32_t n = pattern.edges[e];
      t.adjacency[v].insert(n);
      t.adjacency[n].insert(v);
    }
  return t;
}

TemplateConstraint TemplateConstraint::merge(const TemplateConstraint& a, const TemplateConstraint& b) {
  auto cyc = [](Origin o) { return o == CC || o == TDS_CC; };
  auto pth = [](Origin o) { return o == PC || o == TDS_PC; };
  TemplateConstraint t;
  if (cyc(a.origin) && cyc(b.origin))
    t.origin = TDS_CC;
  else if (pth(a.origin) && pth(b.origin))
    t.origin = TDS_PC;
  else
    t.origin = TDS_SUB;
  t.length = a.length;
  t.adjacency = a.adjacency;
  for (const auto& kv : b.adjacency) {
    if (t.adjacency.count(kv.first) == 0) ++t.length;
    for (uint32_t n : kv.second) t.adjacency[kv.first].insert(n);
  }
  return t;
}

bool TemplateConstraint::shares_vertex(const TemplateConstraint& o) const {
  for (const auto& kv : adjacency)
    if (o.adjacency.count(kv.first)) return true;
  return false;
}

bool TemplateConstraint::shares_edge(const TemplateConstraint& o) const {
  for (const auto& kv : adjacency) {
    auto it = o.adjacency.find(kv.first);
    if (it == o.adjacency.end()) continue;
    for (uint32_t n : kv.second)
      if (it->second.count(n)) return true;
  }
  return false;
}

// generateWalkMap / generateWalk (multiple_label_constraint_template.cc:121-169):
// DFS over the constraint's template graph from every vertex, neighbours in
// ascending order; tree edges become VERTEX moves, non-tree edges CHECK moves,
// and a MOVE_BACK returns to the branching vertex before the next branch.
void TemplateConstraint::build_walks(const Graph& pattern) {
  for (const auto& kv : adjacency) {
    Walk w;
    std::vector<uint32_t> hist{kv.first};
    std::vector<std::pair<uint32_t, uint32_t>> hist_edges;
    w.add_vertex(pattern, kv.first);
    std::function<void(uint32_t)> visit = [&](uint32_t cur) {
      bool at_cur = true;
      const uint32_t cur_index = w.vertex_length - 1;
      for (uint32_t n : adjacency.at(cur)) {
        auto e = std::make_pair(std::min(cur, n), std::max(cur, n));
        if (!contains(hist, n)) {
          if (!at_cur) w.add_move_back(pattern, cur, cur_index);
          hist_edges.push_back(e);
          hist.push_back(n);
          w.add_vertex(pattern, n);
          visit(n);
          at_cur = false;
        } else {
          if (contains(hist_edges, e)) continue;
          if (!at_cur) {
            w.add_move_back(pattern, cur, cur_index);
            at_cur = true;
          }
          uint32_t prev = static_cast<uint32_t>(std::find(hist.begin(), hist.end(), n) - hist.begin());
          hist_edges.push_back(e);
          w.add_check(pattern, n, prev);
        }
      }
    };
    visit(kv.first);
    w.mark_stored();
    walks[kv.first] = w;
  }
}

int ConstraintSet::build(const Graph& pattern, std::string* err) {
  if (!pattern.valued) {
    *err = "pattern has no vertex labels";
    return -1;
  }
  if (pattern.vertex_count == 0 || pattern.vertex_count > PJ_MAX_TEMPLATE_VERTICES) {
    *err = "pattern must have 1..64 vertices";
    return -1;
  }
  const uint32_t P = pattern.vertex_count;
  vertex_count = P;
  vertex_label = pattern.values;

  // Local constraints from the pattern's out-rows.
  local.assign(P, LocalConstraint());
  for (uint32_t v = 0; v < P; ++v) {
    LocalConstraint& lc = local[v];
    lc.origin = v;
    lc.origin_label = pattern.values[v];
    for (uint64_t e = pattern.vertices[v]; e < pattern.vertices[v + 1]; ++e) {
      uint32_t n = pattern.edges[e];
      lc.neighbor_mask |= 1ull << n;
      uint64_t nl = pattern.values[n];
      auto it = std::find_if(lc.need.begin(), lc.need.end(), [&](const auto& p) { return p.first == nl; });
      if (it == lc.need.end())
        lc.need.emplace_back(nl, 1u);
      else
        ++it->second;
    }
  }

  // Cycles of length 3..P from every start vertex, pattern rows in file order.
  {
    std::vector<uint32_t> hist;
    uint64_t visited = 0;
    std::function<void(uint32_t, uint32_t, size_t)> walk = [&](uint32_t src, uint32_t cur, size_t rem) {
      if (rem == 0) {
        if (cur != src) return;
        // Normalise: start at the smallest id, go towards its smaller neighbour
        // (multiple_label_constraint_circular.cc:13-60).
        std::vector<uint32_t> c(hist.begin(), hist.end() - 1);
        const size_t L = c.size();
        size_t m = 0;
        for (size_t i = 1; i < L; ++i)
          if (c[i] < c[m]) m = i;
        bool rev;
        if (m > 0 && m < L - 1)
          rev = c[m - 1] < c[m + 1];
        else if (m == 0)
          rev = c[L - 1] < c[1];
        else
          rev = c[L - 2] < c[0];
        CycleConstraint cc;
        size_t pos = m;
        for (size_t k = 0; k < L; ++k) {
          cc.vertex.push_back(c[pos]);
          cc.label.push_back(pattern.values[c[pos]]);
          pos = rev ? (pos == 0 ? L - 1 : pos - 1) : (pos + 1 == L ? 0 : pos + 1);
        }
        if (!contains(cycles, cc)) cycles.push_back(cc);
        return;
      }
      for (uint64_t e = pattern.vertices[cur]; e < pattern.vertices[cur + 1]; ++e) {
        uint32_t n = pattern.edges[e];
        if (visited >> n & 1) continue;
        hist.push_back(n);
        visited |= 1ull << n;
        walk(src, n, rem - 1);
        visited &= ~(1ull << n);
        hist.pop_back();
      }
    };
    for (uint32_t s = 0; s < P; ++s) {
      hist.assign(1, s);
      for (size_t len = 3; len <= P; ++len) {
        visited = 0;
        walk(s, s, len);
      }
    }
  }

  // Paths between equal-label template vertices at least two edges apart.
  {
    std::vector<uint32_t> hist;
    uint64_t visited = 0;
    std::function<void(uint32_t, uint32_t, size_t)> walk = [&](uint32_t src, uint32_t cur, size_t len) {
      if (len >= 2 && pattern.values[src] == pattern.values[cur] && src != cur) {
        PathConstraint pc;
        pc.initial_label = pattern.values[hist.front()];
        const bool rev = hist.back() > hist.front();
        for (size_t k = 0; k < hist.size(); ++k) {
          uint32_t v = rev ? hist[hist.size() - 1 - k] : hist[k];
          pc.vertex.push_back(v);
          pc.label.push_back(pattern.values[v]);
        }
        if (!contains(paths, pc)) paths.push_back(pc);
        return;
      }
      for (uint64_t e = pattern.vertices[cur]; e < pattern.vertices[cur + 1]; ++e) {
        uint32_t n = pattern.edges[e];
        if (visited >> n & 1) continue;
        hist.push_back(n);
        visited |= 1ull << n;
        walk(src, n, len + 1);
        visited &= ~(1ull << n);
        hist.pop_back();
      }
    };
    for (uint32_t s = 0; s < P; ++s) {
      hist.assign(1, s);
      visited = 1ull << s;
      walk(s, s, 0);
    }
  }

  // Template-driven-search constraints.
  tds.clear();
  {
    std::vector<TemplateConstraint> list;
    for (const auto& c : cycles) list.push_back(TemplateConstraint::from_cycle(c));
    merge_until_stable(&list, &TemplateConstraint::shares_edge);
    for (auto& t : list)
      if (t.origin != TemplateConstraint::CC) tds.push_back(t);
  }
  {
    std::unordered_map<uint64_t, std::vector<TemplateConstraint>> by_label;
    for (const auto& p : paths) by_label[p.initial_label].push_back(TemplateConstraint::from_path(p));
    for (auto& kv : by_label) merge_until_stable(&kv.second, &TemplateConstraint::shares_vertex);
    for (const auto& kv : by_label)
      for (const auto& t : kv.second)
        if (t.origin != TemplateConstraint::PC) tds.push_back(t);
  }
  {
    // Unions of templates that share a vertex (generate_constraint_template.cc:94-130).
    std::vector<TemplateConstraint> list(tds.begin(), tds.end());
    size_t offset = 0;
    while (!list.empty() && offset < list.size() - 1) {
      int merged = 0;
      TemplateConstraint cur = list[offset];
      size_t it = offset + 1;
      while (it < list.size()) {
        if (cur.shares_vertex(list[it])) {
          cur = TemplateConstraint::merge(cur, list[it]);
          list.erase(list.begin() + static_cast<std::ptrdiff_t>(it));
          it = 0;
          ++merged;
        } else {
          ++it;
        }
      }
      if (merged > 0) {
        offset = 0;
        bool present = false;
        for (const auto& t : tds)
          if (t.same_as(cur)) present = true;
        if (!present) tds.push_back(cur);
      } else {
        ++offset;
      }
    }
  }
  for (auto& t : tds) t.build_walks(pattern);

  enumeration = TemplateConstraint::from_pattern(pattern);
  enumeration.build_walks(pattern);
  if (enumeration.walks.empty()) {
    *err = "pattern has no edges";
    return -1;
  }
  flatten();
  return 0;
}

void ConstraintSet::flatten() {
  const uint32_t P = vertex_count;
  f_nbr.assign(P, 0);
  f_need_off.assign(P + 1, 0);
  f_need_label.clear();
  f_need_count.clear();
  for (uint32_t p = 0; p < P; ++p) {
    f_nbr[p] = local[p].neighbor_mask;
    for (const auto& nd : local[p].need) {
      f_need_label.push_back(nd.first);
      f_need_count.push_back(nd.second);
    }
    f_need_off[p + 1] = static_cast<uint32_t>(f_need_label.size());
  }
  f_cycle_off.assign(1, 0);
  f_cycle_vtx.clear();
  for (const auto& c : cycles) {
    f_cycle_vtx.insert(f_cycle_vtx.end(), c.vertex.begin(), c.vertex.end());
    f_cycle_off.push_back(static_cast<uint32_t>(f_cycle_vtx.size()));
  }
  f_path_off.assign(1, 0);
  f_path_vtx.clear();
  f_path_label.clear();
  for (const auto& p : paths) {
    f_path_vtx.insert(f_path_vtx.end(), p.vertex.begin(), p.vertex.end());
    f_path_off.push_back(static_cast<uint32_t>(f_path_vtx.size()));
    f_path_label.push_back(p.initial_label);
  }
  f_walks.clear();
  f_tds_off.assign(1, 0);
  auto walk_desc = [](uint32_t root, const Walk& wk) {
    pj_walk_desc w;
    w.root = root;
    w.length = static_cast<uint32_t>(wk.move.size());
    w.move = wk.move.data();
    w.vertex = wk.vertex.data();
    w.move_back = wk.move_back.data();
    return w;
  };
  for (const auto& t : tds) {
    if (t.order.empty())
      for (const auto& kv : t.walks) f_walks.push_back(walk_desc(kv.first, kv.second));
    else
      for (uint32_t r : t.order) f_walks.push_back(walk_desc(r, t.walks.at(r)));
    f_tds_off.push_back(static_cast<uint32_t>(f_walks.size()));
  }
  pj_template_desc& d = f_desc;
  d = pj_template_desc{};
  d.vertex_count = P;
  d.vertex_label = vertex_label.data();
  d.neighbor_mask = f_nbr.data();
  d.need_offset = f_need_off.data();
  d.need_label = f_need_label.data();
  d.need_count = f_need_count.data();
  d.cycle_count = static_cast<uint32_t>(cycles.size());
  d.cycle_offset = f_cycle_off.data();
  d.cycle_vertex = f_cycle_vtx.data();
  d.path_count = static_cast<uint32_t>(paths.size());
  d.path_offset = f_path_off.data();
  d.path_vertex = f_path_vtx.data();
  d.path_initial_label = f_path_label.data();
  d.tds_count = static_cast<uint32_t>(tds.size());
  d.tds_walk_offset = f_tds_off.data();
  d.walk_count = static_cast<uint32_t>(f_walks.size());
  d.walks = f_walks.data();
  if (!enumeration.walks.empty()) {
    const uint32_t root = enumeration.order.empty() ? enumeration.walks.cbegin()->first : enumeration.order[0];
    d.enumeration_walk = walk_desc(root, enumeration.walks.at(root));
  }
}

std::string ConstraintSet::dump() const {
  std::ostringstream os;
  os << "Local constraints\n";
  for (const auto& lc : local) {
    os << "Vertex " << lc.origin << " : Label " << lc.origin_label << "\n";
    for (const auto& nd : lc.need) os << "\t Sublabel " << nd.first << " : " << nd.second << "\n";
    for (uint32_t n = 0; n < 64; ++n)
      if (lc.neighbor_mask >> n & 1) os << "\t Neighbor " << n << "\n";
  }
  os << "Circular constraint number = " << cycles.size() << "\n";
  for (size_t i = 0; i < cycles.size(); ++i) {
    os << "=== Current constraint = " << i << " ===\nVertex Index : ";
    for (uint32_t v : cycles[i].vertex) os << v << " -> ";
    os << cycles[i].vertex[0] << "\n";
  }
  os << "Path constraint number = " << paths.size() << "\n";
  for (size_t i = 0; i < paths.size(); ++i) {
    os << "=== Current constraint = " << i << " ===\nVertex Index : ";
    for (size_t k = 0; k < paths[i].vertex.size(); ++k) os << (k ? " -> " : "") << paths[i].vertex[k];
    os << "\n";
  }
  os << "Template constraint number = " << tds.size() << "\n";
  for (size_t i = 0; i < tds.size(); ++i) {
    os << "=== Current constraint = " << i << " === [length : " << tds[i].length << "]\n";
    for (const auto& kv : tds[i].walks) os << kv.second.to_string() << "\n";
  }
  os << "=== enumeration constraint ===\n";
  for (const auto& kv : enumeration.walks) {
    os << kv.second.to_string() << "\n";
    break;
  }
  return os.str();
}

}  // namespace pj
