// extern "C" surface of the host library (include/prunejuice_host.h).
#include <getopt.h>

#include <cstring>
#include <iostream>
#include <string>

#include "../../../include/prunejuice_host.h"
#include "pj_constraints.h"
#include "pj_driver.h"
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
  int n = 0;
  int rc = pj::run_step_machine(
      &st, c->c,
      [&](const pj::StepRecord& r) {
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
                            r.deviceMs};
        ++n;
      },
      total_ms, (flags & PJH_STOP_BEFORE_ENUMERATION) != 0, (flags & PJH_STOP_BEFORE_TDS) != 0);
  return rc < 0 ? -1 : n;
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
