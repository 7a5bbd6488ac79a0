// TOTEM formats and synthetic graph generators (host side).
#include "pj_graph.h"

#include <omp.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>

namespace pj {

namespace {

const uint32_t kBinaryMagic = 0x10102048u;  // totem_graph.cu:25

struct FileCloser {
  void operator()(FILE* f) const {
    if (f) fclose(f);
  }
};
using FilePtr = std::unique_ptr<FILE, FileCloser>;

// Splits on the reference's delimiter set " \t\n:" (totem_graph.cu:19).
std::vector<std::string> tokens_of(const std::string& line) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : line) {
    if (ch == ' ' || ch == '\t' || ch == '\n' || ch == ':' || ch == '\r') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(ch);
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

bool is_number(const std::string& s) {
  if (s.empty()) return false;
  for (char ch : s)
    if (!std::isdigit(static_cast<unsigned char>(ch))) return false;
  return true;
}

bool read_line(FILE* f, std::string* line) {
  line->clear();
  int ch;
  while ((ch = fgetc(f)) != EOF) {
    line->push_back(static_cast<char>(ch));
    if (ch == '\n') return true;
  }
  return !line->empty();
}

}  // namespace

int load_totem_text(const std::string& path, Graph* g, std::string* err) {
  FilePtr f(fopen(path.c_str(), "r"));
  if (!f) {
    *err = "cannot open " + path;
    return -1;
  }
  Graph out;
  out.directed = true;
  bool have_nodes = false, have_edges = false;
  std::string line;
  uint64_t line_no = 0;
  // Three metadata lines, each starting with '#'.
  for (int m = 0; m < 3; ++m) {
    if (!read_line(f.get(), &line) || line.empty() || line[0] != '#') {
      *err = "metadata format error in " + path;
      return -1;
    }
    ++line_no;
    auto tok = tokens_of(line.substr(1));
    if (tok.empty()) {
      *err = "metadata format error in " + path;
      return -1;
    }
    std::string key = tok[0];
    for (auto& ch : key) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
    if (key == "NODES") {
      if (tok.size() < 2 || !is_number(tok[1])) {
        *err = "bad node count in " + path;
        return -1;
      }
      out.vertex_count = static_cast<uint32_t>(std::strtoull(tok[1].c_str(), nullptr, 10));
      if (tok.size() >= 3 && std::tolower(static_cast<unsigned char>(tok[2][0])) == 'y')
        out.valued = true;
      have_nodes = true;
    } else if (key == "EDGES") {
      if (tok.size() < 2 || !is_number(tok[1])) {
        *err = "bad edge count in " + path;
        return -1;
      }
      out.edge_count = std::strtoull(tok[1].c_str(), nullptr, 10);
      have_edges = true;
    } else if (key == "DIRECTED") {
      out.directed = true;
    } else if (key == "UNDIRECTED") {
      out.directed = false;
    } else {
      *err = "unknown metadata keyword in " + path;
      return -1;
    }
  }
  if (!have_nodes || !have_edges) {
    *err = "missing # Nodes or # Edges in " + path;
    return -1;
  }
  out.vertices.assign(static_cast<size_t>(out.vertex_count) + 1, 0);
  out.edges.assign(out.edge_count, 0);
  if (out.valued) {
    out.values.assign(out.vertex_count, 0);
    uint32_t vi = 0;
    while (vi < out.vertex_count) {
      if (!read_line(f.get(), &line)) break;
      ++line_no;
      if (line[0] == '#') continue;
      auto tok = tokens_of(line);
      if (tok.size() < 2 || !is_number(tok[0])) {
        *err = "vertex list format error at line " + std::to_string(line_no);
        return -1;
      }
      uint64_t id = std::strtoull(tok[0].c_str(), nullptr, 10);
      uint64_t value = static_cast<uint64_t>(std::atof(tok[1].c_str()));
      if (id != vi) {
        if (!(id > vi && id < out.vertex_count)) {
          *err = "vertex list not increasing at line " + std::to_string(line_no);
          return -1;
        }
        while (vi < id) out.values[vi++] = 0;
      }
      out.values[vi++] = value;
    }
  }
  // Edge list: rows sorted by source.
  uint64_t vi = 0, ei = 0;
  while (read_line(f.get(), &line)) {
    ++line_no;
    if (line[0] == '#') continue;
    auto tok = tokens_of(line);
    if (tok.empty()) continue;
    if (tok.size() < 2 || !is_number(tok[0]) || !is_number(tok[1])) {
      *err = "edge list format error at line " + std::to_string(line_no);
      return -1;
    }
    uint64_t s = std::strtoull(tok[0].c_str(), nullptr, 10);
    uint64_t d = std::strtoull(tok[1].c_str(), nullptr, 10);
    if (s >= 0xFFFFFFFFull || d >= 0xFFFFFFFFull) {
      *err = "vertex id out of 32-bit range at line " + std::to_string(line_no);
      return -1;
    }
    if (vi == 0 || s != vi - 1) {
      if (!(s >= vi && s < out.vertex_count)) {
        *err = "edge list not sorted by source at line " + std::to_string(line_no);
        return -1;
      }
      while (vi <= s) out.vertices[vi++] = ei;
    }
    if (ei >= out.edge_count) {
      *err = "more edges than declared in " + path;
      return -1;
    }
    out.edges[ei++] = static_cast<uint32_t>(d);
  }
  if (ei != out.edge_count) {
    *err = "fewer edges than declared in " + path;
    return -1;
  }
  while (vi <= out.vertex_count) out.vertices[vi++] = ei;
  *g = std::move(out);
  return 0;
}

static int load_totem_binary(FILE* f, const std::string& path, Graph* g, std::string* err) {
  uint32_t vid_size = 0, eid_size = 0;
  if (fread(&vid_size, 4, 1, f) != 1 || fread(&eid_size, 4, 1, f) != 1 || vid_size != 4 ||
      (eid_size != 4 && eid_size != 8)) {
    *err = "invalid binary header in " + path;
    return -1;
  }
  Graph out;
  uint32_t v = 0;
  uint64_t e = 0;
  if (fread(&v, 4, 1, f) != 1) goto bad;
  if (eid_size == 4) {
    uint32_t e32;
    if (fread(&e32, 4, 1, f) != 1) goto bad;
    e = e32;
  } else if (fread(&e, 8, 1, f) != 1) {
    goto bad;
  }
  {
    uint8_t flags[3];
    if (fread(flags, 1, 3, f) != 3) goto bad;
    out.valued = flags[0] != 0;
    out.weighted = flags[1] != 0;
    out.directed = flags[2] != 0;
  }
  out.vertex_count = v;
  out.edge_count = e;
  out.vertices.resize(static_cast<size_t>(v) + 1);
  if (eid_size == 8) {
    if (fread(out.vertices.data(), 8, v + 1ull, f) != v + 1ull) goto bad;
  } else {
    std::vector<uint32_t> tmp(static_cast<size_t>(v) + 1);
    if (fread(tmp.data(), 4, v + 1ull, f) != v + 1ull) goto bad;
    for (size_t i = 0; i <= v; ++i) out.vertices[i] = tmp[i];
  }
  if (out.valued) {
    out.values.resize(v);
    if (fread(out.values.data(), 8, v, f) != v) goto bad;
  }
  out.edges.resize(e);
  if (e && fread(out.edges.data(), 4, e, f) != e) goto bad;
  out.weighted = false;  // weights are not loaded (graph_initialize(..., 0, ...))
  *g = std::move(out);
  return 0;
bad:
  *err = "truncated binary graph " + path;
  return -1;
}

int load_totem_binary_values(const std::string& path, Graph* g, std::string* err) {
  FilePtr f(fopen(path.c_str(), "rb"));
  if (!f) {
    *err = "cannot open " + path;
    return -1;
  }
  uint32_t head[4] = {0, 0, 0, 0};  // magic, vid bytes, eid bytes, vertex count
  if (fread(head, 4, 1, f.get()) != 1 || head[0] != kBinaryMagic) return 1;
  Graph out;
  uint64_t e = 0;
  uint8_t flags[3];
  bool ok = fread(head + 1, 4, 3, f.get()) == 3 && head[1] == 4 && (head[2] == 4 || head[2] == 8);
  if (ok && head[2] == 4) {
    uint32_t e32 = 0;
    ok = fread(&e32, 4, 1, f.get()) == 1;
    e = e32;
  } else if (ok) {
    ok = fread(&e, 8, 1, f.get()) == 1;
  }
  ok = ok && fread(flags, 1, 3, f.get()) == 3;
  out.vertex_count = head[3];
  out.edge_count = e;
  out.valued = ok && flags[0] != 0;
  out.directed = ok && flags[2] != 0;
  if (ok && out.valued) {
    ok = fseeko(f.get(), static_cast<off_t>(head[2]) * (out.vertex_count + 1ull), SEEK_CUR) == 0;
    out.values.resize(out.vertex_count);
    ok = ok && fread(out.values.data(), 8, out.vertex_count, f.get()) == out.vertex_count;
  }
  if (!ok) {
    *err = "truncated binary graph " + path;
    return -1;
  }
  *g = std::move(out);
  return 0;
}

int validate_csr(const Graph& g, std::string* err) {
  const uint64_t V = g.vertex_count, E = g.edge_count;
  if (g.vertices.size() != V + 1 || g.edges.size() != E || g.vertices[0] != 0 || g.vertices[V] != E) {
    *err = "malformed CSR: offsets must run from 0 to the edge count";
    return -1;
  }
  int64_t bad_row = -1;
#pragma omp parallel for reduction(max : bad_row)
  for (int64_t v = 0; v < static_cast<int64_t>(V); ++v)
    if (g.vertices[v + 1] < g.vertices[v]) bad_row = std::max(bad_row, v);
  if (bad_row >= 0) {
    *err = "malformed CSR: offsets decrease at vertex " + std::to_string(bad_row);
    return -1;
  }
  int64_t bad_slot = -1;
#pragma omp parallel for reduction(max : bad_slot)
  for (int64_t e = 0; e < static_cast<int64_t>(E); ++e)
    if (g.edges[e] >= V) bad_slot = std::max(bad_slot, e);
  if (bad_slot >= 0) {
    *err = "malformed CSR: neighbour id " + std::to_string(g.edges[bad_slot]) + " at slot " +
           std::to_string(bad_slot) + " is not a vertex";
    return -1;
  }
  return 0;
}

int load_totem(const std::string& path, Graph* g, std::string* err) {
  FilePtr f(fopen(path.c_str(), "rb"));
  if (!f) {
    *err = "cannot open " + path;
    return -1;
  }
  uint32_t magic = 0;
  if (fread(&magic, 4, 1, f.get()) != 1) {
    *err = "empty file " + path;
    return -1;
  }
  int rc;
  if (magic == kBinaryMagic) {
    rc = load_totem_binary(f.get(), path, g, err);
  } else {
    f.reset();
    rc = load_totem_text(path, g, err);
  }
  if (rc == 0 && validate_csr(*g, err)) {
    *err += " (" + path + ")";
    return -1;
  }
  return rc;
}

int store_totem_binary(const Graph& g, const std::string& path, int eid_bytes, std::string* err) {
  if (eid_bytes != 4 && eid_bytes != 8) {
    *err = "eid_bytes must be 4 or 8";
    return -1;
  }
  if (eid_bytes == 4 && g.edge_count > 0xFFFFFFFFull) {
    *err = "graph too large for 32-bit edge ids";
    return -1;
  }
  FilePtr f(fopen(path.c_str(), "wb"));
  if (!f) {
    *err = "cannot write " + path;
    return -1;
  }
  uint32_t w[3] = {kBinaryMagic, 4u, static_cast<uint32_t>(eid_bytes)};
  bool ok = fwrite(w, 4, 3, f.get()) == 3;
  ok = ok && fwrite(&g.vertex_count, 4, 1, f.get()) == 1;
  if (eid_bytes == 4) {
    uint32_t e32 = static_cast<uint32_t>(g.edge_count);
    ok = ok && fwrite(&e32, 4, 1, f.get()) == 1;
  } else {
    ok = ok && fwrite(&g.edge_count, 8, 1, f.get()) == 1;
  }
  uint8_t flags[3] = {static_cast<uint8_t>(g.valued), 0, static_cast<uint8_t>(g.directed)};
  ok = ok && fwrite(flags, 1, 3, f.get()) == 3;
  if (eid_bytes == 8) {
    ok = ok && fwrite(g.vertices.data(), 8, g.vertex_count + 1ull, f.get()) == g.vertex_count + 1ull;
  } else {
    std::vector<uint32_t> tmp(g.vertices.begin(), g.vertices.end());
    ok = ok && fwrite(tmp.data(), 4, tmp.size(), f.get()) == tmp.size();
  }
  if (g.valued) ok = ok && fwrite(g.values.data(), 8, g.vertex_count, f.get()) == g.vertex_count;
  if (g.edge_count) ok = ok && fwrite(g.edges.data(), 4, g.edge_count, f.get()) == g.edge_count;
  if (!ok) {
    *err = "write failed for " + path;
    return -1;
  }
  return 0;
}

int store_totem_text(const Graph& g, const std::string& path, std::string* err) {
  std::ofstream os(path);
  if (!os) {
    *err = "cannot write " + path;
    return -1;
  }
  os << "#Nodes:" << g.vertex_count << (g.valued ? " Y" : "") << "\n";
  os << "#Edges:" << g.edge_count << "\n";
  os << (g.directed ? "#Directed" : "#Undirected") << "\n";
  if (g.valued)
    for (uint32_t v = 0; v < g.vertex_count; ++v) os << v << " " << g.values[v] << "\n";
  for (uint32_t v = 0; v < g.vertex_count; ++v)
    for (uint64_t e = g.vertices[v]; e < g.vertices[v + 1]; ++e) os << v << " " << g.edges[e] << "\n";
  return os.good() ? 0 : -1;
}

int parse_vertex_file(const std::string& path, Graph* g, std::string* err) {
  FilePtr f(fopen(path.c_str(), "r"));
  if (!f) {
    *err = "cannot open " + path;
    return -1;
  }
  g->valued = true;
  g->values.assign(g->vertex_count, 0);
  std::string line;
  uint32_t vi = 0;
  uint64_t line_no = 0;
  while (vi < g->vertex_count) {
    if (!read_line(f.get(), &line)) break;
    ++line_no;
    if (line[0] == '#') continue;
    auto tok = tokens_of(line);
    if (tok.empty()) continue;
    if (tok.size() < 2 || !is_number(tok[0])) {
      *err = "vertex file format error at line " + std::to_string(line_no);
      return -1;
    }
    uint64_t id = std::strtoull(tok[0].c_str(), nullptr, 10);
    uint64_t value = static_cast<uint64_t>(std::atoll(tok[1].c_str()));
    if (id != vi) {
      if (!(id > vi && id < g->vertex_count)) {
        *err = "vertex file not increasing at line " + std::to_string(line_no);
        return -1;
      }
      while (vi < id) g->values[vi++] = 0;
    }
    g->values[vi++] = value;
  }
  return 0;
}

int store_vertex_file(const Graph& g, const std::string& path, std::string* err) {
  std::ofstream os(path);
  if (!os) {
    *err = "cannot write " + path;
    return -1;
  }
  for (uint32_t v = 0; v < g.vertex_count; ++v) os << v << " " << g.values[v] << "\n";
  return os.good() ? 0 : -1;
}

void rmat_generate(const RmatParams& p, Graph* g) {
  const uint32_t V = 1u << p.scale;
  const uint64_t pairs = static_cast<uint64_t>(p.edge_factor) << p.scale;
  const double two32 = 4294967296.0;
  const uint64_t ta = static_cast<uint64_t>(std::floor(p.a * two32));
  const uint64_t tb = ta + static_cast<uint64_t>(std::floor(p.b * two32));
  const uint64_t tc = tb + static_cast<uint64_t>(std::floor(p.c * two32));
  const int threads = p.threads > 0 ? p.threads : omp_get_max_threads();

  auto pair_at = [&](uint64_t i, uint32_t* s, uint32_t* d) {
    uint32_t src = 0, dst = 0;
    for (int l = 0; l < p.scale; ++l) {
      uint64_t r = rmat_word(p.seed, i, l);
      uint32_t bs = 0, bd = 0;
      if (r < ta) {
      } else if (r < tb) {
        bd = 1;
      } else if (r < tc) {
        bs = 1;
      } else {
        bs = 1;
        bd = 1;
      }
      src = (src << 1) | bs;
      dst = (dst << 1) | bd;
    }
    *s = src;
    *d = dst;
  };

  const bool trace = std::getenv("PJ_TRACE") != nullptr;  // progress of large generations on stderr
  std::vector<uint64_t> deg(static_cast<size_t>(V) + 1, 0);
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(pairs); ++i) {
    uint32_t s, d;
    pair_at(static_cast<uint64_t>(i), &s, &d);
    if (s == d) continue;
    __atomic_fetch_add(&deg[s], 1ull, __ATOMIC_RELAXED);
    __atomic_fetch_add(&deg[d], 1ull, __ATOMIC_RELAXED);
  }
  if (trace) fprintf(stderr, "[rmat] scale %d: degrees of %llu pairs counted\n", p.scale, (unsigned long long)pairs);
  std::vector<uint64_t> off(static_cast<size_t>(V) + 1, 0);
  for (uint32_t v = 0; v < V; ++v) off[v + 1] = off[v] + deg[v];
  std::vector<uint32_t> buf(off[V]);
  std::vector<uint64_t> cur(off.begin(), off.end() - 1);
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t i = 0; i < static_cast<int64_t>(pairs); ++i) {
    uint32_t s, d;
    pair_at(static_cast<uint64_t>(i), &s, &d);
    if (s == d) continue;
    buf[__atomic_fetch_add(&cur[s], 1ull, __ATOMIC_RELAXED)] = d;
    buf[__atomic_fetch_add(&cur[d], 1ull, __ATOMIC_RELAXED)] = s;
  }
  if (trace) fprintf(stderr, "[rmat] scale %d: %llu slots placed\n", p.scale, (unsigned long long)off[V]);
  std::vector<uint64_t> ndeg(V, 0);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256)
  for (int64_t v = 0; v < static_cast<int64_t>(V); ++v) {
    uint32_t* b = buf.data() + off[v];
    uint32_t* e = buf.data() + off[v + 1];
    std::sort(b, e);
    ndeg[v] = static_cast<uint64_t>(std::unique(b, e) - b);
  }
  if (trace) fprintf(stderr, "[rmat] scale %d: rows sorted and deduplicated\n", p.scale);
  Graph out;
  out.vertex_count = V;
  out.directed = false;
  out.vertices.assign(static_cast<size_t>(V) + 1, 0);
  for (uint32_t v = 0; v < V; ++v) out.vertices[v + 1] = out.vertices[v] + ndeg[v];
  out.edge_count = out.vertices[V];
  out.edges.resize(out.edge_count);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1024)
  for (int64_t v = 0; v < static_cast<int64_t>(V); ++v)
    std::memcpy(out.edges.data() + out.vertices[v], buf.data() + off[v], ndeg[v] * sizeof(uint32_t));
  out.valued = g->valued;
  out.values = std::move(g->values);
  *g = std::move(out);
}

static inline uint32_t label_word(uint64_t seed, uint64_t v) {
  return static_cast<uint32_t>(splitmix64(seed ^ splitmix64(v)) >> 32);
}

void labels_uniform(Graph* g, uint32_t n, uint64_t seed) {
  g->valued = true;
  g->values.resize(g->vertex_count);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < static_cast<int64_t>(g->vertex_count); ++v)
    g->values[v] = (static_cast<uint64_t>(label_word(seed, v)) * n) >> 32;
}

std::vector<uint64_t> zipf_thresholds(uint32_t n, double s) {
  std::vector<double> w(n);
  double total = 0.0;
  for (uint32_t k = 0; k < n; ++k) {
    w[k] = std::pow(static_cast<double>(k + 1), -s);
    total += w[k];
  }
  std::vector<uint64_t> t(n);
  double cum = 0.0;
  for (uint32_t k = 0; k < n; ++k) {
    cum += w[k];
    t[k] = static_cast<uint64_t>(std::floor(cum / total * 4294967296.0));
  }
  t[n - 1] = 4294967296ull;
  return t;
}

void labels_zipf(Graph* g, uint32_t n, double s, uint64_t seed) {
  auto t = zipf_thresholds(n, s);
  g->valued = true;
  g->values.resize(g->vertex_count);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < static_cast<int64_t>(g->vertex_count); ++v) {
    uint64_t r = label_word(seed, v);
    uint32_t k = 0;
    while (r >= t[k]) ++k;
    g->values[v] = k;
  }
}

void pattern_from_edges(uint32_t n, const std::vector<uint32_t>& src, const std::vector<uint32_t>& dst,
                        const std::vector<uint64_t>& labels, Graph* g) {
  std::vector<std::vector<uint32_t>> adj(n);
  for (size_t i = 0; i < src.size(); ++i) {
    if (src[i] == dst[i]) continue;
    adj[src[i]].push_back(dst[i]);
    adj[dst[i]].push_back(src[i]);
  }
  Graph out;
  out.vertex_count = n;
  out.directed = false;
  out.vertices.assign(n + 1, 0);
  for (uint32_t v = 0; v < n; ++v) {
    std::sort(adj[v].begin(), adj[v].end());
    adj[v].erase(std::unique(adj[v].begin(), adj[v].end()), adj[v].end());
    out.vertices[v + 1] = out.vertices[v] + adj[v].size();
    out.edges.insert(out.edges.end(), adj[v].begin(), adj[v].end());
  }
  out.edge_count = out.edges.size();
  out.valued = true;
  out.values = labels;
  out.values.resize(n, 0);
  *g = std::move(out);
}

uint64_t count_asymmetric_slots(const Graph& g) {
  uint64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(dynamic, 1024)
  for (int64_t v = 0; v < static_cast<int64_t>(g.vertex_count); ++v) {
    for (uint64_t e = g.vertices[v]; e < g.vertices[v + 1]; ++e) {
      uint32_t u = g.edges[e];
      bool found = false;
      for (uint64_t f = g.vertices[u]; f < g.vertices[u + 1]; ++f)
        if (g.edges[f] == static_cast<uint32_t>(v)) {
          found = true;
          break;
        }
      if (!found) ++bad;
    }
  }
  return bad;
}

}  // namespace pj
