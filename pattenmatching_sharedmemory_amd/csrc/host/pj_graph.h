// Host-side graph container, TOTEM file formats and synthetic generators.
//
// Graph mirrors the fields of TOTEM's graph_t that the pruning path reads
// (src/totem/totem_graph.h:137-180): CSR offsets `vertices` (always 64-bit
// here, the reference's EID64 layout), neighbour ids `edges` (uint32) and
// vertex values `values` (uint64 labels).
#ifndef PJ_GRAPH_H
#define PJ_GRAPH_H

#include <cstdint>
#include <string>
#include <vector>

namespace pj {

struct Graph {
  uint32_t vertex_count = 0;
  uint64_t edge_count = 0;
  std::vector<uint64_t> vertices;  // [V+1]
  std::vector<uint32_t> edges;     // [E]
  std::vector<uint64_t> values;    // [V] when valued
  bool valued = false;
  bool weighted = false;
  bool directed = true;
};

// graph_initialize (src/totem/totem_graph.cu:388-409): binary when the first
// word is the magic 0x10102048, text otherwise.  Binary files written with
// 32-bit or 64-bit edge ids are both accepted (widened to 64-bit offsets).
// Edge weights are skipped.  Returns 0 / -1 (message in *err).
int load_totem(const std::string& path, Graph* g, std::string* err);

// Header and vertex values of a binary TOTEM file, without its offsets and
// edge list (for pj_upload_totem_file, which streams those to the device):
// vertex_count, edge_count, valued and values are set, vertices and edges
// stay empty.  Returns 1 when `path` is not a binary TOTEM file, 0 / -1.
int load_totem_binary_values(const std::string& path, Graph* g, std::string* err);

// CSR well-formedness: offsets start at 0, never decrease, end at the edge
// count, and every neighbour id is a vertex.  Every graph the loaders and
// pjh_graph_from_csr return has passed it (kernels index cand[edges[e]] and
// off[v + 1] without bounds checks).  Returns 0 / -1 (message in *err).
int validate_csr(const Graph& g, std::string* err);

// Text parser only (# Nodes / # Edges / # DIRECTED|UNDIRECTED header,
// optional vertex list, edge list sorted by source), totem_graph.cu:30-350.
int load_totem_text(const std::string& path, Graph* g, std::string* err);

// graph_store_binary (totem_graph.cu:883-923) with `eid_bytes` 4 or 8.
int store_totem_binary(const Graph& g, const std::string& path, int eid_bytes,
                       std::string* err);

// Text writer in the same format the text parser reads.
int store_totem_text(const Graph& g, const std::string& path, std::string* err);

// parseVertexFile (src/patternmatching/utils/common_utils.cc:9-59): "id label"
// lines in increasing id order, '#' comments, gaps filled with 0.  Vertices
// after the last line also get 0 (the reference leaves them unset).
int parse_vertex_file(const std::string& path, Graph* g, std::string* err);

// Writes "id label" lines for every vertex.
int store_vertex_file(const Graph& g, const std::string& path, std::string* err);

// splitmix64 finaliser (Steele, Lea, Flood 2014), the counter-based RNG of
// every generator below.
inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct RmatParams {
  int scale = 14;
  int edge_factor = 16;
  double a = 0.57, b = 0.19, c = 0.19;  // d = 1 - a - b - c
  uint64_t seed = 1;
  int threads = 0;  // 0 = OpenMP default
};

// Undirected R-MAT graph: edge_factor * 2^scale generated pairs, each drawn
// with `scale` quadrant choices (no per-level noise), self-loops dropped,
// both directions stored, duplicates removed, rows in ascending id order.
// Quadrant choice of pair i at level l uses the 32-bit word
// rmat_word(seed, i, l) against integer thresholds floor(p * 2^32).
void rmat_generate(const RmatParams& p, Graph* g);

inline uint32_t rmat_word(uint64_t seed, uint64_t pair, int level) {
  uint64_t h = splitmix64(seed ^ splitmix64(pair));
  uint64_t u = splitmix64(h + 0x632BE59BD9B4E019ull * (uint64_t)(level >> 1));
  return (level & 1) ? (uint32_t)u : (uint32_t)(u >> 32);
}

// Labels drawn per vertex from hi32(splitmix64(seed ^ splitmix64(v))).
// Uniform: label = (word * n) >> 32.
void labels_uniform(Graph* g, uint32_t n_labels, uint64_t seed);
// Zipf(s) over ranks 0..n-1, P(k) proportional to (k+1)^-s: label = first k
// with word < T_k, T_k = floor(CDF(k) * 2^32) (T_{n-1} = 2^32).
void labels_zipf(Graph* g, uint32_t n_labels, double s, uint64_t seed);
std::vector<uint64_t> zipf_thresholds(uint32_t n_labels, double s);

// Small undirected template from an edge list: every edge stored in both
// directions, rows sorted ascending, labels attached.
void pattern_from_edges(uint32_t vertex_count, const std::vector<uint32_t>& src,
                        const std::vector<uint32_t>& dst,
                        const std::vector<uint64_t>& labels, Graph* g);

// Number of slots (u,v) whose reverse (v,u) is missing.
uint64_t count_asymmetric_slots(const Graph& g);

}  // namespace pj

#endif
