// Thin C entry point around the reference's own GUDHI edge collapser, compiled from the
// reference source where it lies (include/Flag_complex_edge_collapser.h) into oracle/_ref.
// Used only by tests to pin oracle/ssq_oracle.c:or_collapse_edges and to write fixtures.
// Edge i gets filtration float(i), exactly as specseq.h:209-212 builds its edge list.
#include <cstdint>
#include <map>
#include <tuple>
#include <utility>
#include <vector>

#include "Flag_complex_edge_collapser.h"

extern "C" int ref_collapse_edges(int64_t ne, const int32_t* u, const int32_t* v, uint8_t* kept) {
  using Filtered_edge = std::tuple<int32_t, int32_t, float>;
  std::vector<Filtered_edge> graph;
  std::map<std::pair<int32_t, int32_t>, int64_t> id;
  graph.reserve((size_t)ne);
  for (int64_t i = 0; i < ne; i++) {
    graph.emplace_back(u[i], v[i], (float)graph.size());
    id[std::make_pair(u[i], v[i])] = i;
    kept[i] = 0;
  }
  auto rem = Gudhi::collapse::flag_complex_collapse_edges(graph);
  for (auto& e : rem) kept[id[std::make_pair(std::get<0>(e), std::get<1>(e))]] = 1;
  return 0;
}
