// Multilevel vertex separators for the nested-dissection ordering.
#pragma once

#include <cstdint>
#include <vector>

namespace vjx {

// Vertex separator of the subgraph of `adj` (symmetric adjacency without self
// loops) induced by `nodes`: part[i] is 0 or 1 for the two halves and 2 for
// the separator of nodes[i]; no edge joins a vertex of half 0 to one of half
// 1.  The refinement moves no vertex into a half that already holds 60 % of
// the vertices (the coarsest bisection aims at one half each).  `loc` is scratch
// indexed by global vertex, all -1 on entry and on return.  Deterministic for
// a given seed.
void vertex_separator(const std::vector<std::vector<int32_t>>& adj,
                      const std::vector<int32_t>& nodes, std::vector<int32_t>& loc, uint64_t seed,
                      std::vector<int8_t>& part);

}  // namespace vjx
