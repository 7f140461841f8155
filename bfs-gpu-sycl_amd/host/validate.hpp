// Level validation without a second search (Graph500-style certificate).
//
// The reference validates GPU levels by re-running the search on the CPU and
// comparing (src/host.cpp:172-188).  A level array is the exact hop-distance
// array from `root` if and only if:
//   (1) level[root] == 0 and no level is below -1;
//   (2) for every edge u -> v with u < n_rows reached, v is reached and
//       level[v] <= level[u] + 1;
//   (3) every reached v != root has an in-edge u -> v with u < n_rows and
//       level[u] == level[v] - 1.
// (3) gives a path of length level[v] (so level >= distance); (2) applied
// along a shortest path gives level <= distance and reaches every reachable
// vertex; a reached vertex has a path, so unreachable ones stay at -1.
// Checking this is O(V + E).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bfsx {

template <class PtrT>
inline std::string validate_levels(uint32_t n_rows, const PtrT *col_ptr, const uint32_t *row_idx,
                                   uint32_t root, const std::vector<int> &lev) {
  const size_t n = lev.size();
  if (root >= n) return "root out of range";
  if (lev[root] != 0) return "root is not at level 0";
  std::vector<uint8_t> has_parent(n, 0);
  for (size_t v = 0; v < n; v++)
    if (lev[v] < -1) return "level below -1 at vertex " + std::to_string(v);
  for (uint32_t u = 0; u < n_rows && u < n; u++) {
    if (lev[u] < 0) continue;
    for (uint64_t e = col_ptr[u]; e < (uint64_t)col_ptr[u + 1]; e++) {
      const uint32_t v = row_idx[e];
      if (v >= n) return "edge target out of range";
      if (lev[v] < 0 || lev[v] > lev[u] + 1)
        return "edge " + std::to_string(u) + "->" + std::to_string(v) + " skips a level";
      if (lev[v] == lev[u] + 1) has_parent[v] = 1;
    }
  }
  for (size_t v = 0; v < n; v++)
    if (lev[v] > 0 && !has_parent[v])
      return "vertex " + std::to_string(v) + " has no parent one level up";
    else if (lev[v] == 0 && v != root)
      return "vertex " + std::to_string(v) + " at level 0 is not the root";
  return "";
}

// The reference's edgesCount (src/run_bfs_cpu.hpp:36,84): out-degree summed
// over reached vertices below n_rows.
template <class PtrT>
inline uint64_t traversed_edges(uint32_t n_rows, const PtrT *col_ptr, const std::vector<int> &lev) {
  uint64_t s = 0;
  for (uint32_t u = 0; u < n_rows && u < lev.size(); u++)
    if (lev[u] >= 0) s += (uint64_t)(col_ptr[u + 1] - col_ptr[u]);
  return s;
}

}  // namespace bfsx
