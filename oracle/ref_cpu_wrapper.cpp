// Compiles the reference's own CPU BFS, src/run_bfs_cpu.hpp, from its source
// file where it lies, into oracle/_ref/libref_cpu.so.  TEST INFRASTRUCTURE
// ONLY: used to pin the oracle restatement and, optionally, as the timed CPU
// baseline.
//
// run_bfs_cpu.hpp needs three things from the rest of the reference:
//   * nlohmann::json  - include/json.hpp (vendored in the reference, compiled
//                       from there);
//   * NUM_GPU/VERBOSE - compile-time macros set by src/CMakeLists.txt:60-62;
//                       defined below to the single-partition, non-verbose
//                       build (NUM_GPU=1, VERBOSE=0);
//   * DeviceInfo      - the 3-field struct of src/functions.hpp:165-169, used
//                       only as the element type of an argument that is
//                       touched under VERBOSE==1.  functions.hpp itself cannot
//                       be included without SYCL (it pulls
//                       include/metaprogramming_utils.hpp:126, which names a
//                       SYCL type), so the struct is declared here with the
//                       same fields.  Nothing the image lacks is stubbed.
//
// The calling sequence follows src/host.cpp:161-176: fresh byte masks of
// numCols entries, root masked, visited and at level 0, levels -1 elsewhere.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

#include "json.hpp"

#define NUM_GPU 1
#define VERBOSE 0
struct DeviceInfo {
  int gpuID;
  long long edgesProcessed;
  double percentage;
};

#include "run_bfs_cpu.hpp"

namespace {
struct RefGraph {
  uint32_t n_rows = 0, n_cols = 0;
  std::vector<uint32_t> indptr, inds;
};
}  // namespace

extern "C" {

// Copies a CSC graph (uint32 col_ptr, as the reference's vectorT holds) into
// the vectors the reference function takes.  Returns an opaque handle.
void *ref_graph_create(uint32_t n_rows, uint32_t n_cols, const uint32_t *col_ptr,
                       const uint32_t *row_idx, uint64_t nnz) {
  RefGraph *g = new RefGraph;
  g->n_rows = n_rows;
  g->n_cols = n_cols;
  g->indptr.assign(col_ptr, col_ptr + (uint64_t)n_cols + 1);
  g->inds.assign(row_idx, row_idx + nnz);
  return g;
}

void ref_graph_destroy(void *h) { delete static_cast<RefGraph *>(h); }

// Runs run_bfs_cpu from `root`; writes numCols levels and the reference's
// edgesCount; *ms is the wall time of the run_bfs_cpu call alone.
int ref_run_bfs_cpu(void *h, uint32_t root, int32_t *levels_out,
                    uint64_t *edges_count, double *ms) {
  RefGraph *g = static_cast<RefGraph *>(h);
  if (!g || root >= g->n_cols) return -1;
  std::vector<uint8_t> host_graph_mask(g->n_cols, 0);
  std::vector<uint8_t> host_updating_graph_mask(g->n_cols, 0);
  std::vector<uint8_t> host_graph_visited(g->n_cols, 0);
  std::vector<int> host_level(g->n_cols, -1);
  host_graph_mask[root] = 1;
  host_graph_visited[root] = 1;
  host_level[root] = 0;
  std::vector<DeviceInfo> host_run_statistics;
  std::vector<uint32_t> h_visit_offsets = {0, g->n_rows};
  nlohmann::json obj;
  auto t0 = std::chrono::steady_clock::now();
  run_bfs_cpu(g->n_rows, g->indptr, g->inds, host_graph_mask,
              host_updating_graph_mask, host_graph_visited, host_level, obj,
              h_visit_offsets, host_run_statistics);
  auto t1 = std::chrono::steady_clock::now();
  if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  std::memcpy(levels_out, host_level.data(), sizeof(int32_t) * g->n_cols);
  if (edges_count) *edges_count = obj["edgesCount"].get<uint64_t>();
  return 0;
}
}
