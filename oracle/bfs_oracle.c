/*
 * CPU oracle for the BFS explore / level-generation path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.
 *
 * Functions and what pins each one (tests/test_oracle.py):
 *   oracle_bfs_levelsync  restatement of src/run_bfs_cpu.hpp:5-85 with the
 *                         initial state of src/host.cpp:161-169.  Pinned by
 *                         the reference's own run_bfs_cpu.hpp compiled from
 *                         its sources (oracle/_ref), closed forms (path, star,
 *                         binary tree, grid), invariants and brute force.
 *   oracle_bfs_queue      exact FIFO BFS, O(V+E); same outputs as the
 *                         level-synchronous form.  Pinned against it and the
 *                         closed forms.
 *   oracle_msbfs_u8       multi-source BFS per bit lane with uint8 levels
 *                         (0xFF = unreached).  Pinned against per-source
 *                         oracle_bfs_queue runs (min over sources).
 *   oracle_check_parents  parent validity: parent is an in-neighbour one
 *                         level up.  Pinned by hand-built cases.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t cp_at(const void *col_ptr, int ptr_bytes, uint64_t i) {
  return ptr_bytes == 8 ? ((const uint64_t *)col_ptr)[i]
                        : (uint64_t)((const uint32_t *)col_ptr)[i];
}

/*
 * Level-synchronous top-down BFS over byte masks, as in
 * src/run_bfs_cpu.hpp:24-75: every level scans tid < n_rows, expands masked
 * vertices (adding their degree to the explored-edge count, :36), marks
 * unvisited neighbours with level+1 (:41-45), then moves the updating mask to
 * the frontier mask (:67-73).  The initial state is src/host.cpp:161-169:
 * levels -1 everywhere, root 0, root masked and visited.
 * Returns the number of level iterations (the reference's final `level`), or
 * -1 on bad arguments / allocation failure.
 */
int64_t oracle_bfs_levelsync(uint32_t n_rows, uint32_t n_cols,
                             const void *col_ptr, int ptr_bytes,
                             const uint32_t *row_idx, uint32_t root,
                             int32_t *levels, uint64_t *edges_count) {
  if (root >= n_cols || n_rows > n_cols || (ptr_bytes != 4 && ptr_bytes != 8))
    return -1;
  uint8_t *mask = calloc(n_cols ? n_cols : 1, 1);
  uint8_t *updating = calloc(n_cols ? n_cols : 1, 1);
  uint8_t *visited = calloc(n_cols ? n_cols : 1, 1);
  if (!mask || !updating || !visited) {
    free(mask); free(updating); free(visited);
    return -1;
  }
  for (uint32_t i = 0; i < n_cols; i++) levels[i] = -1;
  mask[root] = 1;
  visited[root] = 1;
  levels[root] = 0;
  uint64_t explored = 0;
  int level = 0;
  char stop;
  do {
    stop = 0;
    for (uint32_t tid = 0; tid < n_rows; tid++) {
      if (mask[tid] == 1) {
        mask[tid] = 0;
        uint64_t b = cp_at(col_ptr, ptr_bytes, tid);
        uint64_t e = cp_at(col_ptr, ptr_bytes, (uint64_t)tid + 1);
        explored += e - b;
        for (uint64_t i = b; i < e; i++) {
          uint32_t id = row_idx[i];
          if (!visited[id]) {
            levels[id] = level + 1;
            updating[id] = 1;
            visited[id] = 1;
          }
        }
      }
    }
    for (uint32_t tid = 0; tid < n_rows; tid++) {
      if (updating[tid] == 1) {
        mask[tid] = 1;
        stop = 1;
        updating[tid] = 0;
      }
    }
    level++;
  } while (stop);
  free(mask); free(updating); free(visited);
  if (edges_count) *edges_count = explored;
  return level;
}

/*
 * Exact FIFO BFS with the same contract as oracle_bfs_levelsync: only
 * vertices < n_rows are expanded (the reference's loop bound), levels array
 * has n_cols entries, edges_count = sum of degrees of expanded vertices.
 * Returns the number of level iterations the level-synchronous form would
 * report (deepest level holding a vertex < n_rows, plus one), or -1 on error.
 */
int64_t oracle_bfs_queue(uint32_t n_rows, uint32_t n_cols, const void *col_ptr,
                         int ptr_bytes, const uint32_t *row_idx, uint32_t root,
                         int32_t *levels, uint64_t *edges_count) {
  if (root >= n_cols || n_rows > n_cols || (ptr_bytes != 4 && ptr_bytes != 8))
    return -1;
  uint32_t *q = malloc(sizeof(uint32_t) * (n_cols ? n_cols : 1));
  if (!q) return -1;
  for (uint32_t i = 0; i < n_cols; i++) levels[i] = -1;
  levels[root] = 0;
  uint64_t head = 0, tail = 0, explored = 0;
  q[tail++] = root;
  while (head < tail) {
    uint32_t v = q[head++];
    if (v >= n_rows) continue;
    uint64_t b = cp_at(col_ptr, ptr_bytes, v);
    uint64_t e = cp_at(col_ptr, ptr_bytes, (uint64_t)v + 1);
    explored += e - b;
    int32_t nl = levels[v] + 1;
    for (uint64_t i = b; i < e; i++) {
      uint32_t w = row_idx[i];
      if (levels[w] < 0) {
        levels[w] = nl;
        q[tail++] = w;
      }
    }
  }
  free(q);
  if (edges_count) *edges_count = explored;
  /* The level-synchronous loop (run_bfs_cpu.hpp:25-75) runs iteration i to
     expand level i and continues only while level i+1 holds a vertex below
     n_rows (its update loop, :67, stops at n_rows). */
  int32_t deepest_row = -1;
  for (uint32_t i = 0; i < n_rows && i < n_cols; i++)
    if (levels[i] > deepest_row) deepest_row = levels[i];
  if (root >= n_rows) return 1;
  return (int64_t)deepest_row + 1;
}

/*
 * Multi-source BFS per bit lane.  Lane b starts from every source s with
 * bits[s] == b at level 0; levels_out[b * n_cols + v] is the hop distance
 * from the nearest such source, 0xFF when unreached.  Only vertices < n_rows
 * are expanded.  Returns 0, -1 on bad arguments, -2 if a level would exceed
 * 254 (not representable).
 */
int oracle_msbfs_u8(uint32_t n_rows, uint32_t n_cols, const void *col_ptr,
                    int ptr_bytes, const uint32_t *row_idx,
                    const uint32_t *sources, const uint16_t *bits,
                    uint32_t n_sources, uint32_t n_bits, uint8_t *levels_out) {
  if (n_rows > n_cols || (ptr_bytes != 4 && ptr_bytes != 8)) return -1;
  for (uint32_t s = 0; s < n_sources; s++)
    if (sources[s] >= n_cols || bits[s] >= n_bits) return -1;
  uint32_t *q = malloc(sizeof(uint32_t) * (n_cols ? n_cols : 1));
  if (!q) return -1;
  memset(levels_out, 0xFF, (size_t)n_bits * n_cols);
  int rc = 0;
  for (uint32_t b = 0; b < n_bits && rc == 0; b++) {
    uint8_t *lv = levels_out + (size_t)b * n_cols;
    uint64_t head = 0, tail = 0;
    for (uint32_t s = 0; s < n_sources; s++) {
      if (bits[s] == b && lv[sources[s]] == 0xFF) {
        lv[sources[s]] = 0;
        q[tail++] = sources[s];
      }
    }
    while (head < tail) {
      uint32_t v = q[head++];
      if (v >= n_rows) continue;
      uint64_t e0 = cp_at(col_ptr, ptr_bytes, v);
      uint64_t e1 = cp_at(col_ptr, ptr_bytes, (uint64_t)v + 1);
      unsigned nl = (unsigned)lv[v] + 1;
      for (uint64_t i = e0; i < e1; i++) {
        uint32_t w = row_idx[i];
        if (lv[w] == 0xFF) {
          if (nl > 254) { rc = -2; break; }
          lv[w] = (uint8_t)nl;
          q[tail++] = w;
        }
      }
      if (rc) break;
    }
  }
  free(q);
  return rc;
}

/*
 * Checks a parent array against levels: for every reached v != root,
 * parent[v] must be a vertex u < n_rows with levels[u] == levels[v] - 1 and
 * v in u's adjacency list; the root's parent must be itself; unreached
 * vertices must have parent 0xFFFFFFFF.  Returns the number of violations.
 */
uint64_t oracle_check_parents(uint32_t n_rows, uint32_t n_cols,
                              const void *col_ptr, int ptr_bytes,
                              const uint32_t *row_idx, uint32_t root,
                              const int32_t *levels, const uint32_t *parent) {
  /* whether each list is in ascending order (then a binary search finds v;
     otherwise a scan): a hub is the parent of many vertices, and scanning
     its list once per child would be quadratic */
  uint8_t *sorted = calloc(n_rows ? n_rows : 1, 1);
  if (!sorted) return UINT64_MAX;
  for (uint32_t u = 0; u < n_rows; u++) {
    uint64_t b = cp_at(col_ptr, ptr_bytes, u), e = cp_at(col_ptr, ptr_bytes, (uint64_t)u + 1);
    int ok = 1;
    for (uint64_t i = b + 1; i < e && ok; i++) ok = row_idx[i - 1] <= row_idx[i];
    sorted[u] = (uint8_t)ok;
  }
  uint64_t bad = 0;
  for (uint32_t v = 0; v < n_cols; v++) {
    if (levels[v] < 0) { bad += parent[v] != 0xFFFFFFFFu; continue; }
    if (v == root) { bad += parent[v] != root; continue; }
    uint32_t u = parent[v];
    if (u >= n_rows || levels[u] != levels[v] - 1) { bad++; continue; }
    uint64_t b = cp_at(col_ptr, ptr_bytes, u);
    uint64_t e = cp_at(col_ptr, ptr_bytes, (uint64_t)u + 1);
    int found = 0;
    if (sorted[u]) {
      while (b < e) {  /* first entry >= v */
        uint64_t m = b + (e - b) / 2;
        if (row_idx[m] < v) b = m + 1; else e = m;
      }
      found = b < cp_at(col_ptr, ptr_bytes, (uint64_t)u + 1) && row_idx[b] == v;
    } else {
      for (uint64_t i = b; i < e && !found; i++) found = row_idx[i] == v;
    }
    bad += !found;
  }
  free(sorted);
  return bad;
}
