/*
 * CPU oracle for the BFS hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline).
 * The product path (ceregraphs_amd) never calls it.
 *
 * Matrix orientation follows the reference: a nonzero M[r][c] is an edge
 * c -> r ("row = destination, column = source",
 * bfs/implementation/graph_loader.py:76-90).  Two adjacency views are
 * passed in:
 *   out_ptr/out_idx : column c of M = the out-neighbours of c (CSC of M)
 *   in_ptr/in_idx   : row r of M    = the in-neighbours of r (CSR of M),
 *                     sorted ascending inside each row.
 *
 * Round semantics follow bool_pe.csl:
 *   - round r expands the level-r frontier and discovers level r+1
 *     (compute_topdown :865-887, reduce_done :947-1004);
 *   - nf_history[r] = number of vertices discovered in round r
 *     (term_col_bcast_done :1063-1065);
 *   - the search stops after the first round that discovers nothing, so
 *     rounds_completed = max_level + 1 (:1068, reduce_done's rounds++).
 * Parents follow the lowest-id rule (NOT the reference's walk-order rule,
 * bool_pe.csl:78-93): parent[v] = min{u : M[v][u] != 0, level[u] = level[v]-1},
 * parent[source] = source (run_bfs.py:508), -1 when unreached
 * (device_io.py:157-178).
 *
 * Method: a plain FIFO queue BFS for levels, then an independent per-vertex
 * scan of the ascending in-neighbour list for the parent.  This is a
 * deliberately different method from the device kernels (which record
 * parents during traversal with atomicMin / first-hit pulls).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int oracle_bfs(int64_t n, const int64_t *out_ptr, const int32_t *out_idx,
               const int64_t *in_ptr, const int32_t *in_idx, int64_t source,
               int32_t *parents, int32_t *levels, int64_t *nf_history,
               int64_t max_hist, int64_t *rounds_completed) {
  if (n <= 0 || source < 0 || source >= n) return -1;
  int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  if (!queue) return -2;
  for (int64_t v = 0; v < n; v++) {
    levels[v] = -1;
    parents[v] = -1;
  }
  int64_t head = 0, tail = 0;
  levels[source] = 0;
  queue[tail++] = (int32_t)source;
  int64_t level_end = tail;
  int64_t round = 0;
  int64_t found = 0;
  /* Level-synchronous walk so per-round discovery counts fall out. */
  for (;;) {
    found = 0;
    while (head < level_end) {
      int32_t u = queue[head++];
      for (int64_t e = out_ptr[u]; e < out_ptr[u + 1]; e++) {
        int32_t v = out_idx[e];
        if (levels[v] < 0) {
          levels[v] = (int32_t)(round + 1);
          queue[tail++] = v;
          found++;
        }
      }
    }
    if (round < max_hist) nf_history[round] = found;
    round++;
    if (found == 0) break;
    level_end = tail;
  }
  *rounds_completed = round;
  /* Lowest-id parent: first in-neighbour (ascending) one level closer. */
  for (int64_t v = 0; v < n; v++) {
    int32_t lv = levels[v];
    if (lv <= 0) continue;
    for (int64_t e = in_ptr[v]; e < in_ptr[v + 1]; e++) {
      int32_t u = in_idx[e];
      if (levels[u] == lv - 1) {
        parents[v] = u;
        break;
      }
    }
  }
  parents[source] = (int32_t)source;
  free(queue);
  return 0;
}

/*
 * m for one search, bfs/implementation/bfs_timing.py:390-395:
 *   symmetric: #{(r, c) in M : visited[r] and c <= r}  ("undirected_dedup")
 *   directed : #{(r, c) in M : visited[c]}             ("directed_source_visited")
 * Walked over the in-neighbour (CSR of M) view: row r, column c.
 */
int64_t oracle_m_edges(int64_t n, const int64_t *in_ptr, const int32_t *in_idx,
                       const int32_t *levels, int symmetric) {
  int64_t m = 0;
  for (int64_t r = 0; r < n; r++) {
    for (int64_t e = in_ptr[r]; e < in_ptr[r + 1]; e++) {
      int32_t c = in_idx[e];
      if (symmetric) {
        if (levels[r] >= 0 && c <= r) m++;
      } else {
        if (levels[c] >= 0) m++;
      }
    }
  }
  return m;
}
