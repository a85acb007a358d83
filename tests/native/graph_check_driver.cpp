// Drives cgbfs_graph_check (host arrays, no GPU) over valid and malformed
// graphs, for the sanitizer test (tests/test_sanitizers.py).  Prints one
// line per case: the case name and the status code.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "cgbfs.h"

struct Case {
  const char *name;
  int64_t n;
  std::vector<int64_t> ptr;
  std::vector<int32_t> idx;
};

int main() {
  // a 4-vertex undirected graph: edges 0-1, 0-2, 2-3, and defects of it
  const std::vector<Case> cases = {
      {"valid", 4, {0, 2, 3, 5, 6}, {1, 2, 0, 0, 3, 2}},
      {"decreasing_ptr", 4, {0, 2, 1, 5, 6}, {1, 2, 0, 0, 3, 2}},
      {"ptr_past_nnz", 4, {0, 2, 3, 5, 9}, {1, 2, 0, 0, 3, 2}},
      {"ptr_negative", 4, {0, -4, 3, 5, 6}, {1, 2, 0, 0, 3, 2}},
      {"id_out_of_range", 4, {0, 2, 3, 5, 6}, {1, 400000, 0, 0, 3, 2}},
      {"id_negative", 4, {0, 2, 3, 5, 6}, {1, -7, 0, 0, 3, 2}},
      {"unsorted", 4, {0, 2, 3, 5, 6}, {2, 1, 0, 0, 3, 2}},
      {"not_symmetric", 4, {0, 2, 3, 5, 6}, {1, 2, 0, 1, 3, 2}},
      {"empty", 3, {0, 0, 0, 0}, {}},
  };
  for (const Case &c : cases) {
    const int64_t nnz = (int64_t)c.idx.size();
    const int rc = cgbfs_graph_check(c.n, nnz, c.ptr.data(), c.idx.empty() ? nullptr : c.idx.data(),
                                     nullptr, nullptr, 1);
    std::printf("%s %d\n", c.name, rc);
  }
  // a larger valid graph: a ring of 200000 vertices (the threaded path)
  const int64_t n = 200000;
  std::vector<int64_t> ptr(n + 1);
  std::vector<int32_t> idx(2 * n);
  for (int64_t v = 0; v < n; v++) {
    ptr[v] = 2 * v;
    int32_t a = (int32_t)((v + n - 1) % n), b = (int32_t)((v + 1) % n);
    idx[2 * v] = a < b ? a : b;
    idx[2 * v + 1] = a < b ? b : a;
  }
  ptr[n] = 2 * n;
  std::printf("ring %d\n", cgbfs_graph_check(n, 2 * n, ptr.data(), idx.data(), nullptr, nullptr, 1));
  idx[2 * 777 + 1] = (int32_t)n;  // out of range deep inside
  std::printf("ring_bad %d\n", cgbfs_graph_check(n, 2 * n, ptr.data(), idx.data(), nullptr, nullptr, 1));
  return 0;
}
