// Shared declarations for the CereGraphs-MI355X BFS library (gfx950).
//
// Conventions used by every kernel in this library:
//   * A graph is the boolean pattern matrix M of the reference, where a
//     nonzero M[r][c] is the edge c -> r (bfs/implementation/graph_loader.py:76-90).
//     out_ptr/out_idx hold column c of M (the out-neighbours of c, used by the
//     push/top-down kernels); in_ptr/in_idx hold row r of M (the in-neighbours
//     of r, sorted ascending, used by the pull/bottom-up kernels).  For a
//     symmetric graph the two views share one allocation.
//   * Offsets are int64 (s26 ef16 has ~2.0e9 directed entries), column ids
//     int32.
//   * A wavefront is 64 lanes; bitmaps are uint32 words, so one wavefront
//     covers exactly two words when it owns 64 consecutive vertices.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/cgbfs.h"

namespace cg {

constexpr int kWave = 64;
constexpr uint32_t kParentNone = 0xFFFFFFFFu;  // int32 -1 after the cast

// Records a library error message (thread-local) and returns the code.
int set_error(int code, const std::string &msg);
int hip_fail(hipError_t err, const char *what, const char *file, int line);

#define CG_HIP(call)                                                   \
  do {                                                                 \
    hipError_t cg_err__ = (call);                                      \
    if (cg_err__ != hipSuccess)                                        \
      return ::cg::hip_fail(cg_err__, #call, __FILE__, __LINE__);      \
  } while (0)

// Device-resident graph owned by one handle.
struct DeviceGraph {
  int device = 0;
  int64_t n = 0;
  int64_t nnz = 0;
  int symmetric = 0;
  int64_t *out_ptr = nullptr;
  int32_t *out_idx = nullptr;
  int64_t *in_ptr = nullptr;  // == out_ptr when symmetric
  int32_t *in_idx = nullptr;  // == out_idx when symmetric
  // Per-vertex GTEPS weight: #{c <= v in row v} (symmetric, the
  // undirected_dedup rule) or out-degree (directed_source_visited),
  // bfs/implementation/bfs_timing.py:390-395.
  int32_t *m_weight = nullptr;
  int64_t total_out_edges = 0;
  hipStream_t stream = nullptr;
  // Lazily sized single-root workspace.
  void *ws1 = nullptr;
  size_t ws1_bytes = 0;
  // Lazily sized multi-source workspace.
  void *wsm = nullptr;
  size_t wsm_bytes = 0;
  // Per-graph multi-source data (segment lists of high-degree vertices),
  // built on first use; freed through msdata_free.
  void *msdata = nullptr;
  void (*msdata_free)(void *) = nullptr;
  // Pinned host mailbox for the done flag polled between level chunks.
  int32_t *host_flag = nullptr;
};

int graph_finalize(DeviceGraph *g);  // computes m_weight, stream, mailbox

// HIP events owned for the length of one call (destroyed on every exit path).
struct EventPool {
  std::vector<hipEvent_t> ev;
  hipEvent_t make() {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ev.push_back(e);
    return e;
  }
  ~EventPool() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
};

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Appends `count` items (wave-aggregated) to a queue; returns the base index
// for this lane's item, valid only where `want` is true.
__device__ __forceinline__ int64_t wave_append(unsigned long long *tail, bool want) {
  uint64_t mask = ballot64(want);
  if (mask == 0) return -1;
  int leader = __ffsll((unsigned long long)mask) - 1;
  unsigned long long base = 0;
  if (lane_id() == leader) base = atomicAdd(tail, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader);
  uint64_t below = mask & ((1ull << lane_id()) - 1ull);
  return (int64_t)(base + __popcll(below));
}

// Sums a 64-bit value across the wavefront.
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

}  // namespace cg

// The opaque handle of the C ABI; DeviceGraph is its first and only member.
struct cgbfs_graph {
  cg::DeviceGraph g;
};
