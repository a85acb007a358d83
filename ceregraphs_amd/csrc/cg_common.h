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

// The first kRowHead entries of an in-row, one vector load.
#ifndef CG_ROW_HEAD
#define CG_ROW_HEAD 2
#endif
constexpr int kRowHead = CG_ROW_HEAD;
struct alignas(4 * kRowHead) RowHead {
  int32_t id[kRowHead];
};

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
  // Row heads of the in view: the first kRowHead in-neighbours of every
  // vertex (-1 past the row end), in vertex order.  A pull's head probe
  // reads them with the row offsets, in order, instead of one scattered
  // line of in_idx per row.
  RowHead *in_head = nullptr;
  // One bit per vertex without in-edges ([ceil(n/64)] words, bit v & 63 of
  // word v >> 6): no search reaches such a vertex unless it is the source.
  uint64_t *noin = nullptr;
  int64_t total_out_edges = 0;
  int64_t n_in = 0;  // vertices with at least one in-edge
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
  // Single-root graph replay: the last instantiated chunk of level rounds
  // and the launch arguments it was captured with (reused while they match).
  hipGraphExec_t bfs_exec = nullptr;
  std::vector<char> bfs_key;
};

int graph_finalize(DeviceGraph *g);  // computes m_weight, stream, mailbox

// Structural checks of a graph (check.hip): offsets 0..nnz non-decreasing,
// ids in [0, n), rows strictly ascending, in view == transpose of out view.
// The host check reads host arrays; the device check reads g's own arrays
// on g's stream.  Both return CGBFS_OK or CGBFS_EINVAL (message set).
int check_graph_host(int64_t n, int64_t nnz, const int64_t *out_ptr, const int32_t *out_idx,
                     const int64_t *in_ptr, const int32_t *in_idx, int symmetric);
int check_graph_device(const DeviceGraph *g);

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

// Control words (queue tails, counters, flags) are updated by atomics from
// every XCD and read by later kernels, often single-block ones.  A plain load
// may be served from a line the reading XCD's L2 still holds, so they are
// read and written with agent-scope atomic operations only.
template <class T>
__device__ __forceinline__ T ld_ctrl(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_ctrl(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The global-memory view of a pointer.  Accesses through it compile to
// global_* instructions.  Through a generic pointer that was itself read from
// LDS or memory (the per-lane output tables) they compile to flat_* ones,
// which also count against the LDS wait counter: every later wait for an LDS
// read then waits for those stores or atomics to complete as well.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}

// Search outputs (parents, levels) are written once and not read again by
// the search: a non-temporal store keeps them from evicting the lane words
// that the gathers of later rounds reuse.
template <class T>
__device__ __forceinline__ void st_once(T *p, T v) {
  __builtin_nontemporal_store(v, gmem(p));
}

// A counter on a 128-byte line of its own.  Device-scope atomics on one line
// are serialised, so counters that every block updates (round totals, queue
// tails, barrier arrivals) each get a line.
struct alignas(128) LineCtr {
  unsigned long long v;
};

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// A pull hands out its 64-vertex chunks dynamically from kParts cursors, one
// per contiguous part of the vertex range, kept a cache line apart.
#ifndef CG_PARTS
#define CG_PARTS 8
#endif
constexpr int kParts = CG_PARTS;
constexpr int kMaxParts = 64;  // cursor areas are sized for this many parts
static_assert(kParts >= 1 && kParts <= kMaxParts, "kParts out of range");
constexpr int kCursorStride = 16;

// Zeroes the pull's chunk cursors for the next round (one thread).
__device__ __forceinline__ void reset_cursors(unsigned long long *cur) {
  for (int p = 0; p < kParts; ++p) st_ctrl(&cur[p * kCursorStride], 0ull);
}

// Hands out a pull's 64-vertex chunks (the single-root and the multi-source
// pulls).  Part p holds chunks [nch*p/kParts, nch*(p+1)/kParts).  A
// wavefront draws from its home part first, then from the other parts in
// turn, so the parts finish together whatever their costs.  A grab is one
// chunk in rounds with real scanning (a dense round has only a few dozen
// chunks per wavefront, each costing hundreds of microseconds, and needs
// single chunks, taken only when the previous one is done, to balance) and
// several chunks in a round known to be light, where one atomic per chunk
// would cost more than the chunk; there the next grab is issued before the
// current chunks are worked on, so its latency hides behind theirs.  The
// round's decide kernel picks the size.  Blocks reach the XCDs round-robin,
// so with home = blockIdx.x % kParts each XCD mostly walks one contiguous
// stretch of vertices.  Wavefront-uniform.
struct ChunkFeed {
  unsigned long long *cur;
  int32_t nch;
  int home;
  int32_t want;    // chunks per grab
  int tried = 0;   // parts found drained so far
  bool pending = false;  // a grab is in flight
  int32_t issued = 0;    // lane 0: the cursor value it returned
  int32_t size = 0;      // chunks the grab in flight asked for
  // Static split (a round known to have little to do, where most chunks
  // cost less than the atomic that would hand them out): chunk pos, then
  // pos + stride, ..., one wavefront of the grid each.
  bool fixed;
  int32_t pos = 0, stride = 0;

  __device__ ChunkFeed(unsigned long long *c, int32_t n, int32_t grab, bool static_split)
      : cur(c), nch(n), home((int)(blockIdx.x % kParts)), want(grab), fixed(static_split) {
    if (fixed) {
      pos = (int32_t)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave);
      stride = (int32_t)(((int64_t)gridDim.x * blockDim.x) / kWave);
    }
  }

  __device__ int part() const { return (home + tried) % kParts; }
  __device__ void issue() {
    if (lane_id() == 0)
      issued = (int32_t)atomicAdd(&cur[part() * kCursorStride], (unsigned long long)want);
    size = want;
    pending = true;
  }
  // The next chunks, [b, e); false once every part is drained.
  __device__ bool next(int32_t &b, int32_t &e) {
    if (fixed) {
      if (pos >= nch) return false;
      b = pos;
      e = pos + 1;
      pos += stride;
      return true;
    }
    for (;;) {
      if (tried >= kParts) return false;  // drained (asked again after false)
      if (!pending) issue();
      pending = false;
      const int p = part();
      const int64_t lo = (int64_t)nch * p / kParts, hi = (int64_t)nch * (p + 1) / kParts;
      const int64_t k = lo + __shfl(issued, 0);
      if (k < hi) {
        b = (int32_t)k;
        e = (int32_t)(k + size < hi ? k + size : hi);
        if (want > 1) issue();
        return true;
      }
      if (++tried == kParts) return false;
    }
  }
};

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

// Append queue staged in LDS: wavefronts reserve slots with an LDS atomic,
// and the block publishes its entries with one global atomicAdd in flush().
// Entries past the LDS capacity go straight to global memory (still one
// global atomic per wavefront call).  push() is wavefront-uniform; every
// thread of the block must reach init() and flush().
struct StagedQueue {
  int32_t *buf;
  unsigned *cnt;
  unsigned long long *base;
  unsigned cap;
  unsigned long long *gtail;
  int32_t *gq;

  __device__ void init() {
    if (threadIdx.x == 0) *cnt = 0;
    __syncthreads();
  }

  __device__ void push(bool want, int32_t val) {
    uint64_t m = __ballot(want);
    if (m == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)m) - 1;
    unsigned b = 0;
    if (lane == leader) b = atomicAdd(cnt, (unsigned)__popcll(m));
    b = __shfl(b, leader);
    unsigned my = b + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
    bool spill = want && my >= cap;
    if (want && !spill) buf[my] = val;
    uint64_t sm = __ballot(spill);
    if (sm) {
      const int sl = __ffsll((unsigned long long)sm) - 1;
      unsigned long long gb = 0;
      if (lane == sl) gb = atomicAdd(gtail, (unsigned long long)__popcll(sm));
      gb = __shfl(gb, sl);
      if (spill) gq[gb + __popcll(sm & ((1ull << lane) - 1ull))] = val;
    }
  }

  __device__ void flush() {
    __syncthreads();
    unsigned n = *cnt < cap ? *cnt : cap;
    if (threadIdx.x == 0) *base = n ? atomicAdd(gtail, (unsigned long long)n) : 0ull;
    __syncthreads();
    const unsigned long long b = *base;
    for (unsigned i = threadIdx.x; i < n; i += blockDim.x) gq[b + i] = buf[i];
    __syncthreads();
  }
};

// ---------------------------------------------------------------------------
// Cross-lane scans on DPP (VALU, no LDS round trip)
// ---------------------------------------------------------------------------
// GFX9 DPP controls used by the scans.
constexpr int kDppRowShr = 0x110;     // + k: lane j reads lane j-k of its 16-lane row
constexpr int kDppWaveShr1 = 0x138;   // lane j reads lane j-1
constexpr int kDppRowBcast15 = 0x142; // rows 1, 3 (row mask 0xA) read lane 15 of the row before
constexpr int kDppRowBcast31 = 0x143; // rows 2, 3 (row mask 0xC) read lane 31


template <int kCtrl, int kRowMask = 0xF>
__device__ __forceinline__ int32_t dpp32(int32_t x) {
  return __builtin_amdgcn_update_dpp(0, x, kCtrl, kRowMask, 0xF, true);
}

// Inclusive prefix sum / prefix max of non-negative values across the
// wavefront (0 is the identity of both on non-negative values).
__device__ __forceinline__ int32_t wave_scan_add(int32_t x) {
  x += dpp32<kDppRowShr + 1>(x);
  x += dpp32<kDppRowShr + 2>(x);
  x += dpp32<kDppRowShr + 4>(x);
  x += dpp32<kDppRowShr + 8>(x);
  x += dpp32<kDppRowBcast15, 0xA>(x);
  x += dpp32<kDppRowBcast31, 0xC>(x);
  return x;
}
__device__ __forceinline__ int32_t wave_scan_max(int32_t x) {
  x = max(x, dpp32<kDppRowShr + 1>(x));
  x = max(x, dpp32<kDppRowShr + 2>(x));
  x = max(x, dpp32<kDppRowShr + 4>(x));
  x = max(x, dpp32<kDppRowShr + 8>(x));
  x = max(x, dpp32<kDppRowBcast15, 0xA>(x));
  x = max(x, dpp32<kDppRowBcast31, 0xC>(x));
  return x;
}

// Orders a wavefront's LDS writes before its later LDS reads (no block
// barrier: the callers loop with wavefront-divergent trip counts).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
