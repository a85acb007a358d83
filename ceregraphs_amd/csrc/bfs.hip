// Single-root direction-optimising BFS for gfx950.
//
// Replaces the reference's on-device round loop f_spmv_iter
// (bfs/implementation/src/bool_pe.csl:673-1161).  One round = one BFS level:
//   * push round (compute_topdown, bool_pe.csl:865-887): the vertices of the
//     level-d frontier bitmap are expanded over out-neighbours; a vertex is
//     claimed by atomicOr on the visited bitmap (replacing the row OR-reduce,
//     collectives_2d reduce_or), and every frontier predecessor of a vertex
//     new at this level offers itself with atomicMin, so the parent is the
//     lowest-id predecessor;
//   * pull round (compute_bottomup, bool_pe.csl:899-921): every unvisited
//     vertex scans its ascending in-neighbours and stops at the first one in
//     the level-d frontier bitmap -- the first hit is the lowest id;
//   * the decide step (reduce_done + the termination relay + the switch,
//     bool_pe.csl:947-1156) reads device counters and picks the next
//     direction or ends the search, with no host round trip.
// The frontier lives in HBM as a bitmap that both directions read and
// write, so either direction can follow either.  Frontier bitmaps are
// rotated by level: round d reads F[d%3], writes F[(d+1)%3] and zeroes
// F[(d+2)%3] (the level d-1 bitmap, no longer read), so every bitmap is
// clear one round before it is written.  A level of the per-level loop is
// four launches: bfs_level (the push or the pull), bfs_push_scan and
// bfs_push_big (the push's vertices of out-degree >= big_degree, expanded
// by the whole grid) and bfs_decide.
#include <cstring>
#include <mutex>
#include <vector>

#include "cg_common.h"

namespace cg {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kHistCap = 4096;        // rounds recorded in the history arrays
constexpr int64_t kBigDegree = 256;   // push: whole-grid flat expansion at/above this (default)
constexpr int kBigLds = 4096;         // push: big-list prefix entries staged per pass
// push: a wavefront expands each vertex of out-degree at/above this alone
// (default: none below kBigDegree; smaller rows are laid end to end)
constexpr int64_t kWaveDegree = 256;
constexpr int64_t kSpanCap = (int64_t)1 << 24;  // push: rows laid end to end stay below this
#ifndef CG_PULL_HEAD
#define CG_PULL_HEAD CG_ROW_HEAD
#endif
#ifndef CG_PULL_PASS_STEPS
#define CG_PULL_PASS_STEPS 1
#endif
constexpr int kPullUnroll = CG_PULL_HEAD;  // pull: entries of each row the head probe tests
constexpr int kPassSteps = CG_PULL_PASS_STEPS;  // pull: 64-entry steps per tail pass
constexpr int kPass = kPassSteps * kWave;  // pull: tail entries per pass
#ifndef CG_PULL_GRAB
#define CG_PULL_GRAB 8
#endif
constexpr int32_t kPullGrab = CG_PULL_GRAB;  // pull: chunks per grab from the cursors (default)

// The round counters each take a 128-byte line of their own: every block
// of a round adds to them once at its end, and device-scope atomics on one
// line are serialised, with each other and with the flag reads every block
// of the next kernel starts with.
#ifndef CG_CTRL_LINE
#define CG_CTRL_LINE 128
#endif
#ifndef CG_PARENT_FILTER
#define CG_PARENT_FILTER 1
#endif
constexpr bool kParentFilter = CG_PARENT_FILTER;  // push: read a parent before its atomicMin
constexpr int kBarGroups = 64;     // grid barrier: arrival counters, at most
constexpr int kBarGroupSize = 32;  // grid barrier: blocks per counter (consecutive blockIdx)
static_assert(kBarGroups <= 64, "bfs_init_source clears one barrier counter per thread of 64");
constexpr unsigned kBarSpin = 1u << 22;  // grid barrier: polls before it gives up

struct Ctrl {
  alignas(CG_CTRL_LINE) unsigned long long big_tail;  // push: count of whole-grid vertices
  alignas(CG_CTRL_LINE) unsigned long long nf;        // vertices discovered this round
  alignas(CG_CTRL_LINE) unsigned long long mf;        // out-degree sum of the discovered vertices
  alignas(CG_CTRL_LINE) unsigned long long in_sum;    // in-degree sum of the discovered vertices
  alignas(CG_CTRL_LINE) unsigned long long scanned;   // adjacency entries read this round
  alignas(CG_CTRL_LINE) unsigned long long rows;      // adjacency lists opened this round
  // read by every block of a round, written by the decide
  alignas(CG_CTRL_LINE) int dir;   // direction of the current round: 0 push, 1 pull
  int light; // the current round's pull splits its chunks statically
  int done;
  int status;
  unsigned long long big_edges; // push: the big vertices' total degree
  long long mu;                 // in-edges of still-unvisited vertices
  long long prev_nf;
  long long seen;               // vertices reached so far
  long long rounds;
  unsigned long long t0;
  unsigned long long t_end;
  // grid barrier of the persistent search (grid_barrier)
  LineCtr bar_group[kBarGroups];
  LineCtr bar_top;
  LineCtr bar_gen;
  alignas(CG_CTRL_LINE) int bar_abort;
};

// Per-search constants, passed by value to the kernels that need them.
struct BfsParams {
  int mode;
  double alpha, beta;
  long long tau, max_levels;
  long long n_in;  // vertices with at least one in-edge
};

// Per-round statistics, one row per round (a search's rounds are one copy).
struct HistRow {
  long long nf, scanned, rows;
  unsigned long long t;  // device clock at the round's decide
  int dir;
  int pad;
};
struct Hist {
  HistRow r[kHistCap];
};

struct Workspace {
  int64_t n = 0;
  int64_t nwords = 0;  // u32 words per bitmap, rounded up to 64 vertices
  uint32_t *visited = nullptr;
  uint32_t *fb[3] = {nullptr, nullptr, nullptr};
  int32_t *bigq = nullptr;
  int64_t *bigpre = nullptr;  // exclusive degree prefix of bigq
  int32_t *levels = nullptr;   // internal outputs (host-output calls)
  uint32_t *parents = nullptr;
  Ctrl *ctrl = nullptr;
  Hist *hist = nullptr;
  unsigned long long *scalars = nullptr;  // m / visited reductions
  unsigned long long *cursor = nullptr;   // [kParts * kCursorStride] pull chunk cursors
};

static constexpr size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Size of the single-root workspace of an n-vertex graph.  internal_out
// adds the int32 levels / parents a search writes when the caller's outputs
// are host memory (copied out at the end).
static constexpr size_t ws1_size(int64_t n, bool internal_out) {
  return align_up(sizeof(uint32_t) * (size_t)(((n + 63) / 64) * 2)) * 4 +  // visited, fb[3]
         align_up(sizeof(int32_t) * (size_t)n) +                           // bigq
         align_up(sizeof(int64_t) * (size_t)(n + 1)) +                      // bigpre
         (internal_out ? align_up(sizeof(int32_t) * (size_t)n) * 2 : 0) +   // levels, parents
         align_up(sizeof(Ctrl)) + align_up(sizeof(Hist)) + align_up(64) +
         align_up(sizeof(unsigned long long) * kMaxParts * kCursorStride);
}

// Lays the workspace of ws1_size(n, internal_out) bytes out from `base`.
static void ws1_layout(int64_t n, bool internal_out, char *p, Workspace *w) {
  const int64_t nwords = ((n + 63) / 64) * 2;
  const size_t bm = align_up(sizeof(uint32_t) * nwords);
  w->n = n;
  w->nwords = nwords;
  w->visited = (uint32_t *)p; p += bm;
  for (int i = 0; i < 3; i++) { w->fb[i] = (uint32_t *)p; p += bm; }
  w->bigq = (int32_t *)p; p += align_up(sizeof(int32_t) * n);
  w->bigpre = (int64_t *)p; p += align_up(sizeof(int64_t) * (n + 1));
  if (internal_out) {
    w->levels = (int32_t *)p; p += align_up(sizeof(int32_t) * n);
    w->parents = (uint32_t *)p; p += align_up(sizeof(uint32_t) * n);
  } else {
    w->levels = nullptr;
    w->parents = nullptr;
  }
  w->ctrl = (Ctrl *)p; p += align_up(sizeof(Ctrl));
  w->hist = (Hist *)p; p += align_up(sizeof(Hist));
  w->scalars = (unsigned long long *)p; p += align_up(64);
  w->cursor = (unsigned long long *)p;
}

// The graph's own workspace (cgbfs_bfs), allocated on first use and grown
// when a search needs more.
static int ensure_ws1(DeviceGraph *g, Workspace *w) {
  const size_t total = ws1_size(g->n, true);
  if (g->ws1_bytes < total) {
    if (g->ws1) hipFree(g->ws1);
    g->ws1 = nullptr;
    g->ws1_bytes = 0;
    CG_HIP(hipMalloc(&g->ws1, total));
    g->ws1_bytes = total;
  }
  ws1_layout(g->n, true, (char *)g->ws1, w);
  return CGBFS_OK;
}

// The workspace of cgbfs_bfs_noalloc: one file-scope device array per GPU,
// sized for kNoAllocMaxN vertices and held until the process ends.  Host
// code never uses the array's name as an address (on the host it is only a
// stand-in for the symbol): its device address comes from
// hipGetSymbolAddress, once per device, with that device current.
constexpr int64_t kNoAllocMaxN = (int64_t)1 << 21;
constexpr size_t kNoAllocBytes = ws1_size(kNoAllocMaxN, false);
static_assert(kNoAllocBytes == 26396416, "the size stated for cgbfs_bfs_noalloc in cgbfs.h");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[kMaxDevices];

// The current device's address of `scratch` (the caller made `device`
// current).
static int scratch_address(int device, void **out) {
  if (device < 0 || device >= kMaxDevices) return set_error(CGBFS_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    if (!ptr) return set_error(CGBFS_EHIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_dev[device] = ptr;
  }
  *out = g_scratch_dev[device];
  return CGBFS_OK;
}

// ---------------------------------------------------------------------------
// Block-level helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long block_sum(unsigned long long v,
                                                         unsigned long long *lds) {
  v = wave_sum_u64(v);
  int w = threadIdx.x / kWave;
  if (lane_id() == 0) lds[w] = v;
  __syncthreads();
  unsigned long long s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x / kWave); i++) s += lds[i];
  __syncthreads();
  return s;  // valid in thread 0
}

struct LevelArgs {
  const int64_t *out_ptr;
  const int32_t *out_idx;
  const int64_t *in_ptr;
  const int32_t *in_idx;
  const RowHead *in_head;  // first kRowHead entries of every in-row (-1 past the end)
  // vertices without in-edges, as the visited bitmap's words: a search
  // starts with them marked, so no pull looks at them again
  const uint32_t *noin;
  int32_t *levels;
  uint32_t *parents;
  uint32_t *visited;
  uint32_t *fb[3];
  int32_t *bigq;
  int64_t *bigpre;
  Ctrl *ctrl;
  Hist *hist;
  unsigned long long *cursor;  // pull chunk cursors, zero between rounds
  int64_t n;
  int64_t nwords;
  int symmetric;  // in view == out view: a vertex's out-degree is its in-row length
  int64_t wave_degree;  // push: a wavefront expands vertices of out-degree >= this
  int64_t big_degree;   // push: the whole grid expands vertices of out-degree >= this
  int32_t pull_grab;    // pull: chunks per grab in a dense round
};

// ---------------------------------------------------------------------------
// Init
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) bfs_init_fill(LevelArgs A) {
  int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = tid; v < A.n; v += stride) {
    A.levels[v] = -1;
    A.parents[v] = kParentNone;
  }
  for (int64_t w = tid; w < A.nwords; w += stride) {
    A.visited[w] = A.noin[w];
    A.fb[0][w] = 0;
    A.fb[1][w] = 0;
    A.fb[2][w] = 0;
  }
}

__global__ void bfs_init_source(LevelArgs A, int64_t s, BfsParams P, long long total_in) {
  if (blockIdx.x != 0) return;
  // the grid barrier's counters (one block of 64 threads: kBarGroups <= 64)
  if (threadIdx.x < kBarGroups) st_ctrl(&A.ctrl->bar_group[threadIdx.x].v, 0ull);
  if (threadIdx.x != 0) return;
  st_ctrl(&A.ctrl->bar_top.v, 0ull);
  st_ctrl(&A.ctrl->bar_gen.v, 0ull);
  st_ctrl(&A.ctrl->bar_abort, 0);
  A.levels[s] = 0;
  A.parents[s] = (uint32_t)s;
  A.visited[s >> 5] |= 1u << (s & 31);
  A.fb[0][s >> 5] |= 1u << (s & 31);
  Ctrl *c = A.ctrl;
  st_ctrl(&c->big_tail, 0ull);
  st_ctrl(&c->nf, 0ull);
  st_ctrl(&c->mf, 0ull);
  st_ctrl(&c->in_sum, 0ull);
  st_ctrl(&c->scanned, 0ull);
  st_ctrl(&c->rows, 0ull);
  st_ctrl(&c->mu, (long long)(total_in - (A.in_ptr[s + 1] - A.in_ptr[s])));
  st_ctrl(&c->prev_nf, 1ll);
  st_ctrl(&c->seen, 1ll);
  st_ctrl(&c->light, 0);
  st_ctrl(&c->rounds, 0ll);
  // round 0 is push (bool_pe.csl:714-719) unless every round pulls
  st_ctrl(&c->dir, P.mode == CGBFS_DIR_BOTTOMUP ? 1 : 0);
  st_ctrl(&c->done, 0);
  st_ctrl(&c->status, 0);
  reset_cursors(A.cursor);
}

// The search's start time, before anything else of it runs on the stream;
// every decide then records its own time in t_end.
__global__ void bfs_start(Ctrl *c) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const unsigned long long t0 = wall_clock64();
  st_ctrl(&c->t0, t0);
  st_ctrl(&c->t_end, t0);
}

// ---------------------------------------------------------------------------
// Push round
// ---------------------------------------------------------------------------
// Offers frontier vertex u as predecessor of v in round d: the first to set
// v's visited bit claims it (level, next-frontier bit, counts), and every
// frontier in-neighbour offers itself as v's parent with atomicMin, so the
// lowest id wins.  A push keeps no queue: the next frontier is the bitmap.
__device__ __forceinline__ void push_visit_bm(const LevelArgs &A, int d, uint32_t u, int32_t v,
                                              uint32_t *fnext, unsigned long long &mf,
                                              unsigned long long &in_sum,
                                              unsigned long long &found) {
  const uint32_t w = (uint32_t)v >> 5, b = 1u << (v & 31);
  const uint32_t vw = __builtin_nontemporal_load(&A.visited[w]);
  if (!(vw & b)) {
    const uint32_t old = atomicOr(&A.visited[w], b);
    if (!(old & b)) {
      A.levels[v] = d + 1;
      atomicOr(&fnext[w], b);
      found++;
      mf += (unsigned long long)(A.out_ptr[v + 1] - A.out_ptr[v]);
      in_sum += (unsigned long long)(A.in_ptr[v + 1] - A.in_ptr[v]);
    }
    atomicMin(&A.parents[v], u);
  } else {
    const int lv = __hip_atomic_load(&A.levels[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // a parent already at or below u needs no atomic (parents only fall
    // within a round, so a stale read at worst keeps one)
    if ((lv == d + 1 || lv == -1) && (!kParentFilter || u < A.parents[v])) atomicMin(&A.parents[v], u);
  }
}

// Expands the rows of the calling wavefront's lanes (lane l: frontier
// vertex u, out-edges [beg, beg + len), len < kSpanCap; 0 for none) laid
// end to end: each step every lane takes one edge, finds its row (the last
// row starting at or before it: row owners mark their starts in rm, a
// prefix max fills the gaps, carried from step to step) and calls
// visit(u of that row, head of the edge).  A wavefront thus spends steps on
// its rows' edges, not on its longest row, and its lanes' visits are
// independent of each other.  Wavefront-uniform; rm holds 64 entries.
template <class Visit>
__device__ __forceinline__ void push_rows_packed(const LevelArgs &A, int32_t *rm, uint32_t u,
                                                 int64_t beg, int32_t len, Visit visit) {
  const int lane = lane_id();
  const int32_t incl = wave_scan_add(len);
  const int32_t vstart = incl - len;
  const int32_t total = __builtin_amdgcn_readlane(incl, kWave - 1);
  int rcarry = 0;
  for (int32_t s0 = 0; s0 < total; s0 += kWave) {
    rm[lane] = 0;
    wave_sync();
    const int32_t p = vstart - s0;
    if (len && p >= 0 && p < kWave) rm[p] = lane + 1;
    wave_sync();
    const int r1 = max(wave_scan_max(rm[lane]), rcarry);
    rcarry = __builtin_amdgcn_readlane(r1, kWave - 1);
    const int32_t ve = s0 + lane;
    const bool ok = ve < total;
    const int r = ok ? r1 - 1 : 0;
    const int32_t vs = __shfl(vstart, r);
    const int64_t bs = __shfl(beg, r);
    const uint32_t ur = (uint32_t)__shfl((int32_t)u, r);
    if (ok) visit(ur, A.out_idx[bs + (ve - vs)]);
  }
}

// Degree prefix of the push round's big list (one block): bigpre[j] = sum of
// the degrees of bigq[0..j), bigpre[nb] = big_edges.
__global__ void __launch_bounds__(1024) bfs_push_scan(LevelArgs A) {
  Ctrl *c = A.ctrl;
  // the gate's words loaded together, then tested
  const int done = ld_ctrl(&c->done), dir = ld_ctrl(&c->dir);
  const int64_t nb = (int64_t)ld_ctrl(&c->big_tail);
  if (done || dir != 0 || nb == 0) return;
  __shared__ long long wsum[1024 / kWave];
  __shared__ long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = lane_id(), w = threadIdx.x / kWave;
  for (int64_t b0 = 0; b0 < nb; b0 += blockDim.x) {
    int64_t i = b0 + threadIdx.x;
    long long deg = 0;
    if (i < nb) {
      int32_t u = A.bigq[i];
      deg = (long long)(A.out_ptr[u + 1] - A.out_ptr[u]);
    }
    long long x = deg;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      long long y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == kWave - 1) wsum[w] = x;
    __syncthreads();
    long long before = carry;
    for (int k = 0; k < w; k++) before += wsum[k];
    if (i < nb) A.bigpre[i] = before + x - deg;
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) {
      long long t = 0;
      for (int k = 0; k < (int)(blockDim.x / kWave); k++) t += wsum[k];
      carry += t;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    A.bigpre[nb] = carry;
    st_ctrl(&c->big_edges, (unsigned long long)carry);
  }
}

// Whole-grid flat expansion of the big list: the big vertices' edges are
// numbered 0..big_edges-1 in list order and every thread takes flat indices
// gtid, gtid + stride, ...; the owning vertex comes from a binary search
// over the degree prefix staged in LDS (kBigLds entries per pass).
__global__ void __launch_bounds__(kBlock) bfs_push_big(LevelArgs A) {
  Ctrl *c = A.ctrl;
  const int done = ld_ctrl(&c->done), dir = ld_ctrl(&c->dir);
  const int d = (int)ld_ctrl(&c->rounds);
  const int64_t nb = (int64_t)ld_ctrl(&c->big_tail);
  if (done || dir != 0 || nb == 0) return;
  __shared__ unsigned long long red[kWavesPerBlock];
  __shared__ long long pre[kBigLds + 1];
  unsigned long long mf = 0, in_sum = 0, scanned = 0, rows = 0, found = 0;
  uint32_t *fnext = A.fb[(d + 1) % 3];
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j0 = 0; j0 < nb; j0 += kBigLds) {
    const int cnt = (int)((nb - j0) < kBigLds ? (nb - j0) : kBigLds);
    for (int k = threadIdx.x; k <= cnt; k += blockDim.x) pre[k] = A.bigpre[j0 + k];
    __syncthreads();
    const long long lo = pre[0], hi = pre[cnt];
    for (long long f = lo + gtid; f < hi; f += gstride) {
      int a = 0, b = cnt - 1;  // largest j with pre[j] <= f
      while (a < b) {
        int mid = (a + b + 1) >> 1;
        if (pre[mid] <= f) a = mid;
        else b = mid - 1;
      }
      int32_t u = A.bigq[j0 + a];
      int64_t e = A.out_ptr[u] + (f - pre[a]);
      push_visit_bm(A, d, (uint32_t)u, A.out_idx[e], fnext, mf, in_sum, found);
    }
    if (gtid == 0) {
      scanned += (unsigned long long)(hi - lo);
      rows += (unsigned long long)cnt;
    }
    __syncthreads();
  }
  unsigned long long s;
  s = block_sum(found, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->nf, s);
  s = block_sum(mf, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->mf, s);
  s = block_sum(in_sum, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->in_sum, s);
  s = block_sum(scanned, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->scanned, s);
  s = block_sum(rows, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->rows, s);
}

// ---------------------------------------------------------------------------
// Pull round
// ---------------------------------------------------------------------------
// What a chunk needs before its rows can be scanned: its two visited words
// and, in a dense round, the row offsets and row heads (first kRowHead entries)
// of all 64 vertices.  Loaded one chunk ahead, so those loads are in flight
// while the chunk before is being scanned, and a chunk's scan starts with
// the frontier bits of its row heads.
struct PullPre {
  int32_t ch;       // chunk, -1 for none
  uint32_t vlo, vhi;
  int64_t beg, end;
  RowHead hd;
};

// A dense round (most vertices unvisited) loads every lane's offsets and
// row head with the visited words, so the loads overlap; a light round
// (ctrl->light) loads them only for unvisited vertices, once their bits are
// known, since almost every vertex is visited there.
__device__ __forceinline__ void pull_prefetch(const LevelArgs &A, int32_t ch, bool light,
                                              PullPre &p) {
  p.ch = ch;
  if (ch < 0) return;
  const int64_t w0 = (int64_t)ch * 2;
  p.vlo = A.visited[w0];
  p.vhi = A.visited[w0 + 1];
  if (!light) {
    const int64_t v = (int64_t)ch * kWave + lane_id();
    const int64_t vv = v < A.n ? v : A.n - 1;
    p.beg = A.in_ptr[vv];
    p.end = A.in_ptr[vv + 1];
    p.hd = A.in_head[vv];
  }
}

// One pull round over the whole graph by the calling grid (every thread of
// every block must call it): a pull level of bfs_level or of the persistent
// search.
__device__ void pull_level(const LevelArgs &A, int d) {
  Ctrl *c = A.ctrl;
  __shared__ unsigned long long red[kWavesPerBlock];
  const uint32_t *fcur = A.fb[d % 3];
  uint32_t *fnext = A.fb[(d + 1) % 3];
  uint32_t *fclear = A.fb[(d + 2) % 3];
  int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t w = gtid; w < A.nwords; w += gstride) fclear[w] = 0;

  const int lane = lane_id();
  // row map of the current tail pass, [step][position], one per wavefront
  __shared__ int32_t rmap[kBlock / kWave][kPassSteps][kWave];
  int32_t (*rm)[kWave] = rmap[threadIdx.x / kWave];
  const int64_t nchunks = A.nwords / 2;
  const bool light = ld_ctrl(&c->light) != 0;
  const bool sym = A.symmetric != 0;
  unsigned long long mf = 0, in_sum = 0, scanned = 0, rows = 0, found = 0;
  // One chunk of 64 vertices (wavefront-uniform).
  auto chunk = [&](const PullPre &p) {
    const int64_t ch = p.ch;
    const int64_t w0 = ch * 2;
    const uint32_t vlo = p.vlo, vhi = p.vhi;
    const int64_t v = ch * kWave + lane;
    uint32_t myword = (lane < 32) ? vlo : vhi;
    bool todo = v < A.n && !((myword >> (lane & 31)) & 1u);
    if (!__any(todo)) return;
    int64_t beg, end;
    RowHead hd;
    if (light) {
      beg = todo ? A.in_ptr[v] : 0;
      end = todo ? A.in_ptr[v + 1] : 0;
      hd = A.in_head[todo ? v : 0];
    } else {
      beg = todo ? p.beg : 0;
      end = todo ? p.end : 0;
      hd = p.hd;
    }
    bool hit = false;
    int32_t par = -1;
    if (todo) rows++;
    // Head probe: every row's first kPullUnroll entries at once, one lane
    // per row (their frontier-bit loads in flight together); the first hit
    // in row order is the lowest id.  Most rows of a dense level end here.
    // The first kRowHead entries come from the row heads, loaded in order with
    // the offsets; entries past the row end are not tested.
    const int64_t hend = end - beg < kPullUnroll ? end : beg + kPullUnroll;
    if (todo && beg < end) {
      int32_t u[kPullUnroll];
      uint32_t fw[kPullUnroll];
#pragma unroll
      for (int k = 0; k < kPullUnroll; k++) {
        const int32_t h = k < kRowHead ? hd.id[k < kRowHead ? k : 0] : A.in_idx[beg + k < hend ? beg + k : beg];
        u[k] = beg + k < hend ? h : hd.id[0];
      }
#pragma unroll
      for (int k = 0; k < kPullUnroll; k++) fw[k] = fcur[u[k] >> 5];
#pragma unroll
      for (int k = 0; k < kPullUnroll; k++) {
        if (!hit && beg + k < hend) {
          scanned++;
          if ((fw[k] >> (u[k] & 31)) & 1u) {
            hit = true;
            par = u[k];
          }
        }
      }
    }
    // Tails of the rows the probe left open, laid end to end and scanned
    // kPass entries per pass by the whole wavefront, so no lane idles while
    // another walks a long row.  Each pass packs what is still open (rows
    // settled in the last pass drop out), maps every position to its row
    // (row owners mark where their stretch starts; a prefix max fills the
    // rest), loads ids and frontier bits branch-free, and each row's owner
    // takes the first hit of its stretch from the pass's ballot masks: the
    // stretch is the row's next entries in order, so that hit is its
    // lowest-id parent if it has one there.
    int64_t pos = hend;                                      // my row's next entry
    int64_t rem = (todo && !hit) ? end - hend : 0;           // its entries left
    while (__any(rem > 0)) {
      const int32_t len = (int32_t)(rem < kPass ? rem : kPass);  // what a pass can take
      const int32_t incl = wave_scan_add(len);
      const int32_t vstart = incl - len;
      const int32_t total = __builtin_amdgcn_readlane(incl, kWave - 1);
#pragma unroll
      for (int k = 0; k < kPassSteps; k++) rm[k][lane] = 0;
      wave_sync();
      if (len > 0 && vstart < kPass) rm[vstart / kWave][vstart % kWave] = lane + 1;
      wave_sync();
      int32_t u[kPassSteps];
      bool ok[kPassSteps];
      int carry = 0;
#pragma unroll
      for (int k = 0; k < kPassSteps; k++) {
        const int r1 = max(wave_scan_max(rm[k][lane]), carry);  // row + 1 (0: none yet)
        carry = __builtin_amdgcn_readlane(r1, kWave - 1);
        const int32_t ve = k * kWave + lane;
        ok[k] = ve < total;
        const int r = ok[k] ? r1 - 1 : 0;
        const int32_t vs = __shfl(vstart, r);
        const int64_t ps = __shfl(pos, r);
        u[k] = A.in_idx[ok[k] ? ps + (ve - vs) : 0];
      }
      uint64_t H[kPassSteps];
#pragma unroll
      for (int k = 0; k < kPassSteps; k++) {
        const uint32_t w = fcur[u[k] >> 5];
        H[k] = ballot64(ok[k] && ((w >> (u[k] & 31)) & 1u));
      }
      // the owner's stretch in this pass: positions [vstart, vstop)
      const bool mine = len > 0 && vstart < kPass;
      const int32_t vstop = vstart + len < kPass ? vstart + len : kPass;
      int32_t first = -1;  // the stretch's first hit position
#pragma unroll
      for (int k = 0; k < kPassSteps; k++) {
        const int32_t lo = max(vstart, k * kWave) - k * kWave, hi = min(vstop, (k + 1) * kWave) - k * kWave;
        if (first < 0 && mine && lo < hi) {
          const uint64_t inside = (hi == kWave ? ~0ull : ((1ull << hi) - 1ull)) & ~((1ull << lo) - 1ull);
          const uint64_t h = H[k] & inside;
          if (h) first = k * kWave + __ffsll((unsigned long long)h) - 1;
        }
      }
      // the id at the hit position, read from the lane that loaded it
      const int src = first >= 0 ? first % kWave : lane;
      int32_t hid = 0;
#pragma unroll
      for (int k = 0; k < kPassSteps; k++) {
        const int32_t x = __shfl(u[k], src);
        if (first >= 0 && first / kWave == k) hid = x;
      }
      if (mine) {
        if (first >= 0) {
          hit = true;
          par = hid;
          scanned += (unsigned long long)(first - vstart + 1);
          rem = 0;
        } else {
          scanned += (unsigned long long)(vstop - vstart);
          pos += vstop - vstart;
          rem -= vstop - vstart;
        }
      }
    }
    if (hit) {
      A.parents[v] = (uint32_t)par;
      A.levels[v] = d + 1;
      mf += (unsigned long long)(sym ? end - beg : A.out_ptr[v + 1] - A.out_ptr[v]);
      in_sum += (unsigned long long)(end - beg);
      found++;
    }
    // The new frontier is the bitmap alone, which a push reads directly.
    // Appending every found vertex to a queue here made the dense level wait
    // on one shared tail counter.
    uint64_t hm = ballot64(hit);
    if (hm && lane == 0) {
      uint32_t lo = (uint32_t)hm, hi = (uint32_t)(hm >> 32);
      if (lo) { fnext[w0] = lo; A.visited[w0] = vlo | lo; }
      if (hi) { fnext[w0 + 1] = hi; A.visited[w0 + 1] = vhi | hi; }
    }
  };
  // Chunks come from the cursors: their costs differ by orders of magnitude
  // (rows scanned to the end, hub rows, chunks already visited), which a
  // static split over the wavefronts leaves unbalanced.  Once few vertices
  // with in-edges are unreached, nearly every chunk is visited and costs
  // less than the atomic that would hand it out: a static split then.
  // The chunk after the current one is prefetched before the current one is
  // scanned.
  ChunkFeed feed(A.cursor, (int32_t)nchunks, A.pull_grab, light);
  int32_t cb = 0, ce = 0;
  auto next_chunk = [&]() -> int32_t { return (cb < ce || feed.next(cb, ce)) ? cb++ : -1; };
  PullPre cur, nxt;
  pull_prefetch(A, next_chunk(), light, cur);
  while (cur.ch >= 0) {
    pull_prefetch(A, next_chunk(), light, nxt);
    chunk(cur);
    cur = nxt;
  }
  unsigned long long s;
  s = block_sum(found, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->nf, s);
  s = block_sum(mf, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->mf, s);
  s = block_sum(in_sum, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->in_sum, s);
  s = block_sum(scanned, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->scanned, s);
  s = block_sum(rows, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->rows, s);
}


// ---------------------------------------------------------------------------
// Decide: termination, direction, history (one thread)
// ---------------------------------------------------------------------------
// The end of round d (one thread): records the round, resets the round
// counters and chooses the next round's direction, or ends the search.
// nf is the number of vertices the round reached.
// The next round's direction (0 push, 1 pull) after a round in direction cur
// that found nf vertices of out-degree sum mf, with mu in-edges of unvisited
// vertices left and prev_nf found by the round before.
__device__ __forceinline__ int next_direction(const BfsParams &P, int cur, long long nf,
                                              long long mf, long long mu, long long prev_nf,
                                              int64_t n) {
  switch (P.mode) {
    case CGBFS_DIR_TOPDOWN: return 0;
    case CGBFS_DIR_BOTTOMUP: return 1;
    case CGBFS_DIR_TAU:
      return (cur == 0 && nf > P.tau) ? 1 : cur;  // one-way, bool_pe.csl:1078-1088
    default:  // Beamer et al. alpha/beta
      if (cur == 0) return (double)mf > (double)mu / P.alpha ? 1 : 0;
      return ((double)nf < (double)n / P.beta && nf < prev_nf) ? 0 : 1;
  }
}

// Every control word the decide reads, loaded in one batch: the loads are
// independent, so they are in flight together instead of one after another.
__device__ void decide_round(const LevelArgs &A, int d, int64_t n, const BfsParams &P) {
  Ctrl *c = A.ctrl;
  const long long nf = (long long)ld_ctrl(&c->nf);
  const int cur = ld_ctrl(&c->dir);
  const long long mf = (long long)ld_ctrl(&c->mf);
  const long long mu0 = ld_ctrl(&c->mu);
  const long long in_sum = (long long)ld_ctrl(&c->in_sum);
  const long long scanned = (long long)ld_ctrl(&c->scanned);
  const long long rows = (long long)ld_ctrl(&c->rows);
  const long long prev_nf = ld_ctrl(&c->prev_nf);
  const long long seen0 = ld_ctrl(&c->seen);
  const long long mu = mu0 - in_sum;
  unsigned long long now = wall_clock64();
  if (d < kHistCap) {
    HistRow &H = A.hist->r[d];
    H.nf = nf;
    H.scanned = scanned;
    H.rows = rows;
    H.dir = cur;
    H.t = now;
  }
  st_ctrl(&c->rounds, (long long)(d + 1));
  st_ctrl(&c->t_end, now);
  st_ctrl(&c->big_tail, 0ull);
  st_ctrl(&c->mu, mu);
  st_ctrl(&c->nf, 0ull);
  st_ctrl(&c->mf, 0ull);
  st_ctrl(&c->in_sum, 0ull);
  st_ctrl(&c->scanned, 0ull);
  st_ctrl(&c->rows, 0ull);
  reset_cursors(A.cursor);
  if (nf == 0) {
    st_ctrl(&c->done, 1);
    return;
  }
  const int dir = next_direction(P, cur, nf, mf, mu, prev_nf, n);
  st_ctrl(&c->prev_nf, nf);
  // The next pull visits every chunk; once fewer vertices with in-edges are
  // unreached than there are chunks, most chunks have nothing to do.
  const long long seen = seen0 + nf;
  st_ctrl(&c->seen, seen);
  st_ctrl(&c->light, P.n_in - seen < (long long)(A.nwords / 2) ? 1 : 0);
  st_ctrl(&c->dir, dir);
  if (d + 1 >= P.max_levels) {
    st_ctrl(&c->status, (int)CGBFS_ETIMEOUT);
    st_ctrl(&c->done, 1);
  }
}

// The round kernels read their round from the control block (the previous
// decide advanced it), so one launch sequence serves every round and can be
// captured once into a HIP graph.
__global__ void bfs_decide(LevelArgs A, int64_t n, BfsParams P) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Ctrl *c = A.ctrl;
  const int done = ld_ctrl(&c->done);
  const int d = (int)ld_ctrl(&c->rounds);
  if (done) return;
  decide_round(A, d, n, P);
}

// ---------------------------------------------------------------------------
// Whole search in one workgroup (graphs of at most kTinyMaxN vertices)
// ---------------------------------------------------------------------------
// On graphs this small a level is a few hundred row entries, and any
// grid-wide step -- a kernel boundary or a grid barrier -- costs more than
// the level itself.  One workgroup of kTinyBlock threads runs every level:
// the visited and frontier bitmaps, the parents (the atomicMin target) and
// the levels stay in LDS, a level ends at a workgroup barrier, thread 0
// decides with bfs_decide's rules (next_direction), and the outputs are
// written once at the end.  A push lays each 64-vertex chunk's frontier rows
// end to end (push_rows_packed); a pull scans each chunk's open rows in
// packed 64-entry passes, the first hit in a row's stretch being its
// lowest-id frontier in-neighbour.
constexpr int kTinyBlock = 1024;
constexpr int kTinyWaves = kTinyBlock / kWave;
constexpr int64_t kTinyMaxN = 4096;
constexpr int kTinyWords = (int)(kTinyMaxN / 32);

template <class T>
__device__ __forceinline__ T lds_load(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ void lds_store(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__global__ void __launch_bounds__(kTinyBlock) bfs_tiny(LevelArgs A, int64_t s, BfsParams P,
                                                       long long total_in) {
  __shared__ uint32_t vis[kTinyWords], fa[kTinyWords], fz[kTinyWords];
  __shared__ uint32_t par[kTinyMaxN];
  __shared__ int32_t lev[kTinyMaxN];
  __shared__ int32_t rmap[kTinyWaves][kWave];
  __shared__ unsigned long long cnt[5];  // found, mf, in_sum, scanned, rows
  __shared__ int s_dir, s_done;
  Ctrl *c = A.ctrl;
  const int64_t n = A.n;
  const int tid = threadIdx.x, lane = lane_id();
  int32_t *rm = rmap[tid / kWave];
  const int nwords = (int)((n + 31) / 32);
  const int nchunks = (int)((n + 63) / 64);
  const bool sym = A.symmetric != 0;
  for (int w = tid; w < kTinyWords; w += kTinyBlock) {
    vis[w] = 0;
    fa[w] = 0;
    fz[w] = 0;
  }
  for (int v = tid; v < n; v += kTinyBlock) {
    par[v] = kParentNone;
    lev[v] = -1;
  }
  __syncthreads();
  long long mu = 0, prev_nf = 1;  // thread 0's decide state
  if (tid == 0) {
    vis[s >> 5] |= 1u << (s & 31);
    fa[s >> 5] |= 1u << (s & 31);
    par[s] = (uint32_t)s;
    lev[s] = 0;
    mu = total_in - (long long)(A.in_ptr[s + 1] - A.in_ptr[s]);
    s_dir = P.mode == CGBFS_DIR_BOTTOMUP ? 1 : 0;  // round 0 pushes (bool_pe.csl:714-719)
    s_done = 0;
    st_ctrl(&c->status, 0);
    st_ctrl(&c->done, 0);
    st_ctrl(&c->rounds, 0ll);
  }
  uint32_t *cur = fa, *nxt = fz;
  __syncthreads();
  for (long long d = 0; d <= P.max_levels; d++) {
    if (tid < 5) cnt[tid] = 0;
    for (int w = tid; w < nwords; w += kTinyBlock) nxt[w] = 0;
    __syncthreads();
    const int dir = s_dir;
    uint32_t found = 0, mf = 0, in_sum = 0, scanned = 0, rows = 0;
    for (int ch = tid / kWave; ch < nchunks; ch += kTinyWaves) {
      const int64_t v = (int64_t)ch * kWave + lane;
      if (dir == 0) {
        const bool on = v < n && ((cur[v >> 5] >> (v & 31)) & 1u);
        if (!__any(on)) continue;
        const int64_t beg = on ? A.out_ptr[v] : 0, end = on ? A.out_ptr[v + 1] : 0;
        if (on) {
          rows++;
          scanned += (uint32_t)(end - beg);
        }
        push_rows_packed(A, rm, (uint32_t)v, beg, on ? (int32_t)(end - beg) : 0,
                         [&](uint32_t u, int32_t w) {
          const uint32_t b = 1u << (w & 31);
          if (!(lds_load(&vis[w >> 5]) & b)) {
            const uint32_t old = atomicOr(&vis[w >> 5], b);
            if (!(old & b)) {
              lds_store(&lev[w], (int32_t)(d + 1));
              atomicOr(&nxt[w >> 5], b);
              found++;
              mf += (uint32_t)(A.out_ptr[w + 1] - A.out_ptr[w]);
              in_sum += (uint32_t)(A.in_ptr[w + 1] - A.in_ptr[w]);
            }
            atomicMin(&par[w], u);
          } else {
            // claimed this round (level d + 1, or not yet written): offer u
            const int32_t lv = lds_load(&lev[w]);
            if (lv == d + 1 || lv == -1) atomicMin(&par[w], u);
          }
        });
      } else {
        // the chunk's visited words are this wavefront's alone in a pull
        const bool todo = v < n && !((vis[v >> 5] >> (v & 31)) & 1u);
        if (!__any(todo)) continue;
        const int64_t beg = todo ? A.in_ptr[v] : 0, end = todo ? A.in_ptr[v + 1] : 0;
        if (todo) rows++;
        int64_t pos = beg, rem = end - beg;
        bool hit = false;
        uint32_t pu = 0;
        while (__any(rem > 0)) {
          const int32_t len = (int32_t)(rem < kWave ? rem : kWave);
          const int32_t incl = wave_scan_add(len);
          const int32_t vstart = incl - len;
          const int32_t total = __builtin_amdgcn_readlane(incl, kWave - 1);
          rm[lane] = 0;
          wave_sync();
          if (len > 0 && vstart < kWave) rm[vstart] = lane + 1;
          wave_sync();
          const int r1 = wave_scan_max(rm[lane]);  // row + 1 of each position (0: none)
          const bool ok = lane < total;
          const int r = ok ? r1 - 1 : 0;
          const int32_t vs = __shfl(vstart, r);
          const int64_t ps = __shfl(pos, r);
          const int32_t u = A.in_idx[ok ? ps + (lane - vs) : 0];
          const uint64_t H = ballot64(ok && ((cur[u >> 5] >> (u & 31)) & 1u));
          const bool mine = len > 0 && vstart < kWave;
          const int32_t vstop = vstart + len < kWave ? vstart + len : kWave;
          int32_t first = -1;
          if (mine) {
            const uint64_t inside = (vstop == kWave ? ~0ull : ((1ull << vstop) - 1ull)) &
                                    ~((1ull << vstart) - 1ull);
            const uint64_t h = H & inside;
            if (h) first = __ffsll((unsigned long long)h) - 1;
          }
          const int32_t hid = __shfl(u, first >= 0 ? first : lane);
          if (mine) {
            if (first >= 0) {
              hit = true;
              pu = (uint32_t)hid;
              scanned += (uint32_t)(first - vstart + 1);
              rem = 0;
            } else {
              scanned += (uint32_t)(vstop - vstart);
              pos += vstop - vstart;
              rem -= vstop - vstart;
            }
          }
        }
        if (hit) {
          par[v] = pu;
          lev[v] = (int32_t)(d + 1);
          found++;
          mf += (uint32_t)(sym ? end - beg : A.out_ptr[v + 1] - A.out_ptr[v]);
          in_sum += (uint32_t)(end - beg);
        }
        const uint64_t hm = ballot64(hit);
        if (lane == 0 && hm) {
          const int w0 = ch * 2;
          const uint32_t lo = (uint32_t)hm, hi = (uint32_t)(hm >> 32);
          if (lo) { nxt[w0] = lo; vis[w0] |= lo; }
          if (hi) { nxt[w0 + 1] = hi; vis[w0 + 1] |= hi; }
        }
      }
    }
    // the round's counts: wavefront sums (32 bits: a tiny graph's round
    // reads fewer than 2^31 entries), one LDS atomic per wavefront each
    const uint32_t t[5] = {found, mf, in_sum, scanned, rows};
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const int32_t tot = __builtin_amdgcn_readlane(wave_scan_add((int32_t)t[k]), kWave - 1);
      if (lane == 0 && tot) atomicAdd(&cnt[k], (unsigned long long)(uint32_t)tot);
    }
    __syncthreads();
    if (tid == 0) {
      const long long nf = (long long)cnt[0];
      mu -= (long long)cnt[2];
      const unsigned long long now = wall_clock64();
      if (d < kHistCap) {
        HistRow &H = A.hist->r[d];
        H.nf = nf;
        H.scanned = (long long)cnt[3];
        H.rows = (long long)cnt[4];
        H.dir = dir;
        H.t = now;
      }
      st_ctrl(&c->rounds, (long long)(d + 1));
      st_ctrl(&c->t_end, now);
      int done = 0;
      if (nf == 0) {
        done = 1;
      } else {
        s_dir = next_direction(P, dir, nf, (long long)cnt[1], mu, prev_nf, n);
        prev_nf = nf;
        if (d + 1 >= P.max_levels) {
          st_ctrl(&c->status, (int)CGBFS_ETIMEOUT);
          done = 1;
        }
      }
      s_done = done;
    }
    __syncthreads();
    if (s_done) break;
    uint32_t *t2 = cur;
    cur = nxt;
    nxt = t2;
  }
  for (int v = tid; v < n; v += kTinyBlock) {
    A.levels[v] = lev[v];
    A.parents[v] = par[v];
  }
  if (tid == 0) st_ctrl(&c->done, 1);
}

// ---------------------------------------------------------------------------
// Persistent search: every level in one cooperative launch
// ---------------------------------------------------------------------------

// One push round read from the frontier bitmap by the calling grid (every
// thread of every block must call it): wavefronts take 64-vertex chunks of
// the bitmap; the rows of a chunk's frontier vertices of out-degree below
// wave_degree are laid end to end (push_rows_packed), a larger one is
// expanded by the whole wavefront, 64 out-neighbours a step, and those of
// out-degree >= big_degree by the whole grid after the next barrier.
__device__ void push_level_bitmap(const LevelArgs &A, int d) {
  Ctrl *c = A.ctrl;
  __shared__ unsigned long long red[kWavesPerBlock];
  const uint32_t *fcur = A.fb[d % 3];
  uint32_t *fnext = A.fb[(d + 1) % 3];
  uint32_t *fclear = A.fb[(d + 2) % 3];
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t w = gtid; w < A.nwords; w += gstride) fclear[w] = 0;
  const int lane = lane_id();
  const int64_t wave = gtid / kWave, nwaves = gstride / kWave;
  const int64_t nchunks = A.nwords / 2;
  __shared__ int32_t rmap[kWavesPerBlock][kWave];
  int32_t *rm = rmap[threadIdx.x / kWave];
  unsigned long long mf = 0, in_sum = 0, scanned = 0, rows = 0, found = 0;
  for (int64_t ch = wave; ch < nchunks; ch += nwaves) {
    const uint32_t word = fcur[ch * 2 + (lane >> 5)];
    const int64_t v = ch * kWave + lane;
    const bool on = v < A.n && ((word >> (lane & 31)) & 1u);
    if (!__any(on)) continue;
    const int64_t beg = on ? A.out_ptr[v] : 0, end = on ? A.out_ptr[v + 1] : 0;
    const bool big = on && end - beg >= A.big_degree;
    const bool small = on && !big && end - beg < A.wave_degree && end - beg < kSpanCap;
    if (on && !big) {
      rows++;
      scanned += (unsigned long long)(end - beg);
    }
    // the whole grid expands the big ones after the next barrier
    const int64_t slot = wave_append(&c->big_tail, big);
    if (big) A.bigq[slot] = (int32_t)v;
    push_rows_packed(A, rm, (uint32_t)v, beg, small ? (int32_t)(end - beg) : 0,
                     [&](uint32_t pu, int32_t w) { push_visit_bm(A, d, pu, w, fnext, mf, in_sum, found); });
    uint64_t wide = ballot64(on && !small && !big);
    while (wide) {
      const int j = __ffsll((unsigned long long)wide) - 1;
      wide &= wide - 1;
      const int64_t bj = __shfl(beg, j), ej = __shfl(end, j);
      const uint32_t uj = (uint32_t)(ch * kWave + j);
      for (int64_t e = bj + lane; e < ej; e += kWave)
        push_visit_bm(A, d, uj, A.out_idx[e], fnext, mf, in_sum, found);
    }
  }
  unsigned long long s;
  s = block_sum(found, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->nf, s);
  s = block_sum(mf, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->mf, s);
  s = block_sum(in_sum, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->in_sum, s);
  s = block_sum(scanned, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->scanned, s);
  s = block_sum(rows, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->rows, s);
}

// The big frontier vertices a bitmap push listed (out-degree >= big_degree),
// expanded by the whole grid one after another, every thread taking a
// strided share of each row.  Every thread of every block must call it.
__device__ void push_big_rows(const LevelArgs &A, int d) {
  Ctrl *c = A.ctrl;
  __shared__ unsigned long long red[kWavesPerBlock];
  uint32_t *fnext = A.fb[(d + 1) % 3];
  const int64_t nb = (int64_t)ld_ctrl(&c->big_tail);
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gstride = (int64_t)gridDim.x * blockDim.x;
  unsigned long long mf = 0, in_sum = 0, scanned = 0, found = 0;
  for (int64_t j = 0; j < nb; j++) {
    const int32_t u = A.bigq[j];
    const int64_t beg = A.out_ptr[u], end = A.out_ptr[u + 1];
    for (int64_t e = beg + gtid; e < end; e += gstride)
      push_visit_bm(A, d, (uint32_t)u, A.out_idx[e], fnext, mf, in_sum, found);
    if (gtid == 0) scanned += (unsigned long long)(end - beg);
  }
  unsigned long long s;
  s = block_sum(found, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->nf, s);
  s = block_sum(mf, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->mf, s);
  s = block_sum(in_sum, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->in_sum, s);
  s = block_sum(scanned, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->scanned, s);
  if (gtid == 0 && nb) atomicAdd(&c->rows, (unsigned long long)nb);
}

// One level of the per-level loop: the push (read from the frontier bitmap;
// out-degree >= big_degree left to bfs_push_big) or the pull, as the
// previous decide chose.  One launch serves both, so a level costs one
// dispatch less than a kernel per direction.
__global__ void __launch_bounds__(kBlock, 8) bfs_level(LevelArgs A) {
  const Ctrl *c = A.ctrl;
  const int done = ld_ctrl(&c->done), dir = ld_ctrl(&c->dir);
  const int d = (int)ld_ctrl(&c->rounds);
  if (done) return;
  if (dir == 0) push_level_bitmap(A, d);
  else pull_level(A, d);
}

// Grid-wide barrier of the persistent search (every block is resident: the
// launch is cooperative).  A block arrives at the counter of its group
// (kBarGroupSize consecutive blocks, each counter on its own line, so
// arrivals serialise within a group only); the last arrival of a group
// arrives at the top counter (none when the grid is one group), and the
// last there advances the generation that every block polls.  The last
// arrivals reset the counters, so barriers follow one another without
// clearing.
// Memory order: each arrival releases the block's writes (its threads'
// stores reach L2 at the workgroup barrier; the release writes L2 back);
// the last arrivals acquire what the others released before passing it on,
// and every block acquires once it sees the new generation.  Waiting blocks
// poll with relaxed loads and sleep between polls: an acquire per poll
// would invalidate the caches under the blocks still working.
// A barrier that sees no progress for kBarSpin polls, or finds that another
// block gave up, sets bar_abort, done and an error status and returns
// false; the caller leaves its level loop, so a fault ends the search with
// CGBFS_EHIP instead of hanging it.  gen: the generation the caller last
// passed (every thread keeps the same copy).
__device__ bool grid_barrier(Ctrl *c, unsigned long long &gen) {
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned nb = gridDim.x;
    const unsigned g = blockIdx.x / kBarGroupSize;
    const unsigned ng = (nb + kBarGroupSize - 1) / kBarGroupSize;
    const unsigned rest = nb - g * kBarGroupSize;
    const unsigned long long gsize = rest < (unsigned)kBarGroupSize ? rest : kBarGroupSize;
    int ok = 1;
    if (__hip_atomic_fetch_add(&c->bar_group[g].v, 1ull, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT) == gsize - 1) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __hip_atomic_store(&c->bar_group[g].v, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (ng == 1 || __hip_atomic_fetch_add(&c->bar_top.v, 1ull, __ATOMIC_RELEASE,
                                            __HIP_MEMORY_SCOPE_AGENT) == ng - 1) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __hip_atomic_store(&c->bar_top.v, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&c->bar_gen.v, gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    unsigned spins = 0;
    while (__hip_atomic_load(&c->bar_gen.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
      if (ld_ctrl(&c->bar_abort) || ++spins > kBarSpin) {
        st_ctrl(&c->bar_abort, 1);
        st_ctrl(&c->status, (int)CGBFS_EHIP);
        st_ctrl(&c->done, 1);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(8);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    s_ok = ok;
  }
  __syncthreads();
  gen++;
  return s_ok != 0;
}

// The whole level loop in one launch (cooperative: every block is resident,
// so grid barriers are safe).  Each level is one round by the whole grid,
// a grid barrier, the decide by one thread, and a second barrier, after
// which every block reads the same done flag and direction: all blocks run
// the same number of levels and reach every barrier.  The loop is bounded
// by max_levels, and the decide ends the search there at the latest.  Pull
// rounds are pull_level; push rounds read the frontier bitmap, so no queue
// is kept, and expand their big vertices in a second phase behind one more
// barrier (taken by every block: the direction is uniform).  For small graphs, where the per-round launches of the level
// kernels cost more than the rounds themselves.
__global__ void __launch_bounds__(kBlock, 8) bfs_persistent(LevelArgs A, int64_t n, BfsParams P) {
  Ctrl *c = A.ctrl;
  unsigned long long gen = ld_ctrl(&c->bar_gen.v);  // no block passes a barrier before all read it
  for (long long d = 0; d <= P.max_levels; d++) {
    if (ld_ctrl(&c->done)) break;
    if (ld_ctrl(&c->dir) == 0) {  // uniform: every block read the same direction
      push_level_bitmap(A, (int)d);
      if (!grid_barrier(c, gen)) break;
      push_big_rows(A, (int)d);
    } else {
      pull_level(A, (int)d);
    }
    if (!grid_barrier(c, gen)) break;
    if (blockIdx.x == 0 && threadIdx.x == 0) decide_round(A, (int)d, n, P);
    if (!grid_barrier(c, gen)) break;
  }
}

// m (bfs_timing.py:390-395) and the visited count, reduced over levels >= 0.
__global__ void __launch_bounds__(kBlock) bfs_count(const int32_t *levels, const int32_t *w,
                                                    int64_t n, unsigned long long *out) {
  __shared__ unsigned long long red[kWavesPerBlock];
  unsigned long long m = 0, cnt = 0;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (levels[v] >= 0) {
      m += (unsigned long long)w[v];
      cnt++;
    }
  }
  unsigned long long s = block_sum(m, red);
  if (threadIdx.x == 0 && s) atomicAdd(&out[0], s);
  s = block_sum(cnt, red);
  if (threadIdx.x == 0 && s) atomicAdd(&out[1], s);
}

__global__ void zero_u64(unsigned long long *p, int k) {
  if (threadIdx.x < (unsigned)k) p[threadIdx.x] = 0;
}

int grid_blocks(int device, int mult) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
    cus = 256;
  return cus * (mult > 0 ? mult : 8);
}

double wall_clock_hz(int device) {
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || khz <= 0)
    khz = 100000;
  return (double)khz * 1000.0;
}

void fill_options(const cgbfs_options *opt, int64_t n, bool multi_source, cgbfs_options *o) {
  memset(o, 0, sizeof(*o));
  if (opt) *o = *opt;
  if (o->alpha <= 0) o->alpha = 14.0;
  if (o->beta <= 0) o->beta = 24.0;
  if (o->tau <= 0) o->tau = (int64_t)(0.15 * (double)n + 0.5);
  if (o->max_levels <= 0) o->max_levels = n + 1;
  // Blocks per CU of the grid-stride kernels (and the cap of the pull's
  // resident grid): short rounds on small graphs pay for every extra block
  // (each one's end-of-round probes and counter updates), large graphs need
  // the full grid.  Measured on R-MAT ef16 graphs: 64-root passes are best
  // at 2 up to 2^18 vertices, 4 from 2^19 to 2^22, 8 from 2^23; single-root
  // searches (per search, 2 / 4 / 8 blocks per CU) s18 0.24 / 0.30 / 0.41
  // ms, s20 0.30 / 0.34 / 0.45, s21 0.36 / 0.39 / 0.50, s22 0.50 / 0.48 /
  // 0.57, s23 0.81 / 0.67 / 0.73, s24 1.45 / 1.07 / 1.04.
  if (o->grid_mult <= 0) {
    if (multi_source) o->grid_mult = n < (1 << 19) ? 2 : n < (1 << 23) ? 4 : 8;
    else o->grid_mult = n <= (1 << 21) ? 2 : n <= (1 << 24) ? 4 : 8;
  }
  if (o->seg_grab <= 0) o->seg_grab = 32;
  if (o->push_wave_degree <= 0) o->push_wave_degree = (int32_t)kWaveDegree;
  if (o->push_big_degree <= 0)
    o->push_big_degree = o->push_wave_degree > (int32_t)kBigDegree ? o->push_wave_degree : (int32_t)kBigDegree;
}

int check_options(const cgbfs_options *o) {
  if (o->direction < 0 || o->direction > CGBFS_DIR_BEAMER_LANES)
    return set_error(CGBFS_EINVAL, "unknown direction mode");
  if (o->pull_grab < 0 || o->pull_grab > 1024) return set_error(CGBFS_EINVAL, "pull_grab must be 0 (default) or in [1, 1024]");
  if (o->seg_grab < 1 || o->seg_grab > 1024) return set_error(CGBFS_EINVAL, "seg_grab must be in [1, 1024]");
  if (o->grid_mult < 1 || o->grid_mult > 64) return set_error(CGBFS_EINVAL, "grid_mult must be in [1, 64]");
  if (o->push_wave_degree < 1 || o->push_wave_degree > (1 << 30))
    return set_error(CGBFS_EINVAL, "push_wave_degree must be in [1, 2^30]");
  if (o->push_big_degree < o->push_wave_degree || o->push_big_degree > (1 << 30))
    return set_error(CGBFS_EINVAL, "push_big_degree must be in [push_wave_degree, 2^30]");
  if (!(o->alpha > 0) || !(o->beta > 0)) return set_error(CGBFS_EINVAL, "alpha and beta must be positive");
  if (o->loop < CGBFS_LOOP_AUTO || o->loop > CGBFS_LOOP_WORKGROUP)
    return set_error(CGBFS_EINVAL, "unknown loop mode");
  return CGBFS_OK;
}

}  // namespace cg

using namespace cg;

// Graphs up to this many vertices run their levels in one persistent launch
// by default (CGBFS_LOOP_AUTO).  Measured on R-MAT ef16 searches, device time
// per search, per-level kernels / persistent: s10 (balanced 4x4) 0.15 /
// 0.10 ms, s12 0.18 / 0.12, s14 0.19 / 0.15, s15 0.20 / 0.17, s16 0.21 /
// 0.22, s21 0.41 / 0.87: beyond 2^15 vertices the grid barriers (two or
// three per level, over hundreds of blocks) cost more than the four
// launches per level they replace.
constexpr int64_t kPersistentMaxN = (int64_t)1 << 15;

// Caps *blocks at what the device holds resident at once for bfs_persistent
static int persistent_grid(int device, int *blocks) {
  static std::mutex mu;
  static int occ_of[kMaxDevices];
  std::lock_guard<std::mutex> lock(mu);
  if (device < 0 || device >= kMaxDevices) return set_error(CGBFS_EINVAL, "device index out of range");
  if (!occ_of[device]) {
    int occ = 0, cus = 0;
    CG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, bfs_persistent, kBlock, 0));
    CG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    if (occ <= 0 || cus <= 0) return set_error(CGBFS_EHIP, "no resident blocks for the persistent search");
    occ_of[device] = occ * cus;
  }
  // ... and at what the grid barrier's counters cover
  *blocks = std::max(1, std::min({*blocks, occ_of[device], kBarGroups * kBarGroupSize}));
  return CGBFS_OK;
}

// The search itself, shared by cgbfs_bfs and cgbfs_bfs_noalloc: outputs go
// to device arrays par / lev (the caller's or the workspace's).  events: time
// each round's traversal kernels with HIP events; without them a round's time
// is the span between its decide and the one before (device clock).
// One level of the per-level loop: the push or the pull, the push's big
// vertices, the decide.  Each kernel gates itself on the control block.
static void level_round(const LevelArgs &A, int blocks, int64_t n, const BfsParams &P,
                        hipStream_t st) {
  bfs_level<<<blocks, kBlock, 0, st>>>(A);
  bfs_push_scan<<<1, 1024, 0, st>>>(A);
  bfs_push_big<<<blocks, kBlock, 0, st>>>(A);
  bfs_decide<<<1, 64, 0, st>>>(A, n, P);
}
constexpr int kReplayRounds = 12;  // rounds per captured graph

static int bfs_run(DeviceGraph *g, const Workspace &ws, int64_t source, int32_t *lev,
                   uint32_t *par, const cgbfs_options &o, cgbfs_stats *stats, bool events) {
  LevelArgs A;
  A.out_ptr = g->out_ptr;
  A.out_idx = g->out_idx;
  A.in_ptr = g->in_ptr;
  A.in_idx = g->in_idx;
  A.in_head = g->in_head;
  A.noin = reinterpret_cast<const uint32_t *>(g->noin);
  A.levels = lev;
  A.parents = par;
  A.visited = ws.visited;
  for (int i = 0; i < 3; i++) A.fb[i] = ws.fb[i];
  A.bigq = ws.bigq;
  A.bigpre = ws.bigpre;
  A.ctrl = ws.ctrl;
  A.hist = ws.hist;
  A.cursor = ws.cursor;
  A.n = g->n;
  A.nwords = ws.nwords;
  A.symmetric = g->in_ptr == g->out_ptr ? 1 : 0;
  A.wave_degree = o.push_wave_degree;
  A.big_degree = o.push_big_degree;
  A.pull_grab = o.pull_grab > 0 ? o.pull_grab : kPullGrab;

  hipStream_t st = g->stream;
  const int blocks = grid_blocks(g->device, o.grid_mult);
  EventPool pool;
  BfsParams P;
  P.mode = o.direction;
  P.alpha = o.alpha;
  P.beta = o.beta;
  P.tau = (long long)o.tau;
  P.max_levels = (long long)o.max_levels;
  P.n_in = (long long)g->n_in;
  const bool tiny = o.loop == CGBFS_LOOP_WORKGROUP ||
                    (o.loop == CGBFS_LOOP_AUTO && g->n <= kTinyMaxN);
  const bool persistent = o.loop == CGBFS_LOOP_PERSISTENT ||
                          (o.loop == CGBFS_LOOP_AUTO && g->n <= kPersistentMaxN && !tiny);
  bfs_start<<<1, 64, 0, st>>>(A.ctrl);
  if (!tiny) {
    bfs_init_fill<<<blocks, kBlock, 0, st>>>(A);
    bfs_init_source<<<1, 64, 0, st>>>(A, source, P, (long long)g->total_out_edges);
  }
  CG_HIP(hipGetLastError());

  Ctrl hc;
  hipEvent_t rev[2 * 64];  // per-round events of the first 64 rounds
  int nrev = 0;
  if (tiny) {
    // One workgroup runs the whole search and writes every output itself.
    bfs_tiny<<<1, kTinyBlock, 0, st>>>(A, source, P, (long long)g->total_out_edges);
    CG_HIP(hipGetLastError());
    CG_HIP(hipMemcpyAsync(&hc, A.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
    CG_HIP(hipStreamSynchronize(st));
  } else if (persistent) {
    // One cooperative launch: as many blocks as give every wavefront a
    // chunk, at most the usual grid and at most what is resident at once.
    const int64_t nch = A.nwords / 2;
    int pblocks = (int)std::min<int64_t>(blocks, (nch + kWavesPerBlock - 1) / kWavesPerBlock);
    int rc = persistent_grid(g->device, &pblocks);
    if (rc) return rc;
    int64_t n_arg = g->n;
    void *args[] = {(void *)&A, (void *)&n_arg, (void *)&P};
    CG_HIP(hipLaunchCooperativeKernel((const void *)bfs_persistent, dim3(pblocks), dim3(kBlock), args, 0, st));
    CG_HIP(hipMemcpyAsync(&hc, A.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
    CG_HIP(hipStreamSynchronize(st));
  } else if (o.capture && events) {
    // Graph replay: kReplayRounds rounds of the level kernels are captured
    // once (every kernel reads its round and direction from the control
    // block), the instantiated graph is kept on the handle while the launch
    // arguments match, and it is replayed until the done flag is up: one
    // host poll per replay, no per-round events.
    std::vector<char> key(sizeof(A) + sizeof(P) + sizeof(int) + sizeof(int64_t));
    std::memcpy(key.data(), &A, sizeof(A));
    std::memcpy(key.data() + sizeof(A), &P, sizeof(P));
    std::memcpy(key.data() + sizeof(A) + sizeof(P), &blocks, sizeof(int));
    std::memcpy(key.data() + sizeof(A) + sizeof(P) + sizeof(int), &g->n, sizeof(int64_t));
    if (g->bfs_exec && g->bfs_key != key) {
      (void)hipGraphExecDestroy(g->bfs_exec);
      g->bfs_exec = nullptr;
    }
    if (!g->bfs_exec) {
      hipGraph_t graph = nullptr;
      hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
      if (e == hipSuccess) {
        for (int k = 0; k < kReplayRounds; k++) level_round(A, blocks, g->n, P, st);
        const hipError_t ec = hipStreamEndCapture(st, &graph);
        e = ec != hipSuccess ? ec : hipGetLastError();
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&g->bfs_exec, graph, nullptr, nullptr, 0);
      if (graph) (void)hipGraphDestroy(graph);
      if (e != hipSuccess) {
        g->bfs_exec = nullptr;
        return hip_fail(e, "bfs round graph", __FILE__, __LINE__);
      }
      g->bfs_key = key;
    }
    for (int64_t r = 0;; r += kReplayRounds) {
      CG_HIP(hipGraphLaunch(g->bfs_exec, st));
      CG_HIP(hipMemcpyAsync(&hc, A.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
      CG_HIP(hipStreamSynchronize(st));
      if (hc.done || r > o.max_levels + kReplayRounds) break;
    }
  } else {
    // Rounds are enqueued in chunks without host synchronisation; kernels of
    // rounds past the end see ctrl->done and return at once.  The host polls
    // the done flag once per chunk.  The search time is the device clock from
    // bfs_start to the decide of the last round (ctrl->t0 .. ctrl->t_end), so
    // neither the gated launches after the end nor the host poll count.
    int64_t d = 0;
    int chunk = o.rounds_hint > 0 ? o.rounds_hint : 12;
    for (;;) {
      for (int k = 0; k < chunk; k++, d++) {
        hipEvent_t ra = nullptr, rb = nullptr;
        if (events && nrev + 2 <= 2 * 64 && (ra = pool.make()) && (rb = pool.make())) {
          rev[nrev++] = ra;
          rev[nrev++] = rb;
          hipEventRecord(ra, st);
        }
        bfs_level<<<blocks, kBlock, 0, st>>>(A);
        bfs_push_scan<<<1, 1024, 0, st>>>(A);
        bfs_push_big<<<blocks, kBlock, 0, st>>>(A);
        if (rb) hipEventRecord(rb, st);
        bfs_decide<<<1, 64, 0, st>>>(A, g->n, P);
      }
      CG_HIP(hipGetLastError());
      CG_HIP(hipMemcpyAsync(&hc, A.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
      CG_HIP(hipStreamSynchronize(st));
      if (hc.done) break;
      if (d > o.max_levels + chunk) break;
      chunk = 4;
    }
  }
  if (hc.status == CGBFS_EHIP) return set_error(hc.status, "a grid barrier of the persistent search gave up");
  if (hc.status) return set_error(hc.status, "BFS exceeded max_levels");
  const double hz = wall_clock_hz(g->device);

  if (stats) {
    stats->rounds_completed = hc.rounds;
    stats->total_ms = (double)(hc.t_end - hc.t0) / hz * 1e3;
    stats->output_ms = 0;
    int64_t nh = hc.rounds < kHistCap ? hc.rounds : kHistCap;
    if (stats->max_hist < nh) nh = stats->max_hist;
    if (nh > 0) {
      std::vector<HistRow> hr((size_t)nh);
      CG_HIP(hipMemcpy(hr.data(), A.hist->r, sizeof(HistRow) * nh, hipMemcpyDeviceToHost));
      unsigned long long prev = hc.t0;
      for (int64_t r = 0; r < nh; r++) {
        const HistRow &H = hr[(size_t)r];
        if (stats->nf_history) stats->nf_history[r] = H.nf;
        if (stats->pairs_found) stats->pairs_found[r] = H.nf;
        if (stats->edges_scanned) stats->edges_scanned[r] = H.scanned;
        if (stats->rows_scanned) stats->rows_scanned[r] = H.rows;
        if (stats->direction_history) stats->direction_history[r] = H.dir;
        // a round's span: decide to decide on the device clock, or its
        // events (level kernels only) where they were recorded
        double ms = (double)(H.t - prev) / hz * 1e3;
        prev = H.t;
        if (2 * r + 1 < nrev) {
          float rm = 0;
          if (hipEventElapsedTime(&rm, rev[2 * r], rev[2 * r + 1]) == hipSuccess) ms = rm;
        }
        if (stats->level_ms) stats->level_ms[r] = ms;
      }
    }
    unsigned long long sc[2] = {0, 0};
    zero_u64<<<1, 64, 0, st>>>(ws.scalars, 2);
    bfs_count<<<grid_blocks(g->device, 2), kBlock, 0, st>>>(A.levels, g->m_weight, g->n, ws.scalars);
    CG_HIP(hipGetLastError());
    CG_HIP(hipMemcpyAsync(sc, ws.scalars, sizeof(sc), hipMemcpyDeviceToHost, st));
    CG_HIP(hipStreamSynchronize(st));
    stats->m_edges = (int64_t)sc[0];
    stats->visited_count = (int64_t)sc[1];
  }
  return CGBFS_OK;
}

static int bfs_args(cgbfs_graph *h, int64_t source, const void *parents, const void *levels,
                    const cgbfs_options *opt, cgbfs_options *o) {
  if (!h) return set_error(CGBFS_EINVAL, "graph is null");
  if (source < 0 || source >= h->g.n) return set_error(CGBFS_EINVAL, "source out of range");
  if (!parents || !levels) return set_error(CGBFS_EINVAL, "parents/levels must not be null");
  fill_options(opt, h->g.n, false, o);
  const int rc = check_options(o);
  if (rc) return rc;
  if (o->direction == CGBFS_DIR_BEAMER_LANES) o->direction = CGBFS_DIR_BEAMER;  // one lane
  if (o->loop == CGBFS_LOOP_WORKGROUP && h->g.n > kTinyMaxN)
    return set_error(CGBFS_EINVAL, "the workgroup loop takes graphs of at most 4096 vertices");
  return CGBFS_OK;
}

extern "C" int cgbfs_bfs(cgbfs_graph *h, int64_t source, int32_t *parents, int32_t *levels,
                         int32_t on_device, const cgbfs_options *opt, cgbfs_stats *stats) {
  cgbfs_options o;
  int rc = bfs_args(h, source, parents, levels, opt, &o);
  if (rc) return rc;
  DeviceGraph *g = &h->g;
  CG_HIP(hipSetDevice(g->device));
  Workspace ws;
  rc = ensure_ws1(g, &ws);
  if (rc) return rc;
  int32_t *lev = on_device ? levels : ws.levels;
  uint32_t *par = on_device ? (uint32_t *)parents : ws.parents;
  rc = bfs_run(g, ws, source, lev, par, o, stats, true);
  if (rc || on_device) return rc;
  EventPool pool;
  hipEvent_t o0 = pool.make(), o1 = pool.make();
  if (!o0 || !o1) return set_error(CGBFS_EHIP, "hipEventCreate failed");
  hipStream_t st = g->stream;
  CG_HIP(hipEventRecord(o0, st));
  CG_HIP(hipMemcpyAsync(parents, par, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, st));
  CG_HIP(hipMemcpyAsync(levels, lev, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, st));
  CG_HIP(hipEventRecord(o1, st));
  CG_HIP(hipEventSynchronize(o1));
  float oms = 0;
  hipEventElapsedTime(&oms, o0, o1);
  if (stats) stats->output_ms = oms;
  return CGBFS_OK;
}

extern "C" int cgbfs_bfs_noalloc(cgbfs_graph *h, int64_t source, int32_t *parents,
                                 int32_t *levels, const cgbfs_options *opt, cgbfs_stats *stats) {
  cgbfs_options o;
  int rc = bfs_args(h, source, parents, levels, opt, &o);
  if (rc) return rc;
  DeviceGraph *g = &h->g;
  if (g->n > kNoAllocMaxN)
    return set_error(CGBFS_EINVAL, "cgbfs_bfs_noalloc handles graphs of at most 2^21 vertices; "
                                   "use cgbfs_bfs for larger ones");
  CG_HIP(hipSetDevice(g->device));
  void *base = nullptr;
  rc = scratch_address(g->device, &base);
  if (rc) return rc;
  Workspace ws;
  ws1_layout(g->n, false, (char *)base, &ws);
  o.loop = CGBFS_LOOP_KERNELS;  // the cooperative launch is the runtime's; keep to plain launches
  return bfs_run(g, ws, source, levels, (uint32_t *)parents, o, stats, false);
}
