// Multi-source BFS for gfx950: up to 64 roots packed as the bits of one
// uint64 word per vertex, so one adjacency pass serves every root.
//
// Per vertex v the library keeps three lane words: vis[v] (lanes that have
// reached v), and the frontier pair fr[v] / nx[v] (lanes for which v is in
// the current / next level).  Every lane follows exactly the single-root
// semantics of bfs.hip (and of bool_pe.csl's round loop, :865-1156): lane
// l's level is its hop distance from roots[l] and its parent is the lowest-id
// in-neighbour one level closer.
//
// Work tiers (RMAT degrees span 1 .. ~10^6):
//   * degree <  kSeg    a wavefront takes 64 consecutive vertices and walks
//                       the concatenation of their rows, 64 entries per step
//                       (pull; push takes one vertex per wavefront);
//   * degree >= kSeg    the adjacency is cut into kSeg-entry segments listed
//                       once per graph; each segment is one wavefront's work
//                       item, and segments of one vertex combine per lane
//                       with atomicMin / atomicOr in a small scratch area.
//
// Pull round (ms_pull + ms_pull_fin): every vertex with an unreached lane
// ORs the frontier words of its ascending in-neighbours; within a step of
// 64 neighbours an exclusive prefix-OR across the wavefront gives each lane
// of each vertex its first (lowest-id) hit.  Stops once every needed lane is
// found.
// Push round (ms_push x2 + ms_prep): pass 1 ORs frontier words into nx[w]
// of out-neighbours (masked by the level-start vis[w]); the first writer of
// a vertex appends it to the next queue.  ms_prep then marks the new
// (lane, vertex) pairs, and pass 2 offers every frontier predecessor with a
// per-lane atomicMin, so the parent is again the lowest id.
// Outputs parents[l][v] (int32) and levels[l][v] (int16) are written once
// per (lane, vertex): at discovery, or by the final pass for unreached
// pairs.  No initial fill of the 6 B x 64 x V output arrays is needed.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "cg_common.h"

namespace cg {

constexpr int kMsBlock = 256;
constexpr int kMsWaves = kMsBlock / kWave;
constexpr int kMsHist = 4096;
constexpr int64_t kSeg = 512;       // segment length and wavefront-tier cap
constexpr int kMsLdsQueue = 2048;   // per-block staged queue entries
constexpr int kUnroll = 4;          // pull steps whose loads are issued together
constexpr int kFlush = 8;           // lane bits per batch of the pull's output flush

int grid_blocks(int device, int mult);
void fill_options(const cgbfs_options *opt, int64_t n, cgbfs_options *o);

struct MsCtrl {
  unsigned long long qtail[2];
  unsigned long long nf;       // vertices newly active after a pull round
  unsigned long long mf;       // out-degree sum of the next active set
  unsigned long long scanned;  // adjacency entries read this round
  unsigned long long rows;     // adjacency lists opened this round
  unsigned long long pairs;    // (lane, vertex) pairs reached this round
  long long prev_nf;
  long long rounds;
  int dir;
  int done;
  int status;
};

// Per-search constants, passed by value to the kernels that need them.
struct MsParams {
  int mode;
  double alpha, beta;
  long long tau, max_levels, nnz;
};

struct MsHist {
  long long nf[kMsHist];
  long long scanned[kMsHist];
  long long rows[kMsHist];
  long long pairs[kMsHist];
  int dir[kMsHist];
};

// Static segment list of one adjacency view (per graph).
struct SegList {
  int64_t nseg = 0;
  int64_t nbig = 0;
  int32_t *seg_v = nullptr;    // vertex of each segment
  int32_t *seg_bid = nullptr;  // big-vertex index of each segment
  int64_t *seg_beg = nullptr;  // first adjacency entry of each segment
  int32_t *big_v = nullptr;    // big vertices, by index
  uint8_t *is_big = nullptr;   // per vertex
};

struct MsArgs {
  const int64_t *out_ptr;
  const int32_t *out_idx;
  const int64_t *in_ptr;
  const int32_t *in_idx;
  int32_t *parents;  // [nroots][n]
  int16_t *levels;   // [nroots][n]
  uint64_t *vis;
  uint64_t *fw[2];   // frontier words, rotated by level parity
  int32_t *q[2];
  MsCtrl *ctrl;
  MsHist *hist;
  // big-vertex segments, pull (in view) and push (out view)
  SegList pin, pout;
  // Vertex-state slot of each vertex (hub-first order) and the adjacency
  // arrays re-expressed in slots, entry for entry; the lane words vis/fw
  // are indexed by slot, everything else by vertex id.
  const int32_t *slot_of;
  const int32_t *in_slot;
  const int32_t *out_slot;
  uint32_t *cand;    // [pin.nbig][64] per-lane first-hit candidates
  uint64_t *bgot;    // [pin.nbig] lanes found this round
  int64_t n;
  uint64_t act;      // mask of lanes in use
  int nroots;
};

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

struct MsGraphData {
  SegList pin, pout;
  void *mem = nullptr;
  int32_t *slot_of = nullptr;   // [n]
  int32_t *vertex_of = nullptr; // [n], inverse of slot_of
  int32_t *in_slot = nullptr;   // [nnz], in view
  int32_t *out_slot = nullptr;  // [nnz], out view (== in_slot when shared)
};

struct MsWorkspace {
  uint64_t *vis, *fw[2];
  int32_t *q[2];
  MsCtrl *ctrl;
  MsHist *hist;
  unsigned long long *mlane;
  int64_t *roots;
  uint32_t *cand;
  uint64_t *bgot;
};

// ---------------------------------------------------------------------------
// Per-graph segment lists (built on the host from the offset arrays once).
// ---------------------------------------------------------------------------
static void seg_plan(const std::vector<int64_t> &ptr, int64_t n, std::vector<int32_t> &sv,
                     std::vector<int32_t> &sb, std::vector<int64_t> &sg, std::vector<int32_t> &bv,
                     std::vector<uint8_t> &isb) {
  isb.assign(n, 0);
  for (int64_t v = 0; v < n; v++) {
    int64_t deg = ptr[v + 1] - ptr[v];
    if (deg < kSeg) continue;
    int32_t bid = (int32_t)bv.size();
    bv.push_back((int32_t)v);
    isb[v] = 1;
    for (int64_t b = ptr[v]; b < ptr[v + 1]; b += kSeg) {
      sv.push_back((int32_t)v);
      sb.push_back(bid);
      sg.push_back(b);
    }
  }
}

static int build_seglists(DeviceGraph *g, MsGraphData *md) {
  const int64_t n = g->n;
  std::vector<int64_t> ptr_in(n + 1), ptr_out;
  CG_HIP(hipMemcpy(ptr_in.data(), g->in_ptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
  std::vector<int32_t> sv[2], sb[2], bv[2];
  std::vector<int64_t> sg[2];
  std::vector<uint8_t> isb[2];
  seg_plan(ptr_in, n, sv[0], sb[0], sg[0], bv[0], isb[0]);
  const bool shared = g->in_ptr == g->out_ptr;
  if (!shared) {
    ptr_out.resize(n + 1);
    CG_HIP(hipMemcpy(ptr_out.data(), g->out_ptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
    seg_plan(ptr_out, n, sv[1], sb[1], sg[1], bv[1], isb[1]);
  }
  const int views = shared ? 1 : 2;
  size_t total = 0;
  for (int k = 0; k < views; k++)
    total += al(sv[k].size() * 4 + 4) + al(sb[k].size() * 4 + 4) + al(sg[k].size() * 8 + 8) +
             al(bv[k].size() * 4 + 4) + al((size_t)n);
  CG_HIP(hipMalloc(&md->mem, total));
  char *p = (char *)md->mem;
  for (int k = 0; k < views; k++) {
    SegList &L = (k == 0) ? md->pin : md->pout;
    L.nseg = (int64_t)sv[k].size();
    L.nbig = (int64_t)bv[k].size();
    L.seg_v = (int32_t *)p; p += al(sv[k].size() * 4 + 4);
    L.seg_bid = (int32_t *)p; p += al(sb[k].size() * 4 + 4);
    L.seg_beg = (int64_t *)p; p += al(sg[k].size() * 8 + 8);
    L.big_v = (int32_t *)p; p += al(bv[k].size() * 4 + 4);
    L.is_big = (uint8_t *)p; p += al((size_t)n);
    if (L.nseg) {
      CG_HIP(hipMemcpy(L.seg_v, sv[k].data(), sv[k].size() * 4, hipMemcpyHostToDevice));
      CG_HIP(hipMemcpy(L.seg_bid, sb[k].data(), sb[k].size() * 4, hipMemcpyHostToDevice));
      CG_HIP(hipMemcpy(L.seg_beg, sg[k].data(), sg[k].size() * 8, hipMemcpyHostToDevice));
      CG_HIP(hipMemcpy(L.big_v, bv[k].data(), bv[k].size() * 4, hipMemcpyHostToDevice));
    }
    CG_HIP(hipMemcpy(L.is_big, isb[k].data(), (size_t)n, hipMemcpyHostToDevice));
  }
  if (shared) md->pout = md->pin;
  return CGBFS_OK;
}

// Hub-first slots: vertices ordered by out-degree (descending, ties by id).
// An in-neighbour u is gathered once per out-edge of u, so this packs the
// most-gathered lane words into the first few MB, where they stay cached.
__global__ void slot_keys_kernel(const int64_t *__restrict__ out_ptr, int64_t n,
                                 uint64_t *__restrict__ keys) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int64_t deg = out_ptr[v + 1] - out_ptr[v];
  uint64_t rank = (uint64_t)(0x7FFFFFFFll - (deg < 0x7FFFFFFFll ? deg : 0x7FFFFFFFll));
  keys[v] = (rank << 32) | (uint64_t)v;
}

__global__ void slot_scatter_kernel(const uint64_t *__restrict__ sorted, int64_t n,
                                    int32_t *__restrict__ slot_of, int32_t *__restrict__ vertex_of) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int32_t v = (int32_t)(sorted[k] & 0xFFFFFFFFull);
  slot_of[v] = (int32_t)k;
  vertex_of[k] = v;
}

__global__ void slot_map_kernel(const int32_t *__restrict__ idx, int64_t nnz,
                                const int32_t *__restrict__ slot_of, int32_t *__restrict__ out) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  out[e] = slot_of[idx[e]];
}

static int build_slots(DeviceGraph *g, MsGraphData *md) {
  const int64_t n = g->n, nnz = g->nnz;
  hipStream_t st = g->stream;
  uint64_t *keys = nullptr, *sorted = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = CGBFS_OK;
  hipError_t e = hipMalloc(&md->slot_of, sizeof(int32_t) * n);
  if (e == hipSuccess) e = hipMalloc(&md->vertex_of, sizeof(int32_t) * n);
  if (e == hipSuccess) e = hipMalloc(&md->in_slot, sizeof(int32_t) * (nnz > 0 ? nnz : 1));
  if (e == hipSuccess && g->in_idx != g->out_idx)
    e = hipMalloc(&md->out_slot, sizeof(int32_t) * (nnz > 0 ? nnz : 1));
  if (e == hipSuccess) e = hipMalloc(&keys, sizeof(uint64_t) * n);
  if (e == hipSuccess) e = hipMalloc(&sorted, sizeof(uint64_t) * n);
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, keys, sorted, n, 0, 64, st);
  if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes + 16);
  if (e == hipSuccess) {
    slot_keys_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(g->out_ptr, n, keys);
    e = hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, keys, sorted, n, 0, 64, st);
  }
  if (e == hipSuccess) {
    slot_scatter_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(sorted, n, md->slot_of,
                                                                      md->vertex_of);
    if (nnz > 0) {
      slot_map_kernel<<<(unsigned)((nnz + 255) / 256), 256, 0, st>>>(g->in_idx, nnz, md->slot_of, md->in_slot);
      if (md->out_slot)
        slot_map_kernel<<<(unsigned)((nnz + 255) / 256), 256, 0, st>>>(g->out_idx, nnz, md->slot_of, md->out_slot);
    }
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) rc = hip_fail(e, "vertex slot build", __FILE__, __LINE__);
  if (!md->out_slot) md->out_slot = md->in_slot;
  hipFree(keys);
  hipFree(sorted);
  hipFree(tmp);
  return rc;
}

static void free_msdata(void *q) {
  MsGraphData *m = (MsGraphData *)q;
  if (m->mem) hipFree(m->mem);
  if (m->out_slot && m->out_slot != m->in_slot) hipFree(m->out_slot);
  if (m->in_slot) hipFree(m->in_slot);
  if (m->slot_of) hipFree(m->slot_of);
  if (m->vertex_of) hipFree(m->vertex_of);
  delete m;
}

static int ensure_wsm(DeviceGraph *g, MsWorkspace *w, MsGraphData **mdp) {
  const int64_t n = g->n;
  if (!g->msdata) {
    MsGraphData *md = new MsGraphData();
    int rc = build_seglists(g, md);
    if (!rc) rc = build_slots(g, md);
    if (rc) {
      free_msdata(md);
      return rc;
    }
    g->msdata = md;
    g->msdata_free = free_msdata;
  }
  MsGraphData *md = (MsGraphData *)g->msdata;
  *mdp = md;
  const int64_t nbig = md->pin.nbig;
  size_t wb = al(sizeof(uint64_t) * n), qb = al(sizeof(int32_t) * n);
  size_t total = 3 * wb + 2 * qb + al(sizeof(MsCtrl)) + al(sizeof(MsHist)) +
                 al(sizeof(unsigned long long) * 64) + al(sizeof(int64_t) * 64) +
                 al(sizeof(uint32_t) * 64 * (nbig + 1)) + al(sizeof(uint64_t) * (nbig + 1));
  if (g->wsm_bytes < total) {
    if (g->wsm) hipFree(g->wsm);
    g->wsm = nullptr;
    g->wsm_bytes = 0;
    CG_HIP(hipMalloc(&g->wsm, total));
    g->wsm_bytes = total;
  }
  char *p = (char *)g->wsm;
  w->vis = (uint64_t *)p; p += wb;
  w->fw[0] = (uint64_t *)p; p += wb;
  w->fw[1] = (uint64_t *)p; p += wb;
  w->q[0] = (int32_t *)p; p += qb;
  w->q[1] = (int32_t *)p; p += qb;
  w->ctrl = (MsCtrl *)p; p += al(sizeof(MsCtrl));
  w->hist = (MsHist *)p; p += al(sizeof(MsHist));
  w->mlane = (unsigned long long *)p; p += al(sizeof(unsigned long long) * 64);
  w->roots = (int64_t *)p; p += al(sizeof(int64_t) * 64);
  w->cand = (uint32_t *)p; p += al(sizeof(uint32_t) * 64 * (nbig + 1));
  w->bgot = (uint64_t *)p;
  return CGBFS_OK;
}

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t shfl_up64(uint64_t x, int off) {
  uint32_t lo = __shfl_up((uint32_t)x, off), hi = __shfl_up((uint32_t)(x >> 32), off);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t x, int src) {
  uint32_t lo = __shfl((uint32_t)x, src), hi = __shfl((uint32_t)(x >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

// Inclusive prefix-OR across the wavefront.
__device__ __forceinline__ uint64_t wave_scan_or(uint64_t x) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint64_t y = shfl_up64(x, off);
    if (lane >= off) x |= y;
  }
  return x;
}

__device__ __forceinline__ unsigned long long ms_block_sum(unsigned long long v,
                                                            unsigned long long *lds) {
  v = wave_sum_u64(v);
  int w = threadIdx.x / kWave;
  if (lane_id() == 0) lds[w] = v;
  __syncthreads();
  unsigned long long s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x / kWave); i++) s += lds[i];
  __syncthreads();
  return s;
}


// Wavefront scan of in-neighbour entries [beg, end) for lanes `need`
// (wave-uniform arguments): gathers lane words through in_slot and offers
// each lane's first hit in the range (its original id) to cand[] with
// atomicMin.  Returns the lanes found (wave-uniform).
__device__ __forceinline__ uint64_t ms_wave_pull(const MsArgs &A, const uint64_t *fr, int64_t beg,
                                                 int64_t end, uint64_t need, uint32_t *cand,
                                                 unsigned long long &scanned) {
  const int lane = lane_id();
  uint64_t got = 0;
  for (int64_t e0 = beg; e0 < end; e0 += kWave) {
    int64_t e = e0 + lane;
    uint64_t h = 0;
    if (e < end) h = fr[A.in_slot[e]] & need & ~got;
    uint64_t inc = wave_scan_or(h);
    uint64_t before = shfl_up64(inc, 1);
    if (lane == 0) before = 0;
    uint64_t mine = h & ~before;
    if (mine) {
      const uint32_t u = (uint32_t)A.in_idx[e];
      while (mine) {
        int l = __ffsll((unsigned long long)mine) - 1;
        mine &= mine - 1;
        atomicMin(&cand[l], u);
      }
    }
    got |= shfl64(inc, kWave - 1);
    if (lane == 0) scanned += (unsigned long long)((end - e0) < kWave ? (end - e0) : kWave);
    if (got == need) break;
  }
  return got;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kMsBlock) ms_init_fill(MsArgs A) {
  int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = tid; v < A.n; v += stride) {
    A.vis[v] = 0;
    A.fw[0][v] = 0;
    A.fw[1][v] = 0;
  }
  for (int64_t i = tid; i < A.pin.nbig * 64; i += stride) A.cand[i] = 0xFFFFFFFFu;
  for (int64_t i = tid; i < A.pin.nbig; i += stride) A.bgot[i] = 0;
}

__global__ void ms_init_roots(MsArgs A, const int64_t *roots, MsParams P) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned long long nq = 0, mf = 0;
  for (int l = 0; l < A.nroots; l++) {
    int64_t r = roots[l];
    int32_t sr = A.slot_of[r];
    uint64_t bit = 1ull << l;
    if (A.fw[0][sr] == 0) {
      A.q[0][nq++] = (int32_t)r;
      mf += (unsigned long long)(A.out_ptr[r + 1] - A.out_ptr[r]);
    }
    A.vis[sr] |= bit;
    A.fw[0][sr] |= bit;
    A.parents[(int64_t)l * A.n + r] = (int32_t)r;
    A.levels[(int64_t)l * A.n + r] = 0;
  }
  MsCtrl *c = A.ctrl;
  st_ctrl(&c->qtail[0], nq);
  st_ctrl(&c->qtail[1], 0ull);
  st_ctrl(&c->nf, 0ull);
  st_ctrl(&c->mf, 0ull);
  st_ctrl(&c->scanned, 0ull);
  st_ctrl(&c->rows, 0ull);
  st_ctrl(&c->pairs, 0ull);
  st_ctrl(&c->prev_nf, (long long)nq);
  st_ctrl(&c->rounds, 0ll);
  int dir = 0;
  if (P.mode == CGBFS_DIR_BOTTOMUP) dir = 1;
  else if (P.mode == CGBFS_DIR_BEAMER && (double)mf * P.alpha > (double)P.nnz) dir = 1;
  st_ctrl(&c->dir, dir);
  st_ctrl(&c->done, 0);
  st_ctrl(&c->status, 0);
}

// Pull round over non-big vertices (thread and wavefront tiers), then the
// big-vertex segments (one wavefront per segment, per-lane atomicMin into
// cand[], lanes into bgot[]).
__global__ void __launch_bounds__(kMsBlock) ms_pull(MsArgs A, int d) {
  MsCtrl *c = A.ctrl;
  __shared__ unsigned long long red[kMsWaves];
  const uint64_t *fr = A.fw[d & 1];
  uint64_t *nx = A.fw[(d + 1) & 1];
  const int16_t lev = (int16_t)(d + 1);
  unsigned long long scanned = 0, mf = 0, rows = 0, pairs = 0;
  const int lane = lane_id();
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t wave = gtid / kWave, nwaves = stride / kWave;
  // First hits of the current chunk, as offsets into the row (rows here are
  // shorter than kSeg), indexed [lane bit][row] so the per-bit flush below
  // reads one conflict-free row of LDS and stores 64 consecutive vertices.
  __shared__ uint16_t ptile[kMsWaves][kWave][kWave];
  uint16_t (*pt)[kWave] = ptile[threadIdx.x / kWave];
  unsigned long long nf = 0;
  // Chunks of 64 consecutive vertices per wavefront.  The rows of the chunk
  // that still need lanes (big rows excluded) are laid end to end in a
  // virtual edge range [0, total); each step every lane takes one virtual
  // entry, finds its row by binary search over the rows' virtual starts,
  // gathers the frontier word, and a segmented prefix-OR (segments = rows)
  // gives each (row, lane) its first -- lowest-id -- hit.  Row state (got)
  // lives in the lane that owns the row.
  const int64_t nch = (A.n + kWave - 1) / kWave;
  for (int64_t ch = wave; ch < nch; ch += nwaves) {
    const int64_t v = ch * kWave + lane;
    uint64_t vw = 0, need = 0;
    int64_t rb = 0;
    int32_t len = 0, sv_me = 0;
    if (v < A.n && !A.pin.is_big[v]) {
      sv_me = A.slot_of[v];
      vw = A.vis[sv_me];
      need = A.act & ~vw;
      if (need) {
        rb = A.in_ptr[v];
        len = (int32_t)(A.in_ptr[v + 1] - rb);  // < kSeg for non-big rows
      }
    }
    if (len) rows++;
    // exclusive prefix sum of row lengths -> virtual starts
    int32_t incl = len;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      int32_t y = __shfl_up(incl, off);
      if (lane >= off) incl += y;
    }
    const int32_t vstart = incl - len;
    const int32_t total = __shfl(incl, kWave - 1);
    uint64_t got = 0;
    const int32_t my_end = vstart + len;
    // kUnroll steps per pass: their index loads, then their frontier-word
    // gathers, are all in flight together; the scans then run step by step
    // (a row already satisfied when the pass starts issues no loads).
    for (int32_t base0 = 0; base0 < total; base0 += kWave * kUnroll) {
      int rk[kUnroll];
      int32_t rsk[kUnroll], idx[kUnroll];
      uint64_t nrk[kUnroll], raw[kUnroll];
#pragma unroll
      for (int k = 0; k < kUnroll; k++) {
        const int32_t ve = base0 + k * kWave + lane;
        // row r: the largest lane with vstart <= ve (rows of length 0 share
        // the next row's start, so r always has entries)
        int r = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
          int cand = r + step;
          int32_t sv = __shfl(vstart, cand);
          if (sv <= ve) r = cand;
        }
        rk[k] = r;
        rsk[k] = __shfl(vstart, r);
        const int64_t rbase = __shfl(rb, r);
        nrk[k] = shfl64(need, r);
        const uint64_t gr = shfl64(got, r);
        idx[k] = -1;
        if (ve < total && (nrk[k] & ~gr)) idx[k] = A.in_slot[rbase + (ve - rsk[k])];
      }
#pragma unroll
      for (int k = 0; k < kUnroll; k++) {
        raw[k] = 0;
        if (idx[k] >= 0) {
          raw[k] = fr[idx[k]] & nrk[k];
          scanned++;
        }
      }
#pragma unroll
      for (int k = 0; k < kUnroll; k++) {
        const int32_t base = base0 + k * kWave;
        const int r = rk[k], rs = rsk[k];
        uint64_t h = raw[k] & ~shfl64(got, r);
        // segmented inclusive prefix-OR; the segment of lane starts at head
        const int head = rs > base ? rs - base : 0;
        uint64_t x = h;
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
          uint64_t y = shfl_up64(x, off);
          if (lane - off >= head) x |= y;
        }
        uint64_t prev = shfl_up64(x, 1);
        uint64_t mine = h & ~(lane > head ? prev : 0ull);
        const uint16_t off = (uint16_t)(base + lane - rs);
        while (mine) {
          int l = __ffsll((unsigned long long)mine) - 1;
          mine &= mine - 1;
          pt[l][r] = off;
        }
        // each row owner collects its row's OR over this step
        int last = (my_end < base + kWave ? my_end : base + kWave) - 1 - base;
        bool hit_step = len && vstart < base + kWave && my_end > base;
        uint64_t tot = shfl64(x, hit_step ? last : lane);
        if (hit_step) got |= tot;
      }
      if (!__any(len && (need & ~got))) break;  // every row of the chunk satisfied
    }
    // Flush: per lane bit, 64 consecutive vertices' parents and levels.
    uint64_t bits = got;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      uint32_t lo = __shfl_xor((uint32_t)bits, off), hi = __shfl_xor((uint32_t)(bits >> 32), off);
      bits |= ((uint64_t)hi << 32) | lo;
    }
    // kFlush bits at a time: their parent-id loads are in flight together
    // before any store is issued.
    while (bits) {
      int lb[kFlush];
      int32_t pv[kFlush];
#pragma unroll
      for (int k = 0; k < kFlush; k++) {
        lb[k] = bits ? __ffsll((unsigned long long)bits) - 1 : -1;
        bits &= bits - 1;
      }
#pragma unroll
      for (int k = 0; k < kFlush; k++)
        pv[k] = (lb[k] >= 0 && ((got >> lb[k]) & 1ull)) ? A.in_idx[rb + pt[lb[k]][lane]] : 0;
#pragma unroll
      for (int k = 0; k < kFlush; k++) {
        if (lb[k] >= 0 && ((got >> lb[k]) & 1ull)) {
          int64_t o = (int64_t)lb[k] * A.n + v;
          A.parents[o] = pv[k];
          A.levels[o] = lev;
        }
      }
    }
    if (got) {
      nx[sv_me] = got;
      A.vis[sv_me] = vw | got;
      mf += (unsigned long long)(A.out_ptr[v + 1] - A.out_ptr[v]);
      pairs += (unsigned long long)__popcll(got);
      nf++;
    }
  }
  // Big vertices: one segment per wavefront.
  for (int64_t s = wave; s < A.pin.nseg; s += nwaves) {
    int32_t v = A.pin.seg_v[s];
    uint64_t need = A.act & ~A.vis[A.slot_of[v]];
    if (!need) continue;
    int32_t bid = A.pin.seg_bid[s];
    int64_t beg = A.pin.seg_beg[s];
    int64_t rend = A.in_ptr[v + 1];
    int64_t end = beg + kSeg < rend ? beg + kSeg : rend;
    uint64_t gj = ms_wave_pull(A, fr, beg, end, need, &A.cand[(int64_t)bid * 64], scanned);
    if (gj && lane == 0) atomicOr((unsigned long long *)&A.bgot[bid], (unsigned long long)gj);
  }
  unsigned long long s = ms_block_sum(nf, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->nf, s);
  s = ms_block_sum(scanned, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->scanned, s);
  s = ms_block_sum(mf, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->mf, s);
  s = ms_block_sum(rows, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->rows, s);
  s = ms_block_sum(pairs, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->pairs, s);
}

// Big vertices after a pull round: turn per-lane candidates into outputs.
__global__ void __launch_bounds__(kMsBlock) ms_pull_fin(MsArgs A, int d) {
  MsCtrl *c = A.ctrl;
  uint64_t *nx = A.fw[(d + 1) & 1];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < A.pin.nbig; b += stride) {
    {
      const int32_t v = A.pin.big_v[b];
      const uint64_t got = A.bgot[b];
      if (got) {
        uint64_t lanes = got;
        while (lanes) {
          int l = __ffsll((unsigned long long)lanes) - 1;
          lanes &= lanes - 1;
          int64_t o = (int64_t)l * A.n + v;
          A.parents[o] = (int32_t)A.cand[b * 64 + l];
          A.levels[o] = (int16_t)(d + 1);
          A.cand[b * 64 + l] = 0xFFFFFFFFu;
        }
        A.bgot[b] = 0;
        const int32_t sv = A.slot_of[v];
        nx[sv] = got;
        A.vis[sv] |= got;
        atomicAdd(&c->mf, (unsigned long long)(A.out_ptr[v + 1] - A.out_ptr[v]));
        atomicAdd(&c->pairs, (unsigned long long)__popcll(got));
        atomicAdd(&c->nf, 1ull);
      }
    }
  }
}

// Active list from the frontier words (slot order -> vertex ids), for a push
// round that follows a pull round (pull rounds keep no list).
__global__ void __launch_bounds__(kMsBlock) ms_compact(MsArgs A, int d, const int32_t *vertex_of) {
  MsCtrl *c = A.ctrl;
  const uint64_t *fr = A.fw[d & 1];
  __shared__ int32_t lq[kMsLdsQueue];
  __shared__ unsigned lq_n;
  __shared__ unsigned long long lq_base;
  StagedQueue sq{lq, &lq_n, &lq_base, (unsigned)kMsLdsQueue, &c->qtail[d & 1], A.q[d & 1]};
  sq.init();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t nround = ((A.n + stride - 1) / stride) * stride;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nround; k += stride) {
    bool on = k < A.n && fr[k] != 0;
    sq.push(on, on ? vertex_of[k] : 0);
  }
  sq.flush();
}

// Clears a whole frontier-word buffer (after a round whose frontier has no
// list).
__global__ void __launch_bounds__(kMsBlock) ms_clear_all(uint64_t *w, int64_t n) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    w[k] = 0;
}

// Push pass (phase 0: OR into nx and append first writers; phase 1: offer
// predecessors to the new pairs with per-lane atomicMin).
__global__ void __launch_bounds__(kMsBlock) ms_push(MsArgs A, int d, int phase) {
  MsCtrl *c = A.ctrl;
  __shared__ unsigned long long red[kMsWaves];
  const uint64_t *fr = A.fw[d & 1];
  uint64_t *nx = A.fw[(d + 1) & 1];
  const int32_t *Q = A.q[d & 1];
  int32_t *QN = A.q[(d + 1) & 1];
  unsigned long long *tail = &c->qtail[(d + 1) & 1];
  const int64_t nq = (int64_t)ld_ctrl(&c->qtail[d & 1]);
  const int lane = lane_id();
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t wave = gtid / kWave, nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  unsigned long long scanned = 0, rows = 0;
  __shared__ int32_t lq[kMsLdsQueue];
  __shared__ unsigned lq_n;
  __shared__ unsigned long long lq_base;
  StagedQueue sq{lq, &lq_n, &lq_base, (unsigned)kMsLdsQueue, tail, QN};
  if (phase == 0) sq.init();
  // Work items: queue entries (skipping big vertices), then big segments of
  // active big vertices.  Both loops are wavefront-uniform.
  const int64_t nitems = nq + A.pout.nseg;
  for (int64_t it = wave; it < nitems; it += nwaves) {
    int32_t u;
    int64_t beg, end;
    if (it < nq) {
      u = Q[it];
      if (A.pout.is_big[u]) continue;
      beg = A.out_ptr[u];
      end = A.out_ptr[u + 1];
    } else {
      int64_t s = it - nq;
      u = A.pout.seg_v[s];
      beg = A.pout.seg_beg[s];
      int64_t rend = A.out_ptr[u + 1];
      end = beg + kSeg < rend ? beg + kSeg : rend;
    }
    uint64_t f = fr[A.slot_of[u]];
    if (!f) continue;
    if (lane == 0) {
      scanned += (unsigned long long)(end - beg);
      if (it < nq || beg == A.out_ptr[u]) rows++;
    }
    for (int64_t e0 = beg; e0 < end; e0 += kWave) {
      int64_t e = e0 + lane;
      bool first = false;
      int32_t w = 0;
      if (e < end) {
        w = A.out_idx[e];
        const int32_t sw = A.out_slot[e];
        if (phase == 0) {
          uint64_t nv = f & ~A.vis[sw];
          if (nv) {
            uint64_t old = atomicOr((unsigned long long *)&nx[sw], (unsigned long long)nv);
            first = (old == 0);
          }
        } else {
          uint64_t nv = f & nx[sw];
          while (nv) {
            int l = __ffsll((unsigned long long)nv) - 1;
            nv &= nv - 1;
            atomicMin((uint32_t *)&A.parents[(int64_t)l * A.n + w], (uint32_t)u);
          }
        }
      }
      if (phase == 0) sq.push(first, w);
    }
  }
  if (phase == 0) sq.flush();
  unsigned long long s = ms_block_sum(scanned, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->scanned, s);
  s = ms_block_sum(rows, red);
  if (phase == 0 && threadIdx.x == 0 && s) atomicAdd(&c->rows, s);
}

// Between the push passes: mark new pairs (parent slot = +inf for the
// atomicMin of pass 2, level = d+1), fold them into vis.
__global__ void __launch_bounds__(kMsBlock) ms_prep(MsArgs A, int d) {
  MsCtrl *c = A.ctrl;
  __shared__ unsigned long long red[kMsWaves];
  const uint64_t *nx = A.fw[(d + 1) & 1];
  const int32_t *QN = A.q[(d + 1) & 1];
  const int64_t nq = (int64_t)ld_ctrl(&c->qtail[(d + 1) & 1]);
  unsigned long long mf = 0, pairs = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq;
       i += (int64_t)gridDim.x * blockDim.x) {
    int32_t v = QN[i];
    const int32_t sv = A.slot_of[v];
    uint64_t nw = nx[sv];
    uint64_t lanes = nw;
    while (lanes) {
      int l = __ffsll((unsigned long long)lanes) - 1;
      lanes &= lanes - 1;
      int64_t o = (int64_t)l * A.n + v;
      A.parents[o] = -1;  // 0xFFFFFFFF: +inf for the unsigned atomicMin
      A.levels[o] = (int16_t)(d + 1);
    }
    A.vis[sv] |= nw;
    mf += (unsigned long long)(A.out_ptr[v + 1] - A.out_ptr[v]);
    pairs += (unsigned long long)__popcll(nw);
  }
  unsigned long long s = ms_block_sum(mf, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->mf, s);
  s = ms_block_sum(pairs, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->pairs, s);
}

// Clears the level-d frontier words (listed in Q[d]) so that buffer is zero
// when it becomes the next-words buffer of round d+1.
__global__ void __launch_bounds__(kMsBlock) ms_clear(MsArgs A, int d) {
  MsCtrl *c = A.ctrl;
  uint64_t *fr = A.fw[d & 1];
  const int32_t *Q = A.q[d & 1];
  const int64_t nq = (int64_t)ld_ctrl(&c->qtail[d & 1]);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq;
       i += (int64_t)gridDim.x * blockDim.x)
    fr[A.slot_of[Q[i]]] = 0;
}

__global__ void ms_decide(MsArgs A, int d, MsParams P) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  MsCtrl *c = A.ctrl;
  const int cur = ld_ctrl(&c->dir);
  // push rounds list the new active vertices; pull rounds only count them
  const long long nf =
      (long long)(cur == 0 ? ld_ctrl(&c->qtail[(d + 1) & 1]) : ld_ctrl(&c->nf));
  const long long mf = (long long)ld_ctrl(&c->mf);
  if (d < kMsHist) {
    A.hist->nf[d] = nf;
    A.hist->scanned[d] = (long long)ld_ctrl(&c->scanned);
    A.hist->rows[d] = (long long)ld_ctrl(&c->rows);
    A.hist->pairs[d] = (long long)ld_ctrl(&c->pairs);
    A.hist->dir[d] = cur;
  }
  st_ctrl(&c->nf, 0ull);
  st_ctrl(&c->mf, 0ull);
  st_ctrl(&c->scanned, 0ull);
  st_ctrl(&c->rows, 0ull);
  st_ctrl(&c->pairs, 0ull);
  st_ctrl(&c->qtail[d & 1], 0ull);
  st_ctrl(&c->rounds, (long long)(d + 1));
  if (nf == 0) {
    st_ctrl(&c->done, 1);
    return;
  }
  int dir = cur;
  switch (P.mode) {
    case CGBFS_DIR_TOPDOWN: dir = 0; break;
    case CGBFS_DIR_BOTTOMUP: dir = 1; break;
    case CGBFS_DIR_TAU:
      if (dir == 0 && nf > P.tau) dir = 1;
      break;
    default:
      if (dir == 0) {
        if ((double)mf * P.alpha > (double)P.nnz) dir = 1;
      } else {
        if ((double)nf * P.beta < (double)A.n && nf < ld_ctrl(&c->prev_nf)) dir = 0;
      }
  }
  st_ctrl(&c->prev_nf, nf);
  st_ctrl(&c->dir, dir);
  if (d + 1 >= P.max_levels) {
    st_ctrl(&c->status, (int)CGBFS_ETIMEOUT);
    st_ctrl(&c->done, 1);
  }
}

// Final pass: -1 into every (lane, vertex) pair no lane word reached.
__global__ void __launch_bounds__(kMsBlock) ms_finalize(MsArgs A) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < A.n;
       v += (int64_t)gridDim.x * blockDim.x) {
    uint64_t miss = A.act & ~A.vis[A.slot_of[v]];
    while (miss) {
      int l = __ffsll((unsigned long long)miss) - 1;
      miss &= miss - 1;
      int64_t o = (int64_t)l * A.n + v;
      A.parents[o] = -1;
      A.levels[o] = -1;
    }
  }
}

// GTEPS m per lane: sum of the per-vertex weight over each lane's reached set.
__global__ void __launch_bounds__(kMsBlock) ms_count(const uint64_t *vis, const int32_t *slot_of,
                                                     const int32_t *w, int64_t n, int nroots,
                                                     unsigned long long *out) {
  __shared__ unsigned long long acc[64];
  if (threadIdx.x < 64) acc[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    unsigned long long wv = (unsigned long long)w[v];
    if (!wv) continue;
    uint64_t b = vis[slot_of[v]];
    while (b) {
      int l = __ffsll((unsigned long long)b) - 1;
      b &= b - 1;
      atomicAdd(&acc[l], wv);
    }
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)nroots && acc[threadIdx.x]) atomicAdd(&out[threadIdx.x], acc[threadIdx.x]);
}

__global__ void ms_zero(unsigned long long *p, int k) {
  if (threadIdx.x < (unsigned)k) p[threadIdx.x] = 0;
}

}  // namespace cg

using namespace cg;

extern "C" int cgbfs_msbfs(cgbfs_graph *h, const int64_t *roots, int32_t nroots, int32_t *parents,
                           int16_t *levels, int32_t on_device, const cgbfs_options *opt,
                           cgbfs_stats *stats, int64_t *m_lanes) {
  if (!h) return set_error(CGBFS_EINVAL, "graph is null");
  DeviceGraph *g = &h->g;
  if (!roots || nroots < 1 || nroots > 64) return set_error(CGBFS_EINVAL, "nroots must be in [1, 64]");
  if (!parents || !levels) return set_error(CGBFS_EINVAL, "parents/levels must not be null");
  for (int l = 0; l < nroots; l++)
    if (roots[l] < 0 || roots[l] >= g->n) return set_error(CGBFS_EINVAL, "root out of range");
  cgbfs_options o;
  fill_options(opt, g->n, &o);
  if (o.direction < 0 || o.direction > 3) return set_error(CGBFS_EINVAL, "unknown direction mode");
  if (o.max_levels > 32767) o.max_levels = 32767;  // int16 levels
  CG_HIP(hipSetDevice(g->device));
  MsWorkspace ws;
  MsGraphData *md = nullptr;
  int rc = ensure_wsm(g, &ws, &md);
  if (rc) return rc;
  hipStream_t st = g->stream;

  const size_t plane = (size_t)g->n * nroots;
  int32_t *dpar = parents;
  int16_t *dlev = levels;
  void *tmp = nullptr;
  if (!on_device) {
    CG_HIP(hipMalloc(&tmp, al(plane * sizeof(int32_t)) + plane * sizeof(int16_t) + 256));
    dpar = (int32_t *)tmp;
    dlev = (int16_t *)((char *)tmp + al(plane * sizeof(int32_t)));
  }

  MsArgs A;
  A.out_ptr = g->out_ptr;
  A.out_idx = g->out_idx;
  A.in_ptr = g->in_ptr;
  A.in_idx = g->in_idx;
  A.parents = dpar;
  A.levels = dlev;
  A.vis = ws.vis;
  A.fw[0] = ws.fw[0];
  A.fw[1] = ws.fw[1];
  A.q[0] = ws.q[0];
  A.q[1] = ws.q[1];
  A.ctrl = ws.ctrl;
  A.hist = ws.hist;
  A.pin = md->pin;
  A.pout = md->pout;
  A.slot_of = md->slot_of;
  A.in_slot = md->in_slot;
  A.out_slot = md->out_slot;
  A.cand = ws.cand;
  A.bgot = ws.bgot;
  A.n = g->n;
  A.nroots = nroots;
  A.act = nroots == 64 ? ~0ull : ((1ull << nroots) - 1ull);

  const int blocks = grid_blocks(g->device, o.grid_mult);
  std::vector<hipEvent_t> evs;
  auto new_event = [&](hipEvent_t *e) -> hipError_t {
    hipError_t r = hipEventCreate(e);
    if (r == hipSuccess) evs.push_back(*e);
    return r;
  };
  auto cleanup = [&]() {
    for (hipEvent_t e : evs) hipEventDestroy(e);
    if (tmp) hipFree(tmp);
  };
  hipEvent_t e0, e1;
  hipError_t e = hipMemcpyAsync(ws.roots, roots, sizeof(int64_t) * nroots, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = new_event(&e0);
  if (e == hipSuccess) e = new_event(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  if (e != hipSuccess) {
    cleanup();
    return hip_fail(e, "msbfs setup", __FILE__, __LINE__);
  }
  ms_init_fill<<<blocks, kMsBlock, 0, st>>>(A);
  MsParams P;
  P.mode = o.direction;
  P.alpha = o.alpha;
  P.beta = o.beta;
  P.tau = (long long)o.tau;
  P.max_levels = (long long)o.max_levels;
  P.nnz = (long long)g->nnz;
  ms_init_roots<<<1, 64, 0, st>>>(A, ws.roots, P);
  MsCtrl hc;
  e = hipMemcpyAsync(&hc, A.ctrl, sizeof(MsCtrl), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  // One host poll per round: only the chosen direction's kernels launch,
  // and each round's traversal kernels are bracketed by a pair of events.
  std::vector<std::pair<hipEvent_t, hipEvent_t>> rev;
  std::vector<int> rdir;
  int64_t d = 0;
  // listed: the current frontier's vertices are listed in Q[d&1] (the roots,
  // or the output of a push round).  A pull round leaves no list; a push
  // round after it first compacts one from the frontier words.
  bool listed = true;
  while (e == hipSuccess && !hc.done) {
    hipEvent_t ra, rb;
    if ((e = new_event(&ra)) != hipSuccess || (e = new_event(&rb)) != hipSuccess) break;
    hipEventRecord(ra, st);
    if (hc.dir == 1) {
      ms_pull<<<blocks, kMsBlock, 0, st>>>(A, (int)d);
      if (md->pin.nbig) ms_pull_fin<<<(unsigned)((md->pin.nbig + kMsBlock - 1) / kMsBlock), kMsBlock, 0, st>>>(A, (int)d);
    } else {
      if (!listed) ms_compact<<<blocks, kMsBlock, 0, st>>>(A, (int)d, md->vertex_of);
      ms_push<<<blocks, kMsBlock, 0, st>>>(A, (int)d, 0);
      ms_prep<<<blocks, kMsBlock, 0, st>>>(A, (int)d);
      ms_push<<<blocks, kMsBlock, 0, st>>>(A, (int)d, 1);
      listed = true;
    }
    hipEventRecord(rb, st);
    rev.push_back({ra, rb});
    rdir.push_back(hc.dir);
    if (listed) ms_clear<<<blocks, kMsBlock, 0, st>>>(A, (int)d);
    else ms_clear_all<<<blocks, kMsBlock, 0, st>>>(A.fw[d & 1], g->n);
    listed = (hc.dir == 0);  // the next frontier is listed iff this round pushed
    ms_decide<<<1, 64, 0, st>>>(A, (int)d, P);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&hc, A.ctrl, sizeof(MsCtrl), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    d++;
  }
  // The final pass belongs to producing the outputs, so e1 follows it.
  if (e == hipSuccess) {
    ms_finalize<<<blocks, kMsBlock, 0, st>>>(A);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e != hipSuccess || hc.status) {
    cleanup();
    if (e != hipSuccess) return hip_fail(e, "msbfs rounds", __FILE__, __LINE__);
    return set_error(hc.status, "MS-BFS exceeded max_levels (int16 levels cap at 32767)");
  }

  double out_ms = 0;
  if (!on_device) {
    hipEvent_t o0, o1;
    e = new_event(&o0);
    if (e == hipSuccess) e = new_event(&o1);
    if (e == hipSuccess) e = hipEventRecord(o0, st);
    if (e == hipSuccess) e = hipMemcpyAsync(parents, dpar, plane * sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(levels, dlev, plane * sizeof(int16_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(o1, st);
    if (e == hipSuccess) e = hipEventSynchronize(o1);
    float oms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&oms, o0, o1);
    out_ms = oms;
    if (e != hipSuccess) {
      cleanup();
      return hip_fail(e, "msbfs output copy", __FILE__, __LINE__);
    }
  }

  if (stats) {
    stats->rounds_completed = hc.rounds;
    stats->total_ms = ms;
    stats->output_ms = out_ms;
    stats->visited_count = -1;
    stats->m_edges = -1;
    int64_t nh = hc.rounds < kMsHist ? hc.rounds : kMsHist;
    if (stats->max_hist < nh) nh = stats->max_hist;
    if (nh > 0) {
      std::vector<char> hb(sizeof(MsHist));
      e = hipMemcpy(hb.data(), A.hist, sizeof(MsHist), hipMemcpyDeviceToHost);
      if (e != hipSuccess) {
        cleanup();
        return hip_fail(e, "msbfs stats", __FILE__, __LINE__);
      }
      const MsHist *H = (const MsHist *)hb.data();
      for (int64_t r = 0; r < nh; r++) {
        if (stats->nf_history) stats->nf_history[r] = H->nf[r];
        if (stats->direction_history) stats->direction_history[r] = H->dir[r];
        if (stats->edges_scanned) stats->edges_scanned[r] = H->scanned[r];
        if (stats->rows_scanned) stats->rows_scanned[r] = H->rows[r];
        if (stats->pairs_found) stats->pairs_found[r] = H->pairs[r];
        if (stats->level_ms && r < (int64_t)rev.size()) {
          float rm = 0;
          hipEventElapsedTime(&rm, rev[r].first, rev[r].second);
          stats->level_ms[r] = rm;
        }
      }
    }
  }
  if (m_lanes) {
    unsigned long long m[64];
    ms_zero<<<1, 64, 0, st>>>(ws.mlane, 64);
    ms_count<<<grid_blocks(g->device, 2), kMsBlock, 0, st>>>(ws.vis, md->slot_of, g->m_weight, g->n,
                                                             nroots, ws.mlane);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(m, ws.mlane, sizeof(m), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      cleanup();
      return hip_fail(e, "msbfs m count", __FILE__, __LINE__);
    }
    for (int l = 0; l < nroots; l++) m_lanes[l] = (int64_t)m[l];
  }
  cleanup();
  return CGBFS_OK;
}
