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
#include <algorithm>
#include <cstring>
#include <vector>

#include "cg_common.h"

namespace cg {

constexpr int kMsBlock = 256;
constexpr int kMsWaves = kMsBlock / kWave;
constexpr int kMsHist = 4096;
constexpr int64_t kSeg = 1024;      // segment length and wavefront-tier cap

int grid_blocks(int device, int mult);
void fill_options(const cgbfs_options *opt, int64_t n, cgbfs_options *o);

struct MsCtrl {
  unsigned long long qtail[2];
  unsigned long long mf;       // out-degree sum of the next active set
  unsigned long long scanned;  // adjacency entries read this round
  unsigned long long rows;     // adjacency lists opened this round
  unsigned long long pairs;    // (lane, vertex) pairs reached this round
  long long prev_nf;
  long long tau;
  long long max_levels;
  long long rounds;
  double alpha, beta;
  long long nnz;
  int dir;
  int mode;
  int done;
  int status;
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

static int ensure_wsm(DeviceGraph *g, MsWorkspace *w, MsGraphData **mdp) {
  const int64_t n = g->n;
  if (!g->msdata) {
    MsGraphData *md = new MsGraphData();
    int rc = build_seglists(g, md);
    if (rc) {
      if (md->mem) hipFree(md->mem);
      delete md;
      return rc;
    }
    g->msdata = md;
    g->msdata_free = [](void *q) {
      MsGraphData *m = (MsGraphData *)q;
      if (m->mem) hipFree(m->mem);
      delete m;
    };
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

// Writes parents/levels for lanes `lanes` of vertex v whose parent is u.
__device__ __forceinline__ void ms_emit(const MsArgs &A, uint64_t lanes, int32_t u, int64_t v,
                                        int16_t lev) {
  while (lanes) {
    int l = __ffsll((unsigned long long)lanes) - 1;
    lanes &= lanes - 1;
    int64_t o = (int64_t)l * A.n + v;
    A.parents[o] = u;
    A.levels[o] = lev;
  }
}

// Wavefront scan of in-neighbours [beg, end) of v for lanes `need`
// (wave-uniform arguments).  Emits first hits when `emit`, else offers them
// to cand[] with atomicMin.  Returns the lanes found (wave-uniform).
__device__ __forceinline__ uint64_t ms_wave_pull(const MsArgs &A, const uint64_t *fr, int64_t v,
                                                 int64_t beg, int64_t end, uint64_t need,
                                                 int16_t lev, uint32_t *cand,
                                                 unsigned long long &scanned) {
  const int lane = lane_id();
  uint64_t got = 0;
  for (int64_t e0 = beg; e0 < end; e0 += kWave) {
    int64_t e = e0 + lane;
    uint64_t h = 0;
    int32_t u = 0;
    if (e < end) {
      u = A.in_idx[e];
      h = fr[u] & need & ~got;
    }
    uint64_t inc = wave_scan_or(h);
    uint64_t before = shfl_up64(inc, 1);
    if (lane == 0) before = 0;
    uint64_t mine = h & ~before;
    if (mine) {
      if (cand) {
        uint64_t b = mine;
        while (b) {
          int l = __ffsll((unsigned long long)b) - 1;
          b &= b - 1;
          atomicMin(&cand[l], (uint32_t)u);
        }
      } else {
        ms_emit(A, mine, u, v, lev);
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

__global__ void ms_init_roots(MsArgs A, const int64_t *roots, int mode, double alpha, double beta,
                              long long tau, long long max_levels, long long nnz) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned long long nq = 0, mf = 0;
  for (int l = 0; l < A.nroots; l++) {
    int64_t r = roots[l];
    uint64_t bit = 1ull << l;
    if (A.fw[0][r] == 0) {
      A.q[0][nq++] = (int32_t)r;
      mf += (unsigned long long)(A.out_ptr[r + 1] - A.out_ptr[r]);
    }
    A.vis[r] |= bit;
    A.fw[0][r] |= bit;
    A.parents[(int64_t)l * A.n + r] = (int32_t)r;
    A.levels[(int64_t)l * A.n + r] = 0;
  }
  MsCtrl *c = A.ctrl;
  c->qtail[0] = nq;
  c->qtail[1] = 0;
  c->mf = c->scanned = c->rows = c->pairs = 0;
  c->prev_nf = (long long)nq;
  c->tau = tau;
  c->max_levels = max_levels;
  c->rounds = 0;
  c->alpha = alpha;
  c->beta = beta;
  c->nnz = nnz;
  c->mode = mode;
  int dir = 0;
  if (mode == CGBFS_DIR_BOTTOMUP) dir = 1;
  else if (mode == CGBFS_DIR_BEAMER && (double)mf * alpha > (double)nnz) dir = 1;
  c->dir = dir;
  c->done = 0;
  c->status = 0;
}

// Pull round over non-big vertices (thread and wavefront tiers), then the
// big-vertex segments (one wavefront per segment, per-lane atomicMin into
// cand[], lanes into bgot[]).
__global__ void __launch_bounds__(kMsBlock) ms_pull(MsArgs A, int d) {
  MsCtrl *c = A.ctrl;
  __shared__ unsigned long long red[kMsWaves];
  const uint64_t *fr = A.fw[d & 1];
  uint64_t *nx = A.fw[(d + 1) & 1];
  int32_t *QN = A.q[(d + 1) & 1];
  unsigned long long *tail = &c->qtail[(d + 1) & 1];
  const int16_t lev = (int16_t)(d + 1);
  unsigned long long scanned = 0, mf = 0, rows = 0, pairs = 0;
  const int lane = lane_id();
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t wave = gtid / kWave, nwaves = stride / kWave;
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
    int32_t len = 0;
    if (v < A.n && !A.pin.is_big[v]) {
      vw = A.vis[v];
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
    for (int32_t base = 0; base < total; base += kWave) {
      const int32_t ve = base + lane;
      // row r: the largest lane with vstart <= ve (rows of length 0 share
      // the next row's start, so r always has entries)
      int r = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1) {
        int cand = r + step;
        int32_t sv = __shfl(vstart, cand);
        if (sv <= ve) r = cand;
      }
      const int32_t rs = __shfl(vstart, r);
      const int64_t rbase = __shfl(rb, r);
      const uint64_t nr = shfl64(need, r);
      const uint64_t gr = shfl64(got, r);
      uint64_t h = 0;
      int32_t u = 0;
      if (ve < total && (nr & ~gr)) {
        u = A.in_idx[rbase + (ve - rs)];
        h = fr[u] & nr & ~gr;
        scanned++;
      }
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
      if (mine) ms_emit(A, mine, u, ch * kWave + r, lev);
      // each row owner collects its row's OR over this step
      const int32_t my_end = vstart + len;
      int last = (my_end < base + kWave ? my_end : base + kWave) - 1 - base;
      bool hit_step = len && vstart < base + kWave && my_end > base;
      uint64_t tot = shfl64(x, hit_step ? last : lane);
      if (hit_step) got |= tot;
      if (!__any(len && (need & ~got))) break;  // every row of the chunk satisfied
    }
    bool any = got != 0;
    if (any) {
      nx[v] = got;
      A.vis[v] = vw | got;
      mf += (unsigned long long)(A.out_ptr[v + 1] - A.out_ptr[v]);
      pairs += (unsigned long long)__popcll(got);
    }
    int64_t slot = wave_append(tail, any);
    if (any) QN[slot] = (int32_t)v;
  }
  // Big vertices: one segment per wavefront.
  for (int64_t s = wave; s < A.pin.nseg; s += nwaves) {
    int32_t v = A.pin.seg_v[s];
    uint64_t need = A.act & ~A.vis[v];
    if (!need) continue;
    int32_t bid = A.pin.seg_bid[s];
    int64_t beg = A.pin.seg_beg[s];
    int64_t rend = A.in_ptr[v + 1];
    int64_t end = beg + kSeg < rend ? beg + kSeg : rend;
    uint64_t gj = ms_wave_pull(A, fr, v, beg, end, need, lev, &A.cand[(int64_t)bid * 64], scanned);
    if (gj && lane == 0) atomicOr((unsigned long long *)&A.bgot[bid], (unsigned long long)gj);
  }
  unsigned long long s = ms_block_sum(scanned, red);
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
  int32_t *QN = A.q[(d + 1) & 1];
  unsigned long long *tail = &c->qtail[(d + 1) & 1];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t nround = ((A.pin.nbig + stride - 1) / stride) * stride;
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nround; b += stride) {
    uint64_t got = 0;
    int32_t v = 0;
    if (b < A.pin.nbig) {
      v = A.pin.big_v[b];
      got = A.bgot[b];
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
        nx[v] = got;
        A.vis[v] |= got;
        atomicAdd(&c->mf, (unsigned long long)(A.out_ptr[v + 1] - A.out_ptr[v]));
        atomicAdd(&c->pairs, (unsigned long long)__popcll(got));
      }
    }
    int64_t slot = wave_append(tail, got != 0);
    if (got) QN[slot] = v;
  }
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
  const int64_t nq = (int64_t)c->qtail[d & 1];
  const int lane = lane_id();
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t wave = gtid / kWave, nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  unsigned long long scanned = 0, rows = 0;
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
    uint64_t f = fr[u];
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
        if (phase == 0) {
          uint64_t nv = f & ~A.vis[w];
          if (nv) {
            uint64_t old = atomicOr((unsigned long long *)&nx[w], (unsigned long long)nv);
            first = (old == 0);
          }
        } else {
          uint64_t nv = f & nx[w];
          while (nv) {
            int l = __ffsll((unsigned long long)nv) - 1;
            nv &= nv - 1;
            atomicMin((uint32_t *)&A.parents[(int64_t)l * A.n + w], (uint32_t)u);
          }
        }
      }
      if (phase == 0) {
        int64_t slot = wave_append(tail, first);
        if (first) QN[slot] = w;
      }
    }
  }
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
  const int64_t nq = (int64_t)c->qtail[(d + 1) & 1];
  unsigned long long mf = 0, pairs = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq;
       i += (int64_t)gridDim.x * blockDim.x) {
    int32_t v = QN[i];
    uint64_t nw = nx[v];
    uint64_t lanes = nw;
    while (lanes) {
      int l = __ffsll((unsigned long long)lanes) - 1;
      lanes &= lanes - 1;
      int64_t o = (int64_t)l * A.n + v;
      A.parents[o] = -1;  // 0xFFFFFFFF: +inf for the unsigned atomicMin
      A.levels[o] = (int16_t)(d + 1);
    }
    A.vis[v] |= nw;
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
  const int64_t nq = (int64_t)c->qtail[d & 1];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq;
       i += (int64_t)gridDim.x * blockDim.x)
    fr[Q[i]] = 0;
}

__global__ void ms_decide(MsArgs A, int d) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  MsCtrl *c = A.ctrl;
  const long long nf = (long long)c->qtail[(d + 1) & 1];
  if (d < kMsHist) {
    A.hist->nf[d] = nf;
    A.hist->scanned[d] = (long long)c->scanned;
    A.hist->rows[d] = (long long)c->rows;
    A.hist->pairs[d] = (long long)c->pairs;
    A.hist->dir[d] = c->dir;
  }
  c->rounds = d + 1;
  c->qtail[d & 1] = 0;
  const long long mf = (long long)c->mf;
  c->mf = c->scanned = c->rows = c->pairs = 0;
  if (nf == 0) {
    c->done = 1;
    return;
  }
  int dir = c->dir;
  switch (c->mode) {
    case CGBFS_DIR_TOPDOWN: dir = 0; break;
    case CGBFS_DIR_BOTTOMUP: dir = 1; break;
    case CGBFS_DIR_TAU:
      if (dir == 0 && nf > c->tau) dir = 1;
      break;
    default:
      if (dir == 0) {
        if ((double)mf * c->alpha > (double)c->nnz) dir = 1;
      } else {
        if ((double)nf * c->beta < (double)A.n && nf < c->prev_nf) dir = 0;
      }
  }
  c->prev_nf = nf;
  c->dir = dir;
  if (d + 1 >= c->max_levels) {
    c->status = CGBFS_ETIMEOUT;
    c->done = 1;
  }
}

// Final pass: -1 into every (lane, vertex) pair no lane word reached.
__global__ void __launch_bounds__(kMsBlock) ms_finalize(MsArgs A) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < A.n;
       v += (int64_t)gridDim.x * blockDim.x) {
    uint64_t miss = A.act & ~A.vis[v];
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
__global__ void __launch_bounds__(kMsBlock) ms_count(const uint64_t *vis, const int32_t *w,
                                                     int64_t n, int nroots,
                                                     unsigned long long *out) {
  __shared__ unsigned long long acc[64];
  if (threadIdx.x < 64) acc[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    uint64_t b = vis[v];
    unsigned long long wv = (unsigned long long)w[v];
    if (!wv) continue;
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
  ms_init_roots<<<1, 64, 0, st>>>(A, ws.roots, o.direction, o.alpha, o.beta, (long long)o.tau,
                                  (long long)o.max_levels, (long long)g->nnz);
  MsCtrl hc;
  e = hipMemcpyAsync(&hc, A.ctrl, sizeof(MsCtrl), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  // One host poll per round: only the chosen direction's kernels launch,
  // and each round's traversal kernels are bracketed by a pair of events.
  std::vector<std::pair<hipEvent_t, hipEvent_t>> rev;
  std::vector<int> rdir;
  int64_t d = 0;
  while (e == hipSuccess && !hc.done) {
    hipEvent_t ra, rb;
    if ((e = new_event(&ra)) != hipSuccess || (e = new_event(&rb)) != hipSuccess) break;
    hipEventRecord(ra, st);
    if (hc.dir == 1) {
      ms_pull<<<blocks, kMsBlock, 0, st>>>(A, (int)d);
      if (md->pin.nbig) ms_pull_fin<<<(unsigned)((md->pin.nbig + kMsBlock - 1) / kMsBlock), kMsBlock, 0, st>>>(A, (int)d);
    } else {
      ms_push<<<blocks, kMsBlock, 0, st>>>(A, (int)d, 0);
      ms_prep<<<blocks, kMsBlock, 0, st>>>(A, (int)d);
      ms_push<<<blocks, kMsBlock, 0, st>>>(A, (int)d, 1);
    }
    hipEventRecord(rb, st);
    rev.push_back({ra, rb});
    rdir.push_back(hc.dir);
    ms_clear<<<blocks, kMsBlock, 0, st>>>(A, (int)d);
    ms_decide<<<1, 64, 0, st>>>(A, (int)d);
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
    ms_count<<<grid_blocks(g->device, 2), kMsBlock, 0, st>>>(ws.vis, g->m_weight, g->n, nroots, ws.mlane);
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
