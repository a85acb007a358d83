// Multi-source BFS for gfx950: up to 64 roots packed as the bits of one
// uint64 word per vertex, so one adjacency pass serves every root.
//
// Per vertex v the library keeps three lane words: vis[v] (lanes that have
// reached v), and the frontier pair fr[v] / nx[v] (lanes for which v is in
// the current / next level).  Every lane follows exactly the single-root
// semantics of bfs.hip (and of bool_pe.csl's round loop): lane l's level is
// its hop distance from roots[l] and its parent is the lowest-id
// in-neighbour one level closer.
//
// A round is one of:
//   * push: active vertices (queue) OR their frontier word into nx[w] of
//     every out-neighbour w, masked by the level-start vis[w]; the first
//     thread to make nx[w] non-zero appends w to the next queue.  Parents are
//     then resolved by a pull restricted to that queue (no per-lane atomics):
//     each newly reached vertex scans its ascending in-neighbours and gives
//     each new lane the first neighbour whose frontier word has that lane;
//   * pull: every vertex with an unreached lane scans its ascending
//     in-neighbours, ORing frontier words, and stops as soon as every
//     unreached lane has been found.
// Outputs parents[l][v] (int32) and levels[l][v] (int16) are written once
// per (lane, vertex): at discovery, or by the final pass for unreached
// pairs.  No initial fill of the 6 B x 64 x V output arrays is needed.
#include <cstring>
#include <vector>

#include "cg_common.h"

namespace cg {

constexpr int kMsBlock = 256;
constexpr int kMsWaves = kMsBlock / kWave;
constexpr int kMsHist = 4096;

int grid_blocks(int device, int mult);
double wall_clock_hz(int device);
void fill_options(const cgbfs_options *opt, int64_t n, cgbfs_options *o);

struct MsCtrl {
  unsigned long long qtail[2];
  unsigned long long mf;       // out-degree sum of the next active set
  unsigned long long scanned;  // adjacency entries read this round
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
  unsigned long long t0, t_end;
};

struct MsHist {
  long long nf[kMsHist];
  long long scanned[kMsHist];
  unsigned long long t[kMsHist];
  int dir[kMsHist];
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
  int64_t n;
  uint64_t act;      // mask of lanes in use
  int nroots;
};

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

struct MsWorkspace {
  uint64_t *vis, *fw[2];
  int32_t *q[2];
  MsCtrl *ctrl;
  MsHist *hist;
  unsigned long long *mlane;
  int64_t *roots;
};

static int ensure_wsm(DeviceGraph *g, MsWorkspace *w) {
  const int64_t n = g->n;
  size_t wb = al(sizeof(uint64_t) * n), qb = al(sizeof(int32_t) * n);
  size_t total = 3 * wb + 2 * qb + al(sizeof(MsCtrl)) + al(sizeof(MsHist)) +
                 al(sizeof(unsigned long long) * 64) + al(sizeof(int64_t) * 64);
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
  w->roots = (int64_t *)p;
  return CGBFS_OK;
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

__global__ void __launch_bounds__(kMsBlock) ms_init_fill(MsArgs A) {
  int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = tid; v < A.n; v += stride) {
    A.vis[v] = 0;
    A.fw[0][v] = 0;
    A.fw[1][v] = 0;
  }
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
  c->mf = 0;
  c->scanned = 0;
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
  c->t0 = wall_clock64();
  c->t_end = c->t0;
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

// Scans v's ascending in-neighbours for the lanes in `need`; emits
// (lane, first neighbour with that lane).  Returns the lanes found.
__device__ __forceinline__ uint64_t ms_pull_vertex(const MsArgs &A, const uint64_t *fr, int64_t v,
                                                   uint64_t need, int16_t lev,
                                                   unsigned long long &scanned) {
  uint64_t got = 0;
  int64_t beg = A.in_ptr[v], end = A.in_ptr[v + 1];
  for (int64_t e = beg; e < end; e++) {
    int32_t u = A.in_idx[e];
    uint64_t h = fr[u] & need & ~got;
    scanned++;
    if (h) {
      ms_emit(A, h, u, v, lev);
      got |= h;
      if (got == need) break;
    }
  }
  return got;
}

// Push: OR frontier words into the next words of out-neighbours.
__global__ void __launch_bounds__(kMsBlock) ms_push(MsArgs A, int d) {
  MsCtrl *c = A.ctrl;
  if (c->done || c->dir != 0) return;
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
  unsigned long long scanned = 0;
  // One wavefront per active vertex: edges strided over the 64 lanes.
  for (int64_t i = wave; i < nq; i += nwaves) {
    int32_t u = Q[i];
    uint64_t f = fr[u];
    int64_t beg = A.out_ptr[u], end = A.out_ptr[u + 1];
    if (lane == 0) scanned += (unsigned long long)(end - beg);
    for (int64_t e0 = beg; e0 < end; e0 += kWave) {
      int64_t e = e0 + lane;
      bool first = false;
      int32_t w = 0;
      if (e < end) {
        w = A.out_idx[e];
        uint64_t nv = f & ~A.vis[w];
        if (nv) {
          uint64_t old = atomicOr((unsigned long long *)&nx[w], (unsigned long long)nv);
          first = (old == 0);
        }
      }
      int64_t slot = wave_append(tail, first);
      if (first) QN[slot] = w;
    }
  }
  unsigned long long s = ms_block_sum(scanned, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->scanned, s);
}

// Parent/level resolution for the vertices a push round reached.
__global__ void __launch_bounds__(kMsBlock) ms_resolve(MsArgs A, int d) {
  MsCtrl *c = A.ctrl;
  if (c->done || c->dir != 0) return;
  __shared__ unsigned long long red[kMsWaves];
  const uint64_t *fr = A.fw[d & 1];
  const uint64_t *nx = A.fw[(d + 1) & 1];
  const int32_t *QN = A.q[(d + 1) & 1];
  const int64_t nq = (int64_t)c->qtail[(d + 1) & 1];
  unsigned long long scanned = 0, mf = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nq;
       i += (int64_t)gridDim.x * blockDim.x) {
    int32_t v = QN[i];
    uint64_t nw = nx[v];
    ms_pull_vertex(A, fr, v, nw, (int16_t)(d + 1), scanned);
    A.vis[v] |= nw;
    mf += (unsigned long long)(A.out_ptr[v + 1] - A.out_ptr[v]);
  }
  unsigned long long s = ms_block_sum(scanned, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->scanned, s);
  s = ms_block_sum(mf, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->mf, s);
}

// Pull: every vertex with an unreached lane looks for frontier neighbours.
__global__ void __launch_bounds__(kMsBlock) ms_pull(MsArgs A, int d) {
  MsCtrl *c = A.ctrl;
  if (c->done || c->dir != 1) return;
  __shared__ unsigned long long red[kMsWaves];
  const uint64_t *fr = A.fw[d & 1];
  uint64_t *nx = A.fw[(d + 1) & 1];
  int32_t *QN = A.q[(d + 1) & 1];
  unsigned long long *tail = &c->qtail[(d + 1) & 1];
  unsigned long long scanned = 0, mf = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // Whole wavefronts iterate together so wave_append stays convergent.
  const int64_t nround = ((A.n + stride - 1) / stride) * stride;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nround; v += stride) {
    uint64_t got = 0, vw = 0;
    if (v < A.n) {
      vw = A.vis[v];
      uint64_t need = A.act & ~vw;
      if (need) got = ms_pull_vertex(A, fr, v, need, (int16_t)(d + 1), scanned);
    }
    bool any = got != 0;
    if (any) {
      nx[v] = got;
      A.vis[v] = vw | got;
      mf += (unsigned long long)(A.out_ptr[v + 1] - A.out_ptr[v]);
    }
    int64_t slot = wave_append(tail, any);
    if (any) QN[slot] = (int32_t)v;
  }
  unsigned long long s = ms_block_sum(scanned, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->scanned, s);
  s = ms_block_sum(mf, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->mf, s);
}

// Clears the level-d frontier words (listed in Q[d]) so that buffer is zero
// when it becomes the next-words buffer of round d+1.
__global__ void __launch_bounds__(kMsBlock) ms_clear(MsArgs A, int d) {
  MsCtrl *c = A.ctrl;
  if (c->done) return;
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
  if (c->done) return;
  const long long nf = (long long)c->qtail[(d + 1) & 1];
  unsigned long long now = wall_clock64();
  if (d < kMsHist) {
    A.hist->nf[d] = nf;
    A.hist->scanned[d] = (long long)c->scanned;
    A.hist->dir[d] = c->dir;
    A.hist->t[d] = now;
  }
  c->rounds = d + 1;
  c->t_end = now;
  c->qtail[d & 1] = 0;
  const long long mf = (long long)c->mf;
  c->mf = 0;
  c->scanned = 0;
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
  int rc = ensure_wsm(g, &ws);
  if (rc) return rc;
  hipStream_t st = g->stream;

  const size_t plane = (size_t)g->n * nroots;
  int32_t *dpar = parents;
  int16_t *dlev = levels;
  void *tmp = nullptr;
  if (!on_device) {
    CG_HIP(hipMalloc(&tmp, plane * (sizeof(int32_t) + sizeof(int16_t)) + 256));
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
  A.n = g->n;
  A.nroots = nroots;
  A.act = nroots == 64 ? ~0ull : ((1ull << nroots) - 1ull);

  const int blocks = grid_blocks(g->device, o.grid_mult);
  hipError_t e = hipMemcpyAsync(ws.roots, roots, sizeof(int64_t) * nroots, hipMemcpyHostToDevice, st);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  if (e != hipSuccess) {
    if (tmp) hipFree(tmp);
    return hip_fail(e, "msbfs setup", __FILE__, __LINE__);
  }
  ms_init_fill<<<blocks, kMsBlock, 0, st>>>(A);
  ms_init_roots<<<1, 64, 0, st>>>(A, ws.roots, o.direction, o.alpha, o.beta, (long long)o.tau,
                                  (long long)o.max_levels, (long long)g->nnz);
  int64_t d = 0;
  int chunk = o.rounds_hint > 0 ? o.rounds_hint : 12;
  MsCtrl hc;
  for (;;) {
    for (int k = 0; k < chunk; k++, d++) {
      ms_push<<<blocks, kMsBlock, 0, st>>>(A, (int)d);
      ms_resolve<<<blocks, kMsBlock, 0, st>>>(A, (int)d);
      ms_pull<<<blocks, kMsBlock, 0, st>>>(A, (int)d);
      ms_clear<<<blocks, kMsBlock, 0, st>>>(A, (int)d);
      ms_decide<<<1, 64, 0, st>>>(A, (int)d);
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&hc, A.ctrl, sizeof(MsCtrl), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) break;
    if (hc.done) break;
    if (d > o.max_levels + chunk) break;
    chunk = 4;
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
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (e != hipSuccess || hc.status) {
    if (tmp) hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "msbfs rounds", __FILE__, __LINE__);
    return set_error(hc.status, "MS-BFS exceeded max_levels (int16 levels cap at 32767)");
  }

  double out_ms = 0;
  if (!on_device) {
    hipEvent_t o0, o1;
    hipEventCreate(&o0);
    hipEventCreate(&o1);
    hipEventRecord(o0, st);
    e = hipMemcpyAsync(parents, dpar, plane * sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(levels, dlev, plane * sizeof(int16_t), hipMemcpyDeviceToHost, st);
    hipEventRecord(o1, st);
    if (e == hipSuccess) e = hipEventSynchronize(o1);
    float oms = 0;
    hipEventElapsedTime(&oms, o0, o1);
    out_ms = oms;
    hipEventDestroy(o0);
    hipEventDestroy(o1);
    hipFree(tmp);
    if (e != hipSuccess) return hip_fail(e, "msbfs output copy", __FILE__, __LINE__);
  }

  if (stats) {
    const double hz = wall_clock_hz(g->device);
    stats->rounds_completed = hc.rounds;
    stats->total_ms = ms;
    stats->output_ms = out_ms;
    stats->visited_count = -1;
    stats->m_edges = -1;
    int64_t nh = hc.rounds < kMsHist ? hc.rounds : kMsHist;
    if (stats->max_hist < nh) nh = stats->max_hist;
    if (nh > 0) {
      std::vector<char> hb(sizeof(MsHist));
      CG_HIP(hipMemcpy(hb.data(), A.hist, sizeof(MsHist), hipMemcpyDeviceToHost));
      const MsHist *H = (const MsHist *)hb.data();
      for (int64_t r = 0; r < nh; r++) {
        if (stats->nf_history) stats->nf_history[r] = H->nf[r];
        if (stats->direction_history) stats->direction_history[r] = H->dir[r];
        if (stats->edges_scanned) stats->edges_scanned[r] = H->scanned[r];
        if (stats->level_ms) {
          unsigned long long prev = r == 0 ? hc.t0 : H->t[r - 1];
          stats->level_ms[r] = (double)(H->t[r] - prev) * 1e3 / hz;
        }
      }
    }
  }
  if (m_lanes) {
    unsigned long long m[64];
    ms_zero<<<1, 64, 0, st>>>(ws.mlane, 64);
    ms_count<<<grid_blocks(g->device, 2), kMsBlock, 0, st>>>(ws.vis, g->m_weight, g->n, nroots, ws.mlane);
    CG_HIP(hipGetLastError());
    CG_HIP(hipMemcpyAsync(m, ws.mlane, sizeof(m), hipMemcpyDeviceToHost, st));
    CG_HIP(hipStreamSynchronize(st));
    for (int l = 0; l < nroots; l++) m_lanes[l] = (int64_t)m[l];
  }
  return CGBFS_OK;
}
