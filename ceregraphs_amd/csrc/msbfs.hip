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
// Directions: by default (CGBFS_DIR_BEAMER) every lane moves together, by
// Beamer's rule on the union counts of the round.  The per-lane policies
// (CGBFS_DIR_BEAMER_LANES, CGBFS_DIR_TAU) run the single-root rule on each
// lane's own counts (its frontier size and degree sums, tallied per lane
// bit in the kernels), and such a round first pulls the lanes in pull mode,
// then pushes the lanes in push mode.  On R-MAT graphs the mixed rounds
// cost more than the rescans they save, so the union rule is the default.
// Lanes with an empty frontier take no part.
// Pull part (ms_pull_segs + ms_pull): every vertex with an unreached pull
// lane ORs the frontier words of its ascending in-neighbours; within a step
// of 64 neighbours an exclusive prefix-OR across the wavefront gives each
// lane of each vertex its first (lowest-id) hit.  Stops once every needed
// lane is found.
// Push part (ms_push x2 + ms_prep): pass 1 ORs the push lanes of frontier
// words into nx[w] of out-neighbours (masked by the level-start vis[w]); the
// first writer of push bits of a vertex appends it to the next queue.
// ms_prep then marks the new (lane, vertex) pairs, and pass 2 offers every
// frontier predecessor with a per-lane atomicMin, so the parent is again
// the lowest id.
// Outputs parents[l][v] (int32) and levels[l][v] (int16) are written once
// per (lane, vertex): at discovery, or by the final pass for unreached
// pairs.  No initial fill of the 6 B x 64 x V output arrays is needed.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <type_traits>
#include <vector>

#include "cg_common.h"

namespace cg {

constexpr int kMsBlock = 256;
constexpr int kMsWaves = kMsBlock / kWave;
constexpr int kMsHist = 4096;
constexpr int64_t kSeg = 256;       // segment length and wavefront-tier cap
constexpr int kMsLdsQueue = 2048;   // per-block staged queue entries
#ifndef CG_MS_UNROLL
#define CG_MS_UNROLL 2
#endif
constexpr int kUnroll = CG_MS_UNROLL;  // pull steps whose loads are issued together
#ifndef CG_MS_FLUSH
#define CG_MS_FLUSH 4
#endif
constexpr int kFlush = CG_MS_FLUSH;  // lane bits per batch of the pull's output flush
constexpr int kGraphRounds = 8;     // rounds per captured graph (graph replay mode)
#ifndef CG_MS_DENSE_GRAB
#define CG_MS_DENSE_GRAB 1
#endif
constexpr int32_t kGrabDense = CG_MS_DENSE_GRAB;  // chunks per grab of a dense pull
// Which pull round of a search writes the outputs of the vertices without
// in-edges (-1 in every lane but their own seeds'), in the same store
// instructions as its own new pairs.
#ifndef CG_MS_NOIN_PULL
#define CG_MS_NOIN_PULL 2
#endif
constexpr int kNoinPull = CG_MS_NOIN_PULL;
// Waves per SIMD the chunk pull is compiled for (its register budget is
// 512 / kMsPullWaves VGPRs).
#ifndef CG_MS_PULL_WAVES
#define CG_MS_PULL_WAVES 6
#endif
constexpr int kMsPullWaves = CG_MS_PULL_WAVES;
#ifndef CG_MS_PUSH_FILTER
#define CG_MS_PUSH_FILTER 1
#endif
#ifndef CG_MS_PUSH_FILTER2
#define CG_MS_PUSH_FILTER2 1
#endif
constexpr bool kPushFilter = CG_MS_PUSH_FILTER;    // push phase 0: read the word before atomicOr
constexpr bool kPushFilter2 = CG_MS_PUSH_FILTER2;  // push phase 1: read the parent before atomicMin
// First-hit offsets within a wavefront-tier row (< kSeg) fit in a byte.
using TileOff = std::conditional_t<(kSeg <= 256), uint8_t, uint16_t>;

int grid_blocks(int device, int mult);
void fill_options(const cgbfs_options *opt, int64_t n, bool multi_source, cgbfs_options *o);
int check_options(const cgbfs_options *o);

// Round counters come in two parts: the push part (nf .. pairs) and the
// pull part (pnf .. ppairs).  nf counts only vertices the pull did not
// reach, so nf + pnf is the number of vertices with a new lane.
// The round counters and tails each take a 128-byte line: every block of a
// round kernel adds to them (or appends through them) once, and device-scope
// atomics on one line are serialised, with each other and with the reads of
// the read-mostly fields (the last line) that every block starts with.
struct MsCtrl {
  LineCtr qtail[2];
  alignas(128) unsigned long long nf;
  alignas(128) unsigned long long scanned;  // adjacency entries read this round
  alignas(128) unsigned long long rows;     // adjacency lists opened this round
  alignas(128) unsigned long long pairs;    // (lane, vertex) pairs reached this round
  alignas(128) unsigned long long pnf;
  alignas(128) unsigned long long pscanned;
  alignas(128) unsigned long long prows;
  alignas(128) unsigned long long ppairs;
  alignas(128) unsigned long long mf;       // out-degree sum of the vertices with a new lane
  alignas(128) unsigned long long live;     // lanes that reached a vertex this round
  LineCtr segtail[2];  // active segments this round: pull, push
  alignas(128) unsigned long long segtail_late;  // active pull segments past their vertex's first run
  alignas(128) unsigned long long pmask;    // lanes the next round pulls
  unsigned long long qmask;    // lanes the next round pushes
  long long prev_nf;
  long long rounds;            // the round the next kernels work on
  int gdir;                    // union-count policy: 0 push, 1 pull
  int listed;                  // that round's frontier is listed in q[rounds & 1]
  int use_fb;                  // that round's pull reads the frontier bits
  int done;
  int status;
  int grab;                    // chunks per grab of that round's pull; < 0: a static split
  // pull rounds until the one that writes the outputs of the vertices
  // without in-edges (1: the next pull round); 0 once they are written
  int noin_left;
  // Entries of the previous round's new-vertex list (q[rounds & 1]) whose
  // push pairs wait in the side table for their outputs (0: none).
  long long pend_n;
};

// Every round kernel reads its round from the control block (the previous
// ms_decide wrote it), so a fixed sequence of launches serves any round and
// can be captured once into a HIP graph; a kernel whose part of the round
// is off (no pull lanes, no push lanes, ...) returns at once.
struct RoundState {
  int d;
  uint64_t pm, qm;
  bool listed, use_fb;
  int grab;
  bool noin;  // this round's pull also writes the outputs of vertices without in-edges
  int64_t pend_n;  // previous round's push pairs waiting for their outputs (list length)
};

// Per-lane direction state.  cnt/out/in are this round's tallies of the
// vertices each lane newly reached (count, out-degree sum, in-degree sum,
// degrees clamped at kDegClamp); mu is the in-degree sum of the lane's
// unreached vertices, prev the lane's previous count, dir 0 push / 1 pull.
struct MsLane {
  unsigned long long cnt[64], out[64], in[64];
  long long mu[64], prev[64];
  int dir[64];
};
constexpr uint32_t kDegClamp = 1u << 25;  // 64 clamped degrees sum below 2^32

// Where lane l's outputs go: parents par[l][v] (int32) and levels lev[l][v]
// (int16), the level of a vertex reached in round d being d + 1 - base[l].
// A one-call search points lane l at row l of [nroots][n] arrays with base
// 0; a serving engine points each lane at its current query's own buffers
// and sets base to the round the query was seeded in.
struct LaneOut {
  int32_t *par[64];
  int16_t *lev[64];
  int32_t base[64];
};

// Per-search constants, passed by value to the kernels that need them.
struct MsParams {
  int mode;
  double alpha, beta;
  long long tau, max_levels, nnz;
  int dense_grab;  // chunks per grab of a dense pull round
};

// Per-round statistics, one row per round, so a search's rounds are one
// contiguous copy.
struct MsHistRow {
  long long nf, scanned, rows, pairs;
  long long pnf, pscanned, prows, ppairs;
  long long plane;  // lanes pulled
  int dir;          // 0 push only, 1 pull only, 2 both
  int pad;
};
struct MsHist {
  MsHistRow r[kMsHist];
};

// Static segment list of one adjacency view (per graph).
struct SegList {
  int64_t nseg = 0;
  int64_t nbig = 0;
  int32_t *seg_v = nullptr;    // vertex of each segment
  int32_t *seg_bid = nullptr;  // big-vertex index of each segment
  int64_t *seg_beg = nullptr;  // first adjacency entry of each segment
  int32_t *seg_len = nullptr;  // its length (<= kSeg)
  int32_t *big_v = nullptr;    // big vertices, by index
  int32_t *seg_first = nullptr; // [nbig + 1] first segment of each big vertex
  uint8_t *is_big = nullptr;   // per vertex
  int32_t *chunk_big = nullptr; // [n/64 + 1] big vertices below each 64-vertex chunk
};

struct MsArgs {
  const int64_t *out_ptr;
  const int32_t *out_idx;
  const int64_t *in_ptr;
  const int32_t *in_idx;
  const LaneOut *out;  // per-lane output rows and level bases
  uint64_t *vis;
  uint64_t *fw[2];   // frontier words, rotated by level parity
  // One bit per slot: the current frontier word is nonzero.  Rebuilt before
  // each pull part (ms_mark); the pull reads a frontier word only where the
  // bit is set, so most of its gathers hit this small, cached array.
  uint32_t *fb;
  int32_t *q[2];
  MsCtrl *ctrl;
  MsHist *hist;
  MsLane *lanes;
  // big-vertex segments, pull (in view) and push (out view), and this
  // round's active ones (indices into the static lists)
  SegList pin, pout;
  int32_t *segq[2];
  uint64_t *segneed;  // [pin.nseg] lanes needed by each active pull segment
  // The active pull segments past the first kSegRun of their vertex, and
  // their lanes (pulled after every first run has finished).
  int32_t *segq_late;
  uint64_t *segneed_late;
  // [3 * kParts * kCursorStride] pull cursors, zero between rounds: chunks,
  // first-run segments, later segments
  unsigned long long *cursor;
  const uint64_t *noin;  // vertices without in-edges, one bit each
  // Pinned host mailbox (graph replay), or null: [0] the done flag, raised
  // last, [1] the status, [2..3] the rounds (int64).
  int *host_done;
  // Vertex-state slot of each vertex (hub-first order) and the adjacency
  // arrays re-expressed in slots, entry for entry.  The frontier words fw,
  // which the pull gathers through in_slot, are indexed by slot, so the
  // most-gathered words share a few MB; vis and everything else is indexed
  // by vertex id, so per-vertex passes read it in order.
  const int32_t *slot_of;
  const int32_t *in_slot;
  const int32_t *out_slot;
  uint32_t *cand;    // [pin.nbig][64] per-lane first-hit candidates
  uint64_t *bgot;    // [pin.nbig] lanes found this round
  // Deferred push outputs.  A push round writes the parents of its new
  // pairs into a compact side table instead of the output rows: per new
  // vertex, a 64-bit lane mask and one candidate per lane (an even number of
  // 32-bit words, so the mask is 8-byte aligned), taken from one of 64
  // allocation counters (palloc, one 128-byte line each), each owning a
  // 64th of the table.  pend[v] is the vertex's entry, -1 for none.  The
  // next round writes the outputs with its pull's chunk stores (or with
  // ms_push_out when it does not pull), so the scattered partial lines of
  // a push's pairs are not written separately.  A vertex whose entry does
  // not fit takes the direct path: +inf parents in the output rows.
  int32_t *pend;     // [n]
  uint32_t *side;    // [side_cap]
  unsigned long long *palloc;  // [64 * 16]
  int64_t side_cap;
  int64_t n;
  uint64_t act;      // mask of lanes in use
  int nroots;
  int symmetric;     // in view == out view
  int seg_grab;      // hub-row segments per grab of ms_pull_segs
};

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

struct MsGraphData {
  SegList pin, pout;
  void *mem = nullptr;
  // One bit per vertex: no in-edges, so no root but itself ever reaches it
  // ([ceil(n/64)] words).
  uint64_t *noin = nullptr;
  int32_t *slot_of = nullptr;   // [n]
  int32_t *vertex_of = nullptr; // [n], inverse of slot_of
  int32_t *in_slot = nullptr;   // [nnz], in view
  int32_t *out_slot = nullptr;  // [nnz], out view (== in_slot when shared)
  // The last instantiated round graph (graph replay mode) and the launch
  // arguments it was captured with; reused while they stay the same.
  hipGraphExec_t exec = nullptr;
  std::vector<char> exec_key;
  // The call's lane output table, seed vertices and seed lanes, in that
  // order (the table's address stays fixed while the buffer does), and a
  // pinned host copy they are staged in, so one asynchronous copy moves them.
  void *seedbuf = nullptr;
  void *seedstage = nullptr;
  size_t seedcap = 0;
  // Device staging of host outputs (parents, then levels), kept for the next
  // call with host outputs: a 64-root plane at 2^24 vertices is 6.4 GB, and
  // allocating it per call cost more than the search.
  void *outbuf = nullptr;
  size_t outcap = 0;
  bool stage_busy = false;  // a copy from the staging buffer may be in flight
};

struct MsWorkspace {
  uint64_t *vis, *fw[2];
  uint32_t *fb;
  int32_t *q[2];
  MsCtrl *ctrl;
  MsHist *hist;
  MsLane *lanes;
  LaneOut *out;
  unsigned long long *mlane;
  uint32_t *cand;
  uint64_t *bgot;
  int32_t *segq[2];
  uint64_t *segneed;
  int32_t *segq_late;
  uint64_t *segneed_late;
  unsigned long long *cursor;
  int32_t *pend;
  uint32_t *side;
  unsigned long long *palloc;
  int64_t side_cap;
};

// Words of the deferred-output side table: room for four new pairs per
// vertex (a push round reaching every vertex in every lane falls back to
// the direct path for what does not fit), between 2^17 and 2^27 words.
static int64_t side_words(int64_t n) {
  int64_t w = 4 * n;
  if (w < (1ll << 17)) w = 1ll << 17;
  if (w > (1ll << 27)) w = 1ll << 27;
  return w & ~(int64_t)127;  // 64 regions of an even number of words
}
constexpr int kPallocStride = 16;  // u64 per allocation counter: one 128-byte line

// ---------------------------------------------------------------------------
// Per-graph segment lists (built on the host from the offset arrays once).
// ---------------------------------------------------------------------------
static void seg_plan(const std::vector<int64_t> &ptr, int64_t n, std::vector<int32_t> &sv,
                     std::vector<int32_t> &sb, std::vector<int64_t> &sg, std::vector<int32_t> &bv,
                     std::vector<int32_t> &bf, std::vector<uint8_t> &isb,
                     std::vector<int32_t> &cbig) {
  isb.assign(n, 0);
  cbig.assign((size_t)((n + 63) / 64 + 1), 0);
  for (int64_t v = 0; v < n; v++) {
    if ((v & 63) == 0) cbig[(size_t)(v >> 6)] = (int32_t)bv.size();  // big_v ascends by vertex
    int64_t deg = ptr[v + 1] - ptr[v];
    if (deg < kSeg) continue;
    int32_t bid = (int32_t)bv.size();
    bv.push_back((int32_t)v);
    bf.push_back((int32_t)sv.size());
    isb[v] = 1;
    for (int64_t b = ptr[v]; b < ptr[v + 1]; b += kSeg) {
      sv.push_back((int32_t)v);
      sb.push_back(bid);
      sg.push_back(b);
    }
  }
  bf.push_back((int32_t)sv.size());
  cbig.back() = (int32_t)bv.size();
}

static int build_seglists(DeviceGraph *g, MsGraphData *md) {
  const int64_t n = g->n;
  std::vector<int64_t> ptr_in(n + 1), ptr_out;
  CG_HIP(hipMemcpy(ptr_in.data(), g->in_ptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
  std::vector<int32_t> sv[2], sb[2], bv[2], bf[2];
  std::vector<int64_t> sg[2];
  std::vector<uint8_t> isb[2];
  std::vector<int32_t> cbig[2];
  seg_plan(ptr_in, n, sv[0], sb[0], sg[0], bv[0], bf[0], isb[0], cbig[0]);
  const bool shared = g->in_ptr == g->out_ptr;
  if (!shared) {
    ptr_out.resize(n + 1);
    CG_HIP(hipMemcpy(ptr_out.data(), g->out_ptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
    seg_plan(ptr_out, n, sv[1], sb[1], sg[1], bv[1], bf[1], isb[1], cbig[1]);
  }
  std::vector<int32_t> sl[2];
  for (int k = 0; k < (shared ? 1 : 2); k++) {
    const std::vector<int64_t> &ptr = k == 0 ? ptr_in : ptr_out;
    sl[k].resize(sv[k].size());
    for (size_t i = 0; i < sv[k].size(); i++)
      sl[k][i] = (int32_t)std::min<int64_t>(kSeg, ptr[sv[k][i] + 1] - sg[k][i]);
  }
  const int views = shared ? 1 : 2;
  size_t total = 0;
  for (int k = 0; k < views; k++)
    total += al(sv[k].size() * 4 + 4) + al(sb[k].size() * 4 + 4) + al(sg[k].size() * 8 + 8) +
             al(sl[k].size() * 4 + 4) +
             al(bv[k].size() * 4 + 4) + al(bf[k].size() * 4) + al((size_t)n) +
             al(cbig[k].size() * 4);
  CG_HIP(hipMalloc(&md->mem, total));
  char *p = (char *)md->mem;
  for (int k = 0; k < views; k++) {
    SegList &L = (k == 0) ? md->pin : md->pout;
    L.nseg = (int64_t)sv[k].size();
    L.nbig = (int64_t)bv[k].size();
    L.seg_v = (int32_t *)p; p += al(sv[k].size() * 4 + 4);
    L.seg_bid = (int32_t *)p; p += al(sb[k].size() * 4 + 4);
    L.seg_beg = (int64_t *)p; p += al(sg[k].size() * 8 + 8);
    L.seg_len = (int32_t *)p; p += al(sl[k].size() * 4 + 4);
    L.big_v = (int32_t *)p; p += al(bv[k].size() * 4 + 4);
    L.seg_first = (int32_t *)p; p += al(bf[k].size() * 4);
    L.is_big = (uint8_t *)p; p += al((size_t)n);
    L.chunk_big = (int32_t *)p; p += al(cbig[k].size() * 4);
    if (L.nseg) {
      CG_HIP(hipMemcpy(L.seg_v, sv[k].data(), sv[k].size() * 4, hipMemcpyHostToDevice));
      CG_HIP(hipMemcpy(L.seg_bid, sb[k].data(), sb[k].size() * 4, hipMemcpyHostToDevice));
      CG_HIP(hipMemcpy(L.seg_beg, sg[k].data(), sg[k].size() * 8, hipMemcpyHostToDevice));
      CG_HIP(hipMemcpy(L.seg_len, sl[k].data(), sl[k].size() * 4, hipMemcpyHostToDevice));
      CG_HIP(hipMemcpy(L.big_v, bv[k].data(), bv[k].size() * 4, hipMemcpyHostToDevice));
    }
    CG_HIP(hipMemcpy(L.seg_first, bf[k].data(), bf[k].size() * 4, hipMemcpyHostToDevice));
    CG_HIP(hipMemcpy(L.is_big, isb[k].data(), (size_t)n, hipMemcpyHostToDevice));
    CG_HIP(hipMemcpy(L.chunk_big, cbig[k].data(), cbig[k].size() * 4, hipMemcpyHostToDevice));
  }
  if (shared) md->pout = md->pin;
  {
    std::vector<uint64_t> bits((size_t)((n + 63) / 64), 0ull);
    for (int64_t v = 0; v < n; v++)
      if (ptr_in[v + 1] == ptr_in[v]) bits[(size_t)(v >> 6)] |= 1ull << (v & 63);
    CG_HIP(hipMalloc(&md->noin, sizeof(uint64_t) * (bits.size() + 1)));
    CG_HIP(hipMemcpy(md->noin, bits.data(), sizeof(uint64_t) * bits.size(), hipMemcpyHostToDevice));
  }
  return CGBFS_OK;
}

// Hub-first slots: vertices ordered by out-degree (descending, ties by id).
// An in-neighbour u is gathered once per out-edge of u, so this packs the
// most-gathered lane words into the first few MB, where they stay cached.
#ifndef CG_SLOT_HUB_DEGREE
#define CG_SLOT_HUB_DEGREE 0
#endif
// Vertices of out-degree below this keep id order behind the hubs (0: every
// vertex is ordered by degree).
constexpr int64_t kSlotHubDegree = CG_SLOT_HUB_DEGREE;
__global__ void slot_keys_kernel(const int64_t *__restrict__ out_ptr, int64_t n,
                                 uint64_t *__restrict__ keys) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int64_t deg = out_ptr[v + 1] - out_ptr[v];
  if (deg < kSlotHubDegree) deg = 0;
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
  if (m->exec) hipGraphExecDestroy(m->exec);
  if (m->seedbuf) hipFree(m->seedbuf);
  if (m->seedstage) hipHostFree(m->seedstage);
  if (m->outbuf) hipFree(m->outbuf);
  if (m->mem) hipFree(m->mem);
  if (m->noin) hipFree(m->noin);
  if (m->out_slot && m->out_slot != m->in_slot) hipFree(m->out_slot);
  if (m->in_slot) hipFree(m->in_slot);
  if (m->slot_of) hipFree(m->slot_of);
  if (m->vertex_of) hipFree(m->vertex_of);
  delete m;
}

// The three sets of pull cursors (chunks, first-run segments, later
// segments) back to zero.
__device__ __forceinline__ void reset_ms_cursors(unsigned long long *cur) {
  for (int k = 0; k < 3; k++) reset_cursors(cur + k * kParts * kCursorStride);
}

// Lays a multi-source workspace out from `base` (sizes only when base is
// null) and returns its total size.
static size_t ws_layout(int64_t n, const MsGraphData *md, char *base, MsWorkspace *w) {
  const int64_t nbig = md->pin.nbig;
  const size_t wb = al(sizeof(uint64_t) * n), qb = al(sizeof(int32_t) * n);
  const size_t fbb = al(sizeof(uint64_t) * ((n + 63) / 64));  // whole 64-bit words per wavefront
  const size_t sizes[] = {
      wb, wb, wb, fbb, qb, qb, al(sizeof(MsCtrl)), al(sizeof(MsHist)), al(sizeof(MsLane)),
      al(sizeof(LaneOut)), al(sizeof(unsigned long long) * 64),
      al(sizeof(uint32_t) * 64 * (nbig + 1)), al(sizeof(uint64_t) * (nbig + 1)),
      al(sizeof(int32_t) * (md->pin.nseg + 1)), al(sizeof(int32_t) * (md->pout.nseg + 1)),
      al(sizeof(uint64_t) * (md->pin.nseg + 1)), al(sizeof(int32_t) * (md->pin.nseg + 1)),
      al(sizeof(uint64_t) * (md->pin.nseg + 1)), al(sizeof(unsigned long long) * 3 * kParts * kCursorStride),
      qb, al(sizeof(uint32_t) * side_words(n)), al(sizeof(unsigned long long) * 64 * kPallocStride)};
  size_t total = 0;
  for (size_t x : sizes) total += x;
  if (!base) return total;
  char *p = base;
  int i = 0;
  auto take = [&]() { char *q = p; p += sizes[i++]; return q; };
  w->vis = (uint64_t *)take();
  w->fw[0] = (uint64_t *)take();
  w->fw[1] = (uint64_t *)take();
  w->fb = (uint32_t *)take();
  w->q[0] = (int32_t *)take();
  w->q[1] = (int32_t *)take();
  w->ctrl = (MsCtrl *)take();
  w->hist = (MsHist *)take();
  w->lanes = (MsLane *)take();
  w->out = (LaneOut *)take();
  w->mlane = (unsigned long long *)take();
  w->cand = (uint32_t *)take();
  w->bgot = (uint64_t *)take();
  w->segq[0] = (int32_t *)take();
  w->segq[1] = (int32_t *)take();
  w->segneed = (uint64_t *)take();
  w->segq_late = (int32_t *)take();
  w->segneed_late = (uint64_t *)take();
  w->cursor = (unsigned long long *)take();
  w->pend = (int32_t *)take();
  w->side = (uint32_t *)take();
  w->palloc = (unsigned long long *)take();
  w->side_cap = side_words(n);
  return total;
}

// The per-graph multi-source data (segment lists, slots), built on first use.
static int ensure_msdata(DeviceGraph *g, MsGraphData **mdp) {
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
  *mdp = (MsGraphData *)g->msdata;
  return CGBFS_OK;
}

static int ensure_wsm(DeviceGraph *g, MsWorkspace *w, MsGraphData **mdp) {
  int rc = ensure_msdata(g, mdp);
  if (rc) return rc;
  const size_t total = ws_layout(g->n, *mdp, nullptr, w);
  if (g->wsm_bytes < total) {
    if (g->wsm) hipFree(g->wsm);
    g->wsm = nullptr;
    g->wsm_bytes = 0;
    CG_HIP(hipMalloc(&g->wsm, total));
    g->wsm_bytes = total;
  }
  ws_layout(g->n, *mdp, (char *)g->wsm, w);
  return CGBFS_OK;
}

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t shfl64(uint64_t x, int src) {
  uint32_t lo = __shfl((uint32_t)x, src), hi = __shfl((uint32_t)(x >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

// A 64-bit word moved across lanes by one DPP operation (VALU, no LDS
// round trip): kCtrl is the GFX9 DPP control, rows outside kRowMask and
// lanes whose source lies outside the pattern read 0.
template <int kCtrl, int kRowMask = 0xF>
__device__ __forceinline__ uint64_t dpp64(uint64_t x) {
  const uint32_t lo =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)x, kCtrl, kRowMask, 0xF, true);
  const uint32_t hi =
      (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(x >> 32), kCtrl, kRowMask, 0xF, true);
  return ((uint64_t)hi << 32) | lo;
}
// Segmented inclusive prefix-OR across the wavefront: lane j gets the OR of
// x over lanes [head, j], where head <= j is the first lane of j's segment
// (segments are runs of consecutive lanes).  Shifts within 16-lane rows,
// then the row broadcasts carry each row's tail into the rows after it,
// each step gated on the source lane lying in the same segment.  Needs
// every lane of the wavefront active.
__device__ __forceinline__ uint64_t seg_scan_or(uint64_t x, int head, int lane) {
  uint64_t y = dpp64<kDppRowShr + 1>(x);
  if (lane - 1 >= head) x |= y;
  y = dpp64<kDppRowShr + 2>(x);
  if (lane - 2 >= head) x |= y;
  y = dpp64<kDppRowShr + 4>(x);
  if (lane - 4 >= head) x |= y;
  y = dpp64<kDppRowShr + 8>(x);
  if (lane - 8 >= head) x |= y;
  y = dpp64<kDppRowBcast15, 0xA>(x);
  if (head < (lane & ~15)) x |= y;
  y = dpp64<kDppRowBcast31, 0xC>(x);
  if (head < 32) x |= y;
  return x;
}

// The word of lane j-1 (0 for lane 0).
__device__ __forceinline__ uint64_t lane_before(uint64_t x) { return dpp64<kDppWaveShr1>(x); }

// Inclusive prefix-OR across the wavefront.
__device__ __forceinline__ uint64_t wave_scan_or(uint64_t x) { return seg_scan_or(x, 0, lane_id()); }

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

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// OR of x over the wavefront, in every lane (every lane must be active).
__device__ __forceinline__ uint64_t wave_or64(uint64_t x) {
  x = wave_scan_or(x);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, kWave - 1);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), kWave - 1);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t clamp_deg(int64_t d) {
  return d < (int64_t)kDegClamp ? (uint32_t)d : kDegClamp;
}

// Per-lane tallies of newly reached vertices.  Lane l of a wavefront holds
// the totals for lane bit l; tally_flush() folds a block's wavefronts into
// MsLane with one global atomic per lane bit and field.
struct LaneTally {
  unsigned long long cnt = 0, out = 0, in = 0;
};

// Wavefront-uniform: `word` is this lane's vertex's new lane bits (0 for
// none), od/id its clamped out-/in-degree.
__device__ __forceinline__ void tally_words(uint64_t word, uint32_t od, uint32_t id, bool sym,
                                            LaneTally &t) {
  uint64_t bits = wave_or64(word);
  const int lane = lane_id();
  while (bits) {
    const int l = __ffsll((unsigned long long)bits) - 1;
    bits &= bits - 1;
    const bool hit = (word >> l) & 1ull;
    const unsigned c = (unsigned)__popcll(__ballot(hit));
    const uint32_t so = wave_sum_u32(hit ? od : 0u);
    const uint32_t si = sym ? so : wave_sum_u32(hit ? id : 0u);
    if (lane == l) {
      t.cnt += c;
      t.out += so;
      t.in += si;
    }
  }
}

// lds: 192 entries, zeroed (and synchronised) by the caller beforehand.
// Every thread of the block must call it.
__device__ __forceinline__ void tally_flush(const LaneTally &t, MsLane *L,
                                            unsigned long long *lds) {
  const int lane = lane_id();
  if (t.cnt) {
    atomicAdd(&lds[lane], t.cnt);
    atomicAdd(&lds[64 + lane], t.out);
    atomicAdd(&lds[128 + lane], t.in);
  }
  __syncthreads();
  if (threadIdx.x < 64 && lds[threadIdx.x]) {
    atomicAdd(&L->cnt[threadIdx.x], lds[threadIdx.x]);
    atomicAdd(&L->out[threadIdx.x], lds[64 + threadIdx.x]);
    atomicAdd(&L->in[threadIdx.x], lds[128 + threadIdx.x]);
  }
}

// ORs every thread's lane mask into *dst with one global atomic per block.
// Every thread of the block must call it.
__device__ __forceinline__ void live_flush(uint64_t v, unsigned long long *dst) {
  __shared__ unsigned long long blk;
  v = wave_or64(v);
  if (threadIdx.x == 0) blk = 0;
  __syncthreads();
  if (lane_id() == 0 && v) atomicOr(&blk, (unsigned long long)v);
  __syncthreads();
  if (threadIdx.x == 0 && blk) atomicOr(dst, blk);
}

__device__ __forceinline__ void tally_lds_init(unsigned long long *lds) {
  if (threadIdx.x < 192) lds[threadIdx.x] = 0;
  __syncthreads();
}


// Lane j's 64-bit value (j wavefront-uniform).
__device__ __forceinline__ int64_t readlane64(int64_t x, int j) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, j);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)x >> 32), j);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

constexpr int kSegSteps = (int)(kSeg / kWave);
static_assert(kSeg % kWave == 0, "a segment is a whole number of wavefront steps");

// One segment [beg, end) of an in-neighbour list, end - beg <= kSeg, for the
// lanes `need` (wavefront-uniform arguments).  Every index, id and frontier
// word of the segment is loaded before the scan starts, without branches
// (lanes past the end or without a frontier bit read a cached dummy), so the
// loads of all steps are in flight together.  Each lane's first hit in the
// segment (its original id) is offered to cand[] with atomicMin.  Returns
// the lanes found (wavefront-uniform).
// One step of a segment pull: entries [beg + 64 k, ...) below end, first
// hits into cand (atomicMin), returns the lanes found.
__device__ __forceinline__ uint64_t seg_step(const MsArgs &A, int64_t beg, int k, uint64_t h,
                                             uint32_t *cand) {
  const int lane = lane_id();
  const uint64_t inc = wave_scan_or(h);
  const uint64_t before = lane_before(inc);
  uint64_t mine = h & ~before;
  if (mine) {  // the id is read only for a first hit (h != 0: the entry is in range)
    const uint32_t u = (uint32_t)A.in_idx[beg + k * kWave + lane];
    while (mine) {
      int l = __ffsll((unsigned long long)mine) - 1;
      mine &= mine - 1;
      atomicMin(&cand[l], u);
    }
  }
  return shfl64(inc, kWave - 1);
}

// A segment needing at most this many lanes looks at its first step alone
// first: in late rounds the few lanes a hub still needs are mostly found
// among its first, lowest-id in-neighbours (hubs themselves).
constexpr int kSegHeadLanes = 8;

// A pull round reads the frontier bits (one per slot, 2 MB at 2^24
// vertices, so they stay in L2) before it gathers a lane word, and skips
// the words that are zero, when the frontier holds at most n / kFbDenom
// vertices; above that the words are gathered directly.
#ifndef CG_FB_DENOM
#define CG_FB_DENOM 8
#endif
constexpr long long kFbDenom = CG_FB_DENOM;

__device__ __forceinline__ uint64_t ms_seg_pull(const MsArgs &A, const uint64_t *fr, int64_t beg,
                                                int64_t end, uint64_t need, uint32_t *cand,
                                                bool use_fb, uint32_t &scanned) {
  const int lane = lane_id();
  uint64_t got = 0;
  int k0 = 0;
  if (__popcll(need) <= kSegHeadLanes) {
    const int64_t e = beg + lane;
    bool ok = e < end;
    const int32_t sl = A.in_slot[ok ? e : beg];
    if (use_fb) ok = ok && ((A.fb[sl >> 5] >> (sl & 31)) & 1u);
    const uint64_t w = fr[ok ? sl : 0];
    if (lane == 0) scanned += (uint32_t)(end - beg < kWave ? end - beg : kWave);
    const uint64_t h = ok ? w & need : 0ull;
    if (__any(h != 0)) got = seg_step(A, beg, 0, h, cand);
    if (got == need || end - beg <= kWave) return got;
    k0 = 1;
  }
  int32_t sl[kSegSteps];
  bool ok[kSegSteps];
  uint64_t w[kSegSteps];
#pragma unroll
  for (int k = 0; k < kSegSteps; k++) {
    const int64_t e = beg + k * kWave + lane;
    ok[k] = k >= k0 && e < end;
    const int64_t ee = ok[k] ? e : beg;
    sl[k] = A.in_slot[ee];
  }
  if (use_fb) {
#pragma unroll
    for (int k = 0; k < kSegSteps; k++) ok[k] = ok[k] && ((A.fb[sl[k] >> 5] >> (sl[k] & 31)) & 1u);
  }
#pragma unroll
  for (int k = 0; k < kSegSteps; k++) w[k] = fr[ok[k] ? sl[k] : 0];
  if (lane == 0) scanned += (uint32_t)(end - beg - (int64_t)k0 * kWave);
#pragma unroll
  for (int k = 0; k < kSegSteps; k++) {
    const uint64_t h = ok[k] ? w[k] & need & ~got : 0ull;
    if (!__any(h != 0)) continue;  // no frontier in-neighbour in this step
    got |= seg_step(A, beg, k, h, cand);
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
  for (int64_t v = tid; v < A.n; v += stride) A.pend[v] = -1;
  if (tid < 64) A.palloc[tid * kPallocStride] = 0;
}

__device__ __forceinline__ RoundState round_state(MsCtrl *c) {
  RoundState r;
  r.d = (int)ld_ctrl(&c->rounds);
  r.pm = (uint64_t)ld_ctrl(&c->pmask);
  r.qm = (uint64_t)ld_ctrl(&c->qmask);
  r.listed = ld_ctrl(&c->listed) != 0;
  r.use_fb = ld_ctrl(&c->use_fb) != 0;
  r.grab = ld_ctrl(&c->grab);
  r.noin = ld_ctrl(&c->noin_left) == 1;
  r.pend_n = (int64_t)ld_ctrl(&c->pend_n);
  return r;
}

// The lane output table staged in LDS at kernel entry: the per-bit loops
// read pointers and level values with broadcast LDS reads.
struct LaneOutLds {
  int32_t *par[64];
  int16_t *lev[64];
  int16_t level[64];  // the level a vertex reached this round gets in lane l
};

// Every thread of the block must call it (it ends with a block barrier).
__device__ __forceinline__ void stage_out(const LaneOut *out, int d, LaneOutLds &s) {
  if (threadIdx.x < 64) {
    s.par[threadIdx.x] = out->par[threadIdx.x];
    s.lev[threadIdx.x] = out->lev[threadIdx.x];
    s.level[threadIdx.x] = (int16_t)(d + 1 - out->base[threadIdx.x]);
  }
  __syncthreads();
}

// Whether the direction policy needs the per-lane tallies.
__host__ __device__ inline bool lane_policy(int mode) {
  return mode == CGBFS_DIR_TAU || mode == CGBFS_DIR_BEAMER_LANES;
}

// One wavefront.  Thread 0 plants the seeds (vertex sv[k] starts lane
// sl[k]) in the lane words and lists their distinct vertices; thread l sets
// lane l's direction state the way the single-root search starts (round 0
// pushes unless every round pulls, bool_pe.csl:714-719).  The seeds' own
// outputs are written by ms_seed_outputs after the last round.
__global__ void ms_init_seeds(MsArgs A, const int64_t *sv, const int32_t *sl, int64_t nseeds,
                              MsParams P) {
  if (blockIdx.x != 0) return;
  MsCtrl *c = A.ctrl;
  MsLane *L = A.lanes;
  const int l = threadIdx.x;
  __shared__ unsigned long long s_nq, s_deg;
  if (l == 0) {
    s_nq = 0;
    s_deg = 0;
  }
  __syncthreads();
  // the first lane bit planted on a vertex lists it (and counts its edges)
  for (int64_t k = l; k < nseeds; k += blockDim.x) {
    const int64_t r = sv[k];
    const int32_t sr = A.slot_of[r];
    const unsigned long long bit = 1ull << sl[k];
    atomicOr((unsigned long long *)&A.vis[r], bit);
    if (atomicOr((unsigned long long *)&A.fw[0][sr], bit) == 0) {
      A.q[0][atomicAdd(&s_nq, 1ull)] = (int32_t)r;
      atomicAdd(&s_deg, (unsigned long long)(A.out_ptr[r + 1] - A.out_ptr[r]));
    }
  }
  __syncthreads();
  const unsigned long long deg = s_deg;  // out-degree sum of the distinct seed vertices
  if (l == 0) {
    const unsigned long long nq = s_nq;
    st_ctrl(&c->qtail[0].v, nq);
    st_ctrl(&c->qtail[1].v, 0ull);
    st_ctrl(&c->nf, 0ull);
    st_ctrl(&c->scanned, 0ull);
    st_ctrl(&c->rows, 0ull);
    st_ctrl(&c->pairs, 0ull);
    st_ctrl(&c->pnf, 0ull);
    st_ctrl(&c->pscanned, 0ull);
    st_ctrl(&c->prows, 0ull);
    st_ctrl(&c->ppairs, 0ull);
    st_ctrl(&c->rounds, 0ll);
    st_ctrl(&c->done, 0);
    st_ctrl(&c->status, 0);
    st_ctrl(&c->grab, P.dense_grab);
    st_ctrl(&c->noin_left, kNoinPull);
    st_ctrl(&c->pend_n, 0ll);
    reset_ms_cursors(A.cursor);
  }
  bool live = false;
  int dir = 0;
  if (l < 64) {
    st_ctrl(&L->cnt[l], 0ull);
    st_ctrl(&L->out[l], 0ull);
    st_ctrl(&L->in[l], 0ull);
    if (l < A.nroots) {
      long long mu = (long long)P.nnz;
      for (int64_t k = 0; k < nseeds; k++)
        if (sl[k] == l) mu -= (long long)(A.in_ptr[sv[k] + 1] - A.in_ptr[sv[k]]);
      if (P.mode == CGBFS_DIR_BOTTOMUP) dir = 1;
      st_ctrl(&L->mu[l], mu);
      st_ctrl(&L->prev[l], 1ll);
      live = true;
    }
    st_ctrl(&L->dir[l], dir);
  }
  // the union-count policy starts like the per-lane rules, except that 64
  // roots may already hold enough edges to pull from round 0
  const unsigned long long mf = deg;
  int g = 0;
  if (P.mode == CGBFS_DIR_BOTTOMUP) g = 1;
  else if (P.mode == CGBFS_DIR_BEAMER && (double)mf * P.alpha > (double)P.nnz) g = 1;
  uint64_t pm, qm;
  if (lane_policy(P.mode)) {
    pm = __ballot(live && dir == 1);
    qm = __ballot(live && dir == 0);
  } else {
    pm = g ? A.act : 0ull;
    qm = g ? 0ull : A.act;
  }
  if (l == 0) {
    st_ctrl(&c->gdir, g);
    st_ctrl(&c->mf, 0ull);
    st_ctrl(&c->live, 0ull);
    st_ctrl(&c->segtail[0].v, 0ull);
    st_ctrl(&c->segtail[1].v, 0ull);
    st_ctrl(&c->segtail_late, 0ull);
    const long long nq0 = (long long)ld_ctrl(&c->qtail[0].v);
    st_ctrl(&c->prev_nf, nq0);
    st_ctrl(&c->listed, 1);
    st_ctrl(&c->use_fb, nq0 * kFbDenom <= (long long)A.n ? 1 : 0);
    st_ctrl(&c->pmask, (unsigned long long)pm);
    st_ctrl(&c->qmask, (unsigned long long)qm);
  }
}

// Pull round over non-big vertices (thread and wavefront tiers), then the
// big-vertex segments (one wavefront per segment, per-lane atomicMin into
// cand[], lanes into bgot[]).
// The round's pull lanes (only lanes with a non-empty frontier) and whether
// the frontier bits are current (sparse frontiers only: on a dense one nearly
// every bit is set and the extra load only adds latency) come from the
// control block; tally_on keeps the per-lane tallies (per-lane policies).
// pend_on: the instantiation for a round after a push part that deferred
// outputs (it writes them with its chunks); the other one runs every other
// pull round, without the registers that costs.  Both are launched where
// the host does not know the round (graph replay, the serving engine); the
// one that does not match returns at once.
template <bool tally_on, bool pend_on>
__global__ void __launch_bounds__(kMsBlock, kMsPullWaves) ms_pull(MsArgs A) {
  MsCtrl *c = A.ctrl;
  const RoundState R = round_state(c);
  if (!R.pm || (R.pend_n != 0) != pend_on) return;
  const int d = R.d;
  const uint64_t pm = R.pm;
  const bool use_fb = R.use_fb;
  __shared__ unsigned long long red[kMsWaves];
  __shared__ LaneOutLds so;
  stage_out(A.out, d, so);
  // per-lane tallies (the per-lane direction policies only): 1.5 KB of
  // LDS the default pull does without, so one more block fits per CU
  unsigned long long *tl = nullptr;
  if constexpr (tally_on) {
    __shared__ unsigned long long tl_lds[192];
    tl = tl_lds;
    tally_lds_init(tl);
  }
  LaneTally tally;
  const bool sym = A.symmetric != 0;
  const uint64_t *fr = A.fw[d & 1];
  uint64_t *nx = A.fw[(d + 1) & 1];
  // per-thread counts stay far below 2^32 (a thread owns a few rows per
  // chunk, each shorter than kSeg); block sums are 64-bit
  uint32_t scanned = 0, rows = 0, pairs = 0, mf = 0, nf = 0;
  uint64_t live = 0;
  const int lane = lane_id();
  // First hits of the current chunk, as offsets into the row (rows here are
  // shorter than kSeg), indexed [lane bit][row] so the per-bit flush below
  // reads one conflict-free row of LDS and stores 64 consecutive vertices.
  __shared__ __align__(16) TileOff ptile[kMsWaves][kWave][kWave];
  TileOff (*pt)[kWave] = ptile[threadIdx.x / kWave];
  // Row marks of the current pass, [window][position].
  __shared__ int32_t rmap[kMsWaves][kUnroll][kWave];
  int32_t (*rm)[kWave] = rmap[threadIdx.x / kWave];
  // Chunks of 64 consecutive vertices per wavefront.  The rows of the chunk
  // that still need lanes (big rows excluded) are laid end to end in a
  // virtual edge range [0, total); each step every lane takes one virtual
  // entry, finds its row by binary search over the rows' virtual starts,
  // gathers the frontier word, and a segmented prefix-OR (segments = rows)
  // gives each (row, lane) its first -- lowest-id -- hit.  Row state (got)
  // lives in the lane that owns the row.
  const int32_t nch = (int32_t)((A.n + kWave - 1) / kWave);
  // Chunks come from the cursors: their costs differ by orders of magnitude
  // (row lengths, how soon rows are satisfied), which a static split of the
  // chunks over the wavefronts leaves unbalanced.
  ChunkFeed feed(A.cursor, nch, R.grab > 0 ? R.grab : 1, R.grab < 0);
  int32_t cb = 0, ce = 0;
  while (cb < ce || feed.next(cb, ce)) {
    const int32_t ch = cb++;
    const int32_t v = ch * kWave + lane;  // vertex ids are int32
    // branch-free loads (a lane past n reads vertex 0): the per-vertex
    // loads go out together, then the lane word through the slot
    const bool inr = v < A.n;
    const int32_t vv = inr ? v : 0;
    const bool big = A.pin.is_big[vv] != 0;
    const int32_t sv_me = A.slot_of[vv];
    const int64_t rb0 = A.in_ptr[vv], re0 = A.in_ptr[vv + 1];
    const uint64_t vw = A.vis[vv];
    const uint64_t need = (inr && !big) ? pm & ~vw : 0ull;
    const int64_t rb = rb0;
    // Pairs the previous round's push left in the side table (lanes pl,
    // entry pbase): written with this chunk's stores below.
    int32_t pbase = -1;
    uint64_t pl = 0;
    if constexpr (pend_on) {
      pbase = inr ? A.pend[vv] : -1;
      const uint64_t m = *reinterpret_cast<const uint64_t *>(&A.side[pbase >= 0 ? pbase : 0]);
      pl = pbase >= 0 ? m : 0ull;
    }
    // A big vertex's row was pulled by segments (ms_pull_segs, before this
    // kernel): its new lanes and their parents wait in bgot / cand, indexed
    // by its rank among the big vertices (they ascend by vertex id).
    const uint64_t bigm = ballot64(inr && big);
    int32_t bid = 0;
    if (bigm)
      bid = A.pin.chunk_big[ch] + (int32_t)__popcll(bigm & ((1ull << lane) - 1ull));
    const int32_t len = need ? (int32_t)(re0 - rb0) : 0;  // < kSeg for non-big rows
    if (len) rows++;
    // First-hit offsets start at 0 (the row's first entry) for every lane
    // bit and row: what the head probe finds needs no tile write, only the
    // hits of the scan further in do.  Each lane clears one lane bit's row
    // of the tile; the wavefront's LDS operations stay in order, so the
    // previous chunk's flush has read it.
    {
      uint4 *trow = reinterpret_cast<uint4 *>(pt[lane]);
#pragma unroll
      for (int j = 0; j < (int)(kWave * sizeof(TileOff) / sizeof(uint4)); j++)
        trow[j] = make_uint4(0u, 0u, 0u, 0u);
    }
    // Head probe: the first entry of every row that needs lanes, all at
    // once.  In a late round most rows end here: a row's lowest in-neighbours
    // are hubs, reached early.  The rows still open are then laid end to end
    // from their second entry.
    uint64_t got = 0;
    {
      const int32_t hs = A.in_slot[len ? rb : 0];
      bool hok = len != 0;
      if (use_fb) hok = hok && ((A.fb[hs >> 5] >> (hs & 31)) & 1u);
      const uint64_t hw = fr[hok ? hs : 0];
      got = hok ? hw & need : 0ull;
      if (len) scanned++;
    }
    const int32_t lenr = (len && (need & ~got)) ? len - 1 : 0;  // entries left to scan
    const int64_t rbr = rb + 1;
    // exclusive prefix sum of the rest of the rows -> virtual starts
    const int32_t incl = wave_scan_add(lenr);
    const int32_t vstart = incl - lenr;
    const int32_t total = __builtin_amdgcn_readlane(incl, kWave - 1);
    const int32_t my_end = vstart + lenr;
    // Row of the entries before the current window: the windows run in
    // order, so it is the row of the previous window's last lane.
    int rcarry = 0;
    // kUnroll steps per pass: their index loads, then their frontier-word
    // gathers, are all in flight together (no branches around the loads:
    // idle lanes read entry 0 and the word of slot 0, both cached); the
    // scans then run step by step.
    for (int32_t base0 = 0; base0 < total; base0 += kWave * kUnroll) {
      int rk[kUnroll];
      int32_t rsk[kUnroll], idx[kUnroll];
      bool ok[kUnroll];
      uint64_t nrk[kUnroll], raw[kUnroll];
      // Row of each entry of the pass: each row owner marks where its row
      // starts in the pass, and a prefix-max over the marks (carried across
      // windows) gives every entry the last row starting at or before it.
      // Rows of length 0 mark nothing, so the row found always has entries.
#pragma unroll
      for (int k = 0; k < kUnroll; k++) rm[k][lane] = 0;
      wave_sync();
      {
        const int32_t p = vstart - base0;
        if (lenr && p >= 0 && p < kWave * kUnroll) rm[p / kWave][p % kWave] = lane;
      }
      wave_sync();
#pragma unroll
      for (int k = 0; k < kUnroll; k++) {
        const int r = max(wave_scan_max(rm[k][lane]), rcarry);
        rcarry = __builtin_amdgcn_readlane(r, kWave - 1);
        rk[k] = r;
      }
#pragma unroll
      for (int k = 0; k < kUnroll; k++) {
        const int32_t ve = base0 + k * kWave + lane;
        const int r = rk[k];
        rsk[k] = __shfl(vstart, r);
        const int64_t rbase = __shfl(rbr, r);
        nrk[k] = shfl64(need, r);
        const uint64_t gr = shfl64(got, r);
        ok[k] = ve < total && (nrk[k] & ~gr);
        idx[k] = A.in_slot[ok[k] ? rbase + (ve - rsk[k]) : 0];
      }
      if (use_fb) {
        // sparse frontier: the gather is needed only where the bit is set
#pragma unroll
        for (int k = 0; k < kUnroll; k++)
          if (ok[k]) scanned++;
#pragma unroll
        for (int k = 0; k < kUnroll; k++)
          ok[k] = ok[k] && ((A.fb[idx[k] >> 5] >> (idx[k] & 31)) & 1u);
      } else {
#pragma unroll
        for (int k = 0; k < kUnroll; k++)
          if (ok[k]) scanned++;
      }
#pragma unroll
      for (int k = 0; k < kUnroll; k++) raw[k] = fr[ok[k] ? idx[k] : 0];
#pragma unroll
      for (int k = 0; k < kUnroll; k++) raw[k] = ok[k] ? raw[k] & nrk[k] : 0ull;
#pragma unroll
      for (int k = 0; k < kUnroll; k++) {
        if (!__any(raw[k] != 0)) continue;  // no hit in this step: got unchanged
        const int32_t base = base0 + k * kWave;
        const int r = rk[k], rs = rsk[k];
        uint64_t h = raw[k] & ~shfl64(got, r);
        // segmented inclusive prefix-OR; the segment of lane starts at head
        const int head = rs > base ? rs - base : 0;
        const uint64_t x = seg_scan_or(h, head, lane);
        const uint64_t prev = lane_before(x);
        uint64_t mine = h & ~(lane > head ? prev : 0ull);
        const TileOff off = (TileOff)(base + lane - rs + 1);  // from the row start
        // one 32-bit half at a time: a 64-bit find-first and clear take
        // twice the instructions
        for (uint32_t m = (uint32_t)mine; m; m &= m - 1) pt[__builtin_ctz(m)][r] = off;
        for (uint32_t m = (uint32_t)(mine >> 32); m; m &= m - 1) pt[32 + __builtin_ctz(m)][r] = off;
        // each row owner collects its row's OR over this step
        int last = (my_end < base + kWave ? my_end : base + kWave) - 1 - base;
        bool hit_step = lenr && vstart < base + kWave && my_end > base;
        uint64_t tot = shfl64(x, hit_step ? last : lane);
        if (hit_step) got |= tot;
      }
      if (!__any(lenr && (need & ~got))) break;  // every row of the chunk satisfied
    }
    // Vertices without in-edges: no lane but a seed's own ever reaches them.
    // The second pull round writes their -1 outputs (lanes other than their
    // own seeds') in the same store instructions as the chunk's new pairs,
    // so those partial lines fill up instead of being written twice; the
    // first pull round, the heaviest, is left its own stores.
    const uint64_t nmiss = (R.noin && ((A.noin[ch] >> lane) & 1ull)) ? A.act & ~vw : 0ull;
    const bool bg = (bigm >> lane) & 1ull;  // this lane's vertex is big
    if (bg) got = A.bgot[bid];
    // big vertices with new lanes (rare after the round that reaches the hubs)
    const bool bgot_any = __any(bg && got);
    // Flush: per lane bit, 64 consecutive vertices' parents and levels.
    const uint64_t wr = got | nmiss | pl;  // the lane bits this vertex writes
    uint64_t bits = wave_or64(wr);
    // First hits are the lowest ids with the bit, mostly near the row start.
    constexpr int kHead = 4;
    int32_t head[kHead];
#pragma unroll
    for (int j = 0; j < kHead; j++) head[j] = A.in_idx[got && j < len ? rb + j : 0];
    // kFlush bits at a time: their parent-id loads are in flight together
    // before any store is issued.  The values are selected without branches:
    // a hit's parent and this round's level, a deferred push pair's parent
    // and the round before's level, -1 for a vertex without in-edges.
    while (bits) {
      int lb[kFlush];
      int32_t pv[kFlush];
      int16_t lv[kFlush];
      bool st[kFlush];
#pragma unroll
      for (int k = 0; k < kFlush; k++) {
        lb[k] = bits ? __ffsll((unsigned long long)bits) - 1 : -1;
        bits &= bits - 1;
      }
#pragma unroll
      for (int k = 0; k < kFlush; k++) {
        const int lk = lb[k] >= 0 ? lb[k] : 0;
        const uint64_t m = 1ull << lk;
        st[k] = lb[k] >= 0 && (wr & m) != 0;
        const bool hb = (got & m) != 0;
        const bool pb = pend_on && (pl & m) != 0;
        const int off = pt[lk][lane];
        // no branch: loads go out together; a hit among the row's first
        // kHead entries takes its id from the row head loaded once per chunk,
        // a deferred push pair from its side entry
        const int32_t *src =
            pb ? reinterpret_cast<const int32_t *>(A.side) + pbase + 2 + __popcll(pl & (m - 1ull))
               : A.in_idx + (hb && !bg && off >= kHead ? rb + off : 0);
        const int32_t far = *src;
        int32_t p = pb || off >= kHead ? far : off == 0 ? head[0] : off == 1 ? head[1] : off == 2 ? head[2] : head[3];
        if (bgot_any) {  // a big vertex's parent is its candidate
          const uint32_t cv = A.cand[bg && hb ? (int64_t)bid * 64 + lk : 0];
          if (bg && hb) p = (int32_t)cv;
        }
        const int16_t lvl = so.level[lk];
        pv[k] = hb || pb ? p : (int32_t)-1;
        lv[k] = hb ? lvl : pb ? (int16_t)(lvl - 1) : (int16_t)-1;
      }
#pragma unroll
      for (int k = 0; k < kFlush; k++) {
        if (st[k]) {
          st_once(&so.par[lb[k]][v], pv[k]);
          st_once(&so.lev[lb[k]][v], lv[k]);
        }
      }
    }
    if (pend_on && pbase >= 0) A.pend[v] = -1;
    if (bg && got) {  // candidates back to +inf for the next pull
      uint4 *crow = reinterpret_cast<uint4 *>(&A.cand[(int64_t)bid * 64]);
#pragma unroll
      for (int j = 0; j < 16; j++) crow[j] = make_uint4(~0u, ~0u, ~0u, ~0u);
      A.bgot[bid] = 0;
    }
    const int64_t indeg = re0 - rb0;
    int64_t od = 0;
    if (got) {
      nx[sv_me] = got;
      A.vis[v] = vw | got;
      od = sym ? indeg : A.out_ptr[v + 1] - A.out_ptr[v];
      mf += (uint32_t)od;
      live |= got;
      pairs += (uint32_t)__popcll(got);
      nf++;
    }
    if (tally_on) tally_words(got, clamp_deg(od), clamp_deg(indeg), sym, tally);
  }
  if (tally_on) tally_flush(tally, A.lanes, tl);
  live_flush(live, &c->live);
  unsigned long long s = ms_block_sum(nf, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->pnf, s);
  s = ms_block_sum(scanned, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->pscanned, s);
  s = ms_block_sum(rows, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->prows, s);
  s = ms_block_sum(pairs, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->ppairs, s);
  s = ms_block_sum(mf, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->mf, s);
}

// Big vertices of a pull round, before ms_pull: one active segment at a time
// per wavefront (ms_active_segs listed the segments whose vertex needs lanes,
// with those lanes).  First hits go to the vertex's per-lane candidates by
// atomicMin, the lanes found to bgot; ms_pull writes them out with the
// vertex's chunk.  Two launches: late = 0 pulls the first kSegRun segments
// of every vertex, late = 1 the rest, where a lane whose candidate (from
// the finished first launch, or from this one) is already below the
// segment's first id cannot improve and drops out; a stale read of a
// candidate only keeps a lane in.  Segments come from their own cursors,
// seg_grab (32 by default) consecutive ones a grab: a segment stops at its first step that
// finds every needed lane, so costs differ, and a run of one vertex's
// segments in row order lets the later ones skip the lanes the earlier ones
// found.  With fewer than kSegStatic segments per wavefront they are split
// statically instead.
constexpr int32_t kSegRun = 32;
constexpr int32_t kSegStatic = 4;
__global__ void __launch_bounds__(kMsBlock) ms_pull_segs(MsArgs A, int late) {
  MsCtrl *c = A.ctrl;
  const RoundState R = round_state(c);
  if (!R.pm) return;
  const int32_t nact = (int32_t)ld_ctrl(late ? &c->segtail_late : &c->segtail[0].v);
  if (nact == 0) return;  // uniform: written by an earlier kernel
  __shared__ unsigned long long red[kMsWaves];
  const uint64_t *fr = A.fw[R.d & 1];
  const int lane = lane_id();
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  const int32_t *segs = late ? A.segq_late : A.segq[0];
  const uint64_t *needs = late ? A.segneed_late : A.segneed;
  uint32_t scanned = 0;
  ChunkFeed feed(A.cursor + (1 + late) * kParts * kCursorStride, nact, A.seg_grab,
                 nact < kSegStatic * nwaves);
  int32_t gb, ge;
  // A vertex's segments are listed together in row order, so lanes this
  // wavefront found in an earlier segment (lower ids) are not looked for in
  // a later one of the same vertex.
  int32_t prev_bid = -1, prev_i = -1;
  uint64_t prev_got = 0;
  while (feed.next(gb, ge)) for (int32_t b0 = gb; b0 < ge; b0 += kWave) {
    // The descriptors of up to 64 segments of the grab, one per lane, are
    // loaded together (their list entry, then vertex, start and length, and
    // in the second launch the first id), so a segment's scan starts with
    // its slots instead of two dependent loads; the loop reads them from
    // their lanes.
    const int32_t nb = ge - b0 < kWave ? ge - b0 : kWave;
    const int32_t ii = b0 + (lane < nb ? lane : 0);
    const int32_t s_l = segs[ii];
    const uint64_t need_l = needs[ii];
    const int32_t bid_l = A.pin.seg_bid[s_l];
    const int64_t beg_l = A.pin.seg_beg[s_l];
    const int32_t len_l = A.pin.seg_len[s_l];
    const int32_t first_l = late ? A.in_idx[beg_l] : 0;
    for (int32_t j = 0; j < nb; j++) {
      const int32_t i = b0 + j;
      uint64_t need = (uint64_t)readlane64((int64_t)need_l, j);
      const int32_t bid = __builtin_amdgcn_readlane(bid_l, j);
      if (bid == prev_bid && i > prev_i) need &= ~prev_got;
      else prev_got = 0;
      prev_bid = bid;
      prev_i = i;
      const int64_t beg = readlane64(beg_l, j);
      if (late && need) {
        const uint32_t low = A.cand[(int64_t)bid * 64 + lane];
        need &= ~ballot64(low < (uint32_t)__builtin_amdgcn_readlane(first_l, j));
      }
      if (!need) continue;
      const int64_t end = beg + __builtin_amdgcn_readlane(len_l, j);
      uint64_t gj = ms_seg_pull(A, fr, beg, end, need, &A.cand[(int64_t)bid * 64], R.use_fb, scanned);
      prev_got |= gj;
      if (gj && lane == 0) atomicOr((unsigned long long *)&A.bgot[bid], (unsigned long long)gj);
    }
  }
  const unsigned long long t = ms_block_sum(scanned, red);
  if (threadIdx.x == 0 && t) atomicAdd(&c->pscanned, t);
}

// Active list from the frontier words (slot order -> vertex ids), for a push
// round that follows a pull round (pull rounds keep no list).
// Lists the round's frontier (its nonzero words, by vertex) for a push part
// after a pull part, which lists nothing.  Every thread of the block calls it.
__device__ void compact_part(const MsArgs &A, const RoundState &R, const int32_t *vertex_of) {
  MsCtrl *c = A.ctrl;
  const int d = R.d;
  const uint64_t *fr = A.fw[d & 1];
  __shared__ int32_t lq[kMsLdsQueue];
  __shared__ unsigned lq_n;
  __shared__ unsigned long long lq_base;
  StagedQueue sq{lq, &lq_n, &lq_base, (unsigned)kMsLdsQueue, &c->qtail[d & 1].v, A.q[d & 1]};
  sq.init();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t nround = ((A.n + stride - 1) / stride) * stride;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nround; k += stride) {
    bool on = k < A.n && fr[k] != 0;
    sq.push(on, on ? vertex_of[k] : 0);
  }
  sq.flush();
}

// This round's active segments: those of the big vertices that still need
// a pull lane (which = 0, in view) or that are on the push lanes' frontier
// (which = 1, out view).  One wavefront-aggregated append per wavefront.
// Writes each lane's run of consecutive segment ids (and, where needs is
// given, its lane mask) at its place in a list, the whole wavefront writing
// one lane's run at a time: a hub's run is thousands of segments, which
// its own lane alone would write one by one while the others wait.
__device__ __forceinline__ void list_runs(int32_t *ids, uint64_t *needs, unsigned long long base,
                                          int32_t first, int32_t cnt, uint64_t need) {
  const int lane = lane_id();
  uint64_t owners = ballot64(cnt > 0);
  while (owners) {
    const int o = __ffsll((unsigned long long)owners) - 1;
    owners &= owners - 1;
    const unsigned long long ob = (unsigned long long)shfl64((uint64_t)base, o);
    const int32_t of = __shfl(first, o), oc = __shfl(cnt, o);
    const uint64_t on = shfl64(need, o);
    for (int32_t k = lane; k < oc; k += kWave) {
      ids[ob + k] = of + k;
      if (needs) needs[ob + k] = on;
    }
  }
}

__device__ void active_segs_part(const MsArgs &A, const RoundState &R, int which) {
  const uint64_t mask = which ? R.qm : R.pm;
  const int d = R.d;
  const SegList &L = which ? A.pout : A.pin;
  const uint64_t *fr = A.fw[d & 1];
  unsigned long long *tail = &A.ctrl->segtail[which].v;
  int32_t *out = A.segq[which];
  const int lane = lane_id();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x - lane; b0 < L.nbig; b0 += stride) {
    const int64_t b = b0 + lane;
    int32_t first = 0, cnt = 0;
    uint64_t need = 0;
    if (b < L.nbig) {
      const int32_t bv = L.big_v[b];
      const int32_t sv = A.slot_of[bv];
      need = mask & ~A.vis[bv];
      const bool on = which ? (fr[sv] & mask) != 0 : need != 0;
      if (on) {
        first = L.seg_first[b];
        cnt = L.seg_first[b + 1] - first;
      }
    }
    // A pull lists a vertex's first kSegRun segments in segq, the rest in
    // segq_late; both keep row order.
    const int32_t c0 = which ? cnt : (cnt < kSegRun ? cnt : kSegRun);
    const int32_t c1 = cnt - c0;
    const int32_t incl = wave_scan_add(c0);
    const int32_t total = __shfl(incl, kWave - 1);
    if (total) {
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(tail, (unsigned long long)total);
      base = __shfl(base, 0) + (unsigned long long)(incl - c0);
      if (!which) {  // at most kSegRun each: every lane writes its own
        for (int32_t k = 0; k < c0; k++) out[base + k] = first + k;
        for (int32_t k = 0; k < c0; k++) A.segneed[base + k] = need;
      } else {
        list_runs(out, nullptr, base, first, c0, 0ull);
      }
    }
    if (which) continue;
    const int32_t incl1 = wave_scan_add(c1);
    const int32_t total1 = __shfl(incl1, kWave - 1);
    if (total1) {
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(&A.ctrl->segtail_late, (unsigned long long)total1);
      base = __shfl(base, 0) + (unsigned long long)(incl1 - c1);
      list_runs(A.segq_late, A.segneed_late, base, first + c0, c1, need);
    }
  }
}

// Frontier bits from the frontier words: a wavefront ballots 64 consecutive
// slots and stores the two bitmap words (whole words: no atomics, and no
// clearing between rounds).
__device__ void mark_part(const MsArgs &A, const RoundState &R) {
  const uint64_t *__restrict__ fr = A.fw[R.d & 1];
  uint32_t *__restrict__ fb = A.fb;
  const int64_t n = A.n;
  const int lane = lane_id();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t s0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x - lane; s0 < n; s0 += stride) {
    const int64_t s = s0 + lane;
    const uint64_t m = __ballot(s < n && fr[s] != 0);
    if (lane < 2) fb[(s0 >> 5) + lane] = (uint32_t)(m >> (32 * lane));
  }
}

// Whether the level-d frontier words are cleared through their list: only
// when the round's frontier is listed and the list is short.  Otherwise
// ms_clear sweeps every word: a listed word costs a
// gather and a scattered 8-byte store, a swept one a share of a streamed
// line (a 3.3 M-vertex list at 2^24 vertices took 115 us, the sweep 25 us).
__device__ __forceinline__ bool clear_by_list(const MsArgs &A, const RoundState &R) {
  if (!(R.listed || R.qm)) return false;
  return (int64_t)ld_ctrl(&A.ctrl->qtail[R.d & 1].v) * 32 <= A.n;
}


// Push pass for the lanes in qm (phase 0: OR into nx and append the first
// writer of qm bits; phase 1: offer predecessors to the new pairs with
// per-lane atomicMin).  nx may already hold bits the pull found this round;
// they belong to other lanes.
__global__ void __launch_bounds__(kMsBlock) ms_push(MsArgs A, int phase) {
  MsCtrl *c = A.ctrl;
  const RoundState R = round_state(c);
  if (!R.qm) return;
  const int d = R.d;
  const uint64_t qm = R.qm;
  __shared__ unsigned long long red[kMsWaves];
  __shared__ LaneOutLds so;
  stage_out(A.out, d, so);
  const uint64_t *fr = A.fw[d & 1];
  uint64_t *nx = A.fw[(d + 1) & 1];
  const int32_t *Q = A.q[d & 1];
  int32_t *QN = A.q[(d + 1) & 1];
  unsigned long long *tail = &c->qtail[(d + 1) & 1].v;
  const int64_t nq = (int64_t)ld_ctrl(&c->qtail[d & 1].v);
  const int lane = lane_id();
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t wave = gtid / kWave, nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  unsigned long long scanned = 0, rows = 0;
  __shared__ int32_t lq[kMsLdsQueue];
  __shared__ unsigned lq_n;
  __shared__ unsigned long long lq_base;
  StagedQueue sq{lq, &lq_n, &lq_base, (unsigned)kMsLdsQueue, tail, QN};
  if (phase == 0) sq.init();
  // Work items: queue entries (skipping big vertices), then the segments of
  // the big frontier vertices.  Both loops are wavefront-uniform.
  const int64_t nitems = nq + (int64_t)ld_ctrl(&c->segtail[1].v);
  for (int64_t it = wave; it < nitems; it += nwaves) {
    int32_t u;
    int64_t beg, end;
    if (it < nq) {
      u = Q[it];
      if (A.pout.is_big[u]) continue;
      beg = A.out_ptr[u];
      end = A.out_ptr[u + 1];
    } else {
      const int32_t s = A.segq[1][it - nq];
      u = A.pout.seg_v[s];
      beg = A.pout.seg_beg[s];
      end = beg + A.pout.seg_len[s];
    }
    const uint64_t f = fr[A.slot_of[u]] & qm;
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
        // Edges into one vertex meet at its word and its parent slots, and
        // device-scope atomics on one address serialise: a hub's in-edges
        // would queue there.  A plain read first skips the atomics that
        // cannot change anything.  Within a pass bits are only set and
        // parents only lowered, so a stale read at worst keeps an atomic.
        // Pass 0 does not read the target's lane word: lanes of f that had
        // already reached w are set in nx[sw] too, and ms_prep, which reads
        // w's lane word anyway, clears them (and counts a vertex listed only
        // for them as nothing new).
        if (phase == 0) {
          const uint64_t nv = f;
          if (!kPushFilter || (nx[sw] & nv) != nv) {
            uint64_t old = atomicOr((unsigned long long *)&nx[sw], (unsigned long long)nv);
            first = (old & qm) == 0;
          }
        } else {
          // the candidates of w's new lanes: its side entry (one per lane
          // of nw, in lane order) or, without one, its output rows; the
          // entry is loaded beside the word, not after it
          const int32_t b = A.pend[w];
          const uint64_t nxw = nx[sw];
          uint64_t nv = f & nxw;
          if (nv) {
            const uint64_t nw = nxw & qm;
            while (nv) {
              int l = __ffsll((unsigned long long)nv) - 1;
              nv &= nv - 1;
              uint32_t *slot = b >= 0 ? &A.side[b + 2 + __popcll(nw & ((1ull << l) - 1ull))]
                                      : (uint32_t *)&so.par[l][w];
              if (!kPushFilter2 || (uint32_t)u < *gmem(slot))
                __hip_atomic_fetch_min(gmem(slot), (uint32_t)u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
          }
        }
      }
      if (phase == 0) sq.push(first, w);
    }
  }
  if (phase == 0) sq.flush();
  // the entries are counted once, by the first pass
  unsigned long long s = ms_block_sum(scanned, red);
  if (phase == 0 && threadIdx.x == 0 && s) atomicAdd(&c->scanned, s);
  s = ms_block_sum(rows, red);
  if (phase == 0 && threadIdx.x == 0 && s) atomicAdd(&c->rows, s);
}

// Between the push passes: give each new vertex its side-table entry (lane
// mask, +inf candidates for the atomicMin of pass 2) or, where the table is
// full, mark its new qm pairs in the output rows directly (parent +inf,
// level d+1); fold the pairs into vis.  A vertex counts towards nf and mf
// only if the pull reached no lane of it this round.
__global__ void __launch_bounds__(kMsBlock) ms_prep(MsArgs A, int tally_on) {
  MsCtrl *c = A.ctrl;
  const RoundState R = round_state(c);
  if (!R.qm) return;
  const int d = R.d;
  const uint64_t qm = R.qm;
  __shared__ unsigned long long red[kMsWaves];
  __shared__ LaneOutLds so;
  stage_out(A.out, d, so);
  __shared__ unsigned long long tl[192];
  tally_lds_init(tl);
  LaneTally tally;
  uint64_t *nx = A.fw[(d + 1) & 1];
  const int32_t *QN = A.q[(d + 1) & 1];
  const int64_t nq = (int64_t)ld_ctrl(&c->qtail[(d + 1) & 1].v);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long nf = 0, pairs = 0, mf = 0;
  uint64_t live = 0;
  const int lane = lane_id();
  // this wavefront's allocation counter and its region of the side table
  const int64_t wave = gtid / kWave;
  unsigned long long *pa = &A.palloc[(wave & 63) * kPallocStride];
  const int64_t region = A.side_cap / 64;
  const int64_t rbase = (wave & 63) * region;
  // wavefront-uniform trip count (the tally and the allocation are
  // wavefront operations)
  for (int64_t i0 = gtid - lane; i0 < nq; i0 += stride) {
    const int64_t i = i0 + lane;
    uint64_t nw = 0;
    int64_t od = 0, id = 0;
    int32_t v = 0;
    uint64_t all = 0, vw = 0;
    if (i < nq) {
      v = QN[i];
      const int32_t sv = A.slot_of[v];
      all = nx[sv];
      vw = A.vis[v];
      // pushed lanes that had reached v before (pass 0 sets them unchecked)
      const uint64_t stale = all & qm & vw;
      if (stale) {
        all &= ~stale;
        nx[sv] = all;
      }
      nw = all & qm;
    }
    // side entry: the mask's two words and one per new lane, rounded up to
    // an even count; one atomic per wavefront
    const int32_t words = nw ? (2 + (int32_t)__popcll(nw) + 1) & ~1 : 0;
    const int32_t incl = wave_scan_add(words);
    const int32_t total = __builtin_amdgcn_readlane(incl, kWave - 1);
    int64_t b = -1;
    if (total) {
      unsigned long long off = 0;
      if (lane == 0) off = atomicAdd(pa, (unsigned long long)total);
      off = (unsigned long long)shfl64((uint64_t)off, 0) + (unsigned long long)(incl - words);
      if (words && (int64_t)off + words <= region) b = rbase + (int64_t)off;
    }
    if (i < nq) {
      if (b >= 0) {
        // the entry is 8-byte aligned (even offsets in even regions)
        uint2 *e = reinterpret_cast<uint2 *>(&A.side[b]);
        e[0] = make_uint2((uint32_t)nw, (uint32_t)(nw >> 32));
        const int np2 = ((int)__popcll(nw) + 1) >> 1;
        for (int k = 1; k <= np2; k++) e[k] = make_uint2(~0u, ~0u);  // +inf candidates
        A.pend[v] = (int32_t)b;
      } else {
        uint64_t lanes = nw;
        while (lanes) {
          int l = __ffsll((unsigned long long)lanes) - 1;
          lanes &= lanes - 1;
          *gmem(&so.par[l][v]) = -1;  // 0xFFFFFFFF: +inf for the unsigned atomicMin
          *gmem(&so.lev[l][v]) = so.level[l];
        }
      }
      if (nw) A.vis[v] = vw | nw;
      od = A.out_ptr[v + 1] - A.out_ptr[v];
      id = A.symmetric ? od : A.in_ptr[v + 1] - A.in_ptr[v];
      if (nw && (all & ~qm) == 0) {
        nf++;
        mf += (unsigned long long)od;
      }
      live |= nw;
      pairs += (unsigned long long)__popcll(nw);
    }
    if (tally_on) tally_words(nw, clamp_deg(od), clamp_deg(id), A.symmetric != 0, tally);
  }
  if (tally_on) tally_flush(tally, A.lanes, tl);
  live_flush(live, &c->live);
  unsigned long long s = ms_block_sum(nf, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->nf, s);
  s = ms_block_sum(pairs, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->pairs, s);
  s = ms_block_sum(mf, red);
  if (threadIdx.x == 0 && s) atomicAdd(&c->mf, s);
}

// The outputs of the pairs the previous round's push deferred to the side
// table, in a round that does not pull (a pull writes them in its chunk
// flush): one thread per vertex of that round's list, level d for each lane
// of the entry's mask, the entry's candidate as the parent.
// Every thread of the block calls it.
__device__ void push_out_part(const MsArgs &A, const RoundState &R) {
  __shared__ LaneOutLds so;
  stage_out(A.out, R.d, so);
  const int32_t *L = A.q[R.d & 1];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < R.pend_n; i += stride) {
    const int32_t v = L[i];
    const int32_t b = A.pend[v];
    if (b < 0) continue;
    uint64_t m = *reinterpret_cast<const uint64_t *>(&A.side[b]);
    for (int k = 2; m; k++) {
      const int l = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      *gmem(&so.par[l][v]) = (int32_t)A.side[b + k];
      *gmem(&so.lev[l][v]) = (int16_t)(so.level[l] - 1);
    }
    A.pend[v] = -1;
  }
}

// Clears the level-d frontier words so that buffer is zero when it becomes
// the next-words buffer of round d+1: through their list (Q[d]) when it is
// short, else by sweeping every word.
__global__ void __launch_bounds__(kMsBlock) ms_clear(MsArgs A) {
  MsCtrl *c = A.ctrl;
  const RoundState R = round_state(c);
  const int d = R.d;
  uint64_t *fr = A.fw[d & 1];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (clear_by_list(A, R)) {
    const int32_t *Q = A.q[d & 1];
    const int64_t nq = (int64_t)ld_ctrl(&c->qtail[d & 1].v);
    for (int64_t i = tid; i < nq; i += stride) fr[A.slot_of[Q[i]]] = 0;
  } else {
    for (int64_t k = tid; k < A.n; k += stride) fr[k] = 0;
  }
}

// First kernel of a round: the outputs the previous round's push deferred
// (when this round does not pull; a pull writes them with its chunks), the
// frontier bits and the active segments of the pull part.  Each part gates
// itself on the control block.
__global__ void __launch_bounds__(kMsBlock) ms_begin(MsArgs A) {
  const RoundState R = round_state(A.ctrl);
  if (R.pend_n && !R.pm) push_out_part(A, R);
  if (R.pm && R.use_fb) mark_part(A, R);
  if (R.pm && A.pin.nbig) active_segs_part(A, R, 0);
}

// First kernel of a push part: the frontier list after a pull part, and
// the active segments of the push lanes' big frontier vertices.
__global__ void __launch_bounds__(kMsBlock) ms_push_begin(MsArgs A, const int32_t *vertex_of) {
  const RoundState R = round_state(A.ctrl);
  if (!R.qm) return;
  if (!R.listed) compact_part(A, R, vertex_of);
  if (A.pout.nbig) active_segs_part(A, R, 1);
}

// One wavefront; ends round d.  Lanes that reached nothing this round drop
// out of the next one.  CGBFS_DIR_BEAMER moves every lane together by the
// union counts (nf = vertices with a new lane, mf = their out-degree sum);
// CGBFS_DIR_TAU and CGBFS_DIR_BEAMER_LANES apply the single-root rules
// (bfs.hip's bfs_decide) to each lane's own tallies, thread l for lane l.
// Thread 0 records the round and ends the search once nothing was reached.
// relist: the round's new vertices are not all listed (a pull part lists
// none), so the next round's push part rebuilds the list from an empty tail.
__global__ void ms_decide(MsArgs A, MsParams P) {
  if (blockIdx.x != 0) return;
  MsCtrl *c = A.ctrl;
  if (ld_ctrl(&c->done)) return;
  MsLane *L = A.lanes;
  const int l = threadIdx.x;
  const int d = (int)ld_ctrl(&c->rounds);
  const uint64_t live = (uint64_t)ld_ctrl(&c->live) & A.act;
  const long long pnf = (long long)ld_ctrl(&c->pnf);
  const long long nf = pnf + (long long)ld_ctrl(&c->nf);
  int g = ld_ctrl(&c->gdir);
  uint64_t pm = 0, qm = 0;
  if (lane_policy(P.mode)) {
    bool lv = false;
    int dir = 0;
    if (l < A.nroots) {
      const long long cnt = (long long)ld_ctrl(&L->cnt[l]);
      const long long mfl = (long long)ld_ctrl(&L->out[l]);
      const long long mu = ld_ctrl(&L->mu[l]) - (long long)ld_ctrl(&L->in[l]);
      const long long prev = ld_ctrl(&L->prev[l]);
      dir = ld_ctrl(&L->dir[l]);
      st_ctrl(&L->cnt[l], 0ull);
      st_ctrl(&L->out[l], 0ull);
      st_ctrl(&L->in[l], 0ull);
      if (P.mode == CGBFS_DIR_TAU) {
        if (dir == 0 && cnt > P.tau) dir = 1;  // one-way, bool_pe.csl:1078-1088
      } else if (dir == 0) {
        if ((double)mfl > (double)mu / P.alpha) dir = 1;
      } else if ((double)cnt < (double)A.n / P.beta && cnt < prev) {
        dir = 0;
      }
      st_ctrl(&L->mu[l], mu);
      st_ctrl(&L->prev[l], cnt);
      st_ctrl(&L->dir[l], dir);
      lv = (live >> l) & 1ull;
    }
    pm = __ballot(lv && dir == 1);
    qm = __ballot(lv && dir == 0);
  } else {
    if (P.mode == CGBFS_DIR_TOPDOWN) {
      g = 0;
    } else if (P.mode == CGBFS_DIR_BOTTOMUP) {
      g = 1;
    } else if (g == 0) {
      if ((double)ld_ctrl(&c->mf) * P.alpha > (double)P.nnz) g = 1;
    } else if ((double)nf * P.beta < (double)A.n && nf < ld_ctrl(&c->prev_nf)) {
      g = 0;
    }
    pm = g ? live : 0ull;
    qm = g ? 0ull : live;
  }
  // the side table fills again from the start in the next push round; this
  // round's entries are read (by the next round's pull or ms_push_out)
  // before that round's ms_prep allocates
  st_ctrl(&A.palloc[l * kPallocStride], 0ull);
  if (l != 0) return;
  const unsigned long long cur_pm = ld_ctrl(&c->pmask), cur_qm = ld_ctrl(&c->qmask);
  const long long pscanned = (long long)ld_ctrl(&c->pscanned);
  if (d < kMsHist) {
    MsHistRow *H = &A.hist->r[d];
    H->nf = nf;
    H->pnf = pnf;
    H->pscanned = pscanned;
    H->prows = (long long)ld_ctrl(&c->prows);
    H->ppairs = (long long)ld_ctrl(&c->ppairs);
    H->scanned = H->pscanned + (long long)ld_ctrl(&c->scanned);
    H->rows = H->prows + (long long)ld_ctrl(&c->rows);
    H->pairs = H->ppairs + (long long)ld_ctrl(&c->pairs);
    H->plane = __popcll(cur_pm);
    H->dir = cur_pm ? (cur_qm ? 2 : 1) : 0;
  }
  st_ctrl(&c->nf, 0ull);
  st_ctrl(&c->scanned, 0ull);
  st_ctrl(&c->rows, 0ull);
  st_ctrl(&c->pairs, 0ull);
  st_ctrl(&c->pnf, 0ull);
  st_ctrl(&c->pscanned, 0ull);
  st_ctrl(&c->prows, 0ull);
  st_ctrl(&c->ppairs, 0ull);
  st_ctrl(&c->mf, 0ull);
  st_ctrl(&c->live, 0ull);
  st_ctrl(&c->segtail[0].v, 0ull);
  st_ctrl(&c->segtail[1].v, 0ull);
  st_ctrl(&c->segtail_late, 0ull);
  reset_ms_cursors(A.cursor);
  // a pull visits every vertex: the one the count reaches wrote them
  if (cur_pm) {
    const int left = ld_ctrl(&c->noin_left);
    if (left > 0) st_ctrl(&c->noin_left, left - 1);
  }
  // A pull rescans every row it opens from the start, so the next pull reads
  // at most what this one read: below one entry per vertex it is light, and
  // its chunks, most of which have nothing to scan, are split statically.
  st_ctrl(&c->grab, cur_pm && pscanned < (long long)A.n ? -1 : P.dense_grab);
  // A pull part lists none of its new vertices: the next round's frontier is
  // listed only if this round did not pull, and a push part after a pull
  // rebuilds the list from an empty tail.
  const bool listed_next = cur_pm == 0;
  // the vertices this round's push part listed hold the pairs it deferred
  st_ctrl(&c->pend_n, cur_qm ? (long long)ld_ctrl(&c->qtail[(d + 1) & 1].v) : 0ll);
  st_ctrl(&c->qtail[d & 1].v, 0ull);
  if (!listed_next) st_ctrl(&c->qtail[(d + 1) & 1].v, 0ull);
  st_ctrl(&c->rounds, (long long)(d + 1));
  st_ctrl(&c->prev_nf, nf);
  st_ctrl(&c->gdir, g);
  st_ctrl(&c->listed, listed_next ? 1 : 0);
  st_ctrl(&c->use_fb, nf * kFbDenom <= (long long)A.n ? 1 : 0);
  bool done = nf == 0;
  if (!done && d + 1 >= P.max_levels) {
    st_ctrl(&c->status, (int)CGBFS_ETIMEOUT);
    done = true;
  }
  st_ctrl(&c->pmask, done ? 0ull : (unsigned long long)pm);
  st_ctrl(&c->qmask, done ? 0ull : (unsigned long long)qm);
  if (done) {
    st_ctrl(&c->listed, 1);  // nothing left to clear
    st_ctrl(&c->done, 1);
    if (A.host_done) {  // the mailbox: status and rounds, then the flag
      __hip_atomic_store(A.host_done + 1, ld_ctrl(&c->status), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(reinterpret_cast<long long *>(A.host_done + 2), (long long)(d + 1),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      __hip_atomic_store(A.host_done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// Outputs of the vertices without in-edges for a list of lanes: no root but
// the vertex itself reaches them, so they are known when a lane starts (-1;
// a seed's own entry is written afterwards).  The serving engine writes them
// when queries take lanes; a one-call search writes them in its first pull
// round instead (ms_pull), or in ms_finalize if no round pulls.  Plane-major:
// wavefront item k is lane l = k / nch and vertices [64c, 64c + 64),
// c = k % nch, so every output row is written front to back.
// A list of lanes, passed by value.
struct LaneList {
  int32_t count;
  int8_t lane[64];
};

__global__ void __launch_bounds__(kMsBlock) ms_fill_noin(const LaneOut *out, const uint64_t *noin,
                                                         int64_t n, LaneList ll) {
  constexpr int kGroup = 8;  // chunks per work item: their bitmap words load together
  __shared__ int32_t *sp[64];
  __shared__ int16_t *sv[64];
  if ((int)threadIdx.x < ll.count) {
    sp[threadIdx.x] = out->par[ll.lane[threadIdx.x]];
    sv[threadIdx.x] = out->lev[ll.lane[threadIdx.x]];
  }
  __syncthreads();
  const int lane = lane_id();
  const int64_t nch = (n + kWave - 1) / kWave;
  const int64_t ngrp = (nch + kGroup - 1) / kGroup;
  const int64_t total = ngrp * ll.count;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  for (int64_t k = wave; k < total; k += nwaves) {
    const int64_t j = k / ngrp, c0 = (k - j * ngrp) * kGroup;
    uint64_t w[kGroup];
#pragma unroll
    for (int i = 0; i < kGroup; i++) w[i] = c0 + i < nch ? noin[c0 + i] : 0ull;
    int32_t *pp = sp[j];
    int16_t *lp = sv[j];
#pragma unroll
    for (int i = 0; i < kGroup; i++) {
      if ((w[i] >> lane) & 1ull) {  // a set bit implies v < n
        const int64_t v = (c0 + i) * kWave + lane;
        st_once(&pp[v], (int32_t)-1);
        st_once(&lp[v], (int16_t)-1);
      }
    }
  }
}

// The seeds' own outputs (parent = itself, level 0), after the last round.
__global__ void ms_seed_outputs(MsArgs A, const int64_t *sv, const int32_t *sl, int64_t nseeds) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nseeds) return;
  A.out->par[sl[k]][sv[k]] = (int32_t)sv[k];
  A.out->lev[sl[k]][sv[k]] = 0;
}

// Final pass: -1 into every (lane, vertex) pair no lane word reached.  A
// wavefront owns 64 consecutive vertices and walks the lane bits missing in
// any of them, so each store instruction covers 64 consecutive entries of
// one lane's output row.
__global__ void __launch_bounds__(kMsBlock) ms_finalize(MsArgs A) {
  const int lane = lane_id();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // vertices without in-edges were written by a pull round, if one came
  const bool noin_done = ld_ctrl(&A.ctrl->noin_left) == 0;
  for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x - lane; v0 < A.n; v0 += stride) {
    const int64_t v = v0 + lane;
    const bool skip = v >= A.n || (noin_done && ((A.noin[v0 >> 6] >> lane) & 1ull));
    const uint64_t miss = skip ? 0ull : A.act & ~A.vis[v];
    uint64_t bits = wave_or64(miss);
    while (bits) {
      const int l = __ffsll((unsigned long long)bits) - 1;
      bits &= bits - 1;
      if ((miss >> l) & 1ull) {
        st_once(&A.out->par[l][v], (int32_t)-1);
        st_once(&A.out->lev[l][v], (int16_t)-1);
      }
    }
  }
}

// GTEPS m per lane: sum of the per-vertex weight over each lane's reached
// set.  A wavefront stages its 64 vertices' (word, weight) pairs in LDS and
// lane l then walks them for lane bit l: a transpose through LDS with
// broadcast reads, no cross-lane reductions.
__global__ void __launch_bounds__(kMsBlock) ms_count(const uint64_t *vis, const int32_t *w, int64_t n,
                                                     int nroots,
                                                     unsigned long long *out) {
  __shared__ unsigned long long acc[64];
  __shared__ uint64_t sb[kMsWaves][kWave];
  __shared__ uint32_t sw[kMsWaves][kWave];
  if (threadIdx.x < 64) acc[threadIdx.x] = 0;
  __syncthreads();
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  unsigned long long mine = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x - lane; v0 < n; v0 += stride) {
    const int64_t v = v0 + lane;
    const uint32_t wt = v < n ? (uint32_t)w[v] : 0u;
    const uint64_t b = wt ? vis[v] : 0ull;
    if (!__any(b != 0)) continue;
    sb[wv][lane] = b;
    sw[wv][lane] = wt;
    wave_sync();
    for (int r = 0; r < kWave; r++)
      if ((sb[wv][r] >> lane) & 1ull) mine += sw[wv][r];
    wave_sync();
  }
  if (mine) atomicAdd(&acc[lane], mine);
  __syncthreads();
  if (threadIdx.x < (unsigned)nroots && acc[threadIdx.x]) atomicAdd(&out[threadIdx.x], acc[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// Serving engine kernels: a long-lived pass whose lanes take queries as they
// come (continuous batching).  Lanes join and leave at round boundaries.
// ---------------------------------------------------------------------------

// The queries seeded in one engine step, passed by value: lane lane[k]
// starts at vertex root[k].
struct SeedSet {
  int32_t count;
  int8_t lane[64];
  int64_t root[64];
};

// Idle engine: no lane in flight, round counter at 0.  One wavefront.
__global__ void eng_init(MsArgs A) {
  MsCtrl *c = A.ctrl;
  MsLane *L = A.lanes;
  const int l = threadIdx.x;
  if (l < 64) {
    st_ctrl(&L->cnt[l], 0ull);
    st_ctrl(&L->out[l], 0ull);
    st_ctrl(&L->in[l], 0ull);
    st_ctrl(&L->mu[l], 0ll);
    st_ctrl(&L->prev[l], 0ll);
    st_ctrl(&L->dir[l], 0);
  }
  if (l == 0) {
    st_ctrl(&c->qtail[0].v, 0ull);
    st_ctrl(&c->qtail[1].v, 0ull);
    st_ctrl(&c->nf, 0ull);
    st_ctrl(&c->scanned, 0ull);
    st_ctrl(&c->rows, 0ull);
    st_ctrl(&c->pairs, 0ull);
    st_ctrl(&c->pnf, 0ull);
    st_ctrl(&c->pscanned, 0ull);
    st_ctrl(&c->prows, 0ull);
    st_ctrl(&c->ppairs, 0ull);
    st_ctrl(&c->mf, 0ull);
    st_ctrl(&c->live, 0ull);
    st_ctrl(&c->segtail[0].v, 0ull);
    st_ctrl(&c->segtail[1].v, 0ull);
    st_ctrl(&c->segtail_late, 0ull);
    st_ctrl(&c->pmask, 0ull);
    st_ctrl(&c->qmask, 0ull);
    st_ctrl(&c->prev_nf, 0ll);
    st_ctrl(&c->rounds, 0ll);
    st_ctrl(&c->gdir, 0);
    st_ctrl(&c->listed, 1);
    st_ctrl(&c->use_fb, 0);
    st_ctrl(&c->done, 1);
    st_ctrl(&c->status, 0);
    st_ctrl(&c->grab, (int)kGrabDense);
    st_ctrl(&c->noin_left, 0);  // retiring lanes write their own
    st_ctrl(&c->pend_n, 0ll);
    reset_ms_cursors(A.cursor);
  }
}

// Finishes the queries of lanes R: -1 into every vertex they did not reach
// (vertices without in-edges were filled when the query was seeded), their
// GTEPS m into mlane[l], and their bits cleared from the lane words so the
// lanes can take new queries.  Wavefront layout as in ms_finalize (64
// consecutive vertices, one coalesced store per lane bit) and ms_count (m
// through an LDS transpose).
__global__ void __launch_bounds__(kMsBlock) eng_retire(MsArgs A, uint64_t R, const int32_t *weight,
                                                       unsigned long long *mlane) {
  __shared__ LaneOutLds so;
  stage_out(A.out, 0, so);
  __shared__ unsigned long long acc[64];
  __shared__ uint64_t sb[kMsWaves][kWave];
  __shared__ uint32_t sw[kMsWaves][kWave];
  if (threadIdx.x < 64) acc[threadIdx.x] = 0;
  __syncthreads();
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  unsigned long long mine = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x - lane; v0 < A.n; v0 += stride) {
    const int64_t v = v0 + lane;
    const bool inr = v < A.n;
    const bool noin = inr && ((A.noin[v0 >> 6] >> lane) & 1ull);
    const uint64_t w = inr ? A.vis[v] : 0ull;
    const uint64_t miss = (inr && !noin) ? R & ~w : 0ull;
    uint64_t bits = wave_or64(miss);
    while (bits) {
      const int l = __ffsll((unsigned long long)bits) - 1;
      bits &= bits - 1;
      if ((miss >> l) & 1ull) {
        st_once(&so.par[l][v], (int32_t)-1);
        st_once(&so.lev[l][v], (int16_t)-1);
      }
    }
    const uint64_t hit = w & R;
    if (hit) A.vis[v] = w & ~R;
    if (!__any(hit != 0)) continue;
    sb[wv][lane] = hit;
    sw[wv][lane] = hit ? (uint32_t)weight[v] : 0u;
    wave_sync();
    for (int r = 0; r < kWave; r++)
      if ((sb[wv][r] >> lane) & 1ull) mine += sw[wv][r];
    wave_sync();
  }
  if (mine) atomicAdd(&acc[lane], mine);
  __syncthreads();
  if (threadIdx.x < 64 && acc[threadIdx.x]) atomicAdd(&mlane[threadIdx.x], acc[threadIdx.x]);
}

// Seeds the queries of S at the engine's current round: the lane words, the
// lanes' direction state and level base, the roots' own outputs (after the
// in-edge-free fill of their rows), and the control block.  The new lanes
// join this round in their starting direction; the round rebuilds its
// frontier list from the lane words, since the roots are not on it.  One
// wavefront.
__global__ void eng_seed(MsArgs A, LaneOut *out, SeedSet S, MsParams P) {
  MsCtrl *c = A.ctrl;
  MsLane *L = A.lanes;
  const int t = threadIdx.x;
  const int d = (int)ld_ctrl(&c->rounds);
  uint64_t mine = 0;
  int dir = 0;
  if (t < S.count) {
    const int l = S.lane[t];
    const int64_t r = S.root[t];
    const int32_t sr = A.slot_of[r];
    mine = 1ull << l;
    atomicOr((unsigned long long *)&A.vis[r], (unsigned long long)mine);
    atomicOr((unsigned long long *)&A.fw[d & 1][sr], (unsigned long long)mine);
    out->base[l] = d;
    *gmem(&out->par[l][r]) = (int32_t)r;
    *gmem(&out->lev[l][r]) = (int16_t)0;
    dir = P.mode == CGBFS_DIR_BOTTOMUP ? 1 : 0;
    st_ctrl(&L->cnt[l], 0ull);
    st_ctrl(&L->out[l], 0ull);
    st_ctrl(&L->in[l], 0ull);
    st_ctrl(&L->mu[l], (long long)P.nnz - (long long)(A.in_ptr[r + 1] - A.in_ptr[r]));
    st_ctrl(&L->prev[l], 1ll);
    st_ctrl(&L->dir[l], dir);
  }
  const uint64_t added = wave_or64(mine);
  const uint64_t pull = wave_or64(dir ? mine : 0ull);
  if (t == 0 && added) {
    if (ld_ctrl(&c->done)) {
      // An idle engine starts the union-count policy afresh, as a one-call
      // search does: pushing (unless every round pulls), with the seeds as
      // the previous frontier.
      st_ctrl(&c->gdir, P.mode == CGBFS_DIR_BOTTOMUP ? 1 : 0);
      st_ctrl(&c->prev_nf, (long long)S.count);
    }
    st_ctrl(&c->pmask, (unsigned long long)(ld_ctrl(&c->pmask) | pull));
    st_ctrl(&c->qmask, (unsigned long long)(ld_ctrl(&c->qmask) | (added & ~pull)));
    st_ctrl(&c->listed, 0);
    st_ctrl(&c->qtail[d & 1].v, 0ull);
    st_ctrl(&c->done, 0);
  }
}

__global__ void ms_zero(unsigned long long *p, int k) {
  if (threadIdx.x < (unsigned)k) p[threadIdx.x] = 0;
}

}  // namespace cg

using namespace cg;

// The multi-source search behind both entry points: nseeds seeds (vertex
// sv[k] starts lane sl[k]) over nroots lanes.
// Blocks of a pull launch: at most as many as are resident on the device at
// once.  The pull hands out its chunks dynamically, so blocks that could only
// start after others finish would add a tail and nothing else.
static int pull_grid(int device, int blocks, bool tally) {
  int occ = 0, occ2 = 0, cus = 0;
  hipError_t e = tally
      ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, ms_pull<true, false>, kMsBlock, 0)
      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, ms_pull<false, false>, kMsBlock, 0);
  if (e == hipSuccess)
    e = tally ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ2, ms_pull<true, true>, kMsBlock, 0)
              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ2, ms_pull<false, true>, kMsBlock, 0);
  occ = std::min(occ, occ2);
  if (e != hipSuccess || occ <= 0 ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      cus <= 0)
    return blocks;
  return std::min(blocks, occ * cus);
}

static int ms_run(cgbfs_graph *h, const int64_t *sv, const int32_t *sl, int64_t nseeds,
                  int32_t nroots, int32_t *parents, int16_t *levels, int32_t on_device,
                  const cgbfs_options *opt, cgbfs_stats *stats, int64_t *m_lanes) {
  DeviceGraph *g = &h->g;
  cgbfs_options o;
  fill_options(opt, g->n, true, &o);
  int rc0 = check_options(&o);
  if (rc0) return rc0;
  if (o.max_levels > 32767) o.max_levels = 32767;  // int16 levels
  CG_HIP(hipSetDevice(g->device));
  MsWorkspace ws;
  MsGraphData *md = nullptr;
  int rc = ensure_wsm(g, &ws, &md);
  if (rc) return rc;
  hipStream_t st = g->stream;
  // a call that failed after enqueueing its copy may not have waited for it
  if (md->stage_busy) CG_HIP(hipStreamSynchronize(st));
  md->stage_busy = false;
  const size_t sv_at = al(sizeof(LaneOut)), sl_at = sv_at + al(sizeof(int64_t) * nseeds);
  const size_t seed_bytes = sl_at + sizeof(int32_t) * nseeds;
  if (md->seedcap < seed_bytes) {
    if (md->seedbuf) hipFree(md->seedbuf);
    if (md->seedstage) hipHostFree(md->seedstage);
    md->seedbuf = md->seedstage = nullptr;
    md->seedcap = 0;
    CG_HIP(hipMalloc(&md->seedbuf, seed_bytes));
    CG_HIP(hipHostMalloc(&md->seedstage, seed_bytes, hipHostMallocDefault));
    md->seedcap = seed_bytes;
  }
  int64_t *dsv = (int64_t *)((char *)md->seedbuf + sv_at);
  int32_t *dsl = (int32_t *)((char *)md->seedbuf + sl_at);

  const size_t plane = (size_t)g->n * nroots;
  int32_t *dpar = parents;
  int16_t *dlev = levels;
  if (!on_device) {
    const size_t need = al(plane * sizeof(int32_t)) + plane * sizeof(int16_t) + 256;
    if (md->outcap < need) {
      if (md->outbuf) hipFree(md->outbuf);
      md->outbuf = nullptr;
      md->outcap = 0;
      CG_HIP(hipMalloc(&md->outbuf, need));
      md->outcap = need;
    }
    dpar = (int32_t *)md->outbuf;
    dlev = (int16_t *)((char *)md->outbuf + al(plane * sizeof(int32_t)));
  }

  MsArgs A;
  std::memset(&A, 0, sizeof(A));  // padding too: A is part of the graph-cache key
  A.out_ptr = g->out_ptr;
  A.out_idx = g->out_idx;
  A.in_ptr = g->in_ptr;
  A.in_idx = g->in_idx;
  A.out = (LaneOut *)md->seedbuf;
  // staged: the output table, then the seeds (the previous call has
  // finished with the staging buffer: every call ends synchronised)
  char *stage = (char *)md->seedstage;
  LaneOut &lo = *(LaneOut *)stage;
  std::memset(&lo, 0, sizeof(lo));
  for (int l = 0; l < nroots; l++) {
    lo.par[l] = dpar + (size_t)l * g->n;
    lo.lev[l] = dlev + (size_t)l * g->n;
  }
  std::memcpy(stage + sv_at, sv, sizeof(int64_t) * nseeds);
  std::memcpy(stage + sl_at, sl, sizeof(int32_t) * nseeds);
  A.vis = ws.vis;
  A.fw[0] = ws.fw[0];
  A.fw[1] = ws.fw[1];
  A.fb = ws.fb;
  A.q[0] = ws.q[0];
  A.q[1] = ws.q[1];
  A.ctrl = ws.ctrl;
  A.hist = ws.hist;
  A.lanes = ws.lanes;
  A.pin = md->pin;
  A.pout = md->pout;
  A.slot_of = md->slot_of;
  A.in_slot = md->in_slot;
  A.out_slot = md->out_slot;
  A.cand = ws.cand;
  A.bgot = ws.bgot;
  A.segq[0] = ws.segq[0];
  A.segq[1] = ws.segq[1];
  A.segneed = ws.segneed;
  A.segq_late = ws.segq_late;
  A.segneed_late = ws.segneed_late;
  A.cursor = ws.cursor;
  A.pend = ws.pend;
  A.side = ws.side;
  A.palloc = ws.palloc;
  A.side_cap = ws.side_cap;
  A.noin = md->noin;
  A.host_done = nullptr;
  A.n = g->n;
  A.nroots = nroots;
  A.act = nroots == 64 ? ~0ull : ((1ull << nroots) - 1ull);
  A.symmetric = g->in_ptr == g->out_ptr ? 1 : 0;
  A.seg_grab = o.seg_grab;

  const int blocks = grid_blocks(g->device, o.grid_mult);
  const int pblocks = pull_grid(g->device, blocks, lane_policy(o.direction));
  std::vector<hipEvent_t> evs;
  auto new_event = [&](hipEvent_t *e) -> hipError_t {
    hipError_t r = hipEventCreate(e);
    if (r == hipSuccess) evs.push_back(*e);
    return r;
  };
  auto cleanup = [&]() {
    for (hipEvent_t e : evs) hipEventDestroy(e);
  };
  hipEvent_t e0, e1;
  hipError_t e = hipMemcpyAsync(md->seedbuf, stage, seed_bytes, hipMemcpyHostToDevice, st);
  md->stage_busy = true;
  if (e == hipSuccess) e = new_event(&e0);
  if (e == hipSuccess) e = new_event(&e1);
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  if (e != hipSuccess) {
    cleanup();
    return hip_fail(e, "msbfs setup", __FILE__, __LINE__);
  }
  ms_init_fill<<<blocks, kMsBlock, 0, st>>>(A);
  MsParams P;
  std::memset(&P, 0, sizeof(P));
  P.mode = o.direction;
  P.alpha = o.alpha;
  P.beta = o.beta;
  P.tau = (long long)o.tau;
  P.max_levels = (long long)o.max_levels;
  P.nnz = (long long)g->nnz;
  P.dense_grab = o.pull_grab > 0 ? o.pull_grab : kGrabDense;
  ms_init_seeds<<<1, 64, 0, st>>>(A, dsv, dsl, nseeds, P);
  MsCtrl hc;
  std::memset(&hc, 0, sizeof(hc));
  const bool replay = o.capture != 0;
  struct RoundEvents { hipEvent_t a, k, p, b; };
  std::vector<RoundEvents> rev;
  const int tally = lane_policy(P.mode) ? 1 : 0;
  // Every kernel of a round gates itself on the control block, so launching
  // a part that is off costs only an empty dispatch.
  auto pull_kernels = [&]() {
    if (md->pin.nbig) {
      ms_pull_segs<<<blocks, kMsBlock, 0, st>>>(A, 0);
      ms_pull_segs<<<blocks, kMsBlock, 0, st>>>(A, 1);
    }
    // both instantiations when the round is not known here (replay)
    const bool np = hc.pend_n == 0 || replay, pp = hc.pend_n != 0 || replay;
    if (tally) {
      if (np) ms_pull<true, false><<<pblocks, kMsBlock, 0, st>>>(A);
      if (pp) ms_pull<true, true><<<pblocks, kMsBlock, 0, st>>>(A);
    } else {
      if (np) ms_pull<false, false><<<pblocks, kMsBlock, 0, st>>>(A);
      if (pp) ms_pull<false, true><<<pblocks, kMsBlock, 0, st>>>(A);
    }
  };
  auto push_part = [&](bool begin) {
    if (begin) ms_push_begin<<<blocks, kMsBlock, 0, st>>>(A, md->vertex_of);
    ms_push<<<blocks, kMsBlock, 0, st>>>(A, 0);
    ms_prep<<<blocks, kMsBlock, 0, st>>>(A, tally);
    ms_push<<<blocks, kMsBlock, 0, st>>>(A, 1);
  };
  if (o.capture) {
    // Graph replay: kGraphRounds rounds of the full kernel sequence are
    // captured once and replayed until ms_decide raises the pinned done flag
    // (one host wait per replay, none per round; rounds past the end gate
    // themselves off).  No per-round events in this mode.
    volatile int32_t *flag = g->host_flag;
    *flag = 0;
    A.host_done = (int *)g->host_flag;
    // the captured launches depend only on these; a repeated call with the
    // same arguments replays the graph instantiated last time
    std::vector<char> key(sizeof(A) + sizeof(P) + 2 * sizeof(int));
    std::memcpy(key.data(), &A, sizeof(A));
    std::memcpy(key.data() + sizeof(A), &P, sizeof(P));
    std::memcpy(key.data() + sizeof(A) + sizeof(P), &blocks, sizeof(int));
    std::memcpy(key.data() + sizeof(A) + sizeof(P) + sizeof(int), &tally, sizeof(int));
    if (md->exec && md->exec_key != key) {
      hipGraphExecDestroy(md->exec);
      md->exec = nullptr;
    }
    if (!md->exec) {
      hipGraph_t graph = nullptr;
      // Relaxed: searches on other graph handles may run from other host
      // threads meanwhile, and their synchronous calls must not invalidate
      // this capture (it only enqueues this handle's kernels).
      e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
      if (e == hipSuccess) {
        for (int r = 0; r < kGraphRounds; r++) {
          ms_begin<<<blocks, kMsBlock, 0, st>>>(A);
          pull_kernels();
          push_part(true);
          ms_clear<<<blocks, kMsBlock, 0, st>>>(A);
          ms_decide<<<1, 64, 0, st>>>(A, P);
        }
        hipError_t ec = hipStreamEndCapture(st, &graph);
        e = ec != hipSuccess ? ec : hipGetLastError();
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&md->exec, graph, nullptr, nullptr, 0);
      if (e == hipSuccess) md->exec_key = key;
      else md->exec = nullptr;
      if (graph) hipGraphDestroy(graph);
    }
    for (int64_t it = 0; e == hipSuccess && !*flag && it * kGraphRounds <= o.max_levels + 1; it++) {
      e = hipGraphLaunch(md->exec, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    if (e == hipSuccess && *flag) {  // the mailbox holds what the host needs
      std::atomic_thread_fence(std::memory_order_acquire);
      hc.done = 1;
      hc.status = flag[1];
      hc.rounds = (long long)(((uint64_t)(uint32_t)flag[3] << 32) | (uint32_t)flag[2]);
    } else {
      if (e == hipSuccess) e = hipMemcpyAsync(&hc, A.ctrl, sizeof(MsCtrl), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    A.host_done = nullptr;
  } else {
    // One host poll per round: only the parts that are on are launched, and
    // each part is bracketed by events.
    e = hipMemcpyAsync(&hc, A.ctrl, sizeof(MsCtrl), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    while (e == hipSuccess && !hc.done) {
      RoundEvents re;
      if ((e = new_event(&re.a)) != hipSuccess || (e = new_event(&re.k)) != hipSuccess ||
          (e = new_event(&re.p)) != hipSuccess || (e = new_event(&re.b)) != hipSuccess)
        break;
      hipEventRecord(re.a, st);
      if (hc.pmask || hc.pend_n) ms_begin<<<blocks, kMsBlock, 0, st>>>(A);
      hipEventRecord(re.k, st);
      if (hc.pmask) pull_kernels();
      hipEventRecord(re.p, st);
      if (hc.qmask) push_part(!hc.listed || md->pout.nbig);
      hipEventRecord(re.b, st);
      rev.push_back(re);
      ms_clear<<<blocks, kMsBlock, 0, st>>>(A);
      ms_decide<<<1, 64, 0, st>>>(A, P);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(&hc, A.ctrl, sizeof(MsCtrl), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
  }
  // The final pass belongs to producing the outputs, so e1 follows it.
  if (e == hipSuccess) {
    ms_seed_outputs<<<(unsigned)((nseeds + 255) / 256), 256, 0, st>>>(A, dsv, dsl, nseeds);
    ms_finalize<<<blocks, kMsBlock, 0, st>>>(A);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) md->stage_busy = false;
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
      std::vector<MsHistRow> hb((size_t)nh);
      e = hipMemcpy(hb.data(), A.hist->r, sizeof(MsHistRow) * nh, hipMemcpyDeviceToHost);
      if (e != hipSuccess) {
        cleanup();
        return hip_fail(e, "msbfs stats", __FILE__, __LINE__);
      }
      for (int64_t r = 0; r < nh; r++) {
        const MsHistRow *H = &hb[(size_t)r];
        if (stats->nf_history) stats->nf_history[r] = H->nf;
        if (stats->direction_history) stats->direction_history[r] = H->dir;
        if (stats->edges_scanned) stats->edges_scanned[r] = H->scanned;
        if (stats->rows_scanned) stats->rows_scanned[r] = H->rows;
        if (stats->pairs_found) stats->pairs_found[r] = H->pairs;
        if (stats->pull_lanes) stats->pull_lanes[r] = H->plane;
        if (stats->pull_nf) stats->pull_nf[r] = H->pnf;
        if (stats->pull_rows) stats->pull_rows[r] = H->prows;
        if (stats->pull_edges) stats->pull_edges[r] = H->pscanned;
        if (stats->pull_pairs) stats->pull_pairs[r] = H->ppairs;
        if (r < (int64_t)rev.size()) {
          float rm = 0, pms = 0;
          hipEventElapsedTime(&rm, rev[r].a, rev[r].b);
          hipEventElapsedTime(&pms, rev[r].k, rev[r].p);
          if (stats->level_ms) stats->level_ms[r] = rm;
          if (stats->pull_ms) stats->pull_ms[r] = pms;
        }
      }
    }
  }
  if (m_lanes) {
    unsigned long long m[64];
    ms_zero<<<1, 64, 0, st>>>(ws.mlane, 64);
    ms_count<<<grid_blocks(g->device, 8), kMsBlock, 0, st>>>(ws.vis, g->m_weight, g->n, nroots,
                                                             ws.mlane);
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

extern "C" int cgbfs_msbfs(cgbfs_graph *h, const int64_t *roots, int32_t nroots, int32_t *parents,
                           int16_t *levels, int32_t on_device, const cgbfs_options *opt,
                           cgbfs_stats *stats, int64_t *m_lanes) {
  if (!h) return set_error(CGBFS_EINVAL, "graph is null");
  if (!roots || nroots < 1 || nroots > 64) return set_error(CGBFS_EINVAL, "nroots must be in [1, 64]");
  if (!parents || !levels) return set_error(CGBFS_EINVAL, "parents/levels must not be null");
  for (int l = 0; l < nroots; l++)
    if (roots[l] < 0 || roots[l] >= h->g.n) return set_error(CGBFS_EINVAL, "root out of range");
  int32_t lanes[64];
  for (int l = 0; l < nroots; l++) lanes[l] = l;
  return ms_run(h, roots, lanes, nroots, nroots, parents, levels, on_device, opt, stats, m_lanes);
}

extern "C" int cgbfs_msbfs_seeded(cgbfs_graph *h, const int64_t *verts, const int32_t *lanes,
                                  int64_t nseeds, int32_t nlanes, int32_t *parents, int16_t *levels,
                                  int32_t on_device, const cgbfs_options *opt, cgbfs_stats *stats,
                                  int64_t *m_lanes) {
  if (!h) return set_error(CGBFS_EINVAL, "graph is null");
  if (nlanes < 1 || nlanes > 64) return set_error(CGBFS_EINVAL, "nlanes must be in [1, 64]");
  if (!verts || !lanes || nseeds < 1) return set_error(CGBFS_EINVAL, "need at least one seed");
  if (!parents || !levels) return set_error(CGBFS_EINVAL, "parents/levels must not be null");
  for (int64_t k = 0; k < nseeds; k++) {
    if (verts[k] < 0 || verts[k] >= h->g.n) return set_error(CGBFS_EINVAL, "seed vertex out of range");
    if (lanes[k] < 0 || lanes[k] >= nlanes) return set_error(CGBFS_EINVAL, "seed lane out of range");
  }
  return ms_run(h, verts, lanes, nseeds, nlanes, parents, levels, on_device, opt, stats, m_lanes);
}

// ---------------------------------------------------------------------------
// Serving engine (continuous batching): host side
// ---------------------------------------------------------------------------

struct cgbfs_engine {
  cgbfs_graph *h = nullptr;
  MsGraphData *md = nullptr;
  void *mem = nullptr;       // the engine's own multi-source workspace
  MsWorkspace ws;
  MsArgs A;
  MsParams P;
  int blocks = 0;
  int pblocks = 0;  // pull launches
  int tally = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  hipGraphExec_t round = nullptr;  // one round + the copies back, captured once
  struct Pinned {            // read by the device copies of each step
    MsCtrl hc;
    unsigned long long m[64];
    LaneOut out;
  } *pin = nullptr;
  uint64_t busy = 0;         // lanes holding a query (pending, running or retiring)
  uint64_t pending = 0;      // submitted, seeded by the next step
  uint64_t retiring = 0;     // finished, finalised by the next step
  uint64_t retired = 0;      // finalised by the step in flight
  int64_t root[64];
  int64_t base[64];          // round each lane's query was seeded in
  int32_t depth[64];         // rounds each lane's query ran
  int64_t rounds_seen = 0;   // the device round counter at the last poll
  bool inflight = false;
};

// One round for every lane in flight: the gated kernel sequence.
static void eng_round(cgbfs_engine *E) {
  const MsArgs &A = E->A;
  hipStream_t st = E->st;
  const int blocks = E->blocks;
  MsGraphData *md = E->md;
  ms_begin<<<blocks, kMsBlock, 0, st>>>(A);
  if (md->pin.nbig) {
    ms_pull_segs<<<blocks, kMsBlock, 0, st>>>(A, 0);
    ms_pull_segs<<<blocks, kMsBlock, 0, st>>>(A, 1);
  }
  if (E->tally) {
    ms_pull<true, false><<<E->pblocks, kMsBlock, 0, st>>>(A);
    ms_pull<true, true><<<E->pblocks, kMsBlock, 0, st>>>(A);
  } else {
    ms_pull<false, false><<<E->pblocks, kMsBlock, 0, st>>>(A);
    ms_pull<false, true><<<E->pblocks, kMsBlock, 0, st>>>(A);
  }
  ms_push_begin<<<blocks, kMsBlock, 0, st>>>(A, md->vertex_of);
  ms_push<<<blocks, kMsBlock, 0, st>>>(A, 0);
  ms_prep<<<blocks, kMsBlock, 0, st>>>(A, E->tally);
  ms_push<<<blocks, kMsBlock, 0, st>>>(A, 1);
  ms_clear<<<blocks, kMsBlock, 0, st>>>(A);
  ms_decide<<<1, 64, 0, st>>>(A, E->P);
}

extern "C" int cgbfs_engine_create(cgbfs_graph *h, const cgbfs_options *opt, cgbfs_engine **out) {
  if (!h || !out) return set_error(CGBFS_EINVAL, "graph and out must not be null");
  DeviceGraph *g = &h->g;
  cgbfs_options o;
  fill_options(opt, g->n, true, &o);
  const int rc0 = check_options(&o);
  if (rc0) return rc0;
  // lanes run at different depths: the union-count rule becomes the per-lane one
  if (o.direction == CGBFS_DIR_BEAMER) o.direction = CGBFS_DIR_BEAMER_LANES;
  CG_HIP(hipSetDevice(g->device));
  cgbfs_engine *E = new cgbfs_engine();
  E->h = h;
  auto fail = [&](int rc) {
    cgbfs_engine_destroy(E);
    return rc;
  };
  int rc = ensure_msdata(g, &E->md);
  if (rc) return fail(rc);
  const size_t total = ws_layout(g->n, E->md, nullptr, &E->ws);
  hipError_t e = hipMalloc(&E->mem, total);
  if (e == hipSuccess) e = hipHostMalloc((void **)&E->pin, sizeof(cgbfs_engine::Pinned), hipHostMallocDefault);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&E->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&E->ev, hipEventDisableTiming);
  if (e != hipSuccess) return fail(hip_fail(e, "engine create", __FILE__, __LINE__));
  ws_layout(g->n, E->md, (char *)E->mem, &E->ws);
  std::memset(E->pin, 0, sizeof(cgbfs_engine::Pinned));
  MsArgs &A = E->A;
  std::memset(&A, 0, sizeof(A));
  A.out_ptr = g->out_ptr;
  A.out_idx = g->out_idx;
  A.in_ptr = g->in_ptr;
  A.in_idx = g->in_idx;
  A.out = E->ws.out;
  A.vis = E->ws.vis;
  A.fw[0] = E->ws.fw[0];
  A.fw[1] = E->ws.fw[1];
  A.fb = E->ws.fb;
  A.q[0] = E->ws.q[0];
  A.q[1] = E->ws.q[1];
  A.ctrl = E->ws.ctrl;
  A.hist = E->ws.hist;
  A.lanes = E->ws.lanes;
  A.pin = E->md->pin;
  A.pout = E->md->pout;
  A.slot_of = E->md->slot_of;
  A.in_slot = E->md->in_slot;
  A.out_slot = E->md->out_slot;
  A.cand = E->ws.cand;
  A.bgot = E->ws.bgot;
  A.segq[0] = E->ws.segq[0];
  A.segq[1] = E->ws.segq[1];
  A.segneed = E->ws.segneed;
  A.segq_late = E->ws.segq_late;
  A.segneed_late = E->ws.segneed_late;
  A.cursor = E->ws.cursor;
  A.pend = E->ws.pend;
  A.side = E->ws.side;
  A.palloc = E->ws.palloc;
  A.side_cap = E->ws.side_cap;
  A.noin = E->md->noin;
  A.host_done = nullptr;
  A.n = g->n;
  A.nroots = 64;
  A.act = ~0ull;
  A.symmetric = g->in_ptr == g->out_ptr ? 1 : 0;
  A.seg_grab = o.seg_grab;
  MsParams &P = E->P;
  std::memset(&P, 0, sizeof(P));
  P.mode = o.direction;
  P.alpha = o.alpha;
  P.beta = o.beta;
  P.tau = (long long)o.tau;
  P.max_levels = 0x7FFFFFFFFFFFFFFFll;  // queries end on their own; levels are per query
  P.nnz = (long long)g->nnz;
  P.dense_grab = o.pull_grab > 0 ? o.pull_grab : kGrabDense;
  E->tally = lane_policy(P.mode) ? 1 : 0;
  E->blocks = grid_blocks(g->device, o.grid_mult);
  E->pblocks = pull_grid(g->device, E->blocks, E->tally != 0);
  ms_init_fill<<<E->blocks, kMsBlock, 0, E->st>>>(A);
  eng_init<<<1, 64, 0, E->st>>>(A);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(E->st);
  if (e != hipSuccess) return fail(hip_fail(e, "engine init", __FILE__, __LINE__));
  // The round is the same launch sequence every step (each kernel gates
  // itself on the control block): capture it once, with the copies of the
  // control block and the m counters back to pinned memory.
  hipGraph_t graph = nullptr;
  e = hipStreamBeginCapture(E->st, hipStreamCaptureModeRelaxed);
  if (e == hipSuccess) {
    eng_round(E);
    e = hipMemcpyAsync(&E->pin->hc, A.ctrl, sizeof(MsCtrl), hipMemcpyDeviceToHost, E->st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(E->pin->m, E->ws.mlane, sizeof(E->pin->m), hipMemcpyDeviceToHost, E->st);
    hipError_t ec = hipStreamEndCapture(E->st, &graph);
    if (e == hipSuccess) e = ec;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&E->round, graph, nullptr, nullptr, 0);
  if (graph) hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(hip_fail(e, "engine round capture", __FILE__, __LINE__));
  *out = E;
  return CGBFS_OK;
}

extern "C" int cgbfs_engine_submit(cgbfs_engine *E, int32_t lane, int64_t root, int32_t *parents,
                                   int16_t *levels) {
  if (!E) return set_error(CGBFS_EINVAL, "engine is null");
  if (lane < 0 || lane > 63) return set_error(CGBFS_EINVAL, "lane must be in [0, 63]");
  if ((E->busy >> lane) & 1ull) return set_error(CGBFS_EINVAL, "lane is busy");
  if (root < 0 || root >= E->h->g.n) return set_error(CGBFS_EINVAL, "root out of range");
  if (!parents || !levels) return set_error(CGBFS_EINVAL, "parents/levels must not be null");
  if (E->inflight) return set_error(CGBFS_EINVAL, "poll the step in flight first");
  E->busy |= 1ull << lane;
  E->pending |= 1ull << lane;
  E->root[lane] = root;
  E->pin->out.par[lane] = parents;
  E->pin->out.lev[lane] = levels;
  return CGBFS_OK;
}

extern "C" int cgbfs_engine_step(cgbfs_engine *E) {
  if (!E) return set_error(CGBFS_EINVAL, "engine is null");
  if (E->inflight) return set_error(CGBFS_EINVAL, "poll the step in flight first");
  DeviceGraph *g = &E->h->g;
  CG_HIP(hipSetDevice(g->device));
  MsArgs &A = E->A;
  hipStream_t st = E->st;
  const int blocks = E->blocks;
  ms_zero<<<1, 64, 0, st>>>(E->ws.mlane, 64);
  // 1. finish the lanes whose search ended in the last step (the device table
  //    still holds their output rows)
  if (E->retiring) {
    eng_retire<<<blocks, kMsBlock, 0, st>>>(A, E->retiring, g->m_weight, E->ws.mlane);
  }
  E->retired = E->retiring;
  E->retiring = 0;
  // 2. seed the queued queries: output rows, in-edge-free entries, lane state.
  //    The whole table is uploaded, so the host copy carries every lane's
  //    level base too (the device round counter equals rounds_seen here).
  if (E->pending) {
    SeedSet S;
    LaneList ll;
    S.count = 0;
    for (int l = 0; l < 64; l++)
      if ((E->pending >> l) & 1ull) {
        S.lane[S.count] = (int8_t)l;
        S.root[S.count] = E->root[l];
        ll.lane[S.count] = (int8_t)l;
        S.count++;
        E->base[l] = E->rounds_seen;
        E->pin->out.base[l] = (int32_t)E->rounds_seen;
      }
    ll.count = S.count;
    hipError_t e = hipMemcpyAsync(E->ws.out, &E->pin->out, sizeof(LaneOut), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "engine step", __FILE__, __LINE__);
    ms_fill_noin<<<blocks, kMsBlock, 0, st>>>(E->ws.out, E->md->noin, g->n, ll);
    eng_seed<<<1, 64, 0, st>>>(A, E->ws.out, S, E->P);
    E->pending = 0;
  }
  // 3. one round for every lane in flight, and the copies back (captured)
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipGraphLaunch(E->round, st);
  if (e == hipSuccess) e = hipEventRecord(E->ev, st);
  if (e != hipSuccess) return hip_fail(e, "engine step", __FILE__, __LINE__);
  E->inflight = true;
  return CGBFS_OK;
}

extern "C" int cgbfs_engine_poll(cgbfs_engine *E, int32_t wait, int32_t *ready, uint64_t *completed,
                                 uint64_t *busy, int64_t *m_lanes, int32_t *depth_lanes) {
  if (!E || !ready) return set_error(CGBFS_EINVAL, "engine and ready must not be null");
  *ready = 0;
  if (completed) *completed = 0;
  if (busy) *busy = E->busy;
  if (!E->inflight) {
    *ready = 1;
    return CGBFS_OK;
  }
  hipError_t e = wait ? hipEventSynchronize(E->ev) : hipEventQuery(E->ev);
  if (e == hipErrorNotReady) return CGBFS_OK;
  if (e != hipSuccess) return hip_fail(e, "engine poll", __FILE__, __LINE__);
  E->inflight = false;
  *ready = 1;
  const MsCtrl &hc = E->pin->hc;
  if (hc.status) return set_error(hc.status, "engine round failed");
  // lanes retired by this step: outputs final, m counted
  const uint64_t done = E->retired;
  for (int l = 0; l < 64; l++)
    if ((done >> l) & 1ull) {
      if (m_lanes) m_lanes[l] = (int64_t)E->pin->m[l];
      if (depth_lanes) depth_lanes[l] = E->depth[l];
    }
  E->busy &= ~done;
  E->retired = 0;
  // lanes whose frontier emptied this round finish in the next step
  E->rounds_seen = hc.rounds;
  const uint64_t running = E->busy & ~E->pending & ~E->retiring;
  const uint64_t live = hc.done ? 0ull : (uint64_t)(hc.pmask | hc.qmask);
  const uint64_t ended = running & ~live;
  for (int l = 0; l < 64; l++)
    if ((ended >> l) & 1ull) E->depth[l] = (int32_t)(hc.rounds - E->base[l]);
  E->retiring |= ended;
  if (completed) *completed = done;
  if (busy) *busy = E->busy;
  return CGBFS_OK;
}

extern "C" int cgbfs_engine_destroy(cgbfs_engine *E) {
  if (!E) return CGBFS_OK;
  if (E->st) (void)hipStreamSynchronize(E->st);
  if (E->round) (void)hipGraphExecDestroy(E->round);
  if (E->ev) (void)hipEventDestroy(E->ev);
  if (E->st) (void)hipStreamDestroy(E->st);
  if (E->pin) (void)hipHostFree(E->pin);
  if (E->mem) (void)hipFree(E->mem);
  delete E;
  return CGBFS_OK;
}
