// Structural checks of a graph handed to cgbfs_graph_create.
//
// Every search kernel indexes per-vertex arrays with the column ids it reads
// and takes the first frontier hit of an ascending in-row as the lowest-id
// parent, so a graph is accepted only if
//   * offsets run from 0 to nnz and never decrease,
//   * every id lies in [0, n),
//   * every row is strictly ascending (no duplicate entries), and
//   * the two views describe one matrix: for a directed graph the in view is
//     the transpose of the out view; a symmetric graph's one view equals its
//     own transpose.
// The reference asserts none of this: preprocess_bool.py:29-229 trusts the
// scipy CSR/CSC it is given (and run_bfs.py:338-358 uploads it as is).
//
// Host arrays are checked on the host (threads over row ranges) before
// anything touches the GPU; device arrays are checked on the device, on the
// library's own copy, before any search can read them.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "cg_common.h"

namespace cg {

// Defect classes, in the order they are reported.
enum : uint32_t {
  kBadPtr = 1u,     // offsets do not run 0..nnz or decrease
  kBadId = 2u,      // an id outside [0, n)
  kBadOrder = 4u,   // a row not strictly ascending
  kBadTransp = 8u,  // the views disagree (an entry without its transpose)
};

static std::string defect_text(uint32_t bits, const char *view, long long where) {
  char buf[256];
  const char *what = (bits & kBadPtr)     ? "offsets must start at 0, end at nnz and never decrease"
                     : (bits & kBadId)    ? "a column id lies outside [0, n)"
                     : (bits & kBadOrder) ? "a row is not strictly ascending (unsorted or duplicate ids)"
                                          : "the in and out views are not transposes of each other";
  if (where >= 0)
    snprintf(buf, sizeof(buf), "invalid graph (%s view, at %s %lld): %s", view,
             (bits & kBadPtr) ? "vertex" : "entry", where, what);
  else
    snprintf(buf, sizeof(buf), "invalid graph (%s view): %s", view, what);
  return buf;
}

// ---------------------------------------------------------------------------
// Host check
// ---------------------------------------------------------------------------

// Runs fn(lo, hi) over [0, count) in up to 16 threads; returns the first
// (lowest-range) failure position reported by fn, or -1.
template <class Fn>
static long long par_rows(int64_t count, Fn fn) {
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(nt, 16u));
  if (count < (int64_t)1 << 16) nt = 1;
  std::vector<long long> bad(nt, -1);
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++) {
    const int64_t lo = count * t / nt, hi = count * (t + 1) / nt;
    if (nt == 1) bad[t] = fn(lo, hi);
    else th.emplace_back([&, t, lo, hi]() { bad[t] = fn(lo, hi); });
  }
  for (auto &x : th) x.join();
  for (long long b : bad)
    if (b >= 0) return b;
  return -1;
}

// One compressed view: offsets, ids in range, rows strictly ascending.
static uint32_t host_view(int64_t n, int64_t nnz, const int64_t *ptr, const int32_t *idx,
                          long long *where) {
  if (ptr[0] != 0 || ptr[n] != nnz) {
    *where = ptr[0] != 0 ? 0 : n;
    return kBadPtr;
  }
  long long b = par_rows(n, [&](int64_t lo, int64_t hi) -> long long {
    for (int64_t v = lo; v < hi; v++)
      if (ptr[v + 1] < ptr[v]) return v;
    return -1;
  });
  if (b >= 0) {
    *where = b;
    return kBadPtr;
  }
  uint32_t kind = 0;
  b = par_rows(n, [&](int64_t lo, int64_t hi) -> long long {
    for (int64_t v = lo; v < hi; v++) {
      for (int64_t e = ptr[v]; e < ptr[v + 1]; e++) {
        if (idx[e] < 0 || idx[e] >= n) return e;
        if (e > ptr[v] && idx[e] <= idx[e - 1]) return e;
      }
    }
    return -1;
  });
  if (b >= 0) {
    *where = b;
    kind = (idx[b] < 0 || idx[b] >= n) ? kBadId : kBadOrder;
  }
  return kind;
}

// Every entry (r, c) of view a (id c in row r) appears as id r in row c of
// view b (rows ascending, so a binary search).  With equal nnz and no
// duplicates this makes b exactly the transpose of a.
static long long host_transpose(int64_t n, const int64_t *ap, const int32_t *ai, const int64_t *bp,
                                const int32_t *bi) {
  return par_rows(n, [&](int64_t lo, int64_t hi) -> long long {
    for (int64_t r = lo; r < hi; r++)
      for (int64_t e = ap[r]; e < ap[r + 1]; e++) {
        const int32_t c = ai[e];
        if (!std::binary_search(bi + bp[c], bi + bp[c + 1], (int32_t)r)) return e;
      }
    return -1;
  });
}

int check_graph_host(int64_t n, int64_t nnz, const int64_t *out_ptr, const int32_t *out_idx,
                     const int64_t *in_ptr, const int32_t *in_idx, int symmetric) {
  long long where = -1;
  uint32_t bad = host_view(n, nnz, out_ptr, out_idx, &where);
  if (bad) return set_error(CGBFS_EINVAL, defect_text(bad, symmetric ? "shared" : "out", where));
  if (!symmetric) {
    bad = host_view(n, nnz, in_ptr, in_idx, &where);
    if (bad) return set_error(CGBFS_EINVAL, defect_text(bad, "in", where));
  }
  const int64_t *ip = symmetric ? out_ptr : in_ptr;
  const int32_t *ii = symmetric ? out_idx : in_idx;
  where = host_transpose(n, ip, ii, out_ptr, out_idx);
  if (where >= 0)
    return set_error(CGBFS_EINVAL, defect_text(kBadTransp, symmetric ? "shared" : "in", where));
  return CGBFS_OK;
}

// ---------------------------------------------------------------------------
// Device check (arrays already in library-owned HBM)
// ---------------------------------------------------------------------------
struct CheckOut {
  unsigned bits;
  unsigned pad;
  unsigned long long first;  // lowest failing vertex (offsets) or entry
};

__device__ __forceinline__ void report(CheckOut *o, uint32_t bit, unsigned long long at) {
  atomicOr(&o->bits, bit);
  atomicMin(&o->first, at);
}

// Offsets: ptr[0] == 0, ptr[n] == nnz, ptr[v] <= ptr[v+1].
__global__ void __launch_bounds__(256) chk_ptr_kernel(const int64_t *__restrict__ ptr, int64_t n,
                                                      int64_t nnz, CheckOut *o) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > n) return;
  const int64_t p = ptr[v];
  bool bad = (v == 0 && p != 0) || (v == n && p != nnz) || (v < n && ptr[v + 1] < p);
  if (bad) report(o, kBadPtr, (unsigned long long)v);
}

// Ids and row order, one wavefront per row (rows reach millions of entries).
// Runs only once the offsets passed, so every read stays inside [0, nnz).
__global__ void __launch_bounds__(256) chk_rows_kernel(const int64_t *__restrict__ ptr,
                                                       const int32_t *__restrict__ idx, int64_t n,
                                                       CheckOut *o) {
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  for (int64_t r = wave; r < n; r += nwaves) {
    const int64_t b = ptr[r], e = ptr[r + 1];
    for (int64_t k = b + lane; k < e; k += kWave) {
      const int32_t c = idx[k];
      if (c < 0 || c >= n) report(o, kBadId, (unsigned long long)k);
      else if (k > b && idx[k - 1] >= c) report(o, kBadOrder, (unsigned long long)k);
    }
  }
}

// Transposition, one thread per entry (r, c) of view a: r must be an id of
// row c of view b.  Runs only once both views passed the row check.
__global__ void __launch_bounds__(256) chk_transp_kernel(const int64_t *__restrict__ ap,
                                                         const int32_t *__restrict__ ai,
                                                         const int64_t *__restrict__ bp,
                                                         const int32_t *__restrict__ bi, int64_t n,
                                                         CheckOut *o) {
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / kWave;
  for (int64_t r = wave; r < n; r += nwaves) {
    const int64_t b = ap[r], e = ap[r + 1];
    for (int64_t k = b + lane; k < e; k += kWave) {
      const int32_t c = ai[k];
      int64_t lo = bp[c], hi = bp[c + 1];
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (bi[mid] < (int32_t)r) lo = mid + 1;
        else hi = mid;
      }
      if (lo == bp[c + 1] || bi[lo] != (int32_t)r) report(o, kBadTransp, (unsigned long long)k);
    }
  }
}

static int check_result(CheckOut *d, hipStream_t st, const char *view) {
  CheckOut h;
  CG_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, st));
  CG_HIP(hipStreamSynchronize(st));
  if (!h.bits) return CGBFS_OK;
  return set_error(CGBFS_EINVAL, defect_text(h.bits, view, (long long)h.first));
}

static int reset_out(CheckOut *d, hipStream_t st) {
  CheckOut h{0u, 0u, ~0ull};
  CG_HIP(hipMemcpyAsync(d, &h, sizeof(h), hipMemcpyHostToDevice, st));
  return CGBFS_OK;
}

int check_graph_device(const DeviceGraph *g) {
  hipStream_t st = g->stream;
  CheckOut *d = nullptr;
  CG_HIP(hipMalloc(&d, sizeof(CheckOut)));
  const bool shared = g->in_ptr == g->out_ptr;
  const unsigned pblocks = (unsigned)((g->n + 1 + 255) / 256);
  const unsigned rblocks = (unsigned)std::min<int64_t>((g->n + 3) / 4, 4096);
  int rc = CGBFS_OK;
  struct View { const int64_t *p; const int32_t *i; const char *name; };
  const View views[2] = {{g->out_ptr, g->out_idx, shared ? "shared" : "out"},
                         {g->in_ptr, g->in_idx, "in"}};
  for (int k = 0; k < (shared ? 1 : 2) && !rc; k++) {
    rc = reset_out(d, st);
    if (!rc) {
      chk_ptr_kernel<<<pblocks, 256, 0, st>>>(views[k].p, g->n, g->nnz, d);
      rc = check_result(d, st, views[k].name);
    }
    if (!rc && g->nnz > 0) {
      chk_rows_kernel<<<rblocks, 256, 0, st>>>(views[k].p, views[k].i, g->n, d);
      rc = check_result(d, st, views[k].name);
    }
  }
  if (!rc && g->nnz > 0) {
    rc = reset_out(d, st);
    if (!rc) {
      chk_transp_kernel<<<rblocks, 256, 0, st>>>(g->in_ptr, g->in_idx, g->out_ptr, g->out_idx, g->n, d);
      rc = check_result(d, st, shared ? "shared" : "in");
    }
  }
  hipError_t e = hipGetLastError();
  hipFree(d);
  if (!rc && e != hipSuccess) rc = hip_fail(e, "graph check", __FILE__, __LINE__);
  return rc;
}

}  // namespace cg

extern "C" int cgbfs_graph_check(int64_t n, int64_t nnz, const int64_t *out_ptr,
                                 const int32_t *out_idx, const int64_t *in_ptr,
                                 const int32_t *in_idx, int32_t symmetric) {
  using namespace cg;
  if (n <= 0 || n > INT32_MAX || nnz < 0)
    return set_error(CGBFS_EINVAL, "n must be in [1, 2^31-1], nnz >= 0");
  if (!out_ptr || (nnz > 0 && !out_idx)) return set_error(CGBFS_EINVAL, "out_ptr/out_idx null");
  if (!symmetric && (!in_ptr || (nnz > 0 && !in_idx)))
    return set_error(CGBFS_EINVAL, "directed graph needs in_ptr/in_idx");
  return check_graph_host(n, nnz, out_ptr, out_idx, in_ptr, in_idx, symmetric ? 1 : 0);
}
