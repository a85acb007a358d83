// Graph construction ("Kernel 1") for the gfx950 BFS library: host-array
// upload, the on-GPU R-MAT generator, and per-graph derived arrays.
//
// The R-MAT generator reproduces datasets/gen_rmat.py:15-44 edge for edge:
// numpy.random.default_rng(seed) is PCG64 (128-bit LCG, XSL-RR output), and
// rng.random(m) turns each 64-bit output x into (x >> 11) * 2**-53.  Level L
// of the reference loop draws outputs L*m .. L*m+m-1 of that stream, so a
// thread owning edges [i0, i0+CH) jumps its private copy of the LCG to
// position L*m + i0 with one affine map per level (the maps for L*m are
// computed once on the host) and then steps CH times.  The quadrant tests
// compare the same float64 values against the same float64 thresholds, so
// the integer edge endpoints are bit-identical to numpy's.
//
// Symmetrise / drop self-loops / dedup (gen_rmat.py:34-44) becomes: emit
// packed keys (row << scale | col) for both orientations, radix sort,
// unique, and build int64 row offsets from the sorted keys.
#include <hipcub/hipcub.hpp>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "cg_common.h"

namespace cg {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t err, const char *what, const char *file, int line) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(err), file, line);
  return set_error(err == hipErrorOutOfMemory ? CGBFS_ENOMEM : CGBFS_EHIP, buf);
}

// ---------------------------------------------------------------------------
// 128-bit LCG arithmetic (PCG64's state transition).
// ---------------------------------------------------------------------------
struct U128 {
  uint64_t hi, lo;
};

__host__ __device__ inline uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

__host__ __device__ inline U128 mul128(U128 a, U128 b) {
  U128 r;
  r.lo = a.lo * b.lo;
  r.hi = mulhi64(a.lo, b.lo) + a.lo * b.hi + a.hi * b.lo;
  return r;
}

__host__ __device__ inline U128 add128(U128 a, U128 b) {
  U128 r;
  r.lo = a.lo + b.lo;
  r.hi = a.hi + b.hi + (r.lo < a.lo ? 1u : 0u);
  return r;
}

// PCG's default 128-bit multiplier (PCG_DEFAULT_MULTIPLIER_128).
__host__ __device__ inline U128 pcg_mult() {
  return U128{0x2360ED051FC65DA4ull, 0x4385DF649FCCF645ull};
}

// Affine map state -> A*state + C equal to `delta` LCG steps (Brown's
// arbitrary-stride jump: square-and-multiply on the map itself).
__host__ __device__ inline void lcg_jump(uint64_t delta, U128 inc, U128 *A, U128 *C) {
  U128 acc_mult{0, 1}, acc_plus{0, 0};
  U128 cur_mult = pcg_mult(), cur_plus = inc;
  while (delta > 0) {
    if (delta & 1) {
      acc_mult = mul128(acc_mult, cur_mult);
      acc_plus = add128(mul128(acc_plus, cur_mult), cur_plus);
    }
    cur_plus = mul128(add128(cur_mult, U128{0, 1}), cur_plus);
    cur_mult = mul128(cur_mult, cur_mult);
    delta >>= 1;
  }
  *A = acc_mult;
  *C = acc_plus;
}

__device__ inline double pcg_next_double(U128 &s, U128 inc) {
  s = add128(mul128(s, pcg_mult()), inc);
  uint64_t x = s.hi ^ s.lo;
  unsigned rot = (unsigned)(s.hi >> 58);
  uint64_t out = (x >> rot) | (x << ((64u - rot) & 63u));
  return (double)(out >> 11) * (1.0 / 9007199254740992.0);
}

constexpr int kRmatChunk = 8;
constexpr int kMaxScale = 31;

struct RmatParams {
  U128 s0, inc;
  U128 levA[kMaxScale], levC[kMaxScale];  // jump maps for L*m steps
  double a, ab, abc;
  int64_t m;
  int scale;
  int symmetric;
};

// One thread = kRmatChunk consecutive sampled edges.  Writes packed keys:
//   symmetric: keys[2i] = (u<<S)|v, keys[2i+1] = (v<<S)|u   (M[u][v], M[v][u])
//   directed : keys_in[i] = (v<<S)|u (row v of M), keys_out[i] = (u<<S)|v
// Self-loops become the sentinel key of (n-1, n-1), itself a self-loop, so
// dropping the sentinel after the sort removes exactly the self-loops.
__global__ void __launch_bounds__(256) rmat_edges_kernel(const RmatParams *__restrict__ P,
                                                         uint64_t *__restrict__ keys_a,
                                                         uint64_t *__restrict__ keys_b) {
  const int64_t m = P->m;
  const int S = P->scale;
  int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kRmatChunk;
  if (i0 >= m) return;
  int cnt = (int)((m - i0) < kRmatChunk ? (m - i0) : kRmatChunk);
  U128 inc = P->inc;
  U128 JA, JC;
  lcg_jump((uint64_t)i0, inc, &JA, &JC);
  U128 st = add128(mul128(JA, P->s0), JC);  // stream position i0
  uint32_t u[kRmatChunk], v[kRmatChunk];
#pragma unroll
  for (int k = 0; k < kRmatChunk; k++) u[k] = v[k] = 0;
  const double a = P->a, ab = P->ab, abc = P->abc;
  for (int L = 0; L < S; L++) {
    U128 s = add128(mul128(P->levA[L], st), P->levC[L]);  // position L*m + i0
    uint32_t bit = 1u << (S - 1 - L);
#pragma unroll
    for (int k = 0; k < kRmatChunk; k++) {
      if (k < cnt) {
        double r = pcg_next_double(s, inc);
        // (0,1): v gets the bit; (1,0): u; (1,1): both (gen_rmat.py:26-33)
        bool q01 = (r >= a) && (r < ab);
        bool q10 = (r >= ab) && (r < abc);
        bool q11 = (r >= abc);
        if (q10 || q11) u[k] |= bit;
        if (q01 || q11) v[k] |= bit;
      }
    }
  }
  const uint64_t sent = (((uint64_t)1 << S) - 1) << S | (((uint64_t)1 << S) - 1);
#pragma unroll
  for (int k = 0; k < kRmatChunk; k++) {
    if (k < cnt) {
      int64_t i = i0 + k;
      bool loop = (u[k] == v[k]);
      uint64_t uv = ((uint64_t)u[k] << S) | v[k];
      uint64_t vu = ((uint64_t)v[k] << S) | u[k];
      if (P->symmetric) {
        keys_a[2 * i] = loop ? sent : uv;
        keys_a[2 * i + 1] = loop ? sent : vu;
      } else {
        keys_a[i] = loop ? sent : vu;  // in view: row v (dst), col u (src)
        keys_b[i] = loop ? sent : uv;  // out view: column u's rows v
      }
    }
  }
}

// Row offsets from sorted, unique packed keys: ptr[r] = first index with
// row >= r, for r in [0, n].
__global__ void keys_to_csr_kernel(const uint64_t *__restrict__ keys, int64_t nk, int S,
                                   int64_t n, int64_t *__restrict__ ptr,
                                   int32_t *__restrict__ idx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nk) return;
  const uint64_t cmask = ((uint64_t)1 << S) - 1;
  int64_t r = (i < nk) ? (int64_t)(keys[i] >> S) : n;
  int64_t rp = (i == 0) ? -1 : (int64_t)(keys[i - 1] >> S);
  for (int64_t rr = rp + 1; rr <= r; rr++) ptr[rr] = i;
  if (i < nk) idx[i] = (int32_t)(keys[i] & cmask);
}

// Per-vertex GTEPS weight (bfs_timing.py:390-395) and nothing else.
__global__ void m_weight_kernel(const int64_t *__restrict__ in_ptr,
                                const int32_t *__restrict__ in_idx,
                                const int64_t *__restrict__ out_ptr, int64_t n, int symmetric,
                                int32_t *__restrict__ w) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  if (!symmetric) {
    w[v] = (int32_t)(out_ptr[v + 1] - out_ptr[v]);
    return;
  }
  // #{c <= v} in the ascending row v: upper_bound by binary search.
  int64_t lo = in_ptr[v], hi = in_ptr[v + 1];
  int64_t b = lo;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (in_idx[mid] <= v) lo = mid + 1;
    else hi = mid;
  }
  w[v] = (int32_t)(lo - b);
}

// The first kRowHead entries of every in-row (-1 past the row end).
__global__ void row_head_kernel(const int64_t *__restrict__ in_ptr,
                                const int32_t *__restrict__ in_idx, int64_t n,
                                RowHead *__restrict__ head) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int64_t b = in_ptr[v], e = in_ptr[v + 1];
  RowHead h;
#pragma unroll
  for (int k = 0; k < kRowHead; k++) h.id[k] = b + k < e ? in_idx[b + k] : -1;
  head[v] = h;
}

// The bitmap of vertices without in-edges, one 64-bit word per wavefront.
__global__ void noin_kernel(const int64_t *__restrict__ in_ptr, int64_t n,
                            uint64_t *__restrict__ noin) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t m = __ballot(v < n && in_ptr[v + 1] == in_ptr[v]);
  if (lane_id() == 0 && v < n) noin[v >> 6] = m;
}

// Counts the vertices with at least one in-edge (one atomic per wavefront).
__global__ void count_in_kernel(const int64_t *__restrict__ in_ptr, int64_t n,
                                unsigned long long *__restrict__ count) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool has = v < n && in_ptr[v + 1] > in_ptr[v];
  const unsigned long long c = (unsigned long long)__popcll(__ballot(has));
  if (lane_id() == 0 && c) atomicAdd(count, c);
}

static int sort_unique_to_csr(uint64_t *keys, uint64_t *scratch, int64_t count, int S, int64_t n,
                              hipStream_t stream, int64_t **ptr_out, int32_t **idx_out,
                              int64_t *nk_out) {
  size_t tmp_sort = 0, tmp_uniq = 0;
  int end_bit = 2 * S;
  CG_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_sort, keys, scratch, count, 0, end_bit,
                                           stream));
  int64_t *d_nsel = nullptr;
  CG_HIP(hipcub::DeviceSelect::Unique(nullptr, tmp_uniq, scratch, keys, d_nsel, count, stream));
  size_t tmp = tmp_sort > tmp_uniq ? tmp_sort : tmp_uniq;
  void *d_tmp = nullptr;
  CG_HIP(hipMalloc(&d_tmp, tmp + 256));
  d_nsel = (int64_t *)((char *)d_tmp + tmp);
  hipError_t e = hipcub::DeviceRadixSort::SortKeys(d_tmp, tmp_sort, keys, scratch, count, 0,
                                                   end_bit, stream);
  if (e == hipSuccess)
    e = hipcub::DeviceSelect::Unique(d_tmp, tmp_uniq, scratch, keys, d_nsel, count, stream);
  int64_t nsel = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&nsel, d_nsel, sizeof(int64_t), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  uint64_t last = 0;
  if (e == hipSuccess && nsel > 0)
    e = hipMemcpy(&last, keys + nsel - 1, sizeof(uint64_t), hipMemcpyDeviceToHost);
  hipFree(d_tmp);
  if (e != hipSuccess) return hip_fail(e, "sort/unique", __FILE__, __LINE__);
  const uint64_t sent = ((((uint64_t)1 << S) - 1) << S) | (((uint64_t)1 << S) - 1);
  int64_t nk = (nsel > 0 && last == sent) ? nsel - 1 : nsel;
  int64_t *ptr = nullptr;
  int32_t *idx = nullptr;
  CG_HIP(hipMalloc(&ptr, sizeof(int64_t) * (n + 1)));
  e = hipMalloc(&idx, sizeof(int32_t) * (nk > 0 ? nk : 1));
  if (e != hipSuccess) {
    hipFree(ptr);
    return hip_fail(e, "hipMalloc(idx)", __FILE__, __LINE__);
  }
  int64_t threads = nk + 1;
  keys_to_csr_kernel<<<(unsigned)((threads + 255) / 256), 256, 0, stream>>>(keys, nk, S, n, ptr, idx);
  CG_HIP(hipGetLastError());
  *ptr_out = ptr;
  *idx_out = idx;
  *nk_out = nk;
  return CGBFS_OK;
}

int graph_finalize(DeviceGraph *g) {
  if (!g->stream) CG_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  CG_HIP(hipMalloc(&g->m_weight, sizeof(int32_t) * (g->n > 0 ? g->n : 1)));
  CG_HIP(hipMalloc(&g->in_head, sizeof(RowHead) * (g->n > 0 ? g->n : 1)));
  CG_HIP(hipMalloc(&g->noin, sizeof(uint64_t) * (size_t)((g->n + 63) / 64 + 1)));
  if (g->n > 0) {
    m_weight_kernel<<<(unsigned)((g->n + 255) / 256), 256, 0, g->stream>>>(
        g->in_ptr, g->in_idx, g->out_ptr, g->n, g->symmetric, g->m_weight);
    CG_HIP(hipGetLastError());
    row_head_kernel<<<(unsigned)((g->n + 255) / 256), 256, 0, g->stream>>>(g->in_ptr, g->in_idx,
                                                                          g->n, g->in_head);
    CG_HIP(hipGetLastError());
    noin_kernel<<<(unsigned)((g->n + 255) / 256), 256, 0, g->stream>>>(g->in_ptr, g->n, g->noin);
    CG_HIP(hipGetLastError());
  }
  CG_HIP(hipHostMalloc((void **)&g->host_flag, 64, hipHostMallocDefault));
  int64_t last = 0;
  unsigned long long *cnt = nullptr, n_in = 0;
  CG_HIP(hipMalloc(&cnt, sizeof(unsigned long long)));
  hipError_t e = hipMemsetAsync(cnt, 0, sizeof(unsigned long long), g->stream);
  if (e == hipSuccess && g->n > 0) {
    count_in_kernel<<<(unsigned)((g->n + 255) / 256), 256, 0, g->stream>>>(g->in_ptr, g->n, cnt);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&n_in, cnt, sizeof(n_in), hipMemcpyDeviceToHost, g->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&last, g->out_ptr + g->n, sizeof(int64_t), hipMemcpyDeviceToHost,
                                          g->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  hipFree(cnt);
  CG_HIP(e);
  g->total_out_edges = last;
  g->n_in = (int64_t)n_in;
  return CGBFS_OK;
}


static void graph_free(DeviceGraph *g) {
  if (!g) return;
  hipSetDevice(g->device);
  if (g->in_ptr && g->in_ptr != g->out_ptr) hipFree(g->in_ptr);
  if (g->in_idx && g->in_idx != g->out_idx) hipFree(g->in_idx);
  if (g->out_ptr) hipFree(g->out_ptr);
  if (g->out_idx) hipFree(g->out_idx);
  if (g->m_weight) hipFree(g->m_weight);
  if (g->in_head) hipFree(g->in_head);
  if (g->noin) hipFree(g->noin);
  if (g->bfs_exec) hipGraphExecDestroy(g->bfs_exec);
  if (g->ws1) hipFree(g->ws1);
  if (g->wsm) hipFree(g->wsm);
  if (g->msdata && g->msdata_free) g->msdata_free(g->msdata);
  if (g->host_flag) hipHostFree(g->host_flag);
  if (g->stream) hipStreamDestroy(g->stream);
  // DeviceGraph is the first and only member of the handle.
  delete reinterpret_cast<cgbfs_graph *>(g);
}

// 1: device memory of `device`; -1: device memory of another device; 0: host
// memory (pageable, pinned or managed).  A pointer the runtime does not know
// is host memory.
static int pointer_kind(const void *p, int device) {
  hipPointerAttribute_t a;
  std::memset(&a, 0, sizeof(a));
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  if (a.type != hipMemoryTypeDevice || a.isManaged) return 0;
  return a.device == device ? 1 : -1;
}

static int check_device(int32_t device) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0) return set_error(CGBFS_ENODEV, "no HIP device available");
  if (device < 0 || device >= count) return set_error(CGBFS_EINVAL, "device index out of range");
  CG_HIP(hipSetDevice(device));
  return CGBFS_OK;
}

}  // namespace cg

using namespace cg;

extern "C" {

const char *cgbfs_last_error(void) { return g_last_error.c_str(); }

int cgbfs_version(void) { return 1; }

int cgbfs_device_count(int32_t *count) {
  if (!count) return set_error(CGBFS_EINVAL, "count is null");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *count = (e == hipSuccess) ? c : 0;
  return CGBFS_OK;
}

int cgbfs_graph_create(int32_t device, int64_t n, int64_t nnz, const int64_t *out_ptr,
                       const int32_t *out_idx, const int64_t *in_ptr, const int32_t *in_idx,
                       int32_t symmetric, cgbfs_graph **out) {
  if (!out) return set_error(CGBFS_EINVAL, "out is null");
  *out = nullptr;
  if (n <= 0 || n > INT32_MAX || nnz < 0) return set_error(CGBFS_EINVAL, "n must be in [1, 2^31-1], nnz >= 0");
  if (!out_ptr || (nnz > 0 && !out_idx)) return set_error(CGBFS_EINVAL, "out_ptr/out_idx null");
  if (!symmetric && (!in_ptr || (nnz > 0 && !in_idx)))
    return set_error(CGBFS_EINVAL, "directed graph needs in_ptr/in_idx");
  // Host or device arrays: all of one kind.
  const void *arrs[4] = {out_ptr, nnz > 0 ? out_idx : nullptr, symmetric ? nullptr : in_ptr,
                         symmetric || nnz == 0 ? nullptr : in_idx};
  int kinds = 0;  // bit 0: host array seen, bit 1: device array of `device`
  for (const void *a : arrs) {
    if (!a) continue;
    const int k = pointer_kind(a, device);
    if (k < 0) return set_error(CGBFS_EINVAL, "graph arrays live on another device");
    kinds |= k ? 2 : 1;
  }
  if (kinds == 3) return set_error(CGBFS_EINVAL, "graph arrays mix host and device memory");
  const bool on_dev = kinds == 2;
  int rc = CGBFS_OK;
  if (!on_dev) {
    // host arrays: checked before anything touches the GPU
    rc = check_graph_host(n, nnz, out_ptr, out_idx, in_ptr, in_idx, symmetric ? 1 : 0);
    if (rc) return rc;
  }
  rc = check_device(device);
  if (rc) return rc;
  cgbfs_graph *h = new cgbfs_graph();
  DeviceGraph *g = &h->g;
  g->device = device;
  g->n = n;
  g->nnz = nnz;
  g->symmetric = symmetric ? 1 : 0;
  const hipMemcpyKind kind = on_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  size_t pbytes = sizeof(int64_t) * (size_t)(n + 1);
  size_t ibytes = sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1);
  hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&g->out_ptr, pbytes);
  if (e == hipSuccess) e = hipMalloc(&g->out_idx, ibytes);
  if (e == hipSuccess) e = hipMemcpy(g->out_ptr, out_ptr, pbytes, kind);
  if (e == hipSuccess && nnz > 0) e = hipMemcpy(g->out_idx, out_idx, sizeof(int32_t) * nnz, kind);
  if (e == hipSuccess) {
    if (symmetric) {
      g->in_ptr = g->out_ptr;
      g->in_idx = g->out_idx;
    } else {
      e = hipMalloc(&g->in_ptr, pbytes);
      if (e == hipSuccess) e = hipMalloc(&g->in_idx, ibytes);
      if (e == hipSuccess) e = hipMemcpy(g->in_ptr, in_ptr, pbytes, kind);
      if (e == hipSuccess && nnz > 0) e = hipMemcpy(g->in_idx, in_idx, sizeof(int32_t) * nnz, kind);
    }
  }
  if (e != hipSuccess) {
    rc = hip_fail(e, "graph upload", __FILE__, __LINE__);
    graph_free(g);
    return rc;
  }
  // device arrays: checked on the library's own copy, before any kernel
  // that trusts the structure (m_weight below included) reads it
  if (on_dev) rc = check_graph_device(g);
  if (!rc) rc = graph_finalize(g);
  if (rc) {
    graph_free(g);
    return rc;
  }
  *out = h;
  return CGBFS_OK;
}

int cgbfs_graph_create_rmat(int32_t device, int32_t scale, int64_t edgefactor, const uint64_t *pcg,
                            double a, double ab, double abc, int32_t symmetric, cgbfs_graph **out) {
  if (!out || !pcg) return set_error(CGBFS_EINVAL, "null argument");
  *out = nullptr;
  if (scale < 1 || scale > 30) return set_error(CGBFS_EINVAL, "scale must be in [1, 30]");
  if (edgefactor < 1) return set_error(CGBFS_EINVAL, "edgefactor must be >= 1");
  int rc = check_device(device);
  if (rc) return rc;
  const int64_t n = (int64_t)1 << scale;
  const int64_t m = n * edgefactor;
  const int64_t nkeys = symmetric ? 2 * m : m;

  RmatParams P;
  memset(&P, 0, sizeof(P));
  P.s0 = U128{pcg[0], pcg[1]};
  P.inc = U128{pcg[2], pcg[3]};
  for (int L = 0; L < scale; L++) lcg_jump((uint64_t)(L * m), P.inc, &P.levA[L], &P.levC[L]);
  P.a = a;
  P.ab = ab;
  P.abc = abc;
  P.m = m;
  P.scale = scale;
  P.symmetric = symmetric ? 1 : 0;

  cgbfs_graph *h = new cgbfs_graph();
  DeviceGraph *g = &h->g;
  g->device = device;
  g->n = n;
  g->symmetric = symmetric ? 1 : 0;
  hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return hip_fail(e, "hipStreamCreate", __FILE__, __LINE__);
  }
  RmatParams *dP = nullptr;
  uint64_t *ka = nullptr, *kb = nullptr;
  e = hipMalloc(&dP, sizeof(RmatParams));
  if (e == hipSuccess) e = hipMalloc(&ka, sizeof(uint64_t) * nkeys);
  if (e == hipSuccess) e = hipMalloc(&kb, sizeof(uint64_t) * nkeys);
  if (e == hipSuccess) e = hipMemcpyAsync(dP, &P, sizeof(P), hipMemcpyHostToDevice, g->stream);
  if (e == hipSuccess) {
    int64_t threads = (m + kRmatChunk - 1) / kRmatChunk;
    rmat_edges_kernel<<<(unsigned)((threads + 255) / 256), 256, 0, g->stream>>>(dP, ka, kb);
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    rc = hip_fail(e, "rmat generation", __FILE__, __LINE__);
    hipFree(dP);
    hipFree(ka);
    hipFree(kb);
    graph_free(g);
    return rc;
  }
  int64_t nk = 0;
  if (symmetric) {
    rc = sort_unique_to_csr(ka, kb, nkeys, scale, n, g->stream, &g->out_ptr, &g->out_idx, &nk);
    g->in_ptr = g->out_ptr;
    g->in_idx = g->out_idx;
  } else {
    // keys_a = in view (row v); keys_b = out view (column u).  Sort each
    // with the other buffer's half as scratch, reallocating as needed.
    uint64_t *scratch = nullptr;
    e = hipMalloc(&scratch, sizeof(uint64_t) * nkeys);
    if (e != hipSuccess) rc = hip_fail(e, "hipMalloc(scratch)", __FILE__, __LINE__);
    if (!rc) rc = sort_unique_to_csr(ka, scratch, nkeys, scale, n, g->stream, &g->in_ptr, &g->in_idx, &nk);
    int64_t nk2 = 0;
    if (!rc) rc = sort_unique_to_csr(kb, scratch, nkeys, scale, n, g->stream, &g->out_ptr, &g->out_idx, &nk2);
    if (!rc && nk2 != nk) rc = set_error(CGBFS_EHIP, "directed views disagree on nnz");
    hipFree(scratch);
  }
  hipStreamSynchronize(g->stream);
  hipFree(dP);
  hipFree(ka);
  hipFree(kb);
  if (rc) {
    graph_free(g);
    return rc;
  }
  g->nnz = nk;
  rc = graph_finalize(g);
  if (rc) {
    graph_free(g);
    return rc;
  }
  *out = h;
  return CGBFS_OK;
}

int cgbfs_graph_info(const cgbfs_graph *h, int64_t *n, int64_t *nnz, int32_t *symmetric) {
  if (!h) return set_error(CGBFS_EINVAL, "graph is null");
  if (n) *n = h->g.n;
  if (nnz) *nnz = h->g.nnz;
  if (symmetric) *symmetric = h->g.symmetric;
  return CGBFS_OK;
}

int cgbfs_graph_download(const cgbfs_graph *h, int64_t *out_ptr, int32_t *out_idx, int64_t *in_ptr,
                         int32_t *in_idx) {
  if (!h) return set_error(CGBFS_EINVAL, "graph is null");
  const DeviceGraph *g = &h->g;
  CG_HIP(hipSetDevice(g->device));
  size_t pb = sizeof(int64_t) * (g->n + 1), ib = sizeof(int32_t) * g->nnz;
  if (out_ptr) CG_HIP(hipMemcpy(out_ptr, g->out_ptr, pb, hipMemcpyDeviceToHost));
  if (out_idx && ib) CG_HIP(hipMemcpy(out_idx, g->out_idx, ib, hipMemcpyDeviceToHost));
  if (in_ptr) CG_HIP(hipMemcpy(in_ptr, g->in_ptr, pb, hipMemcpyDeviceToHost));
  if (in_idx && ib) CG_HIP(hipMemcpy(in_idx, g->in_idx, ib, hipMemcpyDeviceToHost));
  return CGBFS_OK;
}

int cgbfs_graph_out_degree(const cgbfs_graph *h, int64_t *deg) {
  if (!h || !deg) return set_error(CGBFS_EINVAL, "null argument");
  const DeviceGraph *g = &h->g;
  CG_HIP(hipSetDevice(g->device));
  std::vector<int64_t> ptr(g->n + 1);
  CG_HIP(hipMemcpy(ptr.data(), g->out_ptr, sizeof(int64_t) * (g->n + 1), hipMemcpyDeviceToHost));
  for (int64_t v = 0; v < g->n; v++) deg[v] = ptr[v + 1] - ptr[v];
  return CGBFS_OK;
}

int cgbfs_graph_destroy(cgbfs_graph *h) {
  if (!h) return CGBFS_OK;
  graph_free(&h->g);
  return CGBFS_OK;
}

int cgbfs_sync(cgbfs_graph *h) {
  if (!h) return set_error(CGBFS_EINVAL, "graph is null");
  CG_HIP(hipSetDevice(h->g.device));
  CG_HIP(hipStreamSynchronize(h->g.stream));
  return CGBFS_OK;
}

}  // extern "C"
