/* Vertex-identity-preserving load balancer for a P x P PE grid (host code).
 *
 * Computes the permutation of util/analyze.cpp's iterative_load_balance
 * (:226-372) as its main (:374-728) drives it under --symmetric or
 * --shared-perm, where one permutation q relabels rows and columns alike:
 *
 *   1. Rows are visited in order of decreasing nonzero count.  Rows with
 *      equal counts stay in the order libstdc++'s std::sort leaves them in
 *      when it is handed (row, count) records in row order and compares
 *      counts only, exactly as analyze.cpp:238-245 does; introsort's result
 *      depends on nothing but those comparison outcomes, so ties break alike.
 *   2. Each row goes to the block row whose heaviest PE is currently the
 *      lightest, the lowest such block row on a tie, among block rows that
 *      still have room; the first row goes to block row 0 (:280-334).  Its
 *      nonzeros count towards the block column that the permutation of the
 *      previous pass gives their columns (the identity before the first).
 *   3. The placement of a pass becomes the shared permutation (:355).  A
 *      pass is kept while the ratio nnz / (P*P * heaviest PE) improves; the
 *      first pass is compared against 0 and so always kept unless the matrix
 *      is empty (:259-368).
 *
 * The block size is single-precision ceil(n / P) as in :545-554, so padded
 * vertex counts agree with the reference wherever that value places every
 * row; above 2^24 vertices where it would not, the exact ceiling is used and
 * the report says so.
 *
 * Layout and threading.  The matrix is held once as CSR: int64 offsets,
 * int32 columns.  Because q is shared, one uint16 array gives every vertex
 * its block index under q and serves rows and columns alike.  A pass takes
 * the rows in placement order a chunk at a time: all threads gather the
 * chunk's column blocks into a contiguous buffer, then one thread places the
 * rows, reading that buffer front to back.  Per-PE counts live in a
 * grid x grid array indexed [block row][block column], so one row's updates
 * stay in one stretch of it; the block row to fill next comes from a linear
 * scan for small grids and from a tournament tree for large ones.
 */
#include "cgbalance.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string t_err;

int fail(int code, const std::string &msg) {
  t_err = msg;
  return code;
}

constexpr int32_t kMaxGrid = 4096;         // grid*grid counters stay within 128 MiB
constexpr int32_t kLinearScanMax = 64;     // block-row choice by scan up to this grid
constexpr int64_t kChunkEntries = 1 << 22; // column blocks gathered per placement chunk

int pick_threads(int32_t threads) {
  if (threads > 0) return std::min<int32_t>(threads, 64);
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 1;
  return (int)std::min(h, 16u);
}

/* Runs fn(part, begin, end) on `parts` threads over [0, n), cut where
 * weight(i) (a non-decreasing prefix, weight(0) = 0) crosses equal shares. */
template <class W, class F>
void fork_join(int parts, int64_t n, W &&weight, F &&fn) {
  if (n <= 0) return;
  int64_t total = weight(n);
  if (parts <= 1 || n < 2 || total < (1 << 15)) {
    fn(0, (int64_t)0, n);
    return;
  }
  std::vector<int64_t> cut(parts + 1, n);
  cut[0] = 0;
  for (int p = 1; p < parts; ++p) {
    int64_t goal = total / parts * p, lo = cut[p - 1], hi = n;
    while (lo < hi) {  // first i with weight(i) >= goal
      int64_t mid = lo + (hi - lo) / 2;
      if (weight(mid) < goal) lo = mid + 1; else hi = mid;
    }
    cut[p] = lo;
  }
  std::vector<std::thread> pool;
  pool.reserve(parts - 1);
  for (int p = 1; p < parts; ++p)
    if (cut[p] < cut[p + 1]) pool.emplace_back([&fn, &cut, p] { fn(p, cut[p], cut[p + 1]); });
  if (cut[0] < cut[1]) fn(0, cut[0], cut[1]);
  for (auto &t : pool) t.join();
}

template <class F>
void fork_join_even(int parts, int64_t n, F &&fn) {
  fork_join(parts, n, [](int64_t i) { return i; }, fn);
}

/* Runs fn(t) for t in [0, parts), one thread each. */
template <class F>
void each_part(int parts, F &&fn) {
  std::vector<std::thread> pool;
  pool.reserve(parts);
  for (int t = 1; t < parts; ++t) pool.emplace_back([&fn, t] { fn(t); });
  fn(0);
  for (auto &th : pool) th.join();
}

struct Csr {
  int64_t n = 0;
  const int64_t *ptr = nullptr;
  const int32_t *idx = nullptr;
  std::vector<int64_t> own_ptr;  // storage when the matrix was built or rewritten here
  std::vector<int32_t> own_idx;
  int64_t nnz() const { return ptr[n]; }
  void adopt() {
    ptr = own_ptr.data();
    idx = own_idx.data();
  }
};

bool columns_in_range(const int32_t *idx, int64_t nnz, int64_t n, int threads) {
  std::atomic<bool> ok{true};
  fork_join_even(threads, nnz, [&](int, int64_t b, int64_t e) {
    bool good = true;
    for (int64_t i = b; i < e; ++i) good &= (idx[i] >= 0) & (idx[i] < n);
    if (!good) ok.store(false, std::memory_order_relaxed);
  });
  return ok.load();
}

int csr_from_coo(int64_t n, int64_t nnz, const int32_t *rows, const int32_t *cols, int threads,
                 Csr &A) {
  if (nnz > 0 && (!rows || !cols)) return fail(CGBAL_EINVAL, "rows and cols must not be NULL");
  if (!columns_in_range(rows, nnz, n, threads) || !columns_in_range(cols, nnz, n, threads))
    return fail(CGBAL_EINVAL, "an entry's row or column lies outside [0, n)");
  A.n = n;
  A.own_ptr.assign(n + 1, 0);
  for (int64_t e = 0; e < nnz; ++e) A.own_ptr[rows[e] + 1]++;
  for (int64_t r = 0; r < n; ++r) A.own_ptr[r + 1] += A.own_ptr[r];
  A.own_idx.resize(nnz);
  std::vector<int64_t> head(A.own_ptr.begin(), A.own_ptr.end() - 1);
  for (int64_t e = 0; e < nnz; ++e) A.own_idx[head[rows[e]]++] = cols[e];
  A.adopt();
  return CGBAL_OK;
}

int csr_view(int64_t n, const int64_t *ptr, const int32_t *idx, int threads, Csr &A) {
  if (!ptr) return fail(CGBAL_EINVAL, "ptr must not be NULL");
  if (ptr[0] != 0) return fail(CGBAL_EINVAL, "ptr[0] must be 0");
  for (int64_t r = 0; r < n; ++r)
    if (ptr[r + 1] < ptr[r]) return fail(CGBAL_EINVAL, "ptr must be non-decreasing");
  if (ptr[n] > 0 && !idx) return fail(CGBAL_EINVAL, "idx must not be NULL");
  if (!columns_in_range(idx, ptr[n], n, threads))
    return fail(CGBAL_EINVAL, "a column index lies outside [0, n)");
  A.n = n;
  A.ptr = ptr;
  A.idx = idx;
  return CGBAL_OK;
}

/* Makes A own its arrays and sorts every row's columns. */
void own_sorted(Csr &A, int threads) {
  if (A.own_idx.empty() && A.nnz() > 0) {
    A.own_ptr.assign(A.ptr, A.ptr + A.n + 1);
    A.own_idx.assign(A.idx, A.idx + A.nnz());
  } else if (A.own_ptr.empty()) {
    A.own_ptr.assign(A.ptr, A.ptr + A.n + 1);
  }
  A.adopt();
  int64_t *p = A.own_ptr.data();
  int32_t *x = A.own_idx.data();
  fork_join(threads, A.n, [p](int64_t r) { return p[r]; }, [p, x](int, int64_t b, int64_t e) {
    for (int64_t r = b; r < e; ++r)
      if (!std::is_sorted(x + p[r], x + p[r + 1])) std::sort(x + p[r], x + p[r + 1]);
  });
}

/* --remove_dup (analyze.cpp:504-507 with remove_duplicates :83-102): every
 * (row, column) pair once, and no diagonal entries. */
void remove_duplicates(Csr &A, int threads) {
  own_sorted(A, threads);
  const int64_t n = A.n;
  const int64_t *p = A.own_ptr.data();
  const int32_t *x = A.own_idx.data();
  std::vector<int64_t> np(n + 1, 0);
  fork_join(threads, n, [p](int64_t r) { return p[r]; }, [&](int, int64_t b, int64_t e) {
    for (int64_t r = b; r < e; ++r) {
      int64_t c = 0;
      for (int64_t i = p[r]; i < p[r + 1]; ++i)
        c += (x[i] != r) & (i == p[r] || x[i] != x[i - 1]);
      np[r + 1] = c;
    }
  });
  for (int64_t r = 0; r < n; ++r) np[r + 1] += np[r];
  std::vector<int32_t> nx(np[n]);
  fork_join(threads, n, [p](int64_t r) { return p[r]; }, [&](int, int64_t b, int64_t e) {
    for (int64_t r = b; r < e; ++r) {
      int64_t o = np[r];
      for (int64_t i = p[r]; i < p[r + 1]; ++i)
        if (x[i] != r && (i == p[r] || x[i] != x[i - 1])) nx[o++] = x[i];
    }
  });
  A.own_ptr.swap(np);
  A.own_idx.swap(nx);
  A.adopt();
}

/* --symmetric's check (analyze.cpp:519-541): the entries, repeats counted,
 * are the same multiset as those of the transpose.  With each row's columns
 * sorted, that holds exactly when the transpose, built by scattering rows in
 * ascending order (so its rows come out sorted), equals A array for array.
 * The transpose is built in parallel without atomics: each thread owns a
 * range of column ids and scans every entry, counting and then placing only
 * its own columns. */
bool structurally_symmetric(Csr &A, int threads) {
  own_sorted(A, threads);
  const int64_t n = A.n, nnz = A.nnz();
  const int64_t *p = A.ptr;
  const int32_t *x = A.idx;
  std::vector<int64_t> tp(n + 1, 0), head;
  int parts = (int)std::max<int64_t>(1, std::min<int64_t>(threads, nnz >> 16));
  // thread t owns column ids [n*t/parts, n*(t+1)/parts)
  auto lo_of = [n, parts](int t) { return (int32_t)(n * t / parts); };
  each_part(parts, [&](int t) {
    const int32_t lo = lo_of(t);
    const uint32_t w = (uint32_t)(lo_of(t + 1) - lo);
    for (int64_t i = 0; i < nnz; ++i)
      if ((uint32_t)(x[i] - lo) < w) tp[x[i] + 1]++;
  });
  for (int64_t r = 0; r < n; ++r) tp[r + 1] += tp[r];
  if (!std::equal(tp.begin(), tp.end(), p)) return false;
  std::vector<int32_t> tx(nnz);
  head.assign(tp.begin(), tp.end() - 1);
  each_part(parts, [&](int t) {
    const int32_t lo = lo_of(t);
    const uint32_t w = (uint32_t)(lo_of(t + 1) - lo);
    for (int64_t r = 0; r < n; ++r)
      for (int64_t i = p[r]; i < p[r + 1]; ++i)
        if ((uint32_t)(x[i] - lo) < w) tx[head[x[i]]++] = (int32_t)r;
  });
  std::atomic<bool> ok{true};
  fork_join_even(threads, nnz, [&](int, int64_t b, int64_t e) {
    if (!std::equal(tx.begin() + b, tx.begin() + e, x + b)) ok.store(false, std::memory_order_relaxed);
  });
  return ok.load();
}

/* Per-PE nonzero counts when vertex v sits in block blk[v] (rows and columns
 * share it): out[row_block * grid + col_block].  Returns the largest. */
int64_t count_blocks(const Csr &A, const uint16_t *blk, int grid, int threads,
                     std::vector<int64_t> &out) {
  const int64_t cells = (int64_t)grid * grid;
  out.assign(cells, 0);
  const int64_t *p = A.ptr;
  const int32_t *x = A.idx;
  auto weight = [p](int64_t r) { return p[r]; };
  if (cells * threads * (int64_t)sizeof(int64_t) <= (256 << 20)) {
    std::vector<std::vector<int64_t>> part(threads, std::vector<int64_t>(cells, 0));
    fork_join(threads, A.n, weight, [&](int t, int64_t b, int64_t e) {
      std::vector<int64_t> &c = part[t];
      for (int64_t r = b; r < e; ++r) {
        int64_t *row = c.data() + (int64_t)blk[r] * grid;
        for (int64_t i = p[r]; i < p[r + 1]; ++i) row[blk[x[i]]]++;
      }
    });
    for (auto &c : part)
      for (int64_t i = 0; i < cells; ++i) out[i] += c[i];
  } else {
    std::vector<std::atomic<int64_t>> shared(cells);
    for (auto &a : shared) a.store(0, std::memory_order_relaxed);
    fork_join(threads, A.n, weight, [&](int, int64_t b, int64_t e) {
      for (int64_t r = b; r < e; ++r) {
        std::atomic<int64_t> *row = shared.data() + (int64_t)blk[r] * grid;
        for (int64_t i = p[r]; i < p[r + 1]; ++i)
          row[blk[x[i]]].fetch_add(1, std::memory_order_relaxed);
      }
    });
    for (int64_t i = 0; i < cells; ++i) out[i] = shared[i].load(std::memory_order_relaxed);
  }
  return cells ? *std::max_element(out.begin(), out.end()) : 0;
}

/* Smallest key, lowest index on ties, with point updates: the block row to
 * fill next on large grids.  A full block row has key INT64_MAX; when every
 * row is full the answer is 0, which is what a first-minimum scan gives. */
class MinTree {
 public:
  explicit MinTree(int n) : size_(1) {
    while (size_ < n) size_ <<= 1;
    key_.assign(2 * size_, std::numeric_limits<int64_t>::max());
    at_.assign(2 * size_, 0);
    for (int i = 0; i < size_; ++i) {
      at_[size_ + i] = i;
      if (i < n) key_[size_ + i] = 0;
    }
    for (int v = size_ - 1; v >= 1; --v) pull(v);
  }
  void set(int i, int64_t k) {
    int v = size_ + i;
    key_[v] = k;
    for (v >>= 1; v >= 1; v >>= 1) pull(v);
  }
  int argmin() const { return at_[1]; }

 private:
  void pull(int v) {
    int c = key_[2 * v + 1] < key_[2 * v] ? 2 * v + 1 : 2 * v;
    key_[v] = key_[c];
    at_[v] = at_[c];
  }
  int size_;
  std::vector<int64_t> key_;
  std::vector<int> at_;
};

struct RowCount {
  int64_t row;
  int32_t count;
};

/* One greedy placement pass (analyze.cpp:267-347).  blk gives each column's
 * block under the permutation in force; q receives the new placement. */
void place_rows(const Csr &A, const std::vector<RowCount> &order, const uint16_t *blk,
                int64_t block, int grid, int threads, std::vector<uint16_t> &gathered,
                std::vector<int64_t> &chunk_off, int64_t *q) {
  const int64_t rows = (int64_t)order.size();
  std::vector<int64_t> load((int64_t)grid * grid, 0);  // [block row][block column]
  std::vector<int64_t> heaviest(grid, 0);               // heaviest PE of each block row
  std::vector<int64_t> filled(grid, 0);
  std::vector<char> full(grid, 0);
  const bool scan = grid <= kLinearScanMax;
  MinTree tree(scan ? 1 : grid);
  const int64_t *p = A.ptr;
  const int32_t *x = A.idx;

  int by = 0;  // the first row goes to block row 0
  for (int64_t k0 = 0; k0 < rows;) {
    // the next chunk: rows k0 .. k1 in placement order
    int64_t k1 = k0, got = 0;
    chunk_off.clear();
    chunk_off.push_back(0);
    while (k1 < rows && (k1 == k0 || got + order[k1].count <= kChunkEntries)) {
      got += order[k1].count;
      chunk_off.push_back(got);
      ++k1;
    }
    if ((int64_t)gathered.size() < got) gathered.resize(got);
    uint16_t *g = gathered.data();
    const int64_t *off = chunk_off.data();
    fork_join(threads, k1 - k0, [off](int64_t i) { return off[i]; },
              [&, g, off](int, int64_t b, int64_t e) {
                for (int64_t i = b; i < e; ++i) {
                  const int64_t r = order[k0 + i].row;
                  if (r >= A.n) continue;  // padding rows have no entries
                  uint16_t *dst = g + off[i];
                  for (int64_t j = p[r]; j < p[r + 1]; ++j) *dst++ = blk[x[j]];
                }
              });
    for (int64_t i = 0; i < k1 - k0; ++i) {
      const int64_t r = order[k0 + i].row;
      q[r] = (int64_t)by * block + filled[by];
      if (++filled[by] >= block) full[by] = 1;
      int64_t *row = load.data() + (int64_t)by * grid;
      int64_t mx = heaviest[by];
      for (int64_t j = off[i]; j < off[i + 1]; ++j) {
        const int64_t v = ++row[g[j]];
        mx = v > mx ? v : mx;
      }
      heaviest[by] = mx;
      if (scan) {
        int best = -1;
        int64_t best_load = 0;
        for (int b = 0; b < grid; ++b)
          if (!full[b] && (best < 0 || heaviest[b] < best_load)) {
            best = b;
            best_load = heaviest[b];
          }
        by = best < 0 ? 0 : best;
      } else {
        tree.set(by, full[by] ? std::numeric_limits<int64_t>::max() : mx);
        by = tree.argmin();
      }
    }
    k0 = k1;
  }
}

void blocks_of(const int64_t *q, int64_t n, int64_t block, int threads, std::vector<uint16_t> &blk) {
  blk.resize(n);
  fork_join_even(threads, n, [&](int, int64_t b, int64_t e) {
    for (int64_t v = b; v < e; ++v) blk[v] = (uint16_t)(q[v] / block);
  });
}

struct Shape {
  int grid;
  int64_t block, n_padded;
  bool exact;
};

int shape_of(int64_t n, int32_t fabx, int32_t faby, Shape &s) {
  if (n <= 0 || n > INT32_MAX) return fail(CGBAL_EINVAL, "n must be in [1, 2^31)");
  if (fabx <= 0 || faby <= 0) return fail(CGBAL_EINVAL, "fabx and faby must be positive");
  if (fabx != faby)
    return fail(CGBAL_EINVAL, "a shared row/column permutation needs a square PE grid (fabx == faby)");
  if (fabx > kMaxGrid) return fail(CGBAL_EINVAL, "the PE grid is larger than 4096 x 4096");
  s.grid = (int)std::min<int64_t>(fabx, n);
  // analyze.cpp:550-551 (and again :234-235 from the padded size)
  int64_t b = (int64_t)std::ceil((float)n / (float)s.grid);
  s.exact = false;
  if (b * s.grid < n || (int64_t)std::ceil((float)(b * s.grid) / (float)s.grid) != b) {
    b = (n + s.grid - 1) / s.grid;
    s.exact = true;
  }
  s.block = b;
  s.n_padded = b * s.grid;
  return CGBAL_OK;
}

int balance(Csr &A, int32_t fabx, int32_t faby, int32_t flags, int threads, int64_t *iperm,
            int64_t iperm_len, int64_t *buckets_out, cgbal_report *rep) {
  Shape s;
  if (int rc = shape_of(A.n, fabx, faby, s)) return rc;
  if (iperm_len < s.n_padded || !iperm)
    return fail(CGBAL_EINVAL, "iperm must hold cgbal_padded_n(n, fabx, faby) entries");
  if (flags & CGBAL_REMOVE_DUP) remove_duplicates(A, threads);
  if ((flags & CGBAL_SYMMETRIC) && !structurally_symmetric(A, threads))
    return fail(CGBAL_ENOTSYM,
                "the matrix is not structurally symmetric (A != A^T); a directed graph is "
                "balanced with CGBAL_SHARED_PERM");
  const int64_t n = A.n, nnz = A.nnz(), npad = s.n_padded;
  const int grid = s.grid;

  // analyze.cpp:569, 238-245: rows by decreasing count (padding rows count 0)
  std::vector<RowCount> order(npad);
  for (int64_t r = 0; r < npad; ++r)
    order[r] = {r, r < n ? (int32_t)(A.ptr[r + 1] - A.ptr[r]) : 0};
  std::sort(order.begin(), order.end(),
            [](const RowCount &a, const RowCount &b) { return a.count > b.count; });

  std::vector<int64_t> q(npad), qn(npad);
  for (int64_t v = 0; v < npad; ++v) q[v] = v;
  std::vector<uint16_t> blk, blk_new, gathered;
  std::vector<int64_t> counts, counts_new, chunk_off;
  blocks_of(q.data(), n, s.block, threads, blk);
  const int64_t max_identity = count_blocks(A, blk.data(), grid, threads, counts);

  const double pes = (double)grid * (double)grid;
  double cur = 0.0;
  int passes = 0;
  for (;;) {
    ++passes;
    place_rows(A, order, blk.data(), s.block, grid, threads, gathered, chunk_off, qn.data());
    blocks_of(qn.data(), n, s.block, threads, blk_new);
    const int64_t mx = count_blocks(A, blk_new.data(), grid, threads, counts_new);
    const double prev = cur;
    cur = mx > 0 ? (double)nnz / (pes * (double)mx) : std::numeric_limits<double>::quiet_NaN();
    if (!(cur > prev)) break;
    q.swap(qn);
    blk.swap(blk_new);
    counts.swap(counts_new);
  }

  std::copy(q.begin(), q.end(), iperm);
  if (buckets_out)  // the reference's order: px * grid + py (analyze.cpp:625-631)
    for (int py = 0; py < grid; ++py)
      for (int px = 0; px < grid; ++px)
        buckets_out[(int64_t)px * grid + py] = counts[(int64_t)py * grid + px];
  if (rep) {
    const int64_t mx = *std::max_element(counts.begin(), counts.end());
    rep->n_padded = npad;
    rep->block = s.block;
    rep->nnz = nnz;
    rep->max_identity = max_identity;
    rep->max_balanced = mx;
    rep->lb = mx > 0 ? (double)nnz / (pes * (double)mx) : 0.0;
    rep->grid = grid;
    rep->passes = passes;
    rep->exact_blocks = s.exact ? 1 : 0;
    rep->reserved = 0;
  }
  return CGBAL_OK;
}

int check_flags(int32_t flags) {
  const int mode = flags & (CGBAL_SYMMETRIC | CGBAL_SHARED_PERM);
  if (flags & ~(CGBAL_SYMMETRIC | CGBAL_SHARED_PERM | CGBAL_REMOVE_DUP))
    return fail(CGBAL_EINVAL, "unknown flag bits");
  if (mode == (CGBAL_SYMMETRIC | CGBAL_SHARED_PERM))
    return fail(CGBAL_EINVAL, "CGBAL_SYMMETRIC and CGBAL_SHARED_PERM are mutually exclusive");
  if (mode == 0)
    return fail(CGBAL_EINVAL,
                "one of CGBAL_SYMMETRIC or CGBAL_SHARED_PERM is required: the permutation is "
                "always shared by rows and columns");
  return CGBAL_OK;
}

}  // namespace

extern "C" {

CGBAL_API int64_t cgbal_padded_n(int64_t n, int32_t fabx, int32_t faby) {
  Shape s;
  return shape_of(n, fabx, faby, s) == CGBAL_OK ? s.n_padded : -1;
}

CGBAL_API int cgbal_balance_coo(int64_t n, int64_t nnz, const int32_t *rows, const int32_t *cols,
                                int32_t fabx, int32_t faby, int32_t flags, int32_t threads,
                                int64_t *iperm, int64_t iperm_len, int64_t *buckets,
                                cgbal_report *rep) {
  try {
    if (int rc = check_flags(flags)) return rc;
    Shape s;
    if (int rc = shape_of(n, fabx, faby, s)) return rc;
    if (nnz < 0 || nnz > INT32_MAX) return fail(CGBAL_EINVAL, "nnz must be in [0, 2^31)");
    const int t = pick_threads(threads);
    Csr A;
    if (int rc = csr_from_coo(n, nnz, rows, cols, t, A)) return rc;
    return balance(A, fabx, faby, flags, t, iperm, iperm_len, buckets, rep);
  } catch (const std::bad_alloc &) {
    return fail(CGBAL_ENOMEM, "host allocation failed");
  }
}

CGBAL_API int cgbal_balance_csr(int64_t n, const int64_t *ptr, const int32_t *idx, int32_t fabx,
                                int32_t faby, int32_t flags, int32_t threads, int64_t *iperm,
                                int64_t iperm_len, int64_t *buckets, cgbal_report *rep) {
  try {
    if (int rc = check_flags(flags)) return rc;
    Shape s;
    if (int rc = shape_of(n, fabx, faby, s)) return rc;
    const int t = pick_threads(threads);
    Csr A;
    if (int rc = csr_view(n, ptr, idx, t, A)) return rc;
    if (A.nnz() > INT32_MAX) return fail(CGBAL_EINVAL, "nnz must be below 2^31");
    return balance(A, fabx, faby, flags, t, iperm, iperm_len, buckets, rep);
  } catch (const std::bad_alloc &) {
    return fail(CGBAL_ENOMEM, "host allocation failed");
  }
}

CGBAL_API const char *cgbal_last_error(void) { return t_err.c_str(); }

}  // extern "C"
