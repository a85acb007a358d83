/* cgbalance.h -- C ABI of the vertex-identity-preserving load balancer.
 *
 * Replaces the reference's offline balancing tool util/analyze.cpp
 * (iterative_load_balance :226-372, main :374-728): given a square sparse
 * matrix A (a graph's adjacency matrix, rows and columns in one id space) and
 * a P x P PE grid, it finds ONE permutation q, shared by rows and columns, so
 * that B = A(q, q) spreads its nonzeros evenly over the P*P blocks of
 * ceil(n/P) rows by ceil(n/P) columns. q is the reference's `.operm`:
 * q[original id] = balanced id (analyze.cpp:700-715), 0-based, over the
 * padded vertex count P * ceil(n/P) (analyze.cpp:545-554).
 *
 * The permutation is the reference's, entry for entry: the same greedy row
 * placement, the same acceptance rule between passes and the same ordering of
 * equal-count rows (libstdc++ std::sort with the same comparison, as
 * analyze.cpp:238-245 calls it).
 *
 * Host code only (CPU threads); no device memory is touched.
 */
#ifndef CGBALANCE_H_
#define CGBALANCE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGBAL_API __attribute__((visibility("default")))

/* status codes */
#define CGBAL_OK 0
#define CGBAL_EINVAL 1  /* bad argument (shape, grid, index out of range) */
#define CGBAL_ENOMEM 2  /* host allocation failed */
#define CGBAL_ENOTSYM 3 /* CGBAL_SYMMETRIC asked for and A != A^T */

/* flags: exactly one of SYMMETRIC / SHARED_PERM, optionally REMOVE_DUP */
#define CGBAL_SYMMETRIC 1   /* analyze --symmetric: refuse unless A == A^T (analyze.cpp:519-541) */
#define CGBAL_SHARED_PERM 2 /* analyze --shared-perm: directed input allowed */
#define CGBAL_REMOVE_DUP 4  /* analyze --remove_dup: drop repeated entries and the diagonal first (analyze.cpp:504-507) */

typedef struct {
  int64_t n_padded;     /* grid * block, the length of the permutation */
  int64_t block;        /* rows (and columns) per PE block */
  int64_t nnz;          /* entries balanced (after CGBAL_REMOVE_DUP) */
  int64_t max_identity; /* nonzeros of the heaviest PE under the identity placement */
  int64_t max_balanced; /* nonzeros of the heaviest PE under the returned permutation */
  double lb;            /* nnz / (grid * grid * max_balanced), 1.0 = perfectly even */
  int32_t grid;         /* P after clamping to n (analyze.cpp:545-548) */
  int32_t passes;       /* greedy placement passes run, the rejected last one included */
  int32_t exact_blocks; /* 1 where single-precision ceil(n / P) would leave rows unplaced (n above 2^24) and the exact ceiling is used */
  int32_t reserved;
} cgbal_report;

/* Length of the permutation for an n-vertex matrix on a fabx x faby grid
 * (P * ceil(n / P)), or -1 for a bad shape. */
CGBAL_API int64_t cgbal_padded_n(int64_t n, int32_t fabx, int32_t faby);

/* Balances the n x n matrix given as nnz (row, col) pairs, 0-based, in any
 * order (the pairs of util/mmio.c's mm_read_unsymmetric_sparse, analyze.cpp:
 * 492-501).  fabx must equal faby (analyze.cpp:509-518).  iperm receives
 * cgbal_padded_n(n, fabx, faby) entries (iperm_len must be at least that);
 * buckets, when not NULL, receives grid*grid per-PE nonzero counts of the
 * balanced placement in the reference's order, entry px*grid + py
 * (analyze.cpp:625-631).  threads <= 0 picks the machine's hardware threads,
 * at most 16.  rep may be NULL. */
CGBAL_API int cgbal_balance_coo(int64_t n, int64_t nnz, const int32_t *rows, const int32_t *cols,
                                int32_t fabx, int32_t faby, int32_t flags, int32_t threads,
                                int64_t *iperm, int64_t iperm_len, int64_t *buckets,
                                cgbal_report *rep);

/* The same for a matrix held as CSR: row r's columns are
 * idx[ptr[r] .. ptr[r+1]), in any order (the CSR analyze.cpp:574 builds). */
CGBAL_API int cgbal_balance_csr(int64_t n, const int64_t *ptr, const int32_t *idx, int32_t fabx,
                                int32_t faby, int32_t flags, int32_t threads, int64_t *iperm,
                                int64_t iperm_len, int64_t *buckets, cgbal_report *rep);

/* Message for the last failing call on this thread. */
CGBAL_API const char *cgbal_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CGBALANCE_H_ */
