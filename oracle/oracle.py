"""CPU oracle for the BFS hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or the timed CPU baseline; the product
(ceregraphs_amd) never does.

Contents (each cites the reference behaviour it restates):
  * ``bfs``            -- C restatement (bfs_oracle.c): levels, lowest-id
                          parents, per-round nf history, rounds_completed.
  * ``lowest_parents`` -- the lowest-id parent rule computed a second,
                          independent way (vectorised numpy over all edges).
  * ``validate``       -- Graph500-style structural validation, stricter than
                          the reference's invalid_parents
                          (bfs/plots/bfs_tree_plot.py:140-152).
  * ``spmv_bfs``       -- the reference's round loop restated literally as a
                          boolean SpMV y = OR_j(A[i,j] AND x[j]) with visited
                          masking (bool_pe.csl:865-1004), one scipy sparse
                          product per round; the CPU baseline.
  * ``scipy_levels``   -- the reference's own correctness oracle,
                          scipy breadth_first_order on M^T (run_bfs.py:513-552).
  * ``m_edges``        -- m of bfs_timing.py:390-395.

Pinning: every function here is checked in tests/test_oracle.py against the
golden fixtures written from the reference (tests/golden/), against scipy's
breadth_first_order, against brute force on tiny graphs, and against closed
forms (path, star, two components, self-loops).
"""
import ctypes
import os
import subprocess

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import breadth_first_order

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
  subprocess.check_call(["make", "-s", "-C", HERE, "all"])


def _load():
  global _lib
  if _lib is None:
    if not os.path.exists(LIB):
      build()
    lib = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    lib.oracle_bfs.restype = ctypes.c_int
    lib.oracle_bfs.argtypes = [ctypes.c_int64, P, P, P, P, ctypes.c_int64, P, P, P,
                               ctypes.c_int64, P]
    lib.oracle_m_edges.restype = ctypes.c_int64
    lib.oracle_m_edges.argtypes = [ctypes.c_int64, P, P, P, ctypes.c_int]
    _lib = lib
  return _lib


def _p(a):
  return ctypes.c_void_p(a.ctypes.data)


class Graph:
  """Host CSR pair in the reference orientation (M[r][c] = edge c -> r)."""

  def __init__(self, n, in_ptr, in_idx, out_ptr, out_idx, symmetric):
    self.n = int(n)
    self.in_ptr = np.ascontiguousarray(in_ptr, np.int64)
    self.in_idx = np.ascontiguousarray(in_idx, np.int32)
    self.out_ptr = np.ascontiguousarray(out_ptr, np.int64)
    self.out_idx = np.ascontiguousarray(out_idx, np.int32)
    self.symmetric = bool(symmetric)

  @classmethod
  def from_matrix(cls, M):
    """From a scipy sparse M (row = destination, column = source)."""
    csr = sp.csr_matrix(M, copy=True)
    csr.sum_duplicates()
    csr.sort_indices()
    csc = csr.tocsc()
    csc.sort_indices()
    sym = (csr != csr.T).nnz == 0
    return cls(csr.shape[0], csr.indptr, csr.indices, csc.indptr, csc.indices, sym)

  def matrix(self):
    data = np.ones(len(self.in_idx), dtype=np.int8)
    return sp.csr_matrix((data, self.in_idx, self.in_ptr), shape=(self.n, self.n))

  def coo_rows_cols(self):
    rows = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.in_ptr))
    return rows, self.in_idx.astype(np.int64)


def bfs(g, source, max_hist=4096):
  """(levels int32[n], parents int32[n], nf_history, rounds_completed)."""
  lib = _load()
  lev = np.empty(g.n, np.int32)
  par = np.empty(g.n, np.int32)
  hist = np.zeros(max_hist, np.int64)
  rounds = ctypes.c_int64(0)
  rc = lib.oracle_bfs(g.n, _p(g.out_ptr), _p(g.out_idx), _p(g.in_ptr), _p(g.in_idx), int(source),
                      _p(par), _p(lev), _p(hist), max_hist, ctypes.byref(rounds))
  if rc != 0:
    raise ValueError(f"oracle_bfs failed ({rc})")
  r = rounds.value
  return lev, par, hist[:min(r, max_hist)].copy(), r


def m_edges(g, levels):
  return int(_load().oracle_m_edges(g.n, _p(g.in_ptr), _p(g.in_idx),
                                    _p(np.ascontiguousarray(levels, np.int32)),
                                    1 if g.symmetric else 0))


def lowest_parents(g, levels, source):
  """parent[v] = min{u : M[v][u] != 0, level[u] = level[v]-1} over all edges
  at once (rows ascending, so the first qualifying entry per row is the min)."""
  levels = np.asarray(levels, np.int64)
  rows, cols = g.coo_rows_cols()
  ok = (levels[rows] > 0) & (levels[cols] == levels[rows] - 1)
  r, c = rows[ok], cols[ok]
  par = np.full(g.n, -1, np.int64)
  if len(r):
    first = np.concatenate([[True], r[1:] != r[:-1]])
    par[r[first]] = c[first]
  par[source] = source
  return par.astype(np.int32)


def validate(g, source, parents, levels):
  """Graph500-style checks; returns a list of failure strings (empty = ok).

  1. root: level 0, parent itself;
  2. unreached vertices have level -1 and parent -1;
  3. every tree edge exists in M and joins levels l-1 -> l;
  4. every edge c -> r with c reached has r reached and level[r] <= level[c]+1
     (so the reachable set is complete and levels are shortest);
  5. parents follow the lowest-id rule.
  """
  errs = []
  parents = np.asarray(parents, np.int64)
  levels = np.asarray(levels, np.int64)
  n = g.n
  if levels[source] != 0 or parents[source] != source:
    errs.append("root must have level 0 and parent == root")
  unreached = levels < 0
  if np.any(parents[unreached] != -1) or np.any(levels[unreached] != -1):
    errs.append("unreached vertices must carry -1")
  reached = ~unreached
  if np.any(parents[reached] < 0):
    errs.append("reached vertex without a parent")
  v = np.nonzero(reached)[0]
  v = v[v != source]
  p = parents[v]
  if len(v):
    if np.any((p < 0) | (p >= n)):
      errs.append("parent id out of range")
    else:
      if np.any(levels[p] != levels[v] - 1):
        errs.append("parent is not one level closer")
      # edge p -> v exists: p in row v of M (ascending)
      lo = g.in_ptr[v]
      hi = g.in_ptr[v + 1]
      pos = np.array([np.searchsorted(g.in_idx[a:b], q) for a, b, q in zip(lo, hi, p)]
                     if len(v) < 2048 else _batched_search(g, v, p))
      exists = (pos < hi - lo) & (g.in_idx[np.minimum(lo + pos, max(len(g.in_idx) - 1, 0))] == p)
      if np.any(~exists):
        errs.append("tree edge missing from the graph")
  rows, cols = g.coo_rows_cols()
  src_reached = levels[cols] >= 0
  if np.any(levels[rows[src_reached]] < 0):
    errs.append("reachable set incomplete (edge from a reached vertex to an unreached one)")
  elif np.any(levels[rows[src_reached]] > levels[cols[src_reached]] + 1):
    errs.append("edge spans more than one level")
  if not np.array_equal(lowest_parents(g, levels, source), parents):
    errs.append("parents differ from the lowest-id rule")
  return errs


def _batched_search(g, v, p):
  # position of p within row v, via a global sort key (row * n + col)
  keys = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.in_ptr)) * g.n + g.in_idx
  want = v.astype(np.int64) * g.n + p
  idx = np.searchsorted(keys, want)
  return idx - g.in_ptr[v]


def scipy_levels(g, source):
  """Hop distances from scipy's breadth_first_order on M^T (the reference's
  own oracle, run_bfs.py:518-527), as BFS-tree depths; -1 unreached."""
  fwd = g.matrix().transpose().tocsr()
  order, pred = breadth_first_order(fwd, source, directed=True, return_predecessors=True)
  depth = np.full(g.n, -1, np.int32)
  depth[source] = 0
  for v in order[1:]:
    depth[v] = depth[pred[v]] + 1
  return depth


def spmv_bfs(g, source, tau=None, matrix=None):
  """The reference round loop as boolean SpMV (bool_pe.csl:865-1156): each
  round y = M x over the boolean semiring, newly = y & ~visited, stop when
  newly is empty.  Returns (levels, nf_history, direction_history,
  rounds_completed).  tau: optional one-way switch threshold, recorded in
  direction_history exactly as term_col_bcast_done does (the product is the
  same either way).  matrix: g.matrix().astype(np.int32), prebuilt so a
  timed caller can leave construction out, as the device search does."""
  M = matrix if matrix is not None else g.matrix().astype(np.int32)
  x = np.zeros(g.n, dtype=bool)
  x[source] = True
  visited = x.copy()
  levels = np.full(g.n, -1, np.int32)
  levels[source] = 0
  nf_hist, dir_hist = [], [0]
  bottom_up = False
  r = 0
  while True:
    y = (M @ x.astype(np.int32)) > 0
    newly = y & ~visited
    visited |= newly
    levels[newly] = r + 1
    nf = int(newly.sum())
    nf_hist.append(nf)
    r += 1
    if nf == 0:
      break
    if tau is not None and not bottom_up and nf > tau:
      bottom_up = True
    dir_hist.append(1 if bottom_up else 0)
    x = newly
  return levels, np.array(nf_hist, np.int64), np.array(dir_hist[:r], np.int32), r
