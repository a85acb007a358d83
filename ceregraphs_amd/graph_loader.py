"""Graph ingest: Matrix Market or SNAP edge lists -> the two adjacency views
the device library takes.

Mirrors bfs/implementation/graph_loader.py:27-95: the format is detected
from the file's first line (``%%MatrixMarket`` or not), Matrix Market goes
through scipy.io.mmread (which also expands "symmetric" storage), and an edge
list line ``src dst`` becomes M[dst][src] = 1 (row = destination, column =
source, graph_loader.py:76-90), with the vertex count taken from a
``# Nodes: N`` comment when present.

``HostGraph`` holds row-compressed M (in-neighbours, ascending: the pull
view) and column-compressed M (out-neighbours: the push view), int64 offsets
and int32 indices, duplicates merged.  ``load_csr_cached`` adds a binary
cache next to the source file so large graphs are parsed once
(construction, Graph500 "Kernel 1").
"""
import gzip
import os
import re
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
from scipy.io import mmread

MTX_MAGIC = "%%MatrixMarket"
NODES_DECLARED_RE = re.compile(r"Nodes:\s*(\d+)")
CACHE_SUFFIX = ".cgcsr.npz"


@dataclass
class HostGraph:
  n: int
  in_ptr: np.ndarray   # int64[n+1], row r of M
  in_idx: np.ndarray   # int32, ascending within each row
  out_ptr: np.ndarray  # int64[n+1], column c of M
  out_idx: np.ndarray  # int32, ascending within each column
  symmetric: bool

  @property
  def nnz(self):
    return int(len(self.in_idx))

  def out_degree_no_loops(self):
    """Per-vertex out-degree minus its self-loop (run_graph500.py:119)."""
    deg = np.diff(self.out_ptr)
    cols = np.repeat(np.arange(self.n, dtype=np.int64), deg)
    loops = np.bincount(cols[self.out_idx == cols], minlength=self.n)
    return deg - loops

  def coo(self):
    """(rows, cols) of every nonzero of M, row-major."""
    rows = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.in_ptr))
    return rows, self.in_idx.astype(np.int64)

  def to_device(self, device=0):
    from .engine import DeviceGraph
    return DeviceGraph.from_csr(self.n, self.out_ptr, self.out_idx, self.in_ptr, self.in_idx,
                                symmetric=self.symmetric, device=device)


def _open(path, mode):
  if path.endswith(".gz"):
    return gzip.open(path, mode, encoding="utf-8")
  return open(path, mode, encoding="utf-8")


def _is_mtx(path):
  with _open(path, "rt") as f:
    return f.readline().startswith(MTX_MAGIC)


def _load_edgelist(path):
  src, dst = [], []
  declared = None
  with _open(path, "rt") as f:
    for line in f:
      line = line.strip()
      if not line:
        continue
      if line.startswith("#"):
        m = NODES_DECLARED_RE.search(line)
        if m:
          declared = int(m.group(1))
        continue
      parts = line.split()
      src.append(int(parts[0]))
      dst.append(int(parts[1]))
  inferred = (max(max(src), max(dst)) + 1) if src else 0
  n = max(declared or 0, inferred)
  return sp.coo_matrix((np.ones(len(src)), (np.array(dst, np.int64), np.array(src, np.int64))),
                       shape=(n, n))


def load_graph(path):
  """scipy COO of M (structural; values ignored) from MTX or an edge list."""
  if _is_mtx(path):
    return sp.coo_matrix(mmread(path))
  return _load_edgelist(path)


def views_from_matrix(M):
  """HostGraph from any scipy sparse M (square, row = destination)."""
  M = sp.csr_matrix(M)
  if M.shape[0] != M.shape[1]:
    raise ValueError("BFS needs a square adjacency matrix")
  M = sp.csr_matrix((np.ones(M.nnz, np.int8), M.indices, M.indptr), shape=M.shape)
  M.sum_duplicates()
  M.sort_indices()
  n = M.shape[0]
  C = M.tocsc()
  C.sort_indices()
  symmetric = (M != M.T).nnz == 0
  in_ptr = M.indptr.astype(np.int64)
  in_idx = M.indices.astype(np.int32)
  if symmetric:
    return HostGraph(n, in_ptr, in_idx, in_ptr, in_idx, True)
  return HostGraph(n, in_ptr, in_idx, C.indptr.astype(np.int64), C.indices.astype(np.int32), False)


def load_csr(path):
  return views_from_matrix(load_graph(path))


def load_csr_cached(path, cache=True):
  """load_csr with a binary cache (``<path>.cgcsr.npz``) keyed on the source
  file's size and modification time."""
  cpath = path + CACHE_SUFFIX
  st = os.stat(path)
  key = np.array([st.st_size, int(st.st_mtime_ns)], dtype=np.int64)
  if cache and os.path.exists(cpath):
    with np.load(cpath) as z:
      if np.array_equal(z["key"], key):
        sym = bool(z["symmetric"])
        in_ptr, in_idx = z["in_ptr"], z["in_idx"]
        if sym:
          return HostGraph(int(z["n"]), in_ptr, in_idx, in_ptr, in_idx, True)
        return HostGraph(int(z["n"]), in_ptr, in_idx, z["out_ptr"], z["out_idx"], False)
  g = load_csr(path)
  if cache:
    extra = {} if g.symmetric else {"out_ptr": g.out_ptr, "out_idx": g.out_idx}
    tmp = cpath + ".tmp.npz"
    np.savez(tmp, key=key, n=np.int64(g.n), symmetric=np.bool_(g.symmetric), in_ptr=g.in_ptr,
             in_idx=g.in_idx, **extra)
    os.replace(tmp, cpath)
  return g


def read_operm(path):
  """A balancer's ``.operm``: line i holds original vertex i's id in the
  balanced matrix (util/analyze.cpp:700-715)."""
  return np.loadtxt(path, dtype=np.int64, ndmin=1)


def balanced_source(operm, source_orig):
  """Translate an original vertex id to the balanced id (docs/ERRORS.md:99-103)."""
  if not 0 <= source_orig < len(operm):
    raise ValueError(f"source {source_orig} outside the permutation (n={len(operm)})")
  return int(operm[source_orig])
