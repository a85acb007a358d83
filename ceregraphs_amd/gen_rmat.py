"""R-MAT graph generation (host side) and the PCG64 seed hand-off to the GPU
generator.

``rmat`` restates datasets/gen_rmat.py:15-44: ``m = n * edgefactor`` edges,
``scale`` levels, each level drawing ``rng.random(m)`` from
``numpy.random.default_rng(seed)`` and adding the level's bit to the source
(u), the destination (v) or both by quadrant, followed (symmetric case) by
symmetrise, drop self-loops and dedup.  Draw order and float64 thresholds are
what fix the edge set, so this module keeps both exactly.

``directed=True`` is this project's variant for unsymmetrised graphs (the
reference generator always symmetrises): edge u->v is stored as M[v][u]
(row = destination, bfs/implementation/graph_loader.py:76-90), self-loops
dropped, duplicates merged.

CLI (mirrors gen_rmat.py's):
    python -m ceregraphs_amd.gen_rmat <scale> <edgefactor> <seed> <out.mtx> [--directed]
"""
import argparse

import numpy as np
import scipy.sparse as sp
from scipy.io import mmwrite

GRAPH500_ABCD = (0.57, 0.19, 0.19, 0.05)


def thresholds(a, b, c):
  """Cumulative quadrant thresholds as float64, in the reference's order of
  evaluation (A, A+B, A+B+C; gen_rmat.py:22-23)."""
  ab = a + b
  abc = a + b + c
  return float(a), float(ab), float(abc)


def pcg_words(seed):
  """numpy.random.default_rng(seed)'s PCG64 state as four uint64 words
  (state_hi, state_lo, inc_hi, inc_lo) for the GPU generator."""
  st = np.random.default_rng(seed).bit_generator.state["state"]
  s, inc = int(st["state"]), int(st["inc"])
  mask = (1 << 64) - 1
  return np.array([s >> 64, s & mask, inc >> 64, inc & mask], dtype=np.uint64)


def rmat_endpoints(scale, edgefactor, seed, a=0.57, b=0.19, c=0.19):
  """Sampled (u, v) endpoints, int64, in draw order, before any cleanup."""
  rng = np.random.default_rng(seed)
  n = 1 << scale
  m = n * edgefactor
  cuts = np.array(thresholds(a, b, c), dtype=np.float64)
  u = np.zeros(m, dtype=np.int64)
  v = np.zeros(m, dtype=np.int64)
  for level in range(scale):
    r = rng.random(m)
    # quadrant index: 0 below A, 1 in [A, A+B), 2 in [A+B, A+B+C), 3 above
    q = np.searchsorted(cuts, r, side="right")
    bit = np.int64(1) << np.int64(scale - 1 - level)
    u += (q >= 2) * bit
    v += (q & 1) * bit
  return u, v


def _csr_from_keys(keys, n):
  keys = np.unique(keys)
  rows = keys // n
  cols = (keys % n).astype(np.int32)
  indptr = np.zeros(n + 1, dtype=np.int64)
  np.cumsum(np.bincount(rows, minlength=n), out=indptr[1:])
  return indptr, cols


def rmat_csr(scale, edgefactor, seed, a=0.57, b=0.19, c=0.19, directed=False):
  """Returns (n, in_ptr, in_idx, out_ptr, out_idx): row-compressed M (the
  in-neighbour view) and column-compressed M (the out-neighbour view), int64
  offsets and int32 ascending indices.  Symmetric graphs share one view."""
  u, v = rmat_endpoints(scale, edgefactor, seed, a, b, c)
  n = 1 << scale
  keep = u != v
  u, v = u[keep], v[keep]
  if not directed:
    keys = np.concatenate([u * n + v, v * n + u])
    ptr, idx = _csr_from_keys(keys, n)
    return n, ptr, idx, ptr, idx
  in_ptr, in_idx = _csr_from_keys(v * n + u, n)   # row v of M: sources u
  out_ptr, out_idx = _csr_from_keys(u * n + v, n)  # column u of M: targets v
  return n, in_ptr, in_idx, out_ptr, out_idx


def rmat(scale, edgefactor, seed, a=0.57, b=0.19, c=0.19, directed=False):
  """scipy CSR of M with all-1.0 data, like gen_rmat.rmat's return value."""
  n, in_ptr, in_idx, _, _ = rmat_csr(scale, edgefactor, seed, a, b, c, directed)
  data = np.ones(len(in_idx), dtype=np.float64)
  return sp.csr_matrix((data, in_idx, in_ptr), shape=(n, n))


def main(argv=None):
  p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
  p.add_argument("scale", type=int)
  p.add_argument("edgefactor", type=int)
  p.add_argument("seed", type=int, nargs="?", default=0)
  p.add_argument("outfile")
  p.add_argument("--directed", action="store_true")
  p.add_argument("--abcd", default=None, help="comma list A,B,C,D (default .57,.19,.19,.05)")
  args = p.parse_args(argv)
  a, b, c, _ = GRAPH500_ABCD if args.abcd is None else [float(x) for x in args.abcd.split(",")]
  A = rmat(args.scale, args.edgefactor, args.seed, a, b, c, directed=args.directed)
  deg = np.diff(A.indptr)
  print(f"n={A.shape[0]}, nnz={A.nnz}, avg_degree={A.nnz / A.shape[0]:.2f}, max_degree={deg.max()}")
  mmwrite(args.outfile, A, field="real")


if __name__ == "__main__":
  main()
