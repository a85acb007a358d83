"""Writes the golden fixtures under tests/golden/ (run once, in the build
container; the fixtures are committed, this script is not needed at test
time and never runs on a GPU machine).

What it records, and from where:

* ``rmat_s{S}_ef{E}_seed{K}.npz`` -- the reference generator's own output
  (datasets/gen_rmat.py:15-44, ``rmat(scale, edgefactor, seed)``), imported
  from the reference checkout.  Nothing else pins its exact edge set (the
  invariants -- symmetry, zero diagonal, n = 2**scale -- leave the quadrant
  thresholds and draw order open), so its CSR arrays are stored as vectors.
* ``c1_balanced_4x4.npz`` -- configuration-1's input: the s10 ef16 seed-0
  graph written with ``scipy.io.mmwrite(field='real')`` exactly as
  gen_rmat.py:54 does, then balanced for a 4x4 PE grid by the reference's own
  ``util/analyze --symmetric`` (built from its sources by ``make -C oracle
  ref`` into oracle/_ref/).  Stores the balanced CSR and the ``.operm``
  vector (original id -> balanced id).
* ``c1_scipy_src0.npz`` -- the reference's own correctness oracle on that
  input (run_bfs.py:513-552): scipy ``breadth_first_order`` on M^T from
  balanced source 0, giving the visited set, and BFS-tree depths as levels.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np
import scipy.sparse as sp
from scipy.io import mmread, mmwrite
from scipy.sparse.csgraph import breadth_first_order

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"

GEN_CASES = [(10, 16, 0), (8, 4, 5), (12, 8, 3)]


def _ref_rmat():
  sys.path.insert(0, os.path.join(REF, "datasets"))
  import gen_rmat  # the reference generator, imported here only
  return gen_rmat.rmat


def _tree_depths(pred, order, source):
  depth = np.full(len(pred), -1, dtype=np.int32)
  depth[source] = 0
  for v in order[1:]:
    depth[v] = depth[pred[v]] + 1
  return depth


def main():
  rmat = _ref_rmat()
  for scale, ef, seed in GEN_CASES:
    A = rmat(scale, ef, seed).tocsr()
    A.sort_indices()
    np.savez_compressed(os.path.join(HERE, f"rmat_s{scale}_ef{ef}_seed{seed}.npz"),
                        indptr=A.indptr.astype(np.int64), indices=A.indices.astype(np.int32),
                        n=np.int64(A.shape[0]))
    print(f"rmat s{scale} ef{ef} seed{seed}: n={A.shape[0]} nnz={A.nnz}")

  analyze = os.path.join(REPO, "oracle", "_ref", "analyze")
  subprocess.check_call(["make", "-C", os.path.join(REPO, "oracle"), "ref"])
  with tempfile.TemporaryDirectory() as td:
    raw = os.path.join(td, "rmat_s10_ef16.mtx")
    mmwrite(raw, rmat(10, 16, 0), field="real")
    bal = os.path.join(td, "rmat_s10_ef16.4x4.lb.mtx")
    operm = os.path.join(td, "rmat_s10_ef16.4x4.operm")
    subprocess.check_call([analyze, "--matrix", raw, "--symmetric", "--fabx", "4",
                           "--faby", "4", "--omatrix", bal, "--operm", operm],
                          stdout=subprocess.DEVNULL)
    with open(bal, encoding="utf-8") as f:
      header = f.readline().strip()
    B = mmread(bal).tocsr()
    B.sort_indices()
    perm = np.loadtxt(operm, dtype=np.int64)
  np.savez_compressed(os.path.join(HERE, "c1_balanced_4x4.npz"),
                      indptr=B.indptr.astype(np.int64), indices=B.indices.astype(np.int32),
                      n=np.int64(B.shape[0]), operm=perm.astype(np.int32),
                      header=np.array(header))
  print(f"c1 balanced: n={B.shape[0]} nnz={B.nnz} header={header!r}")

  source = 0
  fwd = B.transpose().tocsr()
  order, pred = breadth_first_order(fwd, source, directed=True, return_predecessors=True)
  visited = np.zeros(B.shape[0], dtype=bool)
  visited[order] = True
  levels = _tree_depths(pred, order, source)
  np.savez_compressed(os.path.join(HERE, "c1_scipy_src0.npz"), visited=visited,
                      levels=levels, source=np.int64(source))
  print(f"c1 scipy: visited={int(visited.sum())} max_level={int(levels.max())}")


if __name__ == "__main__":
  main()
