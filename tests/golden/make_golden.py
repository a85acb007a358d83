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
* ``balance_cases.npz`` -- the reference balancer's ``.operm`` and heaviest
  PE for a set of small matrices (directed with --shared-perm, symmetric with
  --symmetric, padded grids, a 100 x 100 grid, --remove_dup over repeated and
  diagonal entries, n below P, an empty matrix, a star whose equal-degree
  leaves make the sort's tie order visible), each run through
  ``oracle/_ref/analyze``.  The inputs are stored with them.  One balanced
  ``.mtx`` is kept as bytes (``dup_p6_omatrix``) to pin value formatting.

    python tests/golden/make_golden.py            # everything
    python tests/golden/make_golden.py balance    # balance_cases.npz only
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


def write_mtx(path, n, rows, cols, vals=None):
  """A "real general" Matrix Market file, 1-based, values as Python's repr."""
  with open(path, "w", encoding="ascii") as f:
    f.write(f"%%MatrixMarket matrix coordinate real general\n{n} {n} {len(rows)}\n")
    for i in range(len(rows)):
      v = 1.0 if vals is None else float(vals[i])
      f.write(f"{int(rows[i]) + 1} {int(cols[i]) + 1} {v!r}\n")


def balance_inputs():
  """Input matrices by name, (n, rows, cols, vals); and the cases,
  (tag, input name, grid, mode, remove_dup)."""
  g = np.load(os.path.join(HERE, "rmat_s12_ef8_seed3.npz"))
  n12 = int(g["n"])
  r12 = np.repeat(np.arange(n12), np.diff(g["indptr"]))
  c12 = g["indices"].astype(np.int64)
  keep = (r12 > c12) | ((r12 + 2 * c12) % 3 == 0)  # directed: not closed under swap
  rng = np.random.default_rng(7)
  nd = 1000
  rr = rng.integers(0, nd, 8000)
  cc = rng.integers(0, nd, 8000)
  rr = np.concatenate([rr, rr[:500], np.arange(50)])  # repeats and diagonal entries
  cc = np.concatenate([cc, cc[:500], np.arange(50)])
  p = rng.permutation(len(rr))
  rr, cc = rr[p], cc[p]
  vals = rng.choice([1.0, 0.5, 1e-7, 123456.7, -3.0, 1e6, 2.5e10, 999999.0, -0.25, 3.14159265],
                    len(rr))
  ns = 2000
  sr = np.concatenate([np.zeros(ns - 1, np.int64), np.arange(1, ns)])
  sc = np.concatenate([np.arange(1, ns), np.zeros(ns - 1, np.int64)])
  e = np.array([], np.int64)
  inputs = {"dir12": (n12, r12[keep], c12[keep], None), "sym12": (n12, r12, c12, None),
            "rand1000": (nd, rr, cc, vals), "tiny": (3, np.array([0, 1, 2, 2]), np.array([1, 2, 0, 1]), None),
            "star": (ns, sr, sc, None), "empty": (10, e, e, None)}
  cases = [("dir12_p3", "dir12", 3, "shared", False), ("dir12_p7", "dir12", 7, "shared", False),
           ("dir12_p100", "dir12", 100, "shared", False), ("sym12_p9", "sym12", 9, "symmetric", False),
           ("dup_p6", "rand1000", 6, "shared", True), ("nodup_p6", "rand1000", 6, "shared", False),
           ("tiny_p4", "tiny", 4, "shared", False), ("star_p8", "star", 8, "symmetric", False),
           ("empty_p4", "empty", 4, "shared", False)]
  return inputs, cases


def balance_fixtures():
  import re
  analyze = os.path.join(REPO, "oracle", "_ref", "analyze")
  subprocess.check_call(["make", "-C", os.path.join(REPO, "oracle"), "ref"])
  out = {}
  inputs, cases = balance_inputs()
  for name, (n, rows, cols, vals) in inputs.items():
    out[f"in_{name}_n"] = np.int64(n)
    out[f"in_{name}_rows"] = np.asarray(rows, np.int32)
    out[f"in_{name}_cols"] = np.asarray(cols, np.int32)
    if vals is not None:
      out[f"in_{name}_vals"] = np.asarray(vals, np.float64)
  with tempfile.TemporaryDirectory() as td:
    for tag, name, grid, mode, dedup in cases:
      n, rows, cols, vals = inputs[name]
      mtx, bal, operm = (os.path.join(td, tag + x) for x in (".mtx", ".bal.mtx", ".operm"))
      write_mtx(mtx, n, rows, cols, vals)
      cmd = [analyze, "--matrix", mtx, "--" + ("symmetric" if mode == "symmetric" else "shared-perm"),
             "--fabx", str(grid), "--faby", str(grid), "--omatrix", bal, "--operm", operm]
      if dedup:
        cmd.append("--remove_dup")
      log = subprocess.run(cmd, check=True, capture_output=True, text=True).stdout
      heaviest = int(re.findall(r"Permuted Max loaded PE has (\d+) nz", log)[-1])
      out[f"{tag}_input"] = np.array(name)
      out[f"{tag}_grid"] = np.int64(grid)
      out[f"{tag}_mode"] = np.array(mode)
      out[f"{tag}_dedup"] = np.bool_(dedup)
      out[f"{tag}_operm"] = np.loadtxt(operm, dtype=np.int64, ndmin=1)
      out[f"{tag}_max"] = np.int64(heaviest)
      if tag == "dup_p6":
        out[f"{tag}_omatrix"] = np.frombuffer(open(bal, "rb").read(), dtype=np.uint8)
      print(f"balance {tag}: n={n} nnz={len(rows)} grid={grid} {mode} heaviest={heaviest}")
  out["cases"] = np.array([c[0] for c in cases])
  np.savez_compressed(os.path.join(HERE, "balance_cases.npz"), **out)


if __name__ == "__main__":
  if sys.argv[1:] == ["balance"]:
    balance_fixtures()
  else:
    main()
    balance_fixtures()
