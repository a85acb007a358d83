"""The drop-in command lines end to end on the GPU: run_bfs and run_graph500
(one launch per search and --batched) on Matrix Market files, checked
against the oracle through the CSV rows and trees they write.

Mirrors how the reference is driven (bfs/scripts/run_bfs.py:567-711,
run_graph500.py:410-465): same flags, same CSV columns, parents -1 when
unreached and parents[source] = source."""
import csv

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.io import mmwrite

from oracle import oracle as O
from ceregraphs_amd import gen_rmat, timing
from ceregraphs_amd import run_bfs as run_bfs_mod
from ceregraphs_amd import run_graph500 as run_graph500_mod

pytestmark = pytest.mark.gpu


def _graph(directed):
  abcd = (0.45, 0.25, 0.15, 0.15) if directed else gen_rmat.GRAPH500_ABCD
  n, ip, ii, op, oi = gen_rmat.rmat_csr(10, 8 if directed else 16, 5, *abcd[:3], directed=directed)
  return O.Graph(n, ip, ii, op, oi, not directed)


def _write_mtx(path, g):
  # M[dst][src] (graph_loader.py:76-90): row v holds v's in-neighbours
  M = sp.csr_matrix((np.ones(len(g.in_idx)), g.in_idx, g.in_ptr), shape=(g.n, g.n))
  mmwrite(str(path), M.tocoo(), field="real")


def _read_rows(path):
  with open(path, newline="", encoding="utf-8") as f:
    return list(csv.DictReader(f))


@pytest.mark.parametrize("directed", [False, True])
def test_run_bfs_cli_matches_oracle(tmp_path, directed):
  g = _graph(directed)
  mtx = tmp_path / "g.mtx"
  _write_mtx(mtx, g)
  source = int(np.argmax(np.diff(g.out_ptr)))
  out_csv, tree = tmp_path / "t.csv", tmp_path / "tree.npz"
  rc = run_bfs_mod.main(["--infile_mtx", str(mtx), "--num_pe_cols", "4", "--num_pe_rows", "4",
                         "--source", str(source), "--csv", str(out_csv), "--out-tree", str(tree),
                         "--no-cache", "--fabric-dims", "8,6", "--channels", "2"])
  assert rc == 0
  lev, par, hist, rounds = O.bfs(g, source)
  t = np.load(tree)
  np.testing.assert_array_equal(t["levels"], lev)
  np.testing.assert_array_equal(t["parents"], par)
  with open(out_csv, newline="", encoding="utf-8") as f:
    header = next(csv.reader(f))
  assert header == timing.BFS_TIMING_COLUMNS
  row = _read_rows(out_csv)[0]
  assert int(row["rounds_completed"]) == rounds
  prof = min(rounds, 10)  # --max-rounds default
  assert row["nf_history"] == ";".join(str(x) for x in hist[:prof])
  assert int(row["visited_count"]) == int((lev >= 0).sum())
  assert int(row["m_edges_traversed"]) == O.m_edges(g, lev)
  assert row["m_convention"] == ("directed_source_visited" if directed else "undirected_dedup")
  assert float(row["gteps"]) > 0 and float(row["gteps_no_transfer"]) >= float(row["gteps"])


def test_run_graph500_cli_batched_equals_per_search(tmp_path):
  g = _graph(False)
  mtx = tmp_path / "g.mtx"
  _write_mtx(mtx, g)
  common = ["--infile_mtx", str(mtx), "--num_pe_cols", "2", "--num_pe_rows", "2",
            "--num-searches", "12", "--seed", "3", "--direction", "beamer"]
  rows = {}
  for mode in ("single", "batched"):
    c, s = tmp_path / f"{mode}.csv", tmp_path / f"{mode}_sum.csv"
    argv = common + ["--csv", str(c), "--summary-csv", str(s)] + (["--batched"] if mode == "batched" else [])
    assert run_graph500_mod.main(argv) == 0
    rows[mode] = _read_rows(c)
    summary = _read_rows(s)[0]
    assert int(summary["num_searches"]) == 12 and int(summary["num_correctness_fail"]) == 0
    assert float(summary["harmonic_mean_gteps"]) > 0
  for a, b in zip(rows["single"], rows["batched"]):
    assert a["source"] == b["source"] and a["correctness_ok"] == b["correctness_ok"] == "True"
    s = int(a["source"])
    lev, _, _, rounds = O.bfs(g, s)
    for r in (a, b):
      assert int(r["visited_count"]) == int((lev >= 0).sum())
      assert int(r["m_edges_traversed"]) == O.m_edges(g, lev)
      assert int(r["rounds_completed"]) == rounds
