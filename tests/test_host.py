"""Host-side logic of the drop-in surface (no GPU): ingest orientation and
caching, .operm translation, GTEPS accounting, root picking, CSV schema and
the CLI's correctness cross-check."""
import csv
import gzip
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.io import mmwrite

from conftest import golden
from oracle import oracle as O
from ceregraphs_amd import graph_loader, timing
from ceregraphs_amd import run_bfs as run_bfs_mod
from ceregraphs_amd import run_graph500 as run_graph500_mod


def test_edgelist_orientation_and_declared_nodes(tmp_path):
  """'src dst' lines become M[dst][src] (graph_loader.py:76-90); a trailing
  isolated vertex survives through '# Nodes:'."""
  p = tmp_path / "g.txt.gz"
  with gzip.open(p, "wt", encoding="utf-8") as f:
    f.write("# Directed graph\n# Nodes: 6 Edges: 3\n0\t1\n1 2\n\n3 1 extra\n")
  hg = graph_loader.load_csr(str(p))
  assert hg.n == 6 and not hg.symmetric
  M = sp.csr_matrix((np.ones(hg.nnz), hg.in_idx, hg.in_ptr), shape=(6, 6))
  assert M[1, 0] == 1 and M[2, 1] == 1 and M[1, 3] == 1 and M.nnz == 3
  # out view: column 1 of M holds the out-neighbours of 1
  assert list(hg.out_idx[hg.out_ptr[1]:hg.out_ptr[2]]) == [2]
  g = O.Graph(hg.n, hg.in_ptr, hg.in_idx, hg.out_ptr, hg.out_idx, False)
  lev, _, _, _ = O.bfs(g, 0)
  assert list(lev) == [0, 1, 2, -1, -1, -1]


def test_mtx_roundtrip_symmetric_storage_and_duplicates(tmp_path):
  raw = golden("rmat_s10_ef16_seed0.npz")
  A = sp.csr_matrix((np.ones(len(raw["indices"])), raw["indices"], raw["indptr"]), shape=(1024, 1024))
  p = tmp_path / "a.mtx"
  mmwrite(str(p), A, field="real")
  hg = graph_loader.load_csr(str(p))
  assert hg.symmetric and hg.in_ptr is hg.out_ptr
  np.testing.assert_array_equal(hg.in_ptr, raw["indptr"])
  np.testing.assert_array_equal(hg.in_idx, raw["indices"])
  # a small matrix: scipy writes lower-triangle 'symmetric' storage, which
  # mmread expands (the reference's own C reader would not)
  S = sp.csr_matrix(np.array([[0, 1, 1], [1, 0, 0], [1, 0, 0]], dtype=float))
  q = tmp_path / "s.mtx"
  mmwrite(str(q), S, field="real")
  assert "symmetric" in open(q, encoding="utf-8").readline()
  hs = graph_loader.load_csr(str(q))
  assert hs.nnz == 4 and hs.symmetric
  # duplicates merge into one structural entry
  D = sp.coo_matrix((np.ones(3), ([1, 1, 2], [0, 0, 1])), shape=(3, 3))
  hd = graph_loader.views_from_matrix(D)
  assert hd.nnz == 2


def test_binary_cache(tmp_path):
  p = tmp_path / "c.mtx"
  S = sp.random(50, 50, density=0.1, random_state=3, format="csr")
  mmwrite(str(p), S, field="real")
  a = graph_loader.load_csr_cached(str(p))
  assert os.path.exists(str(p) + graph_loader.CACHE_SUFFIX)
  b = graph_loader.load_csr_cached(str(p))
  for x, y in zip((a.in_ptr, a.in_idx, a.out_ptr, a.out_idx), (b.in_ptr, b.in_idx, b.out_ptr, b.out_idx)):
    np.testing.assert_array_equal(x, y)
  assert a.symmetric == b.symmetric


def test_operm_translation(tmp_path):
  g = golden("c1_balanced_4x4.npz")
  p = tmp_path / "x.operm"
  np.savetxt(p, g["operm"], fmt="%d")
  perm = graph_loader.read_operm(str(p))
  src_orig = int(np.nonzero(perm == 0)[0][0])
  assert graph_loader.balanced_source(perm, src_orig) == 0
  with pytest.raises(ValueError):
    graph_loader.balanced_source(perm, len(perm))


def test_m_edges_and_gteps_closed_forms(c1_graph):
  lev, _, _, _ = O.bfs(c1_graph, 0)
  rows, cols = c1_graph.coo_rows_cols()
  m, conv, secs, g = timing.compute_m_and_gteps(rows, cols, lev >= 0, True, 2_000_000)
  assert conv == "undirected_dedup"
  assert m == O.m_edges(c1_graph, lev)
  assert m * 2 == int(np.diff(c1_graph.in_ptr)[lev >= 0].sum())  # loop-free symmetric
  assert secs == pytest.approx(2e-3) and g == pytest.approx(m / 2e-3 / 1e9)
  m2, conv2, _, _ = timing.compute_m_and_gteps(rows, cols, lev >= 0, False, 1)
  assert conv2 == "directed_source_visited" and m2 == int(np.diff(c1_graph.out_ptr)[lev >= 0].sum())


def test_harmonic_mean_and_pick_sources():
  assert timing.harmonic_mean([1.0, 4.0, 4.0]) == pytest.approx(2.0)
  assert timing.harmonic_mean([2.0, float("nan"), 0.0, -1.0]) == pytest.approx(2.0)
  assert np.isnan(timing.harmonic_mean([]))
  deg = np.array([0, 3, 0, 1, 5, 0, 2])
  picked = timing.pick_sources(deg, 3, seed=7)
  assert picked == np.random.default_rng(7).choice(np.array([1, 3, 4, 6]), 3, replace=False).tolist()
  assert timing.pick_sources(deg, 10, seed=0) and len(timing.pick_sources(deg, 10, seed=0)) == 4
  assert timing.pick_sources(deg, 3, 0, sources="4,1") == [4, 1]
  with pytest.raises(ValueError):
    timing.pick_sources(deg, 3, 0, sources="9")


def test_out_degree_excludes_self_loops():
  M = sp.coo_matrix((np.ones(4), ([0, 1, 2, 2], [0, 0, 0, 2])), shape=(3, 3))
  hg = graph_loader.views_from_matrix(M)
  np.testing.assert_array_equal(hg.out_degree_no_loops(), [2, 0, 0])


def test_csv_schema_and_header_guard(tmp_path):
  assert len(timing.BFS_TIMING_COLUMNS) == 37
  assert timing.BFS_TIMING_COLUMNS[0] == "timestamp" and timing.BFS_TIMING_COLUMNS[-1] == "gteps_no_transfer"
  p = tmp_path / "t.csv"
  row = {c: i for i, c in enumerate(timing.BFS_TIMING_COLUMNS)}
  timing.append_csv_row(str(p), row, timing.BFS_TIMING_COLUMNS)
  timing.append_csv_row(str(p), row, timing.BFS_TIMING_COLUMNS)
  with open(p, newline="", encoding="utf-8") as f:
    rows = list(csv.reader(f))
  assert rows[0] == timing.BFS_TIMING_COLUMNS and len(rows) == 3
  with pytest.raises(ValueError):
    timing.append_csv_row(str(p), {"a": 1}, ["a"])
  assert timing.GRAPH500_SEARCH_COLUMNS[:3] == ["timestamp", "infile_mtx", "n"]


def test_cli_cross_check_accepts_oracle_and_flags_bad_parent(c1_graph):
  hg = graph_loader.HostGraph(c1_graph.n, c1_graph.in_ptr, c1_graph.in_idx, c1_graph.out_ptr,
                              c1_graph.out_idx, True)
  lev, par, _, _ = O.bfs(c1_graph, 0)
  mismatch, bad, diff = run_bfs_mod.check_search(hg, 0, par, lev)
  assert mismatch == 0 and bad == [] and diff >= 0
  v = int(np.nonzero(lev == 2)[0][0])
  bad_par = par.copy()
  bad_par[v] = 0  # the root is two levels up
  _, bad, _ = run_bfs_mod.check_search(hg, 0, bad_par, lev)
  assert bad == [v]


def test_cli_argument_surface():
  a = run_bfs_mod.parse_args(["--infile_mtx", "x.mtx", "--num_pe_cols", "4", "--num_pe_rows", "4",
                              "--source", "5", "--directional", "--fabric-dims", "8,6",
                              "--channels", "4", "--parent-resolve-variant", "indexed"])
  assert run_bfs_mod.resolve_direction(a, 1024) == ("tau", 154)
  b = run_bfs_mod.parse_args(["--infile_mtx", "x.mtx", "--num_pe_cols", "2", "--num_pe_rows", "2"])
  assert run_bfs_mod.resolve_direction(b, 100) == ("topdown", None)
  c = run_graph500_mod.parse_args(["--infile_mtx", "x.mtx", "--num_pe_cols", "2", "--num_pe_rows", "2",
                                   "--sources", "1,2", "--batched", "--direction", "beamer"])
  assert c.batched and c.direction == "beamer" and c.max_rounds == 15
