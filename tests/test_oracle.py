"""Pins the CPU oracle (oracle/) and the host generator before anything is
compared against them: golden vectors written from the reference, the
reference's own scipy oracle, brute force on tiny graphs, closed forms, and
planted mistakes the validator must catch."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import golden
from oracle import oracle as O
from ceregraphs_amd import gen_rmat


def _graph(edges, n, directed=True):
  """edges: list of (src, dst); M[dst][src] = 1 (graph_loader.py:76-90)."""
  src = np.array([e[0] for e in edges], dtype=np.int64)
  dst = np.array([e[1] for e in edges], dtype=np.int64)
  if not directed:
    src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
  M = sp.coo_matrix((np.ones(len(src)), (dst, src)), shape=(n, n))
  return O.Graph.from_matrix(M)


# -- the generator, pinned by the reference's own output ----------------------
@pytest.mark.parametrize("scale,ef,seed", [(10, 16, 0), (8, 4, 5), (12, 8, 3)])
def test_generator_matches_reference_vectors(scale, ef, seed):
  g = golden(f"rmat_s{scale}_ef{ef}_seed{seed}.npz")
  n, in_ptr, in_idx, out_ptr, out_idx = gen_rmat.rmat_csr(scale, ef, seed)
  assert n == int(g["n"])
  np.testing.assert_array_equal(in_ptr, g["indptr"])
  np.testing.assert_array_equal(in_idx, g["indices"])
  assert in_ptr is out_ptr  # symmetric: one shared view


def test_directed_generator_symmetrises_to_reference():
  """The directed variant draws the same endpoints; symmetrising it must give
  back the golden-pinned symmetric graph."""
  n, ip, ii, op, oi = gen_rmat.rmat_csr(10, 16, 0, directed=True)
  M = sp.csr_matrix((np.ones(len(ii)), ii, ip), shape=(n, n))
  S = ((M + M.T) > 0).astype(np.int8).tocsr()
  S.sort_indices()
  g = golden("rmat_s10_ef16_seed0.npz")
  np.testing.assert_array_equal(S.indptr, g["indptr"])
  np.testing.assert_array_equal(S.indices, g["indices"])
  # the two directed views are transposes of each other
  Mc = sp.csc_matrix((np.ones(len(oi)), oi, op), shape=(n, n))
  assert (M != Mc).nnz == 0
  assert np.all(M.diagonal() == 0)


def test_generator_invariants_directed_abcd():
  n, ip, ii, op, oi = gen_rmat.rmat_csr(9, 8, 11, 0.45, 0.25, 0.15, directed=True)
  assert n == 512
  assert np.all(np.diff(ip) >= 0) and ip[-1] == len(ii) <= n * 8
  rows = np.repeat(np.arange(n), np.diff(ip))
  assert np.all(rows != ii)
  for r in range(n):  # ascending, unique within each row
    seg = ii[ip[r]:ip[r + 1]]
    assert np.all(np.diff(seg) > 0)


def test_pcg_words_reproduce_numpy_stream():
  """Restates PCG64 (XSL-RR 128/64) from the four words and compares the
  first doubles with numpy's own draws."""
  w = [int(x) for x in gen_rmat.pcg_words(123)]
  s, inc = (w[0] << 64) | w[1], (w[2] << 64) | w[3]
  mult = 0x2360ED051FC65DA44385DF649FCCF645
  m128, m64 = (1 << 128) - 1, (1 << 64) - 1
  out = []
  for _ in range(8):
    s = (s * mult + inc) & m128
    x = ((s >> 64) ^ s) & m64
    rot = s >> 122
    o = ((x >> rot) | (x << ((64 - rot) & 63))) & m64
    out.append((o >> 11) * (1.0 / 9007199254740992.0))
  np.testing.assert_array_equal(np.array(out), np.random.default_rng(123).random(8))


# -- configuration 1 fixture ----------------------------------------------------
def test_c1_fixture_is_the_relabelled_graph(c1_graph):
  """The analyze output is the s10 graph under one shared permutation."""
  g = golden("c1_balanced_4x4.npz")
  perm = g["operm"].astype(np.int64)
  n = int(g["n"])
  assert sorted(perm.tolist()) == list(range(n))
  raw = golden("rmat_s10_ef16_seed0.npz")
  rows = np.repeat(np.arange(n), np.diff(raw["indptr"]))
  M = sp.coo_matrix((np.ones(len(rows)), (perm[rows], perm[raw["indices"]])), shape=(n, n)).tocsr()
  M.sort_indices()
  np.testing.assert_array_equal(M.indptr, g["indptr"])
  np.testing.assert_array_equal(M.indices, g["indices"])
  assert str(g["header"]) == "%%MatrixMarket matrix coordinate real general"


def test_c1_oracle_matches_reference_scipy(c1_graph):
  s = golden("c1_scipy_src0.npz")
  lev, par, hist, rounds = O.bfs(c1_graph, 0)
  np.testing.assert_array_equal(lev, s["levels"])
  np.testing.assert_array_equal(lev >= 0, s["visited"])
  assert rounds == int(s["levels"].max()) + 1
  np.testing.assert_array_equal(hist, [np.sum(s["levels"] == r + 1) for r in range(rounds)])
  assert O.validate(c1_graph, 0, par, lev) == []


def test_levels_are_permutation_covariant(c1_graph):
  g = golden("c1_balanced_4x4.npz")
  perm = g["operm"].astype(np.int64)
  raw = golden("rmat_s10_ef16_seed0.npz")
  G0 = O.Graph(1024, raw["indptr"], raw["indices"], raw["indptr"], raw["indices"], True)
  src_orig = int(np.nonzero(perm == 0)[0][0])
  lev0, _, _, _ = O.bfs(G0, src_orig)
  lev1, _, _, _ = O.bfs(c1_graph, 0)
  np.testing.assert_array_equal(lev1[perm], lev0)


# -- brute force and closed forms -------------------------------------------------
def _brute(n, edges, s):
  """Levels by boolean matrix powers, parents by the definition."""
  A = np.zeros((n, n), dtype=np.int64)
  for a, b in edges:
    A[a, b] = 1
  reach = np.zeros(n, dtype=np.int64)
  reach[s] = 1
  lev = np.full(n, -1)
  lev[s] = 0
  frontier = reach.copy()
  for k in range(1, n + 1):
    frontier = ((frontier @ A) > 0).astype(np.int64)
    new = (frontier > 0) & (lev < 0)
    lev[new] = k
  par = np.full(n, -1)
  par[s] = s
  for v in range(n):
    if lev[v] > 0:
      par[v] = min(u for u in range(n) if A[u, v] and lev[u] == lev[v] - 1)
  return lev, par


@pytest.mark.parametrize("seed", range(40))
def test_oracle_brute_force_tiny(seed):
  rng = np.random.default_rng(seed)
  n = int(rng.integers(1, 9))
  edges = sorted({(int(a), int(b)) for a, b in rng.integers(0, n, size=(int(rng.integers(0, 20)), 2))})
  g = _graph(edges, n)
  s = int(rng.integers(0, n))
  lev, par, hist, rounds = O.bfs(g, s)
  blev, bpar = _brute(n, edges, s)
  np.testing.assert_array_equal(lev, blev)
  np.testing.assert_array_equal(par, bpar)
  assert rounds == blev.max() + 1
  assert O.validate(g, s, par, lev) == []
  np.testing.assert_array_equal(O.lowest_parents(g, lev, s), bpar)
  np.testing.assert_array_equal(O.scipy_levels(g, s), blev)
  slev, nf, _, r2 = O.spmv_bfs(g, s)
  np.testing.assert_array_equal(slev, blev)
  np.testing.assert_array_equal(nf, hist)
  assert r2 == rounds


def test_closed_forms():
  n = 50
  path = _graph([(i, i + 1) for i in range(n - 1)], n)
  lev, par, hist, rounds = O.bfs(path, 0)
  np.testing.assert_array_equal(lev, np.arange(n))
  np.testing.assert_array_equal(par, np.concatenate([[0], np.arange(n - 1)]))
  assert rounds == n and list(hist) == [1] * (n - 1) + [0]
  # directed path from the far end reaches nothing
  lev, par, hist, rounds = O.bfs(path, n - 1)
  assert rounds == 1 and list(hist) == [0] and lev.sum() == -(n - 1)
  # star: every leaf is level 1 with parent 0; leaves reach each other via 0
  star = _graph([(0, i) for i in range(1, 20)], 20, directed=False)
  lev, par, _, _ = O.bfs(star, 5)
  assert lev[0] == 1 and par[0] == 5 and set(lev[1:]) == {0, 2}
  assert all(par[v] == 0 for v in range(1, 20) if v != 5)
  # two components + a source self-loop
  g = _graph([(0, 0), (0, 1), (1, 2), (3, 4)], 5)
  lev, par, _, _ = O.bfs(g, 0)
  np.testing.assert_array_equal(lev, [0, 1, 2, -1, -1])
  np.testing.assert_array_equal(par, [0, 0, 1, -1, -1])


def test_m_edges_closed_form():
  raw = golden("rmat_s10_ef16_seed0.npz")
  G = O.Graph(1024, raw["indptr"], raw["indices"], raw["indptr"], raw["indices"], True)
  lev, _, _, _ = O.bfs(G, 1)
  deg = np.diff(G.in_ptr)
  # loop-free symmetric: each undirected edge of the component counted once
  assert O.m_edges(G, lev) * 2 == int(deg[lev >= 0].sum())
  n, ip, ii, op, oi = gen_rmat.rmat_csr(9, 8, 2, directed=True)
  D = O.Graph(n, ip, ii, op, oi, False)
  lev, _, _, _ = O.bfs(D, int(np.argmax(np.diff(op))))
  assert O.m_edges(D, lev) == int(np.diff(op)[lev >= 0].sum())


def test_validator_catches_planted_mistakes(c1_graph):
  lev, par, _, _ = O.bfs(c1_graph, 0)
  assert O.validate(c1_graph, 0, par, lev) == []
  # a valid but not lowest-id predecessor
  rows, cols = c1_graph.coo_rows_cols()
  cand = [(r, c) for r, c in zip(rows, cols)
          if lev[r] > 0 and lev[c] == lev[r] - 1 and c > par[r]]
  r, c = cand[0]
  bad = par.copy()
  bad[r] = c
  assert "parents differ from the lowest-id rule" in O.validate(c1_graph, 0, bad, lev)
  # a wrong level
  bad_lev = lev.copy()
  bad_lev[r] += 1
  assert O.validate(c1_graph, 0, par, bad_lev)
  # a parent with no edge
  nonadj = [u for u in range(c1_graph.n) if lev[u] == lev[r] - 1 and
            u not in set(c1_graph.in_idx[c1_graph.in_ptr[r]:c1_graph.in_ptr[r + 1]])]
  bad = par.copy()
  bad[r] = nonadj[0]
  assert "tree edge missing from the graph" in O.validate(c1_graph, 0, bad, lev)
  # a dropped vertex
  v = int(np.nonzero(lev == 2)[0][0])
  bad_lev, bad = lev.copy(), par.copy()
  bad_lev[v], bad[v] = -1, -1
  assert O.validate(c1_graph, 0, bad, bad_lev)


def test_spmv_bfs_tau_direction_history(c1_graph):
  lev, nf, dirs, r = O.spmv_bfs(c1_graph, 0, tau=round(0.15 * c1_graph.n))
  # nf = 471 in round 0 exceeds tau = 154, so rounds 1.. run bottom-up
  assert list(nf) == [471, 415, 6, 0] and list(dirs) == [0, 1, 1, 1]
  lev2, _, _, _ = O.bfs(c1_graph, 0)
  np.testing.assert_array_equal(lev, lev2)


def test_scipy_oracle_random_directed():
  for seed in range(3):
    n, ip, ii, op, oi = gen_rmat.rmat_csr(11, 4, seed, 0.45, 0.25, 0.15, directed=True)
    g = O.Graph(n, ip, ii, op, oi, False)
    for s in [0, 7, int(np.argmax(np.diff(op)))]:
      lev, par, _, _ = O.bfs(g, s)
      np.testing.assert_array_equal(lev, O.scipy_levels(g, s))
      assert O.validate(g, s, par, lev) == []


def test_itertools_exhaustive_3_vertices():
  """Every directed graph on 3 vertices, every source."""
  pairs = [(a, b) for a in range(3) for b in range(3)]
  for mask in range(1 << len(pairs)):
    edges = [p for i, p in enumerate(pairs) if mask >> i & 1]
    g = _graph(edges, 3)
    for s in range(3):
      lev, par, _, _ = O.bfs(g, s)
      blev, bpar = _brute(3, edges, s)
      assert list(lev) == list(blev) and list(par) == list(bpar)
  assert itertools is not None
