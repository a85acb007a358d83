"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bar: bit-exact levels and parents (lowest-id rule), identical per-round
discovery counts and round counts, for every direction policy; MS-BFS lanes
identical to single-root searches.  Full-size graphs (s21, s24) are checked
on sampled roots against the oracle and through size-independent properties
(lane == single-root, structural validation on sampled lanes)."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import golden
from oracle import oracle as O
from ceregraphs_amd import gen_rmat
from ceregraphs_amd.engine import DeviceGraph

pytestmark = pytest.mark.gpu

DIRECTIONS = ["beamer", "topdown", "bottomup", "tau"]


def _device(g):
  return DeviceGraph.from_csr(g.n, g.out_ptr, g.out_idx, g.in_ptr, g.in_idx, g.symmetric)


def _edges_graph(edges, n, directed=True):
  src = np.array([e[0] for e in edges], dtype=np.int64)
  dst = np.array([e[1] for e in edges], dtype=np.int64)
  if not directed:
    src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
  M = sp.coo_matrix((np.ones(len(src)), (dst, src)), shape=(n, n))
  return O.Graph.from_matrix(M)


def _check_single(dg, g, s, direction, **kw):
  r = dg.bfs(s, direction=direction, **kw)
  lev, par, hist, rounds = O.bfs(g, s)
  np.testing.assert_array_equal(r.levels, lev, err_msg=f"levels src={s} dir={direction}")
  np.testing.assert_array_equal(r.parents, par, err_msg=f"parents src={s} dir={direction}")
  assert r.stats.rounds_completed == rounds
  np.testing.assert_array_equal(r.stats.nf_history, hist)
  assert r.stats.m_edges == O.m_edges(g, lev)
  assert r.stats.visited_count == int((lev >= 0).sum())
  return r


def test_c1_plumbing(c1_graph):
  """Configuration 1 (s10, 4x4-balanced, source 0) against the reference's
  own scipy result and the lowest-id rule."""
  s = golden("c1_scipy_src0.npz")
  with _device(c1_graph) as dg:
    for d in DIRECTIONS:
      r = _check_single(dg, c1_graph, 0, d)
      np.testing.assert_array_equal(r.levels, s["levels"])
      assert O.validate(c1_graph, 0, r.parents, r.levels) == []
    r = dg.bfs(0, direction="tau")
    # nf = 471 > tau = round(0.15 * 1024) after round 0 (bool_pe.csl:1078-1088)
    assert list(r.stats.direction_history) == [0, 1, 1, 1]


@pytest.mark.parametrize("scale,ef,seed,directed", [
    (10, 16, 0, False), (12, 8, 3, False), (14, 16, 1, False), (13, 8, 10, True),
    (15, 16, 5, True)])
def test_gpu_generator_matches_numpy(scale, ef, seed, directed):
  abcd = (0.45, 0.25, 0.15, 0.15) if directed else gen_rmat.GRAPH500_ABCD
  n, ip, ii, op, oi = gen_rmat.rmat_csr(scale, ef, seed, *abcd[:3], directed=directed)
  with DeviceGraph.from_rmat(scale, ef, seed, abcd=abcd, directed=directed) as dg:
    assert dg.n == n and dg.nnz == len(ii) and dg.symmetric == (not directed)
    gip, gii, gop, goi = dg.download()
    np.testing.assert_array_equal(gip, ip)
    np.testing.assert_array_equal(gii, ii)
    np.testing.assert_array_equal(gop, op)
    np.testing.assert_array_equal(goi, oi)


@pytest.mark.parametrize("scale,ef,seed,directed", [
    (12, 16, 1, False), (14, 8, 2, False), (12, 8, 10, True), (14, 32, 15, True)])
def test_single_root_parity_all_directions(scale, ef, seed, directed):
  abcd = (0.45, 0.25, 0.15, 0.15) if directed else gen_rmat.GRAPH500_ABCD
  n, ip, ii, op, oi = gen_rmat.rmat_csr(scale, ef, seed, *abcd[:3], directed=directed)
  g = O.Graph(n, ip, ii, op, oi, not directed)
  deg = np.diff(g.out_ptr)
  roots = [int(np.argmax(deg)), int(np.nonzero(deg)[0][0]), int(np.nonzero(deg == 0)[0][0]),
           *np.random.default_rng(seed).choice(np.nonzero(deg)[0], 3, replace=False).tolist()]
  with DeviceGraph.from_rmat(scale, ef, seed, abcd=abcd, directed=directed) as dg:
    for s in roots:
      for d in DIRECTIONS:
        _check_single(dg, g, s, d)
      _check_single(dg, g, s, "beamer", alpha=2.0, beta=100.0)


def test_edge_cases():
  """Self-loop at the source, two components, isolated source, a deep path
  (many rounds: exercises the chunked round loop), a star whose hub has more
  than 8192 neighbours (whole-grid push), n not a multiple of 64."""
  cases = []
  cases.append((_edges_graph([(0, 0), (0, 1), (1, 2), (3, 4)], 5), [0, 3, 2]))
  n = 3000
  cases.append((_edges_graph([(i, i + 1) for i in range(n - 1)], n, directed=False), [0, 1500, n - 1]))
  cases.append((_edges_graph([(0, i) for i in range(1, 20001)] + [(20000, 20001)], 20003,
                             directed=False), [0, 7, 20001, 20002]))
  cases.append((_edges_graph([(i, (i * 7 + 3) % 777) for i in range(777)], 777), [0, 5, 776]))
  # 5000 frontier vertices of degree >= 300 in one push round: the flat
  # whole-grid expansion stages its degree prefix in passes of 4096
  rng = np.random.default_rng(5)
  hubs, pool = 5000, 20000
  e = [(0, h) for h in range(1, hubs + 1)]
  leaves = rng.integers(hubs + 1, hubs + 1 + pool, size=(hubs, 300))
  e += [(h, int(x)) for h in range(1, hubs + 1) for x in leaves[h - 1]]
  cases.append((_edges_graph(e, hubs + 1 + pool, directed=False), [0, 17]))
  for g, roots in cases:
    with _device(g) as dg:
      for s in roots:
        for d in DIRECTIONS:
          r = _check_single(dg, g, s, d)
          assert O.validate(g, s, r.parents, r.levels) == []


def test_device_outputs_and_repeat_determinism():
  import torch
  n, ip, ii, op, oi = gen_rmat.rmat_csr(13, 16, 4)
  g = O.Graph(n, ip, ii, op, oi, True)
  with DeviceGraph.from_rmat(13, 16, 4) as dg:
    par = torch.empty(n, dtype=torch.int32, device="cuda")
    lev = torch.empty(n, dtype=torch.int32, device="cuda")
    lev0, par0, _, _ = O.bfs(g, 9)
    for _ in range(3):
      dg.bfs(9, device_out=(par, lev))
      np.testing.assert_array_equal(lev.cpu().numpy(), lev0)
      np.testing.assert_array_equal(par.cpu().numpy(), par0)


@pytest.mark.parametrize("scale,ef,seed,directed,nroots", [
    (12, 16, 1, False, 64), (13, 8, 2, False, 7), (12, 16, 12, True, 64), (11, 8, 3, False, 1)])
def test_msbfs_lanes_equal_oracle(scale, ef, seed, directed, nroots):
  abcd = (0.45, 0.25, 0.15, 0.15) if directed else gen_rmat.GRAPH500_ABCD
  n, ip, ii, op, oi = gen_rmat.rmat_csr(scale, ef, seed, *abcd[:3], directed=directed)
  g = O.Graph(n, ip, ii, op, oi, not directed)
  deg = np.diff(op)
  rng = np.random.default_rng(seed + 100)
  roots = rng.choice(n, nroots, replace=False)
  roots[0] = int(np.argmax(deg))
  if nroots > 2:
    roots[1] = roots[2]  # a repeated root
  oracle_levels = np.stack([O.bfs(g, int(s))[0] for s in roots])
  with DeviceGraph.from_rmat(scale, ef, seed, abcd=abcd, directed=directed) as dg:
    for d, cap in [(d, False) for d in DIRECTIONS + ["beamer_lanes"]] + [("beamer", True),
                                                                       ("beamer_lanes", True)]:
      r = dg.msbfs(roots, direction=d, capture=cap)
      for l, s in enumerate(roots):
        lev, par, _, rounds = O.bfs(g, int(s))
        np.testing.assert_array_equal(r.levels[l], lev.astype(np.int16),
                                      err_msg=f"lane {l} dir {d} capture {cap}")
        np.testing.assert_array_equal(r.parents[l], par, err_msg=f"lane {l} dir {d} capture {cap}")
        assert r.m_lanes[l] == O.m_edges(g, lev)
      _check_ms_stats(r.stats, oracle_levels, d)


@pytest.mark.parametrize("scale,ef,seed,directed", [(12, 16, 1, False), (12, 16, 12, True)])
def test_msbfs_every_output_written(scale, ef, seed, directed):
  """Every (lane, vertex) output is written by the search itself: device
  outputs pre-filled with a sentinel come back equal to the oracle under
  every direction policy, with and without graph replay.  Vertices without
  in-edges are written by the first pull round, or by the final pass when no
  round pulls (topdown); roots without in-edges keep their own entries."""
  import torch
  abcd = (0.45, 0.25, 0.15, 0.15) if directed else gen_rmat.GRAPH500_ABCD
  n, ip, ii, op, oi = gen_rmat.rmat_csr(scale, ef, seed, *abcd[:3], directed=directed)
  g = O.Graph(n, ip, ii, op, oi, not directed)
  outdeg, indeg = np.diff(op), np.diff(ip)
  noin_roots = np.nonzero((outdeg > 0) & (indeg == 0))[0]
  rng = np.random.default_rng(seed + 7)
  roots = rng.choice(np.nonzero(outdeg > 0)[0], 64, replace=False)
  roots[0] = int(np.argmax(outdeg))
  roots[1:1 + min(8, len(noin_roots))] = noin_roots[:8]
  assert not directed or len(noin_roots) > 0
  expect = [O.bfs(g, int(s)) for s in roots]
  par = torch.empty(64 * n, dtype=torch.int32, device="cuda")
  lev = torch.empty(64 * n, dtype=torch.int16, device="cuda")
  with DeviceGraph.from_rmat(scale, ef, seed, abcd=abcd, directed=directed) as dg:
    for d, cap in [(d, False) for d in DIRECTIONS + ["beamer_lanes"]] + [("beamer", True)]:
      par.fill_(0x5A5A5A5A)
      lev.fill_(0x5A5A)
      torch.cuda.synchronize()  # the fill is on torch's stream, the search on the library's
      dg.msbfs(roots, direction=d, capture=cap, device_out=(par, lev))
      P = par.view(64, n).cpu().numpy()
      L = lev.view(64, n).cpu().numpy()
      for l in range(64):
        lev0, par0, _, _ = expect[l]
        np.testing.assert_array_equal(L[l], lev0.astype(np.int16), err_msg=f"lane {l} dir {d} {cap}")
        np.testing.assert_array_equal(P[l], par0, err_msg=f"lane {l} dir {d} {cap}")


def test_msbfs_edge_cases():
  """Roots without in-edges, isolated roots, the same root in several lanes,
  a hub above the segment length, a deep path, n not a multiple of 64."""
  n = 5000
  e = [(i, i + 1) for i in range(1000)]                 # a path: 1000 levels
  e += [(2000, j) for j in range(2001, 2700)]           # a hub of out-degree 699
  e += [(j, 2000) for j in range(2700, 3100)]           # ... and in-degree 400
  e += [(4000, 4001), (4001, 4002), (4002, 4000)]       # a directed cycle
  e += [(4500, 4501)]                                   # 4500: out-edges only
  g = _edges_graph(e, n)
  roots = np.array([0, 4500, 4999, 2000, 2700, 4000, 0, 4999, 999, 3000, 4500])
  levs = np.stack([O.bfs(g, int(s))[0] for s in roots])
  with _device(g) as dg:
    for d, cap in [(d, False) for d in DIRECTIONS + ["beamer_lanes"]] + [("beamer", True),
                                                                       ("tau", True)]:
      r = dg.msbfs(roots, direction=d, capture=cap)
      for l, s in enumerate(roots):
        lev, par, _, _ = O.bfs(g, int(s))
        np.testing.assert_array_equal(r.levels[l], lev.astype(np.int16),
                                      err_msg=f"lane {l} dir {d} capture {cap}")
        np.testing.assert_array_equal(r.parents[l], par, err_msg=f"lane {l} dir {d} capture {cap}")
      _check_ms_stats(r.stats, levs, d)


def test_msbfs_hub_segments_lowest_id():
  """A hub with 20000 in-neighbours: 79 segments of 256 entries, the first
  32 in one list and 47 in the late one.  Lane l's root links to one to four
  of the hub's in-neighbours: lanes 0-15 only below id 8192 (the first 32
  segments), lanes 16-31 only above, lanes 32-63 anywhere, so the hub's
  parent in a pull round (the lowest of them) comes from either launch, and
  lanes found early must be dropped, not lost, in the later segments."""
  hub_deg = 20000
  hub = hub_deg + 64
  rng = np.random.default_rng(11)
  e = [(j, hub) for j in range(hub_deg)]
  roots = np.arange(hub_deg, hub_deg + 64)
  for l in range(64):
    k = 1 + l % 4
    lo, hi = (0, 8192) if l < 16 else (8192, hub_deg) if l < 32 else (0, hub_deg)
    e += [(int(roots[l]), int(p)) for p in rng.integers(lo, hi, size=k)]
  g = _edges_graph(e, hub + 1, directed=False)
  assert g.in_ptr[hub + 1] - g.in_ptr[hub] == hub_deg
  levs = np.stack([O.bfs(g, int(s))[0] for s in roots])
  assert np.all(levs[:, hub] == 2)
  with _device(g) as dg:
    for d, cap in [(d, False) for d in DIRECTIONS + ["beamer_lanes"]] + [("bottomup", True)]:
      r = dg.msbfs(roots, direction=d, capture=cap)
      for l, s in enumerate(roots):
        lev, par, _, _ = O.bfs(g, int(s))
        np.testing.assert_array_equal(r.levels[l], lev.astype(np.int16),
                                      err_msg=f"lane {l} dir {d} capture {cap}")
        np.testing.assert_array_equal(r.parents[l], par, err_msg=f"lane {l} dir {d} capture {cap}")
      _check_ms_stats(r.stats, levs, d)


def test_msbfs_push_side_table_full():
  """Push rounds whose new pairs do not all fit the side table that defers
  their parents: 64 roots each linked to the same 2000 vertices (every one
  new in all 64 lanes in round 0: 2000 x 66 words, more than the 131072 the
  table holds at this size), which in turn link to 4 of another 2000 (new
  in all lanes in round 1, several frontier in-neighbours each, so the
  lowest id decides).  Vertices whose entry does not fit take the direct
  path; the outputs are the oracle's either way, every round a push or
  not."""
  rng = np.random.default_rng(5)
  roots = np.arange(64)
  mid = np.arange(64, 2064)
  far = np.arange(2064, 4064)
  e = [(int(r), int(m)) for r in roots for m in mid]
  e += [(int(m), int(x)) for m in mid for x in rng.choice(far, 4, replace=False)]
  g = _edges_graph(e, 4064, directed=False)
  levs = np.stack([O.bfs(g, int(s))[0] for s in roots])
  with _device(g) as dg:
    for d, cap in [(d, False) for d in DIRECTIONS] + [("topdown", True), ("beamer", True)]:
      r = dg.msbfs(roots, direction=d, capture=cap)
      for l, s in enumerate(roots):
        lev, par, _, _ = O.bfs(g, int(s))
        np.testing.assert_array_equal(r.levels[l], lev.astype(np.int16),
                                      err_msg=f"lane {l} dir {d} capture {cap}")
        np.testing.assert_array_equal(r.parents[l], par, err_msg=f"lane {l} dir {d} capture {cap}")
      _check_ms_stats(r.stats, levs, d)


def _check_ms_stats(st, levs, direction):
  """Round accounting of a multi-source pass against the lanes' oracle
  levels: round r reaches the (lane, vertex) pairs at level r+1, nf counts
  vertices with at least one such lane, the last round finds nothing; the
  pull part of each round is a part of it; the union-count and fixed
  policies never mix directions within a round."""
  maxl = int(levs.max())
  assert st.rounds_completed == maxl + 1
  for r in range(maxl + 1):
    hit = levs == r + 1
    assert st.nf_history[r] == int(hit.any(axis=0).sum()), f"nf round {r}"
    assert st.pairs_found[r] == int(hit.sum()), f"pairs round {r}"
  assert np.all(st.pull_nf <= st.nf_history) and np.all(st.pull_pairs <= st.pairs_found)
  assert np.all(st.pull_edges <= st.edges_scanned) and np.all(st.pull_rows <= st.rows_scanned)
  assert set(st.direction_history.tolist()) <= {0, 1, 2}
  assert np.all((st.pull_lanes > 0) == (st.direction_history > 0))
  if direction in ("beamer", "topdown", "bottomup"):
    assert 2 not in st.direction_history  # these move every lane together
  if direction == "topdown":
    assert not st.pull_lanes.any()
  if direction == "bottomup":
    assert np.all(st.direction_history == 1)


@pytest.mark.timeout(600)
def test_c2_scale21_all_roots():
  """Configuration 2 (s21 ef16 seed 1, 64 roots from seed 2 among vertices
  with degree >= 1, one direction-optimising search per launch): all 64
  searches bit-exact against the oracle (levels, lowest-id parents, nf per
  round, m); the 64 roots in one multi-source pass equal them lane by lane."""
  with DeviceGraph.from_rmat(21, 16, 1) as dg:
    ip, ii, op, oi = dg.download()
    g = O.Graph(dg.n, ip, ii, op, oi, True)
    deg = np.diff(op)
    roots = np.random.default_rng(2).choice(np.nonzero(deg)[0], 64, replace=False)
    singles = [_check_single(dg, g, int(s), "beamer") for s in roots]
    _check_single(dg, g, int(roots[0]), "tau")
    r = dg.msbfs(roots, capture=True)
    for l, single in enumerate(singles):
      np.testing.assert_array_equal(r.levels[l], single.levels.astype(np.int16), err_msg=f"lane {l}")
      np.testing.assert_array_equal(r.parents[l], single.parents, err_msg=f"lane {l}")
      assert r.m_lanes[l] == single.stats.m_edges


@pytest.mark.timeout(600)
def test_c3_scale24_msbfs_lanes_equal_single_root():
  """Configuration 3 (s24 ef16 seed 3, 64 roots from seed 4): the pass the
  benchmark times (graph replay, every level decided on the device) and the
  polled, instrumented pass both give, lane by lane, exactly the 64
  single-root searches; two lanes are checked against the oracle in full."""
  import torch
  with DeviceGraph.from_rmat(24, 16, 3) as dg:
    n = dg.n
    deg = dg.out_degree()
    roots = np.random.default_rng(4).choice(np.nonzero(deg > 0)[0], 64, replace=False)
    outs = {}
    for cap in (True, False):
      par = torch.empty(64 * n, dtype=torch.int32, device="cuda")
      lev = torch.empty(64 * n, dtype=torch.int16, device="cuda")
      r = dg.msbfs(roots, device_out=(par, lev), capture=cap)
      outs[cap] = (par, lev, r.m_lanes)
    par2 = torch.empty(n, dtype=torch.int32, device="cuda")
    lev2 = torch.empty(n, dtype=torch.int32, device="cuda")
    for l in range(64):
      single = dg.bfs(int(roots[l]), device_out=(par2, lev2))
      for cap, (par, lev, m_lanes) in outs.items():
        assert torch.equal(lev[l * n:(l + 1) * n], lev2.to(torch.int16)), f"lane {l} capture {cap}"
        assert torch.equal(par[l * n:(l + 1) * n], par2), f"lane {l} capture {cap}"
        assert m_lanes[l] == single.stats.m_edges
    ip, ii, op, oi = dg.download()
    g = O.Graph(n, ip, ii, op, oi, True)
    par, lev, _ = outs[True]
    for l in (0, 63):
      lev0, par0, _, _ = O.bfs(g, int(roots[l]))
      np.testing.assert_array_equal(lev[l * n:(l + 1) * n].cpu().numpy(), lev0.astype(np.int16))
      np.testing.assert_array_equal(par[l * n:(l + 1) * n].cpu().numpy(), par0)


def test_graph_from_device_arrays():
  """A graph handed over as device arrays (torch tensors) is copied device to
  device, checked on the GPU, and searches exactly like the same graph from
  host arrays; the caller's tensors may be freed right after."""
  import torch
  for directed in (False, True):
    abcd = (0.45, 0.25, 0.15, 0.15) if directed else gen_rmat.GRAPH500_ABCD
    n, ip, ii, op, oi = gen_rmat.rmat_csr(12, 8, 7, *abcd[:3], directed=directed)
    g = O.Graph(n, ip, ii, op, oi, not directed)
    t = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (op, oi, ip, ii)]
    dg = DeviceGraph.from_csr(n, *t, symmetric=not directed)
    del t
    torch.cuda.empty_cache()
    with dg:
      gip, gii, gop, goi = dg.download()
      np.testing.assert_array_equal(gii, ii)
      np.testing.assert_array_equal(goi, oi)
      for s in (0, int(np.argmax(np.diff(op)))):
        _check_single(dg, g, s, "beamer")


@pytest.mark.parametrize("defect", ["id", "order", "ptr", "transpose"])
def test_graph_from_device_arrays_rejects_malformed(defect):
  """The device check catches each defect class before any search kernel
  reads the graph (CGBFS_EINVAL, no handle)."""
  import torch
  from ceregraphs_amd import _lib
  n, ip, ii, op, oi = gen_rmat.rmat_csr(11, 8, 9)
  ii = ii.copy()
  ip = ip.copy()
  v = int(np.argmax(np.diff(ip)))
  if defect == "id":
    ii[ip[v] + 1] = n + 5
  elif defect == "order":
    ii[ip[v]], ii[ip[v] + 1] = ii[ip[v] + 1], ii[ip[v]]
  elif defect == "ptr":
    ip[v + 1] = ip[v] - 1
  else:
    ii[ip[v + 1] - 1] = n - 1 if ii[ip[v + 1] - 1] != n - 1 else n - 2
    ii[ip[v]:ip[v + 1]].sort()
  t = [torch.from_numpy(a).cuda() for a in (ip, ii)]
  with pytest.raises(_lib.CgbfsError) as err:
    DeviceGraph.from_csr(n, *t, symmetric=True)
  assert err.value.code == _lib.CGBFS_EINVAL


def test_tuning_knobs_keep_results_and_are_checked():
  """The work-distribution knobs (chunks per grab, hub segments per grab,
  the push tier thresholds) only move work between wavefronts: any setting
  gives bit-exact results.  Out-of-range values get CGBFS_EINVAL."""
  from ceregraphs_amd import _lib
  n, ip, ii, op, oi = gen_rmat.rmat_csr(14, 16, 3)
  g = O.Graph(n, ip, ii, op, oi, True)
  deg = np.diff(op)
  with DeviceGraph.from_rmat(14, 16, 3) as dg:
    hub = int(np.argmax(deg))
    for kw in [dict(pull_grab=1), dict(pull_grab=64), dict(push_wave_degree=1, push_big_degree=2),
               dict(push_wave_degree=8, push_big_degree=64),
               dict(push_wave_degree=1 << 30, push_big_degree=1 << 30)]:
      for s in (hub, 7):
        for d in ("beamer", "topdown", "bottomup"):
          _check_single(dg, g, s, d, **kw)
    roots = np.random.default_rng(5).choice(np.nonzero(deg)[0], 64, replace=False)
    roots[0] = hub
    expect = [O.bfs(g, int(s)) for s in roots]
    for kw in [dict(pull_grab=3, seg_grab=1), dict(pull_grab=16, seg_grab=1024)]:
      for cap in (False, True):
        r = dg.msbfs(roots, capture=cap, **kw)
        for l in range(64):
          np.testing.assert_array_equal(r.levels[l], expect[l][0].astype(np.int16), err_msg=f"{kw} lane {l}")
          np.testing.assert_array_equal(r.parents[l], expect[l][1], err_msg=f"{kw} lane {l}")
    for bad in [dict(pull_grab=2000), dict(seg_grab=5000), dict(grid_mult=100),
                dict(push_wave_degree=64, push_big_degree=32)]:
      with pytest.raises(_lib.CgbfsError) as e:
        dg.bfs(hub, **bad)
      assert e.value.code == _lib.CGBFS_EINVAL, bad
      with pytest.raises(_lib.CgbfsError) as e:
        dg.msbfs(roots[:4], **bad)
      assert e.value.code == _lib.CGBFS_EINVAL, bad


@pytest.mark.parametrize("loop,capture", [("persistent", False), ("kernels", False),
                                          ("kernels", True)])
def test_level_loops_equal_oracle(loop, capture):
  """Both ways of running the levels -- one cooperative launch with grid
  barriers, or the fixed per-level launch sequence, enqueued or replayed
  from a captured HIP graph -- give the oracle's
  levels, lowest-id parents, per-round counts and m, on undirected and
  directed graphs, under every direction policy, on a graph with a
  1000-level path (1000 rounds of barriers) and with max_levels hit."""
  from ceregraphs_amd import _lib
  for scale, ef, seed, directed in [(10, 16, 0, False), (14, 8, 2, False), (13, 8, 10, True),
                                    (17, 16, 4, False)]:
    abcd = (0.45, 0.25, 0.15, 0.15) if directed else gen_rmat.GRAPH500_ABCD
    n, ip, ii, op, oi = gen_rmat.rmat_csr(scale, ef, seed, *abcd[:3], directed=directed)
    g = O.Graph(n, ip, ii, op, oi, not directed)
    deg = np.diff(op)
    with DeviceGraph.from_rmat(scale, ef, seed, abcd=abcd, directed=directed) as dg:
      for s in (int(np.argmax(deg)), int(np.nonzero(deg)[0][-1]), 0):
        for d in DIRECTIONS:
          _check_single(dg, g, s, d, loop=loop, capture=capture)
  edges = [(i, i + 1) for i in range(1000)] + [(500, j) for j in range(1001, 1400)]
  g = _edges_graph(edges, 1500, directed=False)
  with _device(g) as dg:
    for d in DIRECTIONS:
      _check_single(dg, g, 0, d, loop=loop, capture=capture)
    with pytest.raises(_lib.CgbfsError) as e:
      dg.bfs(0, loop=loop, max_levels=10, capture=capture)
    assert e.value.code == _lib.CGBFS_ETIMEOUT


def test_bfs_replay_reuse_and_rekey():
  """The captured level graph is kept on the handle and replayed for every
  source while the call's arguments match, and captured again when they
  change (other output buffers, host outputs, another direction): every
  search still equals the oracle, and so does a plain search between them."""
  import torch
  n, ip, ii, op, oi = gen_rmat.rmat_csr(15, 16, 6)
  g = O.Graph(n, ip, ii, op, oi, True)
  deg = np.diff(op)
  bufs = [(torch.empty(n, dtype=torch.int32, device="cuda"),
           torch.empty(n, dtype=torch.int32, device="cuda")) for _ in range(2)]
  sources = [int(np.argmax(deg))] + [int(x) for x in np.nonzero(deg)[0][[1, 7, 100, -1]]]
  with DeviceGraph.from_rmat(15, 16, 6) as dg:
    for i, s in enumerate(sources * 2):
      lev0, par0, hist, rounds = O.bfs(g, s)
      for out in (bufs[i % 2], None):
        for d in ("beamer", "topdown") if i % 3 else ("beamer",):
          r = dg.bfs(s, direction=d, loop="kernels", capture=True, device_out=out)
          lev = out[1].cpu().numpy() if out is not None else r.levels
          par = out[0].cpu().numpy() if out is not None else r.parents
          np.testing.assert_array_equal(lev, lev0, err_msg=f"src {s} dir {d} device {out is not None}")
          np.testing.assert_array_equal(par, par0, err_msg=f"src {s} dir {d} device {out is not None}")
          assert r.stats.rounds_completed == rounds
          np.testing.assert_array_equal(r.stats.nf_history, hist)
      _check_single(dg, g, s, "beamer", loop="kernels")


def test_workgroup_loop_equal_oracle():
  """The whole search in one workgroup (graphs of at most 4096 vertices,
  the default up to that size) gives the oracle's levels, lowest-id
  parents, per-round counts and m on undirected and directed graphs under
  every direction policy, on a 1000-level path with a 399-way fan-out, and
  with max_levels hit; a larger graph is refused."""
  from ceregraphs_amd import _lib
  for scale, ef, seed, directed in [(10, 16, 0, False), (12, 8, 3, False), (12, 16, 12, True),
                                    (11, 32, 7, False)]:
    abcd = (0.45, 0.25, 0.15, 0.15) if directed else gen_rmat.GRAPH500_ABCD
    n, ip, ii, op, oi = gen_rmat.rmat_csr(scale, ef, seed, *abcd[:3], directed=directed)
    g = O.Graph(n, ip, ii, op, oi, not directed)
    deg = np.diff(op)
    with DeviceGraph.from_rmat(scale, ef, seed, abcd=abcd, directed=directed) as dg:
      for s in (int(np.argmax(deg)), int(np.nonzero(deg)[0][-1]), 0, n - 1):
        for d in DIRECTIONS:
          for loop in ("workgroup", "auto"):
            _check_single(dg, g, s, d, loop=loop)
  edges = [(i, i + 1) for i in range(1000)] + [(500, j) for j in range(1001, 1400)]
  g = _edges_graph(edges, 1500, directed=False)
  with _device(g) as dg:
    for d in DIRECTIONS:
      _check_single(dg, g, 0, d, loop="workgroup")
    with pytest.raises(_lib.CgbfsError) as e:
      dg.bfs(0, loop="workgroup", max_levels=10)
    assert e.value.code == _lib.CGBFS_ETIMEOUT
  with DeviceGraph.from_rmat(13, 8, 2) as dg:
    with pytest.raises(_lib.CgbfsError) as e:
      dg.bfs(0, loop="workgroup")
    assert e.value.code == _lib.CGBFS_EINVAL


@pytest.mark.parametrize("loop,capture", [("persistent", False), ("kernels", False),
                                          ("kernels", True)])
def test_sources_without_in_edges(loop, capture):
  """A search starts with every vertex without in-edges marked visited, so
  no pull level looks at them again: sources that have out-edges but no
  in-edges (only the source itself is reached among such vertices), a
  vertex with in-edges only, and an isolated one all give the oracle's
  results in every level loop and direction."""
  abcd = (0.45, 0.25, 0.15, 0.15)
  n, ip, ii, op, oi = gen_rmat.rmat_csr(14, 8, 21, *abcd[:3], directed=True)
  g = O.Graph(n, ip, ii, op, oi, False)
  outd, ind = np.diff(op), np.diff(ip)
  srcs = np.nonzero((outd > 0) & (ind == 0))[0]
  assert len(srcs) > 3
  picks = [int(srcs[0]), int(srcs[len(srcs) // 2]), int(srcs[-1]),
           int(np.argmax(outd * (ind == 0))),
           int(np.nonzero((outd == 0) & (ind > 0))[0][0]),
           int(np.nonzero((outd == 0) & (ind == 0))[0][0])]
  with DeviceGraph.from_rmat(14, 8, 21, abcd=abcd, directed=True) as dg:
    for s in picks:
      for d in DIRECTIONS:
        _check_single(dg, g, s, d, loop=loop, capture=capture)


def test_msbfs_host_outputs_reuse_staging():
  """Host outputs are staged in a device buffer kept with the graph: calls
  with more, then fewer, then more lanes reuse or grow it, and every lane
  of every call still equals the oracle."""
  n, ip, ii, op, oi = gen_rmat.rmat_csr(12, 16, 7)
  g = O.Graph(n, ip, ii, op, oi, True)
  deg = np.diff(op)
  cands = np.nonzero(deg)[0]
  rng = np.random.default_rng(7)
  with DeviceGraph.from_rmat(12, 16, 7) as dg:
    for k in (5, 64, 3, 64):
      roots = rng.choice(cands, k, replace=False)
      r = dg.msbfs(roots)
      for l in (0, k // 2, k - 1):
        lev, par, _, _ = O.bfs(g, int(roots[l]))
        np.testing.assert_array_equal(r.levels[l], lev.astype(np.int16), err_msg=f"k={k} lane {l}")
        np.testing.assert_array_equal(r.parents[l], par, err_msg=f"k={k} lane {l}")
