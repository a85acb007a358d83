"""Uneven multi-graph batches (ceregraphs_amd/batch.py).

CPU: the disjoint union is built right (offsets, both views, orientation)
and the property the batch rests on holds on the oracle -- a search from
graph g's root inside the union reaches exactly graph g's single-root
levels and parents, and nothing outside graph g.
GPU: every (graph, root) search of a batch pass is bit-exact against the
oracle on its own graph, directed and undirected, in both round drivers."""
import numpy as np
import pytest

from oracle import oracle as O
from ceregraphs_amd import batch as B
from ceregraphs_amd import gen_rmat

SMALL_DIRECTED = [(8, 8, 10), (10, 32, 11), (9, 16, 12), (11, 8, 13)]


def _parts(specs, directed=True):
  abcd = B.DIRECTED_ABCD if directed else gen_rmat.GRAPH500_ABCD
  out = []
  for s, ef, seed in specs:
    n, ip, ii, op, oi = gen_rmat.rmat_csr(s, ef, seed, *abcd[:3], directed=directed)
    out.append((n, ip, ii, op, oi))
  return out


def test_union_csr_offsets_and_views():
  parts = _parts(SMALL_DIRECTED)
  n, ip, ii, op, oi, offs = B.union_csr(parts)
  assert n == sum(p[0] for p in parts) and offs[-1] == n
  for g, (gn, gip, gii, gop, goi) in enumerate(parts):
    lo, hi = offs[g], offs[g + 1]
    np.testing.assert_array_equal(np.diff(ip[lo:hi + 1]), np.diff(gip))
    np.testing.assert_array_equal(ii[ip[lo]:ip[hi]], gii + lo)
    np.testing.assert_array_equal(oi[op[lo]:op[hi]], goi + lo)
  # rows stay ascending and no edge leaves its graph
  for g in range(len(parts)):
    seg = ii[ip[offs[g]]:ip[offs[g + 1]]]
    assert seg.min() >= offs[g] and seg.max() < offs[g + 1]


def test_search_inside_union_equals_search_on_the_graph():
  parts = _parts(SMALL_DIRECTED)
  n, ip, ii, op, oi, offs = B.union_csr(parts)
  u = O.Graph(n, ip, ii, op, oi, False)
  roots = B.pick_roots([np.diff(p[3]) for p in parts], 3, 18)
  for g, p in enumerate(parts):
    gg = O.Graph(p[0], p[1], p[2], p[3], p[4], False)
    for r in roots[g]:
      lev, par, _, _ = O.bfs(gg, int(r))
      ulev, upar, _, _ = O.bfs(u, int(r) + int(offs[g]))
      lo, hi = offs[g], offs[g + 1]
      np.testing.assert_array_equal(ulev[lo:hi], lev)
      np.testing.assert_array_equal(np.where(upar[lo:hi] >= 0, upar[lo:hi] - lo, -1), par)
      assert (ulev[:lo] == -1).all() and (ulev[hi:] == -1).all()


def test_balanced_part_is_a_relabelling():
  """balance_part gives the same searches under q: levels[q[v]] from q[r]
  equal levels[v] from r; padding vertices stay unreached."""
  part = _parts([(10, 8, 3)])[0]
  bpart, res = B.balance_part(part, 3, symmetric=False)  # 1024 -> 3 * 342 = 1026 vertices
  q = res.operm
  assert bpart[0] == 1026 and res.report["grid"] == 3
  g0 = O.Graph(*part, False)
  g1 = O.Graph(*bpart, False)
  for r in B.pick_roots([np.diff(part[3])], 4, 18)[0]:
    lev0, par0, hist0, _ = O.bfs(g0, int(r))
    lev1, par1, hist1, _ = O.bfs(g1, int(q[r]))
    np.testing.assert_array_equal(lev1[q[:part[0]]], lev0)
    assert (lev1[np.setdiff1d(np.arange(1026), q[:part[0]])] == -1).all()
    np.testing.assert_array_equal(hist1, hist0)
    assert O.validate(g1, int(q[r]), par1, lev1) == []


def test_pick_roots_rule():
  degs = [np.array([0, 2, 1, 0, 5]), np.array([1, 1, 0])]
  roots = B.pick_roots(degs, 2, 18)
  for d, r in zip(degs, roots):
    assert len(set(r.tolist())) == 2 and (d[r] > 0).all()
  again = B.pick_roots(degs, 2, 18)
  for a, b in zip(roots, again):
    np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("directed,specs,per_graph,grid", [
    (True, SMALL_DIRECTED, 32, 0), (False, [(9, 16, 21), (7, 8, 22), (10, 16, 23)], 20, 0),
    (True, SMALL_DIRECTED, 16, 3)])
def test_batch_searches_equal_oracle(directed, specs, per_graph, grid):
  parts = _parts(specs, directed)
  if grid:
    parts = [B.balance_part(p, grid, symmetric=not directed)[0] for p in parts]
  with B.UnevenBatch(parts, symmetric=not directed) as batch:
    roots = B.pick_roots(batch.graph_out_degrees(), per_graph, 18)
    for capture, cfg in [(True, {}), (False, {}), (True, {"direction": "beamer_lanes"}),
                         (True, {"direction": "bottomup"})]:
      res = batch.run(roots, capture=capture, **cfg)
      levels, parents = np.asarray(res.levels), np.asarray(res.parents)
      m = batch.m_table(levels)
      for g, p in enumerate(parts):
        gg = O.Graph(p[0], p[1], p[2], p[3], p[4], not directed)
        for l, r in enumerate(roots[g]):
          lev, par, _, _ = O.bfs(gg, int(r))
          blev, bpar = batch.instance(res, g, l, levels, parents)
          np.testing.assert_array_equal(blev, lev, err_msg=f"graph {g} root {l} {cfg} {capture}")
          np.testing.assert_array_equal(bpar, par, err_msg=f"graph {g} root {l} {cfg} {capture}")
          assert m[g, l] == O.m_edges(gg, lev)


class _FakeBatch:
  """Stands in for an UnevenBatch: a pass costs a known function of the
  configuration, so the tuner's search can be checked without a GPU."""

  def __init__(self, cost, n=1 << 20):
    self.cost, self.n, self.calls = cost, n, 0

  def run(self, roots, device_out=None, **cfg):
    self.calls += 1

    class R:
      pass
    r = R()
    r.total_ms = self.cost(cfg)
    return r


def test_autotune_coordinate_descent_finds_separable_optimum():
  best_vals = {"direction": "beamer_lanes", "alpha": 48.0, "beta": 12.0, "grid_mult": 16,
               "pull_grab": 2, "seg_grab": 128}

  def cost(cfg):  # separable: each knob adds its own penalty
    return 1.0 + sum(0.1 * (cfg[k] != v) + 0.01 * list(B.TUNE_SPACE[k]).index(cfg[k])
                     for k, v in best_vals.items())
  fb = _FakeBatch(cost)
  best, table = B.autotune(fb, None, reps=1)
  assert best == best_vals
  assert table[0][0] == best_vals and table[0][1] == min(t for _, t in table)
  # far fewer passes than the full product of the space
  full = 1
  for v in B.TUNE_SPACE.values():
    full *= len(v)
  assert fb.calls < full / 4


def test_autotune_keeps_defaults_when_nothing_helps():
  fb = _FakeBatch(lambda cfg: 5.0, n=1 << 20)
  best, _ = B.autotune(fb, None, reps=1)
  assert best["direction"] == "beamer" and best["alpha"] == 14.0 and best["beta"] == 24.0
  assert best["grid_mult"] == 4 and best["pull_grab"] == 1 and best["seg_grab"] == 32
