"""The uneven batch at its real size on the GPU: eight directed R-MAT graphs
(A/B/C/D = .45/.25/.15/.15, self-loops removed) at scales 12, 14, 15, 16,
17, 18, 19, 20 with edgefactors 8, 32, 16, 8, 16, 32, 8, 16 and seeds 10-17,
each balanced on a 4x4 grid with --shared-perm semantics, 32 roots per graph
drawn from seed 18 among vertices with out-degree >= 1 (the
run_graph500.py:106-127 rule per graph): 256 searches in one captured,
device-decided pass (cgbfs_msbfs_seeded over the disjoint union).

Every (graph, root) search of the pass is checked against the oracle on its
own balanced graph: levels, lowest-id parents and m.  The autotuner's
configuration gives a pass identical to the default one.  The batch and its
passes are built once (module fixture); each graph is then its own test."""
import numpy as np
import pytest

from oracle import oracle as O
from ceregraphs_amd import batch as B

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


@pytest.fixture(scope="module")
def c4():
  batch = B.UnevenBatch.from_rmat(B.UNEVEN_SPECS, balance_grid=B.UNEVEN_BALANCE_GRID,
                                  keep_host=True, keep_parts=True)
  roots = B.pick_roots(batch.graph_out_degrees(), B.UNEVEN_ROOTS_PER_GRAPH, B.UNEVEN_ROOT_SEED)
  res = batch.run(roots, capture=True)
  levels, parents = np.asarray(res.levels), np.asarray(res.parents)
  best, table = B.autotune(batch, roots, reps=1,
                           space={"direction": ["beamer", "beamer_lanes"], "alpha": [4.0, 48.0],
                                  "grid_mult": [2, 8]})
  tuned = batch.run(roots, capture=True, **best)
  yield {"batch": batch, "roots": roots, "levels": levels, "parents": parents,
         "tuned": (np.asarray(tuned.levels), np.asarray(tuned.parents), best),
         "m": batch.m_table(levels)}
  batch.close()


def test_c4_shape(c4):
  batch = c4["batch"]
  assert len(batch.ns) == 8 and c4["levels"].shape == (32, batch.n)
  # 4x4 balancing pads each graph to a multiple of 4 vertices
  for (s, _, _), n in zip(B.UNEVEN_SPECS, batch.ns):
    assert n == 4 * -(-(1 << s) // 4)
  assert all(len(r) == 32 for r in c4["roots"])


@pytest.mark.parametrize("g", range(8))
def test_c4_searches_equal_oracle(c4, g):
  batch = c4["batch"]
  n, ip, ii, op, oi = batch.host_parts[g]
  gg = O.Graph(n, ip, ii, op, oi, False)
  for l, r in enumerate(c4["roots"][g]):
    lev, par, _, _ = O.bfs(gg, int(r))
    blev, bpar = batch.instance(None, g, l, c4["levels"], c4["parents"])
    np.testing.assert_array_equal(blev, lev, err_msg=f"graph {g} root {l}")
    np.testing.assert_array_equal(bpar, par, err_msg=f"graph {g} root {l}")
    assert c4["m"][g, l] == O.m_edges(gg, lev), f"graph {g} root {l}"


def test_c4_autotuned_pass_equals_default(c4):
  tlev, tpar, cfg = c4["tuned"]
  np.testing.assert_array_equal(tlev, c4["levels"], err_msg=str(cfg))
  np.testing.assert_array_equal(tpar, c4["parents"], err_msg=str(cfg))


def test_c4_concurrent_per_graph_passes_equal_union(c4):
  """Every graph's own pass, all eight in flight at once on their own
  streams, each graph with its own direction policy and alpha / beta, gives
  exactly the union pass's levels and parents."""
  batch, roots = c4["batch"], c4["roots"]
  cfgs = [{"direction": ("beamer", "beamer_lanes", "bottomup", "topdown")[g % 4],
           "alpha": (4.0, 14.0, 48.0)[g % 3], "beta": (12.0, 24.0, 96.0)[(g + 1) % 3]}
          for g in range(len(batch.graphs))]
  conc = B.ConcurrentPasses(batch.graphs, 32)
  try:
    for cap in (True, False):
      conc.run(roots, cfgs, capture=cap)
      for g in range(len(batch.graphs)):
        lo, hi = int(batch.offsets[g]), int(batch.offsets[g + 1])
        lev = conc.outs[g][1].view(32, hi - lo).cpu().numpy()
        par = conc.outs[g][0].view(32, hi - lo).cpu().numpy().astype(np.int64)
        np.testing.assert_array_equal(lev, c4["levels"][:, lo:hi], err_msg=f"graph {g} {cfgs[g]}")
        upar = c4["parents"][:, lo:hi].astype(np.int64)
        np.testing.assert_array_equal(par, np.where(upar >= 0, upar - lo, -1), err_msg=f"graph {g}")
  finally:
    conc.close()
