"""Serving engine (continuous batching) and the serving loop.

GPU: queries that join a running engine at different rounds, take lanes
freed by finished queries and run interleaved with each other produce the
same parents and levels as the same query run alone (cgbfs_bfs) and as the
oracle, with the oracle's m and round count.  CPU: the arrival stream and
the latency statistics of the serving loop."""
import numpy as np
import pytest

from oracle import oracle as O
from ceregraphs_amd import gen_rmat, serve
from ceregraphs_amd.engine import DeviceGraph, ServingEngine


def test_poisson_stream_is_reproducible_and_shaped():
  nonisolated = [np.arange(1, 50), np.arange(0, 200, 2)]
  a = serve.poisson_stream(1000.0, 5000, nonisolated, seed=20)
  b = serve.poisson_stream(1000.0, 5000, nonisolated, seed=20)
  np.testing.assert_array_equal(a.t, b.t)
  np.testing.assert_array_equal(a.graph, b.graph)
  np.testing.assert_array_equal(a.root, b.root)
  assert np.all(np.diff(a.t) >= 0)
  assert abs(np.mean(np.diff(a.t)) - 1e-3) < 1e-4          # rate 1k/s
  for g, cand in enumerate(nonisolated):
    assert np.isin(a.root[a.graph == g], cand).all()
  assert 0.4 < np.mean(a.graph == 0) < 0.6                  # uniform over graphs


def test_latency_summary():
  s = serve.latency_summary(np.array([1.0, 2.0, 3.0, 4.0, 100.0]) * 1e-3)
  assert s["p50_ms"] == 3.0 and s["max_ms"] == 100.0
  assert 4.0 < s["p99_ms"] <= 100.0


def _serve_staggered(dg, g, roots, direction):
  """Runs the queries through one engine, up to 7 joining per step while
  lanes are free, and checks every result, m and round count against the
  oracle; returns the result tensors."""
  import torch
  n, nq = g.n, len(roots)
  par = [torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(nq)]
  lev = [torch.empty(n, dtype=torch.int16, device="cuda") for _ in range(nq)]
  with ServingEngine(dg, direction=direction) as eng:
    free, lane_of, nxt, done = list(range(64)), {}, 0, {}
    step = 0
    while len(done) < nq:
      for _ in range(7):
        if nxt < nq and free:
          lane = free.pop(0)
          eng.submit(lane, int(roots[nxt]), par[nxt], lev[nxt])
          lane_of[lane] = nxt
          nxt += 1
      eng.step()
      ready, completed, busy, m, depth = eng.poll(wait=True)
      assert ready
      for lane in completed:
        q = lane_of.pop(lane)
        done[q] = (int(m[lane]), int(depth[lane]))
        free.append(lane)
      step += 1
      assert step < 100000
  for q in range(nq):
    lev0, par0, _, rounds = O.bfs(g, int(roots[q]))
    np.testing.assert_array_equal(lev[q].cpu().numpy(), lev0.astype(np.int16), err_msg=f"query {q}")
    np.testing.assert_array_equal(par[q].cpu().numpy(), par0, err_msg=f"query {q}")
    assert done[q] == (O.m_edges(g, lev0), rounds), f"query {q}"
  return par, lev


@pytest.mark.gpu
@pytest.mark.parametrize("scale,ef,seed,directed,direction", [
    (13, 16, 21, False, "beamer_lanes"), (12, 8, 22, True, "beamer_lanes"),
    (12, 16, 23, False, "topdown"), (11, 16, 24, False, "bottomup"), (12, 8, 25, True, "tau"),
    (13, 16, 26, False, "beamer"), (12, 8, 27, True, "beamer")])
def test_engine_results_equal_alone_and_in_mix(scale, ef, seed, directed, direction):
  abcd = (0.45, 0.25, 0.15, 0.15) if directed else gen_rmat.GRAPH500_ABCD
  n, ip, ii, op, oi = gen_rmat.rmat_csr(scale, ef, seed, *abcd[:3], directed=directed)
  g = O.Graph(n, ip, ii, op, oi, not directed)
  rng = np.random.default_rng(seed)
  nq = 150  # more queries than lanes: lanes are reused
  roots = rng.integers(0, n, size=nq)
  roots[:3] = [int(np.argmax(np.diff(op))), roots[5], roots[5]]  # a hub, a repeated root
  with DeviceGraph.from_rmat(scale, ef, seed, abcd=abcd, directed=directed) as dg:
    par, lev = _serve_staggered(dg, g, roots, direction)
    for q in range(0, nq, 15):  # the same query alone, single-root kernel
      alone = dg.bfs(int(roots[q]))
      np.testing.assert_array_equal(lev[q].cpu().numpy(), alone.levels.astype(np.int16))
      np.testing.assert_array_equal(par[q].cpu().numpy(), alone.parents)


@pytest.mark.gpu
@pytest.mark.parametrize("direction", ["bottomup", "beamer_lanes"])
def test_engine_hub_with_late_segments(direction):
  """A hub with 20000 in-neighbours (32 segments in the first list, 47 in
  the late one) pulled by queries that join at different rounds: 64 roots
  each linked to one to four of its in-neighbours (below id 8192 only,
  above only, or anywhere), then 40 of those in-neighbours as roots."""
  import scipy.sparse as sp
  hub_deg = 20000
  hub = hub_deg + 64
  rng = np.random.default_rng(11)
  src, dst = [j for j in range(hub_deg)], [hub] * hub_deg
  for l in range(64):
    lo, hi = (0, 8192) if l < 16 else (8192, hub_deg) if l < 32 else (0, hub_deg)
    for p in rng.integers(lo, hi, size=1 + l % 4):
      src.append(hub_deg + l)
      dst.append(int(p))
  src, dst = np.array(src + dst), np.array(dst + src)  # undirected
  M = sp.coo_matrix((np.ones(len(src)), (dst, src)), shape=(hub + 1, hub + 1))
  g = O.Graph.from_matrix(M)
  roots = np.concatenate([np.arange(hub_deg, hub_deg + 64), rng.integers(0, hub_deg, size=40)])
  with DeviceGraph.from_csr(g.n, g.out_ptr, g.out_idx, g.in_ptr, g.in_idx, g.symmetric) as dg:
    _serve_staggered(dg, g, roots, direction)


@pytest.mark.gpu
@pytest.mark.parametrize("admission", ["continuous", "waves"])
def test_serve_stream_admissions_equal_oracle(admission):
  """The serving loop over three resident graphs (two undirected R-MAT, one
  directed) under a Poisson stream fast enough that queries queue: with
  waves, an engine takes queries only while none is in flight (several
  waves per graph), with continuous batching at every round.  Every query's
  parents, levels, m and round count equal the oracle's."""
  import torch
  specs = [(10, 16, 31, False), (12, 8, 32, False), (11, 16, 33, True)]
  dgs, graphs = [], []
  for sc, ef, seed, directed in specs:
    abcd = (0.45, 0.25, 0.15, 0.15) if directed else gen_rmat.GRAPH500_ABCD
    n, ip, ii, op, oi = gen_rmat.rmat_csr(sc, ef, seed, *abcd[:3], directed=directed)
    graphs.append(O.Graph(n, ip, ii, op, oi, not directed))
    dgs.append(DeviceGraph.from_rmat(sc, ef, seed, abcd=abcd, directed=directed))
  slots = [serve.GraphSlot(d, 48, direction=serve.ADMISSION_DIRECTION[admission]) for d in dgs]
  try:
    nonisolated = [np.nonzero(np.diff(g.out_ptr) > 0)[0] for g in graphs]
    stream = serve.poisson_stream(20000.0, 400, nonisolated, seed=7)
    kept = {}

    def keep(q, buf, m, rounds):
      kept[q] = (buf[0].cpu().numpy(), buf[1].cpu().numpy(), m, rounds)

    lat, m, _ = serve.serve_stream(slots, stream, deadline_s=300.0, on_complete=keep,
                                   admission=admission)
    torch.cuda.synchronize()
    assert len(kept) == 400 and np.isfinite(lat).all()
    for q in range(400):
      g = graphs[int(stream.graph[q])]
      lev0, par0, _, rounds0 = O.bfs(g, int(stream.root[q]))
      par, lev, mq, rounds = kept[q]
      np.testing.assert_array_equal(lev, lev0.astype(np.int16), err_msg=f"query {q}")
      np.testing.assert_array_equal(par, par0, err_msg=f"query {q}")
      assert (mq, rounds) == (O.m_edges(g, lev0), rounds0) == (int(m[q]), rounds), f"query {q}"
  finally:
    for s in slots:
      s.close()
    for d in dgs:
      d.close()
