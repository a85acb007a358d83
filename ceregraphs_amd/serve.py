"""BFS serving: a stream of single-root queries over device-resident graphs.

Each graph stays in HBM with one serving engine (cgbfs_engine_*): a 64-lane
multi-source search that never stops.  A query waits until it has arrived,
a lane of its graph's engine is free and a result buffer is free, then
joins the engine at the next round boundary.  It runs interleaved with the
queries already in flight, because one adjacency pass per round serves
every lane, and leaves when its frontier empties; its lane takes the next
query at once.  Engines of different graphs run on their own streams and
overlap on the GPU.  This is continuous batching: queries are admitted
while others are running, not collected into closed batches.

Admission "waves" instead lets an engine take queries only while it has none
in flight: everything that arrived meanwhile (up to 64) starts in one round,
so the lanes stay at one depth and the engine runs the union-count direction
policy (every lane pushes or pulls together).  A round then costs what a
round of a closed 64-root pass costs, about 2.5x less than a round over
lanes at different depths, at the price of waiting for the wave in flight.

Result buffers (parents int32[n] + levels int16[n]) come from per-graph
pools sized within the HBM budget together with the graphs and the engine
workspaces.  A buffer goes back to its pool when the query completes, so
this loop measures serving the search, not holding results.

    python -m ceregraphs_amd.serve --rates 1000,10000,50000 --queries 3000
    python -m ceregraphs_amd.serve --admission waves

The plain loop for comparison runs the same stream one query per launch
(cgbfs_bfs) in arrival order.  Latency = completion - arrival (wall clock);
GTEPS = sum of m over the queries / (last completion - first arrival), with
m as bfs/implementation/bfs_timing.py:390-392 counts it.
"""
import argparse
import json
import sys
import time
from dataclasses import dataclass

import numpy as np

from .engine import DeviceGraph, ServingEngine

# (scale, edgefactor, seed) of the served graphs: undirected Graph500 R-MAT
SERVED_SPECS = [(16, 16, 21), (18, 16, 22), (20, 16, 23), (22, 16, 24), (24, 16, 25),
                (26, 16, 26)]
STREAM_SEED = 20
HBM_BUDGET = 200 << 30
# admission policy -> the engines' direction policy
ADMISSION_DIRECTION = {"continuous": "beamer_lanes", "waves": "beamer"}


@dataclass
class Stream:
  t: np.ndarray      # arrival times (s), ascending
  graph: np.ndarray  # graph index per query
  root: np.ndarray   # root per query (non-isolated vertex of its graph)


def poisson_stream(rate, count, nonisolated, seed=STREAM_SEED):
  """count queries with exponential gaps of mean 1/rate; each picks a graph
  uniformly and then a root uniformly among that graph's non-isolated
  vertices (one numpy Generator, drawn in that order per query)."""
  rng = np.random.default_rng(seed)
  t, graph, root = np.empty(count), np.empty(count, np.int64), np.empty(count, np.int64)
  now = 0.0
  for i in range(count):
    now += rng.exponential(1.0 / rate)
    g = int(rng.integers(len(nonisolated)))
    t[i], graph[i] = now, g
    root[i] = nonisolated[g][int(rng.integers(len(nonisolated[g])))]
  return Stream(t, graph, root)


def latency_summary(lat):
  lat = np.asarray(lat) * 1e3
  return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
          "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}


class GraphSlot:
  """One served graph: its device graph, engine and result-buffer pool."""

  def __init__(self, dg, buffers, **engine_opts):
    import torch
    self.dg = dg
    self.n = dg.n
    self.engine = ServingEngine(dg, **engine_opts)
    self.engine_opts = dict(engine_opts)
    self.free_bufs = [(torch.empty(self.n, dtype=torch.int32, device="cuda"),
                       torch.empty(self.n, dtype=torch.int16, device="cuda")) for _ in range(buffers)]
    self.free_lanes = list(range(ServingEngine.LANES))
    self.lane_query = {}
    self.inflight = False
    self.waiting = []  # arrived query ids, FIFO
    self.admitting = False  # a wave is being admitted (admission "waves")

  def set_engine(self, **engine_opts):
    """Replaces the engine (none may be in flight), e.g. for another direction
    policy; the result pool stays."""
    if self.lane_query or self.inflight:
      raise RuntimeError("the engine has queries in flight")
    if engine_opts != self.engine_opts:
      self.engine.close()
      self.engine = ServingEngine(self.dg, **engine_opts)
      self.engine_opts = dict(engine_opts)

  def close(self):
    self.engine.close()


def graph_bytes(dg):
  """HBM held by a graph and its multi-source data: offsets, ids, slot ids."""
  views = 1 if dg.symmetric else 2
  return views * (8 * (dg.n + 1) + 8 * dg.nnz) + 8 * dg.n


def engine_bytes(n):
  return 3 * 8 * n + 2 * 4 * n + n // 4 + (16 << 20)


def build_slots(specs, budget, buffers_cap=ServingEngine.LANES, **engine_opts):
  """Generates the graphs and sizes the result pools within the budget:
  the graphs and engines first, then equal numbers of result buffers for
  every graph, at most one per lane."""
  dgs = [DeviceGraph.from_rmat(s, ef, seed) for s, ef, seed in specs]
  fixed = sum(graph_bytes(d) + engine_bytes(d.n) for d in dgs)
  per_buf = sum(6 * d.n for d in dgs)
  nbuf = int(min(buffers_cap, max(0, (budget - fixed) // max(per_buf, 1))))
  if nbuf < 1:
    raise RuntimeError(f"the HBM budget ({budget >> 30} GiB) does not hold the graphs")
  slots = [GraphSlot(d, nbuf, **engine_opts) for d in dgs]
  return slots, {"budget_bytes": int(budget), "graphs_and_engines_bytes": int(fixed),
                 "result_buffers_per_graph": nbuf,
                 "result_bytes": int(nbuf * per_buf)}


def serve_stream(slots, stream, deadline_s=120.0, on_complete=None, admission="continuous"):
  """Runs the stream through the engines; returns per-query latency (s),
  m per query, and the wall span from the first arrival to the last
  completion.  on_complete(q, (parents, levels), m, rounds), if given, sees
  each query's result buffers before they go back to the pool.  admission:
  "continuous" (a query takes a free lane at the next round) or "waves" (an
  engine takes queries only while none is in flight)."""
  if admission not in ADMISSION_DIRECTION:
    raise ValueError(f"admission must be one of {sorted(ADMISSION_DIRECTION)}")
  nq = len(stream.t)
  lat = np.full(nq, np.nan)
  m = np.zeros(nq, np.int64)
  done = 0
  nxt = 0
  t0 = time.perf_counter()
  while done < nq:
    now = time.perf_counter() - t0
    if now > deadline_s:
      raise RuntimeError(f"serving did not finish within {deadline_s}s ({done}/{nq} done)")
    while nxt < nq and stream.t[nxt] <= now:
      slots[int(stream.graph[nxt])].waiting.append(nxt)
      nxt += 1
    progressed = False
    for s in slots:
      if s.inflight:
        ready, completed, _, mm, depth = s.engine.poll(wait=False)
        if not ready:
          continue
        s.inflight = False
        progressed = True
        tnow = time.perf_counter() - t0
        for lane in completed:
          q, buf = s.lane_query.pop(lane)
          lat[q] = tnow - stream.t[q]
          m[q] = mm[lane]
          if on_complete is not None:
            on_complete(q, buf, int(mm[lane]), int(depth[lane]))
          s.free_lanes.append(lane)
          s.free_bufs.append(buf)
          done += 1
      # admit arrived queries into free lanes (waves: only into an engine
      # with none in flight), then run a round if any lane is busy
      while (s.waiting and s.free_lanes and s.free_bufs and
             (admission == "continuous" or len(s.free_lanes) == ServingEngine.LANES or
              s.admitting)):
        s.admitting = True
        q = s.waiting.pop(0)
        lane = s.free_lanes.pop(0)
        buf = s.free_bufs.pop()
        s.engine.submit(lane, int(stream.root[q]), buf[0], buf[1])
        s.lane_query[lane] = (q, buf)
      s.admitting = False
      if s.lane_query:
        s.engine.step()
        s.inflight = True
        progressed = True
    if not progressed and nxt < nq and not any(s.lane_query for s in slots):
      time.sleep(max(0.0, min(1e-3, stream.t[nxt] - (time.perf_counter() - t0))))
  span = float(np.nanmax(stream.t + lat) - stream.t[0])
  return lat, m, span


def plain_loop(slots, stream, deadline_s=120.0):
  """The same stream, one query per launch (cgbfs_bfs), in arrival order."""
  import torch
  nq = len(stream.t)
  lat = np.full(nq, np.nan)
  m = np.zeros(nq, np.int64)
  outs = [(torch.empty(s.n, dtype=torch.int32, device="cuda"),
           torch.empty(s.n, dtype=torch.int32, device="cuda")) for s in slots]
  t0 = time.perf_counter()
  for q in range(nq):
    now = time.perf_counter() - t0
    if now > deadline_s:
      raise RuntimeError(f"the plain loop did not finish within {deadline_s}s ({q}/{nq} done)")
    if stream.t[q] > now:
      time.sleep(stream.t[q] - now)
    g = int(stream.graph[q])
    r = slots[g].dg.bfs(int(stream.root[q]), device_out=outs[g], capture=True)
    lat[q] = time.perf_counter() - t0 - stream.t[q]
    m[q] = r.stats.m_edges
  span = float(np.max(stream.t + lat) - stream.t[0])
  return lat, m, span


def main(argv=None):
  p = argparse.ArgumentParser(description="BFS serving with continuous batching")
  p.add_argument("--rates", default="1000,10000,50000", help="queries per second, comma-separated")
  p.add_argument("--queries", type=int, default=3000, help="queries per rate")
  p.add_argument("--max-scale", type=int, default=26, help="serve only graphs up to this scale")
  p.add_argument("--budget-gib", type=float, default=HBM_BUDGET / 2**30)
  p.add_argument("--no-plain", action="store_true")
  p.add_argument("--admission", default="continuous,waves",
                 help="continuous, waves, or both (comma-separated; each with its "
                      "direction policy: continuous beamer_lanes, waves beamer)")
  args = p.parse_args(argv)
  admissions = args.admission.split(",")
  for a in admissions:
    if a not in ADMISSION_DIRECTION:
      p.error(f"unknown admission {a}")
  import torch

  specs = [s for s in SERVED_SPECS if s[0] <= args.max_scale]
  t0 = time.perf_counter()
  slots, mem = build_slots(specs, int(args.budget_gib * 2**30))
  print(f"[serve] {len(slots)} graphs resident ({time.perf_counter() - t0:.1f}s): {mem}",
        file=sys.stderr, flush=True)
  nonisolated = [np.nonzero(s.dg.out_degree() > 0)[0] for s in slots]
  report = {"workload": "serve-rmat-" + "-".join(f"s{s}" for s, _, _ in specs), "memory": mem,
            "rates": {}}
  for rate in [float(x) for x in args.rates.split(",")]:
    stream = poisson_stream(rate, args.queries, nonisolated)
    line = {}
    for a in admissions:
      for s in slots:
        s.set_engine(direction=ADMISSION_DIRECTION[a])
      lat, m, span = serve_stream(slots, stream, admission=a)
      key = "engine" if a == "continuous" else "engine_" + a
      line[key] = dict(latency_summary(lat), gteps=float(m.sum() / span / 1e9),
                       qps=float(len(lat) / span))
    if not args.no_plain:
      plat, pm, pspan = plain_loop(slots, stream)
      if not np.array_equal(pm, m):
        print("[serve] WARNING: m differs between the engine and the plain loop", file=sys.stderr)
      line["plain"] = dict(latency_summary(plat), gteps=float(pm.sum() / pspan / 1e9),
                           qps=float(len(plat) / pspan))
    report["rates"][str(int(rate))] = line
    print(f"[serve] {int(rate)} q/s: {json.dumps(line)}", file=sys.stderr, flush=True)
  print(json.dumps(report), flush=True)
  for s in slots:
    s.close()
  torch.cuda.synchronize()


if __name__ == "__main__":
  main()
