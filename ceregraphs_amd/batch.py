"""Uneven multi-graph batches: many (graph, root) searches in one pass.

The graphs are laid side by side as one disjoint union: graph g's vertex v
becomes v + off[g], and no edge crosses between graphs.  Lane l of one
multi-source pass carries one seed per graph, root l of every graph
(cgbfs_msbfs_seeded).  No path joins two graphs, so graph g's slice of lane
l is exactly the single-root search of graph g from its root l: 8 graphs x
32 roots = 256 searches in one 32-lane pass.  The reference instead runs one
search per launch on one graph at a time (bfs/scripts/run_graph500.py:246-375).

Each graph is first relabelled by the shared-permutation balancer
(balance.py; util/analyze --shared-perm, or --symmetric for undirected
graphs) on a P x P grid, as the reference prepares its inputs, and roots are
drawn in the balanced ids.

The pass runs in graph-replay mode (the round kernels are captured once and
replayed, every level decided on the device).  `autotune` picks the
direction policy, alpha/beta and the blocks per CU for a batch by timing
each candidate on the batch itself.

    python -m ceregraphs_amd.batch [--check K]      # the 8-graph directed workload

GTEPS follow the reference's accounting: m per search is the
directed_source_visited count (bfs/implementation/bfs_timing.py:393-395) or
undirected_dedup for symmetric graphs (:390-392); the batched rate of a
search is m / (t_pass / searches), and the mean is harmonic
(run_graph500.py:130-134).
"""
import argparse
import itertools
import json
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import timing
from .engine import DeviceGraph

DIRECTED_ABCD = (0.45, 0.25, 0.15, 0.15)
# (scale, edgefactor, seed) of the eight directed graphs of the uneven batch
UNEVEN_SPECS = [(12, 8, 10), (14, 32, 11), (15, 16, 12), (16, 8, 13), (17, 16, 14),
                (18, 32, 15), (19, 8, 16), (20, 16, 17)]
UNEVEN_ROOT_SEED = 18
UNEVEN_ROOTS_PER_GRAPH = 32
UNEVEN_BALANCE_GRID = 4  # P of the P x P balancing grid (0: generator ids)

# library defaults (grid_mult 0: by graph size)
DEFAULT_CONFIG = {"direction": "beamer", "alpha": 14.0, "beta": 24.0, "grid_mult": 0,
                  "pull_grab": 1, "seg_grab": 32}


def union_csr(parts):
  """Disjoint union of graphs given as (n, in_ptr, in_idx, out_ptr, out_idx)
  host arrays (int64 offsets, int32 ids).  Returns (n, in_ptr, in_idx,
  out_ptr, out_idx, offsets) with offsets[g] the first id of graph g."""
  ns = [int(p[0]) for p in parts]
  offs = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
  n = int(offs[-1])
  if n >= 2**31:
    raise ValueError("the union has more vertices than int32 ids allow")

  def cat(ptr_i, idx_i):
    ptrs, idxs, base = [np.zeros(1, np.int64)], [], 0
    for g, p in enumerate(parts):
      ptr, idx = np.asarray(p[ptr_i], np.int64), np.asarray(p[idx_i])
      ptrs.append(ptr[1:] + base)
      idxs.append(idx.astype(np.int64) + offs[g])
      base += len(idx)
    return np.concatenate(ptrs), np.concatenate(idxs).astype(np.int32)

  in_ptr, in_idx = cat(1, 2)
  out_ptr, out_idx = cat(3, 4)
  return n, in_ptr, in_idx, out_ptr, out_idx, offs


def balance_part(part, grid, symmetric):
  """The graph relabelled by the balancer's shared permutation q, computed
  on M = CSR(in-view) (M[dst][src], the Matrix Market orientation) with
  --symmetric for an undirected graph, else --shared-perm.  Returns the
  balanced part over grid * ceil(n / grid) vertices (the padding vertices
  have no edges) and the balancer's result (q = .operm)."""
  from . import balance
  n, ip, ii, op, oi = part
  res = balance.balance_csr(n, ip, ii, grid, symmetric=symmetric)
  bip, bii = balance.relabel_csr(ip, ii, res.operm)
  bop, boi = balance.relabel_csr(op, oi, res.operm)
  return (res.n_padded, bip, bii, bop, boi), res


def pick_roots(out_degrees, per_graph, seed):
  """per_graph roots for every graph, among vertices with out-degree >= 1,
  from one numpy Generator(seed) drawn graph after graph in order (the
  run_graph500.py:106-127 rule applied to each graph)."""
  rng = np.random.default_rng(seed)
  return [rng.choice(np.nonzero(d > 0)[0], size=per_graph, replace=False).astype(np.int64)
          for d in out_degrees]


@dataclass
class BatchResult:
  parents: object      # [lanes][n_union] int32 (device tensor or numpy)
  levels: object       # [lanes][n_union] int16
  total_ms: float      # the pass, HIP events on the graph's stream
  rounds: int
  stats: object


class UnevenBatch:
  """Several graphs resident as one disjoint union on the GPU."""

  def __init__(self, parts, symmetric, device=0, graphs=None):
    self.graphs = graphs or []  # the parts as separate device graphs, if kept
    self.balance = []           # the balancer's report per graph, if balanced
    self.ns = [int(p[0]) for p in parts]
    n, ip, ii, op, oi, offs = union_csr(parts)
    self.n, self.offsets, self.symmetric = n, offs, symmetric
    self.out_degree = np.diff(op)
    # GTEPS weight per vertex, as the library's m_weight (bfs_timing.py:390-395)
    if symmetric:
      rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(ip))
      self.weight = np.bincount(rows[ii <= rows], minlength=n).astype(np.int64)
      self.dg = DeviceGraph.from_csr(n, op, oi, symmetric=True, device=device)
    else:
      self.weight = self.out_degree.astype(np.int64)
      self.dg = DeviceGraph.from_csr(n, op, oi, ip, ii, symmetric=False, device=device)

  @classmethod
  def from_rmat(cls, specs, abcd=DIRECTED_ABCD, directed=True, device=0, keep_parts=False,
                balance_grid=UNEVEN_BALANCE_GRID, keep_host=False):
    """Generates the graphs on the GPU, balances each on a balance_grid x
    balance_grid grid (0: keep the generator's ids) and packs the union.
    keep_parts also keeps every graph as a device graph of its own;
    keep_host keeps the (balanced) host arrays of every graph as
    batch.host_parts, (n, in_ptr, in_idx, out_ptr, out_idx) each."""
    parts, graphs, reports = [], [], []
    for scale, ef, seed in specs:
      g = DeviceGraph.from_rmat(scale, ef, seed, abcd=abcd, directed=directed, device=device)
      ip, ii, op, oi = g.download()
      part = (g.n, ip, ii, op, oi)
      if balance_grid:
        g.close()
        part, res = balance_part(part, balance_grid, symmetric=not directed)
        reports.append(res.report)
        if keep_parts:
          g = DeviceGraph.from_csr(part[0], part[3], part[4], part[1], part[2],
                                   symmetric=not directed, device=device)
      parts.append(part)
      if keep_parts:
        graphs.append(g)
      elif not balance_grid:
        g.close()
    batch = cls(parts, symmetric=not directed, device=device, graphs=graphs)
    batch.balance = reports
    batch.host_parts = parts if keep_host else None
    return batch

  def close(self):
    self.dg.close()
    for g in self.graphs:
      g.close()

  def __enter__(self):
    return self

  def __exit__(self, *exc):
    self.close()

  def graph_out_degrees(self):
    return [self.out_degree[self.offsets[g]:self.offsets[g + 1]] for g in range(len(self.ns))]

  def seeds(self, roots):
    """roots[g][l] (ids local to graph g) -> (union vertices, lanes)."""
    lanes = max(len(r) for r in roots)
    verts = np.concatenate([np.asarray(r, np.int64) + self.offsets[g] for g, r in enumerate(roots)])
    lane = np.concatenate([np.arange(len(r), dtype=np.int32) for r in roots])
    return verts, lane, lanes

  def run(self, roots, device_out=None, capture=True, **opts):
    verts, lane, lanes = self.seeds(roots)
    r = self.dg.msbfs_seeded(verts, lane, lanes, device_out=device_out, capture=capture, **opts)
    return BatchResult(r.parents, r.levels, r.stats.total_ms, int(r.stats.rounds_completed), r.stats)

  def instance(self, res, g, l, levels=None, parents=None):
    """(levels, parents) of graph g's root l as graph-local numpy arrays."""
    lo, hi = int(self.offsets[g]), int(self.offsets[g + 1])
    lev = (levels if levels is not None else _np(res.levels))[l, lo:hi].astype(np.int32)
    par = (parents if parents is not None else _np(res.parents))[l, lo:hi].astype(np.int64)
    par = np.where(par >= 0, par - lo, -1).astype(np.int32)
    return lev, par

  def m_table(self, levels):
    """m per (graph, lane) from the pass's levels [lanes][n]."""
    reached = np.asarray(levels) >= 0
    out = np.zeros((len(self.ns), reached.shape[0]), np.int64)
    for g in range(len(self.ns)):
      lo, hi = int(self.offsets[g]), int(self.offsets[g + 1])
      out[g] = reached[:, lo:hi].astype(np.int64) @ self.weight[lo:hi]
    return out


def _np(x):
  return x.cpu().numpy() if hasattr(x, "cpu") else np.asarray(x)


# The tuner's space: the direction policy and Beamer's alpha / beta, blocks
# per CU of the grid-stride kernels, and the pull's work-distribution knobs
# (chunks per grab in a dense round, hub-row segments per grab).  The
# reference fixes its equivalents at compile time (layout_bool.csl:18-38,
# the tau rule of run_bfs.py:239-243).
TUNE_SPACE = {"direction": ["beamer", "beamer_lanes"], "alpha": [4.0, 14.0, 48.0],
              "beta": [12.0, 24.0, 96.0], "grid_mult": [2, 4, 8, 16],
              "pull_grab": [1, 2, 4], "seg_grab": [8, 32, 128]}


def autotune(batch, roots, device_out=None, reps=3, space=None, passes=3):
  """Coordinate descent over `space` (default TUNE_SPACE) from the
  defaults: each knob in turn takes its best value with the others held,
  until a pass changes nothing.  Each configuration is timed on the batch
  itself (best of `reps` passes).  Returns (best config, [(config, ms)...]
  sorted by time)."""
  space = space or TUNE_SPACE
  seen = {}

  def cost(cfg):
    key = tuple(sorted(cfg.items()))
    if key not in seen:
      seen[key] = min(batch.run(roots, device_out=device_out, **cfg).total_ms for _ in range(reps))
    return seen[key]

  start = dict(DEFAULT_CONFIG)
  n = getattr(batch, "n", 0)
  start["grid_mult"] = 2 if n < 2**19 else 4 if n < 2**23 else 8  # the library's choice for n
  best = {k: (start[k] if start.get(k) in v else v[0]) for k, v in space.items()}
  for _ in range(passes):
    changed = False
    for k, values in space.items():
      scores = [(cost(dict(best, **{k: v})), v) for v in values]
      t, v = min(scores, key=lambda x: x[0])
      if v != best[k] and t < cost(best):
        best[k] = v
        changed = True
    if not changed:
      break
  table = sorted(((dict(k), t) for k, t in seen.items()), key=lambda x: x[1])
  return dict(best), table


class ConcurrentPasses:
  """The batch's graphs as separate device graphs, one multi-source pass
  each, all passes in flight at once: one host thread per graph calls into
  the library (the call releases the GIL) and each graph handle runs on its
  own stream, so the passes overlap on the GPU.  Every graph takes its own
  configuration -- its own direction policy and alpha / beta among them --
  which one pass over the union cannot, since that pass decides for all
  graphs together."""

  def __init__(self, graphs, lanes):
    import torch
    self.graphs = graphs
    self.lanes = lanes
    self.outs = [(torch.empty(lanes * g.n, dtype=torch.int32, device="cuda"),
                  torch.empty(lanes * g.n, dtype=torch.int16, device="cuda")) for g in graphs]
    self.pool = None
    self.captured = set()

  def run(self, roots, configs, capture=True):
    """One pass per graph (graph g: roots[g], configs[g]), all at once;
    returns the results in graph order.  In graph-replay mode a graph's
    round kernels are captured on its first pass with a configuration;
    two captures at once from different threads invalidate each other, so
    those first passes run one after another, and only replays overlap."""
    import concurrent.futures
    if self.pool is None:
      self.pool = concurrent.futures.ThreadPoolExecutor(max_workers=len(self.graphs))
    if capture:
      for i, g in enumerate(self.graphs):
        key = (i, tuple(sorted(configs[i].items())))
        if key not in self.captured:
          g.msbfs(roots[i], device_out=self.outs[i], capture=True, **configs[i])
          self.captured.add(key)
    futs = [self.pool.submit(g.msbfs, roots[i], device_out=self.outs[i], capture=capture,
                             **configs[i]) for i, g in enumerate(self.graphs)]
    return [f.result() for f in futs]

  def close(self):
    if self.pool is not None:
      self.pool.shutdown()


def autotune_per_graph(graphs, roots, reps=3, space=None):
  """The coordinate-descent tuner run on each graph alone (its own pass):
  one configuration per graph."""
  import torch
  best = []
  for g, r in zip(graphs, roots):
    out = (torch.empty(len(r) * g.n, dtype=torch.int32, device="cuda"),
           torch.empty(len(r) * g.n, dtype=torch.int16, device="cuda"))

    class One:
      n = g.n

      @staticmethod
      def run(rr, device_out=None, **cfg):
        r = g.msbfs(rr, device_out=out, capture=True, **cfg)
        return BatchResult(r.parents, r.levels, r.stats.total_ms, int(r.stats.rounds_completed),
                           r.stats)
    cfg, _ = autotune(One, r, reps=reps, space=space)
    best.append(cfg)
  return best


def _timed(batch, roots, cfg, device_out, steps):
  import torch
  for _ in range(2):
    batch.run(roots, device_out=device_out, **cfg)
  torch.cuda.synchronize()
  t = time.perf_counter()
  ev = [batch.run(roots, device_out=device_out, **cfg).total_ms for _ in range(steps)]
  torch.cuda.synchronize()
  return (time.perf_counter() - t) / steps * 1e3, float(np.mean(ev))


def main(argv=None):
  p = argparse.ArgumentParser(description="uneven 8-graph directed batch, one pass")
  p.add_argument("--steps", type=int, default=10)
  p.add_argument("--check", type=int, default=8,
                 help="searches checked against single-root GPU searches (0: none)")
  p.add_argument("--no-tune", action="store_true")
  p.add_argument("--balance-grid", type=int, default=UNEVEN_BALANCE_GRID,
                 help="P of the P x P --shared-perm balancing (0: generator ids)")
  args = p.parse_args(argv)
  import torch

  t0 = time.perf_counter()
  batch = UnevenBatch.from_rmat(UNEVEN_SPECS, keep_parts=True, balance_grid=args.balance_grid)
  roots = pick_roots(batch.graph_out_degrees(), UNEVEN_ROOTS_PER_GRAPH, UNEVEN_ROOT_SEED)
  lanes = UNEVEN_ROOTS_PER_GRAPH
  print(f"[batch] {len(UNEVEN_SPECS)} graphs, union n={batch.n} nnz={batch.dg.nnz}, "
        f"{lanes} roots each ({time.perf_counter() - t0:.1f}s to build)", file=sys.stderr)
  par = torch.empty(lanes * batch.n, dtype=torch.int32, device="cuda")
  lev = torch.empty(lanes * batch.n, dtype=torch.int16, device="cuda")
  out = (par, lev)

  res = batch.run(roots, device_out=out)
  levels = lev.view(lanes, batch.n).cpu().numpy()
  m = batch.m_table(levels)
  nsearch = m.size
  if args.check:
    parents = par.view(lanes, batch.n).cpu().numpy()
    rng = np.random.default_rng(0)
    for k in rng.choice(nsearch, size=min(args.check, nsearch), replace=False):
      g, l = divmod(int(k), lanes)
      single = batch.graphs[g].bfs(int(roots[g][l]))
      blev, bpar = batch.instance(None, g, l, levels, parents)
      if not (np.array_equal(blev, single.levels) and np.array_equal(bpar, single.parents)):
        sys.exit(f"graph {g} root {l}: the batched search differs from the single-root one")
    print(f"[batch] {min(args.check, nsearch)} searches equal their single-root searches",
          file=sys.stderr)

  wall_d, ev_d = _timed(batch, roots, DEFAULT_CONFIG, out, args.steps)
  lines = {"defaults": (DEFAULT_CONFIG, wall_d, ev_d)}
  if not args.no_tune:
    best, table = autotune(batch, roots, device_out=out)
    wall_t, ev_t = _timed(batch, roots, best, out, args.steps)
    lines["autotuned"] = (best, wall_t, ev_t)
    print("[batch] autotuner top 5: " + "; ".join(f"{c} {t:.3f} ms" for c, t in table[:5]),
          file=sys.stderr)

  # every graph its own pass with its own tuned configuration, all at once
  if not args.no_tune:
    per_cfg = autotune_per_graph(batch.graphs, roots)
    conc = ConcurrentPasses(batch.graphs, lanes)
    for cfgs, name in [([DEFAULT_CONFIG] * len(batch.graphs), "concurrent_defaults"),
                       (per_cfg, "concurrent_per_graph_tuned")]:
      for _ in range(2):
        conc.run(roots, cfgs)
      torch.cuda.synchronize()
      t = time.perf_counter()
      for _ in range(args.steps):
        rs = conc.run(roots, cfgs)
      torch.cuda.synchronize()
      wall = (time.perf_counter() - t) / args.steps * 1e3
      for g, r in enumerate(rs):  # the same searches as the union pass
        lo, hi = int(batch.offsets[g]), int(batch.offsets[g + 1])
        got = conc.outs[g][1].view(lanes, hi - lo).cpu().numpy()
        if not np.array_equal(got, levels[:, lo:hi]):
          sys.exit(f"graph {g}: the concurrent pass differs from the union pass")
      lines[name] = ([c for c in cfgs], wall, float("nan"))
      print(f"[batch] {name}: {wall:.3f} ms per round of 8 concurrent passes", file=sys.stderr)
    conc.close()

  # for reference: one 32-lane pass per graph, and one launch per search,
  # each on the graph alone (after a warm-up call on each graph)
  t_graph = t_root = 0.0
  for g, dg in enumerate(batch.graphs):
    n = batch.ns[g]
    gp = torch.empty(lanes * n, dtype=torch.int32, device="cuda")
    gl = torch.empty(lanes * n, dtype=torch.int16, device="cuda")
    p1 = torch.empty(n, dtype=torch.int32, device="cuda")
    l1 = torch.empty(n, dtype=torch.int32, device="cuda")
    dg.msbfs(roots[g], capture=True, device_out=(gp, gl))
    dg.bfs(int(roots[g][0]), device_out=(p1, l1), capture=True)
    t_graph += dg.msbfs(roots[g], capture=True, device_out=(gp, gl)).stats.total_ms
    for l in range(lanes):
      t_root += dg.bfs(int(roots[g][l]), device_out=(p1, l1), capture=True).stats.total_ms

  msum = float(m.sum())
  report = {"workload": "uneven-8-directed-rmat", "searches": int(nsearch),
            "union_vertices": batch.n, "union_entries": batch.dg.nnz,
            "rounds": res.rounds, "balance_grid": args.balance_grid,
            "balance_lb": [round(b["lb"], 4) for b in batch.balance]}
  for name, (cfg, wall, ev) in lines.items():
    hm = timing.harmonic_mean(m.ravel() / (wall / 1e3 / nsearch) / 1e9)
    report[name] = {"config": cfg, "ms_per_pass_wall": round(wall, 4),
                    "ms_per_pass_events": None if ev != ev else round(ev, 4),
                    "aggregate_gteps": round(msum / wall * 1e-6, 3),
                    "hmean_gteps": round(hm, 3)}
  report["per_graph_passes_ms"] = round(t_graph, 4)
  report["per_search_launches_ms"] = round(t_root, 4)
  report["per_search_aggregate_gteps"] = round(msum / t_root * 1e-6, 3)
  print(json.dumps(report), flush=True)
  batch.close()


if __name__ == "__main__":
  main()
