#!/usr/bin/env python3
"""Headline benchmark: harmonic-mean GTEPS on RMAT scale-24 edgefactor-16
with 64 roots -- batched multi-source BFS against the same roots run one by
one -- on one MI355X.

    python bench.py                      # the metric (s24 ef16, 64 roots)
    python bench.py --workload c2        # s21 ef16 seed 1, 64 roots one per launch
    python bench.py --workload c1        # s10 ef16 seed 0 balanced 4x4, source 0

A step is one 64-lane multi-source pass over the 64 roots (c3), or one
single-root search (c2: each of the 64 roots in turn; c1: source 0), with
the graph already resident in HBM and parents / levels written to HBM.
GTEPS accounting follows the reference: m per search is the undirected_dedup
count of bfs/implementation/bfs_timing.py:390-392, roots are drawn as in
bfs/scripts/run_graph500.py:106-127 (non-isolated vertices,
default_rng(seed).choice without replacement) and the mean is the harmonic
mean of run_graph500.py:130-134.

Beside the GPU line the reference's CPU path is timed on this host's cores:
the oracle's literal boolean-SpMV restatement of the round loop
(bool_pe.csl:865-1004) and scipy's breadth_first_order, the reference's own
correctness oracle (run_bfs.py:513-552), both single-threaded.  The
Cerebras simulator run (cs_python run_bfs.py) is recorded as unavailable
when the SDK is absent.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0
METRIC_C3 = "harmonic-mean GTEPS, RMAT s24 ef16, 64 roots: batched MS-BFS vs per-root"
WORKLOADS = {
    "c3": {"scale": 24, "edgefactor": 16, "seed": 3, "root_seed": 4, "nroots": 64},
    "c2": {"scale": 21, "edgefactor": 16, "seed": 1, "root_seed": 2, "nroots": 64},
    "c1": {"scale": 10, "edgefactor": 16, "seed": 0, "root_seed": 2, "nroots": 64,
           "balance_grid": 4, "source": 0},
}


def log(msg):
  print(f"[bench] {msg}", file=sys.stderr, flush=True)


def harmonic_mean(values):
  values = [v for v in values if v > 0 and np.isfinite(v)]
  return len(values) / sum(1.0 / v for v in values) if values else float("nan")


def pull_round_bytes(n, rows, edges, pairs, nf, noin_pairs=0):
  """Algorithmic HBM bytes of the pull part of one multi-source round: lane
  words of every vertex (8 V), offsets of the adjacency lists opened (8 U),
  4 B column id + 8 B frontier word per adjacency entry read (12 E), 4 B
  parent + 2 B level per reached (lane, vertex) pair (6 P), and 8 B vis +
  8 B next word per newly active vertex (16 N).  The second pull round of a
  search also writes the -1 outputs of every (lane, vertex) pair whose
  vertex has no in-edges (6 B each, noin_pairs of them)."""
  return 8 * n + 8 * rows + 12 * edges + 6 * pairs + 16 * nf + 6 * noin_pairs


def level_bytes(n, direction, edges, rows, nf, frontier):
  """Algorithmic HBM bytes of one single-root level.  Pull: visited and
  frontier bitmaps (V/8 each), the offsets of every row opened (8 U), 4 B
  per adjacency entry read (4 E) and parent + level of each new vertex
  (8 N).  Push: the frontier bitmap (V/8), the frontier's offsets (8 F),
  4 E and 8 N."""
  if direction == 1:
    return n / 8 + n / 8 + 8 * rows + 4 * edges + 8 * nf
  return n / 8 + 8 * frontier + 4 * edges + 8 * nf


def per_root_roofline(n, results):
  """The dominant level of each single-root search (its longest one, a pull
  level on these graphs) over all searches: algorithmic bytes from the
  level's device counters / the level's event time (its kernels bfs_level,
  bfs_push_scan and bfs_push_big)."""
  byts, secs, dirs = 0.0, 0.0, []
  for r in results:
    st = r.stats
    d = int(np.argmax(st.level_ms))
    front = 1 if d == 0 else int(st.nf_history[d - 1])
    byts += level_bytes(n, int(st.direction_history[d]), int(st.edges_scanned[d]),
                        int(st.rows_scanned[d]), int(st.nf_history[d]), front)
    secs += float(st.level_ms[d]) / 1e3
    dirs.append(int(st.direction_history[d]))
  achieved = byts / secs / 1e9 if secs > 0 else 0.0
  return {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
          "frac": round(achieved / PEAK_HBM_GBS, 4), "kernel": "bfs_level",
          "level": "longest level of each search",
          "pull_levels": int(sum(dirs)), "searches": len(results),
          "avg_level_ms": round(secs * 1e3 / max(len(results), 1), 4),
          "algorithmic_bytes_per_level": int(byts / max(len(results), 1))}


def traffic_from_profile(name, key="hbm_bytes_per_launch"):
  """Measured HBM bytes per launch from a committed rocprof summary
  (profiles/<name>), with the commit its kernels were profiled at."""
  path = os.path.join(HERE, "profiles", name)
  if not os.path.exists(path):
    return None, None
  with open(path, encoding="utf-8") as f:
    prof = json.load(f)
  return prof.get(key), {"profile": f"profiles/{name}", "commit": prof.get("commit")}


def host_info():
  model = "unknown"
  try:
    with open("/proc/cpuinfo", encoding="utf-8") as f:
      for line in f:
        if line.startswith("model name"):
          model = line.split(":", 1)[1].strip()
          break
  except OSError:
    pass
  try:
    avail = len(os.sched_getaffinity(0))
  except AttributeError:
    avail = os.cpu_count()
  return {"cpu_model": model, "nproc": avail}


def csl_simulator_status():
  """The reference's own CPU path is the Cerebras fabric simulator
  (cs_python bfs/scripts/run_bfs.py); it needs the Cerebras SDK."""
  if shutil.which("cs_python") and shutil.which("cslc"):
    return {"status": "available, not run", "reason": "cs_python found"}
  return {"status": "unavailable",
          "reason": "the Cerebras SDK (cslc, cs_python) is not installed on this host"}


def cpu_baseline(dg, roots, spmv_roots, bfo_roots, lanes_check=None):
  """Times the oracle's boolean-SpMV BFS (spmv_roots of the roots) and scipy
  breadth_first_order (bfo_roots) on one host core over the same graph;
  matrix construction stays outside the timed region, as the graph upload
  does on the GPU.  lanes_check(l) -> int16/int32 levels of root l from the
  GPU, compared with the CPU levels."""
  from oracle import oracle as O
  from scipy.sparse.csgraph import breadth_first_order
  n = dg.n
  ip, ii, op, oi = dg.download()
  g = O.Graph(n, ip, ii, op, oi, dg.symmetric)
  out = {"unit": "GTEPS", "cores": 1, "kind": "port"}
  if spmv_roots:
    M = g.matrix().astype(np.int32)
    t, m, bad = 0.0, 0, 0
    for l in range(spmv_roots):
      s = int(roots[l])
      tc = time.perf_counter()
      lev, _, _, _ = O.spmv_bfs(g, s, matrix=M)
      t += time.perf_counter() - tc
      m += O.m_edges(g, lev)
      if lanes_check is not None and not np.array_equal(lanes_check(l), lev):
        bad += 1
      log(f"cpu spmv root {l}: {t:.1f}s so far")
    out.update({"value": round(m / t / 1e9, 5),
                "sample": f"{spmv_roots} of {len(roots)} roots, oracle.spmv_bfs (the round loop "
                          f"as scipy boolean SpMV) over the same graph, {t:.1f}s",
                "levels_differ_from_gpu": bad})
    del M
  if bfo_roots:
    fwd = g.matrix().transpose().tocsr()
    t, m = 0.0, 0
    for l in range(bfo_roots):
      s = int(roots[l])
      tc = time.perf_counter()
      order = breadth_first_order(fwd, s, directed=True, return_predecessors=False)
      t += time.perf_counter() - tc
      visited = np.zeros(n, bool)
      visited[order] = True
      lev = np.where(visited, 0, -1).astype(np.int32)
      m += O.m_edges(g, lev)
    out["breadth_first_order"] = {
        "value": round(m / t / 1e9, 5), "unit": "GTEPS", "cores": 1,
        "sample": f"{bfo_roots} of {len(roots)} roots, scipy breadth_first_order on M^T "
                  f"(the reference's own oracle), {t:.1f}s"}
    log(f"cpu breadth_first_order: {out['breadth_first_order']['value']} GTEPS over {bfo_roots} roots")
  if "value" not in out and "breadth_first_order" in out:
    out["value"] = out["breadth_first_order"]["value"]
    out["sample"] = out["breadth_first_order"]["sample"]
  out.update(host_info())
  out["csl_simulator"] = csl_simulator_status()
  return out


def git_head():
  try:
    return subprocess.run(["git", "-C", HERE, "rev-parse", "--short=12", "HEAD"], capture_output=True,
                          text=True, check=False, timeout=10).stdout.strip() or None
  except (OSError, subprocess.SubprocessError):
    return None


def parse_args():
  p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
  p.add_argument("--gpus", type=int, default=1)
  p.add_argument("--steps", type=int, default=5)
  p.add_argument("--warmup", type=int, default=2)
  p.add_argument("--workload", choices=sorted(WORKLOADS), default="c3")
  p.add_argument("--cpu-roots", type=int, default=None,
                 help="roots timed by the SpMV CPU baseline (default: c3 1, c2 and c1 64)")
  p.add_argument("--cpu-bfo-roots", type=int, default=None,
                 help="roots timed by scipy breadth_first_order (default: c3 2, c2 and c1 64)")
  p.add_argument("--no-per-root", action="store_true", help="c3: skip the one-root-at-a-time leg")
  p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                 help="also run one device-to-device copy of MIB MiB (a known-size "
                      "reference kernel for scaling HBM byte counters)")
  return p.parse_args()


def per_root_leg(dg, roots, device_out, warm=2, capture=True):
  """One search per root.  capture: the per-level loop replays its captured
  HIP graph (the timed leg); without it every level is bracketed by events
  (the instrumented leg, whose level times give the roofline)."""
  import torch
  for s in roots[:warm]:
    dg.bfs(int(s), device_out=device_out, capture=capture)
  torch.cuda.synchronize()
  res = []
  tw = time.perf_counter()
  for s in roots:
    res.append(dg.bfs(int(s), device_out=device_out, capture=capture))
  torch.cuda.synchronize()
  return res, time.perf_counter() - tw


def per_root_summary(n, res, wall, inst=None):
  """Search times and m of the timed leg; the roofline of the dominant
  levels from the instrumented leg inst (the same roots, events per level),
  whose outputs must equal the timed leg's."""
  tms = np.array([r.stats.total_ms for r in res])
  m = np.array([r.stats.m_edges for r in res], dtype=np.float64)
  hm = harmonic_mean(m / (tms / 1e3) / 1e9)
  roof = per_root_roofline(n, inst if inst is not None else res)
  traffic, src = traffic_from_profile("bfs_pull_s24_traffic.json", "dominant_level_hbm_bytes")
  if n == 1 << 24:
    roof["traffic"], roof["traffic_source"] = traffic, src
  return {"hmean_gteps": round(hm, 3), "aggregate_gteps": round(m.sum() / wall / 1e9, 3),
          "ms_mean": round(float(tms.mean()), 4), "wall_s": round(wall, 4), "roofline": roof}, m, tms


def main():
  args = parse_args()
  if args.gpus != 1:
    print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
          "(independent root batches shard across GPUs with no communication; "
          "run one process per GPU)", file=sys.stderr)
    sys.exit(2)
  import torch
  from ceregraphs_amd.engine import DeviceGraph

  W = WORKLOADS[args.workload]
  t0 = time.perf_counter()
  dg = DeviceGraph.from_rmat(W["scale"], W["edgefactor"], W["seed"])
  if W.get("balance_grid"):
    from ceregraphs_amd import batch as B
    ip, ii, op, oi = dg.download()
    dg.close()
    part, _ = B.balance_part((len(ip) - 1, ip, ii, op, oi), W["balance_grid"], symmetric=True)
    dg = DeviceGraph.from_csr(part[0], part[3], part[4], symmetric=True)
  gen_s = time.perf_counter() - t0
  n = dg.n
  log(f"workload {args.workload}: n={n} nnz={dg.nnz} (built in {gen_s:.2f}s)")
  deg = dg.out_degree()
  cands = np.nonzero(deg > 0)[0]  # RMAT output has no self-loops
  roots = np.random.default_rng(W["root_seed"]).choice(cands, size=W["nroots"], replace=False)
  if "source" in W:
    roots = np.concatenate([[W["source"]], roots[roots != W["source"]][:W["nroots"] - 1]])
  spmv_roots = args.cpu_roots if args.cpu_roots is not None else (1 if args.workload == "c3" else 64)
  bfo_roots = args.cpu_bfo_roots if args.cpu_bfo_roots is not None else (2 if args.workload == "c3" else 64)
  head = git_head()

  if args.calibrate_copy:
    src = torch.ones(args.calibrate_copy << 18, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"calibration copy: {args.calibrate_copy} MiB read + {args.calibrate_copy} MiB written")
    del src, dst

  if args.workload in ("c1", "c2"):
    p1 = torch.empty(n, dtype=torch.int32, device="cuda")
    l1 = torch.empty(n, dtype=torch.int32, device="cuda")
    if args.workload == "c2":
      for _ in range(args.warmup):
        per_root_leg(dg, roots, (p1, l1), warm=0)
      res, wall = [], 0.0
      for _ in range(args.steps):
        r, w = per_root_leg(dg, roots, (p1, l1), warm=0)
        res, wall = res + r, wall + w
      inst, _ = per_root_leg(dg, roots, (p1, l1), warm=0, capture=False)
      summary, m, tms = per_root_summary(n, res, wall, inst)
      value = summary["hmean_gteps"]
      ms_step = wall / len(res) * 1e3
      metric = "harmonic-mean GTEPS, RMAT s21 ef16, 64 roots, one search per launch"
      cfg = {"workload": "bfs-rmat-s21-ef16-per-root", "roots": len(roots)}
    else:
      s = int(roots[0])
      for _ in range(args.warmup):
        dg.bfs(s, device_out=(p1, l1))
      res, wall = per_root_leg(dg, [s] * args.steps, (p1, l1), warm=0)
      inst, _ = per_root_leg(dg, [s], (p1, l1), warm=0, capture=False)
      summary, m, tms = per_root_summary(n, res, wall, inst)
      value = summary["hmean_gteps"]
      ms_step = wall / len(res) * 1e3
      metric = "GTEPS, RMAT s10 ef16 balanced 4x4, source 0"
      cfg = {"workload": "bfs-rmat-s10-ef16-balanced4x4-src0", "source": s}
    log(f"{args.workload}: {value:.3f} GTEPS hmean, {summary['ms_mean']:.4f} ms per search "
        f"(device), {ms_step:.4f} ms wall")
    lev_of = {}

    def lanes_check(l):
      if l not in lev_of:
        dg.bfs(int(roots[l]), device_out=(p1, l1))
        lev_of[l] = l1.cpu().numpy()
      return lev_of[l]
    cpu = cpu_baseline(dg, roots, spmv_roots, bfo_roots, lanes_check)
    out = {"metric": metric, "value": value, "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "u32",
           "data": "synthetic (RMAT generated on the GPU, edge-identical to gen_rmat.py)",
           "config": dict(cfg, scale=W["scale"], edgefactor=W["edgefactor"],
                          graph_seed=W["seed"], root_seed=W["root_seed"], n=n, nnz=dg.nnz),
           "roofline": summary["roofline"], "cpu_baseline": cpu, "per_root": summary,
           "head": head}
    print(json.dumps(out), flush=True)
    dg.close()
    return

  plane = len(roots) * n
  par = torch.empty(plane, dtype=torch.int32, device="cuda")
  lev = torch.empty(plane, dtype=torch.int16, device="cuda")
  # Timed passes decide every level on the device (the round kernels are
  # replayed from a captured HIP graph); one more pass with per-round host
  # polls and events gives the per-round breakdown and the pull roofline,
  # and must leave exactly the outputs of the timed passes.  The edge counts
  # m of the GTEPS figures are counted by the warm-up and the instrumented
  # passes, outside the timed region, as the reference counts them on the
  # host after the search (bfs_timing.py:379-398).
  m_warm = None
  for i in range(args.warmup):
    r = dg.msbfs(roots, device_out=(par, lev), capture=True)
    m_warm = r.m_lanes
    log(f"warmup {i}: {r.stats.total_ms:.3f} ms, rounds={r.stats.rounds_completed}")
  torch.cuda.synchronize()
  step_ms, results = [], []
  t1 = time.perf_counter()
  for _ in range(args.steps):
    r = dg.msbfs(roots, device_out=(par, lev), capture=True, count_m=False)
    results.append(r)
    step_ms.append(r.stats.total_ms)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t1
  ms_per_step = wall / args.steps * 1e3
  timed_par, timed_lev = par.clone(), lev.clone()
  last = dg.msbfs(roots, device_out=(par, lev))
  m = last.m_lanes.astype(np.float64)
  same = bool(torch.equal(timed_par, par) and torch.equal(timed_lev, lev)
              and (m_warm is None or np.array_equal(last.m_lanes, m_warm)))
  del timed_par, timed_lev
  if not same:
    log("WARNING: the instrumented pass differs from the timed passes")
  log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step wall, "
      f"{np.mean(step_ms):.3f} ms/step events; instrumented pass {last.stats.total_ms:.3f} ms, "
      f"rounds={last.stats.rounds_completed} dirs={last.stats.direction_history.tolist()}")

  t_step = ms_per_step / 1e3
  batched_hmean = harmonic_mean(m / (t_step / len(roots)) / 1e9)
  batched_aggregate = m.sum() / t_step / 1e9
  lane_hmean = harmonic_mean(m / t_step / 1e9)

  # Roofline of the dominant kernel: the multi-source pull (ms_pull_segs over
  # the rows of degree >= 256, then ms_pull over the 64-vertex chunks), timed
  # by the library's HIP events on the graph's stream.
  st = last.stats
  for i in range(int(st.rounds_completed)):
    log(f"  round {i}: dir {st.direction_history[i]} pull lanes {st.pull_lanes[i]:2d} "
        f"pull {st.pull_ms[i]:.3f} ms ({st.pull_edges[i]} entries) | push "
        f"{st.level_ms[i] - st.pull_ms[i]:.3f} ms ({st.edges_scanned[i] - st.pull_edges[i]} entries)"
        f" | nf {st.nf_history[i]} pairs {st.pairs_found[i]}")
  pr = np.nonzero(st.pull_lanes > 0)[0]
  # The graph is symmetric, so no in-edges means degree 0; roots have
  # degree >= 1, so every lane misses every such vertex.
  noin_pairs = len(roots) * int(np.count_nonzero(deg == 0))
  # the second pull round writes the outputs of the vertices without in-edges
  # (with a single pull round, ms_finalize does, outside the pull)
  noin_round = pr[1] if len(pr) > 1 else None
  byts = sum(pull_round_bytes(n, st.pull_rows[i], st.pull_edges[i], st.pull_pairs[i],
                              st.pull_nf[i], noin_pairs if i == noin_round else 0) for i in pr)
  tpull = float(st.pull_ms[pr].sum()) / 1e3
  npull = len(pr)
  achieved = byts / tpull / 1e9 if tpull > 0 else 0.0
  traffic, tsrc = traffic_from_profile("msbfs_pull_s24_traffic.json")
  roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
              "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
              "traffic_source": tsrc, "kernel": "ms_pull_segs+ms_pull", "launches": npull,
              "avg_launch_ms": round(tpull * 1e3 / max(npull, 1), 4),
              "algorithmic_bytes_per_launch": int(byts / max(npull, 1))}
  log(f"pull rounds: {npull}, {tpull*1e3:.3f} ms, {byts/1e9:.2f} GB algorithmic -> "
      f"{achieved:.0f} GB/s ({achieved/PEAK_HBM_GBS:.1%} of HBM peak)")

  extra = {}
  if not args.no_per_root:
    p1 = torch.empty(n, dtype=torch.int32, device="cuda")
    l1 = torch.empty(n, dtype=torch.int32, device="cuda")
    res, wall_roots = per_root_leg(dg, roots, (p1, l1))
    inst, _ = per_root_leg(dg, roots, (p1, l1), warm=0, capture=False)
    summary, ms_root, tms = per_root_summary(n, res, wall_roots, inst)
    if not np.array_equal(ms_root, m):
      log("WARNING: per-root m differs from the multi-source lanes")
    extra.update({
        "per_root_hmean_gteps": summary["hmean_gteps"],
        "per_root_aggregate_gteps": summary["aggregate_gteps"],
        "per_root_ms_mean": summary["ms_mean"],
        "per_root_wall_s": summary["wall_s"],
        "batched_vs_per_root_aggregate": round(batched_aggregate / summary["aggregate_gteps"], 2),
        "per_root": summary,
    })
    log(f"per-root: hmean {summary['hmean_gteps']:.2f} GTEPS, mean {summary['ms_mean']:.3f} ms/root, "
        f"aggregate {summary['aggregate_gteps']:.2f} GTEPS; dominant level "
        f"{summary['roofline']['achieved']:.0f} GB/s")

  cpu = None
  if spmv_roots or bfo_roots:
    cpu = cpu_baseline(dg, roots, spmv_roots, bfo_roots,
                       lambda l: lev[l * n:(l + 1) * n].cpu().numpy().astype(np.int32))
    log(f"cpu baseline: {cpu['value']} GTEPS")

  out = {
      "metric": METRIC_C3,
      "value": round(batched_hmean, 3),
      "unit": "GTEPS",
      "n_gpus": 1,
      "steps": args.steps,
      "warmup": args.warmup,
      "ms_per_step": round(ms_per_step, 4),
      "higher_is_better": True,
      "scaling": "weak",
      "vs_baseline": None,
      "dtype": "u64",
      "data": "synthetic (RMAT generated on the GPU, edge-identical to gen_rmat.py)",
      "config": {"workload": f"msbfs-rmat-s{W['scale']}-ef{W['edgefactor']}",
                 "scale": W["scale"], "edgefactor": W["edgefactor"], "graph_seed": W["seed"],
                 "root_seed": W["root_seed"], "nroots": int(len(roots)), "n": n,
                 "nnz": dg.nnz},
      "batched_aggregate_gteps": round(batched_aggregate, 3),
      "batched_lane_hmean_gteps": round(lane_hmean, 3),
      "timed_outputs_equal_instrumented": same,
      "msbfs_rounds": int(st.rounds_completed),
      "graph_gen_s": round(gen_s, 3),
      "roofline": roofline,
      "cpu_baseline": cpu,
      "head": head,
  }
  out.update(extra)
  print(json.dumps(out), flush=True)
  dg.close()


if __name__ == "__main__":
  main()
