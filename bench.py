#!/usr/bin/env python3
"""Headline benchmark: harmonic-mean GTEPS on RMAT scale-24 edgefactor-16
with 64 roots -- batched multi-source BFS against the same roots run one by
one -- on one MI355X.

A step is one 64-lane multi-source pass over the 64 roots, inputs (graph)
already resident in HBM, outputs parents int32[64][V] and levels int16[64][V]
written to HBM.  GTEPS accounting follows the reference: m per search is the
undirected_dedup count of bfs/implementation/bfs_timing.py:390-392, roots are
drawn as in bfs/scripts/run_graph500.py:106-127 (non-isolated vertices,
default_rng(seed).choice without replacement) and the mean is the harmonic
mean of run_graph500.py:130-134.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0


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
  8 B next word per newly active vertex (16 N).  The first pull round also
  writes the -1 outputs of every (lane, vertex) pair whose vertex has no
  in-edges (6 B each, noin_pairs of them)."""
  return 8 * n + 8 * rows + 12 * edges + 6 * pairs + 16 * nf + 6 * noin_pairs


def parse_args():
  p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
  p.add_argument("--gpus", type=int, default=1)
  p.add_argument("--steps", type=int, default=5)
  p.add_argument("--warmup", type=int, default=2)
  p.add_argument("--scale", type=int, default=24)
  p.add_argument("--edgefactor", type=int, default=16)
  p.add_argument("--seed", type=int, default=3, help="graph seed")
  p.add_argument("--root-seed", type=int, default=4)
  p.add_argument("--nroots", type=int, default=64)
  p.add_argument("--cpu-roots", type=int, default=1,
                 help="roots timed for the CPU baseline (0 disables it)")
  p.add_argument("--no-per-root", action="store_true", help="skip the one-root-at-a-time leg")
  p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                 help="also run one device-to-device copy of MIB MiB (a known-size "
                      "reference kernel for scaling HBM byte counters)")
  return p.parse_args()


def main():
  args = parse_args()
  if args.gpus != 1:
    print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
          "(independent root batches shard across GPUs with no communication; "
          "run one process per GPU)", file=sys.stderr)
    sys.exit(2)
  import torch
  from ceregraphs_amd.engine import DeviceGraph

  t0 = time.perf_counter()
  dg = DeviceGraph.from_rmat(args.scale, args.edgefactor, args.seed)
  gen_s = time.perf_counter() - t0
  n = dg.n
  log(f"graph s{args.scale} ef{args.edgefactor} seed{args.seed}: n={n} nnz={dg.nnz} "
      f"(generated on GPU in {gen_s:.2f}s)")
  deg = dg.out_degree()
  cands = np.nonzero(deg > 0)[0]  # RMAT output has no self-loops
  roots = np.random.default_rng(args.root_seed).choice(cands, size=args.nroots, replace=False)

  if args.calibrate_copy:
    src = torch.ones(args.calibrate_copy << 18, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"calibration copy: {args.calibrate_copy} MiB read + {args.calibrate_copy} MiB written")
    del src, dst
  par = torch.empty(args.nroots * n, dtype=torch.int32, device="cuda")
  lev = torch.empty(args.nroots * n, dtype=torch.int16, device="cuda")
  # Timed passes decide every level on the device (the round kernels are
  # replayed from a captured HIP graph); one more pass with per-round host
  # polls and events gives the per-round breakdown and the pull roofline.
  for i in range(args.warmup):
    r = dg.msbfs(roots, device_out=(par, lev), capture=True)
    log(f"warmup {i}: {r.stats.total_ms:.3f} ms, rounds={r.stats.rounds_completed}")
  torch.cuda.synchronize()
  step_ms, results = [], []
  t1 = time.perf_counter()
  for _ in range(args.steps):
    r = dg.msbfs(roots, device_out=(par, lev), capture=True)
    results.append(r)
    step_ms.append(r.stats.total_ms)
  torch.cuda.synchronize()
  wall = time.perf_counter() - t1
  ms_per_step = wall / args.steps * 1e3
  m = results[-1].m_lanes.astype(np.float64)
  last = dg.msbfs(roots, device_out=(par, lev))
  if not np.array_equal(last.m_lanes, results[-1].m_lanes):
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
  byts = sum(pull_round_bytes(n, st.pull_rows[i], st.pull_edges[i], st.pull_pairs[i],
                              st.pull_nf[i], noin_pairs if i == pr[0] else 0) for i in pr)
  tpull = float(st.pull_ms[pr].sum()) / 1e3
  npull = len(pr)
  achieved = byts / tpull / 1e9 if tpull > 0 else 0.0
  traffic = None
  prof = os.path.join(HERE, "profiles", "msbfs_pull_s24_traffic.json")
  if os.path.exists(prof):
    with open(prof, encoding="utf-8") as f:
      traffic = json.load(f).get("hbm_bytes_per_launch")
  roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
              "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
              "kernel": "ms_pull_segs+ms_pull", "launches": npull,
              "avg_launch_ms": round(tpull * 1e3 / max(npull, 1), 4),
              "algorithmic_bytes_per_launch": int(byts / max(npull, 1))}
  log(f"pull rounds: {npull}, {tpull*1e3:.3f} ms, {byts/1e9:.2f} GB algorithmic -> "
      f"{achieved:.0f} GB/s ({achieved/PEAK_HBM_GBS:.1%} of HBM peak)")

  extra = {}
  if not args.no_per_root:
    p1 = torch.empty(n, dtype=torch.int32, device="cuda")
    l1 = torch.empty(n, dtype=torch.int32, device="cuda")
    for s in roots[:2]:
      dg.bfs(int(s), device_out=(p1, l1))
    tms, ms_root = [], []
    torch.cuda.synchronize()
    tw = time.perf_counter()
    for s in roots:
      rr = dg.bfs(int(s), device_out=(p1, l1))
      tms.append(rr.stats.total_ms)
      ms_root.append(rr.stats.m_edges)
    torch.cuda.synchronize()
    wall_roots = time.perf_counter() - tw
    tms = np.array(tms)
    ms_root = np.array(ms_root, dtype=np.float64)
    if not np.array_equal(ms_root, m):
      log("WARNING: per-root m differs from the multi-source lanes")
    per_root_hmean = harmonic_mean(ms_root / (tms / 1e3) / 1e9)
    per_root_aggregate = ms_root.sum() / wall_roots / 1e9
    extra.update({
        "per_root_hmean_gteps": round(per_root_hmean, 3),
        "per_root_aggregate_gteps": round(per_root_aggregate, 3),
        "per_root_ms_mean": round(float(tms.mean()), 4),
        "per_root_wall_s": round(wall_roots, 4),
        "batched_vs_per_root_aggregate": round(batched_aggregate / per_root_aggregate, 2),
    })
    log(f"per-root: hmean {per_root_hmean:.2f} GTEPS, mean {tms.mean():.3f} ms/root, "
        f"aggregate {per_root_aggregate:.2f} GTEPS over {wall_roots:.3f}s")

  cpu = None
  if args.cpu_roots > 0:
    from oracle import oracle as O
    ip, ii, op, oi = dg.download()
    g = O.Graph(n, ip, ii, op, oi, True)
    M = g.matrix().astype(np.int32)
    cpu_t, cpu_m = 0.0, 0
    for l in range(args.cpu_roots):
      s = int(roots[l])
      tc = time.perf_counter()
      lev_cpu, _, _, _ = O.spmv_bfs(g, s, matrix=M)
      cpu_t += time.perf_counter() - tc
      cpu_m += O.m_edges(g, lev_cpu)
      lane = lev[l * n:(l + 1) * n].cpu().numpy()
      if not np.array_equal(lane, lev_cpu.astype(np.int16)):
        log(f"WARNING: CPU levels differ from lane {l}")
    cpu = {"value": round(cpu_m / cpu_t / 1e9, 5), "unit": "GTEPS", "cores": 1, "kind": "port",
           "sample": f"{args.cpu_roots} of {args.nroots} roots, scipy boolean-SpMV BFS "
                     f"(oracle.spmv_bfs) over the same s{args.scale} graph, "
                     f"{cpu_t:.1f}s on {os.cpu_count()} visible host cores (1 used)"}
    log(f"cpu baseline: {cpu['value']} GTEPS in {cpu_t:.1f}s")

  out = {
      "metric": "harmonic-mean GTEPS, RMAT s24 ef16, 64 roots: batched MS-BFS vs per-root",
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
      "config": {"workload": f"msbfs-rmat-s{args.scale}-ef{args.edgefactor}",
                 "scale": args.scale, "edgefactor": args.edgefactor, "graph_seed": args.seed,
                 "root_seed": args.root_seed, "nroots": int(len(roots)), "n": n,
                 "nnz": dg.nnz},
      "batched_aggregate_gteps": round(batched_aggregate, 3),
      "batched_lane_hmean_gteps": round(lane_hmean, 3),
      "msbfs_rounds": int(st.rounds_completed),
      "graph_gen_s": round(gen_s, 3),
      "roofline": roofline,
      "cpu_baseline": cpu,
  }
  out.update(extra)
  print(json.dumps(out), flush=True)
  dg.close()


if __name__ == "__main__":
  main()
