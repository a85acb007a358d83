"""Graph500-style harness: one graph upload, many searches, per-search and
summary CSVs, harmonic-mean GTEPS.

    python -m ceregraphs_amd.run_graph500 --infile_mtx G.mtx --num_pe_cols 4 --num_pe_rows 4 \
        [--num-searches 64] [--seed 0] [--sources a,b,c] [--batched] [--csv s.csv] \
        [--summary-csv sum.csv]

Mirrors bfs/scripts/run_graph500.py:64-509: roots by pick_sources
(:106-127), construction (the upload) timed once and kept out of every
search's time (:284-315), one search per root with its own time (:321-440),
the scipy visited-set + parent cross-check per root (:382-395), and the
harmonic mean over searches (:446-448).  ``--batched`` runs the roots as
64-lane multi-source passes instead; each search is then charged its share
(batch time / roots in the batch) of the pass that produced it.
Default direction is top-down, as in the reference harness (which has no
--directional, :405); --direction picks another policy (results are the
same under every policy).
"""
import argparse
import os
import sys
import time
from datetime import datetime, timezone

import numpy as np

from . import graph_loader, timing
from .run_bfs import add_common_args, check_search


def parse_args(argv=None):
  p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
  add_common_args(p)
  p.add_argument("--num-searches", type=int, default=64)
  p.add_argument("--seed", type=int, default=0)
  p.add_argument("--sources", default=None)
  p.add_argument("--max-rounds", type=int, default=15)
  p.add_argument("--nocorrectness", action="store_true")
  p.add_argument("--csv", default=None)
  p.add_argument("--summary-csv", default=None)
  p.add_argument("--batched", action="store_true", help="64-lane multi-source passes")
  return p.parse_args(argv)


def main(argv=None):
  args = parse_args(argv)
  if args.num_pe_cols != args.num_pe_rows:
    sys.exit("--num_pe_cols and --num_pe_rows must match (square grid)")
  print(f"infile_mtx = {args.infile_mtx}, num_searches = {args.num_searches}, "
        f"max_rounds = {args.max_rounds}")
  hg = graph_loader.load_csr_cached(args.infile_mtx, cache=not args.no_cache)
  n = hg.n
  print(f"Load matrix A, {n}-by-{n} with {hg.nnz} nonzeros (structural, boolean)")
  if not hg.symmetric:
    print("[[ NOTE: A is not symmetric -- m uses the directed_source_visited rule ]]")
  sources = timing.pick_sources(hg.out_degree_no_loops(), args.num_searches, args.seed, args.sources)
  if args.operm:
    perm = graph_loader.read_operm(args.operm)
    sources = [graph_loader.balanced_source(perm, s) for s in sources]
  print(f"selected {len(sources)} search roots (seed={args.seed}): {sources}")
  direction = args.direction or "topdown"

  t0 = time.perf_counter()
  dg = hg.to_device(args.device)
  construction_s = time.perf_counter() - t0
  print(f"construction (upload to HBM): {construction_s * 1e6:.2f} us -- not part of any search's time")
  rows, cols = hg.coo()

  results = []  # (source, parents, levels, search_ms, d2h_ms, rounds, hist)
  if args.batched:
    for b0 in range(0, len(sources), 64):
      batch = sources[b0:b0 + 64]
      r = dg.msbfs(batch, direction=direction)
      share = r.stats.total_ms / len(batch)
      d2h = r.stats.output_ms / len(batch)
      for l, s in enumerate(batch):
        results.append((s, r.parents[l], r.levels[l].astype(np.int32), share, d2h,
                        None, r.stats))
  else:
    for s in sources:
      r = dg.bfs(s, direction=direction, capture=True)  # level loop replayed from a HIP graph
      results.append((s, r.parents, r.levels, r.stats.total_ms, r.stats.output_ms,
                      r.stats.rounds_completed, r.stats))

  search_rows, fails = [], 0
  for i, (s, par, lev, ms, d2h_ms, rounds, st) in enumerate(results):
    ok = None
    if not args.nocorrectness:
      mismatch, bad, _ = check_search(hg, s, par, lev)
      ok = mismatch == 0 and not bad
      if not ok:
        fails += 1
        print(f"[[ search {i} (source={s}): CORRECTNESS FAILED -- visited mismatches={mismatch}, "
              f"invalid parents={len(bad)} ]]")
    visited = lev >= 0
    if rounds is None:  # a lane of a batch: its own depth decides its rounds
      rounds = int(lev.max()) + 1
    total = timing.ns(ms) + timing.ns(d2h_ms)
    m, conv, secs, gteps = timing.compute_m_and_gteps(rows, cols, visited, hg.symmetric, total)
    prof = min(rounds, args.max_rounds)
    if args.batched:
      rd = []
    else:
      rd = [timing.ns(x) for x in st.level_ms[:prof]]
    join = lambda xs: ";".join(str(x) for x in xs)  # noqa: E731
    zeros = ";".join("0" for _ in rd)
    row = {
        "timestamp": datetime.now(timezone.utc).isoformat(),
        "infile_mtx": os.path.basename(args.infile_mtx), "n": n, "nnz": hg.nnz,
        "pe_grid": f"{args.num_pe_cols}x{args.num_pe_rows}", "channels": args.channels,
        "matrix_symmetric": hg.symmetric, "search_idx": i, "source": s,
        "rounds_completed": rounds, "profiled_rounds": prof if rd else 0,
        "max_rounds": args.max_rounds, "correctness_ok": ok, "h2d_seed_span_cycles": 0,
        "d2h_span_cycles": timing.ns(d2h_ms), "search_time_cycles": total,
        "visited_count": int(visited.sum()), "m_edges_traversed": m, "m_convention": conv,
        "clock_freq_hz": timing.CLOCK_FREQ_HZ, "search_time_seconds": secs, "gteps": gteps,
        "round_duration_cycles": join(rd),
        "local_compute_min_cycles": join(rd), "local_compute_max_cycles": join(rd),
        "local_compute_avg_cycles": ";".join(f"{x:.1f}" for x in rd),
        "local_term_cond_min_cycles": zeros, "local_term_cond_max_cycles": zeros,
        "local_term_cond_avg_cycles": ";".join("0.0" for _ in rd),
    }
    search_rows.append(row)
    print(f"search {i + 1}/{len(results)}: source={s} visited={row['visited_count']}/{n} "
          f"rounds={rounds} m={m} time={secs * 1e6:.2f}us gteps={gteps:.6f}"
          + ("" if ok is None else (" OK" if ok else " CORRECTNESS-FAILED")))

  g = [r["gteps"] for r in search_rows]
  fin = [x for x in g if np.isfinite(x) and x > 0]
  hmean = timing.harmonic_mean(g)
  summary = {
      "timestamp": datetime.now(timezone.utc).isoformat(),
      "infile_mtx": os.path.basename(args.infile_mtx), "n": n, "nnz": hg.nnz,
      "pe_grid": f"{args.num_pe_cols}x{args.num_pe_rows}", "channels": args.channels,
      "matrix_symmetric": hg.symmetric, "num_searches": len(search_rows),
      "num_correctness_fail": fails, "construction_time_seconds": construction_s,
      "h2d_matrix_span_cycles": timing.ns(construction_s * 1e3),
      "harmonic_mean_gteps": hmean,
      "min_gteps": min(fin) if fin else float("nan"),
      "median_gteps": float(np.median(fin)) if fin else float("nan"),
      "max_gteps": max(fin) if fin else float("nan"),
  }
  print(f"\n[[ {len(search_rows)} searches complete, {fails} correctness FAILURES ]]")
  print(f"[[ GTEPS across searches: harmonic_mean={hmean:.6f}, min={summary['min_gteps']:.6f}, "
        f"median={summary['median_gteps']:.6f}, max={summary['max_gteps']:.6f} ]]")
  csv_path = args.csv or os.path.join("results", "gpu", "graph500_searches.csv")
  timing.append_csv_rows(csv_path, search_rows, timing.GRAPH500_SEARCH_COLUMNS)
  sum_path = args.summary_csv or os.path.join("results", "gpu", "graph500_summary.csv")
  timing.append_csv_row(sum_path, summary, timing.GRAPH500_SUMMARY_COLUMNS)
  print(f"appended {len(search_rows)} per-search rows to {csv_path} and a summary row to {sum_path}")
  dg.close()
  return 0 if fails == 0 else 1


if __name__ == "__main__":
  sys.exit(main())
