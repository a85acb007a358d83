"""Single-source BFS driver with the reference's command line and CSV row.

    python -m ceregraphs_amd.run_bfs --infile_mtx G.mtx --num_pe_cols 4 --num_pe_rows 4 \
        --source 0 [--directional] [--operm G.operm] [--csv out.csv] [--out-tree t.npz]

Mirrors bfs/scripts/run_bfs.py:79-755.  Cerebras-only flags (--fabric-dims,
--channels, --driver, --cmaddr, --arch, --latestlink, --compile-only,
--run-only, --width-*-buf, --dump-pe-timing, --pe-timing-out,
--parent-resolve-variant) are accepted and have no effect on the GPU; the PE
grid is only recorded in the pe_grid column.  The search runs in libcgbfs
(engine.DeviceGraph.bfs).  Differences, all additive:
  --operm      translate --source from an original id through a balancer's
               .operm (the reference leaves this to the user, docs/ERRORS.md:99-103);
  --direction  beamer | topdown | bottomup | tau (default: topdown, or tau
               with --directional, matching the reference's two modes);
  --out-tree   writes parents/levels as .npz (the reference renders a plot).
Parents follow the lowest-id-predecessor rule; the reference cross-check
(scipy breadth_first_order visited set + invalid_parents) is kept and made
stricter (a parent must also be exactly one level closer).
"""
import argparse
import os
import sys
import time
from datetime import datetime, timezone

import numpy as np
from scipy.sparse.csgraph import breadth_first_order

from . import graph_loader, timing
from .engine import DeviceGraph

DEFAULT_TAU_SWITCH_FRAC = 0.15  # run_bfs.py:76


def add_common_args(p):
  p.add_argument("--infile_mtx", required=True, help="the sparse matrix in MTX (or edge-list) format")
  p.add_argument("--num_pe_cols", type=int, required=True, help="recorded as pe_grid only")
  p.add_argument("--num_pe_rows", type=int, required=True, help="recorded as pe_grid only")
  for flag in ("--fabric-dims", "-d", "--cmaddr", "--latestlink"):
    p.add_argument(flag, default=None, help="accepted for compatibility; unused on the GPU")
  p.add_argument("--compile-only", action="store_true", help="build the library and exit")
  p.add_argument("--run-only", action="store_true", help="accepted; unused")
  p.add_argument("--width-west-buf", default=0, type=int)
  p.add_argument("--width-east-buf", default=0, type=int)
  p.add_argument("--channels", default=1, type=int)
  p.add_argument("--arch", choices=["wse3"], default="wse3")
  p.add_argument("--device", type=int, default=0, help="GPU index")
  p.add_argument("--operm", default=None, help="balancer .operm: --source is an original id")
  p.add_argument("--direction", choices=["beamer", "topdown", "bottomup", "tau"], default=None)
  p.add_argument("--no-cache", action="store_true", help="do not write/read the binary CSR cache")


def parse_args(argv=None):
  p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
  add_common_args(p)
  p.add_argument("--source", type=int, default=0)
  p.add_argument("--notree", action="store_true")
  p.add_argument("--notimings", action="store_true")
  p.add_argument("--nocorrectness", action="store_true")
  p.add_argument("--max-rounds", type=int, default=10,
                 help="rounds whose per-round timing is written (all rounds run)")
  p.add_argument("--directional", action="store_true",
                 help=f"one-way push->pull switch once nf > {DEFAULT_TAU_SWITCH_FRAC} * n")
  p.add_argument("--csv", default=None)
  p.add_argument("--out-tree", default=None, help="write parents/levels (.npz)")
  p.add_argument("--out-timing", default=None, help="accepted; no plot is drawn")
  p.add_argument("--no-show-parent-mismatch", dest="show_parent_mismatch", action="store_false")
  p.add_argument("--dump-pe-timing", action="store_true", help="accepted; unused")
  p.add_argument("--pe-timing-out", default=None, help="accepted; unused")
  p.add_argument("--parent-resolve-variant", choices=["dense", "indexed"], default="dense")
  return p.parse_args(argv)


def resolve_direction(args, n):
  if args.direction:
    d = args.direction
  else:
    d = "tau" if getattr(args, "directional", False) else "topdown"
  tau = round(DEFAULT_TAU_SWITCH_FRAC * n) if d == "tau" else None
  return d, tau


def check_search(hg, source, parents, levels):
  """The reference cross-check (run_bfs.py:513-549) plus the level rule:
  visited set vs scipy breadth_first_order on M^T, and every reached
  non-root vertex's parent must be reached, one level closer, and a real
  in-neighbour.  Returns (visited_mismatches, bad_parent_ids, scipy_diff)."""
  import scipy.sparse as sp
  n = hg.n
  M = sp.csr_matrix((np.ones(hg.nnz, np.int8), hg.in_idx, hg.in_ptr), shape=(n, n))
  order, pred = breadth_first_order(M.transpose().tocsr(), source, directed=True,
                                    return_predecessors=True)
  scipy_visited = np.zeros(n, dtype=bool)
  scipy_visited[order] = True
  visited = levels >= 0
  mismatch = int(np.count_nonzero(visited != scipy_visited))
  v = np.nonzero(visited)[0]
  v = v[v != source]
  p = parents[v].astype(np.int64)
  bad = np.zeros(len(v), dtype=bool)
  bad |= (p < 0) | (p >= n)
  ok = ~bad
  bad[ok] |= levels[p[ok]] != levels[v[ok]] - 1
  has_edge = np.zeros(len(v), dtype=bool)
  for k in np.nonzero(~bad)[0]:
    row = hg.in_idx[hg.in_ptr[v[k]]:hg.in_ptr[v[k] + 1]]
    i = np.searchsorted(row, p[k])
    has_edge[k] = i < len(row) and row[i] == p[k]
  bad |= ~has_edge
  scipy_parent = np.where(pred < 0, -1, pred).astype(np.int64)
  scipy_parent[source] = source
  diff = int(np.count_nonzero((parents != scipy_parent) & visited & scipy_visited &
                              (np.arange(n) != source)))
  return mismatch, v[bad].tolist(), diff


def timing_row(args, hg, source, res, h2d_ms):
  is_symmetric = hg.symmetric
  st = res.stats
  prof = min(st.rounds_completed, args.max_rounds)
  rd = [timing.ns(x) for x in st.level_ms[:prof]]
  rows, cols = hg.coo()
  visited = res.levels >= 0
  no_transfer = timing.ns(st.total_ms)
  d2h = timing.ns(st.output_ms)
  total = no_transfer + d2h
  m, conv, secs, gteps = timing.compute_m_and_gteps(rows, cols, visited, is_symmetric, total)
  gteps_nt = timing.compute_m_and_gteps(rows, cols, visited, is_symmetric, no_transfer)[3]
  zeros = ";".join("0" for _ in range(prof))
  join = lambda xs: ";".join(str(x) for x in xs)  # noqa: E731
  row = {
      "timestamp": datetime.now(timezone.utc).isoformat(),
      "infile_mtx": os.path.basename(args.infile_mtx),
      "n": hg.n, "nnz": hg.nnz, "pe_grid": f"{args.num_pe_cols}x{args.num_pe_rows}",
      "source": source, "channels": args.channels,
      "rounds_completed": st.rounds_completed, "max_rounds": args.max_rounds,
      "matrix_symmetric": is_symmetric,
      # parents are resolved inside the rounds: no separate resolve step
      "parent_resolve_min_cycles": 0, "parent_resolve_max_cycles": 0,
      "parent_resolve_avg_cycles": "0.0",
      "h2d_matrix_span_cycles": timing.ns(h2d_ms),
      "h2d_seed_span_cycles": 0,  # the seed is a kernel argument
      "d2h_span_cycles": d2h,
      "round_duration_cycles": join(rd),
      # one device: the per-PE min/max/avg of a round collapse to its time
      "local_compute_min_cycles": join(rd), "local_compute_max_cycles": join(rd),
      "local_compute_avg_cycles": ";".join(f"{x:.1f}" for x in rd),
      "local_term_cond_min_cycles": zeros, "local_term_cond_max_cycles": zeros,
      "local_term_cond_avg_cycles": ";".join("0.0" for _ in range(prof)),
      "direction_history": join(st.direction_history[:prof].tolist()),
      "nf_history": join(st.nf_history[:prof].tolist()),
      # CSC and CSR are both built at construction: no transpose step
      "transpose_min_cycles": zeros, "transpose_max_cycles": zeros,
      "transpose_avg_cycles": ";".join("0.0" for _ in range(prof)),
      "search_time_cycles": total, "visited_count": int(visited.sum()),
      "m_edges_traversed": m, "m_convention": conv, "clock_freq_hz": timing.CLOCK_FREQ_HZ,
      "search_time_seconds": secs, "gteps": gteps,
      "search_time_cycles_no_transfer": no_transfer, "gteps_no_transfer": gteps_nt,
  }
  return row


def main(argv=None):
  args = parse_args(argv)
  if args.num_pe_cols != args.num_pe_rows:
    sys.exit("--num_pe_cols and --num_pe_rows must match (square grid, run_bfs.py:179)")
  if args.compile_only:
    from . import build
    build.build()
    print("COMPILE ONLY: EXIT")
    return 0
  print(f"infile_mtx = {args.infile_mtx}, source = {args.source}, max_rounds = {args.max_rounds}")
  hg = graph_loader.load_csr_cached(args.infile_mtx, cache=not args.no_cache)
  n = hg.n
  source = args.source
  if args.operm:
    source = graph_loader.balanced_source(graph_loader.read_operm(args.operm), args.source)
    print(f"--operm: original source {args.source} -> balanced source {source}")
  if not 0 <= source < n:
    sys.exit(f"--source={source} out of range [0, {n})")
  print(f"Load matrix A, {n}-by-{n} with {hg.nnz} nonzeros (structural, boolean)")
  direction, tau = resolve_direction(args, n)
  if tau is not None:
    print(f"direction tau: tau_switch_count = {tau} ({DEFAULT_TAU_SWITCH_FRAC * 100:.0f}% of n={n})")

  t0 = time.perf_counter()
  dg = hg.to_device(args.device)
  h2d_ms = (time.perf_counter() - t0) * 1e3
  res = dg.bfs(source, direction=direction, tau=tau)
  st = res.stats
  parents, levels = res.parents, res.levels
  print(f"rounds_completed = {st.rounds_completed}, visited = {int((levels >= 0).sum())}/{n}, "
        f"device time {st.total_ms:.3f} ms, d2h {st.output_ms:.3f} ms")
  print(f"  direction per round (TD=top-down, BU=bottom-up): "
        f"{['BU' if d else 'TD' for d in st.direction_history]}")
  print(f"  nf per round: {st.nf_history.tolist()}")

  ok = True
  if not args.nocorrectness or not args.notree:
    mismatch, bad, diff = check_search(hg, source, parents, levels)
    ok = mismatch == 0 and not bad
    if not args.nocorrectness:
      print(f"[[ visited vs scipy mismatches: device={mismatch} ]]")
      print(f"[[ invalid device parents (edge, visited, one level closer): {len(bad)} ]]")
      if bad:
        print(f"  bad device parents at: {bad[:20]}{' ...' if len(bad) > 20 else ''}")
      print(f"[[ parent differs from scipy's own pick (lowest-id rule vs FIFO order): {diff} ]]")
      print(f"[[ scipy cross-check: {'OK' if ok else 'FAILED'} ]]")
  if args.out_tree:
    np.savez_compressed(args.out_tree, parents=parents, levels=levels, source=np.int64(source))
    print(f"wrote parents/levels to {args.out_tree}")

  if not args.notimings:
    row = timing_row(args, hg, source, res, h2d_ms)
    print(f"[[ GTEPS = {row['m_edges_traversed']} edges ({row['m_convention']}) / "
          f"{row['search_time_seconds'] * 1e6:.2f} us = {row['gteps']:.6f} GTEPS ]]")
    print(f"[[ GTEPS w/o d2h = {row['gteps_no_transfer']:.6f} ]]")
    csv_path = args.csv or os.path.join("results", "gpu", "bfs_timing.csv")
    timing.append_csv_row(csv_path, row, timing.BFS_TIMING_COLUMNS)
    print(f"appended timing row to {csv_path}")
  dg.close()
  return 0 if ok else 1


if __name__ == "__main__":
  sys.exit(main())
