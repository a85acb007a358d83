"""GTEPS accounting and the CSV schemas of the reference's drivers.

* ``compute_m_and_gteps`` -- m for one search and GTEPS
  (bfs/implementation/bfs_timing.py:379-398): for a symmetric matrix m is the
  undirected_dedup count #{(r, c) in M : visited[r], c <= r}; for a directed
  one the directed_source_visited count #{(r, c) in M : visited[c]}.
* ``harmonic_mean`` -- over finite positive rates (run_graph500.py:130-134).
* ``pick_sources`` -- roots with out-degree >= 1 excluding self-loops, drawn
  by ``default_rng(seed).choice(..., replace=False)`` (run_graph500.py:106-127).
* ``BFS_TIMING_COLUMNS`` -- the 37 columns of bfs_timing.csv in
  run_bfs.py:567-711 order, and the graph500 per-search / summary columns of
  run_graph500.py:410-465.

Times here are device timer readings: HIP events give nanoseconds, so the
"cycles" columns hold nanoseconds and clock_freq_hz is 1e9.
"""
import csv
import os

import numpy as np

CLOCK_FREQ_HZ = 1.0e9  # HIP event time is reported in ns resolution

BFS_TIMING_COLUMNS = [
    "timestamp", "infile_mtx", "n", "nnz", "pe_grid", "source", "channels",
    "rounds_completed", "max_rounds", "matrix_symmetric",
    "parent_resolve_min_cycles", "parent_resolve_max_cycles", "parent_resolve_avg_cycles",
    "h2d_matrix_span_cycles", "h2d_seed_span_cycles", "d2h_span_cycles",
    "round_duration_cycles",
    "local_compute_min_cycles", "local_compute_max_cycles", "local_compute_avg_cycles",
    "local_term_cond_min_cycles", "local_term_cond_max_cycles", "local_term_cond_avg_cycles",
    "direction_history", "nf_history",
    "transpose_min_cycles", "transpose_max_cycles", "transpose_avg_cycles",
    "search_time_cycles", "visited_count", "m_edges_traversed", "m_convention",
    "clock_freq_hz", "search_time_seconds", "gteps", "search_time_cycles_no_transfer",
    "gteps_no_transfer",
]

GRAPH500_SEARCH_COLUMNS = [
    "timestamp", "infile_mtx", "n", "nnz", "pe_grid", "channels", "matrix_symmetric",
    "search_idx", "source", "rounds_completed", "profiled_rounds", "max_rounds",
    "correctness_ok", "h2d_seed_span_cycles", "d2h_span_cycles", "search_time_cycles",
    "visited_count", "m_edges_traversed", "m_convention", "clock_freq_hz",
    "search_time_seconds", "gteps",
    "round_duration_cycles",
    "local_compute_min_cycles", "local_compute_max_cycles", "local_compute_avg_cycles",
    "local_term_cond_min_cycles", "local_term_cond_max_cycles", "local_term_cond_avg_cycles",
]

GRAPH500_SUMMARY_COLUMNS = [
    "timestamp", "infile_mtx", "n", "nnz", "pe_grid", "channels", "matrix_symmetric",
    "num_searches", "num_correctness_fail", "construction_time_seconds",
    "h2d_matrix_span_cycles", "harmonic_mean_gteps", "min_gteps", "median_gteps", "max_gteps",
]


def m_edges(rows, cols, visited, is_symmetric):
  """m of one search over the nonzeros (rows[k], cols[k]) of M."""
  visited = np.asarray(visited, dtype=bool)
  if is_symmetric:
    return int(np.count_nonzero(visited[rows] & (cols <= rows))), "undirected_dedup"
  return int(np.count_nonzero(visited[cols])), "directed_source_visited"


def compute_m_and_gteps(rows, cols, visited, is_symmetric, search_time_cycles,
                        clock_freq_hz=CLOCK_FREQ_HZ):
  m, conv = m_edges(rows, cols, visited, is_symmetric)
  seconds = search_time_cycles / clock_freq_hz
  gteps = m / seconds / 1e9 if seconds > 0 else float("nan")
  return m, conv, seconds, gteps


def gteps_from_ms(m, ms):
  return m / (ms / 1e3) / 1e9 if ms > 0 else float("nan")


def harmonic_mean(values):
  vals = [float(v) for v in values if v > 0 and np.isfinite(v)]
  if not vals:
    return float("nan")
  return len(vals) / sum(1.0 / v for v in vals)


def pick_sources(out_degree_no_loops, num_searches, seed, sources=None):
  """Explicit ``sources`` (a comma list) win; otherwise a seeded draw
  without replacement among vertices with out-degree >= 1."""
  n = len(out_degree_no_loops)
  if sources is not None:
    picked = [int(s) for s in str(sources).split(",")]
    for s in picked:
      if not 0 <= s < n:
        raise ValueError(f"--sources entry {s} out of range [0, {n})")
    return picked
  cands = np.nonzero(np.asarray(out_degree_no_loops) > 0)[0]
  num = min(num_searches, len(cands))
  rng = np.random.default_rng(seed)
  return rng.choice(cands, size=num, replace=False).tolist()


def ns(ms):
  """Milliseconds -> integer timer ticks (ns)."""
  return int(round(ms * 1e6))


def append_csv_row(path, row, columns=None):
  """Appends one row, writing the header for a new file and refusing to
  append to a file whose header differs (run_bfs.py:734-749)."""
  cols = list(columns or row.keys())
  os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
  write_header = not os.path.exists(path)
  if not write_header:
    with open(path, newline="", encoding="utf-8") as f:
      existing = next(csv.reader(f), [])
    if existing != cols:
      raise ValueError(f"{path}'s header does not match this run's columns; "
                       "use another --csv path")
  with open(path, "a", newline="", encoding="utf-8") as f:
    w = csv.DictWriter(f, fieldnames=cols)
    if write_header:
      w.writeheader()
    w.writerow(row)


def append_csv_rows(path, rows, columns):
  for r in rows:
    append_csv_row(path, r, columns)
