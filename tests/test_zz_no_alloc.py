"""cgbfs_bfs_noalloc (the single-root search whose workspace is a file-scope
device array) against cgbfs_bfs (the same search on the graph's own,
allocated workspace), called through the C ABI.

Two calls in a row share the one device array: a graph of exactly the
largest size it takes (2^21 vertices, the s21 ef16 seed 1 graph), then a
smaller, directed one.  Outputs are pre-filled with a sentinel, so every
entry must be written by the search.  Levels, parents, round counts, nf per
round, m and the visited count must be identical; the small graph is also
checked against the oracle.  A graph above the limit is refused."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from ceregraphs_amd import _lib, gen_rmat
from ceregraphs_amd.engine import DeviceGraph, _StatsBuf, _options

pytestmark = pytest.mark.gpu


def _both(dg, source, **opts):
  """(levels, parents, stats) from cgbfs_bfs and from cgbfs_bfs_noalloc."""
  import torch
  lib = _lib.load()
  out = []
  for fn in (lib.cgbfs_bfs, lib.cgbfs_bfs_noalloc):
    par = torch.full((dg.n,), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
    lev = torch.full((dg.n,), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    sb = _StatsBuf()
    opt = _options(**opts)
    args = [dg._h, int(source), ctypes.c_void_p(par.data_ptr()), ctypes.c_void_p(lev.data_ptr())]
    if fn is lib.cgbfs_bfs:
      args.append(1)
    _lib.check(fn(*args, ctypes.byref(opt), ctypes.byref(sb.s)))
    out.append((lev.cpu().numpy(), par.cpu().numpy(), sb.result()))
  return out


def _same(a, b, what):
  (la, pa, sa), (lb, pb, sb) = a, b
  np.testing.assert_array_equal(lb, la, err_msg=f"levels {what}")
  np.testing.assert_array_equal(pb, pa, err_msg=f"parents {what}")
  assert sb.rounds_completed == sa.rounds_completed, what
  np.testing.assert_array_equal(sb.nf_history, sa.nf_history, err_msg=what)
  np.testing.assert_array_equal(sb.direction_history, sa.direction_history, err_msg=what)
  assert (sb.m_edges, sb.visited_count) == (sa.m_edges, sa.visited_count), what
  assert sb.total_ms > 0 and np.all(sb.level_ms >= 0)


def test_noalloc_equals_cgbfs_bfs_two_calls_second_smaller():
  # call 1: the largest graph the device array holds (2^21 vertices)
  with DeviceGraph.from_rmat(21, 16, 1) as dg:
    assert dg.n == 1 << 21
    deg = dg.out_degree()
    roots = np.random.default_rng(2).choice(np.nonzero(deg)[0], 2, replace=False)
    for s in roots:
      a, b = _both(dg, int(s))
      _same(a, b, f"s21 root {s}")
  # call 2: a smaller, directed graph, after the larger one left its state
  abcd = (0.45, 0.25, 0.15, 0.15)
  n, ip, ii, op, oi = gen_rmat.rmat_csr(12, 8, 10, *abcd[:3], directed=True)
  g = O.Graph(n, ip, ii, op, oi, False)
  with DeviceGraph.from_rmat(12, 8, 10, abcd=abcd, directed=True) as dg:
    s = int(np.argmax(np.diff(op)))
    for direction in ("beamer", "bottomup"):
      a, b = _both(dg, s, direction=direction)
      _same(a, b, f"s12 {direction}")
      lev, par, hist, rounds = O.bfs(g, s)
      np.testing.assert_array_equal(b[0], lev)
      np.testing.assert_array_equal(b[1], par)
      np.testing.assert_array_equal(b[2].nf_history, hist)


def test_noalloc_refuses_graphs_above_its_size():
  import torch
  with DeviceGraph.from_rmat(22, 4, 3) as dg:
    par = torch.empty(dg.n, dtype=torch.int32, device="cuda")
    lev = torch.empty(dg.n, dtype=torch.int32, device="cuda")
    rc = _lib.load().cgbfs_bfs_noalloc(dg._h, 0, ctypes.c_void_p(par.data_ptr()),
                                       ctypes.c_void_p(lev.data_ptr()), None, None)
    assert rc == _lib.CGBFS_EINVAL
    assert b"2^21" in _lib.load().cgbfs_last_error()
