"""The C ABI library loads and exports exactly what include/cgbfs.h declares;
argument checks fail cleanly before any device work (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import REPO
from ceregraphs_amd import _lib

HEADER = os.path.join(REPO, "include", "cgbfs.h")


def _ensure_built():
  if not os.path.exists(_lib.LIB_PATH):
    from ceregraphs_amd import build
    build.build(verbose=False)


def declared_symbols():
  text = open(HEADER, encoding="utf-8").read()
  return set(re.findall(r"CGBFS_API\s+[\w\s\*]*?\b(cgbfs_\w+)\s*\(", text))


def test_header_and_binding_agree():
  assert declared_symbols() == set(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
  _ensure_built()
  lib = ctypes.CDLL(_lib.LIB_PATH)
  for name in declared_symbols():
    assert hasattr(lib, name), name
  assert _lib.load().cgbfs_version() >= 1


def test_invalid_arguments_are_rejected():
  _ensure_built()
  lib = _lib.load()
  h = ctypes.c_void_p()
  ptr = (ctypes.c_int64 * 2)(0, 0)
  rc = lib.cgbfs_graph_create(0, 0, 0, ctypes.cast(ptr, ctypes.c_void_p), None, None, None, 1,
                              ctypes.byref(h))
  assert rc == _lib.CGBFS_EINVAL
  assert b"n must be" in lib.cgbfs_last_error()
  rc = lib.cgbfs_graph_create(0, 4, 0, None, None, None, None, 1, ctypes.byref(h))
  assert rc == _lib.CGBFS_EINVAL
  rc = lib.cgbfs_bfs(None, 0, None, None, 0, None, None)
  assert rc == _lib.CGBFS_EINVAL
  rc = lib.cgbfs_msbfs(None, None, 0, None, None, 0, None, None, None)
  assert rc == _lib.CGBFS_EINVAL
  with pytest.raises(_lib.CgbfsError):
    _lib.check(rc)


def test_options_and_stats_layout_match_header():
  text = open(HEADER, encoding="utf-8").read()
  opts = re.search(r"typedef struct \{([^}]*)\} cgbfs_options;", text, re.S).group(1)
  names = re.findall(r"\b(\w+);", opts)
  assert names == [f for f, _ in _lib.Options._fields_]
  stats = re.search(r"typedef struct \{([^}]*)\} cgbfs_stats;", text, re.S).group(1)
  names = re.findall(r"\*?(\w+);", stats)
  assert names == [f for f, _ in _lib.Stats._fields_]


def _csr_args(ptr, idx):
  ptr = np.ascontiguousarray(ptr, np.int64)
  idx = np.ascontiguousarray(idx, np.int32)
  return ptr, idx, ctypes.c_void_p(ptr.ctypes.data), ctypes.c_void_p(idx.ctypes.data)


def _check(n, out, inn=None, symmetric=True):
  """cgbfs_graph_check on (ptr, idx) host views; returns (rc, message)."""
  lib = _lib.load()
  op, oi, opp, oip = _csr_args(*out)
  if inn is None:
    rc = lib.cgbfs_graph_check(n, len(oi), opp, oip, None, None, 1 if symmetric else 0)
  else:
    ip, ii, ipp, iip = _csr_args(*inn)
    rc = lib.cgbfs_graph_check(n, len(oi), opp, oip, ipp, iip, 1 if symmetric else 0)
  return rc, lib.cgbfs_last_error().decode()


# a 4-vertex undirected graph: edges 0-1, 0-2, 2-3
SYM = ([0, 2, 3, 5, 6], [1, 2, 0, 0, 3, 2])
# a directed graph 0->1, 0->2, 2->3: out view (column c of M = out-neighbours
# of c) and in view (row r = in-neighbours of r)
DIR_OUT = ([0, 2, 2, 3, 3], [1, 2, 3])
DIR_IN = ([0, 0, 1, 2, 3], [0, 0, 2])


def test_graph_check_accepts_valid_graphs():
  _ensure_built()
  assert _check(4, SYM)[0] == _lib.CGBFS_OK
  assert _check(4, DIR_OUT, DIR_IN, symmetric=False)[0] == _lib.CGBFS_OK
  assert _check(3, ([0, 0, 0, 0], []))[0] == _lib.CGBFS_OK  # no edges


@pytest.mark.parametrize("case,ptr,idx,text", [
    ("decreasing ptr", [0, 2, 1, 5, 6], [1, 2, 0, 0, 3, 2], "never decrease"),
    ("ptr not from 0", [1, 2, 3, 5, 6], [1, 2, 0, 0, 3, 2], "start at 0"),
    ("ptr end != nnz", [0, 2, 3, 5, 5], [1, 2, 0, 0, 3, 2], "end at nnz"),
    ("id out of range", [0, 2, 3, 5, 6], [1, 4, 0, 0, 3, 2], "outside [0, n)"),
    ("negative id", [0, 2, 3, 5, 6], [1, -1, 0, 0, 3, 2], "outside [0, n)"),
    ("unsorted row", [0, 2, 3, 5, 6], [2, 1, 0, 0, 3, 2], "strictly ascending"),
    ("duplicate id", [0, 2, 3, 5, 6], [1, 1, 0, 0, 3, 2], "strictly ascending"),
    ("not symmetric", [0, 2, 3, 5, 6], [1, 2, 0, 1, 3, 2], "transposes"),
])
def test_graph_check_rejects_malformed(case, ptr, idx, text):
  """Each defect is caught on the host, before any GPU work: by
  cgbfs_graph_check, and by cgbfs_graph_create on host arrays (which runs
  the same check first, so this needs no GPU either)."""
  _ensure_built()
  nnz = 6
  rc, msg = _check(4, (ptr, idx))
  assert rc == _lib.CGBFS_EINVAL and text in msg, (case, msg)
  p, i, pp, ip = _csr_args(ptr, idx)
  h = ctypes.c_void_p()
  lib = _lib.load()
  rc = lib.cgbfs_graph_create(0, 4, nnz, pp, ip, None, None, 1, ctypes.byref(h))
  assert rc == _lib.CGBFS_EINVAL and text in lib.cgbfs_last_error().decode(), case
  assert not h.value


def test_graph_check_directed_views_must_agree():
  _ensure_built()
  bad_in = ([0, 0, 1, 2, 3], [0, 0, 1])  # 1->3 instead of 2->3
  rc, msg = _check(4, DIR_OUT, bad_in, symmetric=False)
  assert rc == _lib.CGBFS_EINVAL and "transposes" in msg
  rc, msg = _check(4, DIR_OUT, ([0, 0, 1, 3, 3], [0, 0, 2]), symmetric=False)
  assert rc == _lib.CGBFS_EINVAL  # in view with its own defect (ptr end)


def test_graph_check_on_generated_graphs():
  """Graphs from the generator (undirected and directed) pass; the same
  graph with one in-row reversed fails."""
  _ensure_built()
  from ceregraphs_amd import gen_rmat
  n, ip, ii, op, oi = gen_rmat.rmat_csr(10, 8, 3)
  assert _check(n, (op, oi))[0] == _lib.CGBFS_OK
  n, ip, ii, op, oi = gen_rmat.rmat_csr(10, 8, 4, 0.45, 0.25, 0.15, directed=True)
  assert _check(n, (op, oi), (ip, ii), symmetric=False)[0] == _lib.CGBFS_OK
  v = int(np.argmax(np.diff(ip)))
  ii2 = ii.copy()
  ii2[ip[v]:ip[v + 1]] = ii2[ip[v]:ip[v + 1]][::-1]
  rc, msg = _check(n, (op, oi), (ip, ii2), symmetric=False)
  assert rc == _lib.CGBFS_EINVAL and "strictly ascending" in msg
