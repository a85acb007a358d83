"""The C ABI library loads and exports exactly what include/cgbfs.h declares;
argument checks fail cleanly before any device work (no GPU needed)."""
import ctypes
import os
import re

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
