import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
if REPO not in sys.path:
  sys.path.insert(0, REPO)


def pytest_configure(config):
  config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and libcgbfs.so")


def golden(name):
  return np.load(os.path.join(GOLDEN, name))


@pytest.fixture(scope="session")
def c1_graph():
  """Configuration 1: s10 ef16 seed 0, balanced 4x4 by the reference's
  util/analyze --symmetric (tests/golden/make_golden.py)."""
  from oracle import oracle as O
  g = golden("c1_balanced_4x4.npz")
  return O.Graph(int(g["n"]), g["indptr"], g["indices"], g["indptr"], g["indices"], True)


def oracle_graph_from_csr(n, in_ptr, in_idx, out_ptr, out_idx, symmetric):
  from oracle import oracle as O
  return O.Graph(n, in_ptr, in_idx, out_ptr, out_idx, symmetric)
