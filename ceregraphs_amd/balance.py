"""Vertex-identity-preserving load balancing for a P x P PE grid.

In-memory front end of libcgbalance.so (include/cgbalance.h), the native
restatement of the reference's util/analyze.cpp balancer.  The file front
end, with analyze's flags and output files, is the ``cgbalance`` executable
next to this module:

    ceregraphs_amd/cgbalance --matrix G.mtx --symmetric --fabx 4 --faby 4 \\
        --omatrix G.4x4.mtx --operm G.4x4.operm

The permutation q (the ``.operm``) maps an original vertex id to its id in the
balanced matrix B = A(q, q); it covers the padded vertex count P * ceil(n/P)
(analyze.cpp:545-554).  Rows and columns share q, so a BFS source keeps its
identity: source_balanced = q[source_original].

Matrix orientation is the Matrix Market file's: row r, column c.  In this
package's graph convention (graph_loader.py: M[dst][src]) the rows are the
in-view, so a graph is balanced from its CSR(M) = in-view arrays.
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcgbalance.so")
CLI_PATH = os.path.join(HERE, "cgbalance")

SYMMETRIC, SHARED_PERM, REMOVE_DUP = 1, 2, 4
OK, EINVAL, ENOMEM, ENOTSYM = 0, 1, 2, 3
EXPORTED = ("cgbal_padded_n", "cgbal_balance_coo", "cgbal_balance_csr", "cgbal_last_error")


class Report(ctypes.Structure):
  _fields_ = [("n_padded", ctypes.c_int64), ("block", ctypes.c_int64), ("nnz", ctypes.c_int64),
              ("max_identity", ctypes.c_int64), ("max_balanced", ctypes.c_int64),
              ("lb", ctypes.c_double), ("grid", ctypes.c_int32), ("passes", ctypes.c_int32),
              ("exact_blocks", ctypes.c_int32), ("reserved", ctypes.c_int32)]

  def as_dict(self):
    return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class BalanceError(RuntimeError):
  def __init__(self, code, msg):
    super().__init__(f"cgbalance error {code}: {msg}")
    self.code = code


_LIB = None


def load():
  """Loads libcgbalance.so; raises if it has not been built."""
  global _LIB
  if _LIB is not None:
    return _LIB
  if not os.path.exists(LIB_PATH):
    raise RuntimeError(f"{LIB_PATH} is not built; run python -m ceregraphs_amd.build")
  lib = ctypes.CDLL(LIB_PATH)
  P, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
  lib.cgbal_padded_n.argtypes = [i64, i32, i32]
  lib.cgbal_padded_n.restype = i64
  lib.cgbal_balance_coo.argtypes = [i64, i64, P, P, i32, i32, i32, i32, P, i64, P,
                                    ctypes.POINTER(Report)]
  lib.cgbal_balance_coo.restype = ctypes.c_int
  lib.cgbal_balance_csr.argtypes = [i64, P, P, i32, i32, i32, i32, P, i64, P,
                                    ctypes.POINTER(Report)]
  lib.cgbal_balance_csr.restype = ctypes.c_int
  lib.cgbal_last_error.argtypes = []
  lib.cgbal_last_error.restype = ctypes.c_char_p
  _LIB = lib
  return lib


def _check(rc):
  if rc != OK:
    raise BalanceError(rc, load().cgbal_last_error().decode())


@dataclass
class Balanced:
  operm: np.ndarray    # int64[n_padded]: original id -> balanced id
  buckets: np.ndarray  # int64[grid, grid]: nonzeros of PE (px, py) under operm
  report: dict         # n_padded, block, nnz, max_identity, max_balanced, lb, grid, passes, exact_blocks

  @property
  def n_padded(self):
    return len(self.operm)


def padded_n(n, grid):
  v = load().cgbal_padded_n(int(n), int(grid), int(grid))
  if v < 0:
    _check(EINVAL)
  return int(v)


def _flags(symmetric, remove_dup):
  return (SYMMETRIC if symmetric else SHARED_PERM) | (REMOVE_DUP if remove_dup else 0)


def _run(call, n, grid):
  npad = padded_n(n, grid)
  g = min(int(grid), int(n))
  operm = np.empty(npad, dtype=np.int64)
  buckets = np.empty((g, g), dtype=np.int64)
  rep = Report()
  _check(call(operm.ctypes.data, npad, buckets.ctypes.data, ctypes.byref(rep)))
  return Balanced(operm, buckets, rep.as_dict())


def balance_coo(n, rows, cols, grid, symmetric=True, remove_dup=False, threads=0):
  """Balances the n x n matrix with entries (rows[i], cols[i]), 0-based.

  symmetric=True is analyze's --symmetric (refuses a matrix with A != A^T),
  False is --shared-perm; remove_dup is --remove_dup."""
  rows = np.ascontiguousarray(rows, dtype=np.int32)
  cols = np.ascontiguousarray(cols, dtype=np.int32)
  if rows.shape != cols.shape:
    raise ValueError("rows and cols differ in length")
  lib = load()
  return _run(lambda q, ql, b, r: lib.cgbal_balance_coo(
      int(n), len(rows), rows.ctypes.data, cols.ctypes.data, int(grid), int(grid),
      _flags(symmetric, remove_dup), int(threads), q, ql, b, r), n, grid)


def balance_csr(n, indptr, indices, grid, symmetric=True, remove_dup=False, threads=0):
  """Balances the n x n matrix held as CSR (row r's columns are
  indices[indptr[r]:indptr[r+1]], in any order)."""
  indptr = np.ascontiguousarray(indptr, dtype=np.int64)
  indices = np.ascontiguousarray(indices, dtype=np.int32)
  if len(indptr) != int(n) + 1:
    raise ValueError("indptr must have n + 1 entries")
  lib = load()
  return _run(lambda q, ql, b, r: lib.cgbal_balance_csr(
      int(n), indptr.ctypes.data, indices.ctypes.data, int(grid), int(grid),
      _flags(symmetric, remove_dup), int(threads), q, ql, b, r), n, grid)


def relabel_csr(indptr, indices, operm):
  """CSR of B = A(q, q): entry (r, c) of A becomes (q[r], q[c]), over the
  padded size len(q); columns ascending within each row."""
  q = np.asarray(operm, dtype=np.int64)
  npad = len(q)
  indptr = np.asarray(indptr, dtype=np.int64)
  n = len(indptr) - 1
  rows = np.repeat(q[:n], np.diff(indptr))
  cols = q[np.asarray(indices, dtype=np.int64)]
  order = np.lexsort((cols, rows))
  out_ptr = np.zeros(npad + 1, dtype=np.int64)
  np.cumsum(np.bincount(rows, minlength=npad), out=out_ptr[1:])
  return out_ptr, cols[order].astype(np.int32)


def write_operm(path, operm):
  """The reference's .operm bytes: one decimal id per line (analyze.cpp:700-715)."""
  with open(path, "w", encoding="ascii") as f:
    f.write("\n".join(str(int(v)) for v in operm))
    f.write("\n")
