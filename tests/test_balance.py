"""The native balancer (libcgbalance.so, cgbalance) against the reference's
util/analyze: golden .operm vectors and heaviest-PE counts it wrote
(tests/golden/make_golden.py), C1's balanced input, byte-for-byte output
files, hand-worked placements, and the error behaviour.  CPU only."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import REPO, golden
from ceregraphs_amd import balance as B

HEADER = os.path.join(REPO, "include", "cgbalance.h")
REF_ANALYZE = os.path.join(REPO, "oracle", "_ref", "analyze")


@pytest.fixture(scope="module", autouse=True)
def _built():
  if not (os.path.exists(B.LIB_PATH) and os.path.exists(B.CLI_PATH)):
    from ceregraphs_amd import build
    build.build_balancer(verbose=False)


def _cases():
  return [str(t) for t in golden("balance_cases.npz")["cases"]]


def _case(tag):
  g = golden("balance_cases.npz")
  name = str(g[f"{tag}_input"])
  vals = g[f"in_{name}_vals"] if f"in_{name}_vals" in g.files else None
  return dict(n=int(g[f"in_{name}_n"]), rows=g[f"in_{name}_rows"], cols=g[f"in_{name}_cols"],
              vals=vals, grid=int(g[f"{tag}_grid"]), symmetric=str(g[f"{tag}_mode"]) == "symmetric",
              dedup=bool(g[f"{tag}_dedup"]), operm=g[f"{tag}_operm"], heaviest=int(g[f"{tag}_max"]),
              omatrix=g[f"{tag}_omatrix"].tobytes() if f"{tag}_omatrix" in g.files else None)


def _write_mtx(path, n, rows, cols, vals=None):
  with open(path, "w", encoding="ascii") as f:
    f.write(f"%%MatrixMarket matrix coordinate real general\n{n} {n} {len(rows)}\n")
    for i in range(len(rows)):
      v = 1.0 if vals is None else float(vals[i])
      f.write(f"{int(rows[i]) + 1} {int(cols[i]) + 1} {v!r}\n")


def _recount(rows, cols, q, block, grid):
  """Per-PE nonzeros under q, indexed [px, py] (column block, row block)."""
  c = np.zeros((grid, grid), np.int64)
  np.add.at(c, (q[cols] // block, q[rows] // block), 1)
  return c


# -- the C ABI ------------------------------------------------------------------
def test_header_and_binding_agree():
  text = open(HEADER, encoding="utf-8").read()
  declared = set(re.findall(r"CGBAL_API\s+[\w\s\*]*?\b(cgbal_\w+)\s*\(", text))
  assert declared == set(B.EXPORTED)
  lib = ctypes.CDLL(B.LIB_PATH)
  for name in declared:
    assert hasattr(lib, name), name


def test_report_layout_matches_header():
  text = open(HEADER, encoding="utf-8").read()
  body = re.search(r"typedef struct \{([^}]*)\} cgbal_report;", text, re.S).group(1)
  names = re.findall(r"\b(\w+);", body)
  assert names == [f for f, _ in B.Report._fields_]
  assert ctypes.sizeof(B.Report) == 64


# -- against the reference's own outputs -----------------------------------------
@pytest.mark.parametrize("tag", _cases())
def test_operm_matches_reference(tag):
  c = _case(tag)
  r = B.balance_coo(c["n"], c["rows"], c["cols"], c["grid"], symmetric=c["symmetric"],
                    remove_dup=c["dedup"])
  np.testing.assert_array_equal(r.operm, c["operm"])
  assert r.report["max_balanced"] == c["heaviest"]
  assert int(r.buckets.max()) == c["heaviest"]


@pytest.mark.parametrize("tag", ["dir12_p7", "sym12_p9", "dup_p6", "star_p8"])
def test_csr_entry_and_thread_count_agree(tag):
  c = _case(tag)
  order = np.argsort(c["rows"], kind="stable")
  ptr = np.zeros(c["n"] + 1, np.int64)
  np.cumsum(np.bincount(c["rows"], minlength=c["n"]), out=ptr[1:])
  for threads in (1, 3, 8):
    r = B.balance_csr(c["n"], ptr, c["cols"][order], c["grid"], symmetric=c["symmetric"],
                      remove_dup=c["dedup"], threads=threads)
    np.testing.assert_array_equal(r.operm, c["operm"])


def test_c1_input_is_reproduced():
  """C1's balanced graph: the s10 ef16 seed-0 graph through this balancer
  (4x4, --symmetric) equals the reference analyze's .operm and balanced CSR."""
  raw = golden("rmat_s10_ef16_seed0.npz")
  g = golden("c1_balanced_4x4.npz")
  n = int(raw["n"])
  r = B.balance_csr(n, raw["indptr"], raw["indices"], 4, symmetric=True)
  np.testing.assert_array_equal(r.operm, g["operm"])
  ptr, idx = B.relabel_csr(raw["indptr"], raw["indices"], r.operm)
  np.testing.assert_array_equal(ptr, g["indptr"])
  np.testing.assert_array_equal(idx, g["indices"])
  assert r.report["n_padded"] == n and r.report["block"] == 256


def test_cli_writes_the_reference_bytes(tmp_path):
  """cgbalance's .mtx (value formatting included) and .operm, byte for byte."""
  c = _case("dup_p6")
  src = tmp_path / "a.mtx"
  _write_mtx(src, c["n"], c["rows"], c["cols"], c["vals"])
  om, op = tmp_path / "b.mtx", tmp_path / "b.operm"
  subprocess.run([B.CLI_PATH, "--matrix", str(src), "--shared-perm", "--remove_dup", "--fabx", "6",
                  "--faby", "6", "--omatrix", str(om), "--operm", str(op)], check=True,
                 capture_output=True)
  assert om.read_bytes() == c["omatrix"]
  ref_operm = tmp_path / "ref.operm"
  B.write_operm(ref_operm, c["operm"])
  assert op.read_bytes() == ref_operm.read_bytes()


@pytest.mark.skipif(not os.path.exists(REF_ANALYZE),
                    reason="the reference analyze is not built (make -C oracle ref)")
@pytest.mark.parametrize("tag,extra", [("dir12_p3", []), ("star_p8", ["--rand", "3"]),
                                       ("nodup_p6", []), ("tiny_p4", [])])
def test_cli_matches_live_reference(tmp_path, tag, extra):
  c = _case(tag)
  src = tmp_path / "a.mtx"
  _write_mtx(src, c["n"], c["rows"], c["cols"], c["vals"])
  mode = "--symmetric" if c["symmetric"] else "--shared-perm"
  outs = {}
  for tool in (REF_ANALYZE, B.CLI_PATH):
    om, op = tmp_path / f"{os.path.basename(tool)}.mtx", tmp_path / f"{os.path.basename(tool)}.operm"
    subprocess.run([tool, "--matrix", str(src), mode, "--fabx", str(c["grid"]), "--faby",
                    str(c["grid"]), "--omatrix", str(om), "--operm", str(op), *extra], check=True,
                   capture_output=True)
    outs[tool] = (om.read_bytes(), op.read_bytes())
  assert outs[REF_ANALYZE] == outs[B.CLI_PATH]


# -- worked by hand and closed forms ------------------------------------------------
def test_hand_worked_two_by_two():
  """n=4, P=2, entries (0,1) (0,2) (0,3) (1,0) (2,0) (2,3); counts 3,1,2,0.

  Pass 1 (columns in identity blocks 0,0,1,1): row 0 -> block row 0 (its
  heaviest PE becomes 2), row 2 -> block row 1 (heaviest 1), row 1 -> block
  row 1 (lighter; now full), row 3 -> block row 0: q = [0,3,2,1], heaviest
  PE 3, ratio 6/(4*3) = 0.5 > 0, kept.
  Pass 2 (columns in blocks 0,1,1,0 under q): row 0 -> 0 (heaviest 2), row 2
  -> 1 (heaviest 2), row 1 -> 0 (tie: the lower block row), row 3 -> 1:
  q = identity, heaviest 2, ratio 0.75 > 0.5, kept.
  Pass 3 repeats pass 1: ratio 0.5, not better, dropped.  Result: identity."""
  rows = np.array([0, 0, 0, 1, 2, 2])
  cols = np.array([1, 2, 3, 0, 0, 3])
  r = B.balance_coo(4, rows, cols, 2, symmetric=False)
  np.testing.assert_array_equal(r.operm, [0, 1, 2, 3])
  assert r.report["passes"] == 3
  assert r.report["max_identity"] == 2 and r.report["max_balanced"] == 2
  assert r.report["lb"] == pytest.approx(0.75)
  np.testing.assert_array_equal(r.buckets, [[2, 1], [2, 1]])


def test_single_pe_sorts_by_decreasing_degree():
  """With P = 1 every row lands in block row 0 in visiting order, so with
  distinct counts q is the rank by decreasing count; the second pass
  repeats the first and is dropped."""
  rng = np.random.default_rng(1)
  n = 40
  deg = rng.permutation(n)  # distinct out-counts 0..39
  rows = np.repeat(np.arange(n), deg)
  cols = np.concatenate([rng.choice(n, d, replace=False) for d in deg]).astype(np.int64)
  r = B.balance_coo(n, rows, cols, 1, symmetric=False)
  rank = np.empty(n, np.int64)
  rank[np.argsort(-deg, kind="stable")] = np.arange(n)
  np.testing.assert_array_equal(r.operm, rank)
  assert r.report["passes"] == 2 and r.report["lb"] == 1.0


@pytest.mark.parametrize("tag", ["dir12_p7", "dir12_p100", "dup_p6", "tiny_p4", "star_p8"])
def test_permutation_properties(tag):
  """A bijection onto the padded ids, P*ceil(n/P) of them (single-precision
  ceil equals the exact one at these sizes), and per-PE counts that recount
  from the permutation and sum to the entries balanced."""
  c = _case(tag)
  r = B.balance_coo(c["n"], c["rows"], c["cols"], c["grid"], symmetric=c["symmetric"],
                    remove_dup=c["dedup"])
  grid = min(c["grid"], c["n"])
  block = -(-c["n"] // grid)
  assert r.n_padded == grid * block == B.padded_n(c["n"], c["grid"])
  np.testing.assert_array_equal(np.sort(r.operm), np.arange(r.n_padded))
  rows, cols = c["rows"].astype(np.int64), c["cols"].astype(np.int64)
  if c["dedup"]:
    key = np.unique(rows * c["n"] + cols)
    rows, cols = key // c["n"], key % c["n"]
    rows, cols = rows[rows != cols], cols[rows != cols]
  np.testing.assert_array_equal(r.buckets, _recount(rows, cols, r.operm, block, grid))
  assert int(r.buckets.sum()) == r.report["nnz"] == len(rows)


def test_relabel_is_the_permuted_matrix():
  c = _case("dir12_p7")
  n = c["n"]
  ptr = np.zeros(n + 1, np.int64)
  np.cumsum(np.bincount(c["rows"], minlength=n), out=ptr[1:])
  idx = c["cols"][np.argsort(c["rows"], kind="stable")]
  q = c["operm"]
  bp, bi = B.relabel_csr(ptr, idx, q)
  brow = np.repeat(np.arange(len(q)), np.diff(bp))
  inv = np.argsort(q)
  got = sorted(zip(inv[brow].tolist(), inv[bi].tolist()))
  assert got == sorted(zip(c["rows"].tolist(), c["cols"].tolist()))


# -- errors ---------------------------------------------------------------------
def test_errors():
  c = _case("dir12_p3")
  with pytest.raises(B.BalanceError) as e:
    B.balance_coo(c["n"], c["rows"], c["cols"], 4, symmetric=True)
  assert e.value.code == B.ENOTSYM and "not structurally symmetric" in str(e.value)
  lib = B.load()
  q = np.empty(8, np.int64)
  rows = np.array([0, 1], np.int32)
  cols = np.array([1, 0], np.int32)
  args = (4, 2, rows.ctypes.data, cols.ctypes.data)
  for fabx, faby, flags, want in [(2, 2, B.SYMMETRIC | B.SHARED_PERM, "mutually exclusive"),
                                  (2, 2, 0, "is required"), (2, 4, B.SHARED_PERM, "square PE grid"),
                                  (0, 0, B.SHARED_PERM, "positive")]:
    rc = lib.cgbal_balance_coo(*args, fabx, faby, flags, 1, q.ctypes.data, len(q), None, None)
    assert rc == B.EINVAL and want in lib.cgbal_last_error().decode()
  bad = np.array([0, 4], np.int32)
  rc = lib.cgbal_balance_coo(4, 2, bad.ctypes.data, cols.ctypes.data, 2, 2, B.SHARED_PERM, 1,
                             q.ctypes.data, len(q), None, None)
  assert rc == B.EINVAL and "outside" in lib.cgbal_last_error().decode()
  rc = lib.cgbal_balance_coo(4, 2, rows.ctypes.data, cols.ctypes.data, 2, 2, B.SHARED_PERM, 1,
                             q.ctypes.data, 3, None, None)
  assert rc == B.EINVAL and "iperm must hold" in lib.cgbal_last_error().decode()
  assert lib.cgbal_padded_n(0, 4, 4) == -1 and lib.cgbal_padded_n(10, 4, 4) == 12


def test_cli_refusals(tmp_path):
  c = _case("dir12_p3")
  src = tmp_path / "d.mtx"
  _write_mtx(src, c["n"], c["rows"], c["cols"])
  for flags, want in [(["--symmetric"], "not structurally symmetric"),
                      ([], "one of --symmetric or --shared-perm"),
                      (["--symmetric", "--shared-perm"], "mutually exclusive")]:
    p = subprocess.run([B.CLI_PATH, "--matrix", str(src), *flags, "--fabx", "4", "--faby", "4"],
                       capture_output=True, text=True)
    assert p.returncode == 1 and want in p.stderr
  pat = tmp_path / "p.mtx"
  pat.write_text("%%MatrixMarket matrix coordinate pattern general\n2 2 1\n1 2\n")
  p = subprocess.run([B.CLI_PATH, "--matrix", str(pat), "--shared-perm", "--fabx", "2", "--faby",
                      "2"], capture_output=True, text=True)
  assert p.returncode == 1 and "real coordinate matrix is required" in p.stderr
