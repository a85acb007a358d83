"""The host-side native code under AddressSanitizer and UndefinedBehavior-
Sanitizer (host code only; no GPU): the balancer CLI on valid and malformed
Matrix Market files, and the graph check of cgbfs_graph_create on valid and
malformed CSR arrays.  Any sanitizer report fails the test; malformed input
must end in a clean error, never a crash."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import REPO

CSRC = os.path.join(REPO, "ceregraphs_amd", "csrc")
INCLUDE = os.path.join(REPO, "include")
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=undefined"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1")


def _run(cmd, **kw):
  return subprocess.run(cmd, capture_output=True, text=True, env=ENV, timeout=300, **kw)


def _clean(p):
  err = p.stderr
  return "AddressSanitizer" not in err and "runtime error" not in err and "UndefinedBehavior" not in err


@pytest.fixture(scope="module")
def balancer_asan(tmp_path_factory):
  if not shutil.which("g++"):
    pytest.skip("g++ not available")
  out = str(tmp_path_factory.mktemp("asan") / "cgbalance_asan")
  p = _run(["g++", "-std=c++17", "-pthread", *SAN, f"-I{INCLUDE}",
            os.path.join(CSRC, "balance_cli.cpp"), os.path.join(CSRC, "balance.cpp"), "-o", out])
  assert p.returncode == 0, p.stderr
  return out


def _mtx(path, text):
  path.write_text(text)
  return str(path)


def test_balancer_cli_under_sanitizers(balancer_asan, tmp_path):
  from ceregraphs_amd import gen_rmat
  n, ip, ii, _, _ = gen_rmat.rmat_csr(9, 8, 1)
  rows = np.repeat(np.arange(n), np.diff(ip))
  body = "".join(f"{r + 1} {c + 1} 1\n" for r, c in zip(rows, ii))
  good = _mtx(tmp_path / "g.mtx", f"%%MatrixMarket matrix coordinate real general\n{n} {n} {len(ii)}\n{body}")
  for mode, grid in [("--symmetric", 4), ("--shared-perm", 7), ("--symmetric", 1)]:
    p = _run([balancer_asan, "--matrix", good, mode, "--fabx", str(grid), "--faby", str(grid),
              "--omatrix", str(tmp_path / "o.mtx"), "--operm", str(tmp_path / "o.operm"),
              "--threads", "4"])
    assert p.returncode == 0 and _clean(p), p.stderr[-2000:]
  bad = {
      "truncated": "%%MatrixMarket matrix coordinate real general\n4 4 3\n1 2 1\n2 1 1\n",
      "index_past_n": "%%MatrixMarket matrix coordinate real general\n4 4 2\n1 9 1\n9 1 1\n",
      "index_zero": "%%MatrixMarket matrix coordinate real general\n4 4 2\n0 1 1\n1 0 1\n",
      "negative_size": "%%MatrixMarket matrix coordinate real general\n-4 4 2\n1 2 1\n2 1 1\n",
      "no_header": "4 4 2\n1 2 1\n2 1 1\n",
      "garbage_entry": "%%MatrixMarket matrix coordinate real general\n4 4 2\n1 x 1\n2 1 1\n",
      "huge_count": "%%MatrixMarket matrix coordinate real general\n4 4 999999999999\n1 2 1\n",
      "empty_file": "",
      "rectangular": "%%MatrixMarket matrix coordinate real general\n4 5 1\n1 2 1\n",
  }
  for name, text in bad.items():
    p = _run([balancer_asan, "--matrix", _mtx(tmp_path / f"{name}.mtx", text), "--shared-perm",
              "--fabx", "2", "--faby", "2", "--omatrix", str(tmp_path / "x.mtx")])
    assert p.returncode != 0 and _clean(p), (name, p.returncode, p.stderr[-2000:])
  p = _run([balancer_asan, "--matrix", str(tmp_path / "missing.mtx"), "--shared-perm", "--fabx", "2",
            "--faby", "2"])
  assert p.returncode != 0 and _clean(p)
  p = _run([balancer_asan, "--fabx", "abc", "--matrix", good])
  assert p.returncode != 0 and _clean(p)


def test_graph_check_under_sanitizers(tmp_path):
  hipcc = "/opt/rocm/bin/hipcc"
  if not os.path.exists(hipcc):
    pytest.skip("hipcc not available")
  exe = str(tmp_path / "check_asan")
  host_san = []
  for f in SAN:
    host_san += ["-Xarch_host", f] if f.startswith("-fsanitize") or f.startswith("-fno-sanitize") else [f]
  p = _run([hipcc, "--offload-arch=gfx950", "-std=c++17", *host_san, f"-I{INCLUDE}",
            os.path.join(REPO, "tests", "native", "graph_check_driver.cpp"),
            os.path.join(CSRC, "check.hip"), os.path.join(CSRC, "graph.hip"), "-o", exe])
  assert p.returncode == 0, p.stderr[-3000:]
  p = _run([exe])
  assert p.returncode == 0 and _clean(p), p.stderr[-3000:]
  got = dict(line.split() for line in p.stdout.split("\n") if line.strip())
  ok, einval = "0", "-1"
  assert got.pop("valid") == ok and got.pop("empty") == ok and got.pop("ring") == ok
  assert all(v == einval for v in got.values()), got
