"""Builds the package's native code in-tree.

    python -m ceregraphs_amd.build

* libcgbfs.so -- the gfx950 BFS library, every .hip compiled with hipcc;
* libcgbalance.so and the cgbalance executable -- the host-side load
  balancer (balance.cpp, balance_cli.cpp), compiled with g++.

Everything lands next to this file so it travels with the source tree;
nothing is installed into site-packages.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcgbfs.so")
SOURCES = ["graph.hip", "check.hip", "bfs.hip", "msbfs.hip"]
ARCH = os.environ.get("CGBFS_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
BAL_LIB = os.path.join(HERE, "libcgbalance.so")
BAL_CLI = os.path.join(HERE, "cgbalance")
HOST_FLAGS = ["-O3", "-std=c++17", "-pthread", "-Wall", "-Wextra", f"-I{INCLUDE}"]
# Packed-FP32 instructions are disabled for every device compile (the level
# kernels are integer code; this only keeps hipcc from emitting v_pk_*_f32).
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
    "-Wno-unused-command-line-argument", "-fvisibility=hidden",
]


def _compile(src):
  obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
  cmd = [HIPCC, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
  proc = subprocess.run(cmd, capture_output=True, text=True, check=False)
  if proc.returncode != 0:
    raise RuntimeError(f"hipcc failed for {src}:\n{proc.stdout}\n{proc.stderr}")
  noise = "is not a recognized feature"
  warn = "\n".join(l for l in proc.stderr.splitlines() if noise not in l and l.strip())
  return obj, warn


def build_balancer(verbose=True):
  """Compiles the host-side balancer: libcgbalance.so and cgbalance."""
  os.makedirs(BUILD, exist_ok=True)
  jobs = [
      [CXX, *HOST_FLAGS, "-fPIC", "-shared", "-fvisibility=hidden",
       os.path.join(CSRC, "balance.cpp"), "-o", BAL_LIB + ".tmp"],
      [CXX, *HOST_FLAGS, os.path.join(CSRC, "balance_cli.cpp"), os.path.join(CSRC, "balance.cpp"),
       "-o", BAL_CLI + ".tmp"],
  ]
  for cmd in jobs:
    proc = subprocess.run(cmd, capture_output=True, text=True, check=False)
    if proc.returncode != 0:
      raise RuntimeError(f"{CXX} failed:\n{proc.stdout}\n{proc.stderr}")
    if proc.stderr.strip() and verbose:
      print(proc.stderr, file=sys.stderr)
  os.replace(BAL_LIB + ".tmp", BAL_LIB)
  os.replace(BAL_CLI + ".tmp", BAL_CLI)
  if verbose:
    print(f"built {BAL_LIB} and {BAL_CLI}")


def build(verbose=True):
  """Compiles every source, links libcgbfs.so and builds the balancer;
  returns the BFS library path."""
  os.makedirs(BUILD, exist_ok=True)
  build_balancer(verbose)
  with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
    results = list(ex.map(_compile, SOURCES))
  for _, warn in results:
    if warn and verbose:
      print(warn, file=sys.stderr)
  tmp = LIB + ".tmp"
  cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
         *[obj for obj, _ in results]]
  proc = subprocess.run(cmd, capture_output=True, text=True, check=False)
  if proc.returncode != 0:
    raise RuntimeError(f"link failed:\n{proc.stdout}\n{proc.stderr}")
  os.replace(tmp, LIB)
  if verbose:
    print(f"built {LIB}")
  return LIB


if __name__ == "__main__":
  build()
