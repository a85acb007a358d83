"""Device-resident graphs and BFS searches over the libcgbfs C ABI.

This is the host-side replacement for the reference's SdkRuntime session:
``DeviceGraph`` is the compiled-and-loaded program plus its uploaded matrix
(run_bfs.py:288-358), ``DeviceGraph.bfs`` is one ``runner.launch("f_spmv_iter")``
with its seed and parent readback (run_bfs.py:369-409), and
``DeviceGraph.msbfs`` runs up to 64 such searches in one pass.

Outputs come back as numpy arrays (host) or, with ``device=True``, as
torch tensors that stay in HBM.  Every call goes to the GPU library; there is
no CPU path here.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .gen_rmat import GRAPH500_ABCD, pcg_words, thresholds

MAX_HIST = 4096


def _ptr(a):
  return ctypes.c_void_p(a.ctypes.data) if a is not None else None


@dataclass
class SearchStats:
  rounds_completed: int
  total_ms: float
  output_ms: float
  nf_history: np.ndarray
  direction_history: np.ndarray
  edges_scanned: np.ndarray
  rows_scanned: np.ndarray
  pairs_found: np.ndarray
  level_ms: np.ndarray
  visited_count: int = -1
  m_edges: int = -1
  # multi-source: the pull part of each round (the push part is the rest)
  pull_lanes: np.ndarray = None
  pull_nf: np.ndarray = None
  pull_rows: np.ndarray = None
  pull_edges: np.ndarray = None
  pull_pairs: np.ndarray = None
  pull_ms: np.ndarray = None


@dataclass
class BFSResult:
  source: int
  parents: object  # numpy int32[n] or torch.int32 tensor on the GPU
  levels: object
  stats: SearchStats


@dataclass
class MSBFSResult:
  roots: np.ndarray
  parents: object  # int32[nroots][n]
  levels: object   # int16[nroots][n]
  m_lanes: np.ndarray
  stats: SearchStats = field(default=None)


def _options(direction="beamer", alpha=None, beta=None, tau=None, grid_mult=None,
             max_levels=None, rounds_hint=None, capture=False, pull_grab=None, seg_grab=None,
             push_wave_degree=None, push_big_degree=None, loop=None):
  """cgbfs_options from keyword arguments; None or 0 takes the library's
  default for that field (include/cgbfs.h)."""
  o = _lib.Options()
  if isinstance(direction, str):
    direction = _lib.DIRECTIONS[direction]
  o.direction = int(direction)
  o.alpha = float(alpha or 0.0)
  o.beta = float(beta or 0.0)
  o.tau = int(tau or 0)
  o.grid_mult = int(grid_mult or 0)
  o.max_levels = int(max_levels or 0)
  o.rounds_hint = int(rounds_hint or 0)
  o.capture = 1 if capture else 0
  o.pull_grab = int(pull_grab or 0)
  o.seg_grab = int(seg_grab or 0)
  o.push_wave_degree = int(push_wave_degree or 0)
  o.push_big_degree = int(push_big_degree or 0)
  o.loop = _lib.LOOPS[loop] if isinstance(loop, str) else int(loop or 0)
  return o


class _StatsBuf:
  def __init__(self, cap=MAX_HIST):
    self.nf = np.zeros(cap, np.int64)
    self.dir = np.zeros(cap, np.int32)
    self.scanned = np.zeros(cap, np.int64)
    self.rows = np.zeros(cap, np.int64)
    self.pairs = np.zeros(cap, np.int64)
    self.level_ms = np.zeros(cap, np.float64)
    s = _lib.Stats()
    s.max_hist = cap
    s.nf_history = self.nf.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    s.direction_history = self.dir.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    s.edges_scanned = self.scanned.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    s.rows_scanned = self.rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    s.pairs_found = self.pairs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    s.level_ms = self.level_ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    self.pull = {k: np.zeros(cap, np.int64) for k in ("pull_lanes", "pull_nf", "pull_rows",
                                                     "pull_edges", "pull_pairs")}
    self.pull["pull_ms"] = np.zeros(cap, np.float64)
    for k, a in self.pull.items():
      ct = ctypes.c_double if a.dtype == np.float64 else ctypes.c_int64
      setattr(s, k, a.ctypes.data_as(ctypes.POINTER(ct)))
    self.s = s

  def result(self):
    r = min(int(self.s.rounds_completed), len(self.nf))
    return SearchStats(int(self.s.rounds_completed), float(self.s.total_ms),
                       float(self.s.output_ms), self.nf[:r].copy(), self.dir[:r].copy(),
                       self.scanned[:r].copy(), self.rows[:r].copy(), self.pairs[:r].copy(),
                       self.level_ms[:r].copy(),
                       int(self.s.visited_count), int(self.s.m_edges),
                       **{k: a[:r].copy() for k, a in self.pull.items()})


class DeviceGraph:
  """A graph resident in HBM (int64 offsets, int32 ascending indices)."""

  def __init__(self, handle, device=0):
    self._h = handle
    self.device = device
    n = ctypes.c_int64()
    nnz = ctypes.c_int64()
    sym = ctypes.c_int32()
    _lib.check(_lib.load().cgbfs_graph_info(self._h, ctypes.byref(n), ctypes.byref(nnz),
                                            ctypes.byref(sym)))
    self.n, self.nnz, self.symmetric = n.value, nnz.value, bool(sym.value)
    self._last_rounds = 0

  # -- construction --------------------------------------------------------
  @classmethod
  def from_csr(cls, n, out_ptr, out_idx, in_ptr=None, in_idx=None, symmetric=False, device=0):
    """Copies a graph into the library.  out_* = column-compressed M (push
    view), in_* = row-compressed M (pull view, ascending); symmetric graphs
    may pass only the out view.  The arrays are host arrays (numpy, int64
    offsets and int32 ids) or CUDA tensors of those dtypes on `device`, all
    of one kind; device tensors are copied device to device and checked on
    the GPU."""
    if symmetric:
      in_ptr = in_idx = None
    arrays = [out_ptr, out_idx, in_ptr, in_idx]
    on_dev = [a is not None and hasattr(a, "data_ptr") and getattr(a, "is_cuda", False)
              for a in arrays]
    if any(on_dev):
      keep = arrays
      for a, want in zip(arrays, ("int64", "int32", "int64", "int32")):
        if a is not None:
          _check_tensor(a, a.numel(), want)
      _after_torch_stream(out_ptr)
      ptrs = [ctypes.c_void_p(a.data_ptr()) if a is not None else None for a in arrays]
      nnz = int(out_idx.numel())
    else:
      keep = [np.ascontiguousarray(out_ptr, dtype=np.int64),
              np.ascontiguousarray(out_idx, dtype=np.int32), None, None]
      if not symmetric:
        keep[2] = np.ascontiguousarray(in_ptr, dtype=np.int64)
        keep[3] = np.ascontiguousarray(in_idx, dtype=np.int32)
      ptrs = [_ptr(a) for a in keep]
      nnz = len(keep[1])
    h = ctypes.c_void_p()
    _lib.check(_lib.load().cgbfs_graph_create(int(device), int(n), nnz, *ptrs,
                                              1 if symmetric else 0, ctypes.byref(h)))
    del keep
    return cls(h, device)

  @classmethod
  def from_rmat(cls, scale, edgefactor, seed, abcd=GRAPH500_ABCD, directed=False, device=0):
    """Generates the R-MAT graph on the GPU, edge-identical to the numpy
    restatement in gen_rmat.rmat_csr for the same arguments."""
    a, ab, abc = thresholds(abcd[0], abcd[1], abcd[2])
    words = pcg_words(seed)
    h = ctypes.c_void_p()
    _lib.check(_lib.load().cgbfs_graph_create_rmat(
        int(device), int(scale), int(edgefactor),
        words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), a, ab, abc,
        0 if directed else 1, ctypes.byref(h)))
    return cls(h, device)

  def close(self):
    if self._h:
      _lib.load().cgbfs_graph_destroy(self._h)
      self._h = None

  def __del__(self):
    try:
      self.close()
    except Exception:  # pylint: disable=broad-except
      pass

  def __enter__(self):
    return self

  def __exit__(self, *exc):
    self.close()

  # -- inspection -----------------------------------------------------------
  def download(self):
    """(in_ptr, in_idx, out_ptr, out_idx) as host arrays."""
    out_ptr = np.empty(self.n + 1, np.int64)
    out_idx = np.empty(self.nnz, np.int32)
    if self.symmetric:
      _lib.check(_lib.load().cgbfs_graph_download(self._h, _ptr(out_ptr), _ptr(out_idx), None, None))
      return out_ptr, out_idx, out_ptr, out_idx
    in_ptr = np.empty(self.n + 1, np.int64)
    in_idx = np.empty(self.nnz, np.int32)
    _lib.check(_lib.load().cgbfs_graph_download(self._h, _ptr(out_ptr), _ptr(out_idx),
                                                _ptr(in_ptr), _ptr(in_idx)))
    return in_ptr, in_idx, out_ptr, out_idx

  def out_degree(self):
    deg = np.empty(self.n, np.int64)
    _lib.check(_lib.load().cgbfs_graph_out_degree(self._h, _ptr(deg)))
    return deg

  def sync(self):
    _lib.check(_lib.load().cgbfs_sync(self._h))

  # -- searches -------------------------------------------------------------
  def bfs(self, source, direction="beamer", device_out=None, **kw):
    """One search.  device_out: optional (parents, levels) torch int32
    tensors of length n on this graph's GPU to write into; otherwise the
    outputs come back as numpy arrays."""
    sb = _StatsBuf()
    kw.setdefault("rounds_hint", self._last_rounds + 1 if self._last_rounds else 0)
    opt = _options(direction, **kw)
    if device_out is not None:
      par, lev = device_out
      _check_tensor(par, self.n, "int32")
      _check_tensor(lev, self.n, "int32")
      _after_torch_stream(par)
      rc = _lib.load().cgbfs_bfs(self._h, int(source), ctypes.c_void_p(par.data_ptr()),
                                 ctypes.c_void_p(lev.data_ptr()), 1, ctypes.byref(opt),
                                 ctypes.byref(sb.s))
    else:
      par = np.empty(self.n, np.int32)
      lev = np.empty(self.n, np.int32)
      rc = _lib.load().cgbfs_bfs(self._h, int(source), _ptr(par), _ptr(lev), 0, ctypes.byref(opt),
                                 ctypes.byref(sb.s))
    _lib.check(rc)
    self._last_rounds = int(sb.s.rounds_completed)
    return BFSResult(int(source), par, lev, sb.result())

  def msbfs(self, roots, direction="beamer", device_out=None, count_m=True, **kw):
    """Up to 64 searches in one pass; lane l is root roots[l].  count_m:
    also count each lane's GTEPS m on the device (m_lanes; None without)."""
    roots = np.ascontiguousarray(roots, dtype=np.int64)
    k = len(roots)
    if not 1 <= k <= 64:
      raise ValueError("msbfs takes 1..64 roots")
    sb = _StatsBuf()
    opt = _options(direction, **kw)
    m = np.zeros(k, np.int64) if count_m else None
    if device_out is not None:
      par, lev = device_out
      _check_tensor(par, k * self.n, "int32")
      _check_tensor(lev, k * self.n, "int16")
      _after_torch_stream(par)
      rc = _lib.load().cgbfs_msbfs(self._h, _ptr(roots), k, ctypes.c_void_p(par.data_ptr()),
                                   ctypes.c_void_p(lev.data_ptr()), 1, ctypes.byref(opt),
                                   ctypes.byref(sb.s), _ptr(m))
    else:
      par = np.empty((k, self.n), np.int32)
      lev = np.empty((k, self.n), np.int16)
      rc = _lib.load().cgbfs_msbfs(self._h, _ptr(roots), k, _ptr(par), _ptr(lev), 0,
                                   ctypes.byref(opt), ctypes.byref(sb.s), _ptr(m))
    _lib.check(rc)
    return MSBFSResult(roots, par, lev, m, sb.result())


  def msbfs_seeded(self, verts, lanes, nlanes, direction="beamer", device_out=None, **kw):
    """Multi-source pass from seeds: seed k starts lane lanes[k] at vertex
    verts[k]; a lane may hold several seeds (see cgbfs_msbfs_seeded)."""
    verts = np.ascontiguousarray(verts, dtype=np.int64)
    lanes = np.ascontiguousarray(lanes, dtype=np.int32)
    if len(verts) != len(lanes) or not len(verts):
      raise ValueError("verts and lanes must be non-empty and of equal length")
    if not 1 <= nlanes <= 64:
      raise ValueError("nlanes must be in [1, 64]")
    sb = _StatsBuf()
    opt = _options(direction, **kw)
    m = np.zeros(nlanes, np.int64)
    if device_out is not None:
      par, lev = device_out
      _check_tensor(par, nlanes * self.n, "int32")
      _check_tensor(lev, nlanes * self.n, "int16")
      _after_torch_stream(par)
      rc = _lib.load().cgbfs_msbfs_seeded(self._h, _ptr(verts), _ptr(lanes), len(verts), nlanes,
                                          ctypes.c_void_p(par.data_ptr()),
                                          ctypes.c_void_p(lev.data_ptr()), 1, ctypes.byref(opt),
                                          ctypes.byref(sb.s), _ptr(m))
    else:
      par = np.empty((nlanes, self.n), np.int32)
      lev = np.empty((nlanes, self.n), np.int16)
      rc = _lib.load().cgbfs_msbfs_seeded(self._h, _ptr(verts), _ptr(lanes), len(verts), nlanes,
                                          _ptr(par), _ptr(lev), 0, ctypes.byref(opt),
                                          ctypes.byref(sb.s), _ptr(m))
    _lib.check(rc)
    return MSBFSResult(verts, par, lev, m, sb.result())


class ServingEngine:
  """A long-lived 64-lane multi-source search that takes single-root
  queries as lanes free up (cgbfs_engine_*): submit() queues a query on a
  free lane, step() runs one round for every lane in flight (asynchronous),
  poll() reports the queries completed by that step."""

  LANES = 64

  def __init__(self, graph, direction="beamer_lanes", **kw):
    self.graph = graph
    self.n = graph.n
    h = ctypes.c_void_p()
    opt = _options(direction, **kw)
    _lib.check(_lib.load().cgbfs_engine_create(graph._h, ctypes.byref(opt), ctypes.byref(h)))
    self._e = h
    self._m = np.zeros(64, np.int64)
    self._depth = np.zeros(64, np.int32)

  def submit(self, lane, root, parents, levels):
    """parents: int32 CUDA tensor of n elements, levels: int16, both on the
    graph's device; they receive this query's outputs."""
    _check_tensor(parents, self.n, "int32")
    _check_tensor(levels, self.n, "int16")
    _after_torch_stream(parents)
    _lib.check(_lib.load().cgbfs_engine_submit(self._e, int(lane), int(root),
                                               ctypes.c_void_p(parents.data_ptr()),
                                               ctypes.c_void_p(levels.data_ptr())))

  def step(self):
    _lib.check(_lib.load().cgbfs_engine_step(self._e))

  def poll(self, wait=False):
    """(ready, completed lanes, busy lanes, m per lane, rounds per lane);
    the last two are valid for the completed lanes."""
    ready = ctypes.c_int32()
    done = ctypes.c_uint64()
    busy = ctypes.c_uint64()
    _lib.check(_lib.load().cgbfs_engine_poll(self._e, 1 if wait else 0, ctypes.byref(ready),
                                             ctypes.byref(done), ctypes.byref(busy),
                                             _ptr(self._m), _ptr(self._depth)))
    lanes = [l for l in range(64) if (done.value >> l) & 1]
    return bool(ready.value), lanes, busy.value, self._m.copy(), self._depth.copy()

  def close(self):
    if self._e:
      _lib.load().cgbfs_engine_destroy(self._e)
      self._e = None

  def __del__(self):
    try:
      self.close()
    except Exception:  # pylint: disable=broad-except
      pass

  def __enter__(self):
    return self

  def __exit__(self, *exc):
    self.close()


def _after_torch_stream(t):
  """The library writes caller-provided device outputs on its own stream,
  which does not wait for torch's streams: finish the work torch has queued
  on this device (a fill of the same buffers, a reader of a previous result)
  before the call enqueues its writes."""
  import torch
  torch.cuda.current_stream(t.device).synchronize()


def _check_tensor(t, numel, dtype):
  import torch  # plumbing only: device memory for outputs that stay in HBM
  want = {"int64": torch.int64, "int32": torch.int32, "int16": torch.int16}[dtype]
  if t.dtype != want or not t.is_cuda or not t.is_contiguous() or t.numel() != numel:
    raise ValueError(f"expected a contiguous {dtype} CUDA tensor with {numel} elements")
