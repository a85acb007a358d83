"""ctypes binding of libcgbfs.so (declarations in include/cgbfs.h).

The library is loaded from this package directory only.  There is no
fallback: if the shared library is missing or fails to load, every entry
point raises, so a search can never silently run anywhere but the GPU.
"""
import ctypes
import os
import threading

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcgbfs.so")

CGBFS_OK = 0
CGBFS_EINVAL = -1
CGBFS_ENOMEM = -2
CGBFS_EHIP = -3
CGBFS_ETIMEOUT = -4
CGBFS_ENODEV = -5

DIR_BEAMER = 0
DIR_TOPDOWN = 1
DIR_BOTTOMUP = 2
DIR_TAU = 3
DIR_BEAMER_LANES = 4
LOOP_AUTO = 0
LOOP_KERNELS = 1
LOOP_PERSISTENT = 2
LOOP_WORKGROUP = 3
LOOPS = {"auto": LOOP_AUTO, "kernels": LOOP_KERNELS, "persistent": LOOP_PERSISTENT,
         "workgroup": LOOP_WORKGROUP}
DIRECTIONS = {"beamer": DIR_BEAMER, "topdown": DIR_TOPDOWN, "bottomup": DIR_BOTTOMUP,
              "tau": DIR_TAU, "beamer_lanes": DIR_BEAMER_LANES}

# Every symbol include/cgbfs.h declares.
EXPORTED = (
    "cgbfs_last_error", "cgbfs_version", "cgbfs_device_count", "cgbfs_graph_create",
    "cgbfs_graph_check",
    "cgbfs_graph_create_rmat", "cgbfs_graph_info", "cgbfs_graph_download",
    "cgbfs_graph_out_degree", "cgbfs_graph_destroy", "cgbfs_bfs", "cgbfs_bfs_noalloc",
    "cgbfs_msbfs",
    "cgbfs_msbfs_seeded", "cgbfs_sync", "cgbfs_engine_create", "cgbfs_engine_submit",
    "cgbfs_engine_step", "cgbfs_engine_poll", "cgbfs_engine_destroy",
)


class CgbfsError(RuntimeError):
  def __init__(self, code, msg):
    super().__init__(f"cgbfs error {code}: {msg}")
    self.code = code


class Options(ctypes.Structure):
  _fields_ = [
      ("direction", ctypes.c_int32),
      ("rounds_hint", ctypes.c_int32),
      ("alpha", ctypes.c_double),
      ("beta", ctypes.c_double),
      ("tau", ctypes.c_int64),
      ("pull_grab", ctypes.c_int32),
      ("grid_mult", ctypes.c_int32),
      ("max_levels", ctypes.c_int64),
      ("capture", ctypes.c_int32),
      ("seg_grab", ctypes.c_int32),
      ("push_wave_degree", ctypes.c_int32),
      ("push_big_degree", ctypes.c_int32),
      ("loop", ctypes.c_int32),
  ]


class Stats(ctypes.Structure):
  _fields_ = [
      ("rounds_completed", ctypes.c_int64),
      ("max_hist", ctypes.c_int64),
      ("nf_history", ctypes.POINTER(ctypes.c_int64)),
      ("direction_history", ctypes.POINTER(ctypes.c_int32)),
      ("edges_scanned", ctypes.POINTER(ctypes.c_int64)),
      ("rows_scanned", ctypes.POINTER(ctypes.c_int64)),
      ("pairs_found", ctypes.POINTER(ctypes.c_int64)),
      ("level_ms", ctypes.POINTER(ctypes.c_double)),
      ("total_ms", ctypes.c_double),
      ("output_ms", ctypes.c_double),
      ("visited_count", ctypes.c_int64),
      ("m_edges", ctypes.c_int64),
      ("pull_lanes", ctypes.POINTER(ctypes.c_int64)),
      ("pull_nf", ctypes.POINTER(ctypes.c_int64)),
      ("pull_rows", ctypes.POINTER(ctypes.c_int64)),
      ("pull_edges", ctypes.POINTER(ctypes.c_int64)),
      ("pull_pairs", ctypes.POINTER(ctypes.c_int64)),
      ("pull_ms", ctypes.POINTER(ctypes.c_double)),
  ]


_P = ctypes.c_void_p
_I64P = ctypes.POINTER(ctypes.c_int64)
_I32P = ctypes.POINTER(ctypes.c_int32)
_U64P = ctypes.POINTER(ctypes.c_uint64)
_GP = ctypes.c_void_p

_SIGNATURES = {
    "cgbfs_last_error": (ctypes.c_char_p, []),
    "cgbfs_version": (ctypes.c_int, []),
    "cgbfs_device_count": (ctypes.c_int, [_I32P]),
    "cgbfs_graph_create": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, _P, _P,
                                          _P, _P, ctypes.c_int32, ctypes.POINTER(_GP)]),
    "cgbfs_graph_check": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, _P, _P, _P, _P,
                                         ctypes.c_int32]),
    "cgbfs_graph_create_rmat": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                               _U64P, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_int32,
                                               ctypes.POINTER(_GP)]),
    "cgbfs_graph_info": (ctypes.c_int, [_GP, _I64P, _I64P, _I32P]),
    "cgbfs_graph_download": (ctypes.c_int, [_GP, _P, _P, _P, _P]),
    "cgbfs_graph_out_degree": (ctypes.c_int, [_GP, _P]),
    "cgbfs_graph_destroy": (ctypes.c_int, [_GP]),
    "cgbfs_bfs": (ctypes.c_int, [_GP, ctypes.c_int64, _P, _P, ctypes.c_int32,
                                 ctypes.POINTER(Options), ctypes.POINTER(Stats)]),
    "cgbfs_bfs_noalloc": (ctypes.c_int, [_GP, ctypes.c_int64, _P, _P, ctypes.POINTER(Options),
                                         ctypes.POINTER(Stats)]),
    "cgbfs_msbfs": (ctypes.c_int, [_GP, _P, ctypes.c_int32, _P, _P, ctypes.c_int32,
                                   ctypes.POINTER(Options), ctypes.POINTER(Stats), _P]),
    "cgbfs_msbfs_seeded": (ctypes.c_int, [_GP, _P, _P, ctypes.c_int64, ctypes.c_int32, _P, _P,
                                          ctypes.c_int32, ctypes.POINTER(Options),
                                          ctypes.POINTER(Stats), _P]),
    "cgbfs_sync": (ctypes.c_int, [_GP]),
    "cgbfs_engine_create": (ctypes.c_int, [_GP, ctypes.POINTER(Options), ctypes.POINTER(_GP)]),
    "cgbfs_engine_submit": (ctypes.c_int, [_GP, ctypes.c_int32, ctypes.c_int64, _P, _P]),
    "cgbfs_engine_step": (ctypes.c_int, [_GP]),
    "cgbfs_engine_poll": (ctypes.c_int, [_GP, ctypes.c_int32, _I32P, _U64P, _U64P, _P, _P]),
    "cgbfs_engine_destroy": (ctypes.c_int, [_GP]),
}

_lock = threading.Lock()
_lib = None


def load():
  """Returns the loaded library; raises if libcgbfs.so is absent."""
  global _lib
  with _lock:
    if _lib is not None:
      return _lib
    if not os.path.exists(LIB_PATH):
      raise ImportError(f"{LIB_PATH} is missing: build it with `python -m ceregraphs_amd.build`")
    # PyTorch ships its own HIP runtime whose soname matches the one this
    # library links (libamdhip64.so.7).  Loading torch first makes the
    # dynamic linker bind this library to that same runtime, so a process
    # that also uses torch tensors has exactly one HIP runtime.
    try:
      import torch  # noqa: F401  pylint: disable=unused-import,import-outside-toplevel
    except ImportError:
      pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
      fn = getattr(lib, name)
      fn.restype = res
      fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
  if rc != CGBFS_OK:
    msg = load().cgbfs_last_error()
    raise CgbfsError(rc, msg.decode() if msg else "")
  return rc


def device_count():
  c = ctypes.c_int32(0)
  check(load().cgbfs_device_count(ctypes.byref(c)))
  return c.value
