"""CereGraphs-MI355X: a gfx950-native BFS engine behind the CereGraphs host
surface (run_bfs / run_graph500 style drivers, Matrix Market input, balanced
.operm permutations, bfs_timing.csv rows, parent/level outputs).

The searches run in hand-written HIP kernels (libcgbfs.so, C ABI in
include/cgbfs.h) loaded through ctypes; see engine.DeviceGraph.
"""
__version__ = "0.1.0"
