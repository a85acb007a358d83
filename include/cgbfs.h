/*
 * cgbfs.h -- C ABI of the MI355X (gfx950) BFS library libcgbfs.so.
 *
 * This boundary replaces the reference's device/plugin interface for the BFS
 * path: the compiled CSL program behind SdkRuntime, i.e. the compile-time
 * parameters of bfs/implementation/src/layout_bool.csl:18-38, the exported
 * device buffers of layout_bool.csl:84-116 and the entry points
 * f_spmv_iter / f_enable_tsc / f_tic / f_toc / f_sync_hostdevice
 * (layout_bool.csl:118-122) that bfs/scripts/run_bfs.py:288-504 and
 * run_graph500.py:246-375 drive with memcpy_h2d / launch / memcpy_d2h.
 *
 * Plain C types only: pointers, sizes, status codes.  No exceptions cross the
 * ABI.  Every call returns CGBFS_OK (0) or a negative CGBFS_E* code, and
 * cgbfs_last_error() returns a message for the calling thread's last failure.
 *
 * Matrix orientation (kept from bfs/implementation/graph_loader.py:76-90): a
 * nonzero M[r][c] is the edge c -> r.  "out" arrays are column-compressed M
 * (out-neighbours of each vertex, the push view); "in" arrays are
 * row-compressed M (in-neighbours, ascending, the pull view).
 *
 * Threading: calls on one graph handle are serialised by the caller and run
 * on that handle's own HIP stream; distinct handles may be used from
 * distinct threads.
 *
 * Device outputs: a search writes caller-provided device buffers on the
 * handle's own stream, which does not wait for the caller's streams.  Work
 * on other streams that writes or reads those buffers (a fill, a reader of
 * an earlier result) must be complete before the call, e.g. by
 * synchronising that stream.  cgbfs_bfs, cgbfs_msbfs and cgbfs_msbfs_seeded
 * return once their outputs are complete in HBM; a serving engine's query
 * outputs are complete once cgbfs_engine_poll reports the query completed.
 */
#ifndef CGBFS_H
#define CGBFS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGBFS_API __attribute__((visibility("default")))

#define CGBFS_OK 0
#define CGBFS_EINVAL (-1)   /* bad argument (range, size, null pointer)   */
#define CGBFS_ENOMEM (-2)   /* device or pinned-host allocation failed    */
#define CGBFS_EHIP (-3)     /* a HIP runtime call failed                   */
#define CGBFS_ETIMEOUT (-4) /* level cap reached before the search ended  */
#define CGBFS_ENODEV (-5)   /* no usable gfx950 device                     */

/* Direction policies for the level loop. */
#define CGBFS_DIR_BEAMER 0   /* alpha/beta, both ways (default)                  */
#define CGBFS_DIR_TOPDOWN 1  /* push every level                                 */
#define CGBFS_DIR_BOTTOMUP 2 /* pull every level                                 */
#define CGBFS_DIR_TAU 3      /* reference rule: one-way push->pull once nf > tau
                                (bool_pe.csl:1078-1088, run_bfs.py:239-243)      */
#define CGBFS_DIR_BEAMER_LANES 4 /* multi-source: alpha/beta applied to each lane's
                                own counts, a round pulls some lanes and pushes
                                the rest; a single root: CGBFS_DIR_BEAMER        */
/* In a multi-source search CGBFS_DIR_BEAMER decides from the union counts
 * (vertices with a new lane, their out-degree sum) and moves all lanes
 * together; CGBFS_DIR_TAU applies tau to each lane's own count. */

typedef struct cgbfs_graph cgbfs_graph;

/* Per-search options.  Zero-initialised fields take the defaults noted.
 * The last five are the kernels' tuning knobs (what the reference fixes at
 * compile time in layout_bool.csl:18-38); an autotuner may set them per
 * graph.  Values outside the ranges noted give CGBFS_EINVAL. */
typedef struct {
  int32_t direction;  /* CGBFS_DIR_*                                  (0)    */
  int32_t rounds_hint; /* rounds enqueued before the first host poll (12) */
  double alpha;       /* push->pull when m_f > m_u / alpha            (14)   */
  double beta;        /* pull->push when n_f < n / beta               (24)   */
  int64_t tau;        /* CGBFS_DIR_TAU threshold on nf                (0.15n)*/
  int32_t pull_grab;  /* 64-vertex chunks a wavefront takes per grab in a
                         dense pull round, 1..1024 (8 single-root,
                         1 multi-source)                                    */
  int32_t grid_mult;  /* blocks per CU for grid-stride kernels, 1..64
                         (single root: 2 up to 2^21 vertices, 4 up to 2^24,
                         else 8, multi-source: 2 below 2^19, 4 below 2^23,
                         else 8)                                            */
  int64_t max_levels; /* abort with CGBFS_ETIMEOUT beyond this        (n+1)  */
  int32_t capture;    /* capture the round kernels once into a HIP graph and
                         replay it, deciding every level on the device (no
                         per-round host poll, no per-round events); the graph
                         is kept on the handle while the call's arguments
                         match.  Single root: the per-level loop only      (0) */
  int32_t seg_grab;   /* multi-source: 256-entry hub-row segments a wavefront
                         takes per grab, 1..1024                       (32)  */
  int32_t push_wave_degree; /* single root: a wavefront expands each frontier
                         vertex of out-degree >= this alone (smaller rows
                         are laid end to end, 64 edges a step), 1..2^30 (256) */
  int32_t push_big_degree;  /* single root: the whole grid expands frontier
                         vertices of out-degree >= this, >= push_wave_degree
                         and <= 2^30                                    (256) */
  int32_t loop;       /* single root: CGBFS_LOOP_* -- how the levels run  (0) */
} cgbfs_options;

/* How a single-root search runs its levels (cgbfs_options.loop). */
#define CGBFS_LOOP_AUTO 0       /* workgroup up to 4096 vertices, persistent up to
                                   2^15, else kernels                           */
#define CGBFS_LOOP_KERNELS 1    /* a fixed launch sequence per level, each kernel
                                   gated on the device's decision               */
#define CGBFS_LOOP_PERSISTENT 2 /* one cooperative launch runs every level, grid
                                   barriers between a level and its decision    */
#define CGBFS_LOOP_WORKGROUP 3  /* one workgroup runs the whole search, its state
                                   in LDS (graphs of at most 4096 vertices,
                                   larger ones get CGBFS_EINVAL)                */

/* Per-search statistics.  The caller owns the history arrays (capacity
 * max_hist entries each; any of them may be NULL).  Per-round meaning
 * follows bool_pe.csl: nf_history[r] = vertices discovered in round r
 * (:1063-1065); direction_history[r] = 0 push / 1 pull for round r
 * (:1090-1097); rounds_completed = max_level + 1 (:989,1068). */
typedef struct {
  int64_t rounds_completed;
  int64_t max_hist;
  int64_t *nf_history;
  int32_t *direction_history;
  int64_t *edges_scanned;   /* adjacency entries read in round r          */
  int64_t *rows_scanned;    /* vertices whose adjacency round r opened     */
  int64_t *pairs_found;     /* (lane, vertex) pairs reached in round r     */
  double *level_ms;         /* round r's traversal kernels (HIP events)   */
  double total_ms;          /* search time on the graph's stream (HIP events):
                               init -> parents/levels resident in HBM      */
  double output_ms;         /* device->host output copy (host outputs)     */
  int64_t visited_count;    /* single-root only                            */
  int64_t m_edges;          /* GTEPS m for a single root (bfs_timing.py:390-395) */
  /* Multi-source only: a round pulls some lanes and pushes the others
   * (direction_history[r] = 0 push only, 1 pull only, 2 both).  The arrays
   * below give the pull part of round r, and the push part is the total
   * minus it.  Any of them may be NULL. */
  int64_t *pull_lanes;      /* lanes that pulled in round r                */
  int64_t *pull_nf;         /* vertices the pull reached in round r        */
  int64_t *pull_rows;       /* adjacency lists the pull opened             */
  int64_t *pull_edges;      /* adjacency entries the pull read             */
  int64_t *pull_pairs;      /* (lane, vertex) pairs the pull reached       */
  double *pull_ms;          /* round r's pull kernels (HIP events)         */
} cgbfs_stats;

/* --- library ------------------------------------------------------------ */
CGBFS_API const char *cgbfs_last_error(void);
CGBFS_API int cgbfs_version(void);
CGBFS_API int cgbfs_device_count(int32_t *count);

/* --- graph lifecycle (Kernel 1: construction, once per graph) ----------- */

/* Copies a graph into library-owned HBM.  The arrays are either all host
 * memory or all device memory of `device` (e.g. torch tensors' data_ptr);
 * either way the library keeps its own copy and the caller may free theirs
 * on return.  For symmetric != 0 the in_* arrays may be NULL (the out view
 * is shared).  The graph must pass cgbfs_graph_check: offsets run 0..nnz
 * and never decrease, ids lie in [0, n), every row is strictly ascending
 * (the pull takes the first frontier hit of an in-row as the lowest id),
 * and the in view is the transpose of the out view (a symmetric graph's one
 * view equals its transpose); otherwise CGBFS_EINVAL, before any search
 * kernel reads the arrays.  Host arrays are checked on the host before the
 * GPU is touched, device arrays on the device.  Replaces preprocess()
 * (bfs/implementation/preprocess_bool.py:29-229) and the matrix upload of
 * run_bfs.py:338-358. */
CGBFS_API int cgbfs_graph_create(int32_t device, int64_t n, int64_t nnz, const int64_t *out_ptr,
                       const int32_t *out_idx, const int64_t *in_ptr, const int32_t *in_idx,
                       int32_t symmetric, cgbfs_graph **out);

/* The structural check of cgbfs_graph_create on host arrays, without a GPU:
 * CGBFS_OK, or CGBFS_EINVAL with the first defect (view, vertex or entry)
 * in cgbfs_last_error().  The reference trusts the scipy matrices it is
 * given (preprocess_bool.py:29-229); this has no counterpart there. */
CGBFS_API int cgbfs_graph_check(int64_t n, int64_t nnz, const int64_t *out_ptr,
                      const int32_t *out_idx, const int64_t *in_ptr, const int32_t *in_idx,
                      int32_t symmetric);

/* Builds an R-MAT graph on the GPU, edge-identical to
 * datasets/gen_rmat.py:15-44 for the same numpy PCG64 stream.  pcg[0..3] =
 * state_hi, state_lo, inc_hi, inc_lo of numpy.random.default_rng(seed)'s
 * bit generator.  a, ab, abc are the cumulative quadrant thresholds A, A+B,
 * A+B+C as float64.  symmetric != 0: symmetrise + drop self-loops + dedup
 * (gen_rmat.py:34-44).  symmetric == 0: directed edge u->v stored as
 * M[v][u], self-loops dropped, duplicates merged. */
CGBFS_API int cgbfs_graph_create_rmat(int32_t device, int32_t scale, int64_t edgefactor,
                            const uint64_t *pcg, double a, double ab, double abc,
                            int32_t symmetric, cgbfs_graph **out);

CGBFS_API int cgbfs_graph_info(const cgbfs_graph *g, int64_t *n, int64_t *nnz, int32_t *symmetric);

/* Copies the device adjacency back (any pointer may be NULL to skip). */
CGBFS_API int cgbfs_graph_download(const cgbfs_graph *g, int64_t *out_ptr, int32_t *out_idx,
                         int64_t *in_ptr, int32_t *in_idx);

/* Out-degree >= 1 excluding self-loops is the reference's root rule
 * (run_graph500.py:119-120); this returns out-degree per vertex. */
CGBFS_API int cgbfs_graph_out_degree(const cgbfs_graph *g, int64_t *deg);

CGBFS_API int cgbfs_graph_destroy(cgbfs_graph *g);

/* --- search (Kernel 2) ----------------------------------------------------*/

/* Single-root direction-optimising BFS.  parents/levels: int32[n], caller
 * allocated; on_device != 0 means they are device pointers on the graph's
 * device, else host memory.  Unreached = -1, parents[source] = source,
 * parents[v] = lowest-id in-neighbour one level closer.  Replaces
 * runner.launch("f_spmv_iter") plus the parent readback
 * (run_bfs.py:381-409, device_io.py:160-205). */
CGBFS_API int cgbfs_bfs(cgbfs_graph *g, int64_t source, int32_t *parents, int32_t *levels,
              int32_t on_device, const cgbfs_options *opt, cgbfs_stats *stats);

/* cgbfs_bfs with device outputs that allocates nothing: no hipMalloc,
 * hipFree, event or other allocating call, for callers that cannot allocate
 * (code run under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  No counterpart in the reference.  Its workspace (the
 * frontier bitmaps, big-vertex list, control block and round history,
 * 26,396,416 bytes) is one file-scope device array per GPU, held from library load
 * until the process ends and sized for graphs of at most 2^21 vertices:
 * larger graphs get CGBFS_EINVAL (use cgbfs_bfs).  parents/levels: int32[n]
 * device arrays on the graph's device.  Results and statistics equal
 * cgbfs_bfs's for the same source and options; level_ms comes from the
 * device clock (decide to decide), not from events.  Every call on one GPU
 * uses that GPU's one array (each search clears what it needs at its
 * start), so calls that overlap on one GPU -- from different graph handles
 * or threads -- race and must be serialised by the caller. */
CGBFS_API int cgbfs_bfs_noalloc(cgbfs_graph *g, int64_t source, int32_t *parents, int32_t *levels,
                      const cgbfs_options *opt, cgbfs_stats *stats);

/* Multi-source BFS: nroots <= 64 roots packed as the bits of one uint64 word
 * per vertex.  parents: int32[nroots][n]; levels: int16[nroots][n]; lane l
 * is exactly cgbfs_bfs(roots[l]) (levels cast to int16; a search deeper
 * than 32767 levels fails with CGBFS_EINVAL).  m_lanes (may be NULL):
 * int64[nroots], GTEPS m per lane.  stats->nf_history counts vertices with
 * at least one newly reached lane per round.  A round pulls the lanes in
 * pull mode and pushes the rest (see the direction policies); lanes whose
 * frontier is empty take no part. */
CGBFS_API int cgbfs_msbfs(cgbfs_graph *g, const int64_t *roots, int32_t nroots, int32_t *parents,
                int16_t *levels, int32_t on_device, const cgbfs_options *opt,
                cgbfs_stats *stats, int64_t *m_lanes);

/* Multi-source BFS from seeds: seed k starts lane lanes[k] at vertex
 * verts[k] (level 0, parent itself); nlanes <= 64 lanes, any number of
 * seeds per lane.  A lane's levels are distances to its nearest seed.  Seeds
 * of one lane placed in components that no path joins give each component
 * its own single-root search: this is how an uneven batch of several graphs
 * (laid side by side as one disjoint union) runs in one pass, root l of
 * every graph in lane l.  Outputs, stats and m_lanes as for cgbfs_msbfs
 * (m_lanes sums over all of a lane's seeds).  No counterpart in the
 * reference, which runs one search per launch (run_graph500.py:246-375). */
CGBFS_API int cgbfs_msbfs_seeded(cgbfs_graph *g, const int64_t *verts, const int32_t *lanes,
                       int64_t nseeds, int32_t nlanes, int32_t *parents, int16_t *levels,
                       int32_t on_device, const cgbfs_options *opt, cgbfs_stats *stats,
                       int64_t *m_lanes);

/* --- serving engine: continuous batching --------------------------------- */

typedef struct cgbfs_engine cgbfs_engine;

/* A long-lived 64-lane multi-source search over one graph that serves a
 * stream of single-root queries.  A query takes a free lane at the next
 * round boundary, runs interleaved with the queries already in flight (one
 * adjacency pass per round serves them all), and leaves when its frontier
 * empties.  Each query writes its own device buffers, parents int32[n] and
 * levels int16[n], with results identical to cgbfs_bfs from the same root:
 * the lowest-id rule does not depend on the other lanes.  Directions are
 * chosen per lane (CGBFS_DIR_BEAMER acts as CGBFS_DIR_BEAMER_LANES), since
 * lanes are at different depths.  The engine owns its workspace and HIP
 * stream; do not run other searches on the graph while it is live.  No
 * counterpart in the reference, which runs one search per launch
 * (run_bfs.py:288-504); its callers are serving loops. */
CGBFS_API int cgbfs_engine_create(cgbfs_graph *g, const cgbfs_options *opt, cgbfs_engine **out);

/* Queues a query on a free lane (0..63) for the next step.  CGBFS_EINVAL if
 * the lane is busy or a step is in flight (poll it first). */
CGBFS_API int cgbfs_engine_submit(cgbfs_engine *e, int32_t lane, int64_t root, int32_t *parents,
                        int16_t *levels);

/* Enqueues one step on the engine's stream and returns at once: the lanes
 * whose search ended in the previous step are finalised (their outputs
 * become complete), queued queries are seeded, and one BFS round runs for
 * every lane in flight. */
CGBFS_API int cgbfs_engine_step(cgbfs_engine *e);

/* Reports on the step in flight; wait != 0 blocks until it completes.
 * *ready = 0: still running.  Once ready: *completed = lanes whose queries
 * completed in that step (outputs final, lane free again), with their GTEPS
 * m in m_lanes[l] and their round count (max level + 1) in depth_lanes[l]
 * (either array may be NULL, else 64 entries); *busy = lanes still holding
 * a query. */
CGBFS_API int cgbfs_engine_poll(cgbfs_engine *e, int32_t wait, int32_t *ready, uint64_t *completed,
                      uint64_t *busy, int64_t *m_lanes, int32_t *depth_lanes);

CGBFS_API int cgbfs_engine_destroy(cgbfs_engine *e);

/* Blocks until all work queued on the graph's stream has finished. */
CGBFS_API int cgbfs_sync(cgbfs_graph *g);

#ifdef __cplusplus
}
#endif

#endif /* CGBFS_H */
