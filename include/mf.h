/*
 * mf.h - C ABI of the MI355X (gfx950) push-relabel maximum-flow library (libmf.so).
 *
 * The reference solver has no library or FFI layer: its host driver calls C++
 * functions on raw int arrays.  Each entry point below replaces one of those
 * functions (or the CLI step that calls it); the reference file:line it
 * replaces is given beside it.  All paths are relative to the reference tree.
 *
 * Conventions
 *   - Plain C, no exceptions, no exit(): every call returns an mf_status.
 *   - Host memory passed in is owned by the caller and only read during the
 *     call unless documented otherwise.  Device memory is owned by the library.
 *   - All capacities, excesses and heights are int32; flow values are reported
 *     as int64.  Sizes that do not fit the int32 device representation return
 *     MF_ERR_OVERFLOW instead of wrapping.
 *   - One mf_context per host thread; contexts are independent.
 */
#ifndef MF_H
#define MF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MF_VERSION_MAJOR 0
#define MF_VERSION_MINOR 1

/* Replaces the CHECK(x) -> exit(1) error path of maxflow-cuda/include/utils.cuh:18-26
 * and the print-and-continue paths of the loaders (maxflow-cuda/src/graph.cpp:153-156). */
typedef enum mf_status {
  MF_OK = 0,
  MF_ERR_INVALID_ARGUMENT = 1, /* bad pointer, s == t, s/t out of range, bad enum */
  MF_ERR_IO = 2,               /* file missing or unreadable */
  MF_ERR_PARSE = 3,            /* malformed file contents */
  MF_ERR_OVERFLOW = 4,         /* a size or an excess does not fit int32 */
  MF_ERR_DEVICE = 5,           /* a HIP runtime call failed */
  MF_ERR_NO_DEVICE = 6,        /* no usable GPU */
  MF_ERR_OUT_OF_MEMORY = 7,    /* device or host allocation failed */
  MF_ERR_ITERATION_LIMIT = 8,  /* solver exceeded mf_options.max_rounds */
  MF_ERR_INTERNAL = 9,         /* a solver invariant failed (e.g. residual s->t path at exit) */
  MF_ERR_BUDGET = 10,          /* serving: request larger than the whole HBM budget */
  MF_ERR_NOT_READY = 11,       /* serving: polled request has not finished */
  MF_ERR_NO_PATH = 12          /* reserved for callers that treat "t unreachable" as an error */
} mf_status;

/* -a flag of maxflow-cuda/src/main.cu:68-70,96-103. */
typedef enum mf_algo {
  MF_ALGO_TC = 0, /* thread-centric: one lane per active vertex (push_relabel_kernel.cu:14-158) */
  MF_ALGO_VC = 1  /* vertex-centric: a lane tile per active vertex (push_relabel_kernel.cu:400-537) */
} mf_algo;

/* Residual layouts; the reference picks one by directory (maxflow-cuda, maxflow-bcsr, maxflow-bcsr2). */
typedef enum mf_layout {
  MF_LAYOUT_RCSR = 0,  /* forward CSR + reverse CSR with flow index (maxflow-cuda/src/graph.cpp:317-402) */
  MF_LAYOUT_BCSR = 1,  /* one CSR with forward and reverse slots (maxflow-bcsr/src/graph.cpp:322-365);
                          stored as interleaved {dst, residual} pairs + a reverse-slot array */
  MF_LAYOUT_BCSR2 = 2  /* BCSR with interleaved {residual, reverse slot} pairs (maxflow-bcsr2/src/graph.cpp:322-365) */
} mf_layout;

/* Read-only view of a directed graph in CSR form (CSRGraph, maxflow-cuda/include/graph.h:28-58). */
typedef struct mf_csr {
  int32_t n;                /* vertices */
  int64_t m;                /* arcs */
  const int32_t *offsets;   /* n+1 */
  const int32_t *dst;       /* m, 0-based */
  const int32_t *cap;       /* m, >= 0; NULL means unit capacities (SNAP inputs) */
  int32_t s, t;             /* source and sink, 0-based */
} mf_csr;

/* Host graph owned by the library, produced by the loaders and generators. */
typedef struct mf_host_graph {
  int32_t n;
  int64_t m;
  int32_t *offsets;         /* n+1 */
  int32_t *dst;             /* m */
  int32_t *cap;             /* m */
  int32_t s, t;             /* -1 when the file does not name them (SNAP) */
  int64_t m_header;         /* arc count declared by the file header */
  int64_t m_processed;      /* arc lines read (CSRGraph::num_edges_processed) */
} mf_host_graph;

typedef struct mf_options {
  int32_t cycles_per_relabel; /* fixed push-relabel cycles between global relabels; 0 = adaptive */
  int32_t tile_lanes;         /* VC lanes per vertex: 0 = by degree bucket, else 1/8/16/32/64/256 */
  int64_t max_rounds;         /* global-relabel rounds before MF_ERR_ITERATION_LIMIT; 0 = default */
  int32_t verbose;            /* 1: print one line per round like push_relabel.cu:121,194 */
  int32_t profile_kernels;    /* N > 0: every N-th push-relabel launch carries start/end events in
                                 its dispatch packet (1 = every launch; each timed launch costs a
                                 cache flush, so sample for low overhead) */
  int32_t bfs_block_frontier; /* frontier size up to which the one-workgroup BFS runs; 0 = default */
  float relabel_time_ratio;   /* adaptive cadence: after a global relabel taking T ms, cycles run
                                 for about ratio * T ms; 0 = default (2) */
  int32_t batch_algo;         /* mf_batch_solve: 0 = TC (default), 1 = VC */
  int32_t batch_mode;         /* mf_batch_solve: 0 = one workgroup per instance (default),
                                 1 = all instances as one packed graph on the whole grid */
  int32_t bfs_direction;      /* global relabel: 0 = bottom-up levels while the frontier is large
                                 (default), 1 = top-down levels only, 2 = every level that does not
                                 run in the one-workgroup kernel is bottom-up */
  uint16_t class_lanes[5];    /* VC with tile_lanes 0: lanes per vertex for the degree classes
                                 <=8, <=16, <=32, <=1024, larger slots; 0 = default (8, 16, 32, 64,
                                 256); allowed 1, 2, 4, 8, 16, 32, 64, 256 (256 = whole block) */
  uint16_t pad0;
  int32_t forward_hops;       /* push forwarding: a tile whose push activates a vertex may discharge
                                 it in the same cycle, up to this many vertices per cycle;
                                 0 = default (2), -1 = off (every activated vertex waits for the
                                 next cycle) */
  int32_t record_workload;    /* 1: sum the busy clock cycles of every wave of the push-relabel
                                 launches (WORKLOAD of maxflow-cuda/src/push_relabel_kernel.cu:25-40,
                                 146-152); read with mf_graph_read_workload */
  int32_t cycle_graphs;       /* 1: issue the push-relabel cycles in chunks captured as hipGraphs
                                 (replayed, one per chunk length, captured once per solve);
                                 0: one launch per cycle */
  int32_t bfs_cooperative;    /* global relabel: 1 = top-down levels of mid-size frontiers run
                                 inside one cooperative launch of a few workgroups with a grid
                                 barrier per level (where the device supports it; for long-path
                                 graphs); 0 = one grid launch per level (default) */
  int32_t reserved[1];
} mf_options;

typedef struct mf_result {
  int64_t value;              /* max-flow value = excess at the sink */
  int64_t source_outflow;     /* excess pushed out of the source by the preflow */
  double solve_ms;            /* device solve, inputs resident in HBM */
  double kernel_ms;           /* push-relabel cycle kernels (HIP events around each batch of cycles) */
  double relabel_ms;          /* global relabel (BFS) */
  double upload_ms;           /* host layout build + host->device copy */
  int64_t rounds;             /* global relabels */
  int64_t cycles;             /* push-relabel cycle launches */
  int64_t active_total;       /* sum over cycles of active vertices processed */
  int64_t slots_scanned;      /* sum over cycles of residual slots scanned */
  int64_t pushes;
  int64_t relabels;
  int64_t bfs_slots;          /* slots scanned by the global relabels */
  int64_t bytes_algorithmic;  /* algorithmic bytes of the push-relabel kernels */
  int64_t bytes_bfs;          /* algorithmic bytes of the global relabels */
  double profiled_kernel_ms;  /* with profile_kernels: summed per-launch event time */
  int64_t profiled_launches;  /* with profile_kernels: number of timed launches (a sample of cycles) */
  int64_t bfs_levels;         /* BFS levels over all global relabels */
  int64_t bfs_grid_levels;    /* of which run as one grid-wide launch each */
  int32_t status;
  int32_t bfs_bottom_up_levels; /* BFS levels run bottom-up */
  int32_t reserved[6];
  int64_t graph_launches;     /* cycle launches replayed from captured graphs */
} mf_result;

typedef struct mf_context mf_context;
typedef struct mf_graph mf_graph; /* device-resident residual graph */

/* ---- library ---- */
const char *mf_status_string(mf_status s);
const char *mf_version(void);
void mf_options_default(mf_options *opts);

/* ---- context: owns one HIP stream and the autotune cache ----
 * Replaces the device setup in maxflow-cuda/src/push_relabel.cu:66-77. */
mf_status mf_context_create(int device, mf_context **out);
mf_status mf_context_destroy(mf_context *ctx);
mf_status mf_device_synchronize(mf_context *ctx);
/* Number of compute units and device name; for reports. */
mf_status mf_device_info(mf_context *ctx, int32_t *num_cu, char *name, size_t name_len);

/* ---- loaders (CSRGraph, maxflow-cuda/src/graph.cpp) ---- */
/* CSRGraph::buildFromTxtFile, graph.cpp:28-67: '# Nodes: N Edges: M' header, unit capacities,
 * per-vertex sorted adjacency. */
mf_status mf_load_snap_txt(const char *path, mf_host_graph **out);
/* CSRGraph::loadFromBinary, graph.cpp:292-314: int32 [n][m][m_processed][dst*m_processed][off*(n+1)]. */
mf_status mf_load_snap_bin(const char *path, mf_host_graph **out);
/* CSRGraph::saveToBinary, graph.cpp:219-233 (same layout; capacities are not stored). */
mf_status mf_save_snap_bin(const char *path, const mf_csr *g);
/* CSRGraph::buildFromDIMACSFile, graph.cpp:151-214: 1-based, s/t from 'n' lines, file order per vertex. */
mf_status mf_load_dimacs(const char *path, mf_host_graph **out);
/* DIMACS writer ('p max', 'n', 'a' lines), 1-based. */
mf_status mf_save_dimacs(const char *path, const mf_csr *g);
/* Front-ends of the paper's evaluation (WBPR_HPEC2025_CameraReady.pdf p.4, "Applications and
 * datasets"), which the reference snapshot does not ship:
 *   mf_load_konect_bipartite  KONECT bipartite edge list ('%' comments, "u v ..." with 1-based
 *                             left and right ids) -> matching network: source 0 -> left 1..L,
 *                             left -> right L+1..L+R, right -> sink L+R+1, capacities 1;
 *   mf_super_terminals        super-source n -> the given sources, the given sinks -> super-sink
 *                             n+1, with arc capacities that never bind (total out-capacity of a
 *                             source, total in-capacity of a sink);
 *   mf_bfs_terminal_pairs     `pairs` disjoint source/sink pairs, each sink drawn (seeded) from the
 *                             farthest quarter of a BFS from its source, then mf_super_terminals.
 *                             sources_out / sinks_out (pairs entries each, optional) receive them;
 *                             MF_ERR_NO_PATH if not enough pairs are found. */
mf_status mf_load_konect_bipartite(const char *path, mf_host_graph **out);
mf_status mf_super_terminals(const mf_csr *g, const int32_t *sources, int32_t ns,
                             const int32_t *sinks, int32_t nt, mf_host_graph **out);
mf_status mf_bfs_terminal_pairs(const mf_csr *g, int32_t pairs, uint64_t seed,
                                int32_t *sources_out, int32_t *sinks_out, mf_host_graph **out);
void mf_host_graph_free(mf_host_graph *g);
mf_csr mf_host_graph_view(const mf_host_graph *g);

/* ---- seeded generators (network/DIMACS/washington, genrmf, GTgraph/R-MAT) ---- */
/* Washington RLevel(rows, levels, r): washington.c:243-279. */
mf_status mf_gen_washington_rlg(int32_t rows, int32_t levels, int32_t r, uint64_t seed, mf_host_graph **out);
/* Genrmf frames: genrmf.c:76-161. */
mf_status mf_gen_genrmf(int32_t a, int32_t b, int32_t c1, int32_t c2, uint64_t seed, mf_host_graph **out);
/* R-MAT quadrant recursion (GTgraph/R-MAT/graph.c:158-182), self-loops and duplicates dropped,
 * capacities uniform [1, cap_max], s = max out-degree, t = max in-degree. */
mf_status mf_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                      int32_t cap_max, uint64_t seed, mf_host_graph **out);
/* Bipartite matching network: |L| left, |R| right, Poisson(mean_deg) distinct arcs per left
 * vertex, unit capacities; vertex 0 = source, 1..L left, L+1..L+R right, L+R+1 = sink. */
mf_status mf_gen_bipartite(int32_t L, int32_t R, double mean_deg, uint64_t seed, mf_host_graph **out);

/* ---- host-side residual layouts (ResidualGraph::buildFromCSRGraph) ----
 * Exposed for layout tests.  Arrays are allocated by the library; free with mf_layout_free.
 *   RCSR : a = roffsets[n+1], b = rdst[m], c = flow_index[m]    (graph.cpp:317-402)
 *   BCSR / BCSR2 : a = slot offsets[n+1], b = slot dst[S], c = reverse slot[S],
 *                  d = initial residual[S], e = forward slot of each arc[m]   (bcsr graph.cpp:322-365) */
typedef struct mf_layout_arrays {
  int32_t n;
  int64_t slots;
  int32_t *a, *b, *c, *d, *e;
} mf_layout_arrays;
mf_status mf_build_layout(const mf_csr *g, mf_layout layout, mf_layout_arrays *out);
void mf_layout_free(mf_layout_arrays *arr);

/* ---- single-instance device solve ----
 * mf_graph_upload + mf_graph_solve replace main.cu:109-196 (allocation, preflow, H2D copies,
 * push_relabel()).  The graph stays resident; every solve restarts from the capacities. */
mf_status mf_graph_upload(mf_context *ctx, const mf_csr *g, mf_layout layout, mf_graph **out);
mf_status mf_graph_solve(mf_context *ctx, mf_graph *g, mf_algo algo, const mf_options *opts,
                         mf_result *res);
/* Flow on every input arc (cap - residual), m entries, after a solve.  The flow is a maximum
 * preflow; with return_excess != 0 the excess stranded at vertices that cannot reach t is sent
 * back to s on the host so the result satisfies conservation everywhere except s and t.
 * flow_out may be NULL for a graph without arcs. */
mf_status mf_graph_read_flow(mf_context *ctx, mf_graph *g, int32_t *flow_out, int32_t return_excess);
/* Host utility: turns a preflow on g (flow[m], excess >= 0 everywhere but s) into a flow of the
 * same value by cancelling flow cycles and returning every stranded excess to s along
 * flow-carrying arcs.  flow is modified in place (NULL allowed when g has no arcs).  Used by
 * mf_graph_read_flow. */
mf_status mf_flow_return_excess(const mf_csr *g, int32_t *flow);
/* Min-cut side after a solve: side[v] = 1 if v can reach t in the residual graph, else 0. */
mf_status mf_graph_read_cut(mf_context *ctx, mf_graph *g, uint8_t *sink_side);
/* The resident residual layout, downloaded in the mf_build_layout format (free with
 * mf_layout_free).  mf_graph_upload builds it on the device; the two builders give identical
 * arrays. */
mf_status mf_graph_read_layout(mf_context *ctx, mf_graph *g, mf_layout_arrays *out);
/* Per-wave busy clock cycles of the last solve run with record_workload (the per-warp execution
 * times maxflow-cuda prints after "Warp execution time:", maxflow-cuda/src/push_relabel.cu:202-218;
 * a wave here is a 64-lane wavefront) and the degree class of each wave's block.  *num_waves is
 * set (0 if nothing was recorded); cycles / wave_class (either may be NULL) need capacity >=
 * *num_waves entries. */
mf_status mf_graph_read_workload(mf_context *ctx, mf_graph *g, uint64_t *cycles,
                                 uint8_t *wave_class, int64_t capacity, int64_t *num_waves);
/* Heights and excesses after a solve (n entries each; either may be NULL). */
mf_status mf_graph_read_state(mf_context *ctx, mf_graph *g, int32_t *height, int32_t *excess);
mf_status mf_graph_destroy(mf_context *ctx, mf_graph *g);

/* One call: upload, solve, optional flow read-back (flow_out may be NULL), release. */
mf_status mf_solve(mf_context *ctx, const mf_csr *g, mf_algo algo, mf_layout layout,
                   const mf_options *opts, mf_result *res, int32_t *flow_out);

/* Same as mf_solve, allocating nothing on the device (no counterpart in the reference; for callers
 * that cannot allocate: code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Every device array of the solve lives in a static 64 MiB array of the library,
 * one per GPU, held until the process ends; its address is taken when a context is first created
 * on the device.  BCSR only (other layouts: MF_ERR_INVALID_ARGUMENT).  Graphs whose arrays need
 * more than 64 MiB (about 33 n + 36 m bytes, e.g. 262,144 vertices and 1.3M arcs) return
 * MF_ERR_OVERFLOW.  cycle_graphs, record_workload and profile_kernels are ignored (off).  Calls
 * on one GPU share its array: calls that overlap on one GPU race, so issue them from one thread
 * (or serialise them) per device. */
mf_status mf_solve_noalloc(mf_context *ctx, const mf_csr *g, mf_algo algo, mf_layout layout,
                           const mf_options *opts, mf_result *res, int32_t *flow_out);

/* Autotuner (no reference counterpart): trial solves of the resident graph g with candidate tile
 * widths (class_lanes, VC only), relabel cadences (relabel_time_ratio, or a fixed
 * cycles_per_relabel scaled from the cycles per round the starting options ran), forwarding
 * (forward_hops), graph replay (cycle_graphs) and one-workgroup BFS
 * limits (bfs_block_frontier), starting from *base (NULL = defaults), by coordinate descent over
 * the mean of three solves per candidate (the current best's estimate is the mean of all its
 * solves, re-timed before each sweep), with at least two sweeps over all parameters and a third
 * if the second changed one; the winner is then timed against the starting options (six solves
 * each, alternating) and the faster set is kept.  *best receives it, *best_ms its mean solve time,
 * *trials the number of solves run (either may be NULL).  Every trial must produce the same flow
 * value, else MF_ERR_INTERNAL. */
mf_status mf_autotune(mf_context *ctx, mf_graph *g, mf_algo algo, const mf_options *base,
                      mf_options *best, double *best_ms, int32_t *trials);

/* Drop-in for the reference host->solver interface on RCSR arrays:
 * preflow() (maxflow-cuda/src/preflow.cu:3-31) followed by push_relabel()
 * (maxflow-cuda/src/push_relabel.cu:35-221), declared at maxflow-cuda/include/parallel_graph.cuh:44-54.
 * Same host arrays and meaning: on return h, e, fflows, bflows hold the final state and e[t] the
 * max-flow value.  flow_idx is used as given.  Two differences from the reference:
 *   - the solve always starts from the capacities: the caller's cpu_fflows / cpu_bflows, heights
 *     and excesses on entry are ignored (preflow() resets them in the reference too, except that
 *     it keeps the flows of arcs out of other vertices; here every arc starts at residual = cap);
 *   - *excess_total receives e[t], the flow value.  The reference's Excess_total is the preflow's
 *     total minus the excess its global relabel found stranded (push_relabel.cu:114-119,
 *     global_relabel.cu:143-177), which equals e[t] only once every stranded unit is counted;
 *     here stranded excess stays where it is (a maximum preflow) and the value is reported.
 * With E == 0 the per-arc arrays may be NULL. */
mf_status mf_rcsr_push_relabel(mf_context *ctx, int algo_type, int V, int E, int source, int sink,
                               int *cpu_height, int *cpu_excess_flow,
                               const int *cpu_offsets, const int *cpu_destinations,
                               const int *cpu_capacities, int *cpu_fflows, int *cpu_bflows,
                               const int *cpu_roffsets, const int *cpu_rdestinations,
                               const int *cpu_flow_idx, int *excess_total,
                               const mf_options *opts, mf_result *res);

/* ---- batched solve (no reference counterpart) ----
 * k independent instances packed into one concatenated residual graph with per-instance
 * offsets.  opts->batch_mode 0 (default): a persistent kernel whose workgroups pull whole
 * instances from a device queue, largest first, and solve each inside the workgroup;
 * batch_mode 1: the grid-wide kernels run over all instances at once.  opts->batch_algo picks
 * TC (0, default) or VC (1) cycles.  values[k] receives each value. */
mf_status mf_batch_solve(mf_context *ctx, const mf_csr *inst, int32_t k, const mf_options *opts,
                         int64_t *values, mf_result *total);
/* Autotuner of the batched solve: trial batched solves of a sample of at most `sample` instances
 * (the largest ones, which set the tail of a persistent launch, and an even spread of the rest)
 * over batch_algo and relabel_time_ratio; *best receives the fastest options. */
mf_status mf_batch_autotune(mf_context *ctx, const mf_csr *inst, int32_t k, int32_t sample,
                            const mf_options *base, mf_options *best, double *best_ms,
                            int32_t *trials);

/* ---- serving (no reference counterpart) ----
 * Continuous batching.  The server allocates its device arena once, from half of
 * hbm_budget_bytes, and keeps a persistent serving kernel running while requests arrive.
 * Submitted graphs are copied (the caller's arrays may be freed after submit).  A request of at
 * most 16,384 vertices and 262,144 arcs is laid out on the host at submit, admitted as soon as
 * arena ranges and a queue entry are free, copied into its ranges and published into the
 * running kernel's queue, where the next free workgroup solves it inside the running batch; its
 * ranges are reused when it finishes.  A larger request runs alone on the whole grid from one of
 * max_inflight_launches worker contexts (each with its own stream), with device memory charged
 * to the other half of the budget (mf_footprint_bytes); those solves overlap on the device with
 * the serving kernel.  Admission goes through pending requests in arrival order and admits each
 * one that fits now.  Every call takes the server's lock, so calls may come from several
 * threads. */
typedef struct mf_server mf_server;
typedef struct mf_server_stats {
  int64_t submitted, admitted, completed, rejected;
  int64_t launches;               /* serving-kernel launches + requests solved alone */
  int64_t arena_bytes, arena_peak_bytes;  /* budget in use now / at most: arena ranges of
                                             batched requests + device memory of the others */
  int64_t batched, solo;          /* completed requests, by path */
  int64_t serve_kernel_launches;  /* launches of the persistent serving kernel */
} mf_server_stats;
/* Host clock of a completed request, in ms since the server was created. */
typedef struct mf_server_times {
  double submit_ms;   /* mf_server_submit returned */
  double admit_ms;    /* admitted: arena ranges allocated and its copy queued, or its own launch */
  double start_ms;    /* published into the running batch, or its launch started */
  double finish_ms;   /* the server saw its result */
  double device_ms;   /* the request's own solve on the device */
  int32_t batched;    /* 1: solved inside the running batch; 0: alone on the whole grid */
  int32_t reserved;
} mf_server_times;
mf_status mf_server_create(mf_context *ctx, int64_t hbm_budget_bytes, int32_t max_inflight_launches,
                           mf_server **out);
/* Submit a request; *request_id receives its id.  MF_ERR_BUDGET if it can never fit. */
mf_status mf_server_submit(mf_server *srv, const mf_csr *g, int64_t *request_id);
/* Advance the loop: collect finished requests, admit what fits, publish finished copies into
 * the running batch, launch the serving kernel again if no workgroup of it is left.
 * Non-blocking unless wait != 0 (then returns once at least one request completes or nothing is
 * pending or in flight). */
mf_status mf_server_progress(mf_server *srv, int32_t wait);
/* Result of a request: MF_ERR_NOT_READY until it has completed. */
mf_status mf_server_result(mf_server *srv, int64_t request_id, int64_t *value, double *device_ms);
/* Completion queue: up to cap ids of requests that completed (or failed) since the previous
 * call, in completion order; *count receives how many.  A loop over many requests in flight
 * asks this instead of polling every id. */
mf_status mf_server_completed(mf_server *srv, int64_t *ids, int32_t cap, int32_t *count);
/* Times of a completed request (MF_ERR_NOT_READY until it has completed). */
mf_status mf_server_request_times(mf_server *srv, int64_t request_id, mf_server_times *out);
/* Latency (finish - submit) percentiles over the completed requests (linear interpolation
 * between ranks), and completed requests per second from the first submit to the last finish.
 * MF_ERR_NOT_READY while none has completed. */
mf_status mf_server_latency(mf_server *srv, double *p50_ms, double *p99_ms, double *solves_per_s);
mf_status mf_server_stats_get(mf_server *srv, mf_server_stats *st);
/* Finishes the requests already on the GPU, drops pending ones, frees everything. */
mf_status mf_server_destroy(mf_server *srv);
/* Device bytes of a graph solved alone: the larger of the upload's peak (layout, raw CSR copy
 * and device-build buffers) and what stays resident (the admission footprint of large requests). */
int64_t mf_footprint_bytes(int32_t n, int64_t m);

#ifdef __cplusplus
}
#endif
#endif /* MF_H */
