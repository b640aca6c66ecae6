/*
 * tga.h — C ABI of the MI355X temporal-graph engine (libtga.so, gfx950).
 *
 * The reference exposes no library: each algorithm is a standalone CUDA
 * executable whose main() loads one dataset chosen by compile-time macros
 * (gpu_cuda/shared/bfs.cu:1694-2060, gpu_cuda/<Dataset>/<algo>.cu:1-2,
 * gpu_cuda/shared/dataset_config_base.cuh:1-67).  This header keeps the
 * reference's surface — one entry point per algorithm (bfs,
 * connected_components, degree_centrality, graph_coloring, hits,
 * neighbor_access, pagerank; README.md:43) over the unchanged
 * data/compressed/<Dataset>/ files — and replaces the compile-time dataset
 * macros with a runtime descriptor and exit(1) with status codes.
 *
 * Conventions
 *  - Every call returns TGA_OK (0) or a negative TGA_E* status; tga_last_error()
 *    describes the last failure of the calling thread.  Nothing calls exit().
 *  - Vertex ids are int32 in [0, n), n = MAX_LABEL + 1.  Per-vertex outputs are
 *    caller-allocated arrays of n elements: host memory by default, device
 *    memory of the graph's GPU when `flags` has TGA_OUT_DEVICE.
 *  - Windows are inclusive [start, end] on the int32 timestamps;
 *    TGA_FULL_WINDOW covers every contact.
 *  - A vertex v is active when active_begin <= v < active_end
 *    (isTraversableNode, cpu_java/.../BaselineRunner.java:610-614;
 *    ACTIVE_VERTEX_BEGIN, dataset_config_base.cuh:11-13).  Only active
 *    vertices are traversed and only active neighbors are followed.
 *  - A handle owns one HIP stream; calls on it are synchronous and a handle
 *    must not be used from two threads at once.
 *  - Degree centrality, graph coloring, PageRank and HITS build per-call list
 *    state (up to 24 bytes per contact for PageRank/HITS) in device buffers
 *    the handle keeps, when it fits TGA_LIST_MAX_BYTES (environment, bytes;
 *    default: a quarter of the free device memory at first use); otherwise
 *    they decode the records every round or iteration.  Results are the same
 *    either way (PageRank/HITS: FP64 sums in another order).
 */
#ifndef TGA_H
#define TGA_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TGA_OK = 0,
  TGA_EINVAL = -1,   /* bad argument */
  TGA_EIO = -2,      /* dataset file missing or unreadable (tgaOpenFile exits, path_config.cuh:26-36) */
  TGA_EFORMAT = -3,  /* malformed stream or index: out-of-range neighbor, chain deeper than 3, overrun */
  TGA_ERANGE = -4,   /* dataset outside what the reader supports (log2(0) index width, bfs.cu:1785) */
  TGA_EDEVICE = -5,  /* HIP runtime failure */
  TGA_ENOMEM = -6,
  TGA_EBUSY = -7     /* no free lane (BFS server) */
};

#define TGA_OUT_DEVICE 1u
#define TGA_FULL_WINDOW_START (-2147483647LL - 1)
#define TGA_FULL_WINDOW_END 2147483647LL

typedef struct tga_graph tga_graph;

typedef struct tga_window {
  int64_t start; /* inclusive */
  int64_t end;   /* inclusive */
} tga_window;

/* Runtime replacement for the per-dataset macros MAX_LABEL, MAX_Vertex,
 * ACTIVE_VERTEX_BEGIN, minTimestamp (gpu_cuda/<Dataset>/dataset_config.cuh).
 * Fields <= 0 (or TGA_FROM_FILE) take their value from the dataset files. */
#define TGA_FROM_FILE INT64_MIN
typedef struct tga_dataset_desc {
  int64_t max_label;     /* largest vertex id; <= 0: `nodes` of <name>.properties minus 1 */
  int64_t active_begin;  /* first active id (ACTIVE_VERTEX_BEGIN), default 0 */
  int64_t active_end;    /* one past the last active id; <= 0: max_label + 1 */
  int64_t min_timestamp; /* expected .timestamps header, checked; TGA_FROM_FILE: not checked */
} tga_dataset_desc;

typedef struct tga_graph_info {
  int64_t n;             /* vertices (MAX_LABEL + 1) */
  int64_t contacts;      /* sum of outdegrees */
  int64_t min_timestamp; /* .timestamps header */
  uint64_t graph_bytes;  /* size of <name>.graph */
  uint64_t timestamps_bytes;
  uint64_t index_bytes;  /* the four GPU index files */
  int32_t l_graph;       /* Elias-Fano lower widths (bfs.cu:1785-1786) */
  int32_t l_timestamps;
  int64_t active_begin, active_end;
  int64_t device_bytes;  /* device memory held by the handle */
  int64_t max_degree;
  int64_t chunks;        /* decode work items (vertices plus checkpointed chunks) */
} tga_graph_info;

const char* tga_last_error(void);

/* Loads <root>/data/compressed/<dataset>/ (path layout of
 * dataset_config_base.cuh:44-68, TGA_COMPRESS_ROOT of path_config.cuh:11-20)
 * onto `device` and builds the device-side decode directories.  Replaces the
 * load half of every reference main() (bfs.cu:1734-1935). */
int tga_open(const char* root, const char* dataset, const tga_dataset_desc* desc, int device, tga_graph** out);
/* Same, from an explicit directory and basename. */
int tga_open_dir(const char* dir, const char* basename, const tga_dataset_desc* desc, int device, tga_graph** out);
void tga_close(tga_graph* g);
int tga_info(const tga_graph* g, tga_graph_info* info);

/* Neighbor access (RA): decodes every contact of the listed vertices (all
 * active vertices when nverts == 0) and returns their count and an
 * order-independent checksum: the wrapping 64-bit sum over contacts (x, r, t)
 * of h = (t ^ r*0x9E3779B1 ^ x*0x85EBCA77) * 0x7FEB352D (uint32 arithmetic,
 * x, r, t as uint32).
 * Replaces travelGraph/travel (neighbor_access.cu:1438-1642), which decodes and
 * discards. */
int tga_neighbor_access(tga_graph* g, const int32_t* verts, int64_t nverts, uint64_t* checksum, int64_t* contacts);

/* Neighbor access that allocates nothing: the same results as
 * tga_neighbor_access, with every buffer beyond the caller's inputs and
 * outputs in one fixed device array (3.4 MB per GPU, held until the
 * process ends) instead of device memory the handle allocates.  No
 * counterpart in the reference; it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops).  Limits: at most 65536
 * listed vertices, expanding to at most 262144 work items per launch group
 * (an active sub-range counts as its vertices); beyond them it returns
 * TGA_ERANGE.  Calls that overlap on one GPU share the array and race. */
int tga_neighbor_access_noalloc(tga_graph* g, const int32_t* verts, int64_t nverts, uint64_t* checksum,
                                int64_t* contacts);

/* Degree centrality: count[x] = contacts of active x with an active neighbor
 * and t in the window (multi-edges counted).  centrality (nullable) receives
 * the Java formula (int)(0.85 c + 0.15 c (mean - ws) / (we - ws) + 1) with a
 * truncating integer running mean, or c when ws == we
 * (BaselineRunner.java:300-333; replaces count_degree / compute_centre,
 * degree_centrality.cu:1606-1633). */
int tga_degree_centrality(tga_graph* g, tga_window w, int32_t* count, int32_t* centrality, uint32_t flags);

/* Batched degree centrality: nq (1..64) windows in one call.  count (and
 * centrality, nullable) receive nq rows of n: row q is exactly what
 * tga_degree_centrality(g, windows[q], ...) returns.  The contacts are
 * decoded once per handle into a window-independent time cache (8 bytes per
 * contact, under TGA_LIST_MAX_BYTES) that every query of the batch scans; the
 * running means of long records run one wave per (record, query).  No
 * counterpart in the reference (one window per executable run). */
int tga_degree_centrality_batch(tga_graph* g, const tga_window* windows, int32_t nq, int32_t* count,
                                int32_t* centrality, uint32_t flags);

/* BFS levels: hop distance from src over window contacts between active
 * vertices; INT32_MAX when unreached (BaselineRunner.java:339-383; replaces
 * simpleBfs + host level loop, bfs.cu:1657-1692, :1999-2038). */
int tga_bfs(tga_graph* g, tga_window w, int32_t src, int32_t* level, uint32_t flags);

/* Batched BFS: nq (1..64) independent queries, query q from src[q] over
 * windows[q], decoded together -- one pass over the frontier union per level,
 * the queries carried as bit lanes.  levels receives nq rows of n: row q is
 * exactly what tga_bfs(g, windows[q], src[q], row, flags) returns.  No
 * counterpart in the reference (one query per executable run); it is the
 * batched form of bfs (simpleBfs, bfs.cu:1657-1692) for windowed query mixes. */
int tga_bfs_batch(tga_graph* g, const tga_window* windows, const int32_t* src, int32_t nq, int32_t* levels,
                  uint32_t flags);

/* BFS server: continuous batching of BFS queries on one handle (no
 * counterpart in the reference).  Up to 64 queries run at once as bit lanes.
 * tga_bfs_server_admit queues a query into a free lane (TGA_EBUSY when all 64
 * lanes are taken); each tga_bfs_server_step admits the queued queries and
 * advances every running query by one BFS level in one decode pass, setting
 * the bits of the lanes whose query finished in `finished`;
 * tga_bfs_server_result copies a finished lane's levels (exactly tga_bfs's
 * output for that query, whatever ran beside it) and frees the lane.  The
 * server uses its handle's stream and work queues: do not interleave other
 * calls on the same handle from another thread. */
typedef struct tga_bfs_server tga_bfs_server;
int tga_bfs_server_open(tga_graph* g, tga_bfs_server** out);
int tga_bfs_server_admit(tga_bfs_server* s, int32_t src, tga_window w, int32_t* lane);
int tga_bfs_server_step(tga_bfs_server* s, uint64_t* finished);
int tga_bfs_server_result(tga_bfs_server* s, int32_t lane, int32_t* level, uint32_t flags);
void tga_bfs_server_close(tga_bfs_server* s);

/* CC and PageRank servers: continuous admission into the 64 columns of
 * tga_connected_components_batch / tga_pagerank_batch (no counterpart in the
 * reference).  admit / step / result / close work as for the BFS server.  A
 * CC step runs a group of label rounds and finishes the lanes that reached
 * their fixpoint; a PR step runs one iteration of every running lane (each
 * with the iteration count it was admitted with; damping is the server's)
 * and finishes the lanes that reached their count.  Results equal tga_connected_components / tga_pagerank_batch's for
 * the same window, whatever ran beside them.  Server state is allocated on
 * the handle (tga_info's device bytes include it); close servers before
 * their handle (after tga_close their calls fail with TGA_EINVAL). */
typedef struct tga_cc_server tga_cc_server;
int tga_cc_server_open(tga_graph* g, tga_cc_server** out);
int tga_cc_server_admit(tga_cc_server* s, tga_window w, int32_t* lane);
int tga_cc_server_step(tga_cc_server* s, uint64_t* finished);
int tga_cc_server_result(tga_cc_server* s, int32_t lane, int32_t* label, uint32_t flags);
void tga_cc_server_close(tga_cc_server* s);
typedef struct tga_pr_server tga_pr_server;
int tga_pr_server_open(tga_graph* g, float damping, tga_pr_server** out);
int tga_pr_server_admit(tga_pr_server* s, tga_window w, int32_t iters, int32_t* lane);
int tga_pr_server_step(tga_pr_server* s, uint64_t* finished);
int tga_pr_server_result(tga_pr_server* s, int32_t lane, float* rank, uint32_t flags);
void tga_pr_server_close(tga_pr_server* s);

/* Earliest arrival (the reference GPU BFS rule, bfs.cu:1452-1471): arrival[src]
 * = max(ws, 0); the first window contact (x, r, t) of each pair relaxes r when
 * arrival[x] <= t < arrival[r]; INT32_MAX when unreached. */
int tga_earliest_arrival(tga_graph* g, tga_window w, int32_t src, int32_t* arrival, uint32_t flags);

/* Latest departure to `target` (replaces the reference's CPU comparison app
 * third_party/temporal_ligra/apps/temporal/LatestDeparture-Compute.C:86-135,
 * rule :27-51, with instantaneous contacts, start = end = t):
 * departure[target] = the window end clamped to int32; departure[x] = the
 * latest t of a contact x -t-> x1 that starts a path to the target over window
 * contacts between active vertices with non-decreasing times; -(INT32_MAX / 2)
 * (the reference's unset value, :116-119) where the target is unreachable.
 * `rounds` (may be NULL) receives the number of pull passes run, the last one
 * changing nothing. */
int tga_latest_departure(tga_graph* g, tga_window w, int32_t target, int32_t* departure, int32_t* rounds,
                         uint32_t flags);

/* Connected components (min-label): label[v] = smallest active u that reaches
 * v over window contacts, -1 for inactive v (equals the component order of
 * BaselineRunner.java:553-608; replaces connected_components.cu:2091-2154). */
int tga_connected_components(tga_graph* g, tga_window w, int32_t* label, uint32_t flags);

/* Batched connected components: nq (1..64) windows.  label receives nq rows
 * of n: row q is exactly what tga_connected_components(g, windows[q], ...)
 * returns.  Min-label rounds over the window-independent run structure of
 * tga_pagerank_batch, the queries as int32 label columns.  No counterpart in
 * the reference (one window per executable run). */
int tga_connected_components_batch(tga_graph* g, const tga_window* windows, int32_t nq, int32_t* label,
                                   uint32_t flags);

/* Graph coloring: round c colors every uncolored active v that has no
 * uncolored active out-neighbor u over window contacts with priority[u] >
 * priority[v] (ties: u < v); at most `rounds` rounds; INT32_MAX when uncolored.
 * priority == NULL uses java.util.Random(123).nextInt(256) per vertex
 * (BaselineRunner.java:386-430; GC_kernel, graph_coloring.cu:1634-1734). */
int tga_graph_coloring(tga_graph* g, tga_window w, int32_t rounds, const int32_t* priority, int32_t* color,
                       uint32_t flags);

/* Temporal PageRank, FP32 storage with FP64 sums: new[v] = (1-d)/N; each
 * active u with du window contacts pushes d*rank[u]/du*(1-exp(-0.1 t)) to each
 * distinct active neighbor (t = its first window contact); normalised by the
 * sum each iteration; N = active vertices.  Stops after iters or when the L1
 * change < tol (tol > 0).  iters_run (nullable) receives the iterations done
 * (pagerank.cu:1606-1660, :1963-2007; BaselineRunner.java:433-469). */
int tga_pagerank(tga_graph* g, tga_window w, int32_t iters, float damping, float tol, float* rank,
                 int32_t* iters_run, uint32_t flags);

/* Batched PageRank: nq (1..64) windows advanced together, `iters` iterations
 * each (no tolerance stop).  rank receives nq rows of n: row q equals
 * tga_pagerank(g, windows[q], iters, damping, 0, row, ...) up to FP64
 * summation order.  The queries are FP32 columns over a window-independent
 * structure of neighbor runs built once per handle (radix-sorted by
 * destination; with the contact-time cache, under TGA_LIST_MAX_BYTES), so an
 * iteration reads every run once for all queries, with no atomics.  No
 * counterpart in the reference (one window per executable run). */
int tga_pagerank_batch(tga_graph* g, const tga_window* windows, int32_t nq, int32_t iters, float damping, float* rank,
                       uint32_t flags);

/* HITS with per-iteration reset and L2 normalisation over distinct active
 * neighbors with a window contact; FP32 storage, FP64 sums
 * (BaselineRunner.java:472-550; replaces hits.cu:1898-1987, :2284-2349). */
int tga_hits(tga_graph* g, tga_window w, int32_t iters, float* hub, float* auth, uint32_t flags);

/* Batched HITS: nq (1..64) windows, `iters` iterations each; hub and auth
 * receive nq rows of n: row q equals tga_hits(g, windows[q], iters, ...) up
 * to FP64 summation order.  FP32 hub and authority columns over the run
 * structure of tga_pagerank_batch (authority over destination-sorted runs,
 * hub over the same runs grouped by source), no atomics.  No counterpart in
 * the reference (one window per executable run). */
int tga_hits_batch(tga_graph* g, const tga_window* windows, int32_t nq, int32_t iters, float* hub, float* auth,
                   uint32_t flags);

/* Batched graph coloring: up to 64 windows, one priority array for all;
 * row q of `color` (nq rows of MAX_VERTEX, host or device per flags, like
 * `priority`) equals tga_graph_coloring(g, windows[q], rounds, priority, ...).
 * The round rule makes a vertex's color one more than the largest color of
 * its outranking window neighbors (0 without any), so the batch computes
 * all windows in one sweep over priority levels, highest first, instead of
 * `rounds` rounds per window.  No counterpart in the reference (one window
 * per GC_kernel run, graph_coloring.cu:1634-1734). */
int tga_graph_coloring_batch(tga_graph* g, const tga_window* windows, int32_t nq, int32_t rounds,
                             const int32_t* priority, int32_t* color, uint32_t flags);

/* Kernel timing on the handle's own stream: while enabled, every launch of a
 * decode kernel (the kernels that walk the compressed streams) is bracketed by
 * HIP events, and tga_profile_read returns the summed elapsed milliseconds and
 * the number of launches since the last reset.  No counterpart in the
 * reference, which times with host clocks (bfs.cu:1999, :2039-2041). */
int tga_profile(tga_graph* g, int enable);
int tga_profile_read(tga_graph* g, double* kernel_ms, int64_t* launches, int reset);

/* Autotuner: times `passes` full neighbor-access passes (after one warm
 * pass) for each candidate burst length of the decode walk -- the contacts
 * decoded per refill of a lane's ring of neighbor runs -- and keeps the
 * fastest for every later call on the handle (results never depend on it).
 * The report lists each candidate and its average pass time, the burst kept
 * and the flat-record checkpoint stride chosen at open (records longer than
 * it are decoded as chunks by several lanes).  No counterpart in the
 * reference, whose decode granularity is fixed (one thread per vertex,
 * bfs.cu:1485-1633). */
#define TGA_TUNE_MAX 8
typedef struct tga_tune_report {
  int32_t candidates;
  int32_t burst;                  /* burst length kept */
  int32_t flat_stride;            /* flat checkpoint stride in use */
  int32_t setting[TGA_TUNE_MAX];  /* candidate burst lengths */
  double pass_ms[TGA_TUNE_MAX];   /* average pass time of each */
} tga_tune_report;
int tga_autotune(tga_graph* g, int32_t passes, tga_tune_report* report);

/* Timestamp decode methods, timed over the codes of the flat records (whole
 * records without references) with HIP events on the handle's stream: the
 * average of `passes` passes after one warm pass in *ms, and the wrapping
 * sum of the decoded times (as uint32) in *sum, equal for both methods.
 * method 0: one lane per record, through the walk's LDS window over the
 * record's column copy (what every decode pass does); method 1: one wave
 * per record, each lane decoding a code at a speculative bit offset and the
 * true codes chained by shuffles, as the reference's
 * gpu_cuda/Yahoo/neighbor_access_supernode_warp.cu:1697-1725 does.  No
 * counterpart in the reference as an entry point (a measurement of two
 * decode designs); needs the column copies (not TGA_TCOL=0). */
int tga_ts_decode_bench(tga_graph* g, int32_t method, int32_t passes, double* ms, uint64_t* sum);

/* Error-path test hook: the nth device allocation the handle makes from now
 * on fails with TGA_ENOMEM, as if hipMalloc had (nth = 0 turns it off).  A
 * call that fails this way leaves no group of first-use buffers half
 * allocated, so the next call succeeds or fails cleanly.  No counterpart in
 * the reference, which exits on any allocation failure (bfs.cu:23-32). */
int tga_debug_fail_alloc(tga_graph* g, int64_t nth);

#ifdef __cplusplus
}
#endif
#endif
