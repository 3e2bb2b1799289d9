/*
 * wtv_gpupr.h -- C-ABI of the MI355X graph-cut solver (libwtv_gpupr.so).
 *
 * Plain C types only: device buffers are passed as raw pointers, streams as
 * opaque handles (hipStream_t cast to void*, NULL = the null stream).  Every
 * entry point returns a wtv_status code; the message of the last failure on
 * the calling thread is available from wtv_last_error().
 *
 * Reference interfaces these entry points replace:
 *   wtv_pr_solve        <- cellocut_gpu_push_relabel_fast
 *                          (native/gpupr/gpu_push_relabel_fast.h:20-28,
 *                           gpu_push_relabel_fast.cu:406-571)
 *   wtv_graph_cut_fast  <- graph_cut_gpupr_fast_v630, exported to Python as
 *                          wtivo_gpupr.graph_cut_fast and
 *                          wtivo_gpupr.graph_cut_gpupr_fast_v630
 *                          (native/gpupr/wtivo_gpupr_bindings.cpp:100-397,581-596)
 *   wtv_pr_solve_flow   <- cellocut_gpu_push_relabel_fast plus the second phase
 *                          of push-relabel, which the reference leaves out
 *   wtv_batch_solve     <- no single reference function: G independent calls
 *                          of cellocut_gpu_push_relabel_fast solved together
 *   wtv_surface_extract <- surface_extraction_topology_cuda
 *                          (native/gpupr/wtivo_gpupr_bindings.cpp:400-577,597-599)
 *   wtv_server_*        <- no reference counterpart: a serving loop around
 *                          the same solver (continuous batching, HBM admission)
 *   wtv_tet_neighbors   <- the neighbour export of tetrahedralize_neighbors
 *                          (native/core/wtivo_core.cpp:136-154)
 *   wtv_topology_upload <- topology_to_cuda_chunked (wtivo.py:284-309)
 *   wtv_udf_*           <- SparseUDF's grid and SparseUDF.sample_tet_labels
 *                          (native/vdb/wtivo_vdb.cpp:1006-1075,1519-1584)
 */
#ifndef WTV_GPUPR_H
#define WTV_GPUPR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Python mapping (pybind module wtivo_gpupr):
 * WTV_EOVERFLOW -> OverflowError, everything else -> RuntimeError, as the
 * reference's std::overflow_error / std::runtime_error
 * (wtivo_gpupr_bindings.cpp:95-96; gpu_push_relabel_fast.cu:23-29,349-358). */
typedef enum wtv_status {
    WTV_OK = 0,
    WTV_EINVAL = 1,     /* invalid argument or graph                          */
    WTV_EOVERFLOW = 2,  /* capacity outside [0, 2^30)                         */
    WTV_EINVARIANT = 3, /* device invariant (reference codes 2..5 in message) */
    WTV_ENOMEM = 4,     /* insufficient device memory                         */
    WTV_ELIMIT = 5,     /* max_rounds or the 240 s wall limit exceeded        */
    WTV_EDEVICE = 6     /* HIP runtime error                                  */
} wtv_status;

/* Solver statistics.  The first six fields are CellocutGpuPRFastStats
 * (gpu_push_relabel_fast.h:11-18); the rest are this solver's work counters. */
typedef struct wtv_stats {
    double solve_seconds;          /* discharge loop, after the first relabel */
    double final_relabel_seconds;  /* the residual BFS that fixes labels      */
    uint64_t solver_workspace_bytes;
    uint64_t rounds;               /* macro rounds (local_steps sub-rounds)   */
    uint64_t global_relabels;
    int64_t raw_flow;
    /* work counters (device-side, exact) */
    uint64_t visits;     /* discharge vertex-visits                           */
    uint64_t pushes;     /* successful pushes (arc or sink)                   */
    uint64_t relabels;   /* height increases                                  */
    uint64_t bfs_pops;   /* BFS frontier expansions, all global relabels      */
    uint64_t sub_rounds; /* synchronous push/relabel sub-rounds               */
    uint64_t bfs_levels; /* BFS levels expanded, all global relabels          */
    double t_solve_ms;   /* device time: preflow init -> partition written    */
    double t_bfs_ms;     /* device time in global relabels (incl. final)      */
    /* per-kernel device time, filled only when wtv_params.profile != 0 */
    double t_push_ms;    /* sum over discharge push launches                  */
    double t_relabel_ms; /* sum over discharge relabel launches               */
    double t_level_ms;   /* sum over BFS level launches                       */
    uint64_t n_push;     /* launches of each kind                             */
    uint64_t n_relabel;
    uint64_t n_level;
    uint64_t bfs_bu_levels;  /* BFS levels expanded bottom-up                */
    uint64_t bfs_bu_checks;  /* unreached vertices examined bottom-up        */
    double t_tail_ms;        /* one-workgroup discharge launches (profile)   */
    uint64_t n_tail;
    uint64_t tail_sub_rounds; /* sub-rounds run by the one-workgroup kernel  */
    uint64_t bfs_seed_checks; /* term >= 0 vertices examined by the sweep that
                                 expands BFS levels 1 and 2, all relabels    */
    double t_event_overhead_us; /* profile: what the event pair around one launch
                                   adds to it (dispatch and completion latency),
                                   measured on the solver's stream with an empty
                                   kernel of the push launch shape; subtract it
                                   from t_push_ms / n_push for the kernel time  */
    double t_order_ms;          /* the locality test and, when the graph was
                                   renumbered, the renumbering and the copy
                                   back (host wall time)                    */
    uint64_t renumbered;        /* 1: solved as a locality-renumbered copy   */
    uint64_t pull_rounds;       /* macro rounds run as pull-mode sub-rounds
                                   (wtv_params.pull_min)                    */
    uint64_t fused_rounds;      /* macro rounds run as fused push-relabel
                                   sub-rounds (wtv_params.fused)            */
} wtv_stats;

/* Solver knobs.  global_relabel_period, max_rounds and local_steps keep the
 * reference's meaning and clamps (gpu_push_relabel_fast.cu:421-424,476-480):
 * local_steps in [1,32]; period < 0 -> 0 (never); max_rounds 0 -> 2,000,000;
 * a global relabel runs every max(1, ceil(period/local_steps)) macro rounds. */
typedef struct wtv_params {
    int32_t global_relabel_period; /* reference default 1024 */
    int32_t local_steps;           /* reference default 8    */
    uint64_t max_rounds;           /* reference default 2e6  */
    int32_t wg_size;               /* workgroup size of the list kernels (push, relabel,
                                      BFS level, active filter): 0 (= 256), 256, 512 or
                                      1024; other values fail with WTV_EINVAL */
    int32_t grid_wgs;              /* 0 = default (2048)     */
    double time_limit_s;           /* 0 = 240 s (reference)  */
    int32_t profile;               /* 1 = time every kernel with HIP events */
    int32_t no_graph;              /* 1 = launch kernels one by one instead of
                                      replaying a captured macro round      */
    int32_t tail_cap;              /* active lists up to this size run in one
                                      workgroup; 0 = default 512, < 0 = off */
    double relabel_alpha;          /* also relabel globally once the height
                                      increases since the last global relabel
                                      exceed alpha * (active vertices after it)
                                      + 1024 and the time since it has reached
                                      its duration; 0 = default 2.0, < 0 = only
                                      the reference period                    */
    int32_t renumber;              /* single-graph solves: 0 = renumber the nodes
                                      for locality when fewer than half of a
                                      sample of arcs join ids < 4096 apart
                                      (graphs of >= 2^20 nodes), 1 = always,
                                      -1 = never; results are the same      */
    int32_t skip_finished;         /* batched solves: global relabels between
                                      the first and the final one pass over
                                      graphs without an active vertex (0, the
                                      default); < 0 relabels every graph
                                      (same results, more work)             */
    int32_t pull_min;              /* active lists of more than this many
                                      vertices run pull-mode sub-rounds (plain
                                      stores and gathers instead of atomics,
                                      17 B/node more workspace); 0 = default:
                                      off with fused sub-rounds, 4,194,304
                                      with them off (fused < 0); < 0 = off;
                                      results are the same; the
                                      allocation-free solve never uses them */
    int32_t fused;                 /* other multi-workgroup sub-rounds: 0 = one
                                      kernel per sub-round that pushes,
                                      relabels by at most 2 and pushes again
                                      (default); < 0 = a push kernel then a
                                      relabel kernel; results are the same  */
} wtv_params;

/* Fill *p with the reference defaults. */
void wtv_params_default(wtv_params* p);

/*
 * Push-relabel max-flow + final residual BFS on a reduced graph resident in
 * device memory (replaces cellocut_gpu_push_relabel_fast).
 *   n           free nodes (n == 0 returns zero stats)
 *   d_nbr       int32[n*4] neighbour ids, -1 = no arc          (read only)
 *   d_res       uint32[n*4] residual capacities                (in/out)
 *   d_term      int32[n] >0 source supply, <0 sink capacity    (in/out)
 *   partition   uint8[n]: 1 = reaches the sink (outside), 0 = source side.
 *               Host memory when partition_on_device == 0, else device.
 *   stream      hipStream_t or NULL
 */
int wtv_pr_solve(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term,
                 uint8_t* partition, int partition_on_device,
                 const wtv_params* params, wtv_stats* stats, void* stream);

/*
 * wtv_pr_solve that allocates nothing: no device, pinned or host-runtime
 * allocation, no stream, event or graph creation (no reference counterpart;
 * for callers that cannot allocate, such as tight serving loops and harnesses
 * that forbid hipMalloc).  Its workspace is a fixed device array inside the
 * library, one per device, sized for WTV_NOALLOC_MAX_NODES nodes (27.3 MB);
 * a larger n fails with WTV_EINVAL.  The kernels run on `stream` itself and
 * the call returns when they are done.  Results equal wtv_pr_solve's.  The
 * array keeps the last call's contents (each call clears what it reads
 * first), so calls that overlap on one device race: serialise them.
 * Profile mode (params->profile) is not available here (WTV_EINVAL).
 */
#define WTV_NOALLOC_MAX_NODES 1048576u
int wtv_pr_solve_noalloc(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term,
                         uint8_t* partition, int partition_on_device,
                         const wtv_params* params, wtv_stats* stats, void* stream);

/*
 * Batched solve: G independent reduced graphs in one launch sequence.
 *   offsets     uint64[G+1] host: graph g owns nodes [offsets[g], offsets[g+1])
 *               of the concatenated arrays; offsets[0] = 0, total < 2^31
 *   d_nbr       int32[total*4] concatenated neighbour ids, LOCAL to each graph
 *   d_res       uint32[total*4] concatenated residuals (in/out)
 *   d_term      int32[total] concatenated terminals (in/out)
 *   partition   uint8[total] (host or device, as in wtv_pr_solve)
 *   flows       int64[G] host: raw flow of each graph
 * Each graph keeps its own source height (its node count), so every graph's
 * labels and flow equal those of a separate wtv_pr_solve call.
 */
int wtv_batch_solve(uint32_t num_graphs, const uint64_t* offsets, const int32_t* d_nbr,
                    uint32_t* d_res, int32_t* d_term, uint8_t* partition,
                    int partition_on_device, int64_t* flows, const wtv_params* params,
                    wtv_stats* stats, void* stream);

/* The 17 scalars of graph_cut_fast's 18-tuple, in tuple order
 * (wtivo_gpupr_bindings.cpp:377-396); labels are written to labels_out. */
typedef struct wtv_cut_result {
    int64_t adjusted_flow;
    int64_t raw_flow;
    uint64_t graph_bytes;
    uint64_t workspace_bytes;
    uint64_t full_edges;
    uint64_t free_edges;
    uint64_t fixed_source;
    uint64_t fixed_sink;
    uint64_t terminal_edges;
    int64_t threads;
    uint64_t macro_rounds;
    uint64_t global_relabels;
    uint64_t map_bytes;
    double build_upload_seconds;
    int64_t local_steps;
    double solve_seconds;
    double native_total_seconds;
} wtv_cut_result;

/*
 * Full graph cut (replaces graph_cut_gpupr_fast_v630): classify tets, build the
 * reduced graph with the exact capacity formula, fold terminals, solve, and
 * scatter the partition back to tet order.
 *   V          float64[nv*3] host vertices
 *   d_tets     int32[N*4] device tets      (borrowed, not modified)
 *   d_nbr      int32[N*4] device neighbours (borrowed, not modified)
 *   labels     uint8[N] host input labels
 *   labels_out uint8[N] host output labels
 */
int wtv_graph_cut_fast(const double* V, size_t nv, const int32_t* d_tets,
                       const int32_t* d_nbr, size_t N, const uint8_t* labels,
                       double fill_T, int threads, int global_relabel_period,
                       uint64_t max_rounds, int local_steps, uint8_t* labels_out,
                       wtv_cut_result* out, void* stream);

/* The same cut with every array resident in HBM: d_vertices float64[nv*3],
 * d_labels uint8[N] and d_labels_out uint8[N] are device buffers (d_labels_out
 * may equal d_labels: the input labels are read before the output is written).
 * No host copies are made, so a pipeline that keeps its mesh and labels on the
 * device (tet labels from wtv_udf_sample_tet_labels, surface extraction after)
 * skips the vertex upload and both label transfers.  Results and errors are
 * those of wtv_graph_cut_fast; build_upload_seconds then covers the build
 * alone.  No reference counterpart: the reference's interface takes host
 * arrays. */
int wtv_graph_cut_fast_resident(const double* d_vertices, size_t nv, const int32_t* d_tets,
                                const int32_t* d_nbr, size_t N, const uint8_t* d_labels,
                                double fill_T, int global_relabel_period, uint64_t max_rounds,
                                int local_steps, uint8_t* d_labels_out, wtv_cut_result* out,
                                void* stream);

/*
 * wtv_pr_solve followed by an excess return (the second phase of
 * push-relabel, which the reference leaves out: it stops at a maximum
 * preflow).  On return d_res and d_term describe a maximum flow: every vertex
 * conserves, the source side's term is the flow the source still sends into
 * the vertex (>= 0), the sink side's term is as after wtv_pr_solve.  The sink
 * flow (stats->raw_flow), the partition and the cut are exactly
 * wtv_pr_solve's.  *returned (may be NULL) receives the excess sent back to
 * the source.  Uses 25 B per node of extra device memory and about a second
 * solve's time.  WTV_EOVERFLOW when one vertex's excess exceeds int32.
 */
int wtv_pr_solve_flow(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term,
                      uint8_t* partition, int partition_on_device, const wtv_params* params,
                      wtv_stats* stats, int64_t* returned, void* stream);

/* ---------------------------------------------------------------------- */
/* Serving: requests are graphs already resident on the device.  The server  */
/* runs `slots` sessions, each a host thread with its own stream and its own */
/* node arena in HBM (hbm_budget_bytes / slots, at most max_batch_nodes     */
/* nodes).  A session admits waiting requests (oldest first, while arena    */
/* nodes and max_batch graph slots allow) at each global relabel of the     */
/* solve it is running -- continuous batching: a request that arrives while */
/* others are solved joins them on the device -- and each graph leaves at   */
/* the first global relabel after which it has no active vertex, with the   */
/* flow and labels of a lone solve.                                         */
/* ---------------------------------------------------------------------- */
typedef struct wtv_server_config {
    int32_t slots;             /* sessions on the device (default 4)             */
    int32_t max_batch;         /* graphs resident in one session (default 64)    */
    uint64_t max_batch_nodes;  /* arena nodes per session (default 2^27)         */
    uint64_t hbm_budget_bytes; /* all arenas together (default 200 GB)           */
    wtv_params params;         /* solver knobs for every session                 */
} wtv_server_config;

/* Ordering: the solve reads d_nbr, d_res and d_term only after the work that
 * was queued on `stream` when wtv_server_submit was called (the server
 * records an event there and its worker waits for it), so a client may write
 * the graph asynchronously on its own stream and submit without a host sync.
 * With stream == NULL the buffers must be complete at submit (host
 * synchronised).  Either way the client must not touch the four buffers
 * between submit and the return of wtv_server_wait for the ticket. */
typedef struct wtv_request {
    uint32_t n;                /* free nodes                                     */
    int32_t* d_nbr;            /* int32[n*4], read only                          */
    uint32_t* d_res;           /* uint32[n*4], in/out                            */
    int32_t* d_term;           /* int32[n], in/out                               */
    uint8_t* d_partition;      /* uint8[n] device, out                           */
    void* stream;              /* hipStream_t the buffers were written on, or NULL */
} wtv_request;

typedef struct wtv_result {
    int32_t status;            /* wtv_status of the solve                        */
    int32_t batch_size;        /* graphs resident when it joined, itself incl.   */
    int32_t slot;              /* session that solved it                         */
    int32_t pad;
    int64_t raw_flow;
    uint64_t footprint_bytes;  /* arena bytes this request occupied              */
    double submit_s;           /* seconds since server start                     */
    double start_s;
    double end_s;
} wtv_result;

/* Arena bytes a request of n nodes occupies under the default knobs: n rounded
 * up to 1024 nodes, times the arena's bytes per node (graph copy + solver
 * state; a server with pull-mode sub-rounds on adds 17 B per node). */
uint64_t wtv_request_footprint(uint32_t n);

/* Create a server on the current device; NULL on failure (see wtv_last_error). */
void* wtv_server_create(const wtv_server_config* cfg);
/* Queue a request; returns a ticket.  WTV_ENOMEM if it exceeds one session's arena. */
int wtv_server_submit(void* server, const wtv_request* req, uint64_t* ticket);
/* Block until the ticket's request finished and fill *out. */
int wtv_server_wait(void* server, uint64_t ticket, wtv_result* out);
/* Finish all queued work, stop the workers and free the server. */
int wtv_server_destroy(void* server);

/* Counters of the reduced-graph build, as in wtvo's classify/build loops
 * (wtivo_gpupr_bindings.cpp:139-318). */
typedef struct wtv_build_info {
    uint64_t free_n;
    uint64_t fixed_source;
    uint64_t fixed_sink;
    uint64_t full_refs;
    uint64_t free_refs;
    uint64_t terminal_edges;
    int64_t base_flow;
} wtv_build_info;

/*
 * The reduced-graph build of graph_cut_fast on its own, on the device
 * (classify, reduced ids, capacity_exact_formula, terminal folding:
 * wtivo_gpupr_bindings.cpp:139-318).  Outputs are device buffers sized for N
 * nodes; the first free_n entries are written:
 *   d_nbr_out int32[N*4], d_res_out uint32[N*4], d_term_out int32[N],
 *   d_free_old uint32[N] (tet index of each reduced node).
 */
int wtv_build_reduced_graph(const double* V, size_t nv, const int32_t* d_tets,
                            const int32_t* d_nbr, size_t N, const uint8_t* labels,
                            double fill_T, int32_t* d_nbr_out, uint32_t* d_res_out,
                            int32_t* d_term_out, uint32_t* d_free_old,
                            wtv_build_info* info, void* stream);

/* Boundary surface of the label-0 region (surface_extraction_topology_cuda):
 * oriented triangles of label-0 tets on the hull or next to a label != 0 tet,
 * tet-major / slot-minor, with the used vertices compacted in ascending
 * original id.  The library allocates vertices (float64[n_vertices*3]) and
 * faces (int32[n_faces*3]) with malloc; release them with wtv_surface_free. */
typedef struct wtv_surface {
    size_t n_vertices;
    size_t n_faces;
    double* vertices;
    int32_t* faces;
} wtv_surface;

int wtv_surface_extract(const uint8_t* labels, const double* V, size_t nv,
                        const int32_t* d_tets, const int32_t* d_nbr, size_t N,
                        wtv_surface* out, void* stream);
void wtv_surface_free(wtv_surface* s);

/* 4-slot neighbours of a tet mesh, built on the device: d_nbr[i*4+fi] is the
 * tet across face {012,013,023,123}[fi] of tet i (the face opposite local
 * vertex 3-fi, CGAL's neighbor(3-fi) in the reference export) and -1 on the
 * hull.  d_tets int32[N*4] holds vertex ids in [0, nv); d_nbr int32[N*4] is
 * written.  WTV_EINVAL for an out-of-range or repeated vertex id in a tet and
 * for a face held by more than two tets; WTV_ENOMEM when the face table
 * (4 B per entry, a power of two >= 6N entries; 8 B for N >= 2^29) does not
 * fit.  Synchronous on `stream`. */
int wtv_tet_neighbors(const int32_t* d_tets, size_t N, size_t nv, int32_t* d_nbr, void* stream);

/* Upload of an [rows, 4] int32 host matrix into a row-major device array
 * d_out (int32[rows*4]).  Element strides: (4, 1) for C order, (1, rows) for
 * the Fortran order an Eigen MatrixXi exports, anything else is gathered row
 * by row.  Worker threads (threads <= 0: the host's core count, at most 16)
 * pack 262,144-row chunks into pinned buffers kept for the process lifetime
 * (2 x 4 MiB per worker) and copy them on their own streams.  Waits for the
 * work queued on `stream` before writing d_out and returns when every row has
 * arrived. */
int wtv_topology_upload(const int32_t* host, size_t rows, int64_t row_stride, int64_t col_stride,
                        int32_t* d_out, int threads, void* stream);

/* Sparse distance grid in HBM, the grid SparseUDF samples: a float grid with
 * a uniform linear transform (index = world / voxel_size) stored as 8^3
 * leaves.  leaf_origins int32[n_leaves*3] are index-space leaf origins
 * (multiples of 8, distinct); leaf_values float[n_leaves*512] hold each
 * leaf's voxels in OpenVDB order, offset (x&7)<<6 | (y&7)<<3 | (z&7); every
 * voxel outside the leaves reads `background`.  Both arrays are host memory,
 * copied to the device.  The device keeps the values plus a dense int32 index
 * over the leaf cells' bounding box (WTV_ENOMEM beyond 2^31 cells).  On
 * success *udf is a handle for wtv_udf_sample_tet_labels; release it with
 * wtv_udf_destroy. */
int wtv_udf_create(const int32_t* leaf_origins, const float* leaf_values, size_t n_leaves,
                   float background, double voxel_size, void** udf, void* stream);
int wtv_udf_destroy(void* udf);

/* SparseUDF.sample_tet_labels on the device: label = 0 when the trilinear
 * (OpenVDB BoxSampler) sample of the grid at the tet's centroid, mapped into
 * the padded unit box ((c - bbox_min) / (bbox_max - bbox_min) * (1 - 2 *
 * padding) + padding), is <= threshold, else 1.  vertices float64[nv*3] on
 * the host (vertices_on_device = 0, as the reference takes them) or already
 * in HBM (1); d_tets int32[N*4] on the device; labels uint8[N] on the host
 * (labels_on_device = 0) or the device (1).  *seconds (may be NULL) receives
 * the call's wall time, the reference's second result.  WTV_EINVAL for a
 * degenerate bbox or a vertex id outside [0, nv). */
int wtv_udf_sample_tet_labels(const void* udf, const double* vertices, size_t nv, int vertices_on_device,
                              const int32_t* d_tets, size_t N, const double* bbox_min, const double* bbox_max,
                              double padding, double threshold, uint8_t* labels, int labels_on_device,
                              double* seconds, void* stream);

/* Stats of the last wtv_pr_solve / wtv_graph_cut_fast on this thread. */
void wtv_last_stats(wtv_stats* out);

/* Message of the last failure on this thread ("" if none). */
const char* wtv_last_error(void);

/* Library version string. */
const char* wtv_version(void);

#ifdef __cplusplus
}
#endif

#endif /* WTV_GPUPR_H */
