// wtv_internal.h -- shared declarations of libwtv_gpupr (not part of the C-ABI).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "wtv_gpupr.h"

namespace wtv {

// Typed failure carried up to the C-ABI boundary, where it becomes a status
// code plus a thread-local message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

// Message prefix kept from the reference (gpu_push_relabel_fast.cu:26,354) so
// callers that match on it keep working.
constexpr const char* kPrefix = "GPUPr-FAST-v6.30 ";

// Solve a reduced graph already resident on the device, renumbering a
// disordered graph for locality first (wtv_order.hip).
void pr_solve(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term,
              uint8_t* partition, bool partition_on_device, const wtv_params& p,
              wtv_stats& st, void* stream);

// The solve of the graph in the order given (wtv_solver.hip).
void pr_solve_in_order(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term, uint8_t* partition,
                       bool partition_on_device, const wtv_params& p, wtv_stats& st, void* stream);

// pr_solve with its workspace in a file-scope device array: no allocation of
// any kind, n <= WTV_NOALLOC_MAX_NODES, kernels on the caller's stream;
// concurrent calls on one device race (wtv_solver.hip).
void pr_solve_noalloc(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term, uint8_t* partition,
                      bool partition_on_device, const wtv_params& p, wtv_stats& st, void* stream);

// The solve followed by an excess return: residuals and terminals end as a
// maximum flow that conserves at every vertex (wtv_flow.hip).
void pr_solve_flow(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term, uint8_t* partition,
                   bool partition_on_device, const wtv_params& p, wtv_stats& st, int64_t& returned, void* stream);

// Solve G independent graphs as one disjoint union (wtv_solver.hip).
void batch_solve(uint32_t G, const uint64_t* h_off, const int32_t* d_nbr, uint32_t* d_res, int32_t* d_term,
                 uint8_t* partition, bool partition_on_device, int64_t* h_flows, const wtv_params& p,
                 wtv_stats& st, void* stream);

// ------------------------------------------------------------ serving
// One request as an arena session sees it.
struct SessionJob {
    uint64_t ticket = 0;
    wtv_request req{};
    void* ready = nullptr;  // hipEvent_t recorded on the request's stream at submit, or null
    bool alone = false;     // solve with no other graph resident (after a shared solve failed)
};

// The request queue arena sessions pull from (wtv_server.cpp).
class SessionQueue {
  public:
    virtual ~SessionQueue() = default;
    // Oldest-first requests whose node counts, each rounded up to a multiple
    // of 1024, fit one after another into `free_nodes` contiguous nodes and
    // `free_slots` graph slots; an `alone` request is taken only when nothing
    // is resident, and by itself.
    virtual std::vector<SessionJob> take(uint64_t free_nodes, uint32_t free_slots, bool resident_any) = 0;
    // Whether take() would return a request now.
    virtual bool waiting(uint64_t free_nodes, uint32_t free_slots, bool resident_any) = 0;
    // Blocks until a request waits (true) or the queue stops with nothing waiting (false).
    virtual bool wait_nonempty() = 0;
    // A request is done.
    virtual void finish(uint64_t ticket, int status, int64_t flow, int32_t batch_size, int32_t slot, double start_s,
                        double end_s) = 0;
    // Requests a failed shared solve held, back to the front of the queue to be solved alone.
    virtual void requeue_alone(const std::vector<SessionJob>& jobs) = 0;
    // Seconds since the server started.
    virtual double now() const = 0;
};

// Pull-mode sub-rounds run (and their 17 B/node of arrays are allocated) when
// pull_min > 0, or, with the default pull_min 0, when fused sub-rounds are off.
bool pull_enabled(const wtv_params& p);

// Arena bytes per node of a session (rows, workspace, graph ids).
uint64_t arena_bytes_per_node(bool pull);

// Runs an arena session on the calling thread until the queue stops: an
// arena of `capacity` nodes and `slots` graph slots on stream s, fed from q.
void session_serve(SessionQueue& q, const wtv_params& prm, uint32_t capacity, uint32_t slots, void* s,
                   int32_t worker);

// Counters of the reduced-graph build (wtivo_gpupr_bindings.cpp:139-318).
struct BuildInfo {
    uint64_t free_n, fixed_source, fixed_sink, full_refs, free_refs, terminal_edges;
    int64_t base_flow;
};

// Device build of the reduced graph (wtv_graph_build.hip).
void build_reduced_graph(const double* V, size_t nv, const int32_t* dT, const int32_t* dN, size_t N,
                         const uint8_t* labels, double fill_T, int32_t* gN, uint32_t* gR, int32_t* gT,
                         uint32_t* free_old, BuildInfo& info, void* stream);

// Full cut on the device (wtv_graph_build.hip).  resident = vertices, labels
// and output labels are device buffers (else host buffers, as the reference).
void graph_cut_fast(const double* V, size_t nv, const int32_t* dT, const int32_t* dN, size_t N,
                    const uint8_t* lp, double fill_T, int requested_threads, int period, uint64_t max_rounds,
                    int local_steps, uint8_t* out_labels, wtv_cut_result& R, wtv_stats& st, void* stream_v,
                    bool resident = false);

// surface_extraction_topology_cuda on the device (wtv_graph_build.hip).
void surface_extract(const uint8_t* labels, const double* V, size_t nv, const int32_t* dT, const int32_t* dN,
                     size_t N, std::vector<double>& OV, std::vector<int32_t>& OF, void* stream);

// 4-slot tet neighbours by face matching (wtv_topology.hip).
void tet_neighbors(const int32_t* d_tets, size_t N, size_t nv, int32_t* d_nbr, void* stream);

// Pinned, multi-stream upload of an [rows, 4] int32 matrix (wtv_upload.cpp).
void topology_upload(const int32_t* host, size_t rows, int64_t row_stride, int64_t col_stride, int32_t* d_out,
                     int threads, void* stream);

// Sparse distance grid on the device and tet labels from it (wtv_udf.hip).
void* udf_create(const int32_t* origin, const float* values, size_t L, float background, double voxel_size,
                 void* stream);
void udf_destroy(void* udf);
uint64_t udf_bytes(const void* udf);
double udf_sample_tet_labels(const void* udf, const double* V, size_t nv, bool vertices_on_device,
                             const int32_t* d_tets, size_t N, const double* bmin, const double* bmax, double padding,
                             double threshold, uint8_t* labels, bool labels_on_device, void* stream);

// Pinned, multi-stream copies of a byte range (wtv_upload.cpp); synchronous,
// ordered after the work queued on `stream`.
void pinned_to_device(const void* host, size_t bytes, void* d_out, int threads, void* stream);
void pinned_to_host(const void* d_src, size_t bytes, void* host, int threads, void* stream);

// Thread-local stats of the last solve (wtv_capi.cpp).
void set_last_stats(const wtv_stats& st);

}  // namespace wtv
