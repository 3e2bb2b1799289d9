// wtv_capi.cpp -- extern "C" entry points of libwtv_gpupr.so (include/wtv_gpupr.h).
// C++ exceptions never cross this boundary: they become status codes plus a
// thread-local message (wtv_last_error).
#include <cstdlib>
#include <cstring>
#include <exception>
#include <algorithm>
#include <new>
#include <vector>
#include <string>

#include "wtv_internal.h"

namespace wtv {
class Server;
uint64_t request_footprint(uint32_t n, bool pull);
Server* server_new(const wtv_server_config& c);
void server_delete(Server* s);
uint64_t server_submit(Server* s, const wtv_request& r);
wtv_result server_wait(Server* s, uint64_t ticket);
}

namespace {
thread_local std::string g_err;
thread_local wtv_stats g_stats{};

template <class Fn>
int guarded(Fn fn) {
    g_err.clear();
    try {
        fn();
        return WTV_OK;
    } catch (const wtv::Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return WTV_ENOMEM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return WTV_EINVAL;
    } catch (...) {
        g_err = "unknown error";
        return WTV_EINVAL;
    }
}
}  // namespace

namespace wtv {
void set_last_stats(const wtv_stats& st) { g_stats = st; }
}  // namespace wtv

extern "C" {

void wtv_params_default(wtv_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->global_relabel_period = 1024;
    p->local_steps = 8;
    p->max_rounds = 2000000ull;
    p->wg_size = 0;
    p->grid_wgs = 0;
    p->time_limit_s = 0.0;
}

int wtv_pr_solve(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term, uint8_t* partition,
                 int partition_on_device, const wtv_params* params, wtv_stats* stats, void* stream) {
    return guarded([&]() {
        wtv_params p;
        wtv_params_default(&p);
        if (params) p = *params;
        wtv_stats st{};
        wtv::pr_solve(n, d_nbr, d_res, d_term, partition, partition_on_device != 0, p, st, stream);
        wtv::set_last_stats(st);
        if (stats) *stats = st;
    });
}

int wtv_pr_solve_noalloc(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term, uint8_t* partition,
                         int partition_on_device, const wtv_params* params, wtv_stats* stats, void* stream) {
    return guarded([&]() {
        wtv_params p;
        wtv_params_default(&p);
        if (params) p = *params;
        wtv_stats st{};
        wtv::pr_solve_noalloc(n, d_nbr, d_res, d_term, partition, partition_on_device != 0, p, st, stream);
        wtv::set_last_stats(st);
        if (stats) *stats = st;
    });
}

int wtv_batch_solve(uint32_t num_graphs, const uint64_t* offsets, const int32_t* d_nbr, uint32_t* d_res,
                    int32_t* d_term, uint8_t* partition, int partition_on_device, int64_t* flows,
                    const wtv_params* params, wtv_stats* stats, void* stream) {
    return guarded([&]() {
        wtv_params p;
        wtv_params_default(&p);
        if (params) p = *params;
        wtv_stats st{};
        wtv::batch_solve(num_graphs, offsets, d_nbr, d_res, d_term, partition, partition_on_device != 0, flows, p,
                         st, stream);
        wtv::set_last_stats(st);
        if (stats) *stats = st;
    });
}

int wtv_graph_cut_fast(const double* V, size_t nv, const int32_t* d_tets, const int32_t* d_nbr, size_t N,
                       const uint8_t* labels, double fill_T, int threads, int global_relabel_period,
                       uint64_t max_rounds, int local_steps, uint8_t* labels_out, wtv_cut_result* out,
                       void* stream) {
    return guarded([&]() {
        wtv_cut_result r{};
        wtv_stats st{};
        wtv::graph_cut_fast(V, nv, d_tets, d_nbr, N, labels, fill_T, threads, global_relabel_period,
                            max_rounds, local_steps, labels_out, r, st, stream);
        wtv::set_last_stats(st);
        if (out) *out = r;
    });
}

int wtv_pr_solve_flow(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term, uint8_t* partition,
                      int partition_on_device, const wtv_params* params, wtv_stats* stats, int64_t* returned,
                      void* stream) {
    return guarded([&]() {
        wtv_params p{};
        if (params) p = *params;
        else wtv_params_default(&p);
        wtv_stats st{};
        int64_t r = 0;
        wtv::pr_solve_flow(n, d_nbr, d_res, d_term, partition, partition_on_device != 0, p, st, r, stream);
        wtv::set_last_stats(st);
        if (stats) *stats = st;
        if (returned) *returned = r;
    });
}

int wtv_graph_cut_fast_resident(const double* d_vertices, size_t nv, const int32_t* d_tets, const int32_t* d_nbr,
                                size_t N, const uint8_t* d_labels, double fill_T, int global_relabel_period,
                                uint64_t max_rounds, int local_steps, uint8_t* d_labels_out, wtv_cut_result* out,
                                void* stream) {
    return guarded([&]() {
        wtv_cut_result r{};
        wtv_stats st{};
        wtv::graph_cut_fast(d_vertices, nv, d_tets, d_nbr, N, d_labels, fill_T, 1, global_relabel_period,
                            max_rounds, local_steps, d_labels_out, r, st, stream, true);
        wtv::set_last_stats(st);
        if (out) *out = r;
    });
}

int wtv_build_reduced_graph(const double* V, size_t nv, const int32_t* d_tets, const int32_t* d_nbr, size_t N,
                            const uint8_t* labels, double fill_T, int32_t* d_nbr_out, uint32_t* d_res_out,
                            int32_t* d_term_out, uint32_t* d_free_old, wtv_build_info* info, void* stream) {
    return guarded([&]() {
        wtv::BuildInfo bi{};
        wtv::build_reduced_graph(V, nv, d_tets, d_nbr, N, labels, fill_T, d_nbr_out, d_res_out, d_term_out,
                                 d_free_old, bi, stream);
        if (info) {
            info->free_n = bi.free_n;
            info->fixed_source = bi.fixed_source;
            info->fixed_sink = bi.fixed_sink;
            info->full_refs = bi.full_refs;
            info->free_refs = bi.free_refs;
            info->terminal_edges = bi.terminal_edges;
            info->base_flow = bi.base_flow;
        }
    });
}

int wtv_surface_extract(const uint8_t* labels, const double* V, size_t nv, const int32_t* d_tets,
                        const int32_t* d_nbr, size_t N, wtv_surface* out, void* stream) {
    return guarded([&]() {
        if (!out) throw wtv::Error(WTV_EINVAL, "null surface output");
        std::memset(out, 0, sizeof(*out));
        std::vector<double> OV;
        std::vector<int32_t> OF;
        wtv::surface_extract(labels, V, nv, d_tets, d_nbr, N, OV, OF, stream);
        out->n_vertices = OV.size() / 3u;
        out->n_faces = OF.size() / 3u;
        out->vertices = static_cast<double*>(std::malloc(std::max<size_t>(OV.size(), 1) * sizeof(double)));
        out->faces = static_cast<int32_t*>(std::malloc(std::max<size_t>(OF.size(), 1) * sizeof(int32_t)));
        if (!out->vertices || !out->faces) {
            std::free(out->vertices);
            std::free(out->faces);
            std::memset(out, 0, sizeof(*out));
            throw std::bad_alloc();
        }
        if (!OV.empty()) std::memcpy(out->vertices, OV.data(), OV.size() * sizeof(double));
        if (!OF.empty()) std::memcpy(out->faces, OF.data(), OF.size() * sizeof(int32_t));
    });
}

void wtv_surface_free(wtv_surface* s) {
    if (!s) return;
    std::free(s->vertices);
    std::free(s->faces);
    std::memset(s, 0, sizeof(*s));
}

int wtv_tet_neighbors(const int32_t* d_tets, size_t N, size_t nv, int32_t* d_nbr, void* stream) {
    return guarded([&]() { wtv::tet_neighbors(d_tets, N, nv, d_nbr, stream); });
}

int wtv_topology_upload(const int32_t* host, size_t rows, int64_t row_stride, int64_t col_stride,
                        int32_t* d_out, int threads, void* stream) {
    return guarded([&]() { wtv::topology_upload(host, rows, row_stride, col_stride, d_out, threads, stream); });
}

int wtv_udf_create(const int32_t* leaf_origins, const float* leaf_values, size_t n_leaves, float background,
                   double voxel_size, void** udf, void* stream) {
    return guarded([&]() {
        if (!udf) throw wtv::Error(WTV_EINVAL, "null udf handle pointer");
        *udf = nullptr;
        *udf = wtv::udf_create(leaf_origins, leaf_values, n_leaves, background, voxel_size, stream);
    });
}

int wtv_udf_destroy(void* udf) {
    return guarded([&]() { wtv::udf_destroy(udf); });
}

int wtv_udf_sample_tet_labels(const void* udf, const double* vertices, size_t nv, int vertices_on_device,
                              const int32_t* d_tets, size_t N, const double* bbox_min, const double* bbox_max,
                              double padding, double threshold, uint8_t* labels, int labels_on_device,
                              double* seconds, void* stream) {
    return guarded([&]() {
        const double t = wtv::udf_sample_tet_labels(udf, vertices, nv, vertices_on_device != 0, d_tets, N, bbox_min,
                                                    bbox_max, padding, threshold, labels, labels_on_device != 0,
                                                    stream);
        if (seconds) *seconds = t;
    });
}

uint64_t wtv_request_footprint(uint32_t n) {
    wtv_params p;
    wtv_params_default(&p);
    return wtv::request_footprint(n, wtv::pull_enabled(p));
}

void* wtv_server_create(const wtv_server_config* cfg) {
    wtv::Server* srv = nullptr;
    const int rc = guarded([&]() {
        wtv_server_config c{};
        if (cfg) c = *cfg;
        else wtv_params_default(&c.params);
        srv = wtv::server_new(c);
    });
    return rc == WTV_OK ? srv : nullptr;
}

int wtv_server_submit(void* server, const wtv_request* req, uint64_t* ticket) {
    return guarded([&]() {
        if (!server || !req || !ticket) throw wtv::Error(WTV_EINVAL, "null server/request/ticket");
        *ticket = wtv::server_submit(static_cast<wtv::Server*>(server), *req);
    });
}

int wtv_server_wait(void* server, uint64_t ticket, wtv_result* out) {
    return guarded([&]() {
        if (!server || !out) throw wtv::Error(WTV_EINVAL, "null server/result");
        *out = wtv::server_wait(static_cast<wtv::Server*>(server), ticket);
    });
}

int wtv_server_destroy(void* server) {
    return guarded([&]() { wtv::server_delete(static_cast<wtv::Server*>(server)); });
}

void wtv_last_stats(wtv_stats* out) {
    if (out) *out = g_stats;
}

const char* wtv_last_error(void) { return g_err.c_str(); }

const char* wtv_version(void) { return "wtv_gpupr 0.1.0 gfx950"; }

}  // extern "C"
