// pybind_wtivo_gpupr.cpp -- Python module `wtivo_gpupr`, the drop-in for the
// reference's torch extension of the same name (wtivo_gpupr_bindings.cpp:579-600).
//
// graph_cut_fast / graph_cut_gpupr_fast_v630 keep the reference signature,
// defaults, argument checks, messages and 18-tuple result
// (wtivo_gpupr_bindings.cpp:100-137,377-396,581-596).  Tensors are accessed
// through their Python attributes (is_cuda, dtype, is_contiguous(), dim(),
// size(), data_ptr()), so the module needs no libtorch; all device work goes
// through the C-ABI of libwtv_gpupr.so (include/wtv_gpupr.h).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "wtv_gpupr.h"

namespace py = pybind11;

namespace {

// WTV_EOVERFLOW -> OverflowError, any other failure -> RuntimeError.
void check(int rc) {
    if (rc == WTV_OK) return;
    const std::string msg = wtv_last_error();
    if (rc == WTV_EOVERFLOW) {
        PyErr_SetString(PyExc_OverflowError, msg.c_str());
        throw py::error_already_set();
    }
    throw std::runtime_error(msg);
}

bool quiet() {
    const char* e = std::getenv("WTV_QUIET");
    return e && *e && *e != '0';
}

// torch.cuda.current_stream().cuda_stream when torch is loaded, else NULL.
void* current_stream() {
    try {
        py::module_ sys = py::module_::import("sys");
        if (!sys.attr("modules").contains("torch")) return nullptr;
        py::object torch = py::module_::import("torch");
        py::object s = torch.attr("cuda").attr("current_stream")();
        return reinterpret_cast<void*>(s.attr("cuda_stream").cast<uintptr_t>());
    } catch (...) {
        return nullptr;
    }
}

bool is_int32(const py::object& t) { return py::str(t.attr("dtype")).cast<std::string>() == "torch.int32"; }
bool is_uint32(const py::object& t) { return py::str(t.attr("dtype")).cast<std::string>() == "torch.uint32"; }

py::tuple graph_cut_fast(py::array_t<double, py::array::c_style | py::array::forcecast> vertices,
                         py::object tets_cuda, py::object nbr_cuda,
                         py::array_t<std::uint8_t, py::array::c_style | py::array::forcecast> labels,
                         double fill_T, int threads, int global_relabel_period, std::uint64_t max_rounds,
                         int local_steps) {
    // argument checks in the reference's order and wording (:114-132)
    if (vertices.ndim() != 2 || vertices.shape(1) != 3) throw std::runtime_error("vertices must be float64 Nx3");
    if (!tets_cuda.attr("is_cuda").cast<bool>() || !nbr_cuda.attr("is_cuda").cast<bool>())
        throw std::runtime_error("tets/neighbors must be CUDA tensors");
    if (!is_int32(tets_cuda) || !is_int32(nbr_cuda)) throw std::runtime_error("tets/neighbors must be int32");
    if (!tets_cuda.attr("is_contiguous")().cast<bool>() || !nbr_cuda.attr("is_contiguous")().cast<bool>())
        throw std::runtime_error("tets/neighbors must be contiguous");
    if (tets_cuda.attr("dim")().cast<int>() != 2 || nbr_cuda.attr("dim")().cast<int>() != 2 ||
        tets_cuda.attr("size")(1).cast<long long>() != 4 || nbr_cuda.attr("size")(1).cast<long long>() != 4 ||
        tets_cuda.attr("size")(0).cast<long long>() != nbr_cuda.attr("size")(0).cast<long long>())
        throw std::runtime_error("tets/neighbors must be Nx4 and same N");
    const size_t nv = static_cast<size_t>(vertices.shape(0));
    const size_t N = static_cast<size_t>(tets_cuda.attr("size")(0).cast<long long>());
    if (labels.ndim() != 1 || static_cast<size_t>(labels.shape(0)) != N)
        throw std::runtime_error("labels size mismatch");
    const auto* dT = reinterpret_cast<const int32_t*>(tets_cuda.attr("data_ptr")().cast<uintptr_t>());
    const auto* dN = reinterpret_cast<const int32_t*>(nbr_cuda.attr("data_ptr")().cast<uintptr_t>());
    void* stream = current_stream();

    py::array_t<std::uint8_t> out(static_cast<py::ssize_t>(N));
    wtv_cut_result r{};
    int rc;
    {
        const double* V = vertices.data();
        const std::uint8_t* lp = labels.data();
        std::uint8_t* op = out.mutable_data();
        py::gil_scoped_release release;
        rc = wtv_graph_cut_fast(V, nv, dT, dN, N, lp, fill_T, threads, global_relabel_period, max_rounds,
                                local_steps, op, &r, stream);
    }
    check(rc);
    if (!quiet()) {
        std::printf("[GPUPr-FAST-v6.30] solve=%.6gs | macro_rounds=%llu | local_steps=%lld | global_relabels=%llu"
                    " | flow=%lld | graph build+upload=%.6gs | free nodes=%llu\n",
                    r.solve_seconds, static_cast<unsigned long long>(r.macro_rounds),
                    static_cast<long long>(r.local_steps), static_cast<unsigned long long>(r.global_relabels),
                    static_cast<long long>(r.adjusted_flow), r.build_upload_seconds,
                    static_cast<unsigned long long>(r.graph_bytes / 36u));
        std::fflush(stdout);
    }
    return py::make_tuple(out, py::int_(r.adjusted_flow), py::int_(r.raw_flow), py::int_(r.graph_bytes),
                          py::int_(r.workspace_bytes), py::int_(r.full_edges), py::int_(r.free_edges),
                          py::int_(r.fixed_source), py::int_(r.fixed_sink), py::int_(r.terminal_edges),
                          py::int_(r.threads), py::int_(r.macro_rounds), py::int_(r.global_relabels),
                          py::int_(r.map_bytes), py::float_(r.build_upload_seconds), py::int_(r.local_steps),
                          py::float_(r.solve_seconds), py::float_(r.native_total_seconds));
}

// surface_extraction_topology_cuda (wtivo_gpupr_bindings.cpp:400-577): same
// checks and messages (:410-423); returns (vertices float64 [u,3], faces int32 [F,3]).
py::tuple surface_extraction(py::array_t<std::uint8_t, py::array::c_style | py::array::forcecast> labels,
                             py::array_t<double, py::array::c_style | py::array::forcecast> vertices,
                             py::object tets_cuda, py::object nbr_cuda, int threads) {
    if (vertices.ndim() != 2 || vertices.shape(1) != 3) throw std::runtime_error("vertices must be float64 Nx3");
    if (!tets_cuda.attr("is_cuda").cast<bool>() || !nbr_cuda.attr("is_cuda").cast<bool>() || !is_int32(tets_cuda) ||
        !is_int32(nbr_cuda) || !tets_cuda.attr("is_contiguous")().cast<bool>() ||
        !nbr_cuda.attr("is_contiguous")().cast<bool>() || tets_cuda.attr("dim")().cast<int>() != 2 ||
        nbr_cuda.attr("dim")().cast<int>() != 2 || tets_cuda.attr("size")(1).cast<long long>() != 4 ||
        nbr_cuda.attr("size")(1).cast<long long>() != 4 ||
        tets_cuda.attr("size")(0).cast<long long>() != nbr_cuda.attr("size")(0).cast<long long>())
        throw std::runtime_error("tets/neighbors must be contiguous CUDA int32 Nx4 with same N");
    const size_t nv = static_cast<size_t>(vertices.shape(0));
    const size_t N = static_cast<size_t>(tets_cuda.attr("size")(0).cast<long long>());
    if (labels.ndim() != 1 || static_cast<size_t>(labels.shape(0)) != N)
        throw std::runtime_error("labels size mismatch");
    const auto* dT = reinterpret_cast<const int32_t*>(tets_cuda.attr("data_ptr")().cast<uintptr_t>());
    const auto* dN = reinterpret_cast<const int32_t*>(nbr_cuda.attr("data_ptr")().cast<uintptr_t>());
    void* stream = current_stream();
    wtv_surface surf{};
    int rc;
    const auto t0 = std::chrono::steady_clock::now();
    {
        const std::uint8_t* lp = labels.data();
        const double* V = vertices.data();
        py::gil_scoped_release release;
        rc = wtv_surface_extract(lp, V, nv, dT, dN, N, &surf, stream);
    }
    check(rc);
    py::array_t<double> OV({static_cast<py::ssize_t>(surf.n_vertices), py::ssize_t(3)});
    py::array_t<std::int32_t> OF({static_cast<py::ssize_t>(surf.n_faces), py::ssize_t(3)});
    if (surf.n_vertices) std::memcpy(OV.mutable_data(), surf.vertices, surf.n_vertices * 3u * sizeof(double));
    if (surf.n_faces) std::memcpy(OF.mutable_data(), surf.faces, surf.n_faces * 3u * sizeof(std::int32_t));
    wtv_surface_free(&surf);
    if (!quiet()) {
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("[WTiVo-Surface] threads=%d | faces=%zu | topology on device | total=%.6gs\n", std::max(1, threads),
                    static_cast<size_t>(OF.shape(0)), dt);
        std::fflush(stdout);
    }
    return py::make_tuple(OV, OF);
}

py::dict stats_dict(const wtv_stats& s) {
    py::dict d;
    d["solve_seconds"] = s.solve_seconds;
    d["final_relabel_seconds"] = s.final_relabel_seconds;
    d["solver_workspace_bytes"] = s.solver_workspace_bytes;
    d["rounds"] = s.rounds;
    d["global_relabels"] = s.global_relabels;
    d["raw_flow"] = s.raw_flow;
    d["visits"] = s.visits;
    d["pushes"] = s.pushes;
    d["relabels"] = s.relabels;
    d["bfs_pops"] = s.bfs_pops;
    d["sub_rounds"] = s.sub_rounds;
    d["bfs_levels"] = s.bfs_levels;
    d["t_solve_ms"] = s.t_solve_ms;
    d["t_bfs_ms"] = s.t_bfs_ms;
    d["t_push_ms"] = s.t_push_ms;
    d["t_relabel_ms"] = s.t_relabel_ms;
    d["t_level_ms"] = s.t_level_ms;
    d["n_push"] = s.n_push;
    d["n_relabel"] = s.n_relabel;
    d["n_level"] = s.n_level;
    d["bfs_bu_levels"] = s.bfs_bu_levels;
    d["bfs_bu_checks"] = s.bfs_bu_checks;
    d["t_tail_ms"] = s.t_tail_ms;
    d["n_tail"] = s.n_tail;
    d["tail_sub_rounds"] = s.tail_sub_rounds;
    d["bfs_seed_checks"] = s.bfs_seed_checks;
    d["t_event_overhead_us"] = s.t_event_overhead_us;
    d["t_order_ms"] = s.t_order_ms;
    d["renumbered"] = s.renumbered;
    d["pull_rounds"] = s.pull_rounds;
    d["fused_rounds"] = s.fused_rounds;
    return d;
}

// Inner boundary: solve a reduced graph already on the device
// (cellocut_gpu_push_relabel_fast, gpu_push_relabel_fast.h:20-28).  res and
// term are updated in place; returns (partition uint8[n] numpy, stats dict).
py::tuple pr_solve(py::object nbr_cuda, py::object res_cuda, py::object term_cuda, int global_relabel_period,
                   std::uint64_t max_rounds, int local_steps, int grid_wgs, int profile) {
    for (auto* t : {&nbr_cuda, &res_cuda, &term_cuda})
        if (!t->attr("is_cuda").cast<bool>() || !t->attr("is_contiguous")().cast<bool>())
            throw std::runtime_error("nbr/res/term must be contiguous CUDA tensors");
    if (!is_int32(nbr_cuda) || !is_int32(term_cuda) || !(is_int32(res_cuda) || is_uint32(res_cuda)))
        throw std::runtime_error("nbr/term must be int32 and res int32 or uint32");
    const long long n = term_cuda.attr("numel")().cast<long long>();
    if (nbr_cuda.attr("numel")().cast<long long>() != 4 * n || res_cuda.attr("numel")().cast<long long>() != 4 * n)
        throw std::runtime_error("nbr/res must hold 4 entries per node");
    if (n > 0x7FFFFFFFLL) throw std::runtime_error(">INT32_MAX reduced nodes unsupported");
    auto* dN = reinterpret_cast<int32_t*>(nbr_cuda.attr("data_ptr")().cast<uintptr_t>());
    auto* dR = reinterpret_cast<uint32_t*>(res_cuda.attr("data_ptr")().cast<uintptr_t>());
    auto* dT = reinterpret_cast<int32_t*>(term_cuda.attr("data_ptr")().cast<uintptr_t>());
    void* stream = current_stream();
    wtv_params p;
    wtv_params_default(&p);
    p.global_relabel_period = global_relabel_period;
    p.max_rounds = max_rounds;
    p.local_steps = local_steps;
    p.grid_wgs = grid_wgs;
    p.profile = profile;
    py::array_t<std::uint8_t> part(static_cast<py::ssize_t>(n));
    wtv_stats st{};
    int rc;
    {
        std::uint8_t* pp = part.mutable_data();
        py::gil_scoped_release release;
        rc = wtv_pr_solve(static_cast<uint32_t>(n), dN, dR, dT, pp, 0, &p, &st, stream);
    }
    check(rc);
    return py::make_tuple(part, stats_dict(st));
}

py::dict last_stats() {
    wtv_stats s{};
    wtv_last_stats(&s);
    return stats_dict(s);
}

}  // namespace

PYBIND11_MODULE(wtivo_gpupr, m) {
    m.doc() = "WTiVo exact graph cut: MI355X push-relabel max-flow (HIP, gfx950)";
    m.def("graph_cut_fast", &graph_cut_fast, py::arg("vertices"), py::arg("tets_cuda"), py::arg("neighbors_cuda"),
          py::arg("labels"), py::arg("fill_T"), py::arg("threads") = 16, py::arg("global_relabel_period") = 1024,
          py::arg("max_rounds") = 2000000ULL, py::arg("local_steps") = 8);
    m.def("graph_cut_gpupr_fast_v630", &graph_cut_fast, py::arg("vertices"), py::arg("tets_cuda"),
          py::arg("neighbors_cuda"), py::arg("labels"), py::arg("fill_T"), py::arg("threads") = 16,
          py::arg("global_relabel_period") = 1024, py::arg("max_rounds") = 2000000ULL, py::arg("local_steps") = 8);
    m.def("pr_solve", &pr_solve, py::arg("nbr_cuda"), py::arg("res_cuda"), py::arg("term_cuda"),
          py::arg("global_relabel_period") = 1024, py::arg("max_rounds") = 2000000ULL, py::arg("local_steps") = 8,
          py::arg("grid_wgs") = 0, py::arg("profile") = 0);
    m.def("surface_extraction_topology_cuda", &surface_extraction, py::arg("labels"), py::arg("vertices"),
          py::arg("tets_cuda"), py::arg("neighbors_cuda"), py::arg("threads") = 16);
    m.def("last_stats", &last_stats);
    m.def("version", []() { return std::string(wtv_version()); });
}
