"""ctypes binding of libwtv_gpupr.so's C-ABI (include/wtv_gpupr.h).

This is the same boundary a ctypes/cgo/JNI host would bind.  Loading fails
loudly when the HIP library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwtv_gpupr.so")

WTV_OK, WTV_EINVAL, WTV_EOVERFLOW, WTV_EINVARIANT, WTV_ENOMEM, WTV_ELIMIT, WTV_EDEVICE = range(7)

# every extern "C" symbol declared in include/wtv_gpupr.h
EXPORTS = ["wtv_params_default", "wtv_pr_solve", "wtv_pr_solve_noalloc", "wtv_pr_solve_flow", "wtv_batch_solve", "wtv_graph_cut_fast",
           "wtv_graph_cut_fast_resident", "wtv_last_stats",
           "wtv_last_error", "wtv_version", "wtv_build_reduced_graph", "wtv_surface_extract",
           "wtv_surface_free", "wtv_request_footprint", "wtv_server_create", "wtv_server_submit",
           "wtv_server_wait", "wtv_server_destroy", "wtv_tet_neighbors", "wtv_topology_upload",
           "wtv_udf_create", "wtv_udf_destroy", "wtv_udf_sample_tet_labels"]


class WtvStats(C.Structure):
    _fields_ = [
        ("solve_seconds", C.c_double), ("final_relabel_seconds", C.c_double),
        ("solver_workspace_bytes", C.c_uint64), ("rounds", C.c_uint64),
        ("global_relabels", C.c_uint64), ("raw_flow", C.c_int64),
        ("visits", C.c_uint64), ("pushes", C.c_uint64), ("relabels", C.c_uint64),
        ("bfs_pops", C.c_uint64), ("sub_rounds", C.c_uint64), ("bfs_levels", C.c_uint64),
        ("t_solve_ms", C.c_double), ("t_bfs_ms", C.c_double),
        ("t_push_ms", C.c_double), ("t_relabel_ms", C.c_double), ("t_level_ms", C.c_double),
        ("n_push", C.c_uint64), ("n_relabel", C.c_uint64), ("n_level", C.c_uint64),
        ("bfs_bu_levels", C.c_uint64), ("bfs_bu_checks", C.c_uint64),
        ("t_tail_ms", C.c_double), ("n_tail", C.c_uint64), ("tail_sub_rounds", C.c_uint64),
        ("bfs_seed_checks", C.c_uint64), ("t_event_overhead_us", C.c_double),
        ("t_order_ms", C.c_double), ("renumbered", C.c_uint64), ("pull_rounds", C.c_uint64),
        ("fused_rounds", C.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class WtvParams(C.Structure):
    _fields_ = [
        ("global_relabel_period", C.c_int32), ("local_steps", C.c_int32),
        ("max_rounds", C.c_uint64), ("wg_size", C.c_int32), ("grid_wgs", C.c_int32),
        ("time_limit_s", C.c_double), ("profile", C.c_int32), ("no_graph", C.c_int32),
        ("tail_cap", C.c_int32), ("relabel_alpha", C.c_double), ("renumber", C.c_int32),
        ("skip_finished", C.c_int32), ("pull_min", C.c_int32), ("fused", C.c_int32),
    ]


class WtvCutResult(C.Structure):
    _fields_ = [
        ("adjusted_flow", C.c_int64), ("raw_flow", C.c_int64), ("graph_bytes", C.c_uint64),
        ("workspace_bytes", C.c_uint64), ("full_edges", C.c_uint64), ("free_edges", C.c_uint64),
        ("fixed_source", C.c_uint64), ("fixed_sink", C.c_uint64), ("terminal_edges", C.c_uint64),
        ("threads", C.c_int64), ("macro_rounds", C.c_uint64), ("global_relabels", C.c_uint64),
        ("map_bytes", C.c_uint64), ("build_upload_seconds", C.c_double), ("local_steps", C.c_int64),
        ("solve_seconds", C.c_double), ("native_total_seconds", C.c_double),
    ]


class WtvServerConfig(C.Structure):
    _fields_ = [("slots", C.c_int32), ("max_batch", C.c_int32), ("max_batch_nodes", C.c_uint64),
                ("hbm_budget_bytes", C.c_uint64), ("params", WtvParams)]


class WtvRequest(C.Structure):
    _fields_ = [("n", C.c_uint32), ("d_nbr", C.c_void_p), ("d_res", C.c_void_p), ("d_term", C.c_void_p),
                ("d_partition", C.c_void_p), ("stream", C.c_void_p)]


class WtvResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("batch_size", C.c_int32), ("slot", C.c_int32), ("pad", C.c_int32),
                ("raw_flow", C.c_int64), ("footprint_bytes", C.c_uint64), ("submit_s", C.c_double),
                ("start_s", C.c_double), ("end_s", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class WtvBuildInfo(C.Structure):
    _fields_ = [("free_n", C.c_uint64), ("fixed_source", C.c_uint64), ("fixed_sink", C.c_uint64),
                ("full_refs", C.c_uint64), ("free_refs", C.c_uint64), ("terminal_edges", C.c_uint64),
                ("base_flow", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class WtvError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libwtv_gpupr.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.wtv_params_default.restype = None
        L.wtv_params_default.argtypes = [C.POINTER(WtvParams)]
        L.wtv_pr_solve.restype = C.c_int
        L.wtv_pr_solve.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                   C.POINTER(WtvParams), C.POINTER(WtvStats), C.c_void_p]
        L.wtv_pr_solve_noalloc.restype = C.c_int
        L.wtv_pr_solve_noalloc.argtypes = L.wtv_pr_solve.argtypes
        L.wtv_pr_solve_flow.restype = C.c_int
        L.wtv_pr_solve_flow.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                        C.POINTER(WtvParams), C.POINTER(WtvStats), C.POINTER(C.c_int64), C.c_void_p]
        L.wtv_batch_solve.restype = C.c_int
        L.wtv_batch_solve.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_int, C.c_void_p, C.POINTER(WtvParams), C.POINTER(WtvStats), C.c_void_p]
        L.wtv_build_reduced_graph.restype = C.c_int
        L.wtv_build_reduced_graph.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_size_t,
                                              C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.POINTER(WtvBuildInfo), C.c_void_p]
        L.wtv_request_footprint.restype = C.c_uint64
        L.wtv_request_footprint.argtypes = [C.c_uint32]
        L.wtv_server_create.restype = C.c_void_p
        L.wtv_server_create.argtypes = [C.POINTER(WtvServerConfig)]
        L.wtv_server_submit.restype = C.c_int
        L.wtv_server_submit.argtypes = [C.c_void_p, C.POINTER(WtvRequest), C.POINTER(C.c_uint64)]
        L.wtv_server_wait.restype = C.c_int
        L.wtv_server_wait.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(WtvResult)]
        L.wtv_server_destroy.restype = C.c_int
        L.wtv_server_destroy.argtypes = [C.c_void_p]
        L.wtv_graph_cut_fast.restype = C.c_int
        L.wtv_graph_cut_fast.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_size_t,
                                         C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_uint64, C.c_int,
                                         C.c_void_p, C.POINTER(WtvCutResult), C.c_void_p]
        L.wtv_tet_neighbors.restype = C.c_int
        L.wtv_tet_neighbors.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p]
        L.wtv_topology_upload.restype = C.c_int
        L.wtv_topology_upload.argtypes = [C.c_void_p, C.c_size_t, C.c_int64, C.c_int64, C.c_void_p, C.c_int,
                                          C.c_void_p]
        L.wtv_udf_create.restype = C.c_int
        L.wtv_udf_create.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_float, C.c_double,
                                     C.POINTER(C.c_void_p), C.c_void_p]
        L.wtv_udf_destroy.restype = C.c_int
        L.wtv_udf_destroy.argtypes = [C.c_void_p]
        L.wtv_udf_sample_tet_labels.restype = C.c_int
        L.wtv_udf_sample_tet_labels.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p,
                                                C.c_size_t, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                                C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_void_p]
        L.wtv_graph_cut_fast_resident.restype = C.c_int
        L.wtv_graph_cut_fast_resident.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_size_t,
                                                  C.c_void_p, C.c_double, C.c_int, C.c_uint64, C.c_int, C.c_void_p,
                                                  C.POINTER(WtvCutResult), C.c_void_p]
        L.wtv_last_stats.restype = None
        L.wtv_last_stats.argtypes = [C.POINTER(WtvStats)]
        L.wtv_last_error.restype = C.c_char_p
        L.wtv_version.restype = C.c_char_p
        _lib = L
    return _lib


def params(**kw) -> WtvParams:
    p = WtvParams()
    lib().wtv_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def check(rc):
    if rc != WTV_OK:
        raise WtvError(rc, lib().wtv_last_error().decode())


def pr_solve(n, d_nbr, d_res, d_term, d_or_h_partition, partition_on_device, prm=None, stream=0):
    """Call wtv_pr_solve with raw device pointers (ints); returns WtvStats."""
    st = WtvStats()
    prm = prm or params()
    rc = lib().wtv_pr_solve(n, d_nbr, d_res, d_term, d_or_h_partition, int(partition_on_device),
                            C.byref(prm), C.byref(st), stream)
    check(rc)
    return st


NOALLOC_MAX_NODES = 1 << 20  # WTV_NOALLOC_MAX_NODES


def pr_solve_noalloc(n, d_nbr, d_res, d_term, d_or_h_partition, partition_on_device, prm=None, stream=0):
    """Call wtv_pr_solve_noalloc (no allocation; n <= NOALLOC_MAX_NODES); returns WtvStats."""
    st = WtvStats()
    prm = prm or params()
    rc = lib().wtv_pr_solve_noalloc(n, d_nbr, d_res, d_term, d_or_h_partition, int(partition_on_device),
                                    C.byref(prm), C.byref(st), stream)
    check(rc)
    return st


def pr_solve_flow(n, d_nbr, d_res, d_term, d_or_h_partition, partition_on_device, prm=None, stream=0):
    """Call wtv_pr_solve_flow; returns (WtvStats, excess returned to the source)."""
    st = WtvStats()
    ret = C.c_int64(0)
    prm = prm or params()
    rc = lib().wtv_pr_solve_flow(n, d_nbr, d_res, d_term, d_or_h_partition, int(partition_on_device),
                                 C.byref(prm), C.byref(st), C.byref(ret), stream)
    check(rc)
    return st, int(ret.value)


def batch_solve(offsets, d_nbr, d_res, d_term, d_or_h_partition, partition_on_device, prm=None, stream=0):
    """Call wtv_batch_solve; offsets is a host uint64 numpy array of G+1 entries.
    Returns (flows int64[G] numpy, WtvStats)."""
    import numpy as np
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    G = off.shape[0] - 1
    flows = np.zeros(max(G, 1), np.int64)
    st = WtvStats()
    prm = prm or params()
    rc = lib().wtv_batch_solve(G, off.ctypes.data, d_nbr, d_res, d_term, d_or_h_partition,
                               int(partition_on_device), flows.ctypes.data, C.byref(prm), C.byref(st), stream)
    check(rc)
    return flows[:G], st


class Server:
    """Thin wrapper of the native serving loop (wtv_server_*)."""

    def __init__(self, slots=4, max_batch=64, max_batch_nodes=0, hbm_budget_bytes=0, **solver_kw):
        cfg = WtvServerConfig()
        cfg.slots, cfg.max_batch = slots, max_batch
        cfg.max_batch_nodes, cfg.hbm_budget_bytes = max_batch_nodes, hbm_budget_bytes
        cfg.params = params(**solver_kw)
        self._h = lib().wtv_server_create(C.byref(cfg))
        if not self._h:
            raise WtvError(WTV_EINVAL, lib().wtv_last_error().decode())

    def submit(self, n, d_nbr, d_res, d_term, d_partition, stream=None):
        """stream: the hipStream_t (int) the buffers were written on; the solve is
        ordered after the work queued there now.  None: the buffers are complete."""
        r = WtvRequest(n, d_nbr, d_res, d_term, d_partition, stream)
        t = C.c_uint64(0)
        check(lib().wtv_server_submit(self._h, C.byref(r), C.byref(t)))
        return t.value

    def wait(self, ticket):
        out = WtvResult()
        check(lib().wtv_server_wait(self._h, ticket, C.byref(out)))
        return out

    def close(self):
        if self._h:
            check(lib().wtv_server_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
