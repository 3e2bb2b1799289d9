"""ctypes bindings for the CPU oracle (oracle/wtv_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP solver: only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module.  The product path (wtivo_gpupr / libwtv_gpupr.so) never calls it.

Each wrapper names the reference code its C function restates; see the header
of wtv_oracle.c for the full list and for what pins each function.
tet_neighbors is numpy only (a sort of the face triples); it is pinned by the
closed-form Kuhn neighbours, by Qhull's own neighbours of Delaunay meshes, by
equivariance under tet and vertex permutations, and by a brute-force search
on tiny meshes (tests/test_oracle.py).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libwtv_oracle.so")
_lib = None

OK, EINVAL, EOVERFLOW, EINVARIANT, ELIMIT, ENOMEM = 0, 1, 2, 3, 5, 6


def build(force: bool = False) -> str:
    """Compile wtv_oracle.c with oracle/Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH) or (
        os.path.getmtime(_LIB_PATH) < os.path.getmtime(os.path.join(_HERE, "wtv_oracle.c"))
    ):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


class ReduceInfo(C.Structure):
    _fields_ = [
        ("free_n", C.c_uint64),
        ("fixed_source", C.c_uint64),
        ("fixed_sink", C.c_uint64),
        ("full_refs", C.c_uint64),
        ("free_refs", C.c_uint64),
        ("terminal_edges", C.c_uint64),
        ("base_flow", C.c_int64),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("solve_seconds", C.c_double),
        ("final_relabel_seconds", C.c_double),
        ("rounds", C.c_uint64),
        ("global_relabels", C.c_uint64),
        ("raw_flow", C.c_int64),
        ("visits", C.c_uint64),
        ("pushes", C.c_uint64),
        ("relabels", C.c_uint64),
        ("bfs_pops", C.c_uint64),
    ]


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.wtvo_capacity.restype = C.c_int32
        L.wtvo_capacity.argtypes = [C.POINTER(C.c_double), C.c_size_t, C.POINTER(C.c_int32), C.c_int,
                                    C.c_uint8, C.c_uint8, C.c_double, C.POINTER(C.c_int)]
        L.wtvo_reduce.restype = C.c_int
        L.wtvo_reduce.argtypes = [C.c_size_t, C.POINTER(C.c_double), C.c_size_t, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32), C.POINTER(C.c_uint8), C.c_double,
                                  C.POINTER(C.c_int32), C.POINTER(C.c_uint32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(ReduceInfo)]
        L.wtvo_reduce_mt.restype = C.c_int
        L.wtvo_reduce_mt.argtypes = list(L.wtvo_reduce.argtypes) + [C.c_int]
        L.wtvo_scatter.restype = None
        L.wtvo_scatter.argtypes = [C.c_size_t, C.POINTER(C.c_uint8), C.c_uint64, C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)]
        L.wtvo_partition.restype = C.c_int
        L.wtvo_partition.argtypes = [C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_int32), C.POINTER(C.c_uint8)]
        L.wtvo_pr_solve.restype = C.c_int
        L.wtvo_pr_solve.argtypes = [C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_int32), C.POINTER(C.c_uint8), C.c_int, C.c_uint64,
                                    C.c_int, C.POINTER(Stats)]
        L.wtvo_bk_solve.restype = C.c_int
        L.wtvo_bk_solve.argtypes = [C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_int32), C.POINTER(C.c_uint8), C.POINTER(C.c_int64)]
        L.wtvo_certificate.restype = C.c_int
        L.wtvo_certificate.argtypes = [C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_int32), C.POINTER(C.c_uint32), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_uint8), C.c_int64]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__(f"oracle {where} failed with code {code}")
        self.code = code


def capacity(V, tet4, fi, label_i, label_j, fill_T):
    """capacity_exact_formula (wtivo_gpupr_bindings.cpp:65-98). Returns (cap, status)."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    t = np.ascontiguousarray(tet4, dtype=np.int32)
    st = C.c_int(0)
    cap = lib().wtvo_capacity(_p(V, C.c_double), V.shape[0], _p(t, C.c_int32), int(fi),
                              int(label_i), int(label_j), float(fill_T), C.byref(st))
    return int(cap), int(st.value)


@dataclass
class Reduced:
    """The reduced graph of graph_cut_fast (wtivo_gpupr_bindings.cpp:139-318)."""
    n: int
    full_to_free: np.ndarray  # int32[N]: reduced id, -2 fixed source, -3 fixed sink
    free_old: np.ndarray      # uint32[n]
    nbr: np.ndarray           # int32[n,4]
    res: np.ndarray           # uint32[n,4]
    term: np.ndarray          # int32[n]
    info: dict


def reduce(V, tets, nbrs, labels, fill_T, threads: int = 1) -> Reduced:
    """Reduced-graph build (wtivo_gpupr_bindings.cpp:139-318) on `threads`
    threads; the arrays do not depend on the thread count."""
    V = np.ascontiguousarray(V, dtype=np.float64)
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    nbrs = np.ascontiguousarray(nbrs, dtype=np.int32)
    labels = np.ascontiguousarray(labels, dtype=np.uint8).reshape(-1)
    N = tets.shape[0]
    f2f = np.empty(N, np.int32)
    fo = np.empty(max(N, 1), np.uint32)
    nb = np.empty((max(N, 1), 4), np.int32)
    rs = np.empty((max(N, 1), 4), np.uint32)
    tm = np.empty(max(N, 1), np.int32)
    info = ReduceInfo()
    rc = lib().wtvo_reduce_mt(N, _p(V, C.c_double), V.shape[0], _p(tets, C.c_int32), _p(nbrs, C.c_int32),
                              _p(labels, C.c_uint8), float(fill_T), _p(f2f, C.c_int32), _p(fo, C.c_uint32),
                              _p(nb, C.c_int32), _p(rs, C.c_uint32), _p(tm, C.c_int32), C.byref(info),
                              int(threads))
    if rc != OK:
        raise OracleError(rc, "reduce")
    n = int(info.free_n)
    d = {k: getattr(info, k) for k, _ in ReduceInfo._fields_}
    return Reduced(n, f2f, fo[:n].copy(), nb[:n].copy(), rs[:n].copy(), tm[:n].copy(), d)


def scatter(labels, free_old, partition):
    """Label reconstruction (wtivo_gpupr_bindings.cpp:357-364)."""
    labels = np.ascontiguousarray(labels, dtype=np.uint8)
    fo = np.ascontiguousarray(free_old, dtype=np.uint32)
    part = np.ascontiguousarray(partition, dtype=np.uint8)
    out = np.empty_like(labels)
    lib().wtvo_scatter(labels.shape[0], _p(labels, C.c_uint8), fo.shape[0], _p(fo, C.c_uint32),
                       _p(part, C.c_uint8), _p(out, C.c_uint8))
    return out


def _graph_args(nbr, res, term):
    nbr = np.ascontiguousarray(nbr, dtype=np.int32).reshape(-1, 4)
    res = np.array(res, dtype=np.uint32, copy=True).reshape(-1, 4)
    term = np.array(term, dtype=np.int32, copy=True).reshape(-1)
    return nbr, res, term


def partition(nbr, res, term):
    """Residual BFS from the sink: 1 = reaches the sink (gpu_push_relabel_fast.cu:315-324)."""
    nbr, res, term = _graph_args(nbr, res, term)
    n = term.shape[0]
    out = np.empty(max(n, 1), np.uint8)
    rc = lib().wtvo_partition(n, _p(nbr, C.c_int32), _p(res, C.c_uint32), _p(term, C.c_int32),
                              _p(out, C.c_uint8))
    if rc != OK:
        raise OracleError(rc, "partition")
    return out[:n]


@dataclass
class Solution:
    partition: np.ndarray
    raw_flow: int
    res: np.ndarray
    term: np.ndarray
    stats: dict


def pr_solve(nbr, res, term, period=1024, max_rounds=2000000, local_steps=8) -> Solution:
    """Sequential restatement of cellocut_gpu_push_relabel_fast (gpu_push_relabel_fast.cu:406-571)."""
    nbr, res, term = _graph_args(nbr, res, term)
    n = term.shape[0]
    part = np.empty(max(n, 1), np.uint8)
    st = Stats()
    rc = lib().wtvo_pr_solve(n, _p(nbr, C.c_int32), _p(res, C.c_uint32), _p(term, C.c_int32),
                             _p(part, C.c_uint8), int(period), int(max_rounds), int(local_steps),
                             C.byref(st))
    if rc != OK:
        raise OracleError(rc, "pr_solve")
    d = {k: getattr(st, k) for k, _ in Stats._fields_}
    return Solution(part[:n], int(st.raw_flow), res, term, d)


def bk_solve(nbr, res, term) -> Solution:
    """Boykov-Kolmogorov max-flow + residual BFS partition (CPU baseline)."""
    nbr, res, term = _graph_args(nbr, res, term)
    n = term.shape[0]
    part = np.empty(max(n, 1), np.uint8)
    flow = C.c_int64(0)
    rc = lib().wtvo_bk_solve(n, _p(nbr, C.c_int32), _p(res, C.c_uint32), _p(term, C.c_int32),
                             _p(part, C.c_uint8), C.byref(flow))
    if rc != OK:
        raise OracleError(rc, "bk_solve")
    return Solution(part[:n], int(flow.value), res, term, {})


def certificate(nbr, res0, term0, res, term, partition, raw_flow) -> int:
    """0 when (res, term, partition, raw_flow) is a maximum preflow with its minimal sink set."""
    nbr = np.ascontiguousarray(nbr, dtype=np.int32).reshape(-1, 4)
    res0 = np.ascontiguousarray(res0, dtype=np.uint32).reshape(-1, 4)
    res = np.ascontiguousarray(res, dtype=np.uint32).reshape(-1, 4)
    term0 = np.ascontiguousarray(term0, dtype=np.int32).reshape(-1)
    term = np.ascontiguousarray(term, dtype=np.int32).reshape(-1)
    part = np.ascontiguousarray(partition, dtype=np.uint8).reshape(-1)
    n = term0.shape[0]
    return int(lib().wtvo_certificate(n, _p(nbr, C.c_int32), _p(res0, C.c_uint32), _p(term0, C.c_int32),
                                      _p(res, C.c_uint32), _p(term, C.c_int32), _p(part, C.c_uint8),
                                      int(raw_flow)))


def surface(labels, V, tets, nbrs):
    """surface_extraction_topology_cuda (wtivo_gpupr_bindings.cpp:400-577): (OV float64 [u,3], OF int32 [F,3])."""
    L = lib()
    if not hasattr(L, "_surf_ready"):
        L.wtvo_surface.restype = C.c_int
        L.wtvo_surface.argtypes = [C.c_size_t, C.POINTER(C.c_uint8), C.POINTER(C.c_double), C.c_size_t,
                                   C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_size_t),
                                   C.POINTER(C.c_size_t), C.c_void_p, C.c_void_p]
        L._surf_ready = True
    labels = np.ascontiguousarray(labels, dtype=np.uint8).reshape(-1)
    V = np.ascontiguousarray(V, dtype=np.float64)
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    nbrs = np.ascontiguousarray(nbrs, dtype=np.int32)
    nf, nu = C.c_size_t(0), C.c_size_t(0)
    args = (labels.shape[0], _p(labels, C.c_uint8), _p(V, C.c_double), V.shape[0], _p(tets, C.c_int32),
            _p(nbrs, C.c_int32), C.byref(nf), C.byref(nu))
    rc = L.wtvo_surface(*args, None, None)
    if rc != OK:
        raise OracleError(rc, "surface")
    OV = np.empty((nu.value, 3), np.float64)
    OF = np.empty((nf.value, 3), np.int32)
    rc = L.wtvo_surface(*args, OV.ctypes.data, OF.ctypes.data)
    if rc != OK:
        raise OracleError(rc, "surface")
    return OV, OF


_FACE = np.array([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]])


def tet_neighbors(tets, nv=None):
    """4-slot neighbours by face matching: the export of tetrahedralize_neighbors
    (native/core/wtivo_core.cpp:136-154). Slot fi is the tet across face
    {012,013,023,123}[fi], which is CGAL's neighbor(3 - fi); a hull face gets -1.

    Sorts the 4N sorted vertex triples and pairs equal neighbours in that order.
    Raises OracleError(EINVAL) for a vertex id outside [0, nv), a vertex repeated
    within a tet, or a face held by more than two tets."""
    tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
    N = tets.shape[0]
    nbr = np.full((N, 4), -1, np.int32)
    if N == 0:
        return nbr
    t64 = tets.astype(np.int64)
    if nv is not None and ((t64 < 0) | (t64 >= nv)).any():
        raise OracleError(EINVAL, "tet_neighbors: vertex id out of range")
    st = np.sort(t64, axis=1)
    if (st[:, 1:] == st[:, :-1]).any():
        raise OracleError(EINVAL, "tet_neighbors: repeated vertex")
    F = np.sort(t64[:, _FACE], axis=2).reshape(-1, 3)   # row 4i+fi = face fi of tet i
    order = np.lexsort((F[:, 2], F[:, 1], F[:, 0]))
    Fs = F[order]
    eq = (Fs[1:] == Fs[:-1]).all(axis=1)
    if (eq[1:] & eq[:-1]).any():
        raise OracleError(EINVAL, "tet_neighbors: face held by more than two tets")
    k = np.nonzero(eq)[0]
    p, q = order[k], order[k + 1]
    flat = nbr.reshape(-1)
    flat[p] = (q // 4).astype(np.int32)
    flat[q] = (p // 4).astype(np.int32)
    return nbr


def topology_to_cuda_reference(host_array, torch, chunk_rows: int = 262144):
    """Restatement of the reference's upload loop, topology_to_cuda_chunked
    (wtivo.py:284-309): repack 262,144-row slices with np.ascontiguousarray and
    copy each one synchronously into a torch CUDA tensor.  Timed beside
    wtv_topology_upload as the reference's way of doing the same copy."""
    a = np.asarray(host_array)
    n = int(a.shape[0])
    out = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    for i in range(0, n, chunk_rows):
        j = min(n, i + chunk_rows)
        h = np.ascontiguousarray(a[i:j], dtype=np.int32)
        out[i:j].copy_(torch.from_numpy(h), non_blocking=False)
        del h
    torch.cuda.synchronize()
    return out


def sample_labels(origins, values, background, voxel_size, V, tets, bmin, bmax, padding, threshold):
    """SparseUDF.sample_tet_labels (native/vdb/wtivo_vdb.cpp:1519-1584) with
    OpenVDB's BoxSampler, restated in wtv_oracle.c.  Returns (labels uint8[N],
    sampled values float32[N])."""
    L = lib()
    if not hasattr(L, "_udf_ready"):
        L.wtvo_sample_labels.restype = C.c_int
        L.wtvo_sample_labels.argtypes = [C.c_size_t, C.c_void_p, C.c_void_p, C.c_float, C.c_double, C.c_size_t,
                                         C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                         C.c_double, C.c_void_p, C.c_void_p]
        L._udf_ready = True
    origins = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
    values = np.ascontiguousarray(values, dtype=np.float32).reshape(-1, 512)
    V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
    tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
    bmin = np.ascontiguousarray(bmin, dtype=np.float64).reshape(3)
    bmax = np.ascontiguousarray(bmax, dtype=np.float64).reshape(3)
    N = tets.shape[0]
    labels = np.empty(max(N, 1), np.uint8)
    vals = np.empty(max(N, 1), np.float32)
    rc = L.wtvo_sample_labels(origins.shape[0], origins.ctypes.data, values.ctypes.data, float(background),
                              float(voxel_size), V.shape[0], V.ctypes.data, N, tets.ctypes.data, bmin.ctypes.data,
                              bmax.ctypes.data, float(padding), float(threshold), labels.ctypes.data,
                              vals.ctypes.data)
    if rc != OK:
        raise OracleError(rc, "sample_labels")
    return labels[:N], vals[:N]


def flow_certificate(nbr, res0, term0, res, term, partition, raw_flow) -> int:
    """0 when (res, term) is a maximum flow -- every vertex conserves -- with its minimal sink set."""
    L = lib()
    if not hasattr(L, "_flowcert_ready"):
        L.wtvo_flow_certificate.restype = C.c_int
        L.wtvo_flow_certificate.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_int64]
        L._flowcert_ready = True
    nbr = np.ascontiguousarray(nbr, dtype=np.int32).reshape(-1, 4)
    res0 = np.ascontiguousarray(res0, dtype=np.uint32).reshape(-1, 4)
    res = np.ascontiguousarray(res, dtype=np.uint32).reshape(-1, 4)
    term0 = np.ascontiguousarray(term0, dtype=np.int32).reshape(-1)
    term = np.ascontiguousarray(term, dtype=np.int32).reshape(-1)
    part = np.ascontiguousarray(partition, dtype=np.uint8).reshape(-1)
    return int(L.wtvo_flow_certificate(term0.shape[0], nbr.ctypes.data, res0.ctypes.data, term0.ctypes.data,
                                       res.ctypes.data, term.ctypes.data, part.ctypes.data, int(raw_flow)))
