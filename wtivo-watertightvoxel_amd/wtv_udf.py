"""Tet labels from a sparse distance grid on the device, over the C-ABI.

SparseUDFGrid mirrors the part of the reference's SparseUDF that the graph cut
consumes: sample_tet_labels(tet_vertices, tets, bbox_min, bbox_max, padding,
threshold, threads) with the same argument checks, messages and result,
(labels uint8[N], seconds) (native/vdb/wtivo_vdb.cpp:1519-1584).  Building
the grid from a mesh (OpenVDB meshToUnsignedDistanceField) stays with the
caller: the grid comes in as OpenVDB-style 8^3 leaves (index-space origins and
512 values each, offset (x&7)<<6 | (y&7)<<3 | (z&7)), a background value and
the voxel size of the linear transform.

Fails loudly when the HIP library is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import wtv_capi as W


class SparseUDFGrid:
    def __init__(self, leaf_origins, leaf_values, background: float, voxel_size: float):
        o = np.ascontiguousarray(leaf_origins, dtype=np.int32).reshape(-1, 3)
        v = np.ascontiguousarray(leaf_values, dtype=np.float32).reshape(-1, 512)
        if o.shape[0] != v.shape[0]:
            raise RuntimeError("SparseUDF: leaf_origins and leaf_values hold different leaf counts")
        h = C.c_void_p(0)
        W.check(W.lib().wtv_udf_create(o.ctypes.data, v.ctypes.data, o.shape[0], float(background),
                                       float(voxel_size), C.byref(h), None))
        self._h = h
        self.leaf_count = int(o.shape[0])
        self.voxel_size = float(voxel_size)
        self.background = float(background)

    def sample_tet_labels(self, tet_vertices, tets, bbox_min, bbox_max, padding: float, threshold: float,
                          threads: int = 16):
        """(labels uint8[N], seconds).  `tets` is an int32 [N,4] numpy array
        (uploaded first) or a contiguous int32 CUDA tensor."""
        import torch

        if self._h is None:
            raise RuntimeError("SparseUDF.sample_tet_labels: grid is empty")
        V = np.ascontiguousarray(tet_vertices, dtype=np.float64)
        if V.ndim != 2 or V.shape[1] != 3:
            raise RuntimeError("SparseUDF.sample_tet_labels: tet_vertices must be float64 Nx3")
        if isinstance(tets, torch.Tensor):
            if not tets.is_cuda or tets.dtype != torch.int32 or not tets.is_contiguous():
                raise RuntimeError("SparseUDF.sample_tet_labels: tets must be int32 Mx4")
            d_tets = tets
        else:
            t = np.asarray(tets)
            if t.ndim != 2 or t.shape[1] != 4:
                raise RuntimeError("SparseUDF.sample_tet_labels: tets must be int32 Mx4")
            import wtv_topology as T
            d_tets = T.topology_to_cuda_chunked(t, "tets") if t.shape[0] else None
        if d_tets is not None and (d_tets.dim() != 2 or d_tets.size(1) != 4):
            raise RuntimeError("SparseUDF.sample_tet_labels: tets must be int32 Mx4")
        bmin = np.ascontiguousarray(bbox_min, dtype=np.float64)
        bmax = np.ascontiguousarray(bbox_max, dtype=np.float64)
        if bmin.ndim != 1 or bmin.shape[0] != 3 or bmax.ndim != 1 or bmax.shape[0] != 3:
            raise RuntimeError("SparseUDF.sample_tet_labels: bbox must have length 3")
        N = int(d_tets.size(0)) if d_tets is not None else 0
        labels = np.empty(N, np.uint8)
        sec = C.c_double(0.0)
        W.check(W.lib().wtv_udf_sample_tet_labels(self._h, V.ctypes.data, V.shape[0], 0,
                                                  d_tets.data_ptr() if N else None, N, bmin.ctypes.data,
                                                  bmax.ctypes.data, float(padding), float(threshold),
                                                  labels.ctypes.data if N else None, 0, C.byref(sec),
                                                  torch.cuda.current_stream().cuda_stream))
        return labels, float(sec.value)

    def close(self):
        if self._h is not None:
            W.check(W.lib().wtv_udf_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
