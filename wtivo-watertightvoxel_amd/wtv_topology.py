"""Mesh topology on its way into HBM, over the C-ABI of libwtv_gpupr.so.

topology_to_cuda_chunked keeps the name, arguments, checks, log line and
result of the reference's caller-side helper (wtivo.py:284-309): one
C-contiguous int32 [N,4] CUDA tensor from an Nx4 host matrix in any memory
order.  The copy goes through wtv_topology_upload (pinned staging, one stream
per worker thread) instead of synchronous 262,144-row slices.

tet_neighbors_cuda builds, on the device, the 4-slot neighbours that the
reference exports from CGAL (tetrahedralize_neighbors,
native/core/wtivo_core.cpp:136-154) for meshes that come with tets only.

Both fail loudly when the HIP library is missing; there is no CPU fallback.
"""
from __future__ import annotations

import os
import time

import numpy as np

import wtv_capi as W


def _quiet() -> bool:
    e = os.environ.get("WTV_QUIET", "")
    return bool(e) and e != "0"


def _stream(torch):
    return torch.cuda.current_stream().cuda_stream


def topology_to_cuda_chunked(host_array, tag: str, threads: int = 0):
    """Upload an Nx4 int32 topology matrix to one contiguous CUDA tensor.

    Like the reference, other integer dtypes are converted to int32 (here in
    one host pass before the upload).  Fortran-ordered input (Eigen's
    default) is interleaved by the upload's worker threads, without a full
    host copy."""
    import torch

    a = np.asarray(host_array)
    if a.ndim != 2 or a.shape[1] != 4:
        raise RuntimeError(f"{tag}: expected Nx4 topology array, got {a.shape}")
    if a.dtype != np.int32:
        a = a.astype(np.int32)
    if any(s % 4 for s in a.strides):
        a = np.ascontiguousarray(a)
    n = int(a.shape[0])
    out = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    t0 = time.perf_counter()
    if n:
        rs, cs = (int(s) // 4 for s in a.strides)
        W.check(W.lib().wtv_topology_upload(a.ctypes.data, n, rs, cs, out.data_ptr(), int(threads),
                                            _stream(torch)))
    if not _quiet():
        gib = out.numel() * out.element_size() / float(1 << 30)
        print(f"[WTiVo-TOPOGPU] {tag} -> VRAM: rows={n:,} | device={gib:.3f} GiB | "
              f"host staging pinned 2x4 MiB per worker | {time.perf_counter() - t0:.3f}s", flush=True)
    return out


def tet_neighbors_cuda(tets_cuda, num_vertices: int | None = None):
    """4-slot neighbours (int32 [N,4] CUDA tensor) of the tets in `tets_cuda`:
    slot fi is the tet across face {012,013,023,123}[fi], -1 on the hull.
    num_vertices defaults to max(tets) + 1."""
    import torch

    if not tets_cuda.is_cuda or tets_cuda.dtype != torch.int32 or not tets_cuda.is_contiguous():
        raise RuntimeError("tets must be a contiguous int32 CUDA tensor")
    if tets_cuda.dim() != 2 or tets_cuda.size(1) != 4:
        raise RuntimeError("tets must be Nx4")
    N = int(tets_cuda.size(0))
    out = torch.empty_like(tets_cuda)
    if N == 0:
        return out
    if num_vertices is None:
        num_vertices = int(tets_cuda.max().item()) + 1
    if num_vertices < 1:
        raise RuntimeError("num_vertices must be positive")
    W.check(W.lib().wtv_tet_neighbors(tets_cuda.data_ptr(), N, int(num_vertices), out.data_ptr(),
                                      _stream(torch)))
    return out
