"""GPU parity of the sparse-grid tet labels (wtv_udf_*, SparseUDF.sample_tet_labels)
against the oracle's restatement: labels must match bit for bit, on random
sparse grids with holes and on a genus-3 distance grid over a Kuhn mesh, with
and without padding, labels on the host and on the device, and the error paths."""
import ctypes as C

import numpy as np
import pytest

import wtv_capi as W
import wtv_synth as S
import wtv_udf as U

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    W.lib()
    return torch


def _random_grid(rng, cells=6, keep=0.6, scale=3.0, lo_cell=-1):
    c = np.arange(lo_cell, lo_cell + cells)
    cc = np.stack(np.meshgrid(c, c, c, indexing="ij"), -1).reshape(-1, 3)
    cc = cc[rng.random(len(cc)) < keep]
    return (cc * 8).astype(np.int32), (rng.random((len(cc), 512)) * scale).astype(np.float32)


def _mesh_in(rng, bmin, bmax, n):
    """Random tets (4 random vertices each) spread over and a little beyond the box."""
    span = bmax - bmin
    V = bmin - 0.1 * span + rng.random((4 * n, 3)) * 1.2 * span
    return V, rng.permutation(4 * n).astype(np.int32).reshape(-1, 4)


@pytest.mark.parametrize("seed,vs,pad", [(0, 1 / 32, 0.0), (1, 1 / 40.0, 0.05), (2, 1 / 37.0, 0.0)])
def test_random_sparse_grid_labels_equal_oracle(orc, torch_cuda, seed, vs, pad):
    rng = np.random.default_rng(seed)
    o, v = _random_grid(rng)
    bmin, bmax = np.array([-0.5, 0.0, 0.25]), np.array([1.0, 1.3, 1.5])
    V, t = _mesh_in(rng, bmin, bmax, 200_000)
    _, vals = orc.sample_labels(o, v, 2.5, vs, V, t, bmin, bmax, pad, 10.0)
    thr = float(np.median(vals))
    want, _ = orc.sample_labels(o, v, 2.5, vs, V, t, bmin, bmax, pad, thr)
    with U.SparseUDFGrid(o, v, 2.5, vs) as g:
        got, sec = g.sample_tet_labels(V, t, bmin, bmax, pad, thr)
    assert got.dtype == np.uint8 and sec > 0
    assert np.array_equal(got, want)
    assert 0.3 < want.mean() < 0.7


def test_genus3_udf_on_kuhn_mesh(orc, torch_cuda):
    torch = torch_cuda
    k = 40
    m = S.kuhn_mesh(k, jitter=0.2, seed=3)
    shape = S.genus3(3)
    o, v, bg = S.udf_leaves(shape, 1 / 256, 12.0, device="cuda")   # band 0.047 > threshold 0.025
    thr = 1.0 / k
    want, _ = orc.sample_labels(o, v, bg, 1 / 256, m.V, m.tets, np.zeros(3), np.ones(3), 0.0, thr)
    dt = torch.from_numpy(m.tets).cuda()
    with U.SparseUDFGrid(o, v, bg, 1 / 256) as g:
        got, _ = g.sample_tet_labels(m.V, dt, np.zeros(3), np.ones(3), 0.0, thr)
        got_np, _ = g.sample_tet_labels(m.V, m.tets, np.zeros(3), np.ones(3), 0.0, thr)
    assert np.array_equal(got, want) and np.array_equal(got_np, want)
    assert 0.01 < (want == 0).mean() < 0.5


def test_labels_on_device_and_graph_cut(orc, torch_cuda):
    """Device labels feed graph_cut_fast; the cut equals the one on the oracle's labels."""
    torch = torch_cuda
    import wtivo_gpupr as G
    k = 24
    m = S.kuhn_mesh(k, jitter=0.2, seed=4)
    o, v, bg = S.udf_leaves(S.genus3(4), 1 / 128, 8.0, device="cuda")   # band 0.0625 > threshold 0.042
    thr = 1.0 / k
    want, _ = orc.sample_labels(o, v, bg, 1 / 128, m.V, m.tets, np.zeros(3), np.ones(3), 0.0, thr)
    dt = torch.from_numpy(m.tets).cuda()
    dn = torch.from_numpy(m.nbrs).cuda()
    lab_d = torch.empty(m.tets.shape[0], dtype=torch.uint8, device="cuda")
    h = C.c_void_p(0)
    L = W.lib()
    oo, vv = np.ascontiguousarray(o), np.ascontiguousarray(v)
    W.check(L.wtv_udf_create(oo.ctypes.data, vv.ctypes.data, oo.shape[0], float(bg), 1 / 128, C.byref(h), None))
    dV = torch.from_numpy(np.ascontiguousarray(m.V)).cuda()         # vertices resident in HBM
    try:
        bmin, bmax = np.zeros(3), np.ones(3)
        W.check(L.wtv_udf_sample_tet_labels(h, dV.data_ptr(), m.V.shape[0], 1, dt.data_ptr(),
                                            m.tets.shape[0], bmin.ctypes.data, bmax.ctypes.data, 0.0, thr,
                                            lab_d.data_ptr(), 1, None, torch.cuda.current_stream().cuda_stream))
    finally:
        W.check(L.wtv_udf_destroy(h))
    got = lab_d.cpu().numpy()
    assert np.array_equal(got, want)
    a = G.graph_cut_fast(m.V, dt, dn, got, 10.0)
    b = G.graph_cut_fast(m.V, dt, dn, want, 10.0)
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]
    # the resident pipeline: device labels straight into the resident cut
    out_d = torch.empty_like(lab_d)
    res = W.WtvCutResult()
    W.check(L.wtv_graph_cut_fast_resident(dV.data_ptr(), m.V.shape[0], dt.data_ptr(), dn.data_ptr(),
                                          m.tets.shape[0], lab_d.data_ptr(), 10.0, 1024, 2000000, 8,
                                          out_d.data_ptr(), C.byref(res), torch.cuda.current_stream().cuda_stream))
    assert np.array_equal(out_d.cpu().numpy(), a[0]) and res.adjusted_flow == a[1]


def test_udf_error_paths(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(9)
    o, v = _random_grid(rng, cells=2, keep=1.0)
    L = W.lib()
    h = C.c_void_p(0)
    assert L.wtv_udf_create(o.ctypes.data, v.ctypes.data, 0, 1.0, 0.1, C.byref(h), None) == W.WTV_EINVAL
    assert "grid is empty" in L.wtv_last_error().decode()
    oo = np.ascontiguousarray(np.concatenate([o, o]))
    vv = np.ascontiguousarray(np.concatenate([v, v]))
    assert L.wtv_udf_create(oo.ctypes.data, vv.ctypes.data, oo.shape[0], 1.0, 0.1, C.byref(h), None) == W.WTV_EINVAL
    assert "share an origin" in L.wtv_last_error().decode()
    bad = np.ascontiguousarray(o + 4)
    assert L.wtv_udf_create(bad.ctypes.data, v.ctypes.data, o.shape[0], 1.0, 0.1, C.byref(h), None) == W.WTV_EINVAL
    assert L.wtv_udf_create(o.ctypes.data, v.ctypes.data, o.shape[0], 1.0, 0.0, C.byref(h), None) == W.WTV_EINVAL
    g = U.SparseUDFGrid(o, v, 1.0, 0.1)
    V = np.zeros((4, 3))
    t = np.array([[0, 1, 2, 3]], np.int32)
    with pytest.raises(W.WtvError, match="degenerate bbox"):
        g.sample_tet_labels(V, t, np.zeros(3), np.array([1.0, 1.0, 0.0]), 0.0, 0.5)
    with pytest.raises(W.WtvError, match="tet index out of range"):
        g.sample_tet_labels(V, t + 1, np.zeros(3), np.ones(3), 0.0, 0.5)
    with pytest.raises(RuntimeError, match="tet_vertices must be float64 Nx3"):
        g.sample_tet_labels(np.zeros((4, 2)), t, np.zeros(3), np.ones(3), 0.0, 0.5)
    with pytest.raises(RuntimeError, match="bbox must have length 3"):
        g.sample_tet_labels(V, t, np.zeros(2), np.ones(3), 0.0, 0.5)
    lab, _ = g.sample_tet_labels(V, np.zeros((0, 4), np.int32), np.zeros(3), np.ones(3), 0.0, 0.5)
    assert lab.shape == (0,)
    g.close()
