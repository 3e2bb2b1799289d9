"""GPU parity of the topology entry points, called through the C-ABI:

* wtv_tet_neighbors against the closed-form Kuhn neighbours, Qhull's Delaunay
  neighbours and the oracle's face sort (also on permuted meshes, meshes that
  are all hull, and duplicated tets), plus its error paths;
* wtv_topology_upload against np.ascontiguousarray of the host matrix for C,
  Fortran and strided layouts, chunk-edge row counts and worker counts.

Integer work: every array must match exactly."""
import ctypes as C

import numpy as np
import pytest

import wtv_capi as W
import wtv_synth as S
import wtv_topology as T

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    W.lib()
    return torch


def _nbrs(torch, tets_np, nv=None):
    d = torch.from_numpy(np.ascontiguousarray(tets_np, np.int32)).cuda()
    return T.tet_neighbors_cuda(d, nv).cpu().numpy()


@pytest.mark.parametrize("k", [1, 2, 7, 40])
def test_neighbors_kuhn_closed_form(torch_cuda, k):
    tets, nbrs = S.kuhn_topology_t(k, "cuda")
    out = T.tet_neighbors_cuda(tets, (k + 1) ** 3)
    assert torch_cuda.equal(out, nbrs)


def test_neighbors_kuhn_6m_tets_closed_form(torch_cuda):
    k = 100
    tets, nbrs = S.kuhn_topology_t(k, "cuda")
    print(f"kuhn {k}: {tets.shape[0]} tets", flush=True)
    out = T.tet_neighbors_cuda(tets, (k + 1) ** 3)
    assert torch_cuda.equal(out, nbrs)


def test_neighbors_kuhn_vertex_ids_shuffled(torch_cuda):
    """Random vertex numbering takes the hashed-start path; the result is the same."""
    torch = torch_cuda
    k = 60
    tets, nbrs = S.kuhn_topology_t(k, "cuda")
    g = torch.Generator(device="cpu")
    g.manual_seed(5)
    perm = torch.randperm((k + 1) ** 3, generator=g).to(torch.int32).cuda()
    out = T.tet_neighbors_cuda(perm[tets.long()].contiguous(), (k + 1) ** 3)
    assert torch.equal(out, nbrs)


@pytest.mark.parametrize("morton", [False, True])
def test_neighbors_delaunay_match_qhull(orc, torch_cuda, morton):
    pts = S.genus3_point_cloud(200_000, 11)[0]
    m = S.delaunay_mesh(pts, morton=morton)
    got = _nbrs(torch_cuda, m.tets, pts.shape[0])
    assert np.array_equal(got, m.nbrs)
    assert np.array_equal(got, orc.tet_neighbors(m.tets, pts.shape[0]))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_neighbors_permuted_mesh(orc, torch_cuda, seed):
    tets, nbrs = S.kuhn_topology(12)
    t2, n2 = S.permute_topology(tets, nbrs, seed)
    got = _nbrs(torch_cuda, t2)
    assert np.array_equal(got, n2)
    assert np.array_equal(got, orc.tet_neighbors(t2))


def test_neighbors_all_hull_and_duplicated_tets(orc, torch_cuda):
    rng = np.random.default_rng(3)
    loose = rng.permutation(4 * 5000).astype(np.int32).reshape(-1, 4)   # no shared vertices
    assert (_nbrs(torch_cuda, loose) == -1).all()
    # two copies of each loose tet: every face is held by the tet and its twin
    twins = np.repeat(loose[:1000], 2, axis=0)
    got = _nbrs(torch_cuda, twins)
    assert np.array_equal(got, orc.tet_neighbors(twins))
    assert (got == (np.arange(twins.shape[0]) ^ 1)[:, None]).all()


def test_neighbors_single_tet_and_empty(torch_cuda):
    assert (_nbrs(torch_cuda, np.array([[3, 0, 2, 1]], np.int32)) == -1).all()
    empty = torch_cuda.empty((0, 4), dtype=torch_cuda.int32, device="cuda")
    assert T.tet_neighbors_cuda(empty).shape == (0, 4)


@pytest.mark.parametrize("bad,what", [
    (np.array([[0, 1, 2, 3], [0, 1, 2, 4], [0, 1, 2, 5]], np.int32), "more than two tets"),
    (np.array([[0, 1, 2, 3], [0, 1, 1, 4]], np.int32), "repeated vertex"),
    (np.array([[0, 1, 2, 3], [0, 1, -1, 4]], np.int32), "outside [0, nv)"),
    (np.array([[0, 1, 2, 3], [0, 1, 2, 6]], np.int32), "outside [0, nv)"),
])
def test_neighbors_reject_bad_meshes(torch_cuda, bad, what):
    d = torch_cuda.from_numpy(bad).cuda()
    out = torch_cuda.empty_like(d)
    rc = W.lib().wtv_tet_neighbors(d.data_ptr(), d.shape[0], 6, out.data_ptr(),
                                   torch_cuda.cuda.current_stream().cuda_stream)
    assert rc == W.WTV_EINVAL
    assert what in W.lib().wtv_last_error().decode()


# ------------------------------------------------------------------ upload
def _upload(torch, a, threads=0):
    a = np.asarray(a)
    out = torch.full((a.shape[0], 4), -7, dtype=torch.int32, device="cuda")
    rs, cs = (s // 4 for s in a.strides)
    rc = W.lib().wtv_topology_upload(a.ctypes.data, a.shape[0], rs, cs, out.data_ptr(), threads,
                                     torch.cuda.current_stream().cuda_stream)
    W.check(rc)
    return out.cpu().numpy()


@pytest.mark.parametrize("rows", [1, 1000, 262144, 262145, 3_000_001])
@pytest.mark.parametrize("order", ["C", "F"])
def test_upload_layouts_and_chunk_edges(torch_cuda, rows, order):
    rng = np.random.default_rng(rows)
    a = rng.integers(-2**31, 2**31 - 1, size=(rows, 4), dtype=np.int64).astype(np.int32)
    a = np.asfortranarray(a) if order == "F" else a
    assert np.array_equal(_upload(torch_cuda, a), np.ascontiguousarray(a))


@pytest.mark.parametrize("threads", [1, 3, 16])
def test_upload_worker_counts(torch_cuda, threads):
    a = np.asfortranarray(np.arange(4 * 2_100_000, dtype=np.int32).reshape(-1, 4))
    assert np.array_equal(_upload(torch_cuda, a, threads), np.ascontiguousarray(a))


def test_upload_strided_views(torch_cuda):
    big = np.arange(4 * 900_001, dtype=np.int32).reshape(-1, 4)
    for view in (big[::3], big[:, ::-1], big[::-2], np.ascontiguousarray(big.T).T):
        assert np.array_equal(_upload(torch_cuda, view), np.ascontiguousarray(view))


def test_upload_waits_for_the_callers_stream(torch_cuda):
    torch = torch_cuda
    a = np.arange(4 * 1_000_000, dtype=np.int32).reshape(-1, 4)
    out = torch.empty((a.shape[0], 4), dtype=torch.int32, device="cuda")
    for _ in range(20):
        out.fill_(-1)           # queued on the current stream before the upload
    W.check(W.lib().wtv_topology_upload(a.ctypes.data, a.shape[0], 4, 1, out.data_ptr(), 0,
                                        torch.cuda.current_stream().cuda_stream))
    assert np.array_equal(out.cpu().numpy(), a)


def test_topology_to_cuda_chunked_mirror(torch_cuda):
    tets, _ = S.kuhn_topology(20)
    for a in (np.asfortranarray(tets), tets.astype(np.int64), tets):
        t = T.topology_to_cuda_chunked(a, "tets")
        assert t.is_cuda and t.dtype == torch_cuda.int32 and t.is_contiguous() and tuple(t.shape) == tets.shape
        assert np.array_equal(t.cpu().numpy(), tets)
    assert T.topology_to_cuda_chunked(np.zeros((0, 4), np.int32), "empty").shape == (0, 4)


def test_uploaded_topology_drives_graph_cut(orc, torch_cuda):
    """Host topology -> upload -> device neighbours -> graph_cut_fast equals the
    same cut on the closed-form neighbours."""
    import wtivo_gpupr as G
    m = S.kuhn_mesh(14, jitter=0.2, seed=5)
    shape = S.sphere_with_holes(5)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), flip=0.02, seed=5)
    dt = T.topology_to_cuda_chunked(np.asfortranarray(m.tets), "tets")
    dn = T.tet_neighbors_cuda(dt, m.V.shape[0])
    ref_n = torch_cuda.from_numpy(m.nbrs).cuda()
    assert torch_cuda.equal(dn, ref_n)
    a = G.graph_cut_fast(m.V, dt, dn, lab, 10.0)
    b = G.graph_cut_fast(m.V, dt, ref_n, lab, 10.0)
    assert np.array_equal(a[0], b[0]) and a[1] == b[1] and a[2] == b[2]
