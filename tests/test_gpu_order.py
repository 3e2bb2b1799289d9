"""GPU: the locality renumbering of single-graph solves (csrc/wtv_order.hip).
The flow and the partition (the minimal sink set) do not depend on the node
numbering, so a renumbered solve must give BK's flow and labels, and its
residuals and terminals, scattered back to the caller's ids, must certify the
cut on the caller's graph."""
import numpy as np
import pytest

import wtv_capi as W
import wtv_synth as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    W.lib()
    return torch


def _shuffled(nbr, res, term, seed):
    """The same graph with its nodes in a random order (host side)."""
    n = term.shape[0]
    perm = np.random.default_rng(seed).permutation(n)
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    nb = nbr[perm].astype(np.int64)
    nb = np.where(nb >= 0, inv[np.maximum(nb, 0)], -1).astype(np.int32)
    return nb, res[perm], term[perm]


def _solve(torch, nbr, res, term, host_partition=False, **kw):
    dn = torch.from_numpy(np.ascontiguousarray(nbr, np.int32)).cuda()
    dr = torch.from_numpy(np.ascontiguousarray(res, np.uint32).view(np.int32)).cuda()
    dt = torch.from_numpy(np.ascontiguousarray(term, np.int32)).cuda()
    n = term.shape[0]
    stream = torch.cuda.current_stream().cuda_stream
    if host_partition:
        part = np.empty(max(n, 1), np.uint8)
        st = W.pr_solve(n, dn.data_ptr(), dr.data_ptr(), dt.data_ptr(), part.ctypes.data, 0, W.params(**kw), stream)
    else:
        pd = torch.empty(max(n, 1), dtype=torch.uint8, device="cuda")
        st = W.pr_solve(n, dn.data_ptr(), dr.data_ptr(), dt.data_ptr(), pd.data_ptr(), 1, W.params(**kw), stream)
        part = pd.cpu().numpy()
    return part[:n], st, dr.cpu().numpy().view(np.uint32), dt.cpu().numpy()


def _check(orc, nbr, res, term, out):
    part, st, r_out, t_out = out
    bk = orc.bk_solve(nbr, res, term)
    assert st.raw_flow == bk.raw_flow
    assert np.array_equal(part, bk.partition)
    assert orc.certificate(nbr, res, term, r_out, t_out, part, st.raw_flow) == 0


@pytest.mark.parametrize("seed", range(12))
def test_forced_renumbering_small_graphs(orc, torch_cuda, seed):
    """renumber = 1 on random graphs of 2..3000 nodes (isolated nodes and
    several components included: nodes no seed reaches go last)."""
    n = int(np.random.default_rng(seed).integers(2, 3000))
    nbr, res, term = S.random_graph(n, 700 + seed, cap_max=int(3 + 11 * seed))
    out = _solve(torch_cuda, nbr, res, term, renumber=1, host_partition=bool(seed & 1))
    assert out[1].renumbered == 1
    _check(orc, nbr, res, term, out)


def test_shuffled_delaunay_graph_is_renumbered(orc, torch_cuda):
    """A 1.9M-tet Delaunay graph in the tetrahedraliser's spatial order is
    solved as given; the same graph shuffled is detected and renumbered; both
    equal BK on their own numbering (the same flow)."""
    g, _, _ = S.delaunay_genus3_graph(300_000, 31, method="native", device="cpu")
    nbr = g["nbr"].numpy()
    res = g["res"].numpy().view(np.uint32)
    term = g["term"].numpy()
    assert term.shape[0] > (1 << 20)
    a = _solve(torch_cuda, nbr, res, term)
    assert a[1].renumbered == 0 and a[1].t_order_ms > 0
    _check(orc, nbr, res, term, a)
    sn, sr, st_ = _shuffled(nbr, res, term, 5)
    b = _solve(torch_cuda, sn, sr, st_)
    assert b[1].renumbered == 1
    _check(orc, sn, sr, st_, b)
    assert b[1].raw_flow == a[1].raw_flow
    c = _solve(torch_cuda, sn, sr, st_, renumber=-1, host_partition=True)
    assert c[1].renumbered == 0
    assert c[1].raw_flow == a[1].raw_flow and np.array_equal(c[0], b[0])


def test_renumbering_keeps_invalid_ids_invalid(torch_cuda):
    """An id out of range fails the renumbered solve as it fails the solve in
    order (the copy keeps it as it is)."""
    nbr, res, term = S.random_graph(500, 77, cap_max=9)
    nbr = nbr.copy()
    v, k = (int(x) for x in np.argwhere(nbr < 0)[0])
    nbr[v, k] = 500  # an empty slot: the only fault is the id
    for r in (-1, 1):
        with pytest.raises(W.WtvError, match="code=3") as e:
            _solve(torch_cuda, nbr, res, term, renumber=r)
        assert e.value.code == W.WTV_EINVARIANT
