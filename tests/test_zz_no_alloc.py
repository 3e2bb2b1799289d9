"""wtv_pr_solve_noalloc, the allocation-free solve, against wtv_pr_solve through
the C-ABI: two calls in a row, the first at the workspace's maximum size
(1,048,576 nodes), the second smaller, each with the same flow and labels as
wtv_pr_solve and BK, and final residuals that pass the max-flow certificate."""
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


def _soft(k, seed, pad_to=None):
    m = S.kuhn_mesh(k, seed=seed)
    shape = S.genus3(seed)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.02, seed=seed + 1)
    g = S.soft_graph(m, lab, shape, 10.0, 1.0 / k, rev_slots=S.KUHN_REV_SLOT)
    nbr, res, term = g.nbr, g.res, g.term
    if pad_to is not None and pad_to > m.N:
        # isolated nodes (no arcs, no terminals): never active, never reached
        extra = pad_to - m.N
        nbr = np.concatenate([nbr, np.full((extra, 4), -1, np.int32)])
        res = np.concatenate([res, np.zeros((extra, 4), np.uint32)])
        term = np.concatenate([term, np.zeros(extra, np.int32)])
    return nbr, res, term


def _run(torch, fn, nbr, res, term, on_device):
    dn = torch.from_numpy(np.ascontiguousarray(nbr)).cuda()
    dr = torch.from_numpy(np.ascontiguousarray(res).view(np.int32)).cuda()
    dt = torch.from_numpy(np.ascontiguousarray(term)).cuda()
    n = int(term.shape[0])
    stream = torch.cuda.current_stream().cuda_stream
    if on_device:
        pd = torch.empty(n, dtype=torch.uint8, device="cuda")
        st = fn(n, dn.data_ptr(), dr.data_ptr(), dt.data_ptr(), pd.data_ptr(), 1, W.params(), stream)
        part = pd.cpu().numpy()
    else:
        part = np.empty(n, np.uint8)
        st = fn(n, dn.data_ptr(), dr.data_ptr(), dt.data_ptr(), part.ctypes.data, 0, W.params(), stream)
    torch.cuda.synchronize()
    return st, part, dr.cpu().numpy().view(np.uint32), dt.cpu().numpy()


def test_noalloc_equals_pr_solve_two_calls(orc, torch_cuda):
    cases = [(_soft(55, 41, pad_to=W.NOALLOC_MAX_NODES), True), (_soft(20, 43), False)]
    assert cases[0][0][2].shape[0] == W.NOALLOC_MAX_NODES
    for (nbr, res, term), on_device in cases:
        n = term.shape[0]
        print(f"noalloc: n={n} partition_on_device={on_device}", flush=True)
        st_ref, part_ref, _, _ = _run(torch_cuda, W.pr_solve, nbr, res, term, on_device)
        st, part, r_out, t_out = _run(torch_cuda, W.pr_solve_noalloc, nbr, res, term, on_device)
        assert st.raw_flow == st_ref.raw_flow
        assert np.array_equal(part, part_ref)
        assert orc.certificate(nbr, res, term, r_out, t_out, part, st.raw_flow) == 0
        bk = orc.bk_solve(nbr, res, term)
        assert st.raw_flow == bk.raw_flow and np.array_equal(part, bk.partition)
        assert st.global_relabels >= 2 and st.sub_rounds > 0


def test_noalloc_rejects_sizes_beyond_its_workspace(torch_cuda):
    with pytest.raises(W.WtvError, match="exceed its fixed workspace of 1048576 nodes") as e:
        W.pr_solve_noalloc(W.NOALLOC_MAX_NODES + 1, 0, 0, 0, 0, 1)
    assert e.value.code == W.WTV_EINVAL
    with pytest.raises(W.WtvError, match="profile mode"):
        nbr, res, term = S.random_graph(40, 5)
        dn = torch_cuda.from_numpy(nbr).cuda()
        dr = torch_cuda.from_numpy(res.view(np.int32)).cuda()
        dt = torch_cuda.from_numpy(term).cuda()
        pd = torch_cuda.empty(40, dtype=torch_cuda.uint8, device="cuda")
        W.pr_solve_noalloc(40, dn.data_ptr(), dr.data_ptr(), dt.data_ptr(), pd.data_ptr(), 1, W.params(profile=1))
