"""GPU parity at the workloads' own sizes, against Boykov-Kolmogorov on the CPU
(oracle/wtv_oracle.c) graph by graph: flows and labels must be identical.

* 256 uneven soft-seed Kuhn graphs (50k to 4M tets, log-uniform; random shapes;
  lambda U[1,30]) solved in one batched call;
* Poisson-arriving serving requests (100k to 20M tets, lambda 10) through the
  native server, each result equal to BK's and to the same graph solved alone;
* the 78M-tet soft-seed Kuhn 235^3 genus-3 graph (the single-graph stand-in);
* the Delaunay mesh of 2M genus-3 points (12.9M tets), scipy's Qhull;
* config 1 as written: the Delaunay mesh of 12M genus-3 points (~77.5M tets),
  the native tetrahedraliser.
"""
import time

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


def _concat(torch, graphs):
    sizes = [g["N"] for g in graphs]
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    nbr = torch.cat([g["nbr"] for g in graphs])
    res0 = torch.cat([g["res"] for g in graphs])
    term0 = torch.cat([g["term"] for g in graphs])
    return sizes, off, nbr, res0, term0


@pytest.mark.timeout(900)
def test_batched_256_uneven_graphs_vs_bk(orc, torch_cuda):
    torch = torch_cuda
    graphs = S.batch_workload(256, 20260003, 50_000, 4_000_000, device="cuda")
    sizes, off, nbr, res0, term0 = _concat(torch, graphs)
    del graphs
    res, term = res0.clone(), term0.clone()
    total = int(off[-1])
    part = torch.empty(total, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    flows, st = W.batch_solve(off, nbr.data_ptr(), res.data_ptr(), term.data_ptr(), part.data_ptr(), 1,
                              W.params(), stream)
    torch.cuda.synchronize()
    print(f"batched: {len(sizes)} graphs, {total} tets, {st.t_solve_ms:.1f} ms device", flush=True)
    nb_h = nbr.cpu().numpy()
    r0_h = res0.cpu().numpy().view(np.uint32)
    t0_h = term0.cpu().numpy()
    r_h = res.cpu().numpy().view(np.uint32)
    t_h = term.cpu().numpy()
    p_h = part.cpu().numpy()
    assert 200e6 < total < 260e6 and min(sizes) >= 40_000 and max(sizes) <= 4_200_000
    for g in range(len(sizes)):
        a, b = int(off[g]), int(off[g + 1])
        bk = orc.bk_solve(nb_h[a:b], r0_h[a:b], t0_h[a:b])
        assert int(flows[g]) == bk.raw_flow, g
        assert np.array_equal(p_h[a:b], bk.partition), g
        assert orc.certificate(nb_h[a:b], r0_h[a:b], t0_h[a:b], r_h[a:b], t_h[a:b], p_h[a:b], int(flows[g])) == 0, g
    assert int(np.sum(flows)) == st.raw_flow


@pytest.mark.timeout(900)
def test_serving_poisson_requests_vs_bk_and_alone(orc, torch_cuda):
    """48 requests arrive as a Poisson process at 150 per second, faster than
    they are solved one at a time, so they meet on the device; every served
    result equals the lone solve and BK."""
    torch = torch_cuda
    graphs = S.batch_workload(48, 20260004, 100_000, 20_000_000, device="cuda", lam_range=(10.0, 10.0))
    stream = torch.cuda.current_stream().cuda_stream
    reqs = []
    for g in graphs:
        n = g["N"]
        r = {"n": n, "nbr": g["nbr"], "res0": g["res"], "term0": g["term"], "res": g["res"].clone(),
             "term": g["term"].clone(), "part": torch.empty(n, dtype=torch.uint8, device="cuda")}
        st = W.pr_solve(n, r["nbr"].data_ptr(), r["res"].data_ptr(), r["term"].data_ptr(), r["part"].data_ptr(), 1,
                        W.params(), stream)
        r["alone_flow"] = st.raw_flow
        r["alone_part"] = r["part"].clone()
        r["res"].copy_(r["res0"])
        r["term"].copy_(r["term0"])
        r["part"].zero_()
        reqs.append(r)
    del graphs
    torch.cuda.synchronize()
    rng = np.random.Generator(np.random.PCG64(20260005))
    arrivals = np.cumsum(rng.exponential(1.0 / 150.0, size=len(reqs)))
    with W.Server(slots=4, max_batch=16) as srv:
        t0 = time.perf_counter()
        tickets = []
        for r, at in zip(reqs, arrivals):
            delay = at - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(r["n"], r["nbr"].data_ptr(), r["res"].data_ptr(), r["term"].data_ptr(),
                                      r["part"].data_ptr()))
        results = [srv.wait(t) for t in tickets]
    torch.cuda.synchronize()
    print(f"served {len(reqs)} requests, {sum(r['n'] for r in reqs)} tets, "
          f"mean batch {np.mean([x.batch_size for x in results]):.2f}", flush=True)
    for i, (r, x) in enumerate(zip(reqs, results)):
        assert x.status == 0, i
        assert x.raw_flow == r["alone_flow"], i
        assert torch.equal(r["part"], r["alone_part"]), i
        bk = orc.bk_solve(r["nbr"].cpu().numpy(), r["res0"].cpu().numpy().view(np.uint32), r["term0"].cpu().numpy())
        assert x.raw_flow == bk.raw_flow, i
        assert np.array_equal(r["part"].cpu().numpy(), bk.partition), i


@pytest.mark.timeout(900)
def test_kuhn235_genus3_soft_seed_78m_vs_bk(orc, torch_cuda):
    """The single-graph stand-in at full size: Kuhn 235^3 (77,867,250 tets),
    genus-3 solid, 2% flips, soft seeds, lambda 10."""
    torch = torch_cuda
    k, seed = 235, 20260001
    V, T, Nb = S.kuhn_mesh_t(k, seed=seed, device="cuda")
    shape = S.genus3(seed)
    sdf = S.tet_sdf_t(V, T, shape)
    N = int(T.shape[0])
    flips = torch.from_numpy(S.flip_mask(N, 0.02, seed + 1).astype(np.uint8)).cuda()
    labels = (sdf >= 0).to(torch.uint8) ^ flips
    g = S.soft_graph_t(V, T, Nb, labels, shape, 10.0, 1.0 / k, rev_slots=S.KUHN_REV_SLOT)
    del V, T, Nb, sdf, flips, labels
    nbr, res0, term0 = g["nbr"], g["res"], g["term"]
    res, term = res0.clone(), term0.clone()
    part = torch.empty(N, dtype=torch.uint8, device="cuda")
    st = W.pr_solve(N, nbr.data_ptr(), res.data_ptr(), term.data_ptr(), part.data_ptr(), 1, W.params(),
                    torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert N == 77_867_250
    bk = orc.bk_solve(nbr.cpu().numpy(), res0.cpu().numpy().view(np.uint32), term0.cpu().numpy())
    assert st.raw_flow == bk.raw_flow
    assert np.array_equal(part.cpu().numpy(), bk.partition)


@pytest.mark.timeout(900)
def test_delaunay_2m_points_vs_bk(orc, torch_cuda):
    """Delaunay mesh (scipy's Qhull) of 2M genus-3 points: 95% on the fused
    surface with half-voxel noise, 5% uniform; ~12.9M tets in Qhull's own order."""
    torch = torch_cuda
    P, shape, res = S.genus3_point_cloud(2_000_000, 20260006)
    m = S.delaunay_mesh(P, morton=False)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape, device="cuda"), 0.02, seed=20260007)
    g = S.soft_graph(m, lab, shape, 10.0, 1.0 / res, device="cuda")
    assert m.N > 12_000_000
    dn = torch.from_numpy(g.nbr).cuda()
    dr = torch.from_numpy(g.res.view(np.int32)).cuda()
    dt = torch.from_numpy(g.term).cuda()
    part = torch.empty(m.N, dtype=torch.uint8, device="cuda")
    st = W.pr_solve(m.N, dn.data_ptr(), dr.data_ptr(), dt.data_ptr(), part.data_ptr(), 1, W.params(),
                    torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    bk = orc.bk_solve(g.nbr, g.res, g.term)
    p_h = part.cpu().numpy()
    assert st.raw_flow == bk.raw_flow
    assert np.array_equal(p_h, bk.partition)
    assert orc.certificate(g.nbr, g.res, g.term, dr.cpu().numpy().view(np.uint32), dt.cpu().numpy(), p_h,
                           st.raw_flow) == 0


@pytest.mark.timeout(900)
def test_delaunay_12m_points_config1_vs_bk(orc, torch_cuda):
    """Config 1 as written: 12M genus-3 points (95% on the fused surface with
    0.5-voxel noise at resolution 1536, 5% uniform), their Delaunay mesh
    (~77.5M tets, native tetrahedraliser), soft seeds, lambda 10, 2% flips.
    Flow and labels equal BK's (~100 s on one core)."""
    import os
    torch = torch_cuda
    t0 = time.time()
    P, shape, res = S.genus3_point_cloud(12_000_000, 20260001)
    Q, V = S.quantise_points(P)
    del P
    tets, nbrs, _ = S.native_delaunay(Q, min(16, os.cpu_count() or 1))
    del Q
    N = tets.shape[0]
    print(f"config 1 mesh: {N} tets in {time.time() - t0:.1f} s", flush=True)
    assert 76e6 < N < 79e6
    Vd, Td, Nd = (torch.from_numpy(x).cuda() for x in (V, tets, nbrs))
    del tets, nbrs, V
    sdf = S.tet_sdf_t(Vd, Td, shape)
    flips = torch.from_numpy(S.flip_mask(N, 0.02, 20260002).astype(np.uint8)).cuda()
    lab = (sdf >= 0).to(torch.uint8) ^ flips
    g = S.soft_graph_t(Vd, Td, Nd, lab, shape, 10.0, 1.0 / res)
    del Vd, Td, Nd, sdf, flips, lab
    nbr, res0, term0 = g["nbr"], g["res"], g["term"]
    r, t = res0.clone(), term0.clone()
    part = torch.empty(N, dtype=torch.uint8, device="cuda")
    st = W.pr_solve(N, nbr.data_ptr(), r.data_ptr(), t.data_ptr(), part.data_ptr(), 1, W.params(),
                    torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    print(f"config 1 solve: {st.t_solve_ms:.1f} ms device, flow {st.raw_flow}", flush=True)
    nb_h, r0_h, t0_h = nbr.cpu().numpy(), res0.cpu().numpy().view(np.uint32), term0.cpu().numpy()
    del g, nbr, res0, term0, r, t
    bk = orc.bk_solve(nb_h, r0_h, t0_h)
    assert st.raw_flow == bk.raw_flow
    assert np.array_equal(part.cpu().numpy(), bk.partition)
