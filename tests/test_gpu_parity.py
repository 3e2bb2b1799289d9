"""GPU parity: the HIP solver, called through the C-ABI (libwtv_gpupr.so), against
the CPU oracle on the same seeded inputs.  Integer work: flows, labels and final
residual certificates must match exactly."""
import ctypes as C
import os

import numpy as np
import pytest

import wtv_capi as W
import wtv_synth as S

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    W.lib()
    return torch


def gpu_solve(torch, nbr, res, term, device_partition=False, **kw):
    """wtv_pr_solve through ctypes; returns (partition, stats dict, res, term) on the host."""
    dn = torch.from_numpy(np.ascontiguousarray(nbr, np.int32)).cuda()
    dr = torch.from_numpy(np.ascontiguousarray(res, np.uint32).view(np.int32)).cuda()
    dt = torch.from_numpy(np.ascontiguousarray(term, np.int32)).cuda()
    n = int(dt.numel())
    stream = torch.cuda.current_stream().cuda_stream
    prm = W.params(**kw)
    if device_partition:
        part_d = torch.empty(max(n, 1), dtype=torch.uint8, device="cuda")
        st = W.pr_solve(n, dn.data_ptr(), dr.data_ptr(), dt.data_ptr(), part_d.data_ptr(), 1, prm, stream)
        part = part_d.cpu().numpy()[:n]
    else:
        part = np.empty(max(n, 1), np.uint8)
        st = W.pr_solve(n, dn.data_ptr(), dr.data_ptr(), dt.data_ptr(), part.ctypes.data, 0, prm, stream)
        part = part[:n]
    return part, st.as_dict(), dr.cpu().numpy().view(np.uint32), dt.cpu().numpy()


def assert_same(orc, nbr, res, term, gpu, want_flow=None, want_part=None):
    part, st, r_out, t_out = gpu
    if want_flow is None:
        sol = orc.bk_solve(nbr, res, term)
        want_flow, want_part = sol.raw_flow, sol.partition
    assert st["raw_flow"] == want_flow
    assert np.array_equal(part, want_part)
    assert orc.certificate(nbr, res, term, r_out, t_out, part, st["raw_flow"]) == 0


@pytest.mark.parametrize("seed", range(40))
def test_random_graphs(orc, torch_cuda, seed):
    n = int(np.random.default_rng(seed).integers(6, 400))
    nbr, res, term = S.random_graph(n, seed, cap_max=int(2 + seed * 7))
    a = orc.pr_solve(nbr, res, term)
    if orc.certificate(nbr, res, term, a.res, a.term, a.partition, a.raw_flow) != 0:
        pytest.skip("instance hits the reference's source-height-n convention")
    assert_same(orc, nbr, res, term, gpu_solve(torch_cuda, nbr, res, term), a.raw_flow, a.partition)


def test_reference_source_height_quirk(orc, torch_cuda):
    nbr = np.array([[1, -1, -1, -1], [0, -1, -1, -1]], np.int32)
    res = np.array([[5, 0, 0, 0], [0, 0, 0, 0]], np.uint32)
    term = np.array([5, -5], np.int32)
    part, st, _, _ = gpu_solve(torch_cuda, nbr, res, term)
    a = orc.pr_solve(nbr, res, term)
    assert st["raw_flow"] == a.raw_flow == 0
    assert part.tolist() == a.partition.tolist() == [1, 1]


@pytest.mark.parametrize("name", sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz")))
def test_golden(orc, torch_cuda, name):
    import wtivo_gpupr as G
    z = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    if "V" in z.files:
        T = torch_cuda.from_numpy(z["tets"]).cuda()
        Nb = torch_cuda.from_numpy(z["nbrs"]).cuda()
        out = G.graph_cut_fast(z["V"], T, Nb, z["labels"], float(z["fill_T"]), 4)
        assert len(out) == 18
        assert np.array_equal(out[0], z["out_labels"])
        assert out[1] == int(z["adjusted_flow"]) and out[2] == int(z["raw_flow"])
        assert out[3] == int(z["free_n"]) * 36
    else:
        gpu = gpu_solve(torch_cuda, z["nbr"], z["res"], z["term"])
        assert_same(orc, z["nbr"], z["res"], z["term"], gpu, int(z["raw_flow"]), z["partition"])


def _kuhn_soft(k=32, seed=20260000):
    m = S.kuhn_mesh(k, seed=seed)
    shape = S.sphere_with_holes(seed)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.02, seed=seed + 1)
    return m, lab, S.soft_graph(m, lab, shape, 10.0, 1.0 / k)


def test_kuhn32_soft_seed(orc, torch_cuda):
    """196,608-tet jittered Kuhn grid, sphere with 3 holes, 2% flips, soft seeds."""
    m, lab, g = _kuhn_soft()
    a = orc.pr_solve(g.nbr, g.res, g.term)
    gpu = gpu_solve(torch_cuda, g.nbr, g.res, g.term, device_partition=True)
    assert_same(orc, g.nbr, g.res, g.term, gpu, a.raw_flow, a.partition)
    # identical cut energy from the fp32 capacities (same labels -> same energy)
    assert S.cut_energy(g, gpu[0]) == S.cut_energy(g, a.partition)


def test_kuhn32_reference_model_graph_cut_fast(orc, torch_cuda):
    """The drop-in entry point on the Kuhn 32^3 mesh (196,608 tets) with the reference capacity formula."""
    import wtivo_gpupr as G
    m, lab, _ = _kuhn_soft()
    r = orc.reduce(m.V, m.tets, m.nbrs, lab, 10.0)
    sol = orc.pr_solve(r.nbr, r.res, r.term)
    want = orc.scatter(lab, r.free_old, sol.partition)
    T = torch_cuda.from_numpy(m.tets).cuda()
    Nb = torch_cuda.from_numpy(m.nbrs).cuda()
    out = G.graph_cut_fast(m.V, T, Nb, lab, 10.0, 8, 1024, 2000000, 8)
    assert np.array_equal(out[0], want)
    assert out[1] == sol.raw_flow + r.info["base_flow"] and out[2] == sol.raw_flow
    inf = r.info
    assert out[3] == r.n * 36                     # graph_bytes
    assert out[5] == inf["full_refs"] // 2         # full_E
    assert out[6] == inf["free_refs"] // 2         # free_E
    assert out[7] == inf["fixed_source"] and out[8] == inf["fixed_sink"]
    assert out[9] == inf["terminal_edges"] and out[10] == 8 and out[15] == 8
    assert out[13] == m.N * 4 + r.n * 4            # map_bytes
    # the same call through the raw C-ABI
    lab_out = np.empty(m.N, np.uint8)
    res = W.WtvCutResult()
    rc = W.lib().wtv_graph_cut_fast(m.V.ctypes.data, m.V.shape[0], T.data_ptr(), Nb.data_ptr(), m.N,
                                    lab.ctypes.data, 10.0, 4, 1024, 2000000, 8, lab_out.ctypes.data,
                                    C.byref(res), torch_cuda.cuda.current_stream().cuda_stream)
    W.check(rc)
    assert np.array_equal(lab_out, want) and res.adjusted_flow == out[1]


@pytest.mark.parametrize("seed", [0, 1])
def test_graph_cut_fast_three_label_values(orc, torch_cuda, seed):
    """Labels 0, 1 and 2 through the drop-in call.  Any two different nonzero
    labels count as different in the capacities (wtivo_gpupr_bindings.cpp:93),
    a fixed sink tet keeps its input label in the output, and free tets get the
    partition (:357-364): the whole 18-tuple's labels and flows equal the
    oracle's reduce -> solve -> scatter."""
    import wtivo_gpupr as G
    rng = np.random.default_rng(300 + seed)
    m = S.kuhn_mesh(12 + 6 * seed, jitter=0.3, seed=300 + seed)
    lab = rng.choice(np.array([0, 1, 2], np.uint8), size=m.N, p=[0.25, 0.6, 0.15])
    r = orc.reduce(m.V, m.tets, m.nbrs, lab, 7.5)
    assert r.n > 0 and (lab[r.free_old] == 2).any()
    sol = orc.pr_solve(r.nbr, r.res, r.term)
    want = orc.scatter(lab, r.free_old, sol.partition)
    out = G.graph_cut_fast(m.V, torch_cuda.from_numpy(m.tets).cuda(), torch_cuda.from_numpy(m.nbrs).cuda(), lab, 7.5)
    assert np.array_equal(out[0], want)
    assert out[2] == sol.raw_flow and out[1] == sol.raw_flow + r.info["base_flow"]
    assert (out[0] == 2).any()  # fixed sink tets labelled 2 keep their label


def _resident_cut(torch, V, T, Nb, lab_d, out_d, fill_T=10.0):
    res = W.WtvCutResult()
    Vd = torch.from_numpy(np.ascontiguousarray(V)).cuda()
    W.check(W.lib().wtv_graph_cut_fast_resident(Vd.data_ptr(), V.shape[0], T.data_ptr(), Nb.data_ptr(), T.shape[0],
                                                lab_d.data_ptr(), fill_T, 1024, 2000000, 8, out_d.data_ptr(),
                                                C.byref(res), torch.cuda.current_stream().cuda_stream))
    return res


@pytest.mark.parametrize("k", [32, 60])
def test_graph_cut_fast_resident_equals_host_call(orc, torch_cuda, k):
    """Every array in HBM gives the host call's labels, flows and counters, also in place."""
    torch = torch_cuda
    m, lab, _ = _kuhn_soft(k=k)
    T = torch.from_numpy(m.tets).cuda()
    Nb = torch.from_numpy(m.nbrs).cuda()
    host_out = np.empty(m.N, np.uint8)
    hres = W.WtvCutResult()
    W.check(W.lib().wtv_graph_cut_fast(m.V.ctypes.data, m.V.shape[0], T.data_ptr(), Nb.data_ptr(), m.N,
                                       lab.ctypes.data, 10.0, 4, 1024, 2000000, 8, host_out.ctypes.data,
                                       C.byref(hres), torch.cuda.current_stream().cuda_stream))
    lab_d = torch.from_numpy(lab).cuda()
    out_d = torch.empty_like(lab_d)
    rres = _resident_cut(torch, m.V, T, Nb, lab_d, out_d)
    assert np.array_equal(out_d.cpu().numpy(), host_out)
    for f in ("adjusted_flow", "raw_flow", "graph_bytes", "full_edges", "free_edges", "fixed_source", "fixed_sink",
              "terminal_edges", "map_bytes"):
        assert getattr(rres, f) == getattr(hres, f), f
    assert np.array_equal(lab_d.cpu().numpy(), lab)          # input untouched
    inplace = lab_d.clone()
    _resident_cut(torch, m.V, T, Nb, inplace, inplace)
    assert np.array_equal(inplace.cpu().numpy(), host_out)
    if k == 32:
        r = orc.reduce(m.V, m.tets, m.nbrs, lab, 10.0)
        sol = orc.pr_solve(r.nbr, r.res, r.term)
        assert np.array_equal(host_out, orc.scatter(lab, r.free_old, sol.partition))


def test_graph_cut_fast_resident_no_free_tets(torch_cuda):
    torch = torch_cuda
    m, _, _ = _kuhn_soft(k=8)
    lab = np.zeros(m.N, np.uint8)                           # every tet fixed to the source
    T = torch.from_numpy(m.tets).cuda()
    Nb = torch.from_numpy(m.nbrs).cuda()
    lab_d = torch.from_numpy(lab).cuda()
    out_d = torch.full_like(lab_d, 7)
    res = _resident_cut(torch, m.V, T, Nb, lab_d, out_d)
    assert (out_d.cpu().numpy() == 0).all() and res.raw_flow == 0 and res.fixed_source == m.N
    assert W.lib().wtv_graph_cut_fast_resident(None, 0, None, None, 10, None, 10.0, 1024, 0, 8, None,
                                               None, None) == W.WTV_EINVAL


def gpu_solve_flow(torch, nbr, res, term, **kw):
    """wtv_pr_solve_flow through ctypes: (partition, stats, returned excess, res, term) on the host."""
    dn = torch.from_numpy(np.ascontiguousarray(nbr, np.int32)).cuda()
    dr = torch.from_numpy(np.ascontiguousarray(res, np.uint32).view(np.int32)).cuda()
    dt = torch.from_numpy(np.ascontiguousarray(term, np.int32)).cuda()
    n = int(dt.numel())
    part = np.empty(max(n, 1), np.uint8)
    st, ret = W.pr_solve_flow(n, dn.data_ptr(), dr.data_ptr(), dt.data_ptr(), part.ctypes.data, 0, W.params(**kw),
                              torch.cuda.current_stream().cuda_stream)
    return part[:n], st.as_dict(), ret, dr.cpu().numpy().view(np.uint32), dt.cpu().numpy()


def preflow_excess(res0, term0, res, term):
    """Per-vertex excess of a solver state (zero everywhere for a flow)."""
    src = np.where(term0 > 0, term.astype(np.int64), 0)
    sink = np.where(term0 <= 0, term.astype(np.int64) - term0, 0)
    return src - sink - (res0.astype(np.int64) - res.astype(np.int64)).reshape(-1, 4).sum(axis=1)


def _check_flow_mode(orc, torch, nbr, res, term):
    b = orc.bk_solve(nbr, res, term)
    part, st, ret, r_out, t_out = gpu_solve_flow(torch, nbr, res, term)
    assert st["raw_flow"] == b.raw_flow and np.array_equal(part, b.partition)
    assert orc.flow_certificate(nbr, res, term, r_out, t_out, part, st["raw_flow"]) == 0
    # what came back is the excess the plain solve leaves on the source side
    p1, st1, r1, t1 = gpu_solve(torch, nbr, res, term)
    ex = preflow_excess(np.asarray(res, np.uint32).reshape(-1, 4), np.asarray(term, np.int32), r1, t1)
    assert ret == int(ex.sum()) and (ex >= 0).all()
    return ret


@pytest.mark.parametrize("seed", range(6))
def test_flow_mode_random_graphs(orc, torch_cuda, seed):
    rng = np.random.default_rng(seed + 40)
    nbr, res, term = S.random_graph(int(rng.integers(50, 3000)), seed + 40, cap_max=int(5 + 9 * seed))
    _check_flow_mode(orc, torch_cuda, nbr, res, term)


def test_flow_mode_kuhn32_soft_seed(orc, torch_cuda):
    _, _, g = _kuhn_soft()
    assert _check_flow_mode(orc, torch_cuda, g.nbr, g.res, g.term) > 0


def test_flow_mode_kuhn150_integer_shell(orc, torch_cuda):
    k = 150
    m = S.kuhn_mesh(k, seed=20260004)
    shape = S.thin_shell_with_tunnels(20260004, 1.0 / k, tunnels=40)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape, device="cuda"), 0.0)
    g = S.lognormal_int_graph(m, lab, 100.0, seed=20260005, rev_slots=S.KUHN_REV_SLOT)
    print(f"Kuhn 150^3 integer shell, flow mode: {m.N} tets", flush=True)
    _check_flow_mode(orc, torch_cuda, g.nbr, g.res, g.term)


def test_flow_mode_empty_and_no_excess(orc, torch_cuda):
    st, ret = W.pr_solve_flow(0, 0, 0, 0, 0, 0)
    assert st.raw_flow == 0 and ret == 0
    # everything reaches the sink; two isolated vertices keep the path below the
    # source height n (the reference's quirk blocks it at n = 2)
    nbr = np.array([[1, -1, -1, -1], [0, -1, -1, -1], [-1] * 4, [-1] * 4], np.int32)
    res = np.array([[5, 0, 0, 0], [5, 0, 0, 0], [0] * 4, [0] * 4], np.uint32)
    term = np.array([5, -5, 0, 0], np.int32)
    part, st, ret, r_out, t_out = gpu_solve_flow(torch_cuda, nbr, res, term)
    assert st["raw_flow"] == 5 and ret == 0
    assert orc.flow_certificate(nbr, res, term, r_out, t_out, part, 5) == 0


def test_stress_integer_shell(orc, torch_cuda):
    """Scaled-down adversarial stress: thin shell + tunnels, log-normal caps in [1,2^20], lambda 100."""
    k = 40
    m = S.kuhn_mesh(k, seed=20260004)
    shape = S.thin_shell_with_tunnels(20260004, 1.0 / k, tunnels=40)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.0)
    g = S.lognormal_int_graph(m, lab, 100.0, seed=20260005)
    a = orc.bk_solve(g.nbr, g.res, g.term)
    gpu = gpu_solve(torch_cuda, g.nbr, g.res, g.term)
    assert_same(orc, g.nbr, g.res, g.term, gpu, a.raw_flow, a.partition)


def test_kuhn150_integer_shell_full_size(orc, torch_cuda):
    """The adversarial stress at full size: Kuhn 150^3 = 20,250,000 tets, thin shell
    (wall 1.5 cells) with 40 tunnels, log-normal capacities quantised to [1, 2^20],
    lambda 100.  Flow and labels bit-exact against Boykov-Kolmogorov, certificate."""
    k = 150
    m = S.kuhn_mesh(k, seed=20260004)
    shape = S.thin_shell_with_tunnels(20260004, 1.0 / k, tunnels=40)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape, device="cuda"), 0.0)
    g = S.lognormal_int_graph(m, lab, 100.0, seed=20260005, rev_slots=S.KUHN_REV_SLOT)
    print(f"Kuhn 150^3 integer shell: {m.N} tets", flush=True)
    a = orc.bk_solve(g.nbr, g.res, g.term)
    gpu = gpu_solve(torch_cuda, g.nbr, g.res, g.term, device_partition=True)
    assert_same(orc, g.nbr, g.res, g.term, gpu, a.raw_flow, a.partition)


@pytest.mark.parametrize("kw", [dict(local_steps=0), dict(local_steps=100), dict(global_relabel_period=-5),
                                dict(global_relabel_period=1, local_steps=1), dict(grid_wgs=7),
                                dict(max_rounds=0), dict(wg_size=512), dict(wg_size=1024),
                                dict(wg_size=1024, grid_wgs=3, local_steps=2)])
def test_clamped_parameters_same_answer(orc, torch_cuda, kw):
    nbr, res, term = S.random_graph(2000, 77, cap_max=30)
    a = orc.bk_solve(nbr, res, term)
    assert_same(orc, nbr, res, term, gpu_solve(torch_cuda, nbr, res, term, **kw), a.raw_flow, a.partition)


@pytest.mark.parametrize("wg", [512, 1024])
def test_workgroup_size_same_answer_large(orc, torch_cuda, wg):
    """Every list kernel at 512 and 1024 threads per workgroup, in single and
    batched mode, on graphs large enough for the queue flushes to trigger."""
    _, _, g = _kuhn_soft(48, 90 + wg)
    a = orc.bk_solve(g.nbr, g.res, g.term)
    assert_same(orc, g.nbr, g.res, g.term, gpu_solve(torch_cuda, g.nbr, g.res, g.term, wg_size=wg, tail_cap=-1),
                a.raw_flow, a.partition)
    graphs = [(g.nbr, g.res, g.term), S.random_graph(3000, wg, cap_max=40)]
    off, flows, part, _, _, _ = gpu_batch(torch_cuda, graphs, wg_size=wg)
    for i, (nbr, res, term) in enumerate(graphs):
        sol = orc.bk_solve(nbr, res, term)
        assert flows[i] == sol.raw_flow and np.array_equal(part[int(off[i]):int(off[i + 1])], sol.partition)


def test_workgroup_size_rejected(torch_cuda):
    nbr, res, term = S.random_graph(50, 3)
    for wg in (64, 300, 2048):
        with pytest.raises(W.WtvError, match="wg_size"):
            gpu_solve(torch_cuda, nbr, res, term, wg_size=wg)


@pytest.mark.parametrize("seed", range(3))
def test_extreme_capacities(orc, torch_cuda, seed):
    """Arc capacities up to 2^30 - 1 and terminals up to +-(2^31 - 1), the largest
    the reference's build admits: 64-bit excess and flow sums, no overflow."""
    nbr, res, term = S.random_graph(4000, 300 + seed, cap_max=(1 << 30) - 1, term_max=(1 << 31) - 1, p_term=0.9)
    a = orc.bk_solve(nbr, res, term)
    assert a.raw_flow > (1 << 36)  # far beyond 32-bit sums
    assert_same(orc, nbr, res, term, gpu_solve(torch_cuda, nbr, res, term), a.raw_flow, a.partition)


@pytest.mark.parametrize("t", [0, 7, -7])
def test_single_vertex(orc, torch_cuda, t):
    nbr = np.full((1, 4), -1, np.int32)
    res = np.zeros((1, 4), np.uint32)
    term = np.array([t], np.int32)
    part, st, r_out, t_out = gpu_solve(torch_cuda, nbr, res, term)
    a = orc.bk_solve(nbr, res, term)
    assert st["raw_flow"] == a.raw_flow == 0
    assert part.tolist() == a.partition.tolist()


def test_deterministic_labels_and_flow(torch_cuda):
    _, _, g = _kuhn_soft(16, 5)
    p1, s1, r1, t1 = gpu_solve(torch_cuda, g.nbr, g.res, g.term)
    p2, s2, r2, t2 = gpu_solve(torch_cuda, g.nbr, g.res, g.term)
    assert s1["raw_flow"] == s2["raw_flow"] and np.array_equal(p1, p2)


def test_max_rounds_limit(torch_cuda):
    _, _, g = _kuhn_soft(16, 6)
    with pytest.raises(W.WtvError, match="max macro rounds exceeded") as e:
        gpu_solve(torch_cuda, g.nbr, g.res, g.term, max_rounds=1, local_steps=1)
    assert e.value.code == W.WTV_ELIMIT


def test_time_limit_inside_the_discharge_loop(orc, torch_cuda):
    """The wall-clock limit (gpu_push_relabel_fast.cu:538-544) holds inside the
    pipelined macro-round loop: with no periodic or adaptive relabel and the
    one-workgroup tail off, a 1.57M-tet graph never leaves that loop, and a 2 ms
    limit must end the solve with WTV_ELIMIT (max_rounds is only a backstop).
    The kernels still queued finish before the call returns: the next solve on
    the same thread gives the exact answer."""
    k = 64
    m = S.kuhn_mesh(k, seed=9)
    shape = S.genus3(9)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.02, seed=10)
    g = S.soft_graph(m, lab, shape, 10.0, 1.0 / k, rev_slots=S.KUHN_REV_SLOT)
    with pytest.raises(W.WtvError, match=r"solve exceeded 0\.002s safety limit") as e:
        gpu_solve(torch_cuda, g.nbr, g.res, g.term, global_relabel_period=0, relabel_alpha=-1.0, tail_cap=-1,
                  local_steps=1, time_limit_s=0.002, max_rounds=20000)
    assert e.value.code == W.WTV_ELIMIT
    part, st, _, _ = gpu_solve(torch_cuda, g.nbr, g.res, g.term, device_partition=True)
    assert S.quantised_cut(g.nbr, g.res, g.term, part) == st["raw_flow"]


def test_asymmetric_neighbours_report_invariant(torch_cuda):
    nbr, res, term = S.random_graph(50, 3)
    v = int(np.argwhere(nbr >= 0)[0][0])
    k = int(np.argwhere(nbr[v] >= 0)[0][0])
    u = int(nbr[v, k])
    nbr[u, nbr[u] == v] = -1  # drop the reverse arc
    with pytest.raises(W.WtvError, match="code=2") as e:
        gpu_solve(torch_cuda, nbr, res, term)
    assert e.value.code == W.WTV_EINVARIANT


def test_graph_cut_fast_error_paths(torch_cuda):
    import wtivo_gpupr as G
    m = S.kuhn_mesh(6, seed=1)
    lab = np.ones(m.N, np.uint8)
    lab[: m.N // 3] = 0
    T = torch_cuda.from_numpy(m.tets).cuda()
    bad = m.nbrs.copy()
    bad[10, 1] = m.N + 5
    with pytest.raises(RuntimeError, match="neighbor out of range"):
        G.graph_cut_fast(m.V, T, torch_cuda.from_numpy(bad).cuda(), lab, 10.0)
    odd = m.nbrs.copy()
    i = int(np.flatnonzero((odd >= 0).all(1))[0])
    odd[i, 2] = -1
    with pytest.raises(RuntimeError, match="neighbor reference count not even"):
        G.graph_cut_fast(m.V, T, torch_cuda.from_numpy(odd).cuda(), lab, 10.0)
    Nb = torch_cuda.from_numpy(m.nbrs).cuda()
    with pytest.raises(RuntimeError, match="invalid graph reduction/capacity state"):
        G.graph_cut_fast(m.V, T, Nb, lab, 1e15)  # capacities beyond 2^30
    with pytest.raises(RuntimeError, match="tets/neighbors must be int32"):
        G.graph_cut_fast(m.V, T.long(), Nb, lab, 10.0)
    with pytest.raises(RuntimeError, match="labels size mismatch"):
        G.graph_cut_fast(m.V, T, Nb, lab[:-1], 10.0)
    # all label 0 -> no free nodes: labels returned unchanged, full 18-tuple
    z = np.zeros(m.N, np.uint8)
    out = G.graph_cut_fast(m.V, T, Nb, z, 10.0)
    assert len(out) == 18 and np.array_equal(out[0], z) and out[1] == 0 and out[7] == m.N


def test_larger_graph_certificate_and_bk(orc, torch_cuda):
    """1.57M tets (64^3): GPU flow == integer capacity of the GPU cut == BK flow."""
    k = 64
    m = S.kuhn_mesh(k, seed=9)
    shape = S.genus3(9)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.02, seed=10)
    g = S.soft_graph(m, lab, shape, 10.0, 1.0 / k, rev_slots=S.KUHN_REV_SLOT)
    part, st, r_out, t_out = gpu_solve(torch_cuda, g.nbr, g.res, g.term, device_partition=True)
    assert S.quantised_cut(g.nbr, g.res, g.term, part) == st["raw_flow"]
    b = orc.bk_solve(g.nbr, g.res, g.term)
    assert b.raw_flow == st["raw_flow"] and np.array_equal(b.partition, part)


@pytest.mark.parametrize("npts,morton", [(60_000, True), (60_000, False), (150_000, True)])
def test_delaunay_genus3_vs_bk(orc, torch_cuda, npts, morton):
    """Delaunay mesh of a genus-3 point cloud (95% on the surface with half-voxel
    noise, 5% uniform), soft-seed capacities, lambda 10, 2% flips: GPU flow and
    labels equal BK's, and the GPU's final residuals pass the certificate."""
    P, shape, res = S.genus3_point_cloud(npts, 40 + npts // 1000)
    m = S.delaunay_mesh(P, morton=morton)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.02, seed=npts)
    g = S.soft_graph(m, lab, shape, 10.0, 1.0 / res)
    gpu = gpu_solve(torch_cuda, g.nbr, g.res, g.term, device_partition=True)
    assert_same(orc, g.nbr, g.res, g.term, gpu)
    assert S.quantised_cut(g.nbr, g.res, g.term, gpu[0]) == gpu[1]["raw_flow"]


# ------------------------------------------------------------ batched mode
def _batch(graphs):
    off = np.zeros(len(graphs) + 1, np.uint64)
    for i, (nbr, res, term) in enumerate(graphs):
        off[i + 1] = off[i] + term.shape[0]
    nbr = np.concatenate([g[0] for g in graphs]).astype(np.int32)
    res = np.concatenate([g[1] for g in graphs]).astype(np.uint32)
    term = np.concatenate([g[2] for g in graphs]).astype(np.int32)
    return off, nbr, res, term


def gpu_batch(torch, graphs, device_partition=True, **kw):
    off, nbr, res, term = _batch(graphs)
    dn = torch.from_numpy(nbr).cuda()
    dr = torch.from_numpy(res.view(np.int32)).cuda()
    dt = torch.from_numpy(term).cuda()
    n = int(off[-1])
    stream = torch.cuda.current_stream().cuda_stream
    if device_partition:
        part_d = torch.empty(max(n, 1), dtype=torch.uint8, device="cuda")
        flows, st = W.batch_solve(off, dn.data_ptr(), dr.data_ptr(), dt.data_ptr(), part_d.data_ptr(), 1,
                                  W.params(**kw), stream)
        part = part_d.cpu().numpy()[:n]
    else:
        part = np.empty(max(n, 1), np.uint8)
        flows, st = W.batch_solve(off, dn.data_ptr(), dr.data_ptr(), dt.data_ptr(), part.ctypes.data, 0,
                                  W.params(**kw), stream)
    return off, flows, part, dr.cpu().numpy().view(np.uint32), dt.cpu().numpy(), st


def test_batch_matches_individual_solves(orc, torch_cuda):
    graphs = []
    for seed in range(12):
        graphs.append(S.random_graph(int(50 + seed * 37), 500 + seed, cap_max=25))
    # the quirk graph: its answer depends on its own source height (2), not the union's
    graphs.append((np.array([[1, -1, -1, -1], [0, -1, -1, -1]], np.int32),
                   np.array([[5, 0, 0, 0], [0, 0, 0, 0]], np.uint32), np.array([5, -5], np.int32)))
    graphs.append(tuple(np.zeros((0, 4), np.int32) if i < 2 else np.zeros(0, np.int32) for i in range(3)))
    _, _, g = _kuhn_soft(16, 8)
    graphs.append((g.nbr, g.res, g.term))
    off, flows, part, r_out, t_out, st = gpu_batch(torch_cuda, graphs)
    for i, (nbr, res, term) in enumerate(graphs):
        a, b = int(off[i]), int(off[i + 1])
        if a == b:
            assert flows[i] == 0
            continue
        o = orc.pr_solve(nbr, res, term)
        assert flows[i] == o.raw_flow, i
        assert np.array_equal(part[a:b], o.partition), i
        if orc.certificate(nbr, res, term, o.res, o.term, o.partition, o.raw_flow) == 0:
            assert orc.certificate(nbr, res, term, r_out[a:b], t_out[a:b], part[a:b], int(flows[i])) == 0
    assert int(flows.sum()) == st.raw_flow


def test_batch_soft_graphs_vs_bk(orc, torch_cuda):
    rng = np.random.default_rng(3)
    graphs, metas = [], []
    for i in range(6):
        k = int(rng.integers(10, 28))
        m = S.kuhn_mesh(k, seed=100 + i)
        shape = S.random_shape(rng)
        lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.02, seed=200 + i)
        g = S.soft_graph(m, lab, shape, float(rng.uniform(1, 30)), 1.0 / k, rev_slots=S.KUHN_REV_SLOT)
        graphs.append((g.nbr, g.res, g.term))
    off, flows, part, _, _, st = gpu_batch(torch_cuda, graphs, device_partition=False, local_steps=4)
    for i, (nbr, res, term) in enumerate(graphs):
        b = orc.bk_solve(nbr, res, term)
        assert flows[i] == b.raw_flow
        assert np.array_equal(part[int(off[i]):int(off[i + 1])], b.partition)


def test_batch_delaunay_and_kuhn_graphs_vs_bk(orc, torch_cuda):
    """Irregular (Delaunay) and regular (Kuhn) graphs of uneven size in one batch:
    every graph's flow and labels equal BK's on that graph alone."""
    graphs = []
    for i, npts in enumerate((3_000, 20_000, 45_000)):
        P, shape, res = S.genus3_point_cloud(npts, 70 + i)
        m = S.delaunay_mesh(P, morton=bool(i % 2))
        lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.02, seed=71 + i)
        g = S.soft_graph(m, lab, shape, float(5 + 10 * i), 1.0 / res)
        graphs.append((g.nbr, g.res, g.term))
    _, _, g = _kuhn_soft(20, 72)
    graphs.append((g.nbr, g.res, g.term))
    off, flows, part, r_out, t_out, st = gpu_batch(torch_cuda, graphs)
    for i, (nbr, res, term) in enumerate(graphs):
        a, b = int(off[i]), int(off[i + 1])
        sol = orc.bk_solve(nbr, res, term)
        assert flows[i] == sol.raw_flow, i
        assert np.array_equal(part[a:b], sol.partition), i
        assert orc.certificate(nbr, res, term, r_out[a:b], t_out[a:b], part[a:b], int(flows[i])) == 0


def test_batch_relabels_skip_finished_graphs(orc, torch_cuda):
    """Small graphs (sizes not multiples of the 1024-vertex sweep tile) finish
    long before a large one; with a global relabel every 8 sub-rounds, the
    relabels in between pass over the finished graphs' tiles.  Results equal
    BK's and the run with the skip turned off, which examines more vertices."""
    graphs = [S.random_graph(n, 610 + i, cap_max=30) for i, n in enumerate((700, 1500, 2100, 3300, 777))]
    _, _, g = _kuhn_soft(40, 21)
    graphs.insert(2, (g.nbr, g.res, g.term))
    # 196,608 vertices without terminals: no excess, finished from the start,
    # every vertex examined by a seed sweep that does not skip it
    _, _, z = _kuhn_soft(32, 22)
    graphs.insert(4, (z.nbr, z.res, np.zeros_like(z.term)))
    runs = {}
    for skip in (0, -1):
        off, flows, part, r_out, t_out, st = gpu_batch(torch_cuda, graphs, global_relabel_period=8, local_steps=1,
                                                       relabel_alpha=-1.0, skip_finished=skip)
        runs[skip] = (flows.copy(), part.copy(), st)
        for i, (nbr, res, term) in enumerate(graphs):
            a, b = int(off[i]), int(off[i + 1])
            w = orc.bk_solve(nbr, res, term)
            assert flows[i] == w.raw_flow, (skip, i)
            assert np.array_equal(part[a:b], w.partition), (skip, i)
    assert runs[0][2].global_relabels > 4 and runs[-1][2].global_relabels > 4
    # the skip ran: fewer vertices examined by the seed sweeps per relabel
    per = {k: v[2].bfs_seed_checks / v[2].global_relabels for k, v in runs.items()}
    assert per[0] < per[-1] - 100_000, per


def test_batch_rejects_bad_offsets_and_ids(torch_cuda):
    nbr, res, term = S.random_graph(40, 9)
    dn = torch_cuda.from_numpy(nbr).cuda()
    dr = torch_cuda.from_numpy(res.view(np.int32)).cuda()
    dt = torch_cuda.from_numpy(term).cuda()
    part = torch_cuda.empty(40, dtype=torch_cuda.uint8, device="cuda")
    with pytest.raises(W.WtvError, match="offsets"):
        W.batch_solve(np.array([0, 30, 20, 40], np.uint64), dn.data_ptr(), dr.data_ptr(), dt.data_ptr(),
                      part.data_ptr(), 1)
    # ids of the first graph exceed its size: rejected as a device invariant
    # (out-of-range id, and therefore also a missing reverse arc)
    with pytest.raises(W.WtvError, match="device invariant failed at prepare code=[23]"):
        W.batch_solve(np.array([0, 20, 40], np.uint64), dn.data_ptr(), dr.data_ptr(), dt.data_ptr(),
                      part.data_ptr(), 1)


# ------------------------------------------------------------ serving loop
def _dev_graph(torch, nbr, res, term):
    return (torch.from_numpy(np.ascontiguousarray(nbr, np.int32)).cuda(),
            torch.from_numpy(np.ascontiguousarray(res, np.uint32).view(np.int32)).cuda(),
            torch.from_numpy(np.ascontiguousarray(term, np.int32)).cuda(),
            torch.empty(max(term.shape[0], 1), dtype=torch.uint8, device="cuda"))


def test_server_results_equal_alone_and_in_the_mix(orc, torch_cuda):
    graphs = [S.random_graph(int(80 + 29 * i), 900 + i, cap_max=40) for i in range(10)]
    for k, seed in ((12, 1), (20, 2), (16, 3)):
        _, _, g = _kuhn_soft(k, seed)
        graphs.append((g.nbr, g.res, g.term))
    want = [orc.bk_solve(*g) for g in graphs]
    torch_cuda.cuda.synchronize()
    dev = [_dev_graph(torch_cuda, *g) for g in graphs]
    torch_cuda.cuda.synchronize()
    with W.Server(slots=2, max_batch=8) as srv:
        tickets = [srv.submit(g[2].shape[0], d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr())
                   for g, d in zip(graphs, dev)]
        res = [srv.wait(t) for t in tickets]
    sizes = [r.batch_size for r in res]
    assert max(sizes) > 1, "nothing was batched"
    for g, d, r, w in zip(graphs, dev, res, want):
        assert r.status == 0
        assert r.raw_flow == w.raw_flow
        assert np.array_equal(d[3].cpu().numpy()[:g[2].shape[0]], w.partition)
        assert r.submit_s <= r.start_s <= r.end_s
        # the residuals written back are a maximum preflow of this graph alone
        assert orc.certificate(g[0], g[1], g[2], d[1].cpu().numpy().view(np.uint32), d[2].cpu().numpy(),
                               w.partition, r.raw_flow) == 0


def test_server_isolates_a_failed_request(orc, torch_cuda):
    """Requests whose graphs are invalid (a missing reverse arc, an id out of
    range) fail alone at admission, and the requests admitted with them (some
    placed in the arena after them) still return their lone-solve flows and
    labels."""
    _, _, big = _kuhn_soft(24, 11)
    good = [S.random_graph(int(200 + 50 * i), 960 + i, cap_max=40) for i in range(5)]
    bad = S.random_graph(300, 3)
    nbr = bad[0].copy()
    v = int(np.argwhere(nbr >= 0)[0][0])
    u = int(nbr[v, int(np.argwhere(nbr[v] >= 0)[0][0])])
    nbr[u, nbr[u] == v] = -1  # drop the reverse arc
    bad = (nbr, bad[1], bad[2])
    # a second one with a neighbour id out of range (in an empty slot)
    b2 = S.random_graph(250, 4)
    n2 = b2[0].copy()
    r2, k2 = (int(x) for x in np.argwhere(n2 < 0)[0])
    n2[r2, k2] = 250
    graphs = [(big.nbr, big.res, big.term)] + good[:2] + [bad] + good[2:4] + [(n2, b2[1], b2[2])] + good[4:]
    dev = [_dev_graph(torch_cuda, *g) for g in graphs]
    torch_cuda.cuda.synchronize()
    # one slot: the first (large) request runs while the rest queue into one batch
    with W.Server(slots=1, max_batch=16) as srv:
        tickets = [srv.submit(g[2].shape[0], d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr())
                   for g, d in zip(graphs, dev)]
        res = [srv.wait(t) for t in tickets]
    assert res[3].status == W.WTV_EINVARIANT and res[6].status == W.WTV_EINVARIANT
    assert max(r.batch_size for r in res) > 1, "nothing was batched"
    for i, (g, d, r) in enumerate(zip(graphs, dev, res)):
        if i in (3, 6):
            continue
        w = orc.bk_solve(*g)
        assert r.status == 0 and r.raw_flow == w.raw_flow, i
        assert np.array_equal(d[3].cpu().numpy()[:g[2].shape[0]], w.partition), i


def test_server_orders_the_solve_after_the_request_stream(orc, torch_cuda):
    """A client writes its graphs asynchronously on its own stream, behind a
    ~50 ms spin kernel, and submits at once with that stream: the solves wait for
    the writes (the server records an event on the stream at submit), so every
    result equals BK's on the graph as written."""
    torch = torch_cuda
    graphs = [S.random_graph(int(300 + 40 * i), 970 + i, cap_max=40) for i in range(4)]
    _, _, g = _kuhn_soft(16, 12)
    graphs.append((g.nbr, g.res, g.term))
    dev = [_dev_graph(torch, np.zeros_like(a), np.zeros_like(b), np.zeros_like(c)) for a, b, c in graphs]
    host = [tuple(torch.from_numpy(np.ascontiguousarray(x).view(np.int32)).pin_memory() for x in gr)
            for gr in graphs]
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    with W.Server(slots=2, max_batch=4) as srv:
        with torch.cuda.stream(side):
            torch.cuda._sleep(100_000_000)
            for d, h in zip(dev, host):
                for dst, src in zip(d[:3], h):
                    dst.copy_(src.view(dst.shape), non_blocking=True)
        tickets = [srv.submit(gr[2].shape[0], d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr(),
                              stream=side.cuda_stream) for gr, d in zip(graphs, dev)]
        res = [srv.wait(t) for t in tickets]
    torch.cuda.synchronize()
    for gr, d, r in zip(graphs, dev, res):
        w = orc.bk_solve(*gr)
        assert r.status == 0 and r.raw_flow == w.raw_flow
        assert np.array_equal(d[3].cpu().numpy()[:gr[2].shape[0]], w.partition)


def test_server_admission_by_footprint(orc, torch_cuda):
    """One session whose arena (the HBM budget) holds one 3000-node request at a
    time: the requests are solved one after another, each alone, with the same
    results; a request larger than the arena is refused at submit."""
    graphs = [S.random_graph(3000, 950 + i, cap_max=40) for i in range(4)]
    fp = W.lib().wtv_request_footprint(3000)
    assert fp == W.lib().wtv_request_footprint(3072) and fp < W.lib().wtv_request_footprint(3073)
    dev = [_dev_graph(torch_cuda, *g) for g in graphs]
    torch_cuda.cuda.synchronize()
    with W.Server(slots=1, max_batch=8, hbm_budget_bytes=int(fp * 1.5)) as srv:
        tickets = [srv.submit(3000, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr())
                   for d in dev]
        res = [srv.wait(t) for t in tickets]
        assert all(r.batch_size == 1 for r in res)
        spans = sorted((r.start_s, r.end_s) for r in res)
        assert all(spans[i][1] <= spans[i + 1][0] + 1e-6 for i in range(len(spans) - 1)), "budget exceeded"
        with pytest.raises(W.WtvError, match="HBM budget"):
            srv.submit(10000, dev[0][0].data_ptr(), dev[0][1].data_ptr(), dev[0][2].data_ptr(), dev[0][3].data_ptr())
    for g, d, r in zip(graphs, dev, res):
        w = orc.bk_solve(*g)
        assert r.status == 0 and r.raw_flow == w.raw_flow
        assert np.array_equal(d[3].cpu().numpy()[:3000], w.partition)


def test_server_joins_requests_into_a_running_solve(orc, torch_cuda):
    """Continuous batching: with one session, a 20M-tet graph is being solved
    when small requests arrive; they join it on the device at its next global
    relabel and, each graph leaving at the first relabel after its last active
    vertex, finish before it.  The small results equal BK's, the large one the
    same graph solved alone."""
    import time
    big = S.batch_workload(1, 20260011, 20_000_000, 20_000_001, device="cuda", lam_range=(10.0, 10.0))[0]
    small = [S.random_graph(int(400 + 60 * i), 980 + i, cap_max=40) for i in range(6)]
    _, _, k = _kuhn_soft(12, 13)
    small.append((k.nbr, k.res, k.term))
    n_big = big["N"]
    dbig = (big["nbr"], big["res"].clone(), big["term"].clone(), torch_cuda.empty(n_big, dtype=torch_cuda.uint8,
                                                                                   device="cuda"))
    alone_part = torch_cuda.empty(n_big, dtype=torch_cuda.uint8, device="cuda")
    alone = W.pr_solve(n_big, big["nbr"].data_ptr(), big["res"].data_ptr(), big["term"].data_ptr(),
                       alone_part.data_ptr(), 1, W.params(global_relabel_period=8),
                       torch_cuda.cuda.current_stream().cuda_stream)
    dsmall = [_dev_graph(torch_cuda, *g) for g in small]
    torch_cuda.cuda.synchronize()
    # a global relabel every macro round: many admission and exit points
    with W.Server(slots=1, max_batch=16, global_relabel_period=8) as srv:
        tb = srv.submit(n_big, dbig[0].data_ptr(), dbig[1].data_ptr(), dbig[2].data_ptr(), dbig[3].data_ptr())
        time.sleep(0.002)  # the large solve is under way (it takes tens of ms)
        ts = [srv.submit(g[2].shape[0], d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr())
              for g, d in zip(small, dsmall)]
        rs = [srv.wait(t) for t in ts]
        rb = srv.wait(tb)
    print(f"large {rb.start_s:.4f}-{rb.end_s:.4f} small",
          [(r.batch_size, round(r.start_s, 4), round(r.end_s, 4)) for r in rs])
    assert rb.status == 0 and all(r.status == 0 for r in rs)
    assert rs[0].submit_s < rb.end_s, "the large solve ended before the small requests arrived"
    assert all(r.batch_size >= 2 for r in rs), "a small request did not join the running solve"
    assert all(r.end_s < rb.end_s for r in rs), "a small graph did not leave before the large one"
    for g, d, r in zip(small, dsmall, rs):
        w = orc.bk_solve(*g)
        assert r.raw_flow == w.raw_flow
        assert np.array_equal(d[3].cpu().numpy()[:g[2].shape[0]], w.partition)
    assert rb.raw_flow == alone.raw_flow
    assert torch_cuda.equal(dbig[3], alone_part)


# ------------------------------------------------------------ device graph build
@pytest.mark.parametrize("seed", range(8))
def test_device_build_matches_host_formula(orc, torch_cuda, seed):
    """wtv_build_reduced_graph == the oracle's restatement of wtivo_gpupr_bindings.cpp:139-318,
    value by value (FP64 capacities with heavy jitter and several coordinate scales)."""
    rng = np.random.default_rng(seed)
    k = 5 + 3 * seed
    scale = [0.01, 0.3, 1.0, 2.5][seed % 4]
    m = S.kuhn_mesh(k, jitter=0.45, seed=seed, hi=(scale, scale * 0.7, scale * 1.3))
    labels = rng.choice(np.array([0, 1, 2], np.uint8), size=m.N, p=[0.3, 0.6, 0.1])
    fill_T = float(rng.uniform(0.0, 40.0))
    want = orc.reduce(m.V, m.tets, m.nbrs, labels, fill_T)
    T = torch_cuda.from_numpy(m.tets).cuda()
    Nb = torch_cuda.from_numpy(m.nbrs).cuda()
    gn = torch_cuda.empty((m.N, 4), dtype=torch_cuda.int32, device="cuda")
    gr = torch_cuda.empty((m.N, 4), dtype=torch_cuda.int32, device="cuda")
    gt = torch_cuda.empty(m.N, dtype=torch_cuda.int32, device="cuda")
    fo = torch_cuda.empty(m.N, dtype=torch_cuda.int32, device="cuda")
    info = W.WtvBuildInfo()
    W.check(W.lib().wtv_build_reduced_graph(m.V.ctypes.data, m.V.shape[0], T.data_ptr(), Nb.data_ptr(), m.N,
                                            labels.ctypes.data, fill_T, gn.data_ptr(), gr.data_ptr(), gt.data_ptr(),
                                            fo.data_ptr(), C.byref(info), torch_cuda.cuda.current_stream().cuda_stream))
    n = want.n
    assert info.free_n == n
    for key in ("fixed_source", "fixed_sink", "full_refs", "free_refs", "terminal_edges", "base_flow"):
        assert getattr(info, key) == want.info[key], key
    assert np.array_equal(fo.cpu().numpy()[:n].view(np.uint32), want.free_old)
    assert np.array_equal(gn.cpu().numpy()[:n], want.nbr)
    assert np.array_equal(gr.cpu().numpy()[:n].view(np.uint32), want.res)
    assert np.array_equal(gt.cpu().numpy()[:n], want.term)


def test_graph_cut_fast_device_build_large(orc, torch_cuda):
    """A 1.57M-tet reference-model cut, full pipeline on the device, against the oracle."""
    import wtivo_gpupr as G
    k = 64
    m = S.kuhn_mesh(k, seed=31)
    shape = S.sphere_with_holes(31)
    lab = S.band_labels(S.tet_centroids_sdf(m, shape), 1.0 / k)
    r = orc.reduce(m.V, m.tets, m.nbrs, lab, 10.0)
    sol = orc.bk_solve(r.nbr, r.res, r.term)
    want = orc.scatter(lab, r.free_old, sol.partition)
    out = G.graph_cut_fast(m.V, torch_cuda.from_numpy(m.tets).cuda(), torch_cuda.from_numpy(m.nbrs).cuda(), lab, 10.0)
    assert out[1] == sol.raw_flow + r.info["base_flow"] and out[2] == sol.raw_flow
    assert np.array_equal(out[0], want)


def _reference_model_cut(orc, torch, V, tets, nbrs, lab):
    """graph_cut_fast through the pybind module and the device build through
    the C-ABI, both against the oracle's build + BK + scatter."""
    import ctypes as C
    import wtivo_gpupr as G
    r = orc.reduce(V, tets, nbrs, lab, 10.0)
    sol = orc.bk_solve(r.nbr, r.res, r.term)
    want = orc.scatter(lab, r.free_old, sol.partition)
    dT = torch.from_numpy(np.ascontiguousarray(tets, np.int32)).cuda()
    dN = torch.from_numpy(np.ascontiguousarray(nbrs, np.int32)).cuda()
    out = G.graph_cut_fast(V, dT, dN, lab, 10.0)
    assert out[2] == sol.raw_flow and out[1] == sol.raw_flow + r.info["base_flow"]
    assert np.array_equal(out[0], want)
    N = tets.shape[0]
    rn = torch.empty((N, 4), dtype=torch.int32, device="cuda")
    rr = torch.empty_like(rn)
    rt = torch.empty(N, dtype=torch.int32, device="cuda")
    fo = torch.empty_like(rt)
    info = W.WtvBuildInfo()
    Vc = np.ascontiguousarray(V, np.float64)
    lc = np.ascontiguousarray(lab, np.uint8)
    W.check(W.lib().wtv_build_reduced_graph(Vc.ctypes.data, Vc.shape[0], dT.data_ptr(), dN.data_ptr(), N,
                                            lc.ctypes.data, 10.0, rn.data_ptr(), rr.data_ptr(), rt.data_ptr(),
                                            fo.data_ptr(), C.byref(info), torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    F = int(info.free_n)
    assert F == len(r.term)
    assert np.array_equal(rn[:F].cpu().numpy(), r.nbr)
    assert np.array_equal(rr[:F].cpu().numpy().view(np.uint32), r.res)
    assert np.array_equal(rt[:F].cpu().numpy(), r.term)
    assert np.array_equal(fo[:F].cpu().numpy().view(np.uint32), np.asarray(r.free_old, np.uint32))


def test_reference_model_cut_on_delaunay(orc, torch_cuda):
    """The reference capacity model on an irregular mesh: labels 0 within one
    voxel of the genus-3 surface of a Delaunay point cloud."""
    P, shape, res = S.genus3_point_cloud(40_000, 81)
    m = S.delaunay_mesh(P)
    lab = S.band_labels(S.tet_centroids_sdf(m, shape), 1.0 / res)
    _reference_model_cut(orc, torch_cuda, m.V, m.tets, m.nbrs, lab)


def test_reference_model_cut_kuhn96(orc, torch_cuda):
    """The reference capacity model at 5.3M tets (the bench's drop-in leg at a
    size BK finishes in seconds)."""
    k = 96
    m = S.kuhn_mesh(k, seed=20260001)
    shape = S.genus3(20260001)
    lab = S.band_labels(S.tet_centroids_sdf(m, shape), 1.0 / k)
    _reference_model_cut(orc, torch_cuda, m.V, m.tets, m.nbrs, lab)


def test_kuhn150_integer_stress_full_size(orc, torch_cuda):
    """20,250,000-tet (150^3) thin shell with 40 tunnels, lambda 100, log-normal(2)
    capacities quantised to [1, 2^20]: labels and flow bit-exact against BK."""
    k = 150
    V, T, Nb = S.kuhn_mesh_t(k, seed=20260004, device="cuda")
    shape = S.thin_shell_with_tunnels(20260004, 1.0 / k, tunnels=40)
    lab = S.labels_from_sdf(S.tet_sdf_t(V, T, shape).cpu().numpy(), 0.0)
    m = S.TetMesh(k, V.cpu().numpy(), T.cpu().numpy(), Nb.cpu().numpy())
    del V, T, Nb
    g = S.lognormal_int_graph(m, lab, 100.0, seed=20260005, rev_slots=S.KUHN_REV_SLOT)
    assert m.N == 20_250_000 and g.res.max() <= 2 ** 20
    part, st, r_out, t_out = gpu_solve(torch_cuda, g.nbr, g.res, g.term, device_partition=True)
    b = orc.bk_solve(g.nbr, g.res, g.term)
    assert st["raw_flow"] == b.raw_flow
    assert np.array_equal(part, b.partition)
    assert S.quantised_cut(g.nbr, g.res, g.term, part) == st["raw_flow"]


# ------------------------------------------------------------ surface extraction
@pytest.mark.parametrize("seed", range(4))
def test_surface_extraction_matches_oracle(orc, torch_cuda, seed):
    import wtivo_gpupr as G
    k = [8, 20, 33, 48][seed]
    m = S.kuhn_mesh(k, seed=40 + seed)
    shape = S.sphere_with_holes(40 + seed)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.03, seed=50 + seed)
    if seed == 3:
        lab[lab == 1] = 2  # any label != 0 counts as outside
    want_v, want_f = orc.surface(lab, m.V, m.tets, m.nbrs)
    OV, OF = G.surface_extraction_topology_cuda(lab, m.V, torch_cuda.from_numpy(m.tets).cuda(),
                                                torch_cuda.from_numpy(m.nbrs).cuda(), 8)
    assert OF.dtype == np.int32 and OV.dtype == np.float64
    assert np.array_equal(OF, want_f)
    assert np.array_equal(OV, want_v)


def test_surface_after_graph_cut_and_errors(orc, torch_cuda):
    """The pipeline order of wtivo.py:842-859: cut, then surface of the new labels."""
    import wtivo_gpupr as G
    m, lab, _ = _kuhn_soft(24, 60)
    T = torch_cuda.from_numpy(m.tets).cuda()
    Nb = torch_cuda.from_numpy(m.nbrs).cuda()
    new_labels = G.graph_cut_fast(m.V, T, Nb, lab, 10.0)[0]
    OV, OF = G.surface_extraction_topology_cuda(new_labels, m.V, T, Nb)
    want_v, want_f = orc.surface(new_labels, m.V, m.tets, m.nbrs)
    assert np.array_equal(OF, want_f) and np.array_equal(OV, want_v)
    bad = m.nbrs.copy()
    i = int(np.flatnonzero(new_labels == 0)[0])
    bad[i, 0] = m.N + 7
    with pytest.raises(RuntimeError, match="surface neighbor out of range"):
        G.surface_extraction_topology_cuda(new_labels, m.V, T, torch_cuda.from_numpy(bad).cuda())
    with pytest.raises(RuntimeError, match="contiguous CUDA int32 Nx4 with same N"):
        G.surface_extraction_topology_cuda(new_labels, m.V, T.long(), Nb)


def test_graph_replay_equals_plain_launches(orc, torch_cuda):
    """Macro rounds replayed from a captured hipGraph (default) give the same
    answer as kernel-by-kernel launches, over many replays, local_steps that
    start rounds at every list phase (position % 6), and fused or two-kernel
    sub-rounds."""
    _, _, g = _kuhn_soft(24, 70)
    want = orc.bk_solve(g.nbr, g.res, g.term)
    for fused in (0, -1):
        for ls in (1, 3, 4, 8):
            for no_graph in (0, 1):
                gpu = gpu_solve(torch_cuda, g.nbr, g.res, g.term, local_steps=ls, no_graph=no_graph, fused=fused,
                                tail_cap=-1)
                assert_same(orc, g.nbr, g.res, g.term, gpu, want.raw_flow, want.partition)
                assert gpu[1]["rounds"] >= 1


def test_profile_mode_counts_kernels(torch_cuda):
    _, _, g = _kuhn_soft(16, 71)
    part, st, _, _ = gpu_solve(torch_cuda, g.nbr, g.res, g.term, profile=1)
    assert st["sub_rounds"] > 0 and st["n_level"] > 0 and st["t_level_ms"] > 0
    assert st["n_push"] + st["tail_sub_rounds"] >= st["sub_rounds"]
    # fused sub-rounds are one kernel (counted as a push), pull-mode ones two
    assert st["fused_rounds"] > 0 and st["n_relabel"] == 0 and (st["n_tail"] > 0) == (st["tail_sub_rounds"] > 0)
    _, st4, _, _ = gpu_solve(torch_cuda, g.nbr, g.res, g.term, profile=1, fused=-1)
    assert st4["fused_rounds"] == 0 and st4["n_push"] == st4["n_relabel"] > 0
    _, st2, _, _ = gpu_solve(torch_cuda, g.nbr, g.res, g.term, profile=1, tail_cap=-1)
    assert st2["n_tail"] == 0 and st2["n_push"] >= st2["sub_rounds"] and st2["t_push_ms"] > 0
    # the event-pair cost is measured in profile mode only, and is a part of
    # the event-timed launches it is subtracted from
    assert 0.0 <= st2["t_event_overhead_us"] < 1e3 * st2["t_push_ms"] / st2["n_push"]
    _, st3, _, _ = gpu_solve(torch_cuda, g.nbr, g.res, g.term)
    assert st3["t_event_overhead_us"] == 0.0 and st3["n_push"] == 0


@pytest.mark.parametrize("fused", [0, -1])
@pytest.mark.parametrize("pull_min", [1, 64, -1])
def test_pull_mode_same_answer(orc, torch_cuda, pull_min, fused):
    """Pull-mode sub-rounds (plain stores and gathers instead of atomics; by
    default only for lists of millions of vertices, here forced down to every
    list of more than pull_min vertices) give BK's flow and labels and a valid
    maximum preflow, alone, mixed with fused or two-kernel atomic sub-rounds
    and with the one-workgroup tail, in single, batched and served solves."""
    _, _, g = _kuhn_soft(24, 81)
    want = orc.bk_solve(g.nbr, g.res, g.term)
    for tail_cap in (-1, 0):
        gpu = gpu_solve(torch_cuda, g.nbr, g.res, g.term, pull_min=pull_min, tail_cap=tail_cap, fused=fused)
        assert_same(orc, g.nbr, g.res, g.term, gpu, want.raw_flow, want.partition)
        assert (gpu[1]["pull_rounds"] > 0) == (pull_min > 0)
        if fused < 0:
            assert gpu[1]["fused_rounds"] == 0
    # the reference's source-height quirk and random graphs
    for seed in range(6):
        nbr, res, term = S.random_graph(int(40 + 60 * seed), 1300 + seed, cap_max=int(3 + 9 * seed))
        a = orc.pr_solve(nbr, res, term)
        if orc.certificate(nbr, res, term, a.res, a.term, a.partition, a.raw_flow) != 0:
            continue
        gpu = gpu_solve(torch_cuda, nbr, res, term, pull_min=pull_min, tail_cap=-1, fused=fused)
        assert_same(orc, nbr, res, term, gpu, a.raw_flow, a.partition)
    graphs = [S.random_graph(int(200 + 50 * i), 1400 + i, cap_max=30) for i in range(5)]
    graphs.append((g.nbr, g.res, g.term))
    off, flows, part, _, _, st = gpu_batch(torch_cuda, graphs, pull_min=pull_min, tail_cap=-1, fused=fused)
    assert (st.pull_rounds > 0) == (pull_min > 0)
    for i, gr in enumerate(graphs):
        b = orc.bk_solve(*gr)
        assert flows[i] == b.raw_flow
        assert np.array_equal(part[int(off[i]):int(off[i + 1])], b.partition)
    wantb = [orc.bk_solve(*gr) for gr in graphs]
    torch_cuda.cuda.synchronize()
    dev = [_dev_graph(torch_cuda, *gr) for gr in graphs]
    torch_cuda.cuda.synchronize()
    with W.Server(slots=1, max_batch=8, pull_min=pull_min, tail_cap=-1, fused=fused) as srv:
        tickets = [srv.submit(gr[2].shape[0], d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr())
                   for gr, d in zip(graphs, dev)]
        res = [srv.wait(t) for t in tickets]
    for gr, d, r, w in zip(graphs, dev, res, wantb):
        assert r.status == 0 and r.raw_flow == w.raw_flow
        assert np.array_equal(d[3].cpu().numpy()[:gr[2].shape[0]], w.partition)
        assert orc.certificate(gr[0], gr[1], gr[2], d[1].cpu().numpy().view(np.uint32), d[2].cpu().numpy(),
                               w.partition, r.raw_flow) == 0


@pytest.mark.parametrize("tail_cap", [-1, 16, 2048, 1 << 20])
def test_discharge_tail_mode_same_answer(orc, torch_cuda, tail_cap):
    """The one-workgroup discharge (small active lists) and the multi-workgroup
    rounds compute the same flow and labels, whichever handles which phase."""
    _, _, g = _kuhn_soft(24, 80)
    want = orc.bk_solve(g.nbr, g.res, g.term)
    gpu = gpu_solve(torch_cuda, g.nbr, g.res, g.term, tail_cap=tail_cap)
    assert_same(orc, g.nbr, g.res, g.term, gpu, want.raw_flow, want.partition)
    graphs = [S.random_graph(int(200 + 50 * i), 1200 + i, cap_max=30) for i in range(5)]
    off, flows, part, _, _, _ = gpu_batch(torch_cuda, graphs, tail_cap=tail_cap)
    for i, gr in enumerate(graphs):
        b = orc.bk_solve(*gr)
        assert flows[i] == b.raw_flow
        assert np.array_equal(part[int(off[i]):int(off[i + 1])], b.partition)
