"""Pins the CPU oracle (oracle/wtv_oracle.c) before it is trusted as the checker.

* capacity_exact_formula: hand-computed known answers, range errors, degenerate faces
* reduce: a plain-Python brute force of the classification/folding rules on tiny
  meshes, plus the invariants of wtivo_gpupr_bindings.cpp:139-318
* max-flow: scipy.sparse.csgraph.maximum_flow (independent implementation),
  agreement of the restated push-relabel with Boykov-Kolmogorov, certificate
* labels: exhaustive enumeration of every cut on graphs with <= 12 nodes: the
  labels are the smallest sink side among all minimum cuts
* a planted mistake in a copy of a solution is caught by the certificate
* golden fixtures under tests/golden reproduce
"""
import itertools
import os

import numpy as np
import pytest

import wtv_synth as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------ capacity
def _unit_tet(scale=1e-3):
    V = np.array([[0, 0, 0], [scale, 0, 0], [0, scale, 0], [0, 0, scale]], np.float64)
    return V, np.array([0, 1, 2, 3], np.int32)


def test_capacity_known_answers(orc):
    V, t = _unit_tet()
    # face 012: right triangle with legs 1e-3 -> 1e7 * 0.5 * 1e-6 = 5 exactly
    assert orc.capacity(V, t, 0, 0, 1, 10.0) == (5, 0)      # unlike labels: area
    assert orc.capacity(V, t, 0, 1, 1, 10.0) == (55, 0)     # like labels: area * (1 + 10)
    assert orc.capacity(V, t, 1, 1, 1, 0.0) == (5, 0)       # face 013, fill 0
    # face 123: equilateral, side sqrt(2)e-3 -> 0.5*sqrt(3)*10 = 8.66..
    assert orc.capacity(V, t, 3, 0, 1, 10.0) == (8, 0)
    assert orc.capacity(V, t, 3, 1, 1, 10.0) == (95, 0)     # 11 * fp32(8.660254) = 95.26
    # labels compare raw values: 1 vs 2 are "different"
    assert orc.capacity(V, t, 0, 1, 2, 10.0) == (5, 0)


def test_capacity_fp32_rounding_step(orc):
    # area 1e7*0.5*|c| with legs a=b -> 5e6*a^2; a^2 = 0.3e-6 -> 1.5 in fp32 then *(1+1) = 3.0
    a = np.sqrt(0.3e-6)
    V = np.array([[0, 0, 0], [a, 0, 0], [0, a, 0], [0, 0, 1.0]], np.float64)
    cap, st = orc.capacity(V, np.array([0, 1, 2, 3], np.int32), 0, 1, 1, 1.0)
    assert st == 0 and cap in (2, 3)
    want = int(float(np.float32(5e6 * (a * a))) * 2.0)
    assert cap == want


def test_capacity_range_and_errors(orc):
    V, t = _unit_tet()
    # 5 * (1 + f) just below / at 2^30
    f_ok = (2 ** 30 - 1) / 5.0 - 1.0
    cap, st = orc.capacity(V, t, 0, 1, 1, f_ok)
    assert st == 0 and 2 ** 30 - 2 <= cap < 2 ** 30
    cap, st = orc.capacity(V, t, 0, 1, 1, 2 ** 30 / 5.0)
    assert st == 2
    # degenerate (collinear) face has capacity 0
    Vd = np.array([[0, 0, 0], [1e-3, 0, 0], [2e-3, 0, 0], [0, 0, 1e-3]], np.float64)
    assert orc.capacity(Vd, t, 0, 1, 1, 10.0) == (0, 0)
    # invalid vertex index
    assert orc.capacity(V, np.array([0, 1, 7, 3], np.int32), 0, 0, 1, 1.0)[1] == 1


# ------------------------------------------------------------ reduce
def _py_reduce(V, tets, nbrs, labels, fill_T, cap_fn):
    N = tets.shape[0]
    cls = []
    for i in range(N):
        if labels[i] == 0:
            cls.append(-2)
        elif (nbrs[i] < 0).any():
            cls.append(-3)
        else:
            cls.append(-1)
    rid = {}
    for i in range(N):
        if cls[i] == -1:
            rid[i] = len(rid)
    n = len(rid)
    nbr = -np.ones((n, 4), np.int32)
    res = np.zeros((n, 4), np.uint32)
    term = np.zeros(n, np.int32)
    base = 0
    for i, r in rid.items():
        src = snk = 0
        for fi in range(4):
            j = int(nbrs[i, fi])
            c = cap_fn(i, fi, j)
            if cls[j] == -1:
                nbr[r, fi] = rid[j]
                res[r, fi] = c
            elif cls[j] == -2:
                src += c
            else:
                snk += c
        base += min(src, snk)
        term[r] = src - snk
    return n, nbr, res, term, base


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_reduce_matches_bruteforce(orc, seed):
    m = S.kuhn_mesh(5, seed=seed)
    rng = np.random.default_rng(seed)
    labels = (rng.uniform(size=m.N) < 0.6).astype(np.uint8) * rng.integers(1, 3, size=m.N).astype(np.uint8)
    r = orc.reduce(m.V, m.tets, m.nbrs, labels, 10.0)

    def cap_fn(i, fi, j):
        return orc.capacity(m.V, m.tets[i], fi, labels[i], labels[j], 10.0)[0]

    n, nbr, res, term, base = _py_reduce(m.V, m.tets, m.nbrs, labels, 10.0, cap_fn)
    assert r.n == n
    assert np.array_equal(r.nbr, nbr) and np.array_equal(r.res, res) and np.array_equal(r.term, term)
    assert r.info["base_flow"] == base
    inf = r.info
    assert inf["free_n"] + inf["fixed_source"] + inf["fixed_sink"] == m.N
    assert inf["full_refs"] % 2 == 0 and inf["free_refs"] % 2 == 0
    assert inf["terminal_edges"] + inf["free_refs"] == 4 * n
    assert np.all(np.diff(r.free_old.astype(np.int64)) > 0)  # reduced ids in tet order


def test_reduce_errors(orc):
    m = S.kuhn_mesh(3, seed=0)
    lab = np.ones(m.N, np.uint8)
    bad = m.nbrs.copy()
    bad[5, 0] = m.N + 3
    with pytest.raises(orc.OracleError) as e:
        orc.reduce(m.V, m.tets, bad, lab, 10.0)
    assert e.value.code == orc.EINVAL
    odd = m.nbrs.copy()
    i = int(np.flatnonzero((odd >= 0).all(1))[0])
    odd[i, 0] = -1  # one-sided removal -> odd reference count
    with pytest.raises(orc.OracleError):
        orc.reduce(m.V, m.tets, odd, lab, 10.0)
    with pytest.raises(orc.OracleError) as e:
        orc.reduce(m.V, m.tets, m.nbrs, lab, 1e12)  # capacities overflow 2^30
    assert e.value.code == orc.EOVERFLOW


# ------------------------------------------------------------ max-flow
def scipy_flow(nbr, res, term):
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import maximum_flow
    n = term.shape[0]
    s, t = n, n + 1
    rows, cols, caps = [], [], []
    for v in range(n):
        for k in range(4):
            u = int(nbr[v, k])
            if u >= 0:
                rows.append(v); cols.append(u); caps.append(int(res[v, k]))
        if term[v] > 0:
            rows.append(s); cols.append(v); caps.append(int(term[v]))
        if term[v] < 0:
            rows.append(v); cols.append(t); caps.append(int(-term[v]))
    M = csr_matrix((np.array(caps, np.int32), (rows, cols)), shape=(n + 2, n + 2))
    return int(maximum_flow(M, s, t).flow_value)


@pytest.mark.parametrize("seed", range(60))
def test_maxflow_vs_scipy_and_bk(orc, seed):
    n = int(np.random.default_rng(seed).integers(6, 120))
    nbr, res, term = S.random_graph(n, seed, cap_max=int(3 + seed % 40))
    f = scipy_flow(nbr, res, term)
    b = orc.bk_solve(nbr, res, term)
    a = orc.pr_solve(nbr, res, term, period=(seed % 4) * 5, local_steps=1 + seed % 9)
    assert b.raw_flow == f
    assert orc.certificate(nbr, res, term, b.res, b.term, b.partition, b.raw_flow) == 0
    assert a.raw_flow == f
    assert orc.certificate(nbr, res, term, a.res, a.term, a.partition, a.raw_flow) == 0
    assert np.array_equal(a.partition, b.partition)


def _brute_min_cut(nbr, res, term):
    n = term.shape[0]
    best, best_T = None, None
    for mask in range(1 << n):
        T = np.array([(mask >> i) & 1 for i in range(n)], bool)
        e = 0
        for v in range(n):
            if not T[v] and term[v] < 0:
                e += -int(term[v])
            if T[v] and term[v] > 0:
                e += int(term[v])
            if not T[v]:
                for k in range(4):
                    u = nbr[v, k]
                    if u >= 0 and T[u]:
                        e += int(res[v, k])
        if best is None or e < best or (e == best and T.sum() < best_T.sum()):
            best, best_T = e, T
    return best, best_T


@pytest.mark.parametrize("seed", range(25))
def test_labels_are_minimal_sink_side_bruteforce(orc, seed):
    n = int(np.random.default_rng(100 + seed).integers(3, 12))
    nbr, res, term = S.random_graph(n, 100 + seed, edges_per_node=1.8, cap_max=9, term_max=9)
    best, T = _brute_min_cut(nbr, res, term)
    b = orc.bk_solve(nbr, res, term)
    assert b.raw_flow == best
    assert np.array_equal(b.partition.astype(bool), T)
    a = orc.pr_solve(nbr, res, term, period=0)
    code = orc.certificate(nbr, res, term, a.res, a.term, a.partition, a.raw_flow)
    if code == 0:
        assert a.raw_flow == best and np.array_equal(a.partition.astype(bool), T)
    else:
        # the reference's source-height-n convention (see the quirk test) leaves
        # excess on a vertex at distance exactly n: sink side with excess
        assert code == 6 and a.raw_flow < best


def test_reference_source_height_quirk(orc):
    """gpu_push_relabel_fast.cu:426,161,218: source_h = n. With two free nodes on a
    path v0 -> v1 -> sink, v0 sits at distance 2 = n and is never active."""
    nbr = np.array([[1, -1, -1, -1], [0, -1, -1, -1]], np.int32)
    res = np.array([[5, 0, 0, 0], [0, 0, 0, 0]], np.uint32)
    term = np.array([5, -5], np.int32)
    a = orc.pr_solve(nbr, res, term)
    assert a.raw_flow == 0
    assert a.partition.tolist() == [1, 1]
    assert scipy_flow(nbr, res, term) == 5  # the true max flow


def test_certificate_catches_planted_mistakes(orc):
    nbr, res, term = S.random_graph(60, 7, cap_max=15)
    b = orc.bk_solve(nbr, res, term)
    assert orc.certificate(nbr, res, term, b.res, b.term, b.partition, b.raw_flow) == 0
    assert orc.certificate(nbr, res, term, b.res, b.term, b.partition, b.raw_flow + 1) != 0
    p2 = b.partition.copy()
    p2[np.argmax(p2)] ^= 1
    assert orc.certificate(nbr, res, term, b.res, b.term, p2, b.raw_flow) != 0
    r2 = b.res.copy()
    e = np.argwhere((nbr >= 0) & (r2 > 0))[0]
    r2[e[0], e[1]] -= 1
    assert orc.certificate(nbr, res, term, r2, b.term, b.partition, b.raw_flow) != 0


def scipy_flow_state(nbr, res, term):
    """scipy's maximum flow in the solver's representation: residuals after the
    net arc flows, source terminals = flow still sent, sink terminals used."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import maximum_flow
    n = term.shape[0]
    s, t = n, n + 1
    rows, cols, caps = [], [], []
    for v in range(n):
        for k in range(4):
            u = int(nbr[v, k])
            if u >= 0:
                rows.append(v); cols.append(u); caps.append(int(res[v, k]))
        if term[v] > 0:
            rows.append(s); cols.append(v); caps.append(int(term[v]))
        if term[v] < 0:
            rows.append(v); cols.append(t); caps.append(int(-term[v]))
    M = csr_matrix((np.array(caps, np.int32), (rows, cols)), shape=(n + 2, n + 2))
    r = maximum_flow(M, s, t)
    F = r.flow.toarray().astype(np.int64)
    res_f = res.astype(np.int64).copy()
    for v in range(n):
        for k in range(4):
            u = int(nbr[v, k])
            if u >= 0:
                res_f[v, k] -= F[v, u]
    term_f = term.astype(np.int64).copy()
    term_f[term > 0] = F[s, np.nonzero(term > 0)[0]]
    neg = np.nonzero(term < 0)[0]
    term_f[neg] = term[neg] + F[neg, t]
    assert (res_f >= 0).all()
    return int(r.flow_value), res_f.astype(np.uint32), term_f.astype(np.int32)


def flow_excess(nbr, res0, term0, res, term):
    """Per-vertex imbalance of a preflow state (the excess a flow would not have)."""
    src = np.where(term0 > 0, term.astype(np.int64), 0)
    sink = np.where(term0 <= 0, term.astype(np.int64) - term0, 0)
    out = (res0.astype(np.int64) - res.astype(np.int64)).sum(axis=1)
    return src - sink - out


@pytest.mark.parametrize("seed", range(12))
def test_flow_certificate_pinned_by_scipy_flows(orc, seed):
    n = int(np.random.default_rng(seed + 100).integers(6, 100))
    nbr, res, term = S.random_graph(n, seed + 100, cap_max=int(3 + seed * 3))
    val, res_f, term_f = scipy_flow_state(nbr, res, term)
    b = orc.bk_solve(nbr, res, term)
    assert b.raw_flow == val
    assert (flow_excess(nbr, res, term, res_f, term_f) == 0).all()
    assert orc.flow_certificate(nbr, res, term, res_f, term_f, b.partition, val) == 0
    # a preflow with excess left on the source side is not a flow
    ex = flow_excess(nbr, res, term, b.res, b.term)
    if (ex != 0).any():
        assert orc.flow_certificate(nbr, res, term, b.res, b.term, b.partition, b.raw_flow) == 2
    assert orc.flow_certificate(nbr, res, term, res_f, term_f, b.partition, val + 1) != 0
    pos = np.nonzero(term > 0)[0]
    if len(pos):
        t2 = term_f.copy()
        t2[pos[0]] = term[pos[0]] + 1              # more source flow than the arc holds
        assert orc.flow_certificate(nbr, res, term, res_f, t2, b.partition, val) == 3


def test_soft_graph_kuhn32_bk_equals_pr(orc):
    m = S.kuhn_mesh(16, seed=3)
    shape = S.sphere_with_holes(3)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.02, seed=4)
    g = S.soft_graph(m, lab, shape, 10.0, 1.0 / 16)
    a = orc.pr_solve(g.nbr, g.res, g.term)
    b = orc.bk_solve(g.nbr, g.res, g.term)
    assert a.raw_flow == b.raw_flow == S.quantised_cut(g.nbr, g.res, g.term, a.partition)
    assert np.array_equal(a.partition, b.partition)


# ------------------------------------------------------------ golden
def _golden_files():
    return sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))


@pytest.mark.parametrize("name", _golden_files())
def test_golden_fixture_reproduces(orc, name):
    z = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    if "V" in z.files:
        r = orc.reduce(z["V"], z["tets"], z["nbrs"], z["labels"], float(z["fill_T"]))
        sol = orc.pr_solve(r.nbr, r.res, r.term)
        out = orc.scatter(z["labels"], r.free_old, sol.partition)
        assert np.array_equal(out, z["out_labels"])
        assert sol.raw_flow == int(z["raw_flow"])
        assert sol.raw_flow + r.info["base_flow"] == int(z["adjusted_flow"])
    else:
        sol = orc.bk_solve(z["nbr"], z["res"], z["term"])
        assert sol.raw_flow == int(z["raw_flow"])
        assert np.array_equal(sol.partition, z["partition"])


# ------------------------------------------------------------ label scatter
def test_scatter_known_answer(orc):
    """Label reconstruction (wtivo_gpupr_bindings.cpp:357-364): out starts as a copy
    of the input labels, then every free tet free_old[q] takes partition[q] != 0.
    Labels 0, 1 and 7 are mixed; the expected array is written out by hand."""
    labels = np.array([7, 0, 1, 7, 0, 1, 7, 1], np.uint8)
    free_old = np.array([0, 3, 5, 7], np.uint32)   # reduced id q -> tet index
    partition = np.array([0, 1, 0, 5], np.uint8)    # any nonzero byte means "reaches the sink"
    want = np.array([0, 0, 1, 1, 0, 0, 7, 1], np.uint8)
    out = orc.scatter(labels, free_old, partition)
    assert out.tolist() == want.tolist()
    # the fixed tets (label 0 sources; label 1/7 sinks) keep their input label
    fixed = np.setdiff1d(np.arange(8), free_old)
    assert np.array_equal(out[fixed], labels[fixed])
    # plausible mistakes produce a different array, so the literal discriminates
    by_reduced_id = labels.copy()
    by_reduced_id[:4] = partition != 0                      # indexed by q, not free_old[q]
    raw_bytes = labels.copy()
    raw_bytes[free_old] = partition                         # no normalisation to 0/1
    from_zero = np.zeros(8, np.uint8)
    from_zero[free_old] = partition != 0                    # fixed tets not copied from the input
    for wrong in (by_reduced_id, raw_bytes, from_zero):
        assert wrong.tolist() != want.tolist()


def test_scatter_through_reduce_known_answer(orc):
    """Classify -> reduced ids -> scatter on a 3^3 Kuhn mesh (162 tets, tet id =
    cell * 6 + perm).  Every tet has label 0 except the six of the centre cell
    (cell 13: tets 78..83, labels 7 1 7 1 7 1), which have four neighbours each and
    so are the free tets, in increasing tet order (wtivo_gpupr_bindings.cpp:157-188).
    A hand-picked partition then lands on tets 78, 80 and 81 only
    (wtivo_gpupr_bindings.cpp:357-364)."""
    m = S.kuhn_mesh(3, jitter=0.1, seed=5)
    lab = np.zeros(m.N, np.uint8)
    lab[78:84] = [7, 1, 7, 1, 7, 1]
    r = orc.reduce(m.V, m.tets, m.nbrs, lab, 10.0)
    assert r.free_old.tolist() == [78, 79, 80, 81, 82, 83]
    assert r.info["fixed_source"] == 156 and r.info["fixed_sink"] == 0
    out = orc.scatter(lab, r.free_old, np.array([1, 0, 1, 1, 0, 0], np.uint8))
    want = np.zeros(162, np.uint8)
    want[[78, 80, 81]] = 1
    assert out.tolist() == want.tolist()


# ------------------------------------------------------------ surface extraction
def test_surface_single_tet_outward(orc):
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    OV, OF = orc.surface(np.zeros(1, np.uint8), V, np.array([[0, 1, 2, 3]], np.int32), -np.ones((1, 4), np.int32))
    assert np.array_equal(OV, V)
    # every face normal points away from the opposite vertex
    cen = V.mean(0)
    for f in OF:
        a, b, c = V[f]
        nrm = np.cross(b - a, c - a)
        assert np.dot(nrm, a - cen) > 0
    assert sorted(tuple(sorted(f)) for f in OF.tolist()) == [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
    # a label != 0 tet contributes nothing
    OV, OF = orc.surface(np.ones(1, np.uint8), V, np.array([[0, 1, 2, 3]], np.int32), -np.ones((1, 4), np.int32))
    assert OF.shape == (0, 3) and OV.shape == (0, 3)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_surface_is_closed_and_encloses_the_region(orc, seed):
    """Interior blob: each directed edge appears once and its reverse once (closed,
    consistently oriented) and the enclosed volume equals the region's tet volume."""
    m = S.kuhn_mesh(10, jitter=0.1, seed=seed)
    shape = S.Sphere((0.5, 0.5, 0.5), 0.3 + 0.05 * seed)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.0)
    OV, OF = orc.surface(lab, m.V, m.tets, m.nbrs)
    edges = {}
    for f in OF.tolist():
        for a, b in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            edges[(a, b)] = edges.get((a, b), 0) + 1
    assert all(c == 1 for c in edges.values())
    assert all((b, a) in edges for (a, b) in edges)
    P = OV[OF]
    vol_surf = np.einsum("ij,ij->i", P[:, 0], np.cross(P[:, 1], P[:, 2])).sum() / 6.0
    T = m.V[m.tets[lab == 0]]
    vol_tets = np.abs(np.einsum("ij,ij->i", T[:, 1] - T[:, 0], np.cross(T[:, 2] - T[:, 0], T[:, 3] - T[:, 0]))).sum() / 6
    assert abs(vol_surf - vol_tets) < 1e-9 * max(1.0, vol_tets)
    # vertices compacted in ascending original id: map every output vertex back to
    # its original id through its (jittered, hence unique) coordinates
    row_of = {tuple(v): i for i, v in enumerate(m.V.tolist())}
    ids = np.array([row_of[tuple(v)] for v in OV.tolist()])
    lut = np.array([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]])
    on_surface = set()
    for i in np.flatnonzero(lab == 0):
        for fi in range(4):
            j = m.nbrs[i, fi]
            if j < 0 or lab[j] != 0:
                on_surface.update(m.tets[i, lut[fi]].tolist())
    assert ids.tolist() == sorted(on_surface)


def test_surface_known_answer(orc):
    """Three tets, written out by hand (wtivo_gpupr_bindings.cpp:400-546):
        T0 = [4,0,3,2] label 0, T1 = [2,3,4,5] label 7, T2 = [6,2,0,3] label 0,
    slot fi = local face {012,013,023,123}[fi]; T0 meets T1 across {2,3,4} (T0
    slot 2) and T2 across {0,2,3} (T0 slot 3, T2 slot 3).  Faces are taken from
    label-0 tets towards the hull or a label != 0 tet, tet-major then slot-minor;
    each face's ids are sorted and then oriented away from the tet's fourth vertex
    by the sign of the FP64 volume; the used vertices {0,2,3,4,6} are renumbered in
    ascending original id (vertices 1 and 5 are not used)."""
    V = np.array([[0, 0, 0], [9, 9, 9], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [0, 0, -1]], float)
    tets = np.array([[4, 0, 3, 2], [2, 3, 4, 5], [6, 2, 0, 3]], np.int32)
    nbrs = np.array([[-1, -1, 1, 2], [0, -1, -1, -1], [-1, -1, -1, 0]], np.int32)
    labels = np.array([0, 7, 0], np.uint8)
    OV, OF = orc.surface(labels, V, tets, nbrs)
    # original ids, worked by hand: T0 slots 0,1,2 then T2 slots 0,1,2
    raw = [(0, 4, 3), (0, 2, 4), (2, 3, 4), (0, 6, 2), (2, 6, 3), (0, 3, 6)]
    remap = {0: 0, 2: 1, 3: 2, 4: 3, 6: 4}
    want_f = [[remap[x] for x in f] for f in raw]
    assert want_f == [[0, 3, 2], [0, 1, 3], [1, 2, 3], [0, 4, 1], [1, 4, 2], [0, 2, 4]]
    assert OF.tolist() == want_f
    assert OV.tolist() == V[[0, 2, 3, 4, 6]].tolist()
    # plausible mistakes give a different answer
    slot_major = [want_f[i] for i in (0, 3, 1, 4, 2, 5)]
    first_use = {}
    for f in raw:
        for x in f:
            first_use.setdefault(x, len(first_use))
    first_use_f = [[first_use[x] for x in f] for f in raw]
    unoriented = [sorted(f) for f in want_f]
    for wrong in (slot_major, first_use_f, unoriented):
        assert wrong != want_f
    # each face normal points away from its tet's fourth vertex
    fourth = [2, 3, 0, 3, 0, 2]
    for f, q in zip(raw, fourth):
        a, b, c = V[list(f)]
        assert np.dot(np.cross(b - a, c - a), V[q] - a) < 0


def test_extreme_capacities_bk_equals_pr(orc):
    """Arc capacities up to 2^30 - 1 and terminals up to +-(2^31 - 1): both CPU
    solvers agree, and the flow passes the certificate (64-bit sums)."""
    nbr, res, term = S.random_graph(3000, 311, cap_max=(1 << 30) - 1, term_max=(1 << 31) - 1, p_term=0.9)
    b = orc.bk_solve(nbr, res, term)
    p = orc.pr_solve(nbr, res, term)
    assert b.raw_flow == p.raw_flow > (1 << 36)  # far beyond 32-bit sums
    assert np.array_equal(b.partition, p.partition)
    assert orc.certificate(nbr, res, term, p.res, p.term, p.partition, p.raw_flow) == 0


# ------------------------------------------------------------ tet neighbours
def _brute_neighbors(tets):
    """Plain-Python search: the other tet holding all three vertices of each face."""
    faces = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
    sets = [set(int(x) for x in t) for t in tets]
    out = np.full((len(tets), 4), -1, np.int32)
    for i, t in enumerate(tets):
        for fi, f in enumerate(faces):
            key = {int(t[k]) for k in f}
            hits = [j for j, s in enumerate(sets) if j != i and key <= s]
            assert len(hits) <= 1
            if hits:
                out[i, fi] = hits[0]
    return out


@pytest.mark.parametrize("k", [1, 2, 5, 9])
def test_tet_neighbors_closed_form_kuhn(orc, k):
    tets, nbrs = S.kuhn_topology(k)
    assert np.array_equal(orc.tet_neighbors(tets, (k + 1) ** 3), nbrs)


def test_tet_neighbors_brute_force_tiny(orc):
    rng = np.random.default_rng(7)
    tets, _ = S.kuhn_topology(2)
    assert np.array_equal(orc.tet_neighbors(tets), _brute_neighbors(tets))
    m = S.delaunay_mesh(rng.random((40, 3)), morton=False)
    assert np.array_equal(orc.tet_neighbors(m.tets), _brute_neighbors(m.tets))


@pytest.mark.parametrize("seed", [0, 1])
def test_tet_neighbors_match_qhull(orc, seed):
    # Qhull computes neighbours itself; delaunay_mesh maps them to the 4-slot layout
    rng = np.random.default_rng(seed)
    for pts in (rng.random((3000, 3)), S.genus3_point_cloud(20000, seed)[0]):
        m = S.delaunay_mesh(pts, morton=bool(seed))
        assert np.array_equal(orc.tet_neighbors(m.tets, pts.shape[0]), m.nbrs)


def test_tet_neighbors_equivariant_under_permutations(orc):
    tets, nbrs = S.kuhn_topology(6)
    for seed in range(3):
        t2, n2 = S.permute_topology(tets, nbrs, seed)
        assert np.array_equal(orc.tet_neighbors(t2), n2)


def test_tet_neighbors_hull_and_errors(orc):
    one = np.array([[0, 1, 2, 3]], np.int32)
    assert (orc.tet_neighbors(one) == -1).all()
    two = np.array([[0, 1, 2, 3], [4, 2, 1, 0]], np.int32)   # share face {0,1,2}
    n = orc.tet_neighbors(two)
    assert n.tolist() == [[1, -1, -1, -1], [-1, -1, -1, 0]]
    three = np.array([[0, 1, 2, 3], [0, 1, 2, 4], [0, 1, 2, 5]], np.int32)
    with pytest.raises(orc.OracleError):
        orc.tet_neighbors(three)
    with pytest.raises(orc.OracleError):
        orc.tet_neighbors(np.array([[0, 1, 1, 3]], np.int32))
    with pytest.raises(orc.OracleError):
        orc.tet_neighbors(np.array([[0, 1, 2, 9]], np.int32), nv=5)
    assert orc.tet_neighbors(np.zeros((0, 4), np.int32)).shape == (0, 4)


def test_tet_neighbors_planted_slot_mistake_is_caught(orc):
    # a restatement that used the face list {012,013,023,123} reversed (slot fi
    # = CGAL's neighbor(fi)) would disagree with the closed form on Kuhn meshes
    tets, nbrs = S.kuhn_topology(3)
    wrong = orc.tet_neighbors(tets)[:, ::-1]
    assert not np.array_equal(wrong, nbrs)


# ------------------------------------------------------------ tet labels from a sparse grid
def _tets_centred_at(pts, h=2.0 ** -12):
    """One tet per point whose FP64 centroid is exactly the (dyadic) point."""
    off = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0]], np.float64) * h
    V = (pts[:, None, :] + off[None]).reshape(-1, 3)
    return V, np.arange(4 * len(pts), dtype=np.int32).reshape(-1, 4)


def _block_leaves(n_leaf_cells, fn, lo_cell=0):
    """Leaves covering cells [lo_cell, lo_cell + n)^3 with voxel values fn(x, y, z) (index coords)."""
    c = np.arange(lo_cell, lo_cell + n_leaf_cells)
    cells = np.stack(np.meshgrid(c, c, c, indexing="ij"), -1).reshape(-1, 3)
    o = np.arange(8)
    off = np.stack(np.meshgrid(o, o, o, indexing="ij"), -1).reshape(512, 3)   # (x<<6)|(y<<3)|z
    idx = cells[:, None, :] * 8 + off[None]
    vals = fn(idx[..., 0], idx[..., 1], idx[..., 2]).astype(np.float32)
    return (cells * 8).astype(np.int32), vals


def _dense_trilinear(origins, values, bg, p):
    """Independent float64 trilinear over a dense copy of the grid (background outside leaves)."""
    lo = origins.min(0) - 8
    hi = origins.max(0) + 16
    D = np.full(tuple(hi - lo), bg, np.float64)
    o = np.arange(8)
    off = np.stack(np.meshgrid(o, o, o, indexing="ij"), -1).reshape(512, 3)
    for org, val in zip(origins, values):
        q = org - lo + off
        D[q[:, 0], q[:, 1], q[:, 2]] = val
    f = np.floor(p)
    u = p - f
    i = f.astype(np.int64) - lo
    out = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                wgt = (u[:, 0] if dx else 1 - u[:, 0]) * (u[:, 1] if dy else 1 - u[:, 1]) * \
                      (u[:, 2] if dz else 1 - u[:, 2])
                out += wgt * D[i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz]
    return out


UNIT = (np.zeros(3), np.ones(3))


def test_sample_labels_constant_field(orc):
    o, v = _block_leaves(3, lambda x, y, z: np.full(x.shape, 0.5))
    rng = np.random.default_rng(1)
    V, t = _tets_centred_at(rng.integers(-64, 400, (500, 3)) / 256.0)
    for thr, want in ((0.5, 0), (0.4999, 1)):
        lab, val = orc.sample_labels(o, v, 0.5, 1 / 64, V, t, *UNIT, 0.0, thr)
        assert (val == np.float32(0.5)).all() and (lab == want).all()


def test_sample_labels_linear_field_closed_form(orc):
    # trilinear interpolation reproduces x + 2y + 3z; at dyadic points every
    # float step is exact, so a swapped axis or leaf offset shows at once
    o, v = _block_leaves(4, lambda x, y, z: x + 2.0 * y + 3.0 * z)
    rng = np.random.default_rng(2)
    idx = rng.integers(0, 30 * 16, (2000, 3)) / 16.0          # index coords in [0, 30)
    V, t = _tets_centred_at(idx / 64.0)                       # voxel 1/64: world = index / 64
    lab, val = orc.sample_labels(o, v, 0.0, 1 / 64, V, t, *UNIT, 0.0, 40.0)
    want = idx[:, 0] + 2 * idx[:, 1] + 3 * idx[:, 2]
    assert np.array_equal(val.astype(np.float64), want)
    assert np.array_equal(lab, (want > 40.0).astype(np.uint8))


def test_sample_labels_voxel_centres_return_stored_values(orc):
    rng = np.random.default_rng(3)
    o, v = _block_leaves(2, lambda x, y, z: rng.random(x.shape))
    idx = rng.integers(0, 15, (1000, 3)).astype(np.float64)
    V, t = _tets_centred_at(idx / 64.0)
    lab, val = orc.sample_labels(o, v, 9.0, 1 / 64, V, t, *UNIT, 0.0, 0.5)
    flat = {tuple(oo): vv for oo, vv in zip(o, v)}
    want = np.array([flat[tuple((q // 8 * 8).astype(int))][int(q[0] % 8) * 64 + int(q[1] % 8) * 8 + int(q[2] % 8)]
                     for q in idx], np.float32)
    assert np.array_equal(val, want)
    assert np.array_equal(lab, (want > 0.5).astype(np.uint8))
    # threshold is inclusive: a sample equal to it is label 0
    lab2, _ = orc.sample_labels(o, v, 9.0, 1 / 64, V[:4], t[:1], *UNIT, 0.0, float(want[0]))
    assert lab2[0] == 0


def test_sample_labels_sparse_grid_matches_dense_float64(orc):
    rng = np.random.default_rng(4)
    o, v = _block_leaves(4, lambda x, y, z: rng.random(x.shape) * 3)
    keep = rng.random(len(o)) < 0.6                           # holes -> background stencils
    o, v = o[keep], v[keep]
    bmin, bmax, pad = np.array([-1.0, -2.0, 0.5]), np.array([3.0, 2.0, 2.5]), 0.05
    pts = bmin + rng.random((3000, 3)) * (bmax - bmin)
    V, t = _tets_centred_at(pts, h=1e-7)
    vs = 1 / 40.0
    lab, val = orc.sample_labels(o, v, 2.0, vs, V, t, bmin, bmax, pad, 1.0)
    c = V.reshape(-1, 4, 3).sum(1) * 0.25
    p = (((c - bmin) / (bmax - bmin)) * (1 - 2 * pad) + pad) * (1 / vs)
    ref = _dense_trilinear(o, v, 2.0, p)
    assert np.allclose(val, ref, rtol=0, atol=2e-6)
    far = np.abs(ref - 1.0) > 1e-5
    assert np.array_equal(lab[far], (ref[far] > 1.0).astype(np.uint8))


def test_sample_labels_errors(orc):
    o, v = _block_leaves(1, lambda x, y, z: np.zeros(x.shape))
    V, t = _tets_centred_at(np.array([[0.05, 0.05, 0.05]]))
    with pytest.raises(orc.OracleError):
        orc.sample_labels(o, v, 1.0, 1 / 64, V, t, np.zeros(3), np.array([1.0, 0.0, 1.0]), 0.0, 0.5)
    with pytest.raises(orc.OracleError):
        orc.sample_labels(o, v, 1.0, 1 / 64, V, t + 10, *UNIT, 0.0, 0.5)
    with pytest.raises(orc.OracleError):
        orc.sample_labels(np.concatenate([o, o]), np.concatenate([v, v]), 1.0, 1 / 64, V, t, *UNIT, 0.0, 0.5)
    with pytest.raises(orc.OracleError):
        orc.sample_labels(o + 3, v, 1.0, 1 / 64, V, t, *UNIT, 0.0, 0.5)


def test_reduce_threads_give_identical_arrays(orc):
    """The build on 1 and on 7 threads: the same reduced graph, counters and
    errors (ranges are summed; ids are assigned serially in tet order)."""
    m = S.kuhn_mesh(24, seed=5)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, S.sphere_with_holes(5)), 0.05, seed=6)
    a = orc.reduce(m.V, m.tets, m.nbrs, lab, 10.0, threads=1)
    b = orc.reduce(m.V, m.tets, m.nbrs, lab, 10.0, threads=7)
    assert a.n == b.n and a.info == b.info
    for x, y in ((a.full_to_free, b.full_to_free), (a.free_old, b.free_old), (a.nbr, b.nbr), (a.res, b.res),
                 (a.term, b.term)):
        assert np.array_equal(x, y)
    bad = m.nbrs.copy()
    bad[m.N - 3, 1] = m.N + 5
    for t in (1, 7):
        with pytest.raises(orc.OracleError):
            orc.reduce(m.V, m.tets, bad, lab, 10.0, threads=t)
