"""The native tetrahedraliser (csrc/wtv_delaunay.cpp) that makes the config-size
Delaunay meshes: pinned against scipy's Qhull (the same tets on points in general
position), against itself (any thread count gives the same canonical output),
and by closed forms (a lattice splits into exactly 6 tets per cube; a convex
hull's faces are the slots marked -1)."""
import numpy as np
import pytest

import wtv_synth as S

LUT = np.array([[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]])


def _tetset(tets):
    return set(map(tuple, np.sort(tets, axis=1).tolist()))


def _check_neighbours(tets, nbrs):
    """Slot fi of tet t is the tet across face LUT[fi] (the face opposite local
    vertex 3-fi), -1 on the hull; links are symmetric."""
    N = tets.shape[0]
    for fi in range(4):
        u = nbrs[:, fi]
        inner = u >= 0
        face = np.sort(tets[:, LUT[fi]], axis=1)[inner]
        other = tets[u[inner]]
        # every face vertex belongs to the neighbour
        assert all((other == face[:, [k]]).any(axis=1).all() for k in range(3))
        # and the neighbour points back
        back = (nbrs[u[inner]] == np.flatnonzero(inner)[:, None]).any(axis=1)
        assert back.all()
    assert nbrs.max() < N


def _volumes(V, tets):
    P = V[tets]
    return np.einsum("ij,ij->i", P[:, 1] - P[:, 0], np.cross(P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]))


@pytest.mark.parametrize("kind,n,seed", [("genus3", 40_000, 3), ("uniform", 20_000, 4), ("genus3", 5_000, 5)])
def test_same_tets_as_qhull(kind, n, seed):
    from scipy.spatial import Delaunay
    if kind == "genus3":
        P, _, _ = S.genus3_point_cloud(n, seed)
    else:
        P = np.random.default_rng(seed).uniform(size=(n, 3))
    Q, V = S.quantise_points(P)
    tets, nbrs, rep = S.native_delaunay(Q, threads=1, check=True)
    assert rep == 0
    q = Delaunay(V)
    assert _tetset(tets) == _tetset(q.simplices)
    _check_neighbours(tets, nbrs)
    # the hull: faces with no neighbour are exactly Qhull's hull facets
    hull = {tuple(sorted(f)) for f in q.convex_hull.tolist()}
    mine = {tuple(sorted(tets[t, LUT[fi]].tolist())) for t, fi in zip(*np.nonzero(nbrs < 0))}
    assert mine == hull
    assert (_volumes(V, tets) > 0).all()  # each tet keeps a positive orientation


def test_threads_give_identical_output():
    """Above 200k points the insertion runs on several threads; the output is
    canonical, so it equals the one-thread result array for array."""
    P, _, _ = S.genus3_point_cloud(230_000, 11)
    Q, _ = S.quantise_points(P)
    t1, n1, _ = S.native_delaunay(Q, threads=1, check=True)
    t4, n4, _ = S.native_delaunay(Q, threads=4, check=True)
    assert np.array_equal(t1, t4) and np.array_equal(n1, n4)
    assert 6.2 < t1.shape[0] / Q.shape[0] < 6.8


def test_lattice_cospherical_points():
    """A k^3 lattice: every cube's 8 corners lie on one sphere, the most
    degenerate input.  Any Delaunay triangulation of it fills each cube with
    tets of volume summing to the cube: (k-1)^3 cubes, total volume (k-1)^3,
    and the symbolic tie-breaking gives exactly 6 tets per cube."""
    k = 12
    g = np.stack(np.meshgrid(*[np.arange(k)] * 3, indexing="ij"), -1).reshape(-1, 3) * 1000
    tets, nbrs, _ = S.native_delaunay(g.astype(np.int32), threads=1, check=True)
    vol = _volumes(g.astype(np.float64), tets)
    assert (vol > 0).all()
    assert abs(vol.sum() / 6.0 - (k - 1) ** 3 * 1e9) < 1e-3
    assert tets.shape[0] == 6 * (k - 1) ** 3
    _check_neighbours(tets, nbrs)


def test_repeated_points_are_used_once():
    P = np.random.default_rng(8).uniform(size=(3000, 3))
    Q, _ = S.quantise_points(P)
    Qd = np.concatenate([Q, Q[:500]])  # 500 repeats at the end
    t, n, rep = S.native_delaunay(Q, threads=1)
    td, nd, repd = S.native_delaunay(Qd, threads=1)
    assert rep == 0 and repd == 500
    assert np.array_equal(t, td) and np.array_equal(n, nd)


def test_errors():
    with pytest.raises(RuntimeError, match="status 1"):
        S.native_delaunay(np.array([[0, 0, 0], [1 << 21, 0, 0], [0, 1, 0], [0, 0, 1]], np.int32))
    with pytest.raises(RuntimeError, match="status 4"):  # all on one plane
        S.native_delaunay(np.array([[0, 0, 0], [5, 0, 0], [0, 5, 0], [5, 5, 0], [2, 3, 0]], np.int32))
    with pytest.raises(RuntimeError, match="status 2"):
        S.native_delaunay(np.array([[0, 0, 0], [5, 0, 0], [0, 5, 0]], np.int32))


def test_native_mesh_feeds_the_soft_graph():
    """The soft-seed capacity model on a native Delaunay mesh: a valid graph
    (symmetric ids, reverse slots found) whose max flow BK and the restated
    push-relabel agree on."""
    import oracle
    g, mesh, meta = S.delaunay_genus3_graph(20_000, 21, method="native", device="cpu")
    nbr = g["nbr"].numpy()
    res = g["res"].numpy().view(np.uint32)
    term = g["term"].numpy()
    rs = S.reverse_slots(nbr)
    assert ((nbr < 0) | (rs >= 0)).all()
    a = oracle.bk_solve(nbr, res, term)
    b = oracle.pr_solve(nbr, res, term)
    assert a.raw_flow == b.raw_flow and np.array_equal(a.partition, b.partition)
