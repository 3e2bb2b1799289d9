"""CPU checks of the synthetic mesh families the solver is exercised on: the
4-slot neighbour layout must be a consistent facet adjacency (every interior
slot's neighbour points back through a face with the same three vertices), and
the CPU oracle's two max-flow schedules must agree on the resulting graphs."""
import numpy as np
import pytest

import wtv_synth as S


def _check_facets(mesh):
    nb = mesh.nbrs.astype(np.int64)
    rs = S.reverse_slots(mesh.nbrs)
    ok = nb >= 0
    assert ((rs < 0) & ok).sum() == 0, "interior slot without a reverse slot"
    assert (~ok).sum() > 0, "a bounded mesh has hull faces"
    tets = mesh.tets.astype(np.int64)
    lut = np.asarray(S.FACE_LUT)
    for s in range(4):
        i = np.nonzero(ok[:, s])[0]
        j = nb[i, s]
        r = rs[i, s]
        mine = np.sort(tets[i][:, lut[s]], axis=1)
        theirs = np.sort(np.take_along_axis(tets[j], lut[r], axis=1), axis=1)
        assert np.array_equal(mine, theirs), f"slot {s}: shared face vertices differ"


def test_kuhn_mesh_facets():
    _check_facets(S.kuhn_mesh(5, seed=3))


@pytest.mark.parametrize("morton", [False, True])
def test_delaunay_mesh_facets(morton):
    P, _, _ = S.genus3_point_cloud(3000, 5)
    _check_facets(S.delaunay_mesh(P, morton=morton))


def test_genus3_point_cloud_on_surface():
    P, shape, res = S.genus3_point_cloud(20000, 9)
    import torch
    d = shape.sdf(torch.from_numpy(P)).numpy()
    ns = int(round(0.95 * 20000))
    # surface points sit within a few noise sigmas (0.5 voxel) of the surface
    assert np.percentile(np.abs(d[:ns]), 99) < 3 * 0.5 / res
    # the uniform 5% fill the box
    assert P[ns:].min() >= 0.0 and P[ns:].max() <= 1.0
    # all three tori carry points (the fused surface of a genus-3 solid)
    xs = P[:ns, 0]
    for t in shape.parts:
        assert np.sum(np.abs(xs - t.c[0]) < 0.05) > 100


def test_delaunay_graph_pr_equals_bk(orc):
    g, mesh, meta = S.delaunay_genus3_graph(4000, 21)
    nbr = g["nbr"].numpy()
    res = g["res"].numpy().view(np.uint32)
    term = g["term"].numpy()
    b = orc.bk_solve(nbr, res, term)
    p = orc.pr_solve(nbr, res, term)
    assert b.raw_flow == p.raw_flow > 0
    assert np.array_equal(b.partition, p.partition)
    assert orc.certificate(nbr, res, term, p.res, p.term, p.partition, p.raw_flow) == 0
