"""Writes the golden fixtures in tests/golden/ (small .npz inputs + expected outputs).

The reference ships no fixtures for this path, so these vectors are produced by
the oracle and accepted only when three independent results agree: the
restated push-relabel (oracle.pr_solve), Boykov-Kolmogorov (oracle.bk_solve)
and scipy.sparse.csgraph.maximum_flow, plus the max-flow/min-cut certificate.
They freeze that agreement as regression vectors for the GPU tests.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "wtivo-watertightvoxel_amd"))
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
import wtv_synth as S  # noqa: E402
from test_oracle import scipy_flow  # noqa: E402


def agree(nbr, res, term):
    a = oracle.pr_solve(nbr, res, term)
    b = oracle.bk_solve(nbr, res, term)
    f = scipy_flow(nbr, res, term)
    assert a.raw_flow == b.raw_flow == f, (a.raw_flow, b.raw_flow, f)
    assert np.array_equal(a.partition, b.partition)
    assert oracle.certificate(nbr, res, term, b.res, b.term, b.partition, b.raw_flow) == 0
    return b


def main():
    oracle.build()
    # reference model: graph_cut_fast inputs and expected outputs
    for k, seed in ((4, 0), (6, 1), (8, 2)):
        m = S.kuhn_mesh(k, seed=seed)
        shape = S.sphere_with_holes(seed, r=0.3, hole_r=0.08)
        lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.02, seed=seed + 7)
        r = oracle.reduce(m.V, m.tets, m.nbrs, lab, 10.0)
        sol = agree(r.nbr, r.res, r.term)
        out = oracle.scatter(lab, r.free_old, sol.partition)
        np.savez_compressed(os.path.join(HERE, f"ref_kuhn{k}_s{seed}.npz"), V=m.V, tets=m.tets, nbrs=m.nbrs,
                            labels=lab, fill_T=np.float64(10.0), out_labels=out,
                            raw_flow=np.int64(sol.raw_flow),
                            adjusted_flow=np.int64(sol.raw_flow + r.info["base_flow"]),
                            free_n=np.int64(r.n))
    # soft-seed model
    m = S.kuhn_mesh(8, seed=11)
    shape = S.sphere_with_holes(11)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.02, seed=12)
    g = S.soft_graph(m, lab, shape, 10.0, 1.0 / 8)
    sol = agree(g.nbr, g.res, g.term)
    np.savez_compressed(os.path.join(HERE, "soft_kuhn8.npz"), nbr=g.nbr, res=g.res, term=g.term,
                        partition=sol.partition, raw_flow=np.int64(sol.raw_flow))
    # integer stress model
    m = S.kuhn_mesh(6, seed=21)
    shape = S.thin_shell_with_tunnels(21, 1.0 / 6, tunnels=6)
    lab = S.labels_from_sdf(S.tet_centroids_sdf(m, shape), 0.0)
    g = S.lognormal_int_graph(m, lab, 100.0, seed=22)
    sol = agree(g.nbr, g.res, g.term)
    np.savez_compressed(os.path.join(HERE, "stress_kuhn6.npz"), nbr=g.nbr, res=g.res, term=g.term,
                        partition=sol.partition, raw_flow=np.int64(sol.raw_flow))
    # random 4-slot graphs
    for seed in (1, 2, 3):
        nbr, res, term = S.random_graph(300, 1000 + seed, cap_max=50, term_max=60)
        sol = agree(nbr, res, term)
        np.savez_compressed(os.path.join(HERE, f"random_n300_s{seed}.npz"), nbr=nbr, res=res, term=term,
                            partition=sol.partition, raw_flow=np.int64(sol.raw_flow))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
