"""Synthetic tet-graph workloads for the graph-cut solver.

Builds jittered Kuhn (Freudenthal) tetrahedralisations of a k^3 grid with the
reference's 4-slot neighbour layout, shape signed-distance functions, tet
labels, and the two capacity models the solver is exercised with:

* the reference model: graph_cut_fast computes capacities itself from the
  vertices, tets and labels (wtivo_gpupr_bindings.cpp:65-98);
* the soft-seed model: every tet is a free node with a terminal arc of
  lambda_fill * volume on its labelled side and facet arcs of
  area * exp(-|UDF(facet centroid)| / h), quantised to integers with one
  power-of-two scale per graph (the solver's arithmetic is integer,
  gpu_push_relabel_fast.h:23, wtivo_gpupr_bindings.cpp:217-220).

Slot convention (wtivo_core.cpp:68-71, wtivo_gpupr_bindings.cpp:70): slot fi is
the face {012, 013, 023, 123}[fi] of the tet's own vertex order, i.e. the face
opposite local vertex 3-fi; a hull face has neighbour -1.

Kuhn tet (cell c, axis permutation p) has vertices v0=c, v1=v0+e[p0],
v2=v1+e[p1], v3=c+(1,1,1).  Its neighbours:
  slot 0 (opposite v3): cell c-e[p2], perm (p2,p0,p1), their slot 3
  slot 1 (opposite v2): cell c,       perm (p0,p2,p1), their slot 1
  slot 2 (opposite v1): cell c,       perm (p1,p0,p2), their slot 2
  slot 3 (opposite v0): cell c+e[p0], perm (p1,p2,p0), their slot 0

All randomness comes from numpy's PCG64 with an explicit seed; heavy
arithmetic runs in torch (CPU threads, or the GPU when device="cuda").
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

try:  # torch is used as an array engine only
    import torch
except Exception:  # pragma: no cover
    torch = None

PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
_PIDX = {p: i for i, p in enumerate(PERMS)}
FACE_LUT = ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3))
KUHN_REV_SLOT = (3, 1, 2, 0)

# per (slot, perm): neighbour perm index and cell offset
_NB_PERM = np.zeros((4, 6), np.int64)
_NB_OFF = np.zeros((4, 6, 3), np.int64)
for _pi, (_a, _b, _c) in enumerate(PERMS):
    _e = np.eye(3, dtype=np.int64)
    _NB_PERM[0, _pi] = _PIDX[(_c, _a, _b)]
    _NB_OFF[0, _pi] = -_e[_c]
    _NB_PERM[1, _pi] = _PIDX[(_a, _c, _b)]
    _NB_PERM[2, _pi] = _PIDX[(_b, _a, _c)]
    _NB_PERM[3, _pi] = _PIDX[(_b, _c, _a)]
    _NB_OFF[3, _pi] = _e[_a]


@dataclass
class TetMesh:
    k: int
    V: np.ndarray      # float64 [nv,3]
    tets: np.ndarray   # int32 [N,4]
    nbrs: np.ndarray   # int32 [N,4]

    @property
    def N(self) -> int:
        return int(self.tets.shape[0])


def kuhn_mesh(k: int, jitter: float = 0.2, seed: int = 0, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
              device=None) -> TetMesh:
    """Jittered Kuhn 6-tet subdivision of a k^3 grid: 6k^3 tets, (k+1)^3 vertices (numpy arrays)."""
    V, tets, nbrs = kuhn_mesh_t(k, jitter, seed, lo, hi, device=device)
    return TetMesh(k, V.cpu().numpy(), tets.cpu().numpy(), nbrs.cpu().numpy())


def kuhn_vertices(k: int, jitter: float, seed: int, lo, hi) -> np.ndarray:
    """(k+1)^3 lattice points, x fastest, each jittered by U(-jitter, jitter) cells per axis."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k1 = k + 1
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    cell = (hi - lo) / k
    g = np.arange(k1, dtype=np.float64)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    V = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    V += rng.uniform(-jitter, jitter, size=V.shape)
    return lo + V * cell


def kuhn_mesh_t(k: int, jitter: float = 0.2, seed: int = 0, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
                device=None):
    """Torch version of kuhn_mesh: (V float64 [nv,3], tets int32 [N,4], nbrs int32 [N,4]) on `device`."""
    dev = _dev(device)
    V = torch.from_numpy(kuhn_vertices(k, jitter, seed, lo, hi)).to(dev)
    tets, nbrs = kuhn_topology_t(k, dev)
    return V, tets, nbrs


def kuhn_topology(k: int):
    """Tets and 4-slot neighbours of the Kuhn subdivision (numpy), tet id = cell*6 + perm."""
    t, n = kuhn_topology_t(k, "cpu")
    return t.numpy(), n.numpy()


def kuhn_topology_t(k: int, device=None, chunk_cells: int = 1 << 22):
    """Kuhn topology built with torch on `device` (int32 [6k^3,4] tets and neighbours)."""
    dev = _dev(device)
    k1 = k + 1
    ncell = k * k * k
    tets = torch.empty((ncell * 6, 4), dtype=torch.int32, device=dev)
    nbrs = torch.empty((ncell * 6, 4), dtype=torch.int32, device=dev)
    vstride = [1, k1, k1 * k1]
    for c0 in range(0, ncell, chunk_cells):
        c1 = min(ncell, c0 + chunk_cells)
        c = torch.arange(c0, c1, dtype=torch.int64, device=dev)
        cc = [c % k, (c // k) % k, c // (k * k)]
        base_v = cc[0] + cc[1] * k1 + cc[2] * (k1 * k1)
        m = c1 - c0
        T = torch.empty((m, 6, 4), dtype=torch.int32, device=dev)
        Nb = torch.empty((m, 6, 4), dtype=torch.int32, device=dev)
        for pi, p in enumerate(PERMS):
            v1 = base_v + vstride[p[0]]
            v2 = v1 + vstride[p[1]]
            v3 = base_v + sum(vstride)
            T[:, pi, 0] = base_v.to(torch.int32)
            T[:, pi, 1] = v1.to(torch.int32)
            T[:, pi, 2] = v2.to(torch.int32)
            T[:, pi, 3] = v3.to(torch.int32)
            for s in range(4):
                off = _NB_OFF[s, pi]
                ok = torch.ones(m, dtype=torch.bool, device=dev)
                nc = []
                for ax in range(3):
                    q = cc[ax] + int(off[ax])
                    ok &= (q >= 0) & (q < k)
                    nc.append(q)
                nid = (nc[0] + nc[1] * k + nc[2] * (k * k)) * 6 + int(_NB_PERM[s, pi])
                Nb[:, pi, s] = torch.where(ok, nid, torch.full_like(nid, -1)).to(torch.int32)
        tets[c0 * 6:c1 * 6] = T.reshape(-1, 4)
        nbrs[c0 * 6:c1 * 6] = Nb.reshape(-1, 4)
    return tets, nbrs


def permute_topology(tets: np.ndarray, nbrs: np.ndarray, seed: int):
    """The same mesh with its tets renumbered and each tet's vertices reordered,
    with the neighbour table that follows: new slot f of new tet pi[i] leaves
    out new local vertex 3-f, which is old local vertex sig[i, 3-f], i.e. old
    slot 3 - sig[i, 3-f].  Returns (tets, nbrs) of the permuted mesh."""
    rng = np.random.default_rng(seed)
    N = tets.shape[0]
    pi = rng.permutation(N)                       # old id -> new id
    sig = np.argsort(rng.random((N, 4)), axis=1)  # new local k <- old local sig[i, k]
    new_t = np.empty_like(tets)
    new_t[pi] = np.take_along_axis(tets, sig, axis=1)
    old_n = np.take_along_axis(nbrs, 3 - sig[:, ::-1], axis=1)
    new_n = np.empty_like(nbrs)
    new_n[pi] = np.where(old_n >= 0, pi[np.maximum(old_n, 0)], -1)
    return new_t, new_n


def reverse_slots(nbrs: np.ndarray) -> np.ndarray:
    """rs[i,k] = slot of i in nbrs[nbrs[i,k]] (or -1 for hull / missing)."""
    N = nbrs.shape[0]
    rs = np.full((N, 4), -1, np.int8)
    idx = np.arange(N, dtype=np.int64)
    for k in range(4):
        j = nbrs[:, k].astype(np.int64)
        ok = j >= 0
        jj = np.where(ok, j, 0)
        for r in range(4):
            hit = ok & (nbrs[jj, r] == idx) & (rs[:, k] < 0)
            rs[hit, k] = r
    return rs


# --------------------------------------------------------------------------
# shapes (signed distance, negative inside)
# --------------------------------------------------------------------------
class Shape:
    def sdf(self, p):  # p: torch [m,3] float64
        raise NotImplementedError


@dataclass
class Sphere(Shape):
    c: tuple
    r: float

    def sdf(self, p):
        c = p.new_tensor(self.c)
        return torch.linalg.vector_norm(p - c, dim=1) - self.r


@dataclass
class Torus(Shape):
    """Torus around axis `ax` (unit vector) through centre c, radii R, r."""
    c: tuple
    ax: tuple
    R: float
    r: float

    def sdf(self, p):
        c = p.new_tensor(self.c)
        a = p.new_tensor(self.ax)
        a = a / torch.linalg.vector_norm(a)
        d = p - c
        h = d @ a
        radial = torch.linalg.vector_norm(d - h[:, None] * a[None, :], dim=1)
        q = radial - self.R
        return torch.sqrt(q * q + h * h) - self.r


@dataclass
class Cylinder(Shape):
    """Infinite cylinder of radius r along unit direction `ax` through point c."""
    c: tuple
    ax: tuple
    r: float

    def sdf(self, p):
        c = p.new_tensor(self.c)
        a = p.new_tensor(self.ax)
        a = a / torch.linalg.vector_norm(a)
        d = p - c
        h = d @ a
        return torch.linalg.vector_norm(d - h[:, None] * a[None, :], dim=1) - self.r


@dataclass
class Union(Shape):
    parts: list

    def sdf(self, p):
        out = self.parts[0].sdf(p)
        for s in self.parts[1:]:
            out = torch.minimum(out, s.sdf(p))
        return out


@dataclass
class Subtract(Shape):
    a: Shape
    holes: list

    def sdf(self, p):
        out = self.a.sdf(p)
        for h in self.holes:
            out = torch.maximum(out, -h.sdf(p))
        return out


@dataclass
class Shell(Shape):
    """|sdf(a)| - t/2 : a hollow wall of thickness t around a's surface."""
    a: Shape
    t: float

    def sdf(self, p):
        return torch.abs(self.a.sdf(p)) - 0.5 * self.t


def _unit(rng):
    v = rng.normal(size=3)
    return tuple((v / np.linalg.norm(v)).tolist())


def sphere_with_holes(seed: int, r=0.35, hole_r=0.05, holes=3) -> Shape:
    """Sphere of radius r at the centre with `holes` seeded cylinders punched through it."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cyl = []
    for _ in range(holes):
        off = rng.uniform(-0.15, 0.15, size=3)
        cyl.append(Cylinder(tuple((0.5 + off).tolist()), _unit(rng), hole_r))
    return Subtract(Sphere((0.5, 0.5, 0.5), r), cyl)


def genus3(seed: int = 0) -> Shape:
    """Three fused tori in a row along x spanning most of the unit box (a genus-3 solid)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for i in range(3):
        cx = 0.5 + (i - 1) * 0.26 + rng.uniform(-0.01, 0.01)
        parts.append(Torus((cx, 0.5, 0.5), (0.0, 0.0, 1.0), 0.135, 0.065))
    return Union(parts)


def thin_shell_with_tunnels(seed: int, cell: float, tunnels=40) -> Shape:
    """Spherical shell with wall thickness 1.5 cells and seeded tunnels through the wall."""
    rng = np.random.Generator(np.random.PCG64(seed))
    holes = []
    for _ in range(tunnels):
        d = np.asarray(_unit(rng))
        c = 0.5 + 0.35 * d
        holes.append(Cylinder(tuple(c.tolist()), tuple(d.tolist()), 1.5 * cell))
    return Subtract(Shell(Sphere((0.5, 0.5, 0.5), 0.35), 1.5 * cell), holes)


def random_shape(rng: np.random.Generator) -> Shape:
    """Union of 1-4 spheres or tori with 0-5 punched holes."""
    parts = []
    for _ in range(int(rng.integers(1, 5))):
        c = tuple(rng.uniform(0.3, 0.7, size=3).tolist())
        if rng.uniform() < 0.5:
            parts.append(Sphere(c, float(rng.uniform(0.1, 0.25))))
        else:
            R = float(rng.uniform(0.1, 0.2))
            parts.append(Torus(c, _unit(rng), R, float(rng.uniform(0.03, 0.6 * R))))
    holes = [Cylinder(tuple(rng.uniform(0.35, 0.65, size=3).tolist()), _unit(rng),
                      float(rng.uniform(0.02, 0.06))) for _ in range(int(rng.integers(0, 6)))]
    shape = Union(parts) if len(parts) > 1 else parts[0]
    return Subtract(shape, holes) if holes else shape


# --------------------------------------------------------------------------
# geometry helpers (torch, chunked)
# --------------------------------------------------------------------------
def _dev(device):
    return torch.device(device if device is not None else "cpu")


def tet_centroids_sdf(mesh: TetMesh, shape: Shape, device=None, chunk=1 << 22) -> np.ndarray:
    dev = _dev(device)
    Vt = torch.from_numpy(mesh.V).to(dev)
    out = np.empty(mesh.N, np.float64)
    for a in range(0, mesh.N, chunk):
        b = min(mesh.N, a + chunk)
        t = torch.from_numpy(mesh.tets[a:b].astype(np.int64)).to(dev)
        cen = Vt[t].mean(dim=1)
        out[a:b] = shape.sdf(cen).cpu().numpy()
    return out


def labels_from_sdf(sdf: np.ndarray, flip: float = 0.0, seed: int = 0) -> np.ndarray:
    """0 = inside (source side, wtvo_vdb.cpp:1578 polarity), 1 = outside; then seeded flips."""
    lab = (sdf >= 0.0).astype(np.uint8)
    if flip > 0:
        rng = np.random.Generator(np.random.PCG64(seed))
        f = rng.uniform(size=lab.shape[0]) < flip
        lab ^= f.astype(np.uint8)
    return lab


def band_labels(sdf: np.ndarray, band: float) -> np.ndarray:
    """Reference labels: 0 = within the UDF band (|sdf| <= band), else 1 (wtivo_vdb.cpp:1578)."""
    return (np.abs(sdf) > band).astype(np.uint8)


@dataclass
class SoftGraph:
    """Soft-seed graph: every tet is a free node."""
    nbr: np.ndarray        # int32 [N,4]
    res: np.ndarray        # uint32 [N,4] quantised facet capacities
    term: np.ndarray       # int32 [N] quantised terminal (+ source, - sink)
    cap_f32: np.ndarray    # float32 [N,4]
    term_f32: np.ndarray   # float32 [N] (+ source, - sink)
    scale: float
    meta: dict = field(default_factory=dict)


def reverse_slots_t(nbrs):
    """Torch reverse slots: rs[i,k] = slot of i in nbrs[nbrs[i,k]], -1 for hull/missing (int64)."""
    N = nbrs.shape[0]
    dev = nbrs.device
    idx = torch.arange(N, dtype=torch.int64, device=dev)
    rs = torch.full((N, 4), -1, dtype=torch.int64, device=dev)
    nb64 = nbrs.to(torch.int64)
    for k in range(4):
        j = nb64[:, k]
        ok = j >= 0
        row = nb64[torch.where(ok, j, torch.zeros_like(j))]
        for r in range(3, -1, -1):  # first match wins
            hit = ok & (row[:, r] == idx)
            rs[:, k] = torch.where(hit, torch.full_like(j, r), rs[:, k])
    return rs


def tet_sdf_t(V, tets, shape: Shape, chunk=1 << 22):
    """Shape SDF at tet centroids (torch, on the tensors' device)."""
    N = tets.shape[0]
    out = torch.empty(N, dtype=torch.float64, device=V.device)
    for a in range(0, N, chunk):
        b = min(N, a + chunk)
        out[a:b] = shape.sdf(V[tets[a:b].long()].mean(dim=1))
    return out


def flip_mask(N: int, p: float, seed: int) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.uniform(size=N) < p


def soft_graph_t(V, tets, nbrs, labels, shape: Shape, lam: float, h: float, rev_slots=None,
                 chunk=1 << 21, max_cap=(1 << 30) - 1, max_term=(1 << 31) - 1):
    """Soft-seed capacity model on torch tensors (any device).

    terminal = lam * |vol| (fp32) on the labelled side (+ source for label 0,
    - sink otherwise); facet = area * exp(-|UDF(facet centroid)|/h) (fp32),
    one value per facet (the lower tet id's evaluation); all quantised with one
    power-of-two scale S, the largest with every arc < 2^30 and |term| < 2^31,
    cap_int = floor(S * cap_f32).  Returns a dict of tensors on V.device:
    nbr int32 [N,4], res int32 [N,4] (uint32 bit pattern), term int32 [N],
    cap_f32 [N,4], term_f32 [N], and the python float `scale`.
    """
    dev = V.device
    N = tets.shape[0]
    cap = torch.empty((N, 4), dtype=torch.float32, device=dev)
    vol = torch.empty(N, dtype=torch.float32, device=dev)
    for a in range(0, N, chunk):
        b = min(N, a + chunk)
        P = V[tets[a:b].long()]  # [m,4,3]
        d1 = P[:, 1] - P[:, 0]
        d2 = P[:, 2] - P[:, 0]
        d3 = P[:, 3] - P[:, 0]
        vol[a:b] = (torch.abs((torch.cross(d1, d2, dim=1) * d3).sum(1)) / 6.0).float()
        for fi in range(4):
            i0, i1, i2 = FACE_LUT[fi]
            e1 = P[:, i1] - P[:, i0]
            e2 = P[:, i2] - P[:, i0]
            area = 0.5 * torch.linalg.vector_norm(torch.cross(e1, e2, dim=1), dim=1)
            cen = (P[:, i0] + P[:, i1] + P[:, i2]) / 3.0
            udf = torch.abs(shape.sdf(cen))
            cap[a:b, fi] = (area * torch.exp(-udf / h)).float()
        del P
    nb64 = nbrs.to(torch.int64)
    if rev_slots is None:
        rs = reverse_slots_t(nbrs)
    else:
        rs = torch.tensor(rev_slots, dtype=torch.int64, device=dev).expand(N, 4)
    for a in range(0, N, chunk):  # symmetrise: the lower id's value wins
        b = min(N, a + chunk)
        j = nb64[a:b]
        idx = torch.arange(a, b, dtype=torch.int64, device=dev)[:, None]
        take = (j >= 0) & (j < idx)
        jj = torch.where(take, j, torch.zeros_like(j))
        rr = torch.where(take, rs[a:b], torch.zeros_like(j))
        cap[a:b] = torch.where(take, cap[jj, rr], cap[a:b])
    cap[nbrs < 0] = 0.0
    tf = (lam * vol).float()
    term_f32 = torch.where(labels == 0, tf, -tf)
    mc = float(cap.max()) if N else 0.0
    mt = float(term_f32.abs().max()) if N else 0.0
    sc = math.inf
    if mc > 0:
        sc = min(sc, max_cap / mc)
    if mt > 0:
        sc = min(sc, max_term / mt)
    scale = 2.0 ** math.floor(math.log2(sc)) if math.isfinite(sc) else 1.0
    res = torch.floor(cap.double() * scale).to(torch.int64)
    term = torch.trunc(term_f32.double() * scale).to(torch.int64)
    assert int(res.max()) <= max_cap and int(term.abs().max()) <= max_term
    return {"nbr": nbrs.to(torch.int32).contiguous(), "res": res.to(torch.int32).contiguous(),
            "term": term.to(torch.int32).contiguous(), "cap_f32": cap, "term_f32": term_f32,
            "scale": scale}


def soft_graph(mesh: TetMesh, labels: np.ndarray, shape: Shape, lam: float, h: float,
               device=None, rev_slots=None) -> "SoftGraph":
    """numpy wrapper of soft_graph_t."""
    dev = _dev(device)
    V = torch.from_numpy(mesh.V).to(dev)
    T = torch.from_numpy(mesh.tets).to(dev)
    Nb = torch.from_numpy(mesh.nbrs).to(dev)
    L = torch.from_numpy(np.ascontiguousarray(labels, dtype=np.uint8)).to(dev)
    g = soft_graph_t(V, T, Nb, L, shape, lam, h, rev_slots=rev_slots)
    return SoftGraph(g["nbr"].cpu().numpy(), g["res"].cpu().numpy().view(np.uint32),
                     g["term"].cpu().numpy(), g["cap_f32"].cpu().numpy(), g["term_f32"].cpu().numpy(),
                     g["scale"], {"lam": lam, "h": h, "N": mesh.N})


def cut_energy(g: SoftGraph, partition: np.ndarray) -> float:
    """FP64 cut energy of a partition (1 = sink side) from the original fp32 capacities."""
    p = partition.astype(bool)
    t = g.term_f32.astype(np.float64)
    e = float(np.sum(np.where(~p & (t < 0), -t, 0.0)) + np.sum(np.where(p & (t > 0), t, 0.0)))
    j = g.nbr.astype(np.int64)
    ok = j >= 0
    pj = np.where(ok, p[np.where(ok, j, 0)], False)
    cut = ok & (~p)[:, None] & pj
    return e + float(np.sum(np.where(cut, g.cap_f32.astype(np.float64), 0.0)))


def quantised_cut(nbr, res0, term0, partition) -> int:
    """Integer cut capacity of a partition on the quantised graph (== max-flow at optimum)."""
    p = partition.astype(bool)
    t = term0.astype(np.int64)
    e = int(np.sum(np.where(~p & (t < 0), -t, 0)) + np.sum(np.where(p & (t > 0), t, 0)))
    j = nbr.astype(np.int64)
    ok = j >= 0
    pj = np.where(ok, p[np.where(ok, j, 0)], False)
    cut = ok & (~p)[:, None] & pj
    return e + int(np.sum(np.where(cut, res0.astype(np.int64), 0)))


def lognormal_int_graph(mesh: TetMesh, labels: np.ndarray, lam: float, seed: int, sigma=2.0,
                        cap_hi=(1 << 20), rev_slots=None) -> SoftGraph:
    """Integer stress model: facet capacities log-normal(sigma) quantised to [1, cap_hi]
    (one value per facet), terminals lam * round(volume-proportional) integers."""
    rng = np.random.Generator(np.random.PCG64(seed))
    N = mesh.N
    nb = mesh.nbrs
    raw = rng.lognormal(mean=0.0, sigma=sigma, size=(N, 4))
    q = np.clip(np.rint(raw * 64.0), 1, cap_hi).astype(np.int64)
    rs = reverse_slots(nb) if rev_slots is None else np.broadcast_to(np.asarray(rev_slots, np.int64), (N, 4))
    idx = np.arange(N, dtype=np.int64)[:, None]
    j = nb.astype(np.int64)
    take = (j >= 0) & (j < idx)
    q = np.where(take, q[np.where(take, j, 0), np.where(take, rs, 0).astype(np.int64)], q)
    q[nb < 0] = 0
    t = np.full(N, int(lam) * 64, np.int64)
    term = np.where(labels == 0, t, -t)
    return SoftGraph(nb.copy(), q.astype(np.uint32), term.astype(np.int32), q.astype(np.float32),
                     term.astype(np.float32), 1.0, {"lam": lam, "sigma": sigma, "N": N})


def random_graph(n: int, seed: int, edges_per_node: float = 1.6, cap_max: int = 20,
                 term_max: int = 30, p_term: float = 0.6):
    """Random symmetric graph with at most 4 arcs per node in random slots (tests)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nbr = np.full((n, 4), -1, np.int32)
    res = np.zeros((n, 4), np.uint32)
    deg = np.zeros(n, np.int64)
    tries = int(edges_per_node * n * 4) + 8
    target = int(edges_per_node * n)
    m = 0
    for _ in range(tries):
        if m >= target or n < 2:
            break
        a, b = (int(x) for x in rng.integers(0, n, size=2))
        if a == b or deg[a] >= 4 or deg[b] >= 4:
            continue
        if np.any(nbr[a] == b):
            continue
        sa = int(rng.choice(np.flatnonzero(nbr[a] < 0)))
        sb = int(rng.choice(np.flatnonzero(nbr[b] < 0)))
        nbr[a, sa] = b
        nbr[b, sb] = a
        res[a, sa] = int(rng.integers(0, cap_max + 1))
        res[b, sb] = int(rng.integers(0, cap_max + 1))
        deg[a] += 1
        deg[b] += 1
        m += 1
    term = np.zeros(n, np.int32)
    has = rng.uniform(size=n) < p_term
    mag = rng.integers(1, term_max + 1, size=n)
    sign = np.where(rng.uniform(size=n) < 0.5, 1, -1)
    term[has] = (mag * sign)[has]
    return nbr, res, term


def batch_workload(count: int, seed: int, n_min: int = 50_000, n_max: int = 4_000_000, device=None,
                   lam_range=(1.0, 30.0), flip: float = 0.02):
    """Uneven batch of soft-seed Kuhn graphs: tet counts log-uniform in [n_min, n_max]
    (k = round((N/6)^(1/3))), each with a random shape (union of 1-4 spheres or tori,
    0-5 punched holes) and lambda_fill ~ U(lam_range).  Returns a list of dicts with
    device tensors nbr/res/term (local ids) and metadata."""
    dev = _dev(device)
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(count):
        target = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        k = max(4, int(round((target / 6.0) ** (1.0 / 3.0))))
        shape = random_shape(rng)
        lam = float(rng.uniform(*lam_range))
        V, T, Nb = kuhn_mesh_t(k, seed=seed * 1000 + i, device=dev)
        N = T.shape[0]
        sdf = tet_sdf_t(V, T, shape)
        lab = (sdf >= 0).to(torch.uint8)
        if flip > 0:
            lab ^= torch.from_numpy(flip_mask(N, flip, seed * 1000 + i + 7).astype(np.uint8)).to(dev)
        g = soft_graph_t(V, T, Nb, lab, shape, lam, 1.0 / k, rev_slots=KUHN_REV_SLOT)
        out.append({"k": k, "N": N, "lam": lam, "nbr": g["nbr"], "res": g["res"], "term": g["term"]})
        del V, T, Nb, sdf, lab, g
    return out


# --------------------------------------------------------------------------
# Delaunay meshes of a genus-3 point cloud (the production-shape mesh family)
# --------------------------------------------------------------------------
def _torus_frame(t: Torus):
    a = np.asarray(t.ax, np.float64)
    a = a / np.linalg.norm(a)
    u = np.cross(a, [1.0, 0.0, 0.0] if abs(a[0]) < 0.9 else [0.0, 1.0, 0.0])
    u /= np.linalg.norm(u)
    return a, u, np.cross(a, u)


def torus_surface_points(t: Torus, m: int, rng: np.random.Generator) -> np.ndarray:
    """m points uniform by area on a torus surface.  The angle around the tube,
    phi, has density proportional to R + r cos(phi) (rejection sampling)."""
    a, u, w = _torus_frame(t)
    # numpy draws the numbers, torch (vectorised on the host) does the trigonometry
    phis = []
    have = 0
    while have < m:
        cand = torch.from_numpy(rng.uniform(0.0, 2.0 * np.pi, size=2 * (m - have) + 64))
        acc = torch.from_numpy(rng.uniform(size=cand.numel()))
        keep = acc * (t.R + t.r) < t.R + t.r * torch.cos(cand)
        phis.append(cand[keep])
        have += int(phis[-1].numel())
    phi = torch.cat(phis)[:m]
    th = torch.from_numpy(rng.uniform(0.0, 2.0 * np.pi, size=m))
    rad = t.R + t.r * torch.cos(phi)
    out = torch.tensor(t.c, dtype=torch.float64)[None, :].repeat(m, 1)
    out += (rad * torch.cos(th))[:, None] * torch.from_numpy(u)[None, :]
    out += (rad * torch.sin(th))[:, None] * torch.from_numpy(w)[None, :]
    out += (t.r * torch.sin(phi))[:, None] * torch.from_numpy(a)[None, :]
    return out.numpy()


def genus3_point_cloud(npts: int, seed: int, res: float | None = None, surface_frac: float = 0.95):
    """Point cloud of the genus-3 solid (three fused tori, genus3(seed)):
    surface_frac of the points uniform by area on the fused surface (each torus's
    surface minus the parts inside another torus) with N(0, 0.5 voxel) normal
    noise, the rest uniform in the unit box.  The voxel is 1/res; res defaults to
    1536 * sqrt(npts / 12e6), which keeps the noise-to-spacing ratio of 12M points
    at resolution 1536.  Returns (points float64 [npts,3], shape, res)."""
    shape = genus3(seed)
    if res is None:
        res = 1536.0 * math.sqrt(npts / 12e6)
    rng = np.random.Generator(np.random.PCG64(seed + 11))
    ns = int(round(surface_frac * npts))
    tori = shape.parts
    got, need = [], ns
    while need > 0:
        m = int(need * 1.5) + 256
        which = rng.integers(0, len(tori), size=m)
        P = np.empty((m, 3))
        for i, t in enumerate(tori):
            sel = which == i
            P[sel] = torus_surface_points(t, int(sel.sum()), rng)
        pt = torch.from_numpy(P)
        outside_others = torch.ones(m, dtype=torch.bool)
        for i, t in enumerate(tori):
            d = t.sdf(pt)
            outside_others &= torch.from_numpy(which == i) | (d >= 0)
        kept = P[outside_others.numpy()][:need]
        got.append(kept)
        need -= kept.shape[0]
    S = np.concatenate(got)
    S += rng.normal(0.0, 0.5 / res, size=S.shape)
    U = rng.uniform(0.0, 1.0, size=(npts - ns, 3))
    return np.concatenate([S, U]), shape, res


GRID_BITS = 21  # coordinates of tetrahedralised points: multiples of 2^-21 in [0, 1)


def quantise_points(P: np.ndarray):
    """Points of the unit box snapped down to the 2^-21 grid: (integer coordinates
    int32 [n,3], the same points as float64 [n,3]).  Both forms are exact (and
    exact in fp32 too), so a tetrahedralisation of the integers is one of the
    float64 points."""
    scale = float(1 << GRID_BITS)
    Q = np.clip(np.floor(np.asarray(P, np.float64) * scale), 0, scale - 1).astype(np.int32)
    return Q, Q.astype(np.float64) / scale


_DT_LIB = None


def _delaunay_lib():
    global _DT_LIB
    if _DT_LIB is None:
        import ctypes as C
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libwtv_delaunay.so")
        if not os.path.exists(path):
            raise RuntimeError(f"libwtv_delaunay.so is not built ({path}); run __graft_entry__.build()")
        L = C.CDLL(path)
        L.wtvd_triangulate.restype = C.c_void_p
        L.wtvd_triangulate.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int)]
        L.wtvd_export.restype = None
        L.wtvd_export.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.wtvd_check.restype = C.c_int
        L.wtvd_check.argtypes = [C.c_void_p]
        L.wtvd_free.restype = None
        L.wtvd_free.argtypes = [C.c_void_p]
        _DT_LIB = L
    return _DT_LIB


def native_delaunay(Q: np.ndarray, threads: int | None = None, check: bool = False):
    """Delaunay tetrahedralisation of integer points (int32 [n,3] in [0, 2^21),
    as quantise_points returns) by csrc/wtv_delaunay.cpp: (tets int32 [N,4],
    nbrs int32 [N,4], repeated points skipped).  Tets come in a canonical order
    (by their smallest vertex in Morton order); slot s holds the neighbour
    across the face opposite local vertex 3-s, -1 on the hull."""
    import ctypes as C
    import os
    L = _delaunay_lib()
    Q = np.ascontiguousarray(Q, dtype=np.int32)
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    nt, rep, st = C.c_int64(0), C.c_int64(0), C.c_int(0)
    h = L.wtvd_triangulate(Q.ctypes.data, Q.shape[0], int(threads), C.byref(nt), C.byref(rep), C.byref(st))
    if not h:
        raise RuntimeError(f"tetrahedralisation failed (status {st.value})")
    try:
        if check:
            rc = L.wtvd_check(h)
            if rc != 0:
                raise RuntimeError(f"tetrahedralisation check failed ({rc})")
        tets = np.empty((nt.value, 4), np.int32)
        nbrs = np.empty((nt.value, 4), np.int32)
        L.wtvd_export(h, tets.ctypes.data, nbrs.ctypes.data)
    finally:
        L.wtvd_free(h)
    return tets, nbrs, int(rep.value)


def _morton_key(P: np.ndarray, bits: int = 21) -> np.ndarray:
    lo = P.min(0)
    span = max(float((P.max(0) - lo).max()), 1e-300)
    q = np.clip(((P - lo) / span * ((1 << bits) - 1)).astype(np.int64), 0, (1 << bits) - 1)
    key = np.zeros(P.shape[0], np.int64)
    for b in range(bits):
        for ax in range(3):
            key |= ((q[:, ax] >> b) & 1) << (3 * b + ax)
    return key


def delaunay_mesh(points: np.ndarray, morton: bool = True) -> TetMesh:
    """3D Delaunay tetrahedralisation (scipy's Qhull) in the 4-slot layout:
    slot s holds the neighbour across the face opposite local vertex 3-s, -1 on
    the hull (Qhull's neighbours[:, j] is the one opposite vertex j).  With
    morton=True the tets are renumbered in Morton order of their centroids, the
    spatially coherent order a mesher that inserts points along a space-filling
    curve hands over; the solver itself takes any order."""
    from scipy.spatial import Delaunay

    d = Delaunay(points)
    tets = np.ascontiguousarray(d.simplices, dtype=np.int64)
    nbrs = np.ascontiguousarray(d.neighbors[:, ::-1], dtype=np.int64)
    if morton:
        order = np.argsort(_morton_key(points[tets].mean(axis=1)), kind="stable")
        inv = np.empty_like(order)
        inv[order] = np.arange(order.size)
        tets = tets[order]
        nbrs = nbrs[order]
        nbrs = np.where(nbrs >= 0, inv[np.maximum(nbrs, 0)], -1)
    return TetMesh(0, np.ascontiguousarray(points, dtype=np.float64), tets.astype(np.int32),
                   nbrs.astype(np.int32))


def native_delaunay_mesh(points: np.ndarray, threads: int | None = None) -> TetMesh:
    """Delaunay mesh of `points` (unit box) snapped to the 2^-21 grid, by the
    native tetrahedraliser; V holds the snapped points, tets come in canonical
    order (Morton order of their smallest vertex)."""
    Q, V = quantise_points(points)
    tets, nbrs, _ = native_delaunay(Q, threads)
    return TetMesh(0, V, tets, nbrs)


def delaunay_genus3_graph(npts: int, seed: int, lam: float = 10.0, flip: float = 0.02, device=None,
                          res: float | None = None, method: str = "qhull", threads: int | None = None):
    """Soft-seed graph of the Delaunay mesh of genus3_point_cloud(npts, seed):
    labels from the shape SDF at tet centroids with `flip` seeded flips,
    terminal lam * volume, facet area * exp(-|UDF|/h) with h one voxel (1/res).
    method "qhull": scipy's Qhull, Morton-ordered tets; "native": the native
    tetrahedraliser on the points snapped to the 2^-21 grid, canonical order.
    Returns (graph dict as soft_graph_t, mesh, meta)."""
    dev = _dev(device)
    P, shape, res = genus3_point_cloud(npts, seed, res)
    mesh = native_delaunay_mesh(P, threads) if method == "native" else delaunay_mesh(P)
    V = torch.from_numpy(mesh.V).to(dev)
    T = torch.from_numpy(mesh.tets).to(dev)
    Nb = torch.from_numpy(mesh.nbrs).to(dev)
    sdf = tet_sdf_t(V, T, shape)
    lab = (sdf >= 0).to(torch.uint8)
    if flip > 0:
        lab ^= torch.from_numpy(flip_mask(mesh.N, flip, seed + 1).astype(np.uint8)).to(dev)
    g = soft_graph_t(V, T, Nb, lab, shape, lam, 1.0 / res)
    meta = {"points": npts, "tets": mesh.N, "res": res, "h": 1.0 / res, "lam": lam, "flip": flip,
            "inside_frac": float((lab == 0).float().mean())}
    return g, mesh, meta


def udf_leaves(shape: Shape, voxel_size: float, band_voxels: float, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
               device=None, chunk: int = 1 << 21):
    """Sparse unsigned distance grid of `shape`'s surface in the layout the
    device sampler takes: 8^3 leaves whose index-space cells come within the
    band of the surface, voxel (i,j,k) at world (i,j,k) * voxel_size (OpenVDB's
    linear transform), value min(|sdf|, background) with background =
    band_voxels * voxel_size, leaf order offset (x&7)<<6 | (y&7)<<3 | (z&7).
    Returns (origins int32 [L,3], values float32 [L,512], background)."""
    dev = _dev(device)
    bg = float(band_voxels * voxel_size)
    i0 = [int(np.floor(l / voxel_size)) - 1 for l in lo]
    i1 = [int(np.ceil(h / voxel_size)) + 1 for h in hi]
    c0 = [v >> 3 for v in i0]
    c1 = [v >> 3 for v in i1]
    ext = [b - a + 1 for a, b in zip(c0, c1)]
    ncell = ext[0] * ext[1] * ext[2]
    reach = bg + 8 * voxel_size * 0.8661   # band + half the leaf diagonal
    keep = []
    for s0 in range(0, ncell, chunk):
        c = torch.arange(s0, min(ncell, s0 + chunk), dtype=torch.int64, device=dev)
        cx = c % ext[0] + c0[0]
        cy = (c // ext[0]) % ext[1] + c0[1]
        cz = c // (ext[0] * ext[1]) + c0[2]
        ctr = torch.stack([cx, cy, cz], 1).to(torch.float64) * 8.0 + 3.5
        d = shape.sdf(ctr * voxel_size).abs()
        keep.append(torch.stack([cx, cy, cz], 1)[d <= reach])
    cells = torch.cat(keep) if keep else torch.zeros((0, 3), dtype=torch.int64, device=dev)
    o = torch.arange(8, device=dev)
    off = torch.stack(torch.meshgrid(o, o, o, indexing="ij"), -1).reshape(512, 3)   # x-major: (x<<6)|(y<<3)|z
    vals = []
    per = max(1, chunk // 512)
    for s0 in range(0, cells.shape[0], per):
        cc = cells[s0:s0 + per]
        idx = (cc[:, None, :] * 8 + off[None, :, :]).to(torch.float64)
        d = shape.sdf(idx.reshape(-1, 3) * voxel_size).abs().clamp(max=bg)
        vals.append(d.to(torch.float32).reshape(-1, 512))
    values = torch.cat(vals) if vals else torch.zeros((0, 512), dtype=torch.float32, device=dev)
    return ((cells * 8).to(torch.int32).cpu().numpy(), values.cpu().numpy(), np.float32(bg))
