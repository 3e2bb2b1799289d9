"""The fused sub-round's schedule argument, checked on the CPU.

k_push_fused lets every visit of a sub-round run at once: a visit loads its
own state, then its neighbours' heights, pushes, raises itself by at most 2
and pushes again, and its writes (height and terminal stores, residual
atomics, excess atomics) land in any order relative to other visits' loads.
This test replays that schedule in Python with random interleavings of those
steps -- each visit's two loads in order, its three write groups in any order
after them, so another visit may see an excess add before the residual add
that came with it -- and checks after every sub-round what the solver relies on:

* the heights are a valid labelling (h(v) <= h(u) + 1 on every residual arc,
  h(v) <= 1 with sink capacity left);
* a global relabel never lowers a height, so filtering the old active list
  keeps every active vertex;
* every active vertex (excess > 0, height < n) is listed, once;
* the final flow and labels equal the oracle's.

A mistake in the argument -- a raise cap above 2, a plain own-slot store
where a neighbour may push into the same arc pair -- fails the validity or
result checks on these seeds (test_planted_mistakes_fail).
"""
import numpy as np
import pytest

import wtv_synth as S

SENT = 1 << 30


class Graph:
    def __init__(self, nbr, res, term):
        self.nbr = np.asarray(nbr, np.int64)
        self.n = self.nbr.shape[0]
        self.res = np.asarray(res, np.int64).copy()
        self.term = np.asarray(term, np.int64).copy()
        self.back = np.full((self.n, 4), -1, np.int64)
        for v in range(self.n):
            for k in range(4):
                u = self.nbr[v, k]
                if u >= 0:
                    self.back[v, k] = int(np.flatnonzero(self.nbr[u] == v)[0])
        self.excess = np.where(self.term > 0, self.term, 0)
        self.height = np.full(self.n, SENT, np.int64)
        self.sink_flow = 0

    def hval(self, h):
        return self.n + 1 if h >= SENT else h

    def bfs(self):
        """Exact distances to the sink over residual arcs (unreached: SENT)."""
        n = self.n
        d = np.full(n, SENT, np.int64)
        front = [v for v in range(n) if self.term[v] < 0]
        for v in front:
            d[v] = 1
        lvl = 1
        while front:
            nxt = []
            for u in front:
                for k in range(4):
                    v = self.nbr[u, k]
                    if v < 0 or d[v] != SENT:
                        continue
                    if self.res[v, self.back[u, k]] > 0:  # residual arc v -> u
                        d[v] = lvl + 1
                        nxt.append(v)
            front = nxt
            lvl += 1
        return d

    def active(self, v):
        return self.excess[v] > 0 and self.height[v] < self.n

    def check_valid(self):
        for v in range(self.n):
            hv = self.hval(self.height[v])
            if self.term[v] < 0:
                assert hv <= 1, f"vertex {v} with sink capacity at height {hv}"
            for k in range(4):
                u = self.nbr[v, k]
                if u >= 0 and self.res[v, k] > 0:
                    assert hv <= self.hval(self.height[u]) + 1, f"residual arc {v}->{u}: {hv} > {self.height[u]} + 1"


class Visit:
    """One fused visit of v (the arithmetic of fused_vertex)."""

    def __init__(self, g, v, cap, push_again, atomic_own):
        self.g, self.v, self.cap, self.push_again, self.atomic_own = g, v, cap, push_again, atomic_own
        self.keep = False
        self.cand = []

    def load_own(self):
        g, v = self.g, self.v
        self.e = int(g.excess[v])
        self.h = int(g.height[v])
        self.r = [int(x) for x in g.res[v]]
        self.t = int(g.term[v])

    def load_heights(self):
        g, v = self.g, self.v
        self.hu = [int(g.height[u]) if u >= 0 else SENT for u in g.nbr[v]]
        self.compute()

    def compute(self):
        g, v, n = self.g, self.v, self.g.n
        e, h, r, t, hu = self.e, self.h, self.r, self.t, self.hu
        self.act = e > 0 and h < n
        self.dl = [0, 0, 0, 0]
        self.pushed = 0
        self.sdlt = 0
        self.new_t = t
        self.hf = h
        if not self.act:
            return
        if t < 0 and h == 1:
            dlt = min(e, -t)
            self.new_t = t + dlt
            e -= dlt
            self.pushed += dlt
            self.sdlt = dlt
        for k in range(4):
            u = g.nbr[v, k]
            if e > 0 and u >= 0 and r[k] and h == hu[k] + 1:
                d = min(r[k], e)
                self.dl[k] = d
                e -= d
                self.pushed += d
        if e > 0:
            tt = self.new_t
            mh = 0 if tt < 0 else (n if tt > 0 else SENT)
            for k in range(4):
                if g.nbr[v, k] >= 0 and r[k] - self.dl[k] != 0:
                    y = n + 1 if hu[k] >= SENT else hu[k]
                    mh = min(mh, y)
            if mh != SENT:
                self.hf = min(mh + 1, h + self.cap)
                if self.push_again:
                    for k in range(4):
                        u = g.nbr[v, k]
                        rk = r[k] - self.dl[k]
                        if e > 0 and u >= 0 and rk and self.hf == hu[k] + 1:
                            d2 = min(rk, e)
                            self.dl[k] += d2
                            e -= d2
                            self.pushed += d2

    def write_stores(self):
        g, v = self.g, self.v
        if not self.act:
            return
        if self.hf > self.h:
            g.height[v] = self.hf
        if self.sdlt:
            g.term[v] = self.new_t
            g.sink_flow += self.sdlt

    def write_res(self):
        g, v = self.g, self.v
        for k in range(4):
            if self.dl[k]:
                if self.atomic_own:
                    g.res[v, k] -= self.dl[k]
                else:
                    g.res[v, k] = self.r[k] - self.dl[k]  # plain store of the loaded value
                g.res[g.nbr[v, k], g.back[v, k]] += self.dl[k]

    def write_excess(self):
        g, v = self.g, self.v
        for k in range(4):
            if self.dl[k]:
                u = g.nbr[v, k]
                old = int(g.excess[u])
                g.excess[u] += self.dl[k]
                if old == 0:
                    self.cand.append(int(u))
        own = int(g.excess[v])
        if self.pushed:
            g.excess[v] -= self.pushed
        self.keep = self.act and (self.pushed == 0 or own != self.pushed) and self.hf < g.n


def sub_round(g, lst, rng, cap=2, push_again=True, atomic_own=True):
    visits = [Visit(g, v, cap, push_again, atomic_own) for v in lst]
    # per visit: its two loads in order, then its three write groups in a
    # random order of their own; visits interleave at random
    queues = []
    for vi in visits:
        writes = [vi.write_stores, vi.write_res, vi.write_excess]
        rng.shuffle(writes)
        queues.append([vi.load_own, vi.load_heights] + writes)
    live = [i for i in range(len(queues))]
    while live:
        j = int(rng.integers(len(live)))
        q = queues[live[j]]
        q.pop(0)()
        if not q:
            live.pop(j)
    out = []
    for vi in visits:  # list order does not matter to the checks
        if vi.keep:
            out.append(vi.v)
        out.extend(vi.cand)
    return out


def solve(nbr, res, term, seed, cap=2, push_again=True, atomic_own=True, relabel_every=(1, 6)):
    g = Graph(nbr, res, term)
    rng = np.random.default_rng(seed)
    g.height = g.bfs()
    lst = [v for v in range(g.n) if g.active(v)]
    subs = 0
    while lst:
        for _ in range(int(rng.integers(*relabel_every))):
            lst = sub_round(g, lst, rng, cap, push_again, atomic_own)
            subs += 1
            assert subs < 20000, "no termination"
            g.check_valid()
            assert len(lst) == len(set(lst)), "a vertex listed twice"
            listed = set(lst)
            for v in range(g.n):
                if g.active(v):
                    assert v in listed, f"active vertex {v} not listed"
            if not lst:
                break
        # global relabel: heights only rise, then filter the old list
        d = g.bfs()
        for v in range(g.n):
            assert g.hval(d[v]) >= g.hval(g.height[v]), f"relabel lowers {v}: {g.height[v]} -> {d[v]}"
        g.height = d
        lst = [v for v in lst if g.active(v)]
        for v in range(g.n):
            assert not g.active(v) or v in lst
    d = g.bfs()
    part = (d < SENT).astype(np.uint8)
    return g.sink_flow, part, subs


def _cases(orc, count):
    out = []
    for seed in range(count):
        n = 12 + 4 * (seed % 9)
        nbr, res, term = S.random_graph(n, 4100 + seed, cap_max=3 + seed % 7, term_max=8 + seed % 11)
        a = orc.pr_solve(nbr, res, term)
        if orc.certificate(nbr, res, term, a.res, a.term, a.partition, a.raw_flow) != 0:
            continue  # the reference's source-height rule leaves it short of a maximum flow
        out.append((nbr, res, term, a))
    return out


def test_fused_schedule_keeps_a_valid_labelling(orc):
    cases = _cases(orc, 40)
    assert len(cases) >= 20
    for i, (nbr, res, term, a) in enumerate(cases):
        for rep in range(3):
            flow, part, _ = solve(nbr, res, term, seed=1000 * i + rep)
            assert flow == a.raw_flow
            assert np.array_equal(part, a.partition[: len(part)])


def test_planted_mistakes_fail(orc):
    """A raise cap of 3 breaks the labelling; plain own-slot stores lose flow
    when a neighbour pushes into the same arc pair. Each is caught on some
    seed of the same cases (the schedule argument is not vacuous)."""
    cases = _cases(orc, 40)

    def caught(**kw):
        for i, (nbr, res, term, a) in enumerate(cases):
            for rep in range(6):
                try:
                    flow, part, _ = solve(nbr, res, term, seed=7000 * i + rep, **kw)
                except AssertionError:
                    return True
                if flow != a.raw_flow or not np.array_equal(part, a.partition[: len(part)]):
                    return True
        return False

    assert caught(cap=3)
    assert caught(atomic_own=False)
    # without the second push pass plain own-slot stores are safe again
    assert not caught(push_again=False, atomic_own=False)
