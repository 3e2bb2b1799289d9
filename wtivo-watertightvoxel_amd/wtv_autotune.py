"""Autotuner for the solver's schedule knobs.

Tunes, by coordinate descent on measured solve time, the knobs that change
the schedule but never the answer:
  * relabel_alpha  -- global-relabel frequency (relabel work between relabels)
  * local_steps    -- sub-rounds per macro round (the discharge count between
                      host checks, and the replayed graph's length)
  * tail_cap       -- active-list size handed to the one-workgroup discharge
  * grid_wgs       -- workgroups per discharge / BFS launch
  * wg_size        -- threads per workgroup of those launches
Every candidate must reproduce the reference flows exactly (flow and labels do
not depend on the schedule; a mismatch raises).  Results can be cached in a
JSON file keyed by a workload signature.
"""
from __future__ import annotations

import json
import os
import time

DEFAULT = {"relabel_alpha": 2.0, "local_steps": 8, "tail_cap": 512, "grid_wgs": 2048, "wg_size": 256}
SPACE = {
    "relabel_alpha": [2.0, 4.0, 8.0, 16.0, 32.0],
    "local_steps": [4, 8, 16, 32],
    "tail_cap": [-1, 512, 2048, 8192],
    "grid_wgs": [1024, 2048, 4096],
    "wg_size": [256, 512, 1024],
}
CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_params.json")


def tune(run, reference=None, space=None, start=None, reps=1, passes=1, log=None):
    """Coordinate descent.  run(params) -> (seconds, answer); answer is compared
    with `reference` (or the first run's answer).  Returns (best_params,
    best_seconds, trials) where trials is a list of (params, seconds)."""
    space = space or SPACE
    best = dict(start or DEFAULT)
    trials = []

    def measure(p):
        ts = []
        ans = None
        for _ in range(reps):
            t, ans = run(p)
            ts.append(t)
        nonlocal reference
        if reference is None:
            reference = ans
        elif ans != reference:
            raise RuntimeError(f"autotune: answer changed with {p}")
        t = min(ts)
        trials.append((dict(p), t))
        if log:
            log(f"[tune] {p} -> {t * 1e3:.1f} ms")
        return t

    best_t = measure(best)
    for _ in range(passes):
        for key, values in space.items():
            tried = best[key]
            for v in values:
                if v == tried:
                    continue
                cand = dict(best)
                cand[key] = v
                t = measure(cand)
                if t < best_t:
                    best, best_t = cand, t
    return best, best_t, trials


def load(signature: str):
    try:
        with open(CACHE) as f:
            return json.load(f).get(signature)
    except (OSError, ValueError):
        return None


def save(signature: str, params: dict, seconds: float):
    data = {}
    try:
        with open(CACHE) as f:
            data = json.load(f)
    except (OSError, ValueError):
        pass
    data[signature] = {"params": params, "seconds": seconds, "time": time.time()}
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)
