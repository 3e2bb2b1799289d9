"""The autotuner's search logic (CPU, with a synthetic cost model)."""
import pytest

import wtv_autotune as AT


def test_coordinate_descent_finds_separable_minimum():
    best_vals = {"relabel_alpha": 16.0, "local_steps": 32, "tail_cap": 512, "grid_wgs": 2048, "wg_size": 1024}

    def run(p):
        cost = sum(abs(AT.SPACE[k].index(p[k]) - AT.SPACE[k].index(best_vals[k])) for k in p)
        return 1.0 + cost, "same-answer"

    best, t, trials = AT.tune(run)
    assert best == best_vals and t == 1.0
    assert len(trials) == 1 + sum(len(v) - 1 for v in AT.SPACE.values())


def test_answer_change_is_an_error():
    calls = []

    def run(p):
        calls.append(p)
        return 1.0, len(calls) > 3  # the "answer" flips after three runs

    with pytest.raises(RuntimeError, match="answer changed"):
        AT.tune(run)
