"""Host logic of bench.py (no GPU): which frequency each timed step solves."""
import numpy as np
import pytest

import bench


@pytest.mark.parametrize("steps", [1, 20, 40, 80])
def test_step_frequencies(steps):
    nf = 40
    idx = bench.step_frequencies(steps, nf)
    assert len(idx) == steps
    assert idx.min() >= 0 and idx.max() < nf
    if steps <= nf:
        # distinct, increasing, first frequency included, spread over the whole range
        assert np.all(np.diff(idx) > 0)
        assert idx[0] == 0
        assert idx[-1] >= nf - nf // steps
    else:
        # every frequency solved steps // nf times, in order
        assert np.array_equal(np.bincount(idx, minlength=nf), np.full(nf, steps // nf))
        assert np.array_equal(idx[:nf], np.arange(nf))


def test_step_frequencies_twenty_of_forty():
    # the round-end command (--steps 20): every other frequency, 0.05 .. 2.4375 rad/s
    idx = bench.step_frequencies(20, 40)
    assert np.array_equal(idx, np.arange(0, 40, 2))


def test_errors_reported_per_matrix():
    g = [dict(A=np.eye(6), B=np.eye(6) * 2, F=np.ones((6, 2)))]
    r = [dict(A=np.eye(6) * (1 + 1e-7), B=np.eye(6) * 2, F=np.ones((6, 2)) * (1 - 1e-9))]
    e = bench.abf_errs(g, r)
    assert set(e) == {"A", "B", "F"}
    assert abs(e["A"] - 1e-7 / (1 + 1e-7)) < 1e-12 and e["B"] == 0.0 and abs(e["F"] - 1e-9) < 1e-12
