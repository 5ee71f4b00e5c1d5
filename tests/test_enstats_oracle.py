"""Pins of the ensemble-statistics restatement (oracle/enstats_oracle.py)
against what Sources/task_PosEns.F90 fixes on hand-checkable inputs."""
import numpy as np
import pytest

from oracle import enstats_oracle as E


def test_hand_example_five_members():
    # one point, members 4, 1, 5, 2, 3 -> ascending 1 2 3 4 5
    x = np.array([[4.0], [1.0], [5.0], [2.0], [3.0]])
    assert E.mean(x)[0] == 3.0
    assert E.median(x)[0] == 2.0          # ens_rank(int(5/2)) = ens_rank(2): the lower neighbour of the middle
    assert E.stddev(x)[0] == pytest.approx(np.sqrt(10.0 / 4.0), rel=1e-15)
    # int(50/100*5) = 2 -> 2; int(90/100*5) = 4 -> 4; int(10/100*5) = 0 -> clamped to 1; 100 -> 5
    np.testing.assert_array_equal(E.percentiles(x, [50.0, 90.0, 10.0, 100.0])[:, 0], [2.0, 4.0, 1.0, 5.0])
    np.testing.assert_array_equal(E.probabilities(x, [3.0, 3.5, 0.0, 6.0])[:, 0], [60.0, 40.0, 100.0, 0.0])
    assert E.logmean(x)[0] == pytest.approx(120.0 ** 0.2, rel=1e-14)  # geometric mean


def test_even_member_count_median_is_lower_middle():
    x = np.array([[10.0], [40.0], [20.0], [30.0]])
    assert E.median(x)[0] == 20.0  # ens_rank(int(4/2)) = ens_rank(2)


def test_constant_members_and_invariance():
    rng = np.random.default_rng(0)
    x = np.full((7, 11), 2.5)
    assert np.all(E.mean(x) == 2.5) and np.all(E.stddev(x) == 0.0) and np.all(E.median(x) == 2.5)
    y = rng.uniform(0.1, 9.0, (9, 50))
    perm = rng.permutation(9)
    for f in (E.median, E.stddev):
        np.testing.assert_allclose(f(y), f(y[perm]), rtol=1e-13)
    np.testing.assert_array_equal(E.percentiles(y, [30, 70]), E.percentiles(y[perm], [30, 70]))
    np.testing.assert_allclose(E.stddev(y), y.std(axis=0, ddof=1), rtol=1e-12)
    # brute-force rank: the value with exactly k-1 members below it (distinct values)
    k = max(9 // 2, 1)
    brute = np.array([sorted(y[:, p])[k - 1] for p in range(50)])
    np.testing.assert_array_equal(E.median(y), brute)
