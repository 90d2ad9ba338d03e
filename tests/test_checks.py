"""CPU tests of the test helpers themselves and of the oracle's thread rule.

* knn_checks.assert_complete accepts the exact answer and a swap at a
  near-tie, and rejects planted mistakes (a far row in place of a near one, a
  row dropped for a later one, an unsorted row);
* oracle.effective_n_threads follows the reference's
  _openmp_effective_n_threads (src/_pdist_aggregation.pyx:41-49): with
  OMP_NUM_THREADS set it is that value, without it min(OpenMP maximum,
  physical cores) -- checked in child processes so the environment is the
  one OpenMP starts with.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from knn_checks import assert_complete

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _case():
    rs = np.random.RandomState(3)
    Y = rs.rand(2000, 6)
    X = rs.rand(50, 6)
    d, i = oracle.brute_force_knn(X, Y, 12)
    return X, Y, d, i


def test_assert_complete_accepts_exact_answer():
    X, Y, _, i = _case()
    assert assert_complete(X, Y, i, 12, 1e-10) == 50


def test_assert_complete_accepts_near_tie_swap():
    X, Y, _, i = _case()
    Y = Y.copy()
    # make train row j a near-twin of query 0's 12th neighbour; either may be returned
    j = int(np.setdiff1d(np.arange(2000), i[0])[0])
    Y[j] = Y[i[0, 11]] + 1e-9
    d2, i2 = oracle.brute_force_knn(X, Y, 12)
    swapped = i2.copy()
    other = j if i2[0, 11] != j else int(i[0, 11])
    swapped[0, 11] = other
    assert_complete(X, Y, swapped, 12, 1e-6)


def test_assert_complete_rejects_far_row():
    X, Y, d, i = _case()
    far = i.copy()
    dall = ((X[0][None] - Y) ** 2).sum(1)
    far[0, 3:] = np.argsort(dall)[-9:][::-1]  # rows 3..11 replaced by the farthest, ascending
    far[0] = far[0][np.argsort(((X[0][None] - Y[far[0]]) ** 2).sum(1))]
    with pytest.raises(AssertionError, match="miss a train row"):
        assert_complete(X, Y, far, 12, 1e-10)


def test_assert_complete_rejects_skipped_row():
    X, Y, _, i = _case()
    d13, i13 = oracle.brute_force_knn(X, Y, 13)
    skip = i13[:, [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12]]  # the 12th nearest left out
    with pytest.raises(AssertionError, match="miss a train row"):
        assert_complete(X, Y, skip, 12, 1e-10)


def test_assert_complete_rejects_unsorted_row():
    X, Y, _, i = _case()
    bad = i.copy()
    bad[5, [0, 11]] = bad[5, [11, 0]]
    with pytest.raises(AssertionError, match="not sorted"):
        assert_complete(X, Y, bad, 12, 1e-10)


def _threads_in_child(env):
    code = "import oracle; print(oracle.effective_n_threads())"
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, check=True,
                         capture_output=True, text=True)
    return int(out.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("n", [1, 3, 7])
def test_effective_n_threads_follows_omp_num_threads(n):
    env = dict(os.environ, OMP_NUM_THREADS=str(n))
    assert _threads_in_child(env) == n


def test_effective_n_threads_without_omp_num_threads():
    env = {k: v for k, v in os.environ.items() if k != "OMP_NUM_THREADS"}
    from joblib import cpu_count

    expect = max(1, min(len(os.sched_getaffinity(0)), cpu_count(only_physical_cores=True)))
    # OpenMP's default maximum is the usable CPU count of the process
    assert _threads_in_child(env) == expect
