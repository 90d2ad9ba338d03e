"""The uncertified- and widened-query counts returned in the stats (needs a
GPU).

Each query's final merge stores a per-query flag and small kernels count the
flags afterwards.  A query whose sampled guess failed is merged twice (the
first merge stops at the failure check, the re-search's merge certifies or
widens it), so it must be counted exactly once; counts span several counting
workgroups and a tail that is not a multiple of 16 queries.
"""
import numpy as np
import pytest

import pdist_aggregation_amd as pa

pytestmark = pytest.mark.gpu


def _stats(Y, X, k, precision):
    _, _, st = pa.parallel_knn(Y, X, k, return_distance=True, squared=True,
                               precision=precision, return_stats=True)
    return st


def test_exact_many_queries_all_certified(require_gpu):
    rs = np.random.RandomState(5)
    Y = rs.rand(2000, 32).astype(np.float32)
    X = rs.rand(70001, 32).astype(np.float32)
    st = _stats(Y, X, 5, "exact")
    assert st["n_uncertified"] == 0


def test_fp8_count_once_per_query(require_gpu, monkeypatch):
    """FP8 E4M3 at d=128 on uniform data: the rounding-error bound is wider
    than the gap between the k-th and k'-th distances, so (nearly) every
    query is widened to its bound-following row set, after which none is
    uncertified; forcing every third guess to fail sends those queries
    through the re-search and its merge without changing either count."""
    rs = np.random.RandomState(6)
    Y = rs.rand(40000, 128).astype(np.float32)
    X = rs.rand(3001, 128).astype(np.float32)
    st = _stats(Y, X, 10, "fp8")
    n = st["n_widened"]
    assert 0.9 * X.shape[0] <= n <= X.shape[0]
    assert st["n_uncertified"] == 0
    assert st["n_rescued"] == 0
    monkeypatch.setenv("PDIST_KNN_GUESS", "fail:3")
    st2 = _stats(Y, X, 10, "fp8")
    assert st2["n_rescued"] == (X.shape[0] + 2) // 3
    assert st2["n_widened"] == n
    assert st2["n_uncertified"] == 0
