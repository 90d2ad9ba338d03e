"""The reference test's whole parameter grid on the GPU (needs a GPU).

tests/test_correctness.py:8-13 of the reference sweeps n in {100, 1000,
10000} train rows, d in {2, 5, 10, 100}, n_train / n_test in {10, 2, 1, 0.5}
and k in {1, 10, 100, 1000}, skipping n < k (176 points), with the data
recipe np.random.seed(1), train drawn first (:29-33), float64, and compares
parallel_knn(..., return_distance=True) with sklearn's brute force: indices
exactly, Euclidean distances to 7 decimals (:35-49).

Here every point runs in float64 and in float32 (the same values rounded):
* the first 1,000 queries against the FP64 oracle (oracle.exact_knn_f64):
  indices identical outside near-ties, distances to 7 decimals (float64) or
  1e-5 relative (float32);
* every query through size-independent properties (knn_checks.assert_complete:
  rows sorted by exact distance, no train row outside the answer clearly
  closer than the k-th);
* on a subset of points, every chunk size of the reference grid {512, 1024,
  2048, 4096} x every strategy gives the same rows and the reference's
  n_parallel_chunks closed form.
"""
import sys

import numpy as np
import pytest

import oracle
import pdist_aggregation_amd as pa
from knn_checks import EPS_F32, EPS_F64, assert_complete, assert_knn_parity

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

NS = (100, 1000, 10000)
DS = (2, 5, 10, 100)
RATIOS = (10, 2, 1, 0.5)
KS = (1, 10, 100, 1000)
POINTS = [(n, d, r, k) for n in NS for d in DS for r in RATIOS for k in KS if n >= k]
assert len(POINTS) == 176
ORACLE_QUERIES = 1000
# every chunk size x strategy on every point with n = 1000 and d = 10
SWEEP = {p for p in POINTS if p[0] == 1000 and p[1] == 10}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _inputs(n, d, ratio):
    """tests/test_correctness.py:29-33 of the reference."""
    np.random.seed(1)
    X_train = np.random.rand(int(n * d)).astype(np.float64).reshape((-1, d))
    X_test = np.random.rand(int(n * d / ratio)).astype(np.float64).reshape((-1, d))
    return X_train, X_test


@pytest.mark.parametrize("n,d,ratio,k", POINTS,
                         ids=[f"n{n}-d{d}-r{r}-k{k}" for n, d, r, k in POINTS])
def test_reference_grid_point(require_gpu, n, d, ratio, k):
    X_train, X_test = _inputs(n, d, ratio)
    _say(f"[grid] n={n} d={d} ratio={ratio} k={k}: {X_train.shape} x {X_test.shape}")
    for dtype, eps in ((np.float64, EPS_F64), (np.float32, EPS_F32)):
        Y = X_train.astype(dtype)
        X = X_test.astype(dtype)
        (dist, idx), n_chunks = pa.parallel_knn(Y, X, k, return_distance=True)
        assert dist.dtype == dtype and idx.dtype == np.int64
        assert dist.shape == (X.shape[0], k)
        sub = slice(0, ORACLE_QUERIES)
        ref = oracle.exact_knn_f64(X[sub], Y, k, return_boundary=True)
        assert_knn_parity(X[sub], Y, dist[sub], idx[sub], k, eps, squared=False,
                          dist_rtol=1e-12 if dtype == np.float64 else 1e-5, ref=ref)
        if dtype == np.float64:  # the reference test's own distance check
            mine = idx[sub] == ref[1]
            np.testing.assert_almost_equal(dist[sub][mine], np.sqrt(ref[0])[mine])
        assert_complete(X, Y, idx, k, eps)
        assert n_chunks == pa.n_parallel_chunks("auto", Y.shape[0], X.shape[0], 256,
                                                pa.effective_n_threads())
        if (n, d, ratio, k) in SWEEP:
            for chunk in (512, 1024, 2048, 4096):
                for strategy in ("auto", "chunk_on_train", "chunk_on_test"):
                    (d2, i2), nc = pa.parallel_knn(Y, X, k, chunk_size=chunk, strategy=strategy,
                                                   return_distance=True)
                    np.testing.assert_array_equal(i2, idx)
                    np.testing.assert_array_equal(d2, dist)
                    assert nc == oracle.n_parallel_chunks(strategy, Y.shape[0], X.shape[0],
                                                          chunk, pa.effective_n_threads())
