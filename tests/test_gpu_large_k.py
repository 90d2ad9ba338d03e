"""k above the merge limit of 1,984 (needs a GPU).

The reference takes any k <= n_train (src/_pdist_aggregation.pyx:503-507,
its heap has k slots whatever k is).  Here k > 1,984 runs the large-k path
(exact FP64 distance rows per chunk of queries, radix select of the k-th
(distance, index) key, rank sort), compared with the FP64 oracle at k =
1,985, 2,000, 5,000 and k = n_train = 10,000, in float32 and float64, with
exact ties straddling the k-th rank; and low-precision requests beyond the
BF16/FP8 kernels (k' > 478 or d > 256) run the exact search instead of
failing.
"""
import sys

import numpy as np
import pytest

import oracle
import pdist_aggregation_amd as pa
from knn_checks import EPS_F32, EPS_F64, assert_knn_parity, assert_sorted_unique

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("k,d", [(1985, 8), (2000, 64), (5000, 5), (10000, 32)])
def test_large_k_matches_oracle(require_gpu, dtype, k, d):
    rs = np.random.RandomState(k + d)
    Y = rs.rand(10000, d).astype(dtype)
    X = rs.rand(300, d).astype(dtype)
    _say(f"[large k] k={k} d={d} {np.dtype(dtype).name}")
    (dist, idx), n_chunks, st = pa.parallel_knn(Y, X, k, return_distance=True, squared=True,
                                                return_stats=True)
    assert st["kernel_path"] == 5 and st["n_uncertified"] == 0
    assert dist.shape == (300, k) and dist.dtype == dtype
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    eps = EPS_F32 if dtype == np.float32 else EPS_F64
    assert_knn_parity(X, Y, dist, idx, k, eps)
    if k == Y.shape[0]:
        assert np.all(np.sort(idx, axis=1) == np.arange(k)[None, :])
    # Euclidean by default, like the reference
    (de, ie), _ = pa.parallel_knn(Y, X, k, return_distance=True)
    np.testing.assert_array_equal(ie, idx)
    np.testing.assert_allclose(de, np.sqrt(dist.astype(np.float64)), rtol=1e-6)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_large_k_ties_lower_index_first(require_gpu, dtype):
    """Rows repeated so that exact ties straddle rank k = 2,500."""
    rs = np.random.RandomState(3)
    base = rs.rand(1500, 6).astype(dtype)
    Y = base[rs.randint(0, 1500, 8000)]  # every row repeated about 5 times
    X = rs.rand(64, 6).astype(dtype)
    k = 2500
    (dist, idx), _ = pa.parallel_knn(Y, X, k, return_distance=True, squared=True)
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    ref = oracle.exact_knn_f64(X, Y, k, return_boundary=True)
    assert_knn_parity(X, Y, dist, idx, k, 1e-12, ref=ref)
    np.testing.assert_array_equal(idx, ref[1])


@pytest.mark.parametrize("precision,d,k", [("bf16", 64, 600), ("fp8", 128, 200),
                                           ("bf16", 300, 10), ("fp8", 2048, 5)])
def test_low_precision_beyond_kernel_limits_runs_exact(require_gpu, precision, d, k):
    rs = np.random.RandomState(d + k)
    Y = rs.rand(6000, d).astype(np.float32)
    X = rs.rand(100, d).astype(np.float32)
    (dist, idx), _, st = pa.parallel_knn(Y, X, k, return_distance=True, squared=True,
                                         precision=precision, return_stats=True)
    assert st["kernel_path"] in (0, 1)
    assert st["n_uncertified"] == 0
    assert_knn_parity(X, Y, dist, idx, k, EPS_F32)
