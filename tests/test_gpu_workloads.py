"""The low-precision workloads at full size (needs a GPU; marked slow).

* BF16: a mixture of 100 Gaussians (centres U[-1,1]^128, sigma 0.1), 1e5
  queries x 1e6 train rows, k=32, k'=128.
* FP8 E4M3: 1,000 Gaussians (centres U[-1,1]^256, sigma 0.05), 5e4 queries x
  2e6 train rows with 20,000 train rows overwritten by copies of other rows
  (exact ties), k=100, k'=400, plus the float32 backward dL/dX.  The 3-bit
  mantissa makes the error bound wider than a Gaussian's spread, so no query
  is certified from its k'=400 candidates (their recall@100 is about 0.986):
  every query is widened to all rows within the bound of its k-th exact
  distance (about 2,000 rows, the query's whole Gaussian) and re-ranked
  exactly.

Both draw from numpy default_rng(seed) in the order centres, train labels,
train noise, query labels, query noise.  Checked on every row: ascending,
unique, in range; on a query subsample: recall@k against the FP64 oracle,
distances equal to the exact FP64 recomputation, index parity on every
query of the subsample (near-ties exempt), and no query uncertified.
"""
import sys

import numpy as np
import pytest

import oracle
import pdist_aggregation_amd as pa
from knn_checks import EPS_F32, assert_knn_parity
from pdist_aggregation_amd.autograd import knn_distances

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mixture(seed, n_centres, d, sigma, n_train, n_test):
    g = np.random.default_rng(seed)
    centres = g.uniform(-1.0, 1.0, (n_centres, d)).astype(np.float32)
    Y = centres[g.integers(0, n_centres, n_train)]
    Y += sigma * g.standard_normal((n_train, d), dtype=np.float32)
    X = centres[g.integers(0, n_centres, n_test)]
    X += sigma * g.standard_normal((n_test, d), dtype=np.float32)
    return g, Y, X


def _check_rows(dist, idx, n_train):
    assert idx.min() >= 0 and idx.max() < n_train
    assert np.all(np.diff(dist, axis=1) >= 0), "rows not sorted ascending"
    s = np.sort(idx, axis=1)
    assert np.all(s[:, 1:] != s[:, :-1]), "duplicate index in a row"


def _check_subsample(X, Y, dist, idx, k, sub, st):
    ref = oracle.exact_knn_f64(X[sub], Y, k, return_boundary=True)
    rec = oracle.recall_at_k(idx[sub], ref[1])
    got = oracle._direct_sqdist(X[sub], Y, idx[sub])
    np.testing.assert_allclose(dist[sub], got, rtol=1e-6)
    err = np.max(np.abs(dist[sub].astype(np.float64) - ref[0]) / np.maximum(ref[0], 1e-30))
    _say(f"    recall@{k} on {len(sub)} queries = {rec:.5f}, max rel dist err {err:.2e}; "
         f"stats {st}")
    assert st["n_uncertified"] == 0
    assert_knn_parity(X[sub], Y, dist[sub], idx[sub], k, EPS_F32, ref=ref)
    assert rec >= 0.999


def test_bf16_mixture_1e5_by_1e6(require_gpu):
    torch = require_gpu
    _, Y, X = _mixture(4, 100, 128, 0.1, 1_000_000, 100_000)
    _say("[bf16 workload] 100000 x 1000000 x 128, k=32, k'=128")
    Yd, Xd = torch.from_numpy(Y).cuda(), torch.from_numpy(X).cuda()
    (dist, idx), _, st = pa.parallel_knn(Yd, Xd, 32, return_distance=True, squared=True,
                                         precision="bf16", oversample=128, return_stats=True)
    dist, idx = dist.cpu().numpy(), idx.cpu().numpy()
    _check_rows(dist, idx, Y.shape[0])
    _check_subsample(X, Y, dist, idx, 32, np.arange(0, 100_000, 200), st)


def test_fp8_mixture_with_duplicates_5e4_by_2e6(require_gpu):
    torch = require_gpu
    g, Y, X = _mixture(5, 1000, 256, 0.05, 2_000_000, 50_000)
    dst = g.choice(Y.shape[0], 20_000, replace=False)
    Y[dst] = Y[g.integers(0, Y.shape[0], 20_000)]
    _say("[fp8 workload] 50000 x 2000000 x 256 with 20000 duplicated rows, k=100, k'=400")
    Yd, Xd = torch.from_numpy(Y).cuda(), torch.from_numpy(X).cuda()
    (dist, idx), _, st = pa.parallel_knn(Yd, Xd, 100, return_distance=True, squared=True,
                                         precision="fp8", oversample=400, return_stats=True)
    dist, idx = dist.cpu().numpy(), idx.cpu().numpy()
    _check_rows(dist, idx, Y.shape[0])
    _check_subsample(X, Y, dist, idx, 100, np.arange(0, 50_000, 250), st)
    # float32 backward: dL/dX of L = sum of squared k-NN distances is
    # 2 * sum_j (x - y_j) over the returned neighbours (the same search).
    Xg = Xd.clone().requires_grad_(True)
    dist_t, idx_t = knn_distances(Yd, Xg, 100, squared=True, precision="fp8", oversample=400)
    dist_t.sum().backward()
    np.testing.assert_array_equal(idx_t.cpu().numpy(), idx)
    sub = np.arange(0, 50_000, 97)
    diff = X[sub].astype(np.float64)[:, None, :] - Y.astype(np.float64)[idx[sub]]
    np.testing.assert_allclose(Xg.grad.cpu().numpy()[sub], 2.0 * diff.sum(1), rtol=1e-4,
                               atol=1e-4)
