"""Parity of the HIP path with the oracle, through the C ABI (needs a GPU).

* the reference's CPU-runnable workload (1,000 x 10,000 x 50, float64, k=10)
  and a subset of the reference test grid, against the committed sklearn
  fixtures, for every strategy and several chunk sizes (the result must not
  depend on them, as in tests/test_correctness.py of the reference);
* float32 inputs against the exact FP64 oracle, near-ties exempt and counted;
* edge cases: exact ties from duplicate rows, k = n_train, one query, d from 1
  to 300, sizes that are not multiples of any tile;
* the 10,000 x 1,000,000 x 128 float32 workload: a 1,000-query subsample
  against the FP64 oracle and size-independent properties on every row.
"""
import sys

import numpy as np
import pytest

import oracle
import pdist_aggregation_amd as pa
from conftest import golden_manifest, load_golden
from knn_checks import EPS_F32, EPS_F64, assert_knn_parity, assert_sorted_unique

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", sorted(golden_manifest()))
def test_f64_matches_sklearn_fixture(require_gpu, name):
    X_train, X_test, k, sk_dist, sk_idx = load_golden(name)
    _say(f"[f64 fixture] {name} train={X_train.shape} test={X_test.shape} k={k}")
    ref = None
    for strategy in ("auto", "chunk_on_train", "chunk_on_test"):
        for chunk in (256, 4096):
            (dist, idx), n_chunks = pa.parallel_knn(X_train, X_test, k, chunk_size=chunk,
                                                    strategy=strategy, return_distance=True)
            assert dist.dtype == np.float64 and idx.dtype == np.int64
            np.testing.assert_almost_equal(dist, sk_dist)
            if not np.array_equal(idx, sk_idx):
                if ref is None:
                    ref = oracle.exact_knn_f64(X_test, X_train, k, return_boundary=True)
                assert_knn_parity(X_test, X_train, dist, idx, k, EPS_F64, squared=False,
                                  ref=ref)
            assert n_chunks == pa.n_parallel_chunks(strategy, X_train.shape[0],
                                                    X_test.shape[0], chunk,
                                                    pa.effective_n_threads())


def test_workload_cpu_case_squared_f64(require_gpu):
    """1,000 x 10,000 x 50 float64, k=10, chunk 256, squared distances."""
    X_train, X_test, k, _, sk_idx = load_golden("workload_1000x10000x50_k10")
    (dist, idx), _ = pa.parallel_knn(X_train, X_test, k, chunk_size=256, return_distance=True,
                                     squared=True)
    np.testing.assert_array_equal(idx, sk_idx)
    ref_d, ref_i = oracle.exact_knn_f64(X_test, X_train, k)
    np.testing.assert_allclose(dist, ref_d, rtol=1e-12)
    # the C++ restatement of the reference engine agrees too
    (pd, pi), _ = oracle.port_parallel_knn(X_train, X_test, k, chunk_size=256,
                                           return_distance=True)
    np.testing.assert_array_equal(idx, pi)
    np.testing.assert_allclose(np.sqrt(dist), pd, rtol=1e-12)


@pytest.mark.parametrize("name", sorted(golden_manifest()))
def test_f32_matches_oracle(require_gpu, name):
    X_train, X_test, k, _, _ = load_golden(name)
    Y = X_train.astype(np.float32)
    X = X_test.astype(np.float32)
    _say(f"[f32] {name}")
    (dist, idx), _ = pa.parallel_knn(Y, X, k, return_distance=True, squared=True)
    assert dist.dtype == np.float32
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    amb = assert_knn_parity(X, Y, dist, idx, k, EPS_F32, squared=True)
    _say(f"    ambiguous ranks: {amb}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 2, 3, 4, 7, 8, 12, 16, 50, 128, 129, 300])
def test_dims_and_ragged_sizes(require_gpu, dtype, d):
    rs = np.random.RandomState(d)
    Y = rs.rand(1237, d).astype(dtype)
    X = rs.rand(131, d).astype(dtype)
    _say(f"[dims] d={d} {np.dtype(dtype).name}")
    for k in (1, 5, 64):
        (dist, idx), _ = pa.parallel_knn(Y, X, k, return_distance=True, squared=True)
        assert_knn_parity(X, Y, dist, idx, k, EPS_F32 if dtype == np.float32 else EPS_F64)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_exact_ties_lower_index_first(require_gpu, dtype):
    rs = np.random.RandomState(0)
    Y = rs.rand(5000, 8).astype(dtype)
    dup = np.arange(11, 5000, 37)
    Y[dup] = Y[7]
    X = rs.rand(70, 8).astype(dtype)
    X[0] = Y[7]
    X[1] = Y[7] + dtype(1e-3)
    k = 40
    (dist, idx), _ = pa.parallel_knn(Y, X, k, return_distance=True, squared=True)
    expect = np.concatenate([[7], dup])[:k]
    np.testing.assert_array_equal(idx[0], expect)
    assert np.all(dist[0] == 0)
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dist, idx, k, 1e-12)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_k_equals_n_train_and_single_query(require_gpu, dtype):
    rs = np.random.RandomState(5)
    Y = rs.rand(300, 5).astype(dtype)
    X = rs.rand(1, 5).astype(dtype)
    (dist, idx), _ = pa.parallel_knn(Y, X, 300, return_distance=True, squared=True)
    assert sorted(idx[0]) == list(range(300))
    assert_knn_parity(X, Y, dist, idx, 300, 1e-6)
    (dist, idx), _ = pa.parallel_knn(Y, X, 1, return_distance=True)
    bd, bi = oracle.brute_force_knn(X, Y, 1)
    assert idx[0, 0] == bi[0, 0]


@pytest.mark.parametrize("dtype,precision,d", [(np.float32, "exact", 64), (np.float32, "exact", 3),
                                               (np.float64, "exact", 50),
                                               (np.float32, "bf16", 128),
                                               (np.float32, "fp8", 128)])
def test_no_queries(require_gpu, dtype, precision, d):
    """An empty query set returns empty (0, k) results on every path."""
    rs = np.random.RandomState(1)
    Y = rs.rand(500, d).astype(dtype)
    X = np.empty((0, d), dtype=dtype)
    (dist, idx), _ = pa.parallel_knn(Y, X, 7, return_distance=True, precision=precision)
    assert dist.shape == (0, 7) and idx.shape == (0, 7)


def test_large_k(require_gpu):
    rs = np.random.RandomState(9)
    Y = rs.rand(3000, 10)
    X = rs.rand(257, 10)
    (dist, idx), _ = pa.parallel_knn(Y, X, 1000, return_distance=True, squared=True)
    assert_knn_parity(X, Y, dist, idx, 1000, EPS_F64)


def test_device_tensors_in_and_out(require_gpu):
    torch = require_gpu
    rs = np.random.RandomState(2)
    Y = rs.rand(4000, 33).astype(np.float32)
    X = rs.rand(300, 33).astype(np.float32)
    Yd = torch.from_numpy(Y).cuda()
    Xd = torch.from_numpy(X).cuda()
    (dist, idx), _ = pa.parallel_knn(Yd, Xd, 7, return_distance=True)
    assert dist.is_cuda and idx.is_cuda and idx.dtype == torch.int64
    (hd, hi), _ = pa.parallel_knn(Y, X, 7, return_distance=True)
    np.testing.assert_array_equal(idx.cpu().numpy(), hi)
    np.testing.assert_array_equal(dist.cpu().numpy(), hd)


def test_estimators_and_sycl_surface(require_gpu):
    from pdist_aggregation_amd import sycl_compat

    rs = np.random.RandomState(4)
    Y = rs.rand(2000, 6)
    X = rs.rand(100, 6)
    ref_d, ref_i = oracle.exact_knn_f64(X, Y, 4)
    for cls in (pa.NearestNeighborsParallelAuto, pa.NearestNeighborsParallelXtrain,
                pa.NearestNeighborsParallelXtest):
        (d, i), n = cls(n_neighbors=4).fit(Y).kneighbors(X, return_distance=True)
        np.testing.assert_array_equal(i, ref_i)
        np.testing.assert_allclose(d, np.sqrt(ref_d), rtol=1e-12)
        i2, n2 = cls(n_neighbors=4).fit(Y).kneighbors(X)
        np.testing.assert_array_equal(i2, ref_i)
        assert n2 == n
    i, d = sycl_compat.NearestNeighbors(n_neighbors=4).fit(Y).kneighbors(X, return_distance=True)
    np.testing.assert_array_equal(i, ref_i)
    np.testing.assert_allclose(d, ref_d, rtol=1e-12)  # squared, as the SYCL backend
    with pytest.raises(ValueError):
        sycl_compat.parallel_knn(Y, X, 0)
    with pytest.raises(ValueError):
        sycl_compat.parallel_knn(Y, X, 3, strategy="chunk_on_train")


def test_input_validation(require_gpu):
    Y = np.random.rand(50, 4)
    X = np.random.rand(5, 4)
    with pytest.raises(ValueError):
        pa.parallel_knn(Y, X, 51)
    with pytest.raises(ValueError):
        pa.parallel_knn(Y, np.random.rand(5, 3), 2)
    with pytest.raises(TypeError):
        pa.parallel_knn(Y, X.astype(np.float32), 2)
    with pytest.raises(ValueError):
        pa.parallel_knn(Y, np.asfortranarray(np.random.rand(5, 4)), 2)
    with pytest.raises(RuntimeError):
        pa.parallel_knn(Y, X, 2, strategy="chunk_on_nothing")


def _bench_scale_inputs():
    """10,000 queries and 1,000,000 train rows, d=128, U[0,1) float32,
    default_rng(0), train drawn first."""
    g = np.random.default_rng(0)
    Y = g.random((1_000_000, 128), dtype=np.float32)
    X = g.random((10_000, 128), dtype=np.float32)
    return Y, X


@pytest.fixture(scope="module")
def headline():
    """The 1e4 x 1e6 x 128 inputs on the GPU and the FP64 oracle's answer
    for queries 0..999 (k = 10, squared distances, boundary distance)."""
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    Y, X = _bench_scale_inputs()
    _say("[full] oracle for queries 0..999 of 10000 x 1000000 x 128 f32, k=10")
    ref = oracle.exact_knn_f64(X[:1000], Y, 10, return_boundary=True)
    return Y, X, torch.from_numpy(Y).cuda(), torch.from_numpy(X).cuda(), ref


@pytest.mark.slow
@pytest.mark.parametrize("precision", ["exact", "bf16", "fp8"])
def test_full_scale_against_oracle(headline, precision):
    """The benchmark workload in every precision: queries 0..999 against the
    FP64 oracle (index parity outside near-ties on every one of them, recall
    1.0 there), every row sorted, unique and equal to its exact distance, and
    no query left uncertified."""
    Y, X, Yd, Xd, ref = headline
    _say(f"[full] 10000 x 1000000 x 128 f32, k=10, {precision}")
    (dist, idx), _, st = pa.parallel_knn(Yd, Xd, 10, chunk_size=4096, return_distance=True,
                                         squared=True, precision=precision, return_stats=True)
    _say(f"    stats {st}")
    dist = dist.cpu().numpy()
    idx = idx.cpu().numpy()
    # every row: sorted, unique, in range, distance = exact recomputation
    assert_sorted_unique(idx, dist, Y.shape[0])
    sub = np.arange(0, 10_000, 97)
    diff = Y[idx[sub]].astype(np.float64) - X[sub][:, None, :].astype(np.float64)
    np.testing.assert_allclose(dist[sub], (diff ** 2).sum(-1), rtol=1e-6)
    q = np.arange(1000)
    amb = assert_knn_parity(X[q], Y, dist[q], idx[q], 10, EPS_F32, squared=True, ref=ref)
    rec = oracle.recall_at_k(idx[q], ref[1])
    _say(f"    recall@10 on 1000 queries = {rec:.5f}; ambiguous ranks {amb}; "
         f"uncertified {st['n_uncertified']}, widened {st['n_widened']}")
    assert st["n_uncertified"] == 0
    assert rec >= 0.999


@pytest.mark.parametrize("d,expect_path", [(16, 1), (32, 1), (48, 1), (64, 1), (100, 1),
                                           (128, 1), (132, 1), (192, 1), (200, 1), (256, 1),
                                           (50, 1), (130, 1), (17, 1), (255, 1), (15, 4),
                                           (260, 0), (301, 0)])
def test_f32_kernel_path_selection(require_gpu, d, expect_path):
    """Float32 with 16 <= d <= 256 runs on the MFMA kernel (d % 4 != 0 through
    zero-padded copies of the rows), d < 16 on the low-d kernel, d > 256 on the
    generic one."""
    rs = np.random.RandomState(d)
    Y = rs.rand(2111, d).astype(np.float32)
    X = rs.rand(301, d).astype(np.float32)
    (dist, idx), _, st = pa.parallel_knn(Y, X, 9, return_distance=True, squared=True,
                                         return_stats=True)
    assert st["kernel_path"] == expect_path
    assert_knn_parity(X, Y, dist, idx, 9, EPS_F32)
    assert st["n_uncertified"] == 0


@pytest.mark.parametrize("d,k", [(256, 100), (256, 400), (192, 250)])
def test_mfma_f32_wide_rows_large_k(require_gpu, d, k):
    """float32 at search widths 192 / 256 with lists near their largest size."""
    rs = np.random.RandomState(d + k)
    Y = rs.rand(12000, d).astype(np.float32)
    X = rs.rand(300, d).astype(np.float32)
    (dist, idx), _, st = pa.parallel_knn(Y, X, k, return_distance=True, squared=True,
                                         return_stats=True)
    assert st["kernel_path"] == 1
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dist, idx, k, EPS_F32)


@pytest.mark.parametrize("k", [1, 10, 100, 400])
def test_mfma_f32_many_k(require_gpu, k):
    rs = np.random.RandomState(k)
    Y = rs.rand(20000, 64).astype(np.float32)
    X = rs.rand(513, 64).astype(np.float32)
    (dist, idx), _, st = pa.parallel_knn(Y, X, k, return_distance=True, squared=True,
                                         return_stats=True)
    assert st["kernel_path"] == 1
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dist, idx, k, EPS_F32)


def _gaussian_mixture(rs, n, d, centres, sigma):
    lab = rs.randint(0, centres.shape[0], n)
    return (centres[lab] + sigma * rs.randn(n, d)).astype(np.float32)


@pytest.mark.parametrize("d", [32, 50, 128, 200])
def test_bf16_rerank_exact_when_certified(require_gpu, d):
    """BF16 candidates + exact re-rank: certified rows equal the FP64 answer;
    recall@k over all rows >= 0.999."""
    rs = np.random.RandomState(100 + d)
    Y = rs.rand(30000, d).astype(np.float32)
    X = rs.rand(400, d).astype(np.float32)
    k = 10
    (dist, idx), _, st = pa.parallel_knn(Y, X, k, return_distance=True, squared=True,
                                         precision="bf16", return_stats=True)
    assert st["kernel_path"] == 2 and st["k_candidates"] == 4 * k
    ref = oracle.exact_knn_f64(X, Y, k, return_boundary=True)
    rec = oracle.recall_at_k(idx, ref[1])
    _say(f"    bf16 d={d}: recall@{k}={rec:.5f} uncertified={st['n_uncertified']}")
    assert rec >= 0.999
    assert st["n_uncertified"] == 0
    assert_knn_parity(X, Y, dist, idx, k, EPS_F32, ref=ref)


def test_bf16_gaussian_mixture_k32(require_gpu):
    """Mixture of Gaussians (centres U[-1,1]^128, sigma 0.1), k=32, k'=128."""
    rs = np.random.RandomState(7)
    centres = rs.uniform(-1, 1, (100, 128))
    Y = _gaussian_mixture(rs, 50000, 128, centres, 0.1)
    X = _gaussian_mixture(rs, 300, 128, centres, 0.1)
    (dist, idx), _, st = pa.parallel_knn(Y, X, 32, return_distance=True, squared=True,
                                         precision="bf16", oversample=128, return_stats=True)
    ref = oracle.exact_knn_f64(X, Y, 32, return_boundary=True)
    rec = oracle.recall_at_k(idx, ref[1])
    err = np.max(np.abs(dist.astype(np.float64) - ref[0]) / np.maximum(ref[0], 1e-30))
    _say(f"    bf16 mixture: recall@32={rec:.5f} max rel dist err={err:.2e} "
         f"uncertified={st['n_uncertified']}")
    assert rec >= 0.999
    # returned distances are exact FP64 recomputations of the returned rows
    got = oracle._direct_sqdist(X, Y, idx)
    np.testing.assert_allclose(dist, got, rtol=1e-6)
    assert st["n_uncertified"] == 0
    assert_knn_parity(X, Y, dist, idx, 32, EPS_F32, ref=ref)


@pytest.mark.parametrize("d", [64, 100, 256])
def test_fp8_rerank_exact_when_certified(require_gpu, d):
    """FP8 E4M3 candidates (power-of-two row scales) + exact re-rank."""
    rs = np.random.RandomState(200 + d)
    Y = rs.rand(30000, d).astype(np.float32)
    X = rs.rand(300, d).astype(np.float32)
    k = 10
    (dist, idx), _, st = pa.parallel_knn(Y, X, k, return_distance=True, squared=True,
                                         precision="fp8", return_stats=True)
    assert st["kernel_path"] == 3
    ref = oracle.exact_knn_f64(X, Y, k, return_boundary=True)
    rec = oracle.recall_at_k(idx, ref[1])
    _say(f"    fp8 d={d}: recall@{k}={rec:.5f} uncertified={st['n_uncertified']} "
         f"widened={st['n_widened']}")
    got = oracle._direct_sqdist(X, Y, idx)
    np.testing.assert_allclose(dist, got, rtol=1e-6)
    assert st["n_uncertified"] == 0
    assert_knn_parity(X, Y, dist, idx, k, EPS_F32, ref=ref)
    assert rec >= 0.999


def test_fp8_mixture_with_duplicates_k100(require_gpu):
    """Mixture of 1,000 Gaussians (sigma 0.05) in d=256 with 1% duplicated
    train rows, k=100, k'=400."""
    rs = np.random.RandomState(11)
    centres = rs.uniform(-1, 1, (1000, 256))
    Y = _gaussian_mixture(rs, 40000, 256, centres, 0.05)
    dup = rs.choice(40000, 400, replace=False)
    Y[dup] = Y[rs.randint(0, 40000, 400)]
    X = _gaussian_mixture(rs, 200, 256, centres, 0.05)
    (dist, idx), _, st = pa.parallel_knn(Y, X, 100, return_distance=True, squared=True,
                                         precision="fp8", oversample=400, return_stats=True)
    ref = oracle.exact_knn_f64(X, Y, 100, return_boundary=True)
    rec = oracle.recall_at_k(idx, ref[1])
    _say(f"    fp8 mixture: recall@100={rec:.5f} uncertified={st['n_uncertified']} "
         f"widened={st['n_widened']}")
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert st["n_uncertified"] == 0
    assert_knn_parity(X, Y, dist, idx, 100, EPS_F32, ref=ref)
    assert rec >= 0.999


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("d", [1, 7, 50, 129])
def test_distance_squared_matrix(require_gpu, dtype, d):
    """Dense squared distances (the SYCL backend's distance_squared_matrix),
    direct differences, against numpy in FP64 (sycl_ext/check_distances.py
    compares the same matrix with numpy)."""
    from pdist_aggregation_amd import sycl_compat

    rs = np.random.RandomState(d)
    Y = rs.rand(517, d).astype(dtype)
    X = rs.rand(131, d).astype(dtype)
    D = sycl_compat.distance_squared_matrix(Y, X)
    assert D.shape == (131, 517) and D.dtype == dtype
    ref = ((X.astype(np.float64)[:, None, :] - Y.astype(np.float64)[None]) ** 2).sum(-1)
    np.testing.assert_allclose(D, ref, rtol=1e-5 if dtype == np.float32 else 1e-13, atol=1e-12)


@pytest.mark.parametrize("precision", ["exact", "bf16", "fp8"])
def test_repeated_calls_identical(require_gpu, precision):
    """Two calls on the same inputs return bit-identical rows (no order-
    dependent floating-point reductions, e.g. in the centring vector)."""
    rs = np.random.RandomState(17)
    centres = rs.uniform(-1, 1, (200, 128))
    Y = _gaussian_mixture(rs, 120000, 128, centres, 0.05)
    Y[rs.choice(120000, 2000, replace=False)] = Y[rs.randint(0, 120000, 2000)]
    X = _gaussian_mixture(rs, 2000, 128, centres, 0.05)
    outs = [pa.parallel_knn(Y, X, 50, return_distance=True, squared=True, precision=precision,
                            oversample=200)[0] for _ in range(3)]
    for dist, idx in outs[1:]:
        np.testing.assert_array_equal(idx, outs[0][1])
        np.testing.assert_array_equal(dist, outs[0][0])
