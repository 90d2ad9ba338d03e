"""Sampled-threshold start and re-search of the MFMA, low-d and generic paths
(needs a GPU).

The search first sweeps 1/32 of the train rows, keeps per-lane minima
and starts every query from a guessed threshold (the j-th smallest sampled
minimum).  A query left with fewer than k' candidates (MFMA kernels: fewer
than k; with k to k'-1 the merge certifies it against the guess) is searched
again from an infinite threshold.  PDIST_KNN_GUESS forces the sampling on at
small sizes ("1"), forces chosen guesses to fail ("fail", "fail:S" = every
S-th query) and forces the guess's rank ("j:J"), so the normal path, the
short candidate sets and the re-search path are compared with the FP64
oracle here.
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


def _run(Y, X, k, precision, **kw):
    (dist, idx), _, st = pa.parallel_knn(Y, X, k, return_distance=True, squared=True,
                                         precision=precision, return_stats=True, **kw)
    return dist, idx, st


def _check(X, Y, dist, idx, k, precision, st):
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    if precision == "exact":
        assert_knn_parity(X, Y, dist, idx, k, EPS_F32)
        return
    ref = oracle.exact_knn_f64(X, Y, k, return_boundary=True)
    rec = oracle.recall_at_k(idx, ref[1])
    _say(f"    {precision}: recall@{k}={rec:.5f} uncertified={st['n_uncertified']} "
         f"widened={st['n_widened']}")
    got = oracle._direct_sqdist(X, Y, idx)
    np.testing.assert_allclose(dist, got, rtol=1e-6)
    assert st["n_uncertified"] == 0
    assert_knn_parity(X, Y, dist, idx, k, EPS_F32, ref=ref)
    assert rec >= 0.999


CASES = [("exact", 64, 10), ("exact", 128, 10), ("exact", 32, 1), ("bf16", 128, 10),
         ("bf16", 50, 32), ("fp8", 128, 10),
         # float32 search widths 192 and 256 (one or two workgroups per CU)
         ("exact", 160, 10), ("exact", 256, 10),
         # low-d kernel (one lane per query, padded to 4 / 8 / 16 features)
         ("exact", 3, 10), ("exact", 7, 16), ("exact", 12, 5),
         # float32 MFMA through zero-padded row copies (d % 4 != 0)
         ("exact", 17, 10), ("exact", 50, 10), ("exact", 130, 10), ("exact", 255, 16)]


@pytest.mark.parametrize("precision,d,k", CASES)
def test_sampled_guess_matches_oracle(require_gpu, monkeypatch, precision, d, k):
    monkeypatch.setenv("PDIST_KNN_GUESS", "1")
    rs = np.random.RandomState(d * 7 + k)
    Y = rs.rand(40000, d).astype(np.float32)
    X = rs.rand(700, d).astype(np.float32)
    _say(f"[guess] {precision} d={d} k={k} 40000 x 700")
    dist, idx, st = _run(Y, X, k, precision)
    _say(f"    stats {st}")
    assert st["kernel_path"] in ((4,) if d < 16 else (1, 2, 3))
    _check(X, Y, dist, idx, k, precision, st)


@pytest.mark.parametrize("precision,d,k", CASES)
def test_every_guess_failing_is_searched_again(require_gpu, monkeypatch, precision, d, k):
    monkeypatch.setenv("PDIST_KNN_GUESS", "fail")
    rs = np.random.RandomState(d * 11 + k)
    Y = rs.rand(30000, d).astype(np.float32)
    X = rs.rand(513, d).astype(np.float32)
    _say(f"[guess fail] {precision} d={d} k={k} 30000 x 513")
    dist, idx, st = _run(Y, X, k, precision)
    assert st["n_rescued"] == X.shape[0]
    _check(X, Y, dist, idx, k, precision, st)


@pytest.mark.parametrize("precision,d", [("exact", 64), ("exact", 128), ("bf16", 128),
                                         ("bf16", 64), ("fp8", 128)])
@pytest.mark.parametrize("j", [1, 2])
def test_tight_guess_short_candidate_sets(require_gpu, monkeypatch, precision, d, j):
    """MFMA kernels with the guess forced to the 1st / 2nd smallest sampled
    minimum, about 32 / 64 rows deep: most queries end with k to k'-1
    candidates (k = 10, k' = 40 for BF16 / FP8, 18 for float32) and are
    certified against the guess or widened, the rest (fewer than k) are
    searched again; every answer is the oracle's."""
    monkeypatch.setenv("PDIST_KNN_GUESS", f"j:{j}")
    rs = np.random.RandomState(d * 3 + j)
    Y = rs.rand(40000, d).astype(np.float32)
    X = rs.rand(700, d).astype(np.float32)
    _say(f"[guess j:{j}] {precision} d={d} k=10 40000 x 700")
    dist, idx, st = _run(Y, X, 10, precision)
    _say(f"    stats {st}")
    assert st["kernel_path"] in (1, 2, 3)
    assert 0 < st["n_rescued"] < X.shape[0]
    _check(X, Y, dist, idx, 10, precision, st)


@pytest.mark.parametrize("stride", [500, 3, 64])
def test_some_guesses_failing(require_gpu, monkeypatch, stride):
    """Only every stride-th query fails: the re-search must skip whole query
    blocks without one (stride 500 leaves the last block out) and leave the
    other queries of a selected block untouched."""
    monkeypatch.setenv("PDIST_KNN_GUESS", f"fail:{stride}")
    rs = np.random.RandomState(stride)
    Y = rs.rand(25000, 64).astype(np.float32)
    X = rs.rand(700, 64).astype(np.float32)
    _say(f"[guess fail:{stride}] exact d=64 25000 x 700")
    dist, idx, st = _run(Y, X, 10, "exact")
    assert st["n_rescued"] == len(range(0, 700, stride))
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dist, idx, 10, EPS_F32)


def test_duplicated_rows_with_guess(require_gpu, monkeypatch):
    """Every distinct train row repeated 40 times along the set: a query's k'
    nearest are exact ties, sampled copies put the guess on the tie value."""
    monkeypatch.setenv("PDIST_KNN_GUESS", "1")
    rs = np.random.RandomState(5)
    base = rs.rand(1000, 64).astype(np.float32)
    Y = np.tile(base, (40, 1))
    X = (base[rs.randint(0, 1000, 300)] + 1e-3 * rs.randn(300, 64)).astype(np.float32)
    dist, idx, st = _run(Y, X, 10, "exact")
    _say(f"    duplicates: stats {st}")
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dist, idx, 10, EPS_F32)


def test_guess_default_off_below_size_and_repeatable(require_gpu, monkeypatch):
    """Without the override small searches take no sample; two calls on the
    same inputs give identical results (re-search or not)."""
    monkeypatch.delenv("PDIST_KNN_GUESS", raising=False)
    rs = np.random.RandomState(9)
    Y = rs.rand(20000, 64).astype(np.float32)
    X = rs.rand(300, 64).astype(np.float32)
    d0, i0, st0 = _run(Y, X, 10, "exact")
    monkeypatch.setenv("PDIST_KNN_GUESS", "1")
    d1, i1, st1 = _run(Y, X, 10, "exact")
    d2, i2, _ = _run(Y, X, 10, "exact")
    assert st0["n_rescued"] == 0
    np.testing.assert_array_equal(i1, i2)
    np.testing.assert_array_equal(d1, d2)
    np.testing.assert_array_equal(i0, i1)


# Generic kernel (float32 with d > 256, float64 with d >= 16):
# its sampling launch keeps the kSampleJ smallest reduced distances of each
# region (float64 ones rounded up to float) and its re-search runs 64-query
# blocks.
GENERIC_CASES = [(np.float32, 261, 10), (np.float32, 300, 32), (np.float64, 50, 10),
                 (np.float64, 100, 100)]


@pytest.mark.parametrize("mode", ["1", "fail", "fail:7"])
@pytest.mark.parametrize("dtype,d,k", GENERIC_CASES)
def test_generic_guess_matches_oracle(require_gpu, monkeypatch, mode, dtype, d, k):
    monkeypatch.setenv("PDIST_KNN_GUESS", mode)
    rs = np.random.RandomState(d * 13 + k)
    Y = rs.rand(30000, d).astype(dtype)
    X = rs.rand(300, d).astype(dtype)
    _say(f"[generic guess {mode}] {np.dtype(dtype).name} d={d} k={k} 30000 x 300")
    dist, idx, st = _run(Y, X, k, "exact")
    _say(f"    stats {st}")
    assert st["kernel_path"] == 0
    if mode == "fail":
        assert st["n_rescued"] == X.shape[0]
    elif mode == "fail:7":
        assert st["n_rescued"] == len(range(0, X.shape[0], 7))
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dist, idx, k, EPS_F64 if dtype == np.float64 else EPS_F32)


def test_generic_guess_on_by_default_and_repeatable(require_gpu, monkeypatch):
    """Float64 with 40,000 train rows takes the sample by default; the answer
    equals the one searched from +inf (sampling forced off)."""
    monkeypatch.delenv("PDIST_KNN_GUESS", raising=False)
    rs = np.random.RandomState(21)
    Y = rs.rand(40000, 50)
    X = rs.rand(500, 50)
    d1, i1, st1 = _run(Y, X, 10, "exact")
    monkeypatch.setenv("PDIST_KNN_GUESS", "0")
    d0, i0, _ = _run(Y, X, 10, "exact")
    _say(f"    float64 default: stats {st1}")
    np.testing.assert_array_equal(i0, i1)
    np.testing.assert_array_equal(d0, d1)
    assert_knn_parity(X, Y, d1, i1, 10, EPS_F64)


@pytest.mark.parametrize("mode", ["1", "fail", "fail:7"])
@pytest.mark.parametrize("d,k", [(3, 10), (7, 16), (12, 5)])
def test_lowd_float64_guess_matches_oracle(require_gpu, monkeypatch, mode, d, k):
    """Low-d kernel on float64: sampled minima rounded up to float, re-search
    of 256-query blocks."""
    monkeypatch.setenv("PDIST_KNN_GUESS", mode)
    rs = np.random.RandomState(d * 17 + k)
    Y = rs.rand(30000, d)
    X = rs.rand(600, d)
    _say(f"[low-d float64 guess {mode}] d={d} k={k} 30000 x 600")
    dist, idx, st = _run(Y, X, k, "exact")
    _say(f"    stats {st}")
    assert st["kernel_path"] == 4
    if mode == "fail":
        assert st["n_rescued"] == X.shape[0]
    elif mode == "fail:7":
        assert st["n_rescued"] == len(range(0, X.shape[0], 7))
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dist, idx, k, EPS_F64)
