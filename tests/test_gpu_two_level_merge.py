"""Two-level merge (needs a GPU): with more train splits than one merge
workgroup takes (n_splits * k' > 2048), groups of split lists are first
reduced to one sorted list each, then merged.  PDIST_KNN_SPLITS forces the
split count (clamped to what two levels take), so the pre-merge runs at test
sizes; results are compared with the FP64 oracle, also with every query
widened and its set taken from the split lists."""
import sys

import numpy as np
import pytest

import oracle
import pdist_aggregation_amd as pa
from knn_checks import EPS_F32, assert_knn_parity, assert_sorted_unique

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    # precision, n_train, n_test, d, k, forced splits
    ("exact", 60000, 300, 64, 100, 150),   # k'=112: 18 lists per merge -> groups of 9
    ("exact", 60000, 257, 128, 10, 200),   # k'=18: 113 lists -> groups of 2
    ("bf16", 60000, 300, 128, 32, 100),    # k'=128: 16 lists -> groups of 7
    ("fp8", 60000, 200, 128, 10, 120),     # k'=40: 51 lists -> groups of 3
    ("exact", 40000, 100, 32, 50, 17),     # below the limit: one level (control)
    ("exact", 60000, 300, 50, 20, 150),    # zero-padded rows (d % 4 != 0), k'=28: groups of 3
]


@pytest.mark.parametrize("precision,nt,nq,d,k,splits", CASES)
def test_forced_splits_match_oracle(require_gpu, monkeypatch, precision, nt, nq, d, k, splits):
    monkeypatch.setenv("PDIST_KNN_SPLITS", str(splits))
    monkeypatch.setenv("PDIST_KNN_GUESS", "0")
    rs = np.random.RandomState(nt + k + splits)
    Y = rs.rand(nt, d).astype(np.float32)
    X = rs.rand(nq, d).astype(np.float32)
    _say(f"[two-level merge] {precision} {nq} x {nt} x {d} k={k} splits={splits}")
    (dist, idx), _, st = pa.parallel_knn(Y, X, k, return_distance=True, squared=True,
                                         precision=precision, return_stats=True)
    _say(f"    stats {st}")
    two_level = splits not in (17,)
    assert (st["n_splits"] * st["k_candidates"] > 2048) == two_level, st
    assert_sorted_unique(idx, dist, nt, X, Y)
    if precision == "exact":
        assert_knn_parity(X, Y, dist, idx, k, EPS_F32)
        return
    ref = oracle.exact_knn_f64(X, Y, k, return_boundary=True)
    rec = oracle.recall_at_k(idx, ref[1])
    _say(f"    recall@{k}={rec:.5f}")
    np.testing.assert_allclose(dist, oracle._direct_sqdist(X, Y, idx), rtol=1e-6)
    assert st["n_uncertified"] == 0
    assert_knn_parity(X, Y, dist, idx, k, EPS_F32, ref=ref)
    assert rec >= 0.999


def test_forced_splits_with_sampled_guess_and_failures(require_gpu, monkeypatch):
    """Pre-merge before the re-search's merge too: every guess fails."""
    monkeypatch.setenv("PDIST_KNN_SPLITS", "160")
    monkeypatch.setenv("PDIST_KNN_GUESS", "fail:3")
    rs = np.random.RandomState(77)
    Y = rs.rand(50000, 64).astype(np.float32)
    X = rs.rand(333, 64).astype(np.float32)
    (dist, idx), _, st = pa.parallel_knn(Y, X, 20, return_distance=True, squared=True,
                                         precision="exact", return_stats=True)
    _say(f"    stats {st}")
    assert st["n_rescued"] == len(range(0, 333, 3))
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dist, idx, 20, EPS_F32)


@pytest.mark.parametrize("precision", ["bf16", "fp8"])
def test_forced_splits_widened_from_lists(require_gpu, monkeypatch, precision):
    """Two-level merge with every query widened and a deep sampled guess
    (j = 24): the widened sets come from the search's own split lists (the
    pre-merge reduces copies of them) and give the collect sweep's answers
    bit for bit, and the oracle's."""
    monkeypatch.setenv("PDIST_KNN_SPLITS", "120")
    monkeypatch.setenv("PDIST_KNN_GUESS", "j:24")
    rs = np.random.RandomState(91)
    Y = rs.rand(60000, 128).astype(np.float32)
    X = rs.rand(200, 128).astype(np.float32)
    out = {}
    for route in ("all", "all,sweep"):
        monkeypatch.setenv("PDIST_KNN_WIDEN", route)
        _say(f"[two-level merge, widen {route}] {precision} 200 x 60000 x 128 k=10 splits=120")
        (dist, idx), _, st = pa.parallel_knn(Y, X, 10, return_distance=True, squared=True,
                                             precision=precision, return_stats=True)
        _say(f"    stats {st}")
        out[route] = (dist, idx, st)
    dl, il, sl = out["all"]
    ds, is_, ss = out["all,sweep"]
    assert sl["n_splits"] * sl["k_candidates"] > 2048, sl
    assert sl["n_widened"] == ss["n_widened"] == X.shape[0]
    assert sl["n_widened_from_lists"] > 0 and ss["n_widened_from_lists"] == 0
    assert sl["n_uncertified"] == 0 and ss["n_uncertified"] == 0
    np.testing.assert_array_equal(il, is_)
    np.testing.assert_array_equal(dl, ds)
    assert_sorted_unique(il, dl, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dl, il, 10, EPS_F32)
