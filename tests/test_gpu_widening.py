"""Bound-following re-rank sets of the MFMA paths (needs a GPU).

A query whose k' candidates do not prove the answer exact (the error-bound
certificate of the merge fails) is widened: the merge sets its threshold T to
the largest value the search can compute for any train row within its k-th
exact distance, a collect launch gathers every row at or below T, and all of
them are re-ranked in FP64.  PDIST_KNN_WIDEN=all (a testing switch) widens
every query, so the path is compared with the FP64 oracle here on small
inputs for every MFMA format, with exact ties at the k-th distance, together
with failing sampled guesses (the re-search's merge widens too), and with a
collect set larger than its capacity (the query keeps its first answer and
stays counted as uncertified).  PDIST_KNN_WIDEN=every:S widens every S-th
query: fewer than half of them run the collect launch in packed query
blocks, more run it in the search's own blocks.  A widened query whose
collect threshold lies below the search's final shared threshold takes its
row set from the search's split lists instead of the collect launch; "sweep"
in PDIST_KNN_WIDEN sends every widened query through the collect launch, and
the two routes must give the same answers.
"""
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


def _run(Y, X, k, precision, **kw):
    (dist, idx), _, st = pa.parallel_knn(Y, X, k, return_distance=True, squared=True,
                                         precision=precision, return_stats=True, **kw)
    return dist, idx, st


CASES = [("exact", 64, 10), ("exact", 130, 7), ("bf16", 128, 10), ("bf16", 50, 32),
         ("fp8", 128, 10), ("fp8", 256, 40)]


@pytest.mark.parametrize("precision,d,k", CASES)
def test_widen_every_query_matches_oracle(require_gpu, monkeypatch, precision, d, k):
    monkeypatch.setenv("PDIST_KNN_WIDEN", "all")
    rs = np.random.RandomState(d + k)
    Y = rs.rand(30000, d).astype(np.float32)
    X = rs.rand(301, d).astype(np.float32)
    _say(f"[widen all] {precision} d={d} k={k} 30000 x 301")
    dist, idx, st = _run(Y, X, k, precision)
    _say(f"    stats {st}")
    assert st["n_widened"] == X.shape[0]
    assert st["n_uncertified"] == 0
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dist, idx, k, EPS_F32)


@pytest.mark.parametrize("precision,d", [("exact", 64), ("bf16", 128), ("fp8", 128),
                                         ("fp8", 256)])
@pytest.mark.parametrize("stride", [1000, 7, 2])
def test_widen_some_queries_packed_or_not(require_gpu, monkeypatch, precision, d, stride):
    """Every stride-th query widened: one (stride 1000) or 1/7 of the queries
    (packed blocks), or half of them (the search's blocks); 1,100 queries
    span several query blocks.  All of them take the collect launch."""
    monkeypatch.setenv("PDIST_KNN_WIDEN", f"every:{stride},sweep")
    rs = np.random.RandomState(d + stride)
    Y = rs.rand(30000, d).astype(np.float32)
    X = rs.rand(1100, d).astype(np.float32)
    _say(f"[widen every:{stride}] {precision} d={d} k=10 30000 x 1100")
    dist, idx, st = _run(Y, X, 10, precision)
    _say(f"    stats {st}")
    assert st["n_widened"] >= len(range(0, X.shape[0], stride))
    assert st["n_widened_from_lists"] == 0
    assert st["n_uncertified"] == 0
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dist, idx, 10, EPS_F32)


@pytest.mark.parametrize("precision,d,k", [("exact", 64, 10), ("bf16", 128, 10),
                                           ("fp8", 128, 10), ("fp8", 256, 40)])
def test_widened_sets_from_lists_match_the_sweep(require_gpu, monkeypatch, precision, d, k):
    """A deep sampled guess (j = 24) leaves every row up to about 24 x 32
    rows deep in the split lists, so widened queries take their row sets
    from there; the answers equal those of the collect launch bit for bit
    (the same rows lie within the k-th exact distance either way) and the
    oracle's."""
    rs = np.random.RandomState(40 + d + k)
    Y = rs.rand(60000, d).astype(np.float32)
    X = rs.rand(700, d).astype(np.float32)
    monkeypatch.setenv("PDIST_KNN_GUESS", "j:24")
    out = {}
    for route in ("all", "all,sweep"):
        monkeypatch.setenv("PDIST_KNN_WIDEN", route)
        _say(f"[widen {route}] {precision} d={d} k={k} 60000 x 700, guess j=24")
        out[route] = _run(Y, X, k, precision)
        _say(f"    stats {out[route][2]}")
    dl, il, sl = out["all"]
    ds, is_, ss = out["all,sweep"]
    assert sl["n_widened"] == ss["n_widened"] == X.shape[0]
    assert ss["n_widened_from_lists"] == 0
    assert sl["n_widened_from_lists"] > 0
    assert sl["n_uncertified"] == 0 and ss["n_uncertified"] == 0
    np.testing.assert_array_equal(il, is_)
    np.testing.assert_array_equal(dl, ds)
    assert_sorted_unique(il, dl, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dl, il, k, EPS_F32)


@pytest.mark.parametrize("precision", ["exact", "bf16", "fp8"])
def test_widen_with_exact_ties_at_the_cut(require_gpu, monkeypatch, precision):
    """Train rows duplicated so that exact ties straddle the k-th rank: the
    widened answer lists the lower indices first, like the oracle."""
    monkeypatch.setenv("PDIST_KNN_WIDEN", "all")
    rs = np.random.RandomState(21)
    Y = rs.rand(20000, 64).astype(np.float32)
    X = rs.rand(97, 64).astype(np.float32)
    # for every query, 6 copies of one row placed near it: ranks 1-6 tie, k = 4
    for q in range(X.shape[0]):
        base = X[q] + np.float32(0.02) * rs.randn(64).astype(np.float32)
        rows = rs.choice(20000, 6, replace=False)
        Y[rows] = base
    dist, idx, st = _run(Y, X, 4, precision)
    assert st["n_uncertified"] == 0
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dist, idx, 4, 1e-12)


def test_widen_after_failed_guesses(require_gpu, monkeypatch):
    """Every second guess fails: those queries are widened by the re-search's
    merge, the others by the first merge."""
    monkeypatch.setenv("PDIST_KNN_WIDEN", "all")
    monkeypatch.setenv("PDIST_KNN_GUESS", "fail:2")
    rs = np.random.RandomState(8)
    Y = rs.rand(50000, 128).astype(np.float32)
    X = rs.rand(515, 128).astype(np.float32)
    dist, idx, st = _run(Y, X, 10, "bf16")
    _say(f"    stats {st}")
    assert st["n_rescued"] == (X.shape[0] + 1) // 2
    assert st["n_widened"] == X.shape[0]
    assert st["n_uncertified"] == 0
    assert_knn_parity(X, Y, dist, idx, 10, EPS_F32)


@pytest.mark.parametrize("precision", ["exact", "bf16", "fp8"])
def test_ties_overflowing_the_merge_staging(require_gpu, precision):
    """600 copies of one row next to query 0: its k'-th candidate value is
    tied 600 times, more than the merge stages (256), so the staging keeps
    the ties of lowest row index and the merge's answer is the oracle's (the
    ten lowest indices of the copies).  The query is widened (its bound does
    not certify ties at the k-th distance) and its 600 rows at the k-th
    exact distance exceed the re-rank's 256 staging slots, so it keeps that
    first answer and is the one query counted as uncertified; every other
    query matches the oracle."""
    rs = np.random.RandomState(12)
    Y = rs.rand(40000, 64).astype(np.float32)
    X = rs.rand(150, 64).astype(np.float32)
    rep = np.sort(rs.choice(40000, 600, replace=False))
    Y[rep] = X[0] + np.float32(0.01)
    dist, idx, st = _run(Y, X, 10, precision)
    _say(f"    {precision} staging ties: stats {st}")
    assert st["n_uncertified"] == 1
    np.testing.assert_array_equal(idx[0], rep[:10])
    d0 = oracle._direct_sqdist(X[:1], Y, idx[:1])
    np.testing.assert_allclose(dist[0], d0[0], rtol=1e-6)
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert_knn_parity(X[1:], Y, dist[1:], idx[1:], 10, EPS_F32)


def test_collect_overflow_keeps_first_answer(require_gpu, monkeypatch):
    """Query 0 sits on a train row repeated 3,000 times (exact ties at
    distance 0): its row set exceeds the collect capacity (1,024 at k'=18),
    so it keeps the first answer (the lowest indices of the repeated row, the
    correct one) and is the one query counted as uncertified."""
    monkeypatch.setenv("PDIST_KNN_WIDEN", "all")
    rs = np.random.RandomState(9)
    Y = rs.rand(40000, 32).astype(np.float32)
    rep = np.sort(rs.choice(40000, 3000, replace=False))
    Y[rep] = Y[rep[0]]
    X = rs.rand(200, 32).astype(np.float32)
    X[0] = Y[rep[0]]
    dist, idx, st = _run(Y, X, 10, "exact")
    _say(f"    stats {st}")
    assert st["n_widened"] == X.shape[0]
    assert st["n_uncertified"] == 1
    np.testing.assert_array_equal(idx[0], rep[:10])
    assert np.all(dist[0] == 0)
    assert_knn_parity(X, Y, dist, idx, 10, EPS_F32)


def test_cluster_ordered_rerank_matches_oracle(require_gpu, monkeypatch):
    """FP8 with k' >= 256 (collect lists of 4,096 rows and more) re-ranks the
    widened queries in cluster order: keys (the smallest row index each
    collected) are bucketed over the range in use, so queries of one cluster
    run side by side.  24 Gaussian clusters, every query widened: the
    answers are the oracle's."""
    rs = np.random.RandomState(5)
    centres = rs.uniform(-1, 1, (24, 64)).astype(np.float32)
    Y = (centres[rs.randint(0, 24, 40000)] + 0.05 * rs.randn(40000, 64)).astype(np.float32)
    X = (centres[rs.randint(0, 24, 300)] + 0.05 * rs.randn(300, 64)).astype(np.float32)
    k = 64  # k' = 256: FP8 collect lists of 4,096 rows
    monkeypatch.setenv("PDIST_KNN_WIDEN", "all")
    _say(f"[widen all] fp8 d=64 k={k} 40000 x 300, 24 clusters")
    dist, idx, st = _run(Y, X, k, "fp8")
    _say(f"    stats {st}")
    assert st["n_widened"] == X.shape[0]
    assert st["n_uncertified"] == 0
    assert_sorted_unique(idx, dist, Y.shape[0], X, Y)
    assert_knn_parity(X, Y, dist, idx, k, EPS_F32)
