"""Parity assertions shared by the test modules."""
from __future__ import annotations

import numpy as np

import oracle

# Near-tie half-width, relative to max(1, |distance^2|).  FP32 GEMM-expansion
# candidates can swap two train rows whose exact squared distances differ by
# less than a few FP32 ulps of |x|^2 + |y|^2; such swaps are exempt (and
# counted), everything else must match index for index.
EPS_F32 = 2e-5
EPS_F64 = 1e-10
DIST_RTOL = 1e-5


def assert_sorted_unique(idx, dist, n_train, X=None, Y=None):
    """Rows ascending, indices unique and in range.  With the inputs given,
    equal exact FP64 distances must list the lower index first (float32
    outputs can round two different distances to one value, so the tie rule
    is checked on recomputed exact distances)."""
    idx = np.asarray(idx)
    dist = np.asarray(dist)
    assert idx.min() >= 0 and idx.max() < n_train
    assert np.all(np.diff(dist, axis=1) >= 0), "rows not sorted ascending"
    s = np.sort(idx, axis=1)
    assert np.all(s[:, 1:] != s[:, :-1]), "duplicate index in a row"
    if X is not None:
        exact = oracle._direct_sqdist(np.asarray(X, np.float64), np.asarray(Y, np.float64), idx)
        assert np.all(np.diff(exact, axis=1) >= 0), "rows not sorted by exact distance"
        eq = np.diff(exact, axis=1) == 0
        assert np.all(np.diff(idx, axis=1)[eq] > 0), "tie not broken by lower index"


def assert_knn_parity(X_test, X_train, dist, idx, k, eps_rel, squared=True, dist_rtol=DIST_RTOL,
                      ref=None):
    """Compare an answer with the exact FP64 oracle.

    Indices must be identical except at near-ties (|gap| <= eps_rel); distances
    within dist_rtol relative of the exact ones.  Returns the number of
    ambiguous ranks (reported, not failed).
    """
    X_test = np.asarray(X_test, np.float64)
    X_train = np.asarray(X_train, np.float64)
    if ref is None:
        ref = oracle.exact_knn_f64(X_test, X_train, k, return_boundary=True)
    rd, ri, rb = ref
    dist = np.asarray(dist, np.float64)
    idx = np.asarray(idx)
    assert idx.shape == ri.shape and dist.shape == rd.shape
    got_sq = dist if squared else dist ** 2
    # distances: within dist_rtol, or (at a swapped near-tie) within the tie width
    same = idx == ri
    tol = np.where(same, dist_rtol * np.abs(rd), eps_rel * np.maximum(1.0, np.abs(rd)))
    tol = np.maximum(tol, 4 * dist_rtol * np.abs(rd)) + 1e-12
    err = np.abs(got_sq - rd)
    assert np.all(err <= tol), f"distance error {np.max(err / np.maximum(np.abs(rd), 1e-30))}"
    bad, ambiguous, _ = oracle.near_tie_report(X_test, X_train, idx, rd, ri, rb, eps_rel)
    assert bad == 0, f"{bad} index mismatches outside near-ties"
    return ambiguous


def assert_complete(X_test, X_train, idx, k, eps_rel, block=1 << 24):
    """Every row of every query, at any size: the returned rows are sorted by
    their exact FP64 distance, and no train row outside them lies below the
    k-th returned distance by more than the near-tie width.

    For each block of queries the FP64 expansion |x|^2 + |y|^2 - 2 x.y of
    every (query, train row) pair is compared with c = kth - eps_rel *
    max(1, kth) - e, where e bounds the expansion's rounding error: the count
    of all rows below c must equal the count of returned rows below c, i.e.
    every row whose exact distance is below kth - eps_rel * max(1, kth) is in
    the answer.  Returns the number of queries checked.
    """
    X = np.asarray(X_test, np.float64)
    Y = np.asarray(X_train, np.float64)
    idx = np.asarray(idx)
    nq, nt = X.shape[0], Y.shape[0]
    assert idx.shape == (nq, k)
    assert idx.min() >= 0 and idx.max() < nt
    ynorm = np.einsum("ij,ij->i", Y, Y)
    qb = max(1, block // nt)
    for q0 in range(0, nq, qb):
        Xb = X[q0:q0 + qb]
        ib = idx[q0:q0 + qb]
        exact = oracle._direct_sqdist(Xb, Y, ib)
        assert np.all(np.diff(exact, axis=1) >= 0), "rows not sorted by exact distance"
        xnorm = np.einsum("ij,ij->i", Xb, Xb)
        D = xnorm[:, None] + ynorm[None, :] - 2.0 * (Xb @ Y.T)
        kth = exact[:, -1]
        e = 1e-12 * (xnorm + ynorm.max() + 1.0)
        c = kth - eps_rel * np.maximum(1.0, kth) - e
        n_all = np.sum(D < c[:, None], axis=1)
        n_ret = np.sum(np.take_along_axis(D, ib, axis=1) < c[:, None], axis=1)
        bad = np.nonzero(n_all != n_ret)[0]
        assert bad.size == 0, (f"{bad.size} queries miss a train row clearly below their k-th "
                               f"distance (first: query {q0 + bad[0]})")
    return nq
