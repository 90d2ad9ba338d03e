"""Pin the CPU oracle before trusting it (no GPU needed).

The oracle is checked against (a) the sklearn brute-force fixtures, the
reference test's own oracle (tests/test_correctness.py:35-49 of the
reference), (b) direct-difference brute force, (c) closed forms, and (d) the
reference's chunk/strategy rules.  Mutated copies of the oracle (a dropped
term, a wrong sign, swapped operands, an index offset) must fail (a).
"""
import numpy as np
import pytest

import oracle
from conftest import golden_manifest, load_golden

NAMES = sorted(golden_manifest())


@pytest.fixture(scope="module", autouse=True)
def _built():
    import subprocess

    subprocess.run(["make", "-s", "-C", str(oracle._HERE)], check=True)


@pytest.mark.parametrize("name", NAMES)
def test_exact_oracle_matches_sklearn_fixture(name):
    X_train, X_test, k, sk_dist, sk_idx = load_golden(name)
    dist, idx = oracle.exact_knn_f64(X_test, X_train, k)
    np.testing.assert_array_equal(idx, sk_idx)
    np.testing.assert_almost_equal(np.sqrt(dist), sk_dist)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("strategy", ["auto", "chunk_on_train", "chunk_on_test"])
def test_port_matches_sklearn_fixture(name, strategy):
    X_train, X_test, k, sk_dist, sk_idx = load_golden(name)
    for chunk_size in (512, 4096) if X_test.shape[0] * X_train.shape[0] > 2e7 else (256, 512,
                                                                                    1024):
        (dist, idx), _ = oracle.port_parallel_knn(X_train, X_test, k, chunk_size=chunk_size,
                                                  strategy=strategy, return_distance=True)
        np.testing.assert_array_equal(idx, sk_idx)
        np.testing.assert_almost_equal(dist, sk_dist)


def test_port_float32_both_strategies():
    rs = np.random.RandomState(3)
    Y = rs.rand(3000, 17).astype(np.float32)
    X = rs.rand(700, 17).astype(np.float32)
    ref_d, ref_i, ref_b = oracle.exact_knn_f64(X, Y, 12, return_boundary=True)
    for strategy in ("chunk_on_train", "chunk_on_test"):
        (d, i), _ = oracle.port_parallel_knn(Y, X, 12, chunk_size=300, strategy=strategy,
                                             return_distance=True)
        bad, _, _ = oracle.near_tie_report(X, Y, i, ref_d, ref_i, ref_b, 1e-5)
        assert bad == 0
        np.testing.assert_allclose(d.astype(np.float64) ** 2, ref_d, rtol=1e-5)


def _random_with_duplicates(seed, n_train, n_test, d):
    rs = np.random.RandomState(seed)
    Y = rs.rand(n_train, d)
    Y[5::7] = Y[3]  # many exact duplicates of row 3
    X = rs.rand(n_test, d)
    X[0] = Y[3]  # a query sitting on the duplicated point
    return Y, X


@pytest.mark.parametrize("d", [1, 2, 3, 7, 50])
def test_exact_oracle_equals_brute_force(d):
    Y, X = _random_with_duplicates(d, 400, 60, d)
    k = 25
    bd, bi = oracle.brute_force_knn(X, Y, k)
    ed, ei = oracle.exact_knn_f64(X, Y, k, y_chunk=97)
    np.testing.assert_array_equal(ei, bi)
    np.testing.assert_allclose(ed, bd, rtol=1e-12, atol=1e-15)


def test_ties_lower_index_first():
    Y, X = _random_with_duplicates(0, 200, 3, 4)
    d, i = oracle.exact_knn_f64(X, Y, 10)
    dup = np.flatnonzero(np.all(Y == Y[3], axis=1))
    np.testing.assert_array_equal(i[0][: len(dup)], dup[:10])
    assert np.all(d[0][: len(dup)] == 0.0)


def test_closed_form_on_a_line():
    """Train points at 0..19 on a line, query at 6.3: neighbours 6, 7, 5, 8, 4,
    ... with squared distances (j - 6.3)^2."""
    Y = np.arange(20, dtype=np.float64)[:, None]
    X = np.array([[6.3]])
    d, i = oracle.exact_knn_f64(X, Y, 6)
    np.testing.assert_array_equal(i[0], [6, 7, 5, 8, 4, 9])
    np.testing.assert_allclose(d[0], (np.array([6, 7, 5, 8, 4, 9]) - 6.3) ** 2)
    # Port: Euclidean distances, same order
    (pd, pi), _ = oracle.port_parallel_knn(Y, X, 6, chunk_size=20, return_distance=True)
    np.testing.assert_array_equal(pi[0], [6, 7, 5, 8, 4, 9])
    np.testing.assert_allclose(pd[0], np.abs(np.array([6, 7, 5, 8, 4, 9]) - 6.3))


@pytest.mark.parametrize("n_train,n_test,chunk", [(10000, 1000, 256), (1000000, 10000, 4096),
                                                  (100, 5, 512), (1000, 3000, 20), (57, 57, 1)])
@pytest.mark.parametrize("threads", [1, 8, 64])
def test_strategy_and_chunk_rules(n_train, n_test, chunk, threads):
    s = oracle.choose_strategy("auto", chunk, threads, n_test)
    assert s == ("chunk_on_test" if 4 * chunk * threads < n_test else "chunk_on_train")
    per, n = oracle.chunk_geometry(n_train, chunk)
    assert per == min(n_train, max(20, chunk))
    assert n == -(-n_train // per)
    # the workload of the reference's CPU-runnable case resolves to chunk_on_train
    assert oracle.choose_strategy("auto", 256, 8, 1000) == "chunk_on_train"


def test_port_n_parallel_chunks_closed_form():
    rs = np.random.RandomState(0)
    Y = rs.rand(1013, 3)
    X = rs.rand(333, 3)
    for strategy in ("chunk_on_train", "chunk_on_test"):
        for chunk in (7, 20, 100, 5000):
            _, n = oracle.port_parallel_knn(Y, X, 3, chunk_size=chunk, strategy=strategy)
            assert n == oracle.n_parallel_chunks(strategy, 1013, 333, chunk, 4)


def test_port_quirk_k_above_n_train():
    """Reference quirk: with k > n_train the unfilled slots keep index 0
    (chunk_on_test: the output initialised to 0 at :503 is the heap) or a
    mix of 0 and -1 (chunk_on_train: thread heaps start at -1, :398, and an
    equal 1e36 value replaces the root, :66)."""
    rs = np.random.RandomState(1)
    Y = rs.rand(5, 2)
    X = rs.rand(3, 2)
    i_test, _ = oracle.port_parallel_knn(Y, X, 8, chunk_size=256, strategy="chunk_on_test")
    i_train, _ = oracle.port_parallel_knn(Y, X, 8, chunk_size=256, strategy="chunk_on_train")
    assert np.all(i_test[:, 5:] == 0)
    assert np.all(np.isin(i_train[:, 5:], [0, -1]))
    for row_t, row_r in zip(i_test, i_train):
        assert sorted(row_t[:5]) == sorted(row_r[:5]) == list(range(5))


def test_port_rejects_unknown_strategy():
    with pytest.raises(RuntimeError, match="not supported"):
        oracle.port_parallel_knn(np.zeros((3, 2)), np.zeros((2, 2)), 1, strategy="bogus")


# --- the checks must catch plausible mistakes planted in the oracle ----------
def _mutant(kind):
    def run(X, Y, k):
        X = np.asarray(X, np.float64)
        Y = np.asarray(Y, np.float64)
        if kind == "drop_norm":
            v = -2.0 * X @ Y.T
        elif kind == "wrong_sign":
            v = (Y * Y).sum(1)[None, :] + 2.0 * X @ Y.T
        elif kind == "swapped":
            n = min(X.shape[0], Y.shape[0])
            v = (X[:n] * X[:n]).sum(1)[None, :] - 2.0 * Y[:n] @ X[:n].T
            v = np.pad(v, ((0, X.shape[0] - v.shape[0]), (0, Y.shape[0] - v.shape[1])),
                       constant_values=np.inf)
        elif kind == "offset":
            v = (Y * Y).sum(1)[None, :] - 2.0 * X @ Y.T
            o = np.argsort(v, axis=1, kind="stable")[:, :k]
            return (np.minimum(o + (o >= 256), Y.shape[0] - 1)).astype(np.int64)
        return np.argsort(v, axis=1, kind="stable")[:, :k]

    return run


@pytest.mark.parametrize("kind", ["drop_norm", "wrong_sign", "swapped", "offset"])
def test_planted_oracle_mistakes_are_caught(kind):
    X_train, X_test, k, _, sk_idx = load_golden("workload_1000x10000x50_k10")
    idx = _mutant(kind)(X_test, X_train, k)
    assert not np.array_equal(idx, sk_idx)
    ref_d, ref_i, ref_b = oracle.exact_knn_f64(X_test, X_train, k, return_boundary=True)
    bad, _, _ = oracle.near_tie_report(X_test, X_train, idx, ref_d, ref_i, ref_b, 2e-5)
    assert bad > 0


def test_recall_at_k():
    a = np.array([[1, 2, 3], [4, 5, 6]])
    b = np.array([[3, 2, 9], [4, 5, 6]])
    assert oracle.recall_at_k(a, b) == pytest.approx(5 / 6)
