"""CPU oracle for brute-force k-NN argkmin (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It is the checker, never the product:
``pdist_aggregation_amd`` does not import it and has no CPU fallback.

Contents
--------
* :func:`brute_force_knn` - direct-difference FP64 brute force with the
  lexicographic ``(distance, index)`` order (lower index wins a tie).  The
  mathematical truth for small inputs.
* :func:`exact_knn_f64` - the same answer at large sizes: FP64 GEMM-expansion
  candidates, then exact direct-difference FP64 re-rank of those candidates.
* :func:`port_parallel_knn` - ctypes binding of ``knn_port.cpp``, the
  C++/OpenMP restatement of the reference's Cython engine
  (``src/_pdist_aggregation.pyx``), used as a second oracle and as the CPU
  baseline.
* :func:`effective_n_threads`, :func:`n_parallel_chunks`,
  :func:`choose_strategy` - the reference's thread, chunk and strategy rules.
* :func:`recall_at_k`, :func:`near_tie_report` - comparison helpers.

Every function is pinned by tests in ``tests/test_oracle.py`` against the
sklearn brute-force fixtures (the reference test's own oracle,
``tests/test_correctness.py:35-38``), closed forms and brute force.
"""
from __future__ import annotations

import ctypes
import os
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent

CHUNK_SIZE = 256  # src/_pdist_aggregation.pyx:24
MIN_CHUNK_SAMPLES = 20  # :26
FLOAT_INF = 1e36  # :28

STRATEGIES = {"auto": 0, "chunk_on_train": 1, "chunk_on_test": 2}


# ----------------------------------------------------------------------------
# Mathematical truth
# ----------------------------------------------------------------------------
def _lexsort_topk(d: np.ndarray, idx: np.ndarray, k: int):
    """Sort each row of 2-D (d, idx) by (distance, index) and keep k."""
    idx = np.broadcast_to(idx, d.shape)
    order = np.lexsort((idx, d), axis=-1)[:, :k]
    return (np.take_along_axis(d, order, 1),
            np.ascontiguousarray(np.take_along_axis(idx, order, 1), dtype=np.int64))


def brute_force_knn(X, Y, k):
    """Exact k nearest rows of Y for each row of X (squared Euclidean, FP64).

    Direct differences sum((x - y)^2) over every pair; ties broken by the
    lower index.  Memory is O(n_x * n_y * d): small inputs only.
    """
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    D = ((X[:, None, :] - Y[None, :, :]) ** 2).sum(-1)
    idx = np.broadcast_to(np.arange(Y.shape[0]), D.shape)
    return _lexsort_topk(D, idx, k)


def _direct_sqdist(X, Y, idx):
    """sum((X[i] - Y[idx[i, j]])^2) in FP64 for every (i, j), in blocks of
    queries of at most about 128 MiB of gathered rows."""
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    idx = np.asarray(idx)
    out = np.empty(idx.shape, np.float64)
    per = max(1, idx.shape[1] * X.shape[1])
    b = max(1, (1 << 24) // per)
    for r0 in range(0, idx.shape[0], b):
        diff = Y[idx[r0:r0 + b]] - X[r0:r0 + b, None, :]
        out[r0:r0 + b] = np.einsum("ijk,ijk->ij", diff, diff)
    return out


def exact_knn_f64(X, Y, k, extra=None, y_chunk=1 << 16, return_boundary=False):
    """Exact FP64 k-NN at any size (squared distances, lower index wins ties).

    Candidates are the ``k + extra`` smallest FP64 expansion values
    ``|y|^2 - 2 x.y`` per query (error ~1e-13 relative), then every candidate
    is re-ranked with the direct-difference FP64 distance and the rows are
    ordered by (distance, index).  ``extra`` (default max(16, k)) absorbs
    exact ties between duplicate rows and expansion rounding.  Queries are
    processed in blocks so that memory stays bounded at any size.

    With ``return_boundary`` also returns the (k+1)-th exact distance of each
    row (``inf`` when Y has only k rows), used for near-tie analysis.
    """
    X = np.ascontiguousarray(X, np.float64)
    Y = np.ascontiguousarray(Y, np.float64)
    nq, nt = X.shape[0], Y.shape[0]
    d = X.shape[1]
    kk = min(nt, k + (max(16, k) if extra is None else extra))
    m = min(nt, k + 1)
    ynorm = np.einsum("ij,ij->i", Y, Y)
    dist = np.empty((nq, m))
    idx = np.empty((nq, m), np.int64)
    # query block: the value matrix (block x y_chunk) and the gathered
    # candidate rows (block x ~kk x d) each stay within about 256 MiB
    qb = max(1, min((1 << 25) // min(nt, y_chunk), (1 << 25) // max(1, 2 * kk * d)))
    for q0 in range(0, nq, qb):
        Xb = X[q0:q0 + qb]
        dist[q0:q0 + qb], idx[q0:q0 + qb] = _exact_block(Xb, Y, ynorm, kk, m, y_chunk)
    boundary = dist[:, k].copy() if m > k else np.full(nq, np.inf)
    dist, idx = dist[:, :k], idx[:, :k]
    if return_boundary:
        return dist, idx, boundary
    return dist, idx


def _exact_block(Xb, Y, ynorm, kk, m, y_chunk):
    """exact_knn_f64 for one block of queries: the m best (distance, index)."""
    nb, nt = Xb.shape[0], Y.shape[0]
    xnorm = np.einsum("ij,ij->i", Xb, Xb)

    def values(t0):
        Yc = Y[t0:t0 + y_chunk]
        return ynorm[None, t0:t0 + Yc.shape[0]] - 2.0 * (Xb @ Yc.T)

    # Pass 1: the kk-th smallest expansion value of every query.
    best = np.full((nb, 0), np.inf)
    for t0 in range(0, nt, y_chunk):
        v = np.concatenate([best, values(t0)], axis=1)
        best = np.partition(v, kk - 1, axis=1)[:, :kk] if v.shape[1] > kk else v
    thr = best.max(axis=1)
    # Pass 2: every row within rounding of that value is a candidate, so exact
    # ties at the cut (duplicate rows) are all kept, whatever their number.
    margin = 1e-9 * (1.0 + np.abs(thr) + xnorm + ynorm.max())
    rows, cols = [], []
    for t0 in range(0, nt, y_chunk):
        r, c = np.nonzero(values(t0) <= (thr + margin)[:, None])
        rows.append(r)
        cols.append(c + t0)
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    diff = Y[cols] - Xb[rows]
    cd = np.einsum("ij,ij->i", diff, diff)
    # order by (query, exact distance, index); every query has >= kk >= m
    # candidates, so its first m entries are its answer
    o = np.lexsort((cols, cd, rows))
    rows, cols, cd = rows[o], cols[o], cd[o]
    start = np.searchsorted(rows, np.arange(nb))
    take = start[:, None] + np.arange(m)[None, :]
    return cd[take], cols[take].astype(np.int64)


# ----------------------------------------------------------------------------
# Reference rules (src/_pdist_aggregation.pyx)
# ----------------------------------------------------------------------------
def effective_n_threads():
    """_openmp_effective_n_threads (:41-49): OMP_NUM_THREADS wins if set,
    else min(omp max threads, physical cores)."""
    lib = _port_lib()
    omp_max = int(lib.port_omp_max_threads())
    if os.getenv("OMP_NUM_THREADS"):
        return omp_max
    try:
        from joblib import cpu_count

        phys = cpu_count(only_physical_cores=True)
    except Exception:  # pragma: no cover - joblib is installed in this image
        phys = os.cpu_count() or 1
    return max(1, min(omp_max, phys))


def chunk_geometry(n, chunk_size):
    """(samples per chunk, number of chunks) of :247-269."""
    nsc = max(MIN_CHUNK_SAMPLES, chunk_size)
    per = min(n, nsc)
    full = n // per
    return per, full + (n != full * per)


def choose_strategy(strategy, chunk_size, n_threads, n_test):
    """The 'auto' rule of :511-515."""
    if strategy == "auto":
        return "chunk_on_test" if 4 * chunk_size * n_threads < n_test else "chunk_on_train"
    return strategy


def n_parallel_chunks(strategy, n_train, n_test, chunk_size, n_threads):
    """Second return value of parallel_knn (:330, :448)."""
    s = choose_strategy(strategy, chunk_size, n_threads, n_test)
    return chunk_geometry(n_test if s == "chunk_on_test" else n_train, chunk_size)[1]


# ----------------------------------------------------------------------------
# C++/OpenMP restatement of the reference engine
# ----------------------------------------------------------------------------
_LIB = None


def _port_lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "libknn_port.so"
        if not path.exists():
            raise RuntimeError(f"{path} is not built: run `make -C oracle`")
        lib = ctypes.CDLL(str(path))
        for name, ct in (("port_parallel_knn_f64", ctypes.c_double),
                         ("port_parallel_knn_f32", ctypes.c_float)):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int64
            fn.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                           ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                           ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.c_void_p]
        lib.port_omp_max_threads.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def port_parallel_knn(X_train, X_test, k, chunk_size=CHUNK_SIZE, strategy="auto",
                      return_distance=False, n_threads=None):
    """Same call and return shape as the reference ``parallel_knn`` (:492-540).

    Returns ``((dist, idx), n_parallel_chunks)`` or ``(idx, n_parallel_chunks)``;
    distances are Euclidean (exact recompute, rows not re-sorted), exactly as
    the reference.
    """
    X_train = np.asarray(X_train)
    X_test = np.asarray(X_test)
    if X_train.dtype != X_test.dtype or X_train.dtype not in (np.float32, np.float64):
        raise TypeError("X_train and X_test must both be float32 or both float64")
    if strategy not in STRATEGIES:
        raise RuntimeError(f"strategy '{strategy}' not supported.")
    X_train = np.ascontiguousarray(X_train)
    X_test = np.ascontiguousarray(X_test)
    lib = _port_lib()
    if n_threads is None:
        n_threads = effective_n_threads()
    n_test, d = X_test.shape
    idx = np.empty((n_test, k), np.int64)
    dist = np.empty((n_test, k), X_train.dtype)
    used = ctypes.c_int(0)
    fn = lib.port_parallel_knn_f64 if X_train.dtype == np.float64 else lib.port_parallel_knn_f32
    n_chunks = fn(X_train.ctypes.data, X_train.shape[0], X_test.ctypes.data, n_test, d, k,
                  chunk_size, STRATEGIES[strategy], int(bool(return_distance)), int(n_threads),
                  idx.ctypes.data, dist.ctypes.data, ctypes.byref(used))
    if return_distance:
        return (dist, idx), int(n_chunks)
    return idx, int(n_chunks)


# ----------------------------------------------------------------------------
# Comparison helpers
# ----------------------------------------------------------------------------
def recall_at_k(idx, ref_idx):
    """Fraction of reference neighbours found (row-wise set intersection)."""
    idx = np.asarray(idx)
    ref_idx = np.asarray(ref_idx)
    hits = 0
    for a, b in zip(idx, ref_idx):
        hits += len(np.intersect1d(a, b, assume_unique=False))
    return hits / ref_idx.size


def near_tie_report(X, Y, idx, ref_dist, ref_idx, boundary, eps_rel):
    """Compare an answer with the exact one, exempting near-ties.

    ``ref_dist``/``ref_idx`` are exact squared distances and indices (sorted),
    ``boundary`` the exact (k+1)-th distance.  A rank is *ambiguous* when the
    exact distance of the reference neighbour there lies within
    ``eps_rel * max(1, |d|)`` of another candidate distance (its neighbours in
    the list or the boundary).  Returns (n_mismatch_outside_ties, n_ambiguous,
    max_abs_gap_of_swapped) where a mismatch at an ambiguous rank is allowed
    only if the returned index's exact distance is within eps of the
    reference distance at that rank.
    """
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    idx = np.asarray(idx)
    nq, k = ref_idx.shape
    full = np.concatenate([ref_dist, boundary[:, None]], axis=1)
    tol = eps_rel * np.maximum(1.0, np.abs(full))
    gap_lo = np.full((nq, k), np.inf)
    gap_hi = np.full((nq, k), np.inf)
    gap_lo[:, 1:] = full[:, 1:k] - full[:, :k - 1]
    gap_hi[:, :] = full[:, 1:k + 1] - full[:, :k]
    ambiguous = (gap_lo <= tol[:, :k]) | (gap_hi <= tol[:, :k])
    safe_idx = np.clip(idx, 0, Y.shape[0] - 1)
    diff = idx != ref_idx
    rows = np.nonzero(diff.any(axis=1))[0]
    bad = 0
    worst = 0.0
    if rows.size:
        got = _direct_sqdist(X[rows], Y, safe_idx[rows])
        d = diff[rows]
        g = np.abs(got - ref_dist[rows])
        in_range = (idx[rows] >= 0) & (idx[rows] < Y.shape[0])
        ok = ambiguous[rows] & (g <= tol[rows, :k]) & in_range
        bad = int(np.sum(d & ~ok))
        if np.any(d & ok):
            worst = float(g[d & ok].max())
    return bad, int(ambiguous.sum()), worst
