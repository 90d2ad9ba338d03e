"""Call surface of the reference's SYCL backend, on the MI355X kernels.

Mirrors sycl_ext/_pdist_aggregation_sycl.pyx and sycl_ext/nn.py:
``parallel_knn(X_train, X_test, k, queue, strategy="auto",
return_distance=False)`` returns ``(indices, distances)`` or ``indices``
(no chunk count) with SQUARED distances, and raises ``ValueError`` for a
strategy other than "auto", k == 0 or a dimension mismatch (:48-55).  The
``queue`` argument (a dpctl.SyclQueue there) is an optional torch CUDA stream
here.
"""
from __future__ import annotations

from .knn import parallel_knn as _parallel_knn
from .knn import sqdist_matrix as _sqdist_matrix


def parallel_knn(X_train, X_test, k, queue=None, strategy="auto", return_distance=False):
    if strategy != "auto":
        raise ValueError("Keyword strategy's only supported value is 'auto'")
    if k == 0:
        raise ValueError("Number requested nearest neighbors must be non-zero")
    if X_train.shape[1] != X_test.shape[1]:
        raise ValueError("Dimensionality of test and train datasets must be the same")
    (dist, idx), _ = _parallel_knn(X_train, X_test, k, strategy="auto", return_distance=True,
                                   squared=True, stream=queue)
    return (idx, dist) if return_distance else idx


def distance_squared_matrix(X_train, X_test, queue=None):
    """(n_test x n_train) squared distances (:73-99), direct differences."""
    if X_train.shape[1] != X_test.shape[1]:
        raise ValueError("Dimensionality of test and train datasets must be the same")
    return _sqdist_matrix(X_train, X_test, stream=queue)


class NearestNeighbors:
    """sycl_ext/nn.py:8-33 with a torch stream in place of the SYCL queue."""

    def __init__(self, n_neighbors=1, algorithm="hip", queue=None):
        self.n_neighbors = n_neighbors
        self.algorithm = algorithm
        self.queue = queue

    def fit(self, X):
        from .nn import _check_array

        self._X_train = _check_array(X)
        return self

    def kneighbors(self, X, return_distance=False):
        from .nn import _check_array

        X = _check_array(X)
        return parallel_knn(self._X_train, X, k=self.n_neighbors, queue=self.queue,
                            strategy="auto", return_distance=return_distance)
