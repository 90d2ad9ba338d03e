"""sklearn-style estimator shells over :func:`parallel_knn`.

Mirror of the reference's src/nn.py:9-49: ``NearestNeighbors`` with a
``_strategy`` hook and three concrete classes, ``fit`` storing the train set
as given (:15-17) and ``kneighbors(X, chunk_size=4096, return_distance=False)``
validating X as a C-ordered array (:24) and returning whatever
``parallel_knn`` returns, including the trailing ``n_parallel_chunks``.
"""
from __future__ import annotations

import abc

import numpy as np

from .knn import parallel_knn


def _check_array(X):
    """sklearn's check_array(X, order="C") when sklearn is importable."""
    try:
        from sklearn.utils.validation import check_array
    except ImportError:  # pragma: no cover - sklearn is present in this image
        return np.ascontiguousarray(X)
    return check_array(X, order="C")


class NearestNeighbors:
    def __init__(self, n_neighbors=1, algorithm="brute"):
        self.n_neighbors = n_neighbors
        self.algorithm = algorithm

    def fit(self, X):
        self._X_train = X
        return self

    @abc.abstractmethod
    def _strategy(self):
        pass

    def kneighbors(self, X, chunk_size=4096, return_distance=False, **kwargs):
        """Extra keywords (precision, oversample, squared, ...) are passed on."""
        try:
            import torch

            is_dev = isinstance(X, torch.Tensor) and X.is_cuda
        except ImportError:  # pragma: no cover
            is_dev = False
        if not is_dev:
            X = _check_array(X)
        return parallel_knn(
            self._X_train,
            X,
            k=self.n_neighbors,
            chunk_size=chunk_size,
            strategy=self._strategy(),
            return_distance=return_distance,
            **kwargs,
        )


class NearestNeighborsParallelXtrain(NearestNeighbors):
    def _strategy(self):
        return "chunk_on_train"


class NearestNeighborsParallelXtest(NearestNeighbors):
    def _strategy(self):
        return "chunk_on_test"


class NearestNeighborsParallelAuto(NearestNeighbors):
    def _strategy(self):
        return "auto"
