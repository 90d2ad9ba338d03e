"""MI355X-native brute-force k-NN argkmin, a drop-in for pdist_aggregation.

Public surface (same names and meaning as the reference):

* :func:`parallel_knn` - src/_pdist_aggregation.pyx:492-540
* :class:`NearestNeighborsParallelAuto`, :class:`NearestNeighborsParallelXtrain`,
  :class:`NearestNeighborsParallelXtest` - src/nn.py:9-49
* :mod:`pdist_aggregation_amd.sycl_compat` - the SYCL backend's surface
  (sycl_ext/_pdist_aggregation_sycl.pyx, sycl_ext/nn.py)

All compute runs in hand-written HIP kernels for gfx950 behind the C ABI of
include/pdist_knn.h (libpdist_knn.so).
"""
from .knn import effective_n_threads, n_parallel_chunks, parallel_knn  # noqa: F401
from .nn import (  # noqa: F401
    NearestNeighbors,
    NearestNeighborsParallelAuto,
    NearestNeighborsParallelXtest,
    NearestNeighborsParallelXtrain,
)

__all__ = [
    "parallel_knn",
    "effective_n_threads",
    "n_parallel_chunks",
    "NearestNeighbors",
    "NearestNeighborsParallelAuto",
    "NearestNeighborsParallelXtest",
    "NearestNeighborsParallelXtrain",
]
