"""Differentiable k-NN distances on the MI355X kernels.

``knn_distances(X_train, X_test, k)`` runs the same search as
:func:`pdist_aggregation_amd.parallel_knn` and returns ``(dist, idx)`` as GPU
tensors; ``dist`` carries gradients to both point sets.  The neighbour
indices are fixed by the forward pass; the backward kernel
(``knn_argkmin_backward_*``) recomputes the differences x_i - y_j instead of
storing them and scatters into the train-set gradient with atomics.
"""
from __future__ import annotations

import ctypes

import torch

from . import _capi
from .knn import parallel_knn


class _KnnDistances(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X_train, X_test, k, squared, precision, oversample):
        (dist, idx), _ = parallel_knn(X_train.detach(), X_test.detach(), k,
                                      return_distance=True, squared=squared,
                                      precision=precision, oversample=oversample)
        ctx.save_for_backward(X_train, X_test, idx)
        ctx.squared = squared
        ctx.mark_non_differentiable(idx)
        return dist, idx

    @staticmethod
    def backward(ctx, grad_dist, _grad_idx):
        X_train, X_test, idx = ctx.saved_tensors
        need_y, need_x = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_x or need_y):
            return None, None, None, None, None, None
        grad = grad_dist.contiguous().to(X_train.dtype)
        gy = torch.empty_like(X_train) if need_y else None
        gx = torch.empty_like(X_test) if need_x else None
        lib = _capi.lib()
        fn = lib.knn_argkmin_backward_f64 if X_train.dtype == torch.float64 \
            else lib.knn_argkmin_backward_f32
        stream = torch.cuda.current_stream(X_train.device)
        with torch.cuda.device(X_train.device):
            code = fn(X_train.data_ptr(), X_train.shape[0], X_test.data_ptr(), X_test.shape[0],
                      X_train.shape[1], idx.shape[1], idx.data_ptr(), grad.data_ptr(),
                      _capi.FLAG_SQUARED if ctx.squared else 0,
                      gy.data_ptr() if gy is not None else None,
                      gx.data_ptr() if gx is not None else None,
                      ctypes.c_void_p(stream.cuda_stream))
        _capi.check(code)
        return gy, gx, None, None, None, None


def knn_distances(X_train, X_test, k, squared=True, precision="exact", oversample=None):
    """k-NN distances (n_test x k) and indices of ``X_test`` rows in ``X_train``.

    Both inputs must be contiguous float32/float64 tensors on the GPU;
    gradients flow to whichever of them requires grad.
    """
    for name, t in (("X_train", X_train), ("X_test", X_test)):
        if not (isinstance(t, torch.Tensor) and t.is_cuda):
            raise TypeError(f"{name} must be a torch tensor on the GPU")
        if not t.is_contiguous():
            raise ValueError(f"{name} is not C-contiguous")
    return _KnnDistances.apply(X_train, X_test, int(k), bool(squared), precision, oversample)
