"""``parallel_knn``: the reference's Python entry point, on the MI355X kernels.

Mirrors ``parallel_knn`` of the reference's Cython module
(src/_pdist_aggregation.pyx:492-540): same positional order (train set
first), same keywords and defaults, same return shape
``((distances, indices), n_parallel_chunks)`` or ``(indices,
n_parallel_chunks)``, Euclidean distances by default.  Extra keyword-only
arguments select the candidate precision, the oversampling, squared
distances and the stream.

Inputs may be numpy arrays (copied to the GPU and back) or torch tensors
already on the GPU (no host copies; results come back as GPU tensors).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _capi

CHUNK_SIZE = 256  # src/_pdist_aggregation.pyx:24 (default of parallel_knn)


def effective_n_threads():
    """The reference's _openmp_effective_n_threads (:41-49).

    OMP_NUM_THREADS wins if set; otherwise min(CPU count, physical cores).
    It only feeds the "auto" strategy rule and n_parallel_chunks.
    """
    env = os.getenv("OMP_NUM_THREADS")
    if env:
        try:
            return max(1, int(env.split(",")[0]))
        except ValueError:
            pass
    n = os.cpu_count() or 1
    try:
        from joblib import cpu_count

        return max(1, min(n, cpu_count(only_physical_cores=True)))
    except Exception:  # joblib absent: every CPU counts
        return n


def resolve_strategy(strategy, chunk_size, n_threads, n_test):
    """The "auto" rule of src/_pdist_aggregation.pyx:511-515."""
    if strategy == "auto":
        return "chunk_on_test" if 4 * chunk_size * n_threads < n_test else "chunk_on_train"
    return strategy


def n_parallel_chunks(strategy, n_train, n_test, chunk_size, n_threads):
    """Second return value of the reference's parallel_knn: the number of
    chunks of the axis its strategy parallelises (:247-269, :330, :448)."""
    s = resolve_strategy(strategy, chunk_size, n_threads, n_test)
    n = n_test if s == "chunk_on_test" else n_train
    per = min(n, max(20, int(chunk_size)))
    if per <= 0:
        return 0
    full = n // per
    return full + (n != full * per)


def _torch():
    import torch

    return torch


_UPLOAD_MIN_BYTES = 4 << 20  # host arrays from this size go through pinned staging
_UPLOAD_CHUNK_BYTES = 8 << 20


def _upload(src, torch, device, stream):
    """Copy a contiguous host tensor to ``device``, ordered on ``stream``.

    From _UPLOAD_MIN_BYTES on, the array is copied in 8 MiB chunks into two
    alternating page-locked buffers (torch's caching host allocator) and each
    chunk is sent with an asynchronous copy while the CPU fills the other
    buffer: about 50 GB/s on the MI355X's PCIe link against about 28 GB/s for
    a pageable copy (400 MiB float64: 14.7 -> 8.1 ms).
    """
    src = src.detach()  # the search is not differentiated through this path
    if src.numel() * src.element_size() < _UPLOAD_MIN_BYTES:
        with torch.cuda.stream(stream):
            return src.to(device)
    flat = src.reshape(-1)
    with torch.cuda.stream(stream):
        out = torch.empty(src.shape, dtype=src.dtype, device=device)
    dst = out.view(-1)
    n = flat.numel()
    step = _UPLOAD_CHUNK_BYTES // flat.element_size()
    bufs = [torch.empty(step, dtype=flat.dtype, pin_memory=True) for _ in range(2)]
    done = [None, None]  # event after the last copy out of each buffer
    with torch.cuda.stream(stream):
        for i, s0 in enumerate(range(0, n, step)):
            j = i & 1
            if done[j] is not None:
                done[j].synchronize()
            m = min(step, n - s0)
            bufs[j][:m].copy_(flat[s0:s0 + m])
            dst[s0:s0 + m].copy_(bufs[j][:m], non_blocking=True)
            done[j] = torch.cuda.Event()
            done[j].record(stream)
    return out


def _call_stream(torch, device, stream):
    """The stream a call runs on.  A stream other than the current one first
    waits for the work already queued on the current stream (which may be
    producing the inputs)."""
    cur = torch.cuda.current_stream(device)
    if stream is None or stream == cur:
        return cur
    stream.wait_stream(cur)
    return stream


def _as_device(a, name, torch, device, stream):
    """Return (tensor on GPU, came_from_host)."""
    if isinstance(a, torch.Tensor):
        if a.dim() != 2:
            raise ValueError(f"{name} must be a 2-D array")
        if not a.is_contiguous():
            raise ValueError(f"{name} is not C-contiguous")
        if a.device.type != "cuda":
            return _upload(a, torch, device, stream), True
        if stream != torch.cuda.current_stream(a.device):
            # the caller may free it on its own stream while `stream` reads it
            a.record_stream(stream)
        return a, False
    arr = np.asarray(a)
    if arr.ndim != 2:
        raise ValueError(f"{name} must be a 2-D array")
    if not arr.flags.c_contiguous:  # the reference's floating[:, ::1] memoryview
        raise ValueError(f"{name} is not C-contiguous")
    return _upload(torch.from_numpy(arr), torch, device, stream), True


def _to_host(tensors, torch, stream):
    """numpy copies of device tensors, through pinned host buffers.

    The copies are queued on ``stream`` (after the search).  Pageable copies
    of the results ran at a few GB/s (1e5 x 100 float64 distances + int64
    indices: about 40 ms).  Results up to _UPLOAD_CHUNK_BYTES land in
    page-locked buffers of their own size from torch's caching host
    allocator (kept for reuse by later calls); larger ones stream through two
    such buffers of _UPLOAD_CHUNK_BYTES into pageable numpy arrays, so the
    page-locked memory a process keeps stays bounded whatever the result size.
    """
    outs = [None] * len(tensors)
    small = [i for i, t in enumerate(tensors)
             if t.numel() * t.element_size() <= _UPLOAD_CHUNK_BYTES]
    hosts = {i: torch.empty(tensors[i].shape, dtype=tensors[i].dtype, pin_memory=True)
             for i in small}
    with torch.cuda.stream(stream):
        for i, h in hosts.items():
            h.copy_(tensors[i], non_blocking=True)
    for i, t in enumerate(tensors):
        if i not in hosts:
            outs[i] = _download_chunked(t, torch, stream)
    stream.synchronize()
    for i, h in hosts.items():
        outs[i] = h.numpy()
    return outs


def _download_chunked(t, torch, stream):
    """Copy a contiguous device tensor to a new numpy array through two
    alternating page-locked buffers of _UPLOAD_CHUNK_BYTES: the device copy of
    one chunk overlaps the host copy of the previous one."""
    flat = t.reshape(-1)
    out = np.empty(t.shape, dtype=torch.empty(0, dtype=t.dtype).numpy().dtype)
    dst = out.reshape(-1)
    step = _UPLOAD_CHUNK_BYTES // flat.element_size()
    bufs = [torch.empty(step, dtype=flat.dtype, pin_memory=True) for _ in range(2)]
    pending = []  # (event, buffer index, start, length), oldest first
    with torch.cuda.stream(stream):
        for i, s0 in enumerate(range(0, flat.numel(), step)):
            j = i & 1
            if len(pending) == 2:  # buffer j holds the chunk before the last: drain it
                ev, jb, a, m = pending.pop(0)
                ev.synchronize()
                dst[a:a + m] = bufs[jb][:m].numpy()
            m = min(step, flat.numel() - s0)
            bufs[j][:m].copy_(flat[s0:s0 + m], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
            pending.append((ev, j, s0, m))
    for ev, jb, a, m in pending:
        ev.synchronize()
        dst[a:a + m] = bufs[jb][:m].numpy()
    return out


def parallel_knn(X_train, X_test, k, chunk_size=CHUNK_SIZE, strategy="auto",
                 return_distance=False, *, squared=False, precision="exact",
                 oversample=None, stream=None, return_stats=False, n_threads=None):
    """k nearest rows of ``X_train`` for every row of ``X_test``.

    Parameters follow src/_pdist_aggregation.pyx:492-499.  ``strategy`` and
    ``chunk_size`` do not change the GPU schedule; they determine the
    returned ``n_parallel_chunks`` exactly as the reference computes it.

    Keyword-only extensions
    -----------------------
    squared : return squared distances (the SYCL backend's convention).
    precision : "exact" (search in the input type), "bf16" or "fp8" (low
        precision candidates, then exact re-rank; float32 inputs only).
    oversample : k' candidates kept before the exact re-rank.
    stream : a torch.cuda.Stream (default: the current stream).
    return_stats : append a dict of call statistics to the return tuple.

    Rows are sorted by exact distance; equal distances list the lower train
    index first.
    """
    if strategy not in _capi.STRATEGY:
        raise RuntimeError(f"strategy '{strategy}' not supported.")
    if precision not in _capi.PRECISION:
        raise ValueError(f"precision '{precision}' not supported")
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("pdist_aggregation_amd needs a visible AMD GPU (no CPU fallback)")
    lib = _capi.lib()
    device = X_train.device if isinstance(X_train, torch.Tensor) and X_train.is_cuda \
        else torch.device("cuda", torch.cuda.current_device())
    s = _call_stream(torch, device, stream)
    Ytr, host_in = _as_device(X_train, "X_train", torch, device, s)
    Xte, host_q = _as_device(X_test, "X_test", torch, device, s)
    if Ytr.dtype != Xte.dtype or Ytr.dtype not in (torch.float32, torch.float64):
        raise TypeError("X_train and X_test must both be float32 or both float64")
    if Ytr.shape[1] != Xte.shape[1]:
        raise ValueError("Dimensionality of test and train datasets must be the same")
    k = int(k)
    n_train, dim = Ytr.shape
    n_test = Xte.shape[0]
    if k <= 0:
        raise ValueError("Number requested nearest neighbors must be non-zero")
    if k > n_train:
        raise ValueError(f"k (={k}) must not exceed the number of train rows (={n_train})")
    is64 = Ytr.dtype == torch.float64
    prec = _capi.PRECISION[precision]
    ov = int(oversample) if oversample else 0
    ws_bytes = lib.knn_argkmin_workspace_size(8 if is64 else 4, n_train, n_test, dim, k,
                                              prec, ov)
    if ws_bytes == 0:
        _capi.check(_capi.KNN_EINVAL)
    # Allocated on the call's stream: torch's caching allocator hands a freed
    # block only to later work on the stream it was allocated on, so the
    # workspace freed at return cannot be reused while the kernels on `s`
    # still run.
    with torch.cuda.stream(s):
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=device)
        dist = torch.empty((n_test, k), dtype=Ytr.dtype, device=device)
        idx = torch.empty((n_test, k), dtype=torch.int64, device=device)
    stats = _capi.KnnStats()
    fn = lib.knn_argkmin_f64 if is64 else lib.knn_argkmin_f32
    threads = effective_n_threads() if n_threads is None else int(n_threads)
    with torch.cuda.device(device):
        # Passing stats makes the call wait for the GPU; only on request.
        code = fn(Ytr.data_ptr(), n_train, Xte.data_ptr(), n_test, dim, k, int(chunk_size),
                  _capi.STRATEGY[strategy], prec, ov, _capi.FLAG_SQUARED if squared else 0,
                  threads, dist.data_ptr(), idx.data_ptr(), ws.data_ptr(), ws_bytes,
                  ctypes.c_void_p(s.cuda_stream),
                  ctypes.byref(stats) if return_stats else None)
    _capi.check(code)
    n_chunks = n_parallel_chunks(strategy, n_train, n_test, chunk_size, threads)
    if not (isinstance(X_test, torch.Tensor) and X_test.is_cuda):
        if return_distance:
            dist_o, idx_o = _to_host([dist, idx], torch, s)
        else:
            dist_o, (idx_o,) = None, _to_host([idx], torch, s)
    else:
        dist_o, idx_o = dist, idx
    out = ((dist_o, idx_o), n_chunks) if return_distance else (idx_o, n_chunks)
    if return_stats:
        st = stats.as_dict()
        st["strategy_used"] = _capi.STRATEGY_NAME.get(st["strategy_used"], "?")
        return out + (st,)
    return out


def sqdist_matrix(X_train, X_test, stream=None):
    """Dense (n_test x n_train) squared distances by direct differences.

    The SYCL backend's ``distance_squared_matrix``
    (sycl_ext/_pdist_aggregation_sycl.pyx:73-99) on the MI355X.  numpy in,
    numpy out; GPU tensors in, GPU tensor out.
    """
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("pdist_aggregation_amd needs a visible AMD GPU (no CPU fallback)")
    lib = _capi.lib()
    device = X_train.device if isinstance(X_train, torch.Tensor) and X_train.is_cuda \
        else torch.device("cuda", torch.cuda.current_device())
    s = _call_stream(torch, device, stream)
    Ytr, _ = _as_device(X_train, "X_train", torch, device, s)
    Xte, _ = _as_device(X_test, "X_test", torch, device, s)
    if Ytr.dtype != Xte.dtype or Ytr.dtype not in (torch.float32, torch.float64):
        raise TypeError("X_train and X_test must both be float32 or both float64")
    if Ytr.shape[1] != Xte.shape[1]:
        raise ValueError("Dimensionality of test and train datasets must be the same")
    with torch.cuda.stream(s):
        out = torch.empty((Xte.shape[0], Ytr.shape[0]), dtype=Ytr.dtype, device=device)
    fn = lib.knn_sqdist_matrix_f64 if Ytr.dtype == torch.float64 else lib.knn_sqdist_matrix_f32
    with torch.cuda.device(device):
        code = fn(Ytr.data_ptr(), Ytr.shape[0], Xte.data_ptr(), Xte.shape[0], Ytr.shape[1],
                  out.data_ptr(), ctypes.c_void_p(s.cuda_stream))
    _capi.check(code)
    if not (isinstance(X_test, torch.Tensor) and X_test.is_cuda):
        return _to_host([out], torch, s)[0]
    return out
