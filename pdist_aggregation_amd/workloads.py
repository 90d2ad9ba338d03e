"""Timed runs of the reference configurations beyond the headline benchmark.

``bench.py`` measures the headline (1e4 x 1e6 x 128, FP32 vs BF16 + re-rank).
This module times the other configurations on the same synthetic inputs the
GPU tests check (same seeds, same draw order), through the C ABI with the
library's per-stage HIP-event timing (``KNN_FLAG_TIMING``), inputs resident
in device memory:

* ``plain-f64``: X 1,000 x 50 and Y 10,000 x 50 float64 U[0,1) from
  numpy RandomState(0) (Y drawn first), k = 10 -- the reference's
  CPU-runnable case, the inputs of tests/golden/make_golden.py;
* ``grid-f64-d500``: Y 100,000 x 500 and X 1,000 x 500 float64 U[0,1) from
  RandomState(1), k = 10 -- the largest float64 shape of the reference's
  benchmark grid, on the 16 x 16 x 4 float64 MFMA kernel;
* ``lowd``: 65,536 x 65,536 points on the unit sphere with N(0, 0.01^2)
  radial jitter, d = 3, k = 16, float32; forward, then forward + backward
  through ``knn_distances``;
* ``bf16-mixture``: 100 Gaussians (centres U[-1,1]^128, sigma 0.1),
  1e5 queries x 1e6 train rows, k = 32, k' = 128;
* ``fp8-mixture``: 1,000 Gaussians (centres U[-1,1]^256, sigma 0.05),
  5e4 x 2e6 with 20,000 train rows overwritten by copies of other rows,
  k = 100, k' = 400.

Usage: python -m pdist_aggregation_amd.workloads [NAME ...] [--reps R]
One line per timed call: call wall time, prep / search / merge stage times,
search throughput, kernel path, split count and stats counters.
"""
from __future__ import annotations

import argparse
import ctypes
import time

import numpy as np

from . import _capi

PRECISION = {"exact": 0, "bf16": 1, "fp8": 2}


def _mixture(seed, n_centres, d, sigma, n_train, n_test):
    g = np.random.default_rng(seed)
    centres = g.uniform(-1.0, 1.0, (n_centres, d)).astype(np.float32)
    Y = centres[g.integers(0, n_centres, n_train)]
    Y += sigma * g.standard_normal((n_train, d), dtype=np.float32)
    X = centres[g.integers(0, n_centres, n_test)]
    X += sigma * g.standard_normal((n_test, d), dtype=np.float32)
    return g, Y, X


def _sphere(rs, n):
    u = rs.standard_normal((n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    return u * (1.0 + 0.01 * rs.standard_normal((n, 1)))


def _inputs(name):
    """(Y, X, k, precision, oversample) of a workload, host arrays."""
    if name == "plain-f64":
        rs = np.random.RandomState(0)
        Y = rs.rand(10_000, 50)
        X = rs.rand(1_000, 50)
        return Y, X, 10, "exact", 0
    if name == "grid-f64-d500":
        # the largest float64 shape of the reference benchmark grid
        # (benchmarks/config.yml of the reference: 100,000 train x 1,000 test x 500)
        rs = np.random.RandomState(1)
        Y = rs.rand(100_000, 500)
        X = rs.rand(1_000, 500)
        return Y, X, 10, "exact", 0
    if name == "lowd":
        rs = np.random.RandomState(0)
        X = _sphere(rs, 65536).astype(np.float32)
        Y = _sphere(rs, 65536).astype(np.float32)
        return Y, X, 16, "exact", 0
    if name == "bf16-mixture":
        _, Y, X = _mixture(4, 100, 128, 0.1, 1_000_000, 100_000)
        return Y, X, 32, "bf16", 128
    if name == "fp8-mixture":
        g, Y, X = _mixture(5, 1000, 256, 0.05, 2_000_000, 50_000)
        dst = g.choice(Y.shape[0], 20_000, replace=False)
        Y[dst] = Y[g.integers(0, Y.shape[0], 20_000)]
        return Y, X, 100, "fp8", 400
    raise ValueError(f"unknown workload {name!r}")


def run(name, reps=3):
    import torch

    lib = _capi.lib()
    Y, X, k, precision, over = _inputs(name)
    Yd, Xd = torch.from_numpy(Y).cuda(), torch.from_numpy(X).cuda()
    del Y, X
    n_t, d = Yd.shape
    n_q = Xd.shape[0]
    is64 = Yd.dtype == torch.float64
    prec = PRECISION[precision]
    ws_b = lib.knn_argkmin_workspace_size(8 if is64 else 4, n_t, n_q, d, k, prec, over)
    ws = torch.empty(ws_b, dtype=torch.uint8, device="cuda")
    dist = torch.empty((n_q, k), dtype=Yd.dtype, device="cuda")
    idx = torch.empty((n_q, k), dtype=torch.int64, device="cuda")
    fn = lib.knn_argkmin_f64 if is64 else lib.knn_argkmin_f32
    stream = torch.cuda.current_stream()
    flags = _capi.FLAG_SQUARED | _capi.FLAG_TIMING
    for it in range(reps + 1):  # the first call warms up and is not reported
        st = _capi.KnnStats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        code = fn(Yd.data_ptr(), n_t, Xd.data_ptr(), n_q, d, k, 4096, 0, prec, over, flags, 0,
                  dist.data_ptr(), idx.data_ptr(), ws.data_ptr(), ws_b,
                  ctypes.c_void_p(stream.cuda_stream), ctypes.byref(st))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if code != 0:
            raise RuntimeError(lib.knn_last_error().decode())
        if it:
            rate = 2.0 * n_q * n_t * d / (st.search_ms * 1e-3) / 1e12
            print(f"{name}: call {1e3 * (t1 - t0):.3f} ms (prep {st.prep_ms:.3f}, search "
                  f"{st.search_ms:.3f} = {rate:.1f} TFLOP/s, merge {st.merge_ms:.3f}); "
                  f"path {st.kernel_path}, splits {st.n_splits}, k' {st.k_candidates}, "
                  f"uncertified {st.n_uncertified}, widened {st.n_widened} "
                  f"({st.n_widened_from_lists} from the split lists), "
                  f"re-searched {st.n_rescued}", flush=True)
    del ws
    if name == "grid-f64-d500":
        # the largest float64 shape of the reference benchmark grid
        # (benchmarks/config.yml of the reference: 100,000 train x 1,000 test x 500)
        rs = np.random.RandomState(1)
        Y = rs.rand(100_000, 500)
        X = rs.rand(1_000, 500)
        return Y, X, 10, "exact", 0
    if name == "lowd":
        from .autograd import knn_distances

        for it in range(reps + 1):
            Yg = Yd.clone().requires_grad_(True)
            Xg = Xd.clone().requires_grad_(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            dd, _ = knn_distances(Yg, Xg, k, squared=True)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            dd.sum().backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if it:
                print(f"{name}: autograd forward {1e3 * (t1 - t0):.3f} ms, backward "
                      f"{1e3 * (t2 - t1):.3f} ms ({n_q * k} (query, neighbour) pairs)", flush=True)
    del Yd, Xd, dist, idx
    torch.cuda.empty_cache()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("names", nargs="*",
                    default=["plain-f64", "grid-f64-d500", "lowd", "bf16-mixture",
                             "fp8-mixture"])
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)
    for name in args.names:
        run(name, args.reps)


if __name__ == "__main__":
    main()
