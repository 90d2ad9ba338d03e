"""Benchmark grid over datasets x estimators x k x chunk_size.

Restates the reference's harness (benchmarks/benchmark.py:99-219 with
benchmarks/config.yml) for the MI355X implementation:

* same YAML schema: ``estimators`` (name, import path, ``chunk`` flag),
  ``datasets`` (n_samples_train, n_samples_test, n_features), ``chunk_size``,
  ``n_neighbors``, ``n_trials``, ``return_distance``;
* same row fields: implementation, shapes, chunk_size, n_parallel_chunks
  (the estimator's trailing return value), n_neighbors, time_elapsed and
  ``throughput = (X_train.nbytes + X_test.nbytes) / time / 1e9`` GB/s
  (benchmark.py:140, :201), plus the reference's datastructure-size model
  (benchmark.py:41-96) next to this library's actual device workspace;
* the memory column is the peak device allocation of the call (torch's
  allocator statistics) instead of memory_profiler's host RSS;
* results go to ``<results>/<name>.csv`` (rewritten after every row, as the
  reference does) and ``<name>.json`` (environment: commit, device, config).

The reference's seaborn/PDF report is not restated.

Usage: python -m pdist_aggregation_amd.benchmark_grid NAME [--config FILE]
"""
from __future__ import annotations

import argparse
import csv
import importlib
import json
import os
import pathlib
import subprocess
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
DEFAULT_CONFIG = HERE / "bench_config.yml"


def datastructures_sizes(n, d, k, fdtype=np.float64, idtype=np.int64, parallel_on_xtrain=True):
    """The reference's per-thread memory model (benchmarks/benchmark.py:41-96),
    in bytes, for chunk size n."""
    sf = 8 if np.dtype(fdtype) == np.float64 else 4
    si = 8 if np.dtype(idtype) == np.int64 else 4
    sizes = {
        "dist_middle_terms_chunks": n * n * sf,
        "X_train_sq_norms": n * sf,
        "Xc_and_Yc": 2 * n * d * sf,
        "chunk_heaps": n * k * (sf + si),
        "heaps_indices_chunks": n * k * si if parallel_on_xtrain else 0,
    }
    sizes["total"] = sum(sizes.values())
    return sizes


def _commit():
    try:
        return subprocess.check_output(["git", "rev-parse", "--short", "HEAD"],
                                       cwd=HERE, stderr=subprocess.DEVNULL).decode().strip()
    except Exception:
        return "unknown"


def benchmark(config, results_folder, bench_name, seed=0):
    import torch
    import yaml  # noqa: F401  (config already parsed; kept for parity of deps)

    from . import _capi

    datasets = config["datasets"]
    chunk_sizes = config["chunk_size"]
    n_neighbors = config["n_neighbors"]
    estimators = config["estimators"]
    n_trials = config.get("n_trials", 3)
    return_distance = config.get("return_distance", False)
    os.makedirs(results_folder, exist_ok=True)
    env = {
        "commit": _commit(),
        "device": torch.cuda.get_device_name(0) if torch.cuda.is_available() else None,
        "torch": torch.__version__,
        "hip": getattr(torch.version, "hip", None),
        "config": config,
    }
    with open(f"{results_folder}/{bench_name}.json", "w") as f:
        json.dump(env, f, indent=1)
    rows = []
    rng = np.random.default_rng(seed)
    lib = _capi.lib()
    for dataset in datasets:
        ns_train, ns_test, nf = (int(float(v)) for v in dataset.values())
        for trial in range(n_trials):
            X_train = rng.random((ns_train, nf))
            X_test = rng.random((ns_test, nf))
            nbytes = X_train.nbytes + X_test.nbytes
            for est in estimators:
                name, path, chunk = est["name"], est["estimator"], est.get("chunk", True)
                module, cls = path.rsplit(".", 1)
                estim_class = getattr(importlib.import_module(module), cls)
                for k in n_neighbors:
                    if k > ns_train:
                        continue
                    for chunk_size in (chunk_sizes if chunk else [0]):
                        nn = estim_class(n_neighbors=k, algorithm="brute").fit(X_train)
                        kwargs = {"X": X_test, "return_distance": return_distance}
                        if chunk:
                            kwargs["chunk_size"] = chunk_size
                        torch.cuda.synchronize()
                        torch.cuda.reset_peak_memory_stats()
                        t0 = time.perf_counter()
                        res = nn.kneighbors(**kwargs)
                        torch.cuda.synchronize()
                        elapsed = time.perf_counter() - t0
                        n_parallel_chunks = res[-1] if isinstance(res, tuple) else None
                        sizes = datastructures_sizes(max(chunk_size, 1), nf, k, X_train.dtype,
                                                     np.int64, "train" in cls.lower())
                        ws = lib.knn_argkmin_workspace_size(X_train.dtype.itemsize, ns_train,
                                                            ns_test, nf, k, 0, 0)
                        row = dict(trial=trial, implementation=name, n_samples_train=ns_train,
                                   n_samples_test=ns_test, n_features=nf, chunk_size=chunk_size,
                                   n_parallel_chunks=n_parallel_chunks, n_neighbors=k,
                                   max_device_mem=torch.cuda.max_memory_allocated(),
                                   device_workspace=int(ws), time_elapsed=round(elapsed, 6),
                                   throughput=nbytes / elapsed / 1e9, **sizes)
                        rows.append(row)
                        print(row, flush=True)
                        with open(f"{results_folder}/{bench_name}.csv", "w", newline="") as f:
                            w = csv.DictWriter(f, fieldnames=list(rows[0]))
                            w.writeheader()
                            w.writerows(rows)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser("benchmark_grid")
    ap.add_argument("bench_name")
    ap.add_argument("--config", default=str(DEFAULT_CONFIG))
    ap.add_argument("--results", default=None)
    args = ap.parse_args(argv)
    import yaml

    with open(args.config) as f:
        config = yaml.safe_load(f)
    folder = args.results or f"benchmark_results/{args.bench_name}"
    benchmark(config, folder, args.bench_name)
    print(f"Benchmark results wrote in {folder}")


if __name__ == "__main__":
    main()
