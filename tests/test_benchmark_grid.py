"""The restated benchmark harness (reference benchmarks/benchmark.py)."""
import csv

import numpy as np
import pytest
import yaml

from pdist_aggregation_amd import benchmark_grid


def test_datastructure_sizes_model():
    """benchmarks/benchmark.py:41-96 of the reference, for float64/int64."""
    s = benchmark_grid.datastructures_sizes(256, 50, 10, np.float64, np.int64, True)
    assert s["dist_middle_terms_chunks"] == 256 * 256 * 8
    assert s["X_train_sq_norms"] == 256 * 8
    assert s["Xc_and_Yc"] == 2 * 256 * 50 * 8
    assert s["chunk_heaps"] == 256 * 10 * 16
    assert s["heaps_indices_chunks"] == 256 * 10 * 8
    assert s["total"] == sum(v for k, v in s.items() if k != "total")
    s2 = benchmark_grid.datastructures_sizes(256, 50, 10, np.float32, np.int64, False)
    assert s2["heaps_indices_chunks"] == 0 and s2["dist_middle_terms_chunks"] == 256 * 256 * 4


def test_default_config_schema():
    cfg = yaml.safe_load(open(benchmark_grid.DEFAULT_CONFIG))
    assert {"estimators", "datasets", "chunk_size", "n_neighbors"} <= set(cfg)
    for d in cfg["datasets"]:
        assert list(d) == ["n_samples_train", "n_samples_test", "n_features"]


@pytest.mark.gpu
def test_grid_runs_and_writes_csv(require_gpu, tmp_path):
    cfg = {"estimators": [{"name": "auto",
                           "estimator": "pdist_aggregation_amd.NearestNeighborsParallelAuto",
                           "chunk": True}],
           "datasets": [{"n_samples_train": 3000, "n_samples_test": 100, "n_features": 20}],
           "chunk_size": [64, 256], "n_neighbors": [1, 10], "n_trials": 1,
           "return_distance": True}
    rows = benchmark_grid.benchmark(cfg, str(tmp_path), "t")
    assert len(rows) == 4
    with open(tmp_path / "t.csv") as f:
        got = list(csv.DictReader(f))
    assert len(got) == 4
    for r in got:
        assert float(r["throughput"]) > 0
        chunk = int(r["chunk_size"])
        assert int(r["n_parallel_chunks"]) == -(-3000 // min(3000, max(20, chunk)))
