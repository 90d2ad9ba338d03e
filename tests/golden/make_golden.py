"""Write the golden k-NN fixtures used by the tests.

The reference holds no stored vectors: its test (tests/test_correctness.py)
computes sklearn's brute-force NearestNeighbors at run time and compares
indices exactly and distances to 7 decimals.  This script runs that same
oracle (sklearn, third party, installed in this image) once on a subset of the
reference's parameter grid plus the 1,000 x 10,000 x 50 float64 workload, and
stores the outputs.  Inputs are NOT stored: they are regenerated from the
reference's seeding recipe, and their sha256 is stored so a changed generator
is caught.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent

# (n, d, ratio_train_test, k) points of the reference grid
# (tests/test_correctness.py:8-13), chosen to cover every n, d, ratio and k.
GRID = [
    (100, 2, 10, 1),
    (100, 5, 2, 10),
    (100, 10, 1, 100),
    (100, 100, 0.5, 10),
    (1000, 5, 10, 1000),
    (1000, 10, 2, 100),
    (1000, 100, 0.5, 10),
    (1000, 2, 1, 1),
    (10000, 2, 10, 10),
    (10000, 100, 10, 100),
    (10000, 5, 1, 1),
    (10000, 10, 0.5, 1),
]


def grid_inputs(n, d, ratio):
    """The reference test's data recipe (tests/test_correctness.py:29-33)."""
    np.random.seed(1)
    X_train = np.random.rand(int(n * d)).astype(np.float64).reshape((-1, d))
    X_test = np.random.rand(int(n * d / ratio)).astype(np.float64).reshape((-1, d))
    return X_train, X_test


def workload_inputs():
    """1,000 queries x 10,000 train rows x 50 features, float64, U[0,1),
    RandomState(0), train drawn first."""
    rs = np.random.RandomState(0)
    Y = rs.rand(10000, 50)
    X = rs.rand(1000, 50)
    return Y, X


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def grid_name(n, d, ratio, k):
    return f"grid_n{n}_d{d}_r{ratio}_k{k}".replace(".", "p")


def main():
    from sklearn.neighbors import NearestNeighbors

    manifest = {}
    cases = [(grid_name(*g), grid_inputs(*g[:3]), g[3]) for g in GRID]
    cases.append(("workload_1000x10000x50_k10", workload_inputs(), 10))
    for name, (X_train, X_test), k in cases:
        nn = NearestNeighbors(n_neighbors=k, algorithm="brute").fit(X_train)
        dist, idx = nn.kneighbors(X_test, return_distance=True)
        np.savez_compressed(HERE / f"{name}.npz", dist=dist.astype(np.float64),
                            idx=idx.astype(np.int32))
        manifest[name] = {"k": k, "train_shape": list(X_train.shape),
                          "test_shape": list(X_test.shape), "inputs_sha256": sha(X_train, X_test),
                          "oracle": "sklearn.neighbors.NearestNeighbors(algorithm='brute')"}
    (HERE / "manifest.json").write_text(json.dumps(manifest, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
