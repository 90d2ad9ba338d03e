import json
import os
import pathlib
import sys

# The library reads its testing switches (PDIST_KNN_GUESS, PDIST_KNN_SPLITS,
# PDIST_KNN_WIDEN) only with this opt-in; the tests set them per case.
os.environ["PDIST_KNN_TESTING"] = "1"

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
GOLDEN = ROOT / "tests" / "golden"
if str(GOLDEN) not in sys.path:
    sys.path.insert(0, str(GOLDEN))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: full-size workloads")


def golden_manifest():
    return json.loads((GOLDEN / "manifest.json").read_text())


def load_golden(name):
    """(X_train, X_test, k, sklearn_dist, sklearn_idx) of a committed fixture."""
    import make_golden

    meta = golden_manifest()[name]
    if name.startswith("grid_"):
        g = next(g for g in make_golden.GRID if make_golden.grid_name(*g) == name)
        X_train, X_test = make_golden.grid_inputs(*g[:3])
    else:
        X_train, X_test = make_golden.workload_inputs()
    assert make_golden.sha(X_train, X_test) == meta["inputs_sha256"], name
    z = np.load(GOLDEN / f"{name}.npz")
    return X_train, X_test, meta["k"], z["dist"], z["idx"].astype(np.int64)


@pytest.fixture
def require_gpu():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch
