"""Host arrays in and out of parallel_knn (needs a GPU).

Inputs from 4 MiB are uploaded in 8 MiB chunks through two alternating pinned
buffers, and results come back through pinned buffers on the call's stream.
These tests check that both are byte-exact: the uploaded copy equals the array,
including a last partial chunk, and answers from host arrays equal answers
from device tensors.
"""
import sys

import numpy as np
import pytest

import pdist_aggregation_amd as pa
from pdist_aggregation_amd import knn as knn_mod

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dtype,rows,cols", [(np.float64, 50_341, 50), (np.float32, 100_003, 37),
                                             (np.float64, 1000, 50)])
def test_upload_is_exact(require_gpu, dtype, rows, cols):
    import torch

    arr = np.random.RandomState(rows).rand(rows, cols).astype(dtype)
    _say(f"[upload] {np.dtype(dtype).name} {rows} x {cols} ({arr.nbytes / 2**20:.1f} MiB)")
    s = torch.cuda.current_stream()
    got = knn_mod._upload(torch.from_numpy(arr), torch, torch.device("cuda", 0), s)
    s.synchronize()
    assert got.shape == arr.shape and got.dtype == torch.from_numpy(arr).dtype
    assert torch.equal(got.cpu(), torch.from_numpy(arr))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_host_arrays_match_device_tensors(require_gpu, dtype):
    import torch

    rs = np.random.RandomState(17)
    Y = rs.rand(40_000, 37).astype(dtype)
    X = rs.rand(3_000, 37).astype(dtype)
    _say(f"[host io] {np.dtype(dtype).name} 40000 x 3000 x 37, k=10")
    (d1, i1), _ = pa.parallel_knn(Y, X, 10, return_distance=True)
    (d2, i2), _ = pa.parallel_knn(torch.from_numpy(Y).cuda(), torch.from_numpy(X).cuda(), 10,
                                  return_distance=True)
    assert isinstance(i1, np.ndarray) and isinstance(d1, np.ndarray)
    np.testing.assert_array_equal(i1, i2.cpu().numpy())
    np.testing.assert_array_equal(d1, d2.cpu().numpy())
    i3, _ = pa.parallel_knn(Y, X, 10)
    np.testing.assert_array_equal(i1, i3)
