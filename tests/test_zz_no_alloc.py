"""knn_argkmin_f32_noalloc against knn_argkmin_f32 (needs a GPU).

The allocation-free entry point runs the same schedule and kernels on a
static per-GPU scratch array instead of a workspace.  Through the C ABI,
two calls in a row (the second with a smaller input, so it reads scratch the
first one left behind) must give the same rows and distances, bit for bit,
as knn_argkmin_f32 with a workspace, for the exact, BF16 and FP8 searches and
the generic and low-d kernels; each result also matches the FP64 oracle.  A
request whose workspace exceeds the scratch size and one asking for stage
timing are rejected before any GPU work.
"""
import ctypes
import sys

import numpy as np
import pytest

from pdist_aggregation_amd import _capi
from knn_checks import EPS_F32, assert_knn_parity

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _call(torch, lib, Y, X, k, precision, noalloc):
    Yd = torch.from_numpy(Y).cuda()
    Xd = torch.from_numpy(X).cuda()
    nt, d = Y.shape
    nq = X.shape[0]
    dist = torch.full((nq, k), -1.0, dtype=torch.float32, device="cuda")
    idx = torch.full((nq, k), -7, dtype=torch.int64, device="cuda")
    st = _capi.KnnStats()
    stream = torch.cuda.current_stream()
    prec = _capi.PRECISION[precision]
    if noalloc:
        code = lib.knn_argkmin_f32_noalloc(Yd.data_ptr(), nt, Xd.data_ptr(), nq, d, k, 256, 0,
                                           prec, 0, _capi.FLAG_SQUARED, 8, dist.data_ptr(),
                                           idx.data_ptr(), ctypes.c_void_p(stream.cuda_stream),
                                           ctypes.byref(st))
    else:
        wsb = lib.knn_argkmin_workspace_size(4, nt, nq, d, k, prec, 0)
        ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
        code = lib.knn_argkmin_f32(Yd.data_ptr(), nt, Xd.data_ptr(), nq, d, k, 256, 0, prec, 0,
                                   _capi.FLAG_SQUARED, 8, dist.data_ptr(), idx.data_ptr(),
                                   ws.data_ptr(), wsb, ctypes.c_void_p(stream.cuda_stream),
                                   ctypes.byref(st))
    _capi.check(code)
    torch.cuda.synchronize()
    return dist.cpu().numpy(), idx.cpu().numpy(), st.as_dict()


# (precision, d, k, (n_train, n_test) of the first call, of the second call)
CASES = [("exact", 64, 10, (36000, 200), (25000, 130)),
         ("bf16", 128, 10, (33000, 250), (20000, 100)),
         ("fp8", 128, 10, (35000, 200), (20000, 129)),
         ("exact", 300, 16, (20000, 400), (9000, 100)),   # generic kernel
         ("exact", 3, 16, (30000, 500), (20000, 300))]    # low-d kernel


@pytest.mark.parametrize("precision,d,k,first,second", CASES)
def test_noalloc_matches_workspace_entry_point(require_gpu, precision, d, k, first, second):
    torch = require_gpu
    lib = _capi.lib()
    rs = np.random.RandomState(d * 31 + k)
    for nt, nq in (first, second):
        Y = rs.rand(nt, d).astype(np.float32)
        X = rs.rand(nq, d).astype(np.float32)
        wsb = lib.knn_argkmin_workspace_size(4, nt, nq, d, k, _capi.PRECISION[precision], 0)
        assert 0 < wsb <= _capi.NOALLOC_SCRATCH_BYTES, wsb
        _say(f"[noalloc] {precision} {nq} x {nt} x {d} k={k} workspace {wsb} bytes")
        d0, i0, s0 = _call(torch, lib, Y, X, k, precision, noalloc=True)
        d1, i1, s1 = _call(torch, lib, Y, X, k, precision, noalloc=False)
        np.testing.assert_array_equal(i0, i1)
        np.testing.assert_array_equal(d0, d1)
        for key in ("n_uncertified", "n_widened", "n_widened_from_lists", "n_rescued", "kernel_path",
                    "n_splits"):
            assert s0[key] == s1[key], (key, s0, s1)
        assert s0["n_uncertified"] == 0
        assert_knn_parity(X, Y, d0, i0, k, EPS_F32)


def test_noalloc_rejects_what_it_cannot_hold(require_gpu):
    lib = _capi.lib()
    nt, nq, d, k = 1_000_000, 10_000, 128, 10
    assert lib.knn_argkmin_workspace_size(4, nt, nq, d, k, 1, 0) > _capi.NOALLOC_SCRATCH_BYTES
    # pointers are never dereferenced: the size check comes first
    code = lib.knn_argkmin_f32_noalloc(16, nt, 16, nq, d, k, 256, 0, 1, 0, 0, 8, 16, 16, None,
                                       None)
    assert code == _capi.KNN_EWORKSPACE
    assert b"scratch" in lib.knn_last_error()
    code = lib.knn_argkmin_f32_noalloc(16, 1000, 16, 10, 8, 5, 256, 0, 0, 0, _capi.FLAG_TIMING,
                                       8, 16, 16, None, None)
    assert code == _capi.KNN_EINVAL
