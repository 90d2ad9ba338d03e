"""The C ABI library without a GPU: it loads, exports every symbol the header
declares, sizes workspaces, and rejects bad arguments before touching the
device.  The Python surface fails loudly when no GPU is present."""
import ctypes
import pathlib
import re

import numpy as np
import pytest

from pdist_aggregation_amd import _capi
import pdist_aggregation_amd as pa

HEADER = pathlib.Path(__file__).resolve().parents[1] / "include" / "pdist_knn.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(knn_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(_capi.EXPORTED)


def test_stats_struct_matches_header():
    """knn_stats: the ctypes mirror has the header's fields in the header's
    order, with the same widths (int64_t / double), so a library and binding
    of one ABI version read the same bytes."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    body = re.search(r"typedef struct knn_stats \{(.*?)\} knn_stats;", text, re.S).group(1)
    fields = re.findall(r"\b(int64_t|double)\s+(\w+)\s*;", body)
    ctype = {"int64_t": ctypes.c_int64, "double": ctypes.c_double}
    assert [(n, ctype[t]) for t, n in fields] == list(_capi.KnnStats._fields_)
    assert ctypes.sizeof(_capi.KnnStats) == 8 * len(fields)


def test_library_exports_every_declared_symbol():
    lib = _capi.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.knn_abi_version() == _capi.ABI_VERSION


def test_workspace_size():
    lib = _capi.lib()
    ws = lib.knn_argkmin_workspace_size
    a = ws(4, 1_000_000, 10_000, 128, 10, 0, 0)
    b = ws(4, 1_000_000, 200_000, 128, 10, 0, 0)
    assert 0 < a < b
    assert ws(8, 10_000, 1000, 50, 10, 0, 0) > 0
    assert ws(4, 10, 3, 4, 11, 0, 0) == 0  # k > n_train
    assert ws(4, 10, 3, 4, 0, 0, 0) == 0  # k == 0
    assert ws(3, 10, 3, 4, 1, 0, 0) == 0  # bad dtype
    assert ws(8, 100, 3, 4, 1, 1, 0) == 0  # f64 + bf16
    # k above the merge limit (1,984): the large-k path's dense rows + selections
    big = ws(4, 10_000, 300, 32, 5000, 0, 0)
    assert big >= 300 * 10_000 * 8 + 300 * 5000 * 12
    assert ws(8, 10_000, 300, 32, 10_000, 0, 0) > 0  # k = n_train
    # bf16 / fp8 beyond their kernels (k' > 478, d > 256) size the exact search
    assert ws(4, 6000, 100, 64, 600, 1, 0) == ws(4, 6000, 100, 64, 600, 0, 0)
    assert ws(4, 6000, 100, 300, 10, 2, 0) == ws(4, 6000, 100, 300, 10, 0, 0)


def _call(**over):
    lib = _capi.lib()
    args = dict(Xtr=1, ntr=100, Xte=1, nte=10, dim=8, k=5, chunk=256, strategy=0, prec=0,
                ovs=0, flags=0, threads=1, dist=1, idx=1, ws=None, wsb=0, stream=None,
                stats=None)
    args.update(over)
    return lib.knn_argkmin_f32(*args.values())


@pytest.mark.parametrize("over,needle", [
    ({"k": 0}, "k must be positive"),
    ({"k": 101}, "k must not exceed n_train"),
    ({"strategy": 7}, "strategy"),
    ({"chunk": 0}, "chunk_size"),
    ({"dim": 0}, "dim"),
    ({"ovs": 3}, "oversample"),
    ({"Xtr": None}, "null"),
    ({"ws": 1, "wsb": 16}, "workspace"),
])
def test_bad_arguments_are_rejected_before_any_device_work(over, needle):
    code = _call(**over)
    assert code in (_capi.KNN_EINVAL, _capi.KNN_EWORKSPACE)
    assert needle in _capi.lib().knn_last_error().decode()


def test_f64_rejects_low_precision():
    lib = _capi.lib()
    code = lib.knn_argkmin_f64(1, 100, 1, 10, 8, 5, 256, 0, 1, 0, 0, 1, 1, 1, None, 0, None,
                               None)
    assert code == _capi.KNN_EINVAL


def test_python_surface_rejects_bad_strategy_like_the_reference():
    with pytest.raises(RuntimeError, match="strategy 'bogus' not supported"):
        pa.parallel_knn(np.zeros((4, 2)), np.zeros((2, 2)), 1, strategy="bogus")


def test_no_cpu_fallback():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="needs a visible AMD GPU"):
        pa.parallel_knn(np.random.rand(10, 3), np.random.rand(4, 3), 2)


@pytest.mark.parametrize("strategy", ["auto", "chunk_on_train", "chunk_on_test"])
def test_n_parallel_chunks_matches_port(strategy):
    import oracle

    rs = np.random.RandomState(0)
    Y = rs.rand(777, 3)
    X = rs.rand(1500, 3)
    for chunk in (1, 20, 64, 256, 5000):
        _, n = oracle.port_parallel_knn(Y, X, 2, chunk_size=chunk, strategy=strategy,
                                        n_threads=2)
        assert pa.n_parallel_chunks(strategy, 777, 1500, chunk, 2) == n


def test_effective_n_threads_env(monkeypatch):
    monkeypatch.setenv("OMP_NUM_THREADS", "3")
    assert pa.effective_n_threads() == 3
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert 1 <= pa.effective_n_threads() <= (__import__("os").cpu_count() or 1)


def test_rate_probe_library_loads_and_rejects_bad_arguments():
    """libpdist_probe.so (the matrix-core rate bench.py reports beside the
    datasheet peak) loads without a GPU; argument checks fail before any HIP
    call."""
    path = pathlib.Path(_capi.LIB_PATH).with_name("libpdist_probe.so")
    lib = ctypes.CDLL(str(path))
    fn = lib.pdist_probe_mfma_tflops
    fn.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_double)] * 2
    tf, ms = ctypes.c_double(), ctypes.c_double()
    assert fn(2, 1, 10, 1, ctypes.byref(tf), ctypes.byref(ms)) == -1  # unknown format
    assert fn(0, 0, 10, 1, ctypes.byref(tf), ctypes.byref(ms)) == -1  # per_cu out of range
    assert fn(0, 5, 10, 1, ctypes.byref(tf), ctypes.byref(ms)) == -1
    assert fn(0, 1, 0, 1, ctypes.byref(tf), ctypes.byref(ms)) == -1  # no iterations


def test_workload_inputs_cpu():
    """The timed workloads build the inputs the GPU tests check (CPU only)."""
    from pdist_aggregation_amd import workloads

    Y, X, k, precision, over = workloads._inputs("plain-f64")
    assert Y.shape == (10_000, 50) and X.shape == (1_000, 50) and Y.dtype == np.float64
    assert (k, precision, over) == (10, "exact", 0)
    Y, X, k, precision, over = workloads._inputs("lowd")
    assert Y.shape == X.shape == (65536, 3) and Y.dtype == np.float32 and k == 16
    r = np.linalg.norm(X.astype(np.float64), axis=1)
    assert abs(r.mean() - 1.0) < 1e-3 and 0.005 < r.std() < 0.02
    with pytest.raises(ValueError):
        workloads._inputs("nope")
