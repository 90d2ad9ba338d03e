"""The matrix-core rate probe bench.py reports beside the datasheet peak
(needs a GPU): it runs, and its rates are positive and below the MI355X
datasheet figures (bf16 2.5 PF/s, fp32 157.3 TF/s dense)."""
import ctypes
import pathlib
import sys

import pytest

from pdist_aggregation_amd import _capi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt,peak", [(0, 2500.0), (1, 157.3)])
def test_probe_rate_is_plausible(require_gpu, fmt, peak):
    torch = require_gpu
    torch.cuda.init()
    lib = ctypes.CDLL(str(pathlib.Path(_capi.LIB_PATH).with_name("libpdist_probe.so")))
    fn = lib.pdist_probe_mfma_tflops
    fn.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_double)] * 2
    for per_cu in (1, 4):
        tf, ms = ctypes.c_double(), ctypes.c_double()
        print(f"[probe] fmt={fmt} per_cu={per_cu}", flush=True)
        sys.stdout.flush()
        assert fn(fmt, per_cu, 50, 2, ctypes.byref(tf), ctypes.byref(ms)) == 0
        print(f"    {tf.value:.1f} TFLOP/s, {ms.value:.3f} ms per launch", flush=True)
        assert 0.0 < tf.value < peak * 1.02
        assert ms.value > 0.0
