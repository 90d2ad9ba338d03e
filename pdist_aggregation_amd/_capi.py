"""ctypes binding of libpdist_knn.so (declared in include/pdist_knn.h).

The library is the only compute path of this package: if it is missing, or no
GPU is visible when a search is requested, calls raise instead of falling back
to anything else.
"""
from __future__ import annotations

import ctypes
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libpdist_knn.so"

KNN_OK = 0
KNN_EINVAL = -1
KNN_EHIP = -2
KNN_EWORKSPACE = -3
KNN_EUNSUPPORTED = -4

STRATEGY = {"auto": 0, "chunk_on_train": 1, "chunk_on_test": 2}
STRATEGY_NAME = {v: k for k, v in STRATEGY.items()}
PRECISION = {"exact": 0, "fp32": 0, "fp64": 0, "bf16": 1, "fp8": 2}
FLAG_SQUARED = 1
FLAG_TIMING = 2
NOALLOC_SCRATCH_BYTES = 128 << 20  # KNN_NOALLOC_SCRATCH_BYTES of include/pdist_knn.h
KERNEL_PATH = {0: "mfma-16x16x4", 1: "mfma-fp32", 2: "mfma-bf16", 3: "mfma-fp8", 4: "lowd-simt",
               5: "large-k"}
ABI_VERSION = 4  # knn_abi_version() of the library this binding matches

# Every symbol include/pdist_knn.h declares.
EXPORTED = (
    "knn_argkmin_f32",
    "knn_argkmin_f64",
    "knn_argkmin_f32_noalloc",
    "knn_argkmin_backward_f32",
    "knn_argkmin_backward_f64",
    "knn_sqdist_matrix_f32",
    "knn_sqdist_matrix_f64",
    "knn_argkmin_workspace_size",
    "knn_last_error",
    "knn_abi_version",
)


class KnnStats(ctypes.Structure):
    _fields_ = [
        ("n_parallel_chunks", ctypes.c_int64),
        ("n_uncertified", ctypes.c_int64),
        ("k_candidates", ctypes.c_int64),
        ("n_splits", ctypes.c_int64),
        ("strategy_used", ctypes.c_int64),
        ("kernel_path", ctypes.c_int64),
        ("prep_ms", ctypes.c_double),
        ("search_ms", ctypes.c_double),
        ("merge_ms", ctypes.c_double),
        ("n_rescued", ctypes.c_int64),
        ("n_widened", ctypes.c_int64),
        ("n_widened_from_lists", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: (float if t is ctypes.c_double else int)(getattr(self, name))
                for name, t in self._fields_}


class KnnError(RuntimeError):
    """A non-zero return code of the C ABI."""

    def __init__(self, code, message):
        super().__init__(f"pdist_knn error {code}: {message}")
        self.code = code


_LIB = None


def _argkmin_argtypes():
    v, i64, i = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    return [v, i64, v, i64, i64, i64, i64, i, i, i64, i, i, v, v, v, ctypes.c_size_t, v,
            ctypes.POINTER(KnnStats)]


def lib():
    """Load the shared library once; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m pdist_aggregation_amd.build`")
        h = ctypes.CDLL(str(LIB_PATH))
        for name in ("knn_argkmin_f32", "knn_argkmin_f64"):
            fn = getattr(h, name)
            fn.restype = ctypes.c_int
            fn.argtypes = _argkmin_argtypes()
        v, i64, i = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        h.knn_argkmin_f32_noalloc.restype = ctypes.c_int
        h.knn_argkmin_f32_noalloc.argtypes = [v, i64, v, i64, i64, i64, i64, i, i, i64, i, i,
                                              v, v, v, ctypes.POINTER(KnnStats)]
        for name in ("knn_argkmin_backward_f32", "knn_argkmin_backward_f64"):
            fn = getattr(h, name)
            fn.restype = ctypes.c_int
            v, i64 = ctypes.c_void_p, ctypes.c_int64
            fn.argtypes = [v, i64, v, i64, i64, i64, v, v, ctypes.c_int, v, v, v]
        for name in ("knn_sqdist_matrix_f32", "knn_sqdist_matrix_f64"):
            fn = getattr(h, name)
            fn.restype = ctypes.c_int
            v, i64 = ctypes.c_void_p, ctypes.c_int64
            fn.argtypes = [v, i64, v, i64, i64, v, v]
        h.knn_argkmin_workspace_size.restype = ctypes.c_size_t
        h.knn_argkmin_workspace_size.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                                 ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                                 ctypes.c_int64]
        h.knn_last_error.restype = ctypes.c_char_p
        h.knn_last_error.argtypes = []
        h.knn_abi_version.restype = ctypes.c_int
        h.knn_abi_version.argtypes = []
        if h.knn_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} has ABI version {h.knn_abi_version()}, this binding "
                               f"expects {ABI_VERSION}: rebuild with python -m "
                               f"pdist_aggregation_amd.build")
        _LIB = h
    return _LIB


def check(code):
    if code != KNN_OK:
        msg = lib().knn_last_error().decode(errors="replace")
        if code == KNN_EINVAL:
            raise ValueError(msg)
        raise KnnError(code, msg)
