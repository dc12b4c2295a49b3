"""ctypes binding of libflagsparse_bench.so (declared in
csrc/bench/fsp_bench.h): the measurement kernels bench.py times beside the
operators, kept out of the operator library."""

import ctypes
import os
import threading

import torch  # noqa: F401  (bind to the HIP runtime torch loaded)

from ._build import BENCH_LIB_PATH

_I32, _VP = ctypes.c_int32, ctypes.c_void_p

SIGNATURES = {
    "fsp_bench_last_error": (ctypes.c_char_p, []),
    "fsp_slab_spmm_probe": (_I32, [_I32] + [_VP] * 8 + [_I32] * 4 + [_VP, _I32, _VP, _VP]),
    "fsp_stream_read_probe": (_I32, [_VP, _VP, ctypes.c_int64, _I32, _I32, _I32, _VP, _VP]),
    "fsp_gather_rows_probe": (ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                               ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
}

_lock = threading.Lock()
_lib = None


def load():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(BENCH_LIB_PATH):
                raise RuntimeError(f"{BENCH_LIB_PATH} is missing: build it with `python -m flagsparse_amd._build`")
            lib = ctypes.CDLL(BENCH_LIB_PATH)
            for name, (restype, argtypes) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = restype
                fn.argtypes = argtypes
            _lib = lib
    return _lib


def check(status, what):
    if status != 0:
        msg = load().fsp_bench_last_error()
        raise RuntimeError(f"{what} failed with status {status}: {msg.decode('utf-8', 'replace') if msg else ''}")
