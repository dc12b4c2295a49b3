"""ctypes binding of libflagsparse_hip.so (declared in include/flagsparse_hip.h).

The library is loaded lazily on the first GPU call, always after ``import
torch`` so that it binds to the HIP runtime torch has loaded.  There is no
fallback: if the library is missing or fails to load, every operator raises.
"""

import ctypes
import os
import threading

import torch

from ._build import LIB_PATH

FSP_F32, FSP_F64, FSP_F16, FSP_BF16, FSP_F8E4M3 = 0, 1, 2, 3, 4
FSP_I32, FSP_I64 = 0, 1
FSP_MAX_BINS = 8
ABI_VERSION = 15
NOALLOC_SCRATCH_BYTES = 16 << 20

_VALUE_CODES = {
    torch.float32: FSP_F32,
    torch.float64: FSP_F64,
    torch.float16: FSP_F16,
    torch.bfloat16: FSP_BF16,
}
_INDEX_CODES = {torch.int32: FSP_I32, torch.int64: FSP_I64}

_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64
_c_vp = ctypes.c_void_p
_c_dbl = ctypes.c_double


class FspBin(ctypes.Structure):
    _fields_ = [
        ("lanes", _c_i32),
        ("reserved", _c_i32),
        ("count", _c_i64),
        ("rows", _c_vp),
    ]


class FspCsrPlan(ctypes.Structure):
    _fields_ = [
        ("n_bins", _c_i32),
        ("unroll", _c_i32),
        ("bins", FspBin * FSP_MAX_BINS),
        ("n_segments", _c_i64),
        ("seg_row_dev", _c_vp),
        ("seg_begin_dev", _c_vp),
        ("seg_end_dev", _c_vp),
        ("n_long_rows", _c_i64),
        ("long_rows_dev", _c_vp),
        ("long_seg_ptr_dev", _c_vp),
        ("partials_dev", _c_vp),
        ("partials_bytes", _c_i64),
        ("max_blocks", _c_i64),
        ("xcd_groups", _c_i64),
        ("variant", _c_i64),
        ("row_len", _c_i64),
        ("indptr_base", _c_i64),
    ]


# name -> (restype, argtypes); mirrors include/flagsparse_hip.h.
SIGNATURES = {
    "fsp_abi_version": (_c_i32, []),
    "fsp_last_error": (ctypes.c_char_p, []),
    "fsp_device_arch": (_c_i32, [ctypes.c_char_p, _c_i64]),
    "fsp_csr_row_bins_workspace": (_c_i32, [_c_i64, _c_i32, ctypes.POINTER(_c_i64)]),
    "fsp_csr_inspect": (_c_i32, [_c_i32, _c_vp, _c_i32, _c_vp, _c_i64, _c_i64, _c_vp, _c_vp]),
    "fsp_csr_row_bins": (
        _c_i32,
        [_c_i32, _c_vp, _c_i64, _c_i32, ctypes.POINTER(_c_i64), _c_vp, _c_vp, _c_vp, _c_i64, _c_vp],
    ),
    "fsp_csr_to_csc_workspace": (_c_i32, [_c_i64, _c_i64, _c_i64, _c_i32, ctypes.POINTER(_c_i64)]),
    "fsp_csr_to_csc": (
        _c_i32,
        [_c_i32, _c_i32, _c_vp, _c_vp, _c_i64, _c_i64, _c_i64, _c_i32, _c_vp, _c_i32, _c_vp,
         _c_i32, _c_vp, _c_vp, _c_i64, _c_vp],
    ),
    "fsp_gather": (_c_i32, [_c_i32, _c_i32, _c_vp, _c_vp, _c_vp, _c_i64, _c_vp]),
    "fsp_spmv_csr": (
        _c_i32,
        [_c_i32, _c_i32, _c_i32, _c_vp, _c_vp, _c_vp, _c_i32, _c_vp, _c_vp, _c_vp, _c_i64, _c_i64, _c_i64,
         ctypes.POINTER(FspCsrPlan), _c_vp],
    ),
    "fsp_spmv_csr_parts": (
        _c_i32,
        [_c_i32, _c_i32, _c_i32, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_i64, _c_i64, _c_vp, _c_i32, _c_i64, _c_i64,
         _c_vp],
    ),
    "fsp_spmm_csr": (
        _c_i32,
        [_c_i32, _c_i32, _c_i32, _c_vp, _c_vp, _c_vp, _c_i32, _c_vp, _c_vp, _c_i64, _c_vp, _c_i64,
         _c_i64, _c_i64, _c_i64, _c_i64, _c_i32, ctypes.POINTER(FspCsrPlan), _c_vp],
    ),
    "fsp_spmm_csr_trans_sddmm": (
        _c_i32,
        [_c_i32, _c_i32, _c_i32, _c_vp, _c_vp, _c_vp, _c_i32, _c_vp, _c_vp, _c_i64, _c_vp, _c_i64, _c_vp, _c_i64,
         _c_vp, _c_i64, _c_i64, _c_i64, _c_i64, _c_i32, ctypes.POINTER(FspCsrPlan), _c_vp],
    ),
    "fsp_spmm_csr_noalloc": (
        _c_i32,
        [_c_i32, _c_i32, _c_i32, _c_vp, _c_vp, _c_vp, _c_i32, _c_vp, _c_vp, _c_i64, _c_vp, _c_i64,
         _c_i64, _c_i64, _c_i64, _c_i64, _c_i32, ctypes.POINTER(FspCsrPlan), _c_vp],
    ),
    "fsp_sddmm_csr": (
        _c_i32,
        [_c_i32, _c_i32, _c_i32, _c_vp, _c_vp, _c_vp, _c_i64, _c_vp, _c_i64, _c_vp, _c_vp, _c_i64,
         _c_i64, _c_i64, _c_i64, _c_dbl, _c_dbl, _c_i32, ctypes.POINTER(FspCsrPlan), _c_vp],
    ),
    "fsp_csr_to_bsr_workspace": (_c_i32, [_c_i64, _c_i64, _c_i64, ctypes.POINTER(_c_i64)]),
    "fsp_csr_to_bsr_count": (_c_i32, [_c_i32, _c_i32, _c_vp, _c_vp, _c_i64, _c_i64, _c_i64, _c_vp, _c_i64, _c_vp,
                                      _c_vp]),
    "fsp_csr_to_bsr_fill": (_c_i32, [_c_i32, _c_i32, _c_i32, _c_vp, _c_vp, _c_vp, _c_i64, _c_i64, _c_i64, _c_i64,
                                     _c_i32, _c_vp, _c_vp, _c_vp, _c_vp, _c_i64, _c_vp]),
    "fsp_absmax": (_c_i32, [_c_i32, _c_vp, _c_i64, _c_i64, _c_i64, _c_vp, _c_vp]),
    "fsp_bsr_pack_b": (_c_i32, [_c_i32, _c_i32, _c_vp, _c_i64, _c_i64, _c_i64, _c_vp, _c_vp, _c_vp]),
    "fsp_bsr_quantize_fp8": (_c_i32, [_c_i32, _c_vp, _c_i64, _c_vp, _c_vp, _c_vp]),
    "fsp_bsr_spmm": (_c_i32, [_c_i32, _c_i32, _c_vp, _c_vp, _c_vp, _c_vp, _c_i64, _c_i64, _c_vp, _c_vp, _c_i64,
                              _c_i64, _c_vp, _c_i64, _c_i64, _c_vp]),
    "fsp_bsr_spmv": (_c_i32, [_c_i32, _c_i32, _c_i32, _c_vp, _c_vp, _c_i32, _c_vp, _c_vp, _c_i32, _c_vp, _c_vp, _c_i64,
                              _c_vp]),
    "fsp_coo_to_csr_workspace": (_c_i32, [_c_i64, _c_i64, _c_i64, _c_i32, _c_vp]),
    "fsp_coo_to_csr": (_c_i32, [_c_i32, _c_vp, _c_vp, _c_i64, _c_i64, _c_i64, _c_i32, _c_vp, _c_i32, _c_vp, _c_i32,
                                _c_vp, _c_vp, _c_vp, _c_i64, _c_vp]),
    "fsp_batch_pack": (_c_i32, [_c_i32, _c_i32, _c_i64, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_i64, _c_i64, _c_i32,
                                _c_vp, _c_i32, _c_vp, _c_vp, _c_vp]),
}

_lock = threading.Lock()
_lib = None


class FlagSparseLibraryError(RuntimeError):
    """The native HIP library is missing, stale or failed a call."""


def library_path():
    return os.environ.get("FLAGSPARSE_AMD_LIB", LIB_PATH)


def load(path=None):
    """Load (once) and return the ctypes handle of libflagsparse_hip.so."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or library_path()
        if not os.path.exists(p):
            raise FlagSparseLibraryError(
                f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `python -m flagsparse_amd._build` (the operators have no non-HIP path)"
            )
        lib = ctypes.CDLL(p)
        for name, (restype, argtypes) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        version = lib.fsp_abi_version()
        if version != ABI_VERSION:
            raise FlagSparseLibraryError(
                f"{p} has ABI version {version}, the Python side expects {ABI_VERSION}; rebuild it"
            )
        if path is None:
            _lib = lib
        return lib


def check(status, what):
    if status != 0:
        msg = load().fsp_last_error()
        msg = msg.decode("utf-8", "replace") if msg else ""
        raise FlagSparseLibraryError(f"{what} failed with status {status}: {msg}")


def value_code(dtype):
    try:
        return _VALUE_CODES[dtype]
    except KeyError:
        raise TypeError(f"unsupported value dtype for the HIP kernels: {dtype}") from None


def index_code(dtype):
    try:
        return _INDEX_CODES[dtype]
    except KeyError:
        raise TypeError(f"unsupported index dtype for the HIP kernels: {dtype}") from None


def ptr(t):
    """Device address of a tensor's first element (None for an empty tensor)."""
    if t is None:
        return None
    return t.data_ptr() if t.numel() > 0 else None


def stream_ptr(device=None):
    return torch.cuda.current_stream(device).cuda_stream


def device_arch():
    buf = ctypes.create_string_buffer(128)
    check(load().fsp_device_arch(buf, 128), "fsp_device_arch")
    return buf.value.decode()
