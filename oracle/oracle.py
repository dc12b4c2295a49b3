"""numpy front end of the C oracle (oracle/oracle.c -> oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker; the flagsparse_amd package never
imports it.  Each function states the reference semantics it restates
(src/flagsparse/sparse_operations/ of FlagSparse).

Pinning (tests/test_oracle.py): every function is checked against scipy.sparse
(an independent implementation), dense numpy brute force on small inputs and
closed-form invariants (linearity, transpose identities, permutation
bijectivity); the reference ships no golden vectors for this path.
"""

import ctypes
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "liboracle.so")

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_c_i64 = ctypes.c_int64
_c_dbl = ctypes.c_double

_lib = None


def build():
    """Compile liboracle.so with gcc (make)."""
    subprocess.run(["make", "-s", "-C", _DIR, "liboracle.so"], check=True)
    return LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        lib.orc_spmv_csr.argtypes = [_c_i64, _i64p, _i64p, _f64p, _f64p, _f64p]
        lib.orc_spmv_csr_trans.argtypes = [_c_i64, _c_i64, _i64p, _i64p, _f64p, _f64p, _f64p]
        lib.orc_spmm_csr.argtypes = [_c_i64, _c_i64, _i64p, _i64p, _f64p, _f64p, _f64p]
        lib.orc_spmm_csr_trans.argtypes = [_c_i64, _c_i64, _c_i64, _i64p, _i64p, _f64p, _f64p, _f64p]
        lib.orc_sddmm_csr.argtypes = [_c_i64, _c_i64, _i64p, _i64p, _f64p, _f64p, ctypes.c_void_p, _c_dbl, _c_dbl,
                                      _f64p]
        lib.orc_csr_to_csc.argtypes = [_c_i64, _c_i64, _i64p, _i64p, _i64p, _i64p, _i64p]
        for fn in (lib.orc_spmv_csr, lib.orc_spmv_csr_trans, lib.orc_spmm_csr, lib.orc_spmm_csr_trans,
                   lib.orc_sddmm_csr, lib.orc_csr_to_csc):
            fn.restype = None
        _lib = lib
    return _lib


def _i64(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.int64)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.float64)


def spmv(indptr, indices, data, x, shape, op="non"):
    """y = op(A) x (spmv_csr.py:179-205; op per spmv_csr.py:43-57). Real data:
    conj == trans."""
    n_rows, n_cols = int(shape[0]), int(shape[1])
    indptr, indices, data, x = _i64(indptr), _i64(indices), _f64(data), _f64(x)
    lib = _load()
    if op in ("non", 0):
        y = np.empty(n_rows)
        lib.orc_spmv_csr(n_rows, indptr, indices, data, x, y)
    else:
        y = np.empty(n_cols)
        lib.orc_spmv_csr_trans(n_rows, n_cols, indptr, indices, data, x, y)
    return y


def spmm(indptr, indices, data, B, shape, op="non"):
    """C = op(A) B (spmm_csr.py:144-200, 407-456)."""
    n_rows, n_cols = int(shape[0]), int(shape[1])
    indptr, indices, data, B = _i64(indptr), _i64(indices), _f64(data), _f64(B)
    n_dense = int(B.shape[1]) if B.ndim == 2 else 0
    lib = _load()
    if op in ("non", 0):
        C = np.empty((n_rows, n_dense))
        lib.orc_spmm_csr(n_rows, n_dense, indptr, indices, data, B, C)
    else:
        C = np.empty((n_cols, n_dense))
        lib.orc_spmm_csr_trans(n_rows, n_cols, n_dense, indptr, indices, data, B, C)
    return C


def sddmm(indptr, indices, X, Y, data=None, alpha=1.0, beta=0.0):
    """out[p] = alpha <X[row p], Y[col p]> + beta data[p] (sddmm_csr.py:404, 471-483)."""
    indptr, indices, X, Y = _i64(indptr), _i64(indices), _f64(X), _f64(Y)
    n_rows = indptr.size - 1
    k = int(X.shape[1])
    out = np.empty(indices.size)
    d = None if data is None else _f64(data)
    _load().orc_sddmm_csr(n_rows, k, indptr, indices, X, Y, None if d is None else d.ctypes.data, float(alpha),
                          float(beta), out)
    return out


def csr_to_csc(indptr, indices, shape):
    """(colptr, rowind, perm) of the stable bucketing by column
    (spmv_csr.py:498-536)."""
    n_rows, n_cols = int(shape[0]), int(shape[1])
    indptr, indices = _i64(indptr), _i64(indices)
    colptr = np.empty(n_cols + 1, dtype=np.int64)
    rowind = np.empty(indices.size, dtype=np.int64)
    perm = np.empty(indices.size, dtype=np.int64)
    _load().orc_csr_to_csc(n_rows, n_cols, indptr, indices, colptr, rowind, perm)
    return colptr, rowind, perm


def max_relative_error(candidate, reference):
    """max|cand - ref| / max|ref| (spmm_csr.py:86-125)."""
    cand = np.asarray(candidate, dtype=np.float64)
    ref = np.asarray(reference, dtype=np.float64)
    if cand.size == 0:
        return 0.0
    diff = float(np.max(np.abs(cand - ref)))
    mag = float(np.max(np.abs(ref)))
    if mag == 0.0:
        return 0.0 if diff == 0.0 else float("inf")
    return diff / mag
