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
        lib.orc_csr_to_bsr.argtypes = [_c_i64, _c_i64, _c_i64, _i64p, _i64p, _f64p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]
        lib.orc_csr_to_bsr.restype = _c_i64
        lib.orc_bsr_spmm.argtypes = [_c_i64, _c_i64, _c_i64, _i64p, _i64p, _f64p, _f64p, _f64p]
        for fn in (lib.orc_spmv_csr, lib.orc_spmv_csr_trans, lib.orc_spmm_csr, lib.orc_spmm_csr_trans,
                   lib.orc_sddmm_csr, lib.orc_csr_to_csc, lib.orc_bsr_spmm):
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


def csr_to_bsr(indptr, indices, data, shape, bs=16):
    """(block_ptr, block_col, block_data[nnzb, bs, bs]) with block columns
    ascending inside each block row; the block storage convention of the
    reference's BSR SpMV (spmv_bsr.py:451-452)."""
    n_rows, n_cols = int(shape[0]), int(shape[1])
    indptr, indices, data = _i64(indptr), _i64(indices), _f64(data)
    lib = _load()
    nnzb = lib.orc_csr_to_bsr(n_rows, n_cols, bs, indptr, indices, data, None, None, None)
    nbr = -(-n_rows // bs)
    block_ptr = np.zeros(nbr + 1, dtype=np.int64)
    block_col = np.zeros(max(nnzb, 1), dtype=np.int64)
    block_data = np.zeros((max(nnzb, 1), bs, bs), dtype=np.float64)
    lib.orc_csr_to_bsr(n_rows, n_cols, bs, indptr, indices, data, block_ptr.ctypes.data, block_col.ctypes.data,
                       block_data.ctypes.data)
    return block_ptr, block_col[:nnzb], block_data[:nnzb]


def bsr_spmm(block_ptr, block_col, block_data, B, bs=16):
    """C = BSR(A) @ B with long-double accumulation."""
    block_ptr, block_col = _i64(block_ptr), _i64(block_col)
    data = _f64(np.asarray(block_data).reshape(-1))
    B = _f64(B)
    nbr = block_ptr.size - 1
    C = np.empty((nbr * bs, B.shape[1]))
    _load().orc_bsr_spmm(nbr, bs, int(B.shape[1]), block_ptr, block_col, data, B, C)
    return C


# ---- OCP FP8 E4M3 (finite-only "fn" variant): bias 7, 3 mantissa bits,
# largest finite 448, no infinities, 0x7F/0xFF = NaN.  Restated from the
# Open Compute Project OFP8 specification (round to nearest, ties to even).
E4M3_MAX = 448.0


def e4m3_encode(x):
    """float array -> uint8 E4M3FN codes (RNE; |x| > 448 saturates to 448)."""
    x = np.asarray(x, dtype=np.float64)
    sign = (np.signbit(x)).astype(np.uint8) << 7
    a = np.minimum(np.abs(x), E4M3_MAX)
    code = np.zeros(a.shape, dtype=np.int64)
    normal = a >= 2.0 ** -6
    # subnormals: value = m * 2^-9, m in 0..8 (m == 8 carries into the first normal)
    m_sub = np.rint(a / 2.0 ** -9).astype(np.int64)
    code = np.where(normal, 0, m_sub)
    e = np.floor(np.log2(np.where(normal, a, 1.0))).astype(np.int64)
    mant = np.rint(np.where(normal, a, 1.0) / np.exp2(e - 3)).astype(np.int64)  # 8..16
    carry = mant == 16
    e = np.where(carry, e + 1, e)
    mant = np.where(carry, 8, mant)
    code = np.where(normal, ((e + 7) << 3) | (mant - 8), code)
    return (code.astype(np.uint8) | sign).astype(np.uint8)


def e4m3_decode(code):
    code = np.asarray(code, dtype=np.uint8)
    sign = np.where(code & 0x80, -1.0, 1.0)
    e = (code >> 3) & 0xF
    m = (code & 0x7).astype(np.float64)
    val = np.where(e == 0, m * 2.0 ** -9, (1.0 + m / 8.0) * np.exp2(e.astype(np.float64) - 7))
    val = np.where((code & 0x7F) == 0x7F, np.nan, val)
    return sign * val


def fp8_block_quantize(block_data):
    """Per-block absmax scaling to E4M3: scale = absmax / 448 (1 for an
    all-zero block), codes = e4m3(block / scale)."""
    b = np.asarray(block_data, dtype=np.float64)
    absmax = np.abs(b.reshape(b.shape[0], -1)).max(axis=1) if b.size else np.zeros(0)
    scale = np.where(absmax > 0, absmax / E4M3_MAX, 1.0)
    codes = e4m3_encode(b / scale[:, None, None])
    return codes, scale


def fp8_block_quantize_f32(block_data):
    """fp8_block_quantize in float32 arithmetic, as the device computes it:
    scale = f32(absmax) / 448 (IEEE f32 division), codes = e4m3(f32(x) / scale)."""
    b = np.asarray(block_data, dtype=np.float32)
    n = b.shape[0]
    absmax = np.abs(b.reshape(n, -1)).max(axis=1) if b.size else np.zeros(0, np.float32)
    scale = np.where(absmax > 0, absmax / np.float32(E4M3_MAX), np.float32(1.0)).astype(np.float32)
    x = (b / scale[:, None, None]).astype(np.float32)
    return e4m3_encode(x.astype(np.float64)), scale


def fp8_tensor_quantize_f32(x):
    """One absmax scale for a whole tensor, float32 arithmetic:
    codes = e4m3(f32(x) * f32(448 / absmax)), scale = f32(absmax / 448)."""
    a = np.asarray(x, dtype=np.float32)
    am = np.float32(np.abs(a).max()) if a.size else np.float32(0.0)
    inv = np.float32(E4M3_MAX) / am if am > 0 else np.float32(1.0)
    codes = e4m3_encode((a * np.float32(inv)).astype(np.float32).astype(np.float64))
    scale = am / np.float32(E4M3_MAX) if am > 0 else np.float32(1.0)
    return codes, np.float32(scale)


def coo_to_csr(row, col, data, shape):
    """(indptr, indices, data) of a COO matrix: nonzeros stably ordered by
    (row, col), duplicates kept in input order (coo_to_csr_for_spmv,
    spmv_csr.py:958-981, with a stable sort)."""
    row, col = _i64(row), _i64(col)
    order = np.lexsort((col, row))  # stable
    indptr = np.zeros(int(shape[0]) + 1, dtype=np.int64)
    np.cumsum(np.bincount(row, minlength=int(shape[0])), out=indptr[1:])
    return indptr, col[order], np.asarray(data)[order]


def mtx_to_csr_loop(lines):
    """Matrix Market lines -> (values float64, indices, indptr, shape) by the
    reference loader's own procedure (tests/test_spmv.py:116-189): one dict
    per row, every entry (and its symmetric / skew mirror) added in file
    order starting from 0.0, columns sorted per row.  Pure Python: small
    files only."""
    field, symmetry = "real", "general"
    header = None
    data_lines = []
    for line in lines:
        s = line.strip()
        if s.startswith("%%MatrixMarket"):
            t = s.split()
            if len(t) >= 5:
                field, symmetry = t[3].lower(), t[4].lower()
            continue
        if s.startswith("%"):
            continue
        if header is None and s:
            p = s.split()
            header = (int(p[0]), int(p[1]), int(p[2]) if len(p) > 2 else 0)
            continue
        if s:
            data_lines.append(s)
    if header is None:
        raise ValueError("Cannot parse .mtx header")
    n_rows, n_cols, nnz = header
    if nnz == 0:
        return np.zeros(0), np.zeros(0, np.int64), np.zeros(n_rows + 1, np.int64), (n_rows, n_cols)
    sym = symmetry in ("symmetric", "hermitian")
    skew = symmetry == "skew-symmetric"
    maps = [dict() for _ in range(n_rows)]
    for line in data_lines[:nnz]:
        p = line.split()
        r, c = int(p[0]) - 1, int(p[1]) - 1
        v = 1.0 if field == "pattern" else float(p[2])
        if 0 <= r < n_rows and 0 <= c < n_cols:
            maps[r][c] = maps[r].get(c, 0.0) + v
            if r != c and (sym or skew) and 0 <= c < n_rows and 0 <= r < n_cols:
                maps[c][r] = maps[c].get(r, 0.0) + (-v if skew else v)
    vals, cols, ptr = [], [], [0]
    for m in maps:
        for c in sorted(m):
            cols.append(c)
            vals.append(m[c])
        ptr.append(len(cols))
    return np.array(vals, dtype=np.float64), np.array(cols, dtype=np.int64), np.array(ptr, dtype=np.int64), \
        (n_rows, n_cols)


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
