"""BSR (16 x 16 blocks) SpMM on the MI355X matrix cores.

New in this package (the reference computes no BSR SpMM); the block storage
follows the reference's BSR SpMV (spmv_bsr.py:448-470): ``data`` of shape
(nnzb, block_dim, block_dim), row-major inside a block, ``indices`` = block
columns ascending inside a block row, ``indptr`` over block rows, and its
validation messages are reused.

* ``csr_to_bsr``             -- GPU conversion, bit-exact (values moved, not
                               recomputed; untouched block entries are 0).
* ``quantize_bsr_fp8``       -- per-block absmax scaling to OCP FP8 e4m3.
* ``flagsparse_spmm_bsr``    -- C (float32) = A_bsr @ B with bf16 / f16 blocks
                               (MFMA 16x16x32, f32 accumulation) or FP8 codes +
                               per-block scales (B quantised per tensor).
"""

import ctypes

import torch

from . import _lib, _plan

BLOCK_DIM = 16
_BSR_KERNEL_DTYPES = (torch.bfloat16, torch.float16, torch.uint8)  # uint8 = FP8 e4m3 codes


def _padded_cols(n):
    return -(-int(n) // 128) * 128


def _validate_bsr(data, indices, indptr, shape, block_dim):
    """Validation contract of the reference's _prepare_spmv_bsr_matrix
    (spmv_bsr.py:448-490): same checks, same messages, for block_dim 16.
    Returns (data, block columns as int32, indptr, n_rows, n_cols,
    n_block_rows, n_block_cols); int64 block columns are narrowed to int32
    (after the range check) because the MFMA kernel reads int32."""
    if not all(torch.is_tensor(t) for t in (data, indices, indptr)):
        raise TypeError("data, indices, indptr must all be torch.Tensor")
    if data.ndim != 3:
        raise ValueError("data must have shape (nnzb, block_dim, block_dim)")
    if indices.ndim != 1 or indptr.ndim != 1:
        raise ValueError("indices and indptr must be 1D tensors")
    n_rows, n_cols = int(shape[0]), int(shape[1])
    block_dim = int(block_dim)
    if block_dim <= 1:
        raise ValueError("block_dim must be greater than 1 for BSR SpMV")
    if block_dim != BLOCK_DIM:
        raise ValueError(f"block_dim must be {BLOCK_DIM} for the MFMA BSR SpMM")
    if data.shape[1] != block_dim or data.shape[2] != block_dim:
        raise ValueError("data block dimensions must match block_dim")
    n_block_rows = (n_rows + block_dim - 1) // block_dim
    n_block_cols = (n_cols + block_dim - 1) // block_dim
    if indptr.numel() != n_block_rows + 1:
        raise ValueError(f"indptr length must be n_block_rows+1={n_block_rows + 1}, got {indptr.numel()}")
    if data.shape[0] != indices.numel():
        raise ValueError("data.shape[0] and indices length must both equal nnzb")
    if not all(t.is_cuda for t in (data, indices, indptr)):
        raise ValueError("data, indices, indptr must be CUDA tensors")
    if not all(t.device == data.device for t in (indices, indptr)):
        raise ValueError("data, indices, indptr must be on the same CUDA device")
    if data.dtype not in _BSR_KERNEL_DTYPES:
        raise TypeError("BSR data dtype must be torch.bfloat16, torch.float16 or torch.uint8 (fp8 e4m3 codes)")
    if indices.dtype not in (torch.int32, torch.int64):
        raise TypeError("indices dtype must be torch.int32 or torch.int64")
    if indptr.dtype not in (torch.int32, torch.int64):
        raise TypeError("indptr dtype must be torch.int32 or torch.int64")
    data = data.contiguous()
    indices = indices.contiguous()
    indptr = indptr.contiguous()
    # every check from one device inspection and one host readback
    nnzb = int(indices.numel())
    stats, _ = _plan.inspect_csr(indptr, indices, n_block_rows, nnzb)
    if stats["indptr_first"] != 0:
        raise ValueError("indptr must start at zero")
    if stats["indptr_last"] != data.shape[0]:
        raise ValueError("indptr[-1] must equal nnzb")
    if stats["n_decreasing"]:
        raise ValueError("indptr must be non-decreasing")
    if nnzb > 0 and (stats["min_col"] < 0 or stats["max_col"] >= n_block_cols):
        raise IndexError("indices out of range for n_block_cols")
    if indices.dtype != torch.int32:
        indices = indices.to(torch.int32)
    return data, indices, indptr, n_rows, n_cols, n_block_rows, n_block_cols


def csr_to_bsr(data, indices, indptr, shape, block_dim=BLOCK_DIM):
    """CSR -> BSR on the GPU: (bsr_data [nnzb, 16, 16], bsr_indices int32,
    bsr_indptr of indptr's dtype).

    Bit-exact: values are moved, never recomputed; block entries with no CSR
    nonzero are 0; block columns ascend inside a block row (the order of
    scipy's ``tobsr`` followed by ``sort_indices``).  The input must be
    canonical CSR (no duplicate (row, column) pairs) and is validated like the
    CSR operators (spmm_csr.py:342-404)."""
    from .spmm_csr import _prepare_spmm_csr_matrix

    if int(block_dim) != BLOCK_DIM:
        raise ValueError(f"block_dim must be {BLOCK_DIM}")
    data, indices, indptr, n_rows, n_cols, _, _ = _prepare_spmm_csr_matrix(data, indices, indptr, shape)
    if data.dtype not in (torch.float32, torch.float64, torch.float16, torch.bfloat16):
        raise TypeError("csr_to_bsr supports float16, bfloat16, float32 and float64 values")
    device = data.device
    nnz = int(data.numel())
    lib = _lib.load()
    ws_bytes = ctypes.c_int64(0)
    _lib.check(lib.fsp_csr_to_bsr_workspace(n_rows, n_cols, nnz, ctypes.byref(ws_bytes)), "fsp_csr_to_bsr_workspace")
    ws = torch.empty(max(1, ws_bytes.value), dtype=torch.uint8, device=device)
    nnzb_dev = torch.empty(1, dtype=torch.int64, device=device)
    stream = _lib.stream_ptr(device)
    _lib.check(
        lib.fsp_csr_to_bsr_count(_lib.index_code(indptr.dtype), _lib.index_code(indices.dtype), _lib.ptr(indptr),
                                 _lib.ptr(indices), n_rows, n_cols, nnz, ws.data_ptr(), ws.numel(),
                                 nnzb_dev.data_ptr(), stream),
        "fsp_csr_to_bsr_count",
    )
    nnzb = int(nnzb_dev.item())
    nbr = -(-n_rows // BLOCK_DIM)
    bsr_indptr = torch.empty(nbr + 1, dtype=indptr.dtype, device=device)
    bsr_indices = torch.empty(nnzb, dtype=torch.int32, device=device)
    bsr_data = torch.empty((nnzb, BLOCK_DIM, BLOCK_DIM), dtype=data.dtype, device=device)
    _lib.check(
        lib.fsp_csr_to_bsr_fill(_lib.value_code(data.dtype), _lib.index_code(indptr.dtype),
                                _lib.index_code(indices.dtype), _lib.ptr(indptr), _lib.ptr(indices), _lib.ptr(data),
                                n_rows, n_cols, nnz, nnzb, _lib.index_code(bsr_indptr.dtype), bsr_indptr.data_ptr(),
                                _lib.ptr(bsr_indices), _lib.ptr(bsr_data), ws.data_ptr(), ws.numel(), stream),
        "fsp_csr_to_bsr_fill",
    )
    return bsr_data, bsr_indices, bsr_indptr


def quantize_bsr_fp8(bsr_data):
    """Per-block absmax FP8 (OCP e4m3): (codes uint8 [nnzb,16,16], scales f32 [nnzb])
    with block ~= e4m3(codes) * scale, scale = absmax / 448 in float32 (1 for
    an all-zero block) and codes = e4m3(float32(block) / scale), rounded to
    nearest even."""
    if not torch.is_tensor(bsr_data) or bsr_data.ndim != 3 or tuple(bsr_data.shape[1:]) != (BLOCK_DIM, BLOCK_DIM):
        raise ValueError("bsr_data must have shape (nnzb, 16, 16)")
    if not bsr_data.is_cuda:
        raise ValueError("bsr_data must be a CUDA tensor")
    if bsr_data.dtype not in (torch.float32, torch.float64, torch.float16, torch.bfloat16):
        raise TypeError("bsr_data dtype must be float16, bfloat16, float32 or float64")
    src = bsr_data.contiguous()
    nnzb = int(src.shape[0])
    device = src.device
    codes = torch.empty((nnzb, BLOCK_DIM, BLOCK_DIM), dtype=torch.uint8, device=device)
    scales = torch.empty(nnzb, dtype=torch.float32, device=device)
    _lib.check(
        _lib.load().fsp_bsr_quantize_fp8(_lib.value_code(src.dtype), _lib.ptr(src), nnzb, _lib.ptr(codes),
                                         _lib.ptr(scales), _lib.stream_ptr(device)),
        "fsp_bsr_quantize_fp8",
    )
    return codes, scales


def pack_dense_operand(B, out_dtype):
    """B [K, N] -> the packed MFMA operand (and, for fp8, B's absmax)."""
    K, N = int(B.shape[0]), int(B.shape[1])
    device = B.device
    B = B.contiguous()
    k8 = -(-K // BLOCK_DIM) * 2
    npad = _padded_cols(N)
    lib = _lib.load()
    stream = _lib.stream_ptr(device)
    absmax = None
    if out_dtype == torch.uint8:
        absmax = torch.empty(1, dtype=torch.float32, device=device)
        _lib.check(lib.fsp_absmax(_lib.value_code(B.dtype), _lib.ptr(B), K, N, N, absmax.data_ptr(), stream),
                   "fsp_absmax")
        Bp = torch.empty(k8 * npad * 8, dtype=torch.uint8, device=device)
        code = _lib.FSP_F8E4M3
    else:
        Bp = torch.empty(k8 * npad * 8, dtype=out_dtype, device=device)
        code = _lib.value_code(out_dtype)
    _lib.check(
        lib.fsp_bsr_pack_b(_lib.value_code(B.dtype), code, _lib.ptr(B), N, K, N,
                           None if absmax is None else absmax.data_ptr(), Bp.data_ptr(), stream),
        "fsp_bsr_pack_b",
    )
    return Bp, absmax


def flagsparse_spmm_bsr(data, indices, indptr, B, shape, block_dim=BLOCK_DIM, scales=None, out=None):
    """C (float32, [n_rows, N]) = A_bsr @ B on the matrix cores.

    data: bf16 / f16 blocks, or uint8 FP8 e4m3 codes together with per-block
    ``scales`` (see quantize_bsr_fp8).  B: dense [n_cols, N] (float32, bf16
    or f16); for bf16/f16 blocks B is converted to the block dtype, for FP8 it
    is quantised with one absmax scale per tensor.  Accumulation is float32.
    """
    data, indices, indptr, n_rows, n_cols, n_block_rows, _ = _validate_bsr(data, indices, indptr, shape, block_dim)
    fp8 = data.dtype == torch.uint8
    if fp8 and (scales is None or scales.numel() != data.shape[0] or scales.dtype != torch.float32
                or not scales.is_cuda or scales.device != data.device):
        raise ValueError("fp8 BSR data needs float32 per-block scales of length nnzb on the same device")
    if not fp8 and scales is not None:
        raise ValueError("per-block scales apply to fp8 BSR data only")
    if not torch.is_tensor(B) or B.ndim != 2:
        raise ValueError("B must be a 2D dense tensor")
    if B.shape[0] != n_cols:
        raise ValueError(f"B.shape[0] must be n_cols={n_cols}, got {B.shape[0]}")
    if not B.is_cuda or B.device != data.device:
        raise ValueError("B must be a CUDA tensor on the same device as the sparse matrix")
    if B.dtype not in (torch.float32, torch.bfloat16, torch.float16):
        raise TypeError("B dtype must be float32, bfloat16 or float16")
    N = int(B.shape[1])
    device = data.device
    if out is None:
        out = torch.empty((n_rows, N), dtype=torch.float32, device=device)
    elif tuple(out.shape) != (n_rows, N) or out.dtype != torch.float32 or not out.is_contiguous() \
            or out.device != device:
        raise ValueError("out must be a contiguous float32 tensor of shape (n_rows, N) on the matrix's device")
    if n_rows == 0 or N == 0:
        return out
    Bp, absmax = pack_dense_operand(B, torch.uint8 if fp8 else data.dtype)
    scales = None if scales is None else scales.contiguous()
    vcode = _lib.FSP_F8E4M3 if fp8 else _lib.value_code(data.dtype)
    _lib.check(
        _lib.load().fsp_bsr_spmm(vcode, _lib.index_code(indptr.dtype), _lib.ptr(indptr), _lib.ptr(indices),
                                 _lib.ptr(data), None if scales is None else _lib.ptr(scales),
                                 n_block_rows, int(data.shape[0]), Bp.data_ptr(),
                                 None if absmax is None else absmax.data_ptr(), n_cols, N, out.data_ptr(), N, n_rows,
                                 _lib.stream_ptr(device)),
        "fsp_bsr_spmm",
    )
    return out
