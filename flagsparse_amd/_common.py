"""Shared dtype tables, op codes, tolerances and error metrics.

Restates ops/_common.py:36-46 and :113-122 (dtype tables, tolerances) and
ops/spmm_csr.py:36-141 (op normalisation, relative thresholds, error metrics)
of the reference (paths relative to src/flagsparse/sparse_operations/).
"""

import time

import torch

SUPPORTED_VALUE_DTYPES = (
    torch.float16,
    torch.bfloat16,
    torch.float32,
    torch.float64,
    torch.complex64,
    torch.complex128,
)
SUPPORTED_INDEX_DTYPES = (torch.int32, torch.int64)
_INDEX_LIMIT_INT32 = 2**31 - 1

# Value dtypes the HIP kernels compute; complex dtypes are accepted by the
# reference's validators but have no HIP path here.
KERNEL_VALUE_DTYPES = (torch.float16, torch.bfloat16, torch.float32, torch.float64)

OP_NON = 0
OP_TRANS = 1
OP_CONJ_TRANS = 2
OP_NAMES = {OP_NON: "non", OP_TRANS: "trans", OP_CONJ_TRANS: "conj"}
_OP_NAME_TO_CODE = {name: code for code, name in OP_NAMES.items()}
_OP_ERROR = "op must be one of: 0=non, 1=trans, 2=conj"


def _is_complex_dtype(value_dtype):
    return value_dtype in (torch.complex64, torch.complex128)


def normalize_op(op=None, transpose=False):
    """ops/spmv_csr.py:43-57 and ops/spmm_csr.py:36-50."""
    if op is None:
        return OP_TRANS if bool(transpose) else OP_NON
    if isinstance(op, str):
        token = op.strip().lower()
        if token not in _OP_NAME_TO_CODE:
            raise ValueError(_OP_ERROR)
        return _OP_NAME_TO_CODE[token]
    try:
        op_code = int(op)
    except (TypeError, ValueError) as exc:
        raise ValueError(_OP_ERROR) from exc
    if op_code not in OP_NAMES:
        raise ValueError(_OP_ERROR)
    return op_code


def op_to_name(op):
    return OP_NAMES[normalize_op(op)]


def op_transposes(op):
    return normalize_op(op) in (OP_TRANS, OP_CONJ_TRANS)


def _tolerance_for_dtype(value_dtype):
    """(atol, rtol) of ops/_common.py:113-122."""
    if value_dtype == torch.float16:
        return 2e-3, 2e-3
    if value_dtype == torch.bfloat16:
        return 1e-1, 1e-1
    if value_dtype in (torch.float32, torch.complex64):
        return 1e-6, 1e-5
    if value_dtype in (torch.float64, torch.complex128):
        return 1e-10, 1e-8
    return 1e-6, 1e-5


def relative_threshold(value_dtype):
    """ops/spmm_csr.py:62-71 (threshold on max|cand-ref| / max|ref|)."""
    if value_dtype == torch.float16:
        return 5e-3
    if value_dtype == torch.bfloat16:
        return 1e-2
    if value_dtype in (torch.float32, torch.complex64):
        return 1e-6
    if value_dtype in (torch.float64, torch.complex128):
        return 1e-12
    return 1e-6


def error_metrics(candidate, reference):
    """ops/spmm_csr.py:86-125: max/sum absolute and relative errors."""
    if candidate.shape != reference.shape:
        raise ValueError(
            f"candidate and reference must have the same shape, got {candidate.shape} vs {reference.shape}"
        )
    if candidate.numel() == 0:
        return {
            "max_abs_error": 0.0,
            "max_relative_error": 0.0,
            "sum_relative_error": 0.0,
            "reference_max_magnitude": 0.0,
            "reference_sum_magnitude": 0.0,
        }
    ref = reference.to(torch.float64) if not _is_complex_dtype(reference.dtype) else reference
    cand = candidate.to(ref.dtype).to(ref.device)
    reference_compare = torch.abs(ref)
    abs_diff = torch.abs(cand - ref)
    max_abs_error = float(torch.max(abs_diff).item())
    reference_max_magnitude = float(torch.max(reference_compare).item())
    sum_abs_error = float(torch.sum(abs_diff).item())
    reference_sum_magnitude = float(torch.sum(reference_compare).item())
    if reference_max_magnitude == 0.0:
        max_relative_error = 0.0 if max_abs_error == 0.0 else float("inf")
    else:
        max_relative_error = max_abs_error / reference_max_magnitude
    if reference_sum_magnitude == 0.0:
        sum_relative_error = 0.0 if sum_abs_error == 0.0 else float("inf")
    else:
        sum_relative_error = sum_abs_error / reference_sum_magnitude
    return {
        "max_abs_error": max_abs_error,
        "max_relative_error": max_relative_error,
        "sum_relative_error": sum_relative_error,
        "reference_max_magnitude": reference_max_magnitude,
        "reference_sum_magnitude": reference_sum_magnitude,
    }


def validation_metrics(candidate, reference, value_dtype=None):
    """ops/spmm_csr.py:128-141 (threshold judged against the tested dtype)."""
    metrics = error_metrics(candidate, reference)
    dtype = value_dtype or candidate.dtype
    threshold = relative_threshold(dtype)
    atol, rtol = _tolerance_for_dtype(dtype)
    cand = candidate.to(torch.float64) if not _is_complex_dtype(candidate.dtype) else candidate
    ref = reference.to(cand.dtype).to(cand.device)
    metrics.update(
        {
            "relative_threshold": threshold,
            "matches_threshold": metrics["max_relative_error"] <= threshold,
            "strict_allclose_match": bool(torch.allclose(cand, ref, atol=atol, rtol=rtol)),
        }
    )
    return metrics


def require_kernel_dtype(dtype, what="data"):
    if dtype not in KERNEL_VALUE_DTYPES:
        raise TypeError(
            f"{what} dtype {str(dtype).replace('torch.', '')} has no HIP kernel; "
            "supported: float16, bfloat16, float32, float64"
        )


class _Timer:
    """Host wall-clock timing between device synchronisations (the reference's
    return_time convention, ops/spmv_csr.py:902-923)."""

    def __init__(self, enabled):
        self.enabled = bool(enabled)
        self.t = None

    def mark(self):
        if not self.enabled:
            return None
        torch.cuda.synchronize()
        now = time.perf_counter()
        prev, self.t = self.t, now
        return None if prev is None else (now - prev) * 1000.0
