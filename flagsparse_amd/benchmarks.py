"""Benchmark harness of the SpMV / SpMM / SDDMM operators (reference harness
semantics: ops/_common.py:186-218, :340-405; ops/benchmarks.py:411-563;
ops/spmm_csr.py:4440-4793; ops/sddmm_csr.py:520-625).

The ``benchmark_*_case`` functions keep the reference's signatures and the
layout of the dict they return (``parameters`` / ``performance`` /
``verification`` / ``backend_status`` / ``samples``), so report readers keyed
on it keep working.  Key names are the reference's: ``triton_*`` entries hold
this package's HIP kernels, ``pytorch_*`` entries hold ``torch.sparse`` on the
GPU -- which on ROCm builds of torch is rocSPARSE, the vendor library -- and
``cusparse_*`` entries stay None with the reason set, since CuPy/cuSPARSE do
not exist on ROCm.

Inputs are the reference's: unseeded, rows drawn uniformly (duplicate
coordinates possible, summed by every path), values ``randn``, indptr int64.
The golden result is torch.sparse in float64 on the CPU; the rocSPARSE leg
runs on the coalesced (sorted, duplicate-free) CSR of the same matrix.
Timing is the reference's too: host wall clock between device
synchronisations, averaged over ``iters`` launches after ``warmup``.
"""

import statistics
import time

import torch

from ._common import _tolerance_for_dtype, normalize_op, op_to_name, op_transposes, validation_metrics

_VENDOR_NOTE = "CuPy/cuSPARSE is not available on ROCm; torch.sparse (rocSPARSE) is timed as pytorch_ms"


# ------------------------------------------------------------------ inputs
def _build_random_dense(dense_size, value_dtype, device):
    """ops/_common.py:186-194 (real dtypes; complex has no HIP kernel here)."""
    if value_dtype in (torch.float16, torch.bfloat16, torch.float32, torch.float64):
        return torch.randn(dense_size, dtype=value_dtype, device=device)
    raise TypeError(f"Unsupported value dtype: {value_dtype}")


def _build_random_csr(n_rows, n_cols, nnz, value_dtype, index_dtype, device):
    """ops/_common.py:203-218: nnz row ids drawn uniformly and sorted, column
    ids drawn uniformly (unsorted within a row, duplicates possible)."""
    if nnz <= 0 or n_rows <= 0 or n_cols <= 0:
        indptr = torch.zeros(n_rows + 1, dtype=torch.int64, device=device)
        return (torch.empty(0, dtype=value_dtype, device=device), torch.empty(0, dtype=index_dtype, device=device),
                indptr)
    rows, _ = torch.sort(torch.randint(0, n_rows, (nnz,), device=device))
    indptr = torch.zeros(n_rows + 1, dtype=torch.int64, device=device)
    indptr[1:] = torch.cumsum(torch.bincount(rows, minlength=n_rows), dim=0)
    indices = torch.randint(0, n_cols, (nnz,), dtype=index_dtype, device=device)
    return _build_random_dense(nnz, value_dtype, device), indices, indptr


# ------------------------------------------------------------------ timers
def _benchmark_cuda_op(op, warmup, iters):
    """ops/_common.py:340-358: (last output, mean ms per call)."""
    warmup = max(0, int(warmup))
    iters = max(1, int(iters))
    output = None
    for _ in range(warmup):
        output = op()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        output = op()
    torch.cuda.synchronize()
    return output, (time.perf_counter() - t0) * 1000.0 / iters


def _benchmark_cuda_graph_op(op, *, graph_batch=100, warmup=20, repeats=10, capture_setup=None):
    """ops/_common.py:361-405: median ms per call of `op` replayed from a
    captured graph of `graph_batch` calls (op must not allocate or sync)."""
    graph_batch = max(1, int(graph_batch))
    warmup = max(0, int(warmup))
    repeats = max(1, int(repeats))
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(stream):
        if capture_setup is not None:
            capture_setup()
        op()
    stream.synchronize()
    with torch.cuda.graph(graph, stream=stream):
        for _ in range(graph_batch):
            op()
    stream.synchronize()
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            graph.replay()
    stream.synchronize()
    samples = []
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    for _ in range(repeats):
        with torch.cuda.stream(stream):
            start.record()
            graph.replay()
            end.record()
        end.synchronize()
        samples.append(float(start.elapsed_time(end)) / graph_batch)
    return statistics.median(samples)


def _coalesced(data, indices, indptr, shape, op_code):
    """op(A) as a coalesced COO tensor (sorted, duplicates summed)."""
    n_rows = int(indptr.numel()) - 1
    rows = torch.repeat_interleave(torch.arange(n_rows, dtype=torch.int64, device=indices.device),
                                   indptr[1:] - indptr[:-1])
    coo = torch.sparse_coo_tensor(torch.stack([rows, indices.to(torch.int64)]), data, shape).coalesce()
    return coo if op_code == 0 else coo.t().coalesce()


def _golden(data, indices, indptr, shape, op_code, dense):
    """op(A) @ dense in float64 on the CPU (torch.sparse), back on the device."""
    a = _coalesced(data.double().cpu(), indices.cpu(), indptr.cpu(), shape, op_code)
    return torch.sparse.mm(a, dense.double().cpu()).to(dense.device)


def _vendor_operand(data, indices, indptr, shape, op_code):
    """op(A) as a canonical CSR tensor on the device for torch.sparse.mm
    (rocSPARSE on ROCm builds of torch)."""
    return _coalesced(data, indices, indptr, shape, op_code).to_sparse_csr()


def _speedup(other_ms, ours_ms):
    return other_ms / ours_ms if (other_ms is not None and ours_ms and ours_ms > 0) else None


def _max_abs(a, b):
    return float(torch.max(torch.abs(a.to(torch.float64) - b.to(torch.float64))).item()) if a.numel() else 0.0


# -------------------------------------------------------------------- SpMV
def benchmark_spmv_case(n_rows=4096, n_cols=4096, nnz=65536, value_dtype=torch.float32, index_dtype=torch.int32,
                        warmup=20, iters=200, block_nnz=256, max_segments=None, run_cusparse=True, transpose=False,
                        op=None, index_fallback_policy="auto"):
    """HIP CSR SpMV vs torch.sparse (rocSPARSE); ops/benchmarks.py:411-563."""
    from .spmv_csr import flagsparse_spmv_csr, prepare_spmv_csr  # noqa: PLC0415

    device = torch.device("cuda")
    data, indices, indptr = _build_random_csr(n_rows, n_cols, nnz, value_dtype, index_dtype, device)
    op_code = normalize_op(op, transpose=transpose)
    if op is not None and bool(transpose) and op_code == 0:
        raise ValueError("transpose=True conflicts with op=non")
    transpose = op_transposes(op_code)
    x = _build_random_dense(n_rows if transpose else n_cols, value_dtype, device)
    shape = (n_rows, n_cols)
    prepared = prepare_spmv_csr(data, indices, indptr, shape, block_nnz=block_nnz, max_segments=max_segments,
                                op=op_code, index_fallback_policy=index_fallback_policy)
    ours, ours_ms = _benchmark_cuda_op(lambda: flagsparse_spmv_csr(x=x, prepared=prepared), warmup, iters)
    expected = _golden(data, indices, indptr, shape, op_code, x.unsqueeze(1)).squeeze(1)
    pytorch_ms, pytorch_reason, pytorch_values = None, None, None
    try:
        a = _vendor_operand(data, indices, indptr, shape, op_code)
        xv = x.unsqueeze(1)
        pytorch_values, pytorch_ms = _benchmark_cuda_op(lambda: torch.sparse.mm(a, xv), warmup, iters)
        pytorch_values = pytorch_values.squeeze(1)
    except RuntimeError as exc:  # a dtype rocSPARSE lacks
        pytorch_reason = str(exc)
    atol, rtol = _tolerance_for_dtype(value_dtype)
    exp_cast = expected.to(value_dtype)
    return {
        "parameters": {"n_rows": n_rows, "n_cols": n_cols, "nnz": nnz, "value_dtype": str(value_dtype),
                       "index_dtype": str(index_dtype), "op": op_to_name(op_code), "transpose": transpose,
                       "index_fallback_policy": str(index_fallback_policy).lower(), "warmup": warmup,
                       "iters": iters},
        "performance": {"triton_ms": ours_ms, "pytorch_ms": pytorch_ms, "cusparse_ms": None,
                        "triton_speedup_vs_pytorch": _speedup(pytorch_ms, ours_ms),
                        "triton_speedup_vs_cusparse": None},
        "verification": {
            "triton_match_reference": bool(torch.allclose(ours, exp_cast, atol=atol, rtol=rtol)),
            "triton_max_error": _max_abs(ours, expected),
            "triton_max_relative_error": validation_metrics(ours, expected, value_dtype)["max_relative_error"],
            "pytorch_match_reference": (None if pytorch_values is None else
                                        bool(torch.allclose(pytorch_values, exp_cast, atol=atol, rtol=rtol))),
            "cusparse_match_reference": None,
            "cusparse_max_error": None,
        },
        "backend_status": {"pytorch_unavailable_reason": pytorch_reason,
                           "cusparse_unavailable_reason": _VENDOR_NOTE if run_cusparse else None},
        "samples": {"triton": ours, "reference": expected, "pytorch": pytorch_values},
    }


# -------------------------------------------------------------------- SpMM
def benchmark_spmm_case(n_rows=4096, n_cols=4096, nnz=65536, n_dense_cols=32, value_dtype=torch.float32,
                        index_dtype=torch.int32, warmup=20, iters=200, block_n=None, block_nnz=None,
                        max_segments=None, run_cusparse=True, op="non"):
    """HIP CSR SpMM vs torch.sparse.mm (rocSPARSE); ops/spmm_csr.py:4440-4793."""
    from .spmm_csr import flagsparse_spmm_csr  # noqa: PLC0415

    op_code = normalize_op(op)
    device = torch.device("cuda")
    data, indices, indptr = _build_random_csr(n_rows, n_cols, nnz, value_dtype, index_dtype, device)
    b_rows = n_rows if op_transposes(op_code) else n_cols
    B = _build_random_dense((b_rows, n_dense_cols), value_dtype, device)
    shape = (n_rows, n_cols)
    kwargs = dict(block_n=block_n, block_nnz=block_nnz, max_segments=max_segments, op=op_to_name(op_code))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    flagsparse_spmm_csr(data, indices, indptr, B, shape, **kwargs)
    torch.cuda.synchronize()
    first_ms = (time.perf_counter() - t0) * 1000.0
    ours, ours_ms = _benchmark_cuda_op(lambda: flagsparse_spmm_csr(data, indices, indptr, B, shape, **kwargs),
                                       warmup, iters)
    expected = _golden(data, indices, indptr, shape, op_code, B)
    pytorch_ms, pytorch_reason, pytorch_values = None, None, None
    try:
        a = _vendor_operand(data, indices, indptr, shape, op_code)
        pytorch_values, pytorch_ms = _benchmark_cuda_op(lambda: torch.sparse.mm(a, B), warmup, iters)
    except RuntimeError as exc:
        pytorch_reason = str(exc)
    m = validation_metrics(ours, expected, value_dtype)
    pm = validation_metrics(pytorch_values, expected, value_dtype) if pytorch_values is not None else None
    return {
        "parameters": {"format": "csr", "n_rows": n_rows, "n_cols": n_cols, "nnz": nnz,
                       "n_dense_cols": n_dense_cols, "value_dtype": str(value_dtype),
                       "index_dtype": str(index_dtype), "op": op_to_name(op_code), "warmup": warmup,
                       "iters": iters, "block_n": block_n, "block_nnz": block_nnz, "max_segments": max_segments,
                       "run_cusparse": run_cusparse},
        "performance": {"pytorch_ms": pytorch_ms, "triton_ms": ours_ms, "triton_first_call_ms": first_ms,
                        "cusparse_ms": None, "triton_speedup_vs_pytorch": _speedup(pytorch_ms, ours_ms),
                        "triton_speedup_vs_cusparse": None},
        "verification": {
            "triton_match_reference": m["strict_allclose_match"],
            "triton_max_error": m["max_abs_error"],
            "triton_max_abs_error": m["max_abs_error"],
            "triton_max_relative_error": m["max_relative_error"],
            "triton_sum_relative_error": m["sum_relative_error"],
            "triton_relative_threshold": m["relative_threshold"],
            "triton_strict_allclose_match": m["strict_allclose_match"],
            "pytorch_match_reference": None if pm is None else pm["strict_allclose_match"],
            "pytorch_max_relative_error": None if pm is None else pm["max_relative_error"],
            "cusparse_match_reference": None,
        },
        "backend_status": {"pytorch_unavailable_reason": pytorch_reason, "pytorch_sparse_format": "CSR",
                           "cusparse_unavailable_reason": _VENDOR_NOTE if run_cusparse else None,
                           "flagsparse_internal_route": "hip-row-bins"},
        "samples": {"pytorch": pytorch_values, "triton": ours, "reference": expected, "cusparse": None},
    }


# ------------------------------------------------------------------- SDDMM
def _sddmm_gather(indices, indptr, x, y, data, alpha, beta):
    """The reference's torch formulation (ops/sddmm_csr.py:471-483)."""
    n_rows = int(indptr.numel()) - 1
    row_ids = torch.repeat_interleave(torch.arange(n_rows, dtype=torch.int64, device=indices.device),
                                      indptr[1:] - indptr[:-1])
    if row_ids.numel() == 0:
        return torch.empty(0, dtype=x.dtype, device=x.device)
    vals = alpha * torch.sum(x[row_ids] * y[indices.to(torch.int64)], dim=1)
    return vals + beta * data if data is not None else vals


def benchmark_sddmm_case(n_rows=1024, n_cols=1024, nnz=16384, k_dim=64, value_dtype=torch.float32, warmup=10,
                         iters=30, alpha=1.0, beta=0.0, run_cusparse=False):
    """HIP SDDMM vs the torch gather formulation; ops/sddmm_csr.py:520-625."""
    from .sddmm_csr import SUPPORTED_SDDMM_VALUE_DTYPES, flagsparse_sddmm_csr, prepare_sddmm_csr  # noqa: PLC0415

    if value_dtype not in SUPPORTED_SDDMM_VALUE_DTYPES:
        raise TypeError("value_dtype must be torch.float32 or torch.float64")
    device = torch.device("cuda")
    data, indices, indptr = _build_random_csr(n_rows, n_cols, nnz, value_dtype, torch.int32, device)
    x = _build_random_dense((n_rows, k_dim), value_dtype, device)
    y = _build_random_dense((n_cols, k_dim), value_dtype, device)
    prepared = prepare_sddmm_csr(indices, indptr, (n_rows, n_cols), k_hint=k_dim)
    ours, ours_ms = _benchmark_cuda_op(
        lambda: flagsparse_sddmm_csr(data=data, x=x, y=y, alpha=alpha, beta=beta, prepared=prepared), warmup, iters)
    ref, pytorch_ms = _benchmark_cuda_op(lambda: _sddmm_gather(indices, indptr, x, y, data, alpha, beta), warmup,
                                         iters)
    # float64 golden of the same formula: both f32 candidates carry rounding
    # error, so elementwise allclose between them fails near zero; the
    # north-star metric is max|cand - golden| / max|golden|
    golden = _sddmm_gather(indices, indptr, x.double(), y.double(), data.double(), alpha, beta)
    atol, rtol = _tolerance_for_dtype(value_dtype)
    return {
        "parameters": {"n_rows": n_rows, "n_cols": n_cols, "nnz": nnz, "k_dim": k_dim,
                       "value_dtype": str(value_dtype), "warmup": warmup, "iters": iters, "alpha": alpha,
                       "beta": beta},
        "performance": {"triton_ms": ours_ms, "pytorch_ms": pytorch_ms, "cusparse_ms": None,
                        "triton_speedup_vs_pytorch": _speedup(pytorch_ms, ours_ms),
                        "triton_speedup_vs_cusparse": None},
        "verification": {"triton_match_pytorch": bool(torch.allclose(ours, ref, atol=atol, rtol=rtol)),
                         "triton_max_abs_error": _max_abs(ours, ref),
                         "triton_max_relative_error": validation_metrics(ours, golden, value_dtype)[
                             "max_relative_error"],
                         "pytorch_max_relative_error": validation_metrics(ref, golden, value_dtype)[
                             "max_relative_error"],
                         "cusparse_match_pytorch": None},
        "backend_status": {"cusparse_unavailable_reason": _VENDOR_NOTE if run_cusparse else None},
        "samples": {"triton": ours, "pytorch": ref, "reference": golden},
    }


# ------------------------------------------------------- SpMM opt / alg2 cases
# (atol, rtol) of the reference's SpMM reference checks (spmm_csr.py:74-83)
_OPT_TOLERANCE = {
    torch.float16: (1e-3, 2e-3),
    torch.bfloat16: (0.016, 1e-1),
    torch.float32: (1.3e-6, 1e-3),
    torch.float64: (1e-7, 1e-5),
}


def _opt_reference_error(candidate, reference, value_dtype):
    """max |c - r| / (atol + rtol |r|): the opt cases' error figure
    (spmm_csr.py:4304-4310, spmm_csr_opt_alg2.py:1243-1249); <= 1 passes."""
    atol, rtol = _OPT_TOLERANCE.get(value_dtype, (1e-6, 1e-5))
    if candidate.numel() == 0:
        return 0.0
    diff = torch.abs(candidate.to(torch.float64) - reference.to(torch.float64))
    return float(torch.max(diff / (atol + rtol * torch.abs(reference.to(torch.float64)))).item())


def _opt_case_inputs(n_rows, n_cols, nnz, n_dense_cols, value_dtype, index_dtype):
    device = torch.device("cuda")
    data, indices, indptr = _build_random_csr(n_rows, n_cols, nnz, value_dtype, index_dtype, device)
    B = _build_random_dense((n_cols, n_dense_cols), value_dtype, device)
    shape = (n_rows, n_cols)
    reference = _golden(data, indices, indptr, shape, 0, B).to(value_dtype)
    return data, indices, indptr, B, shape, reference


def _first_call_ms(fn):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1000.0


def _rocsparse_ms(data, indices, indptr, shape, B, warmup, iters):
    try:
        a = _vendor_operand(data, indices, indptr, shape, 0)
        return _benchmark_cuda_op(lambda: torch.sparse.mm(a, B), warmup, iters)[1]
    except RuntimeError:
        return None


def benchmark_spmm_opt_case(n_rows=4096, n_cols=4096, nnz=65536, n_dense_cols=32, value_dtype=torch.float32,
                            index_dtype=torch.int32, warmup=20, iters=200, run_cusparse=True):
    """SpMM base vs opt against the float64 reference (spmm_csr.py:4313-4437):
    base is flagsparse_spmm_csr, opt flagsparse_spmm_csr_opt over a prepared
    object, pt_ms torch.sparse.mm (rocSPARSE); cu_ms stays None (no
    CuPy/cuSPARSE on ROCm)."""
    from .spmm_csr import flagsparse_spmm_csr, flagsparse_spmm_csr_opt, prepare_spmm_csr_opt  # noqa: PLC0415

    if value_dtype not in (torch.float32, torch.float64):
        raise TypeError("benchmark_spmm_opt_case only supports float32 and float64")
    data, indices, indptr, B, shape, ref = _opt_case_inputs(n_rows, n_cols, nnz, n_dense_cols, value_dtype,
                                                            index_dtype)
    prepared = prepare_spmm_csr_opt(data, indices, indptr, shape)
    base_first_call_ms = _first_call_ms(lambda: flagsparse_spmm_csr(data, indices, indptr, B, shape))
    base_values, base_ms = _benchmark_cuda_op(lambda: flagsparse_spmm_csr(data, indices, indptr, B, shape),
                                              warmup, iters)
    opt_first_call_ms = _first_call_ms(lambda: flagsparse_spmm_csr_opt(B=B, prepared=prepared))
    opt_values, opt_ms = _benchmark_cuda_op(lambda: flagsparse_spmm_csr_opt(B=B, prepared=prepared), warmup, iters)
    pt_ms = _rocsparse_ms(data, indices, indptr, shape, B, warmup, iters)
    err_base = _opt_reference_error(base_values, ref, value_dtype)
    err_opt = _opt_reference_error(opt_values, ref, value_dtype)
    return {
        "parameters": {"n_rows": n_rows, "n_cols": n_cols, "nnz": nnz, "n_dense_cols": n_dense_cols,
                       "value_dtype": str(value_dtype), "index_dtype": str(index_dtype)},
        "performance": {
            "base_ms": base_ms, "base_first_call_ms": base_first_call_ms, "opt_ms": opt_ms,
            "opt_first_call_ms": opt_first_call_ms, "pt_ms": pt_ms, "cu_ms": None,
            "opt_vs_base": _speedup(base_ms, opt_ms), "opt_vs_pt": _speedup(pt_ms, opt_ms), "opt_vs_cu": None,
        },
        "verification": {"err_base": err_base, "err_opt": err_opt, "base_ok": err_base <= 1.0,
                         "opt_ok": err_opt <= 1.0, "status": "PASS" if err_opt <= 1.0 else "FAIL"},
        "backend_status": {"long_row_fallback_used": bool(prepared.long_row_fallback_only),
                           "flagsparse_internal_route": "hip-row-bins",
                           "cusparse_unavailable_reason": _VENDOR_NOTE if run_cusparse else None},
        "samples": {"base": base_values, "opt": opt_values, "reference": ref},
    }


def benchmark_spmm_opt_alg2_case(n_rows=4096, n_cols=4096, nnz=65536, n_dense_cols=32, value_dtype=torch.float32,
                                 index_dtype=torch.int32, warmup=20, iters=200, run_cusparse=True):
    """SpMM base, opt and opt-alg2 against the float64 reference
    (spmm_csr_opt_alg2.py:1252-1400); torch_ms is torch.sparse.mm
    (rocSPARSE), the cusparse entries stay None with the reason set."""
    from .spmm_csr import (  # noqa: PLC0415
        flagsparse_spmm_csr,
        flagsparse_spmm_csr_opt,
        flagsparse_spmm_csr_opt_alg2,
        prepare_spmm_csr_opt,
        prepare_spmm_csr_opt_alg2,
    )

    if value_dtype not in (torch.float32, torch.float64):
        raise TypeError("benchmark_spmm_opt_alg2_case only supports float32 and float64")
    data, indices, indptr, B, shape, ref = _opt_case_inputs(n_rows, n_cols, nnz, n_dense_cols, value_dtype,
                                                            index_dtype)
    prepared_opt = prepare_spmm_csr_opt(data, indices, indptr, shape)
    prepared_alg2 = prepare_spmm_csr_opt_alg2(data, indices, indptr, shape)
    base_values, base_ms = _benchmark_cuda_op(lambda: flagsparse_spmm_csr(data, indices, indptr, B, shape),
                                              warmup, iters)
    opt_values, opt_ms = _benchmark_cuda_op(lambda: flagsparse_spmm_csr_opt(B=B, prepared=prepared_opt), warmup,
                                            iters)
    alg2_values, alg2_ms = _benchmark_cuda_op(lambda: flagsparse_spmm_csr_opt_alg2(B=B, prepared=prepared_alg2),
                                              warmup, iters)
    torch_ms = _rocsparse_ms(data, indices, indptr, shape, B, warmup, iters)
    return {
        "parameters": {"n_rows": n_rows, "n_cols": n_cols, "nnz": nnz, "n_dense_cols": n_dense_cols,
                       "value_dtype": str(value_dtype), "index_dtype": str(index_dtype)},
        "performance": {"base_ms": base_ms, "opt_ms": opt_ms, "opt_alg2_ms": alg2_ms, "torch_ms": torch_ms,
                        "cusparse_ms": None},
        "verification": {
            "base_vs_torch_err": _opt_reference_error(base_values, ref, value_dtype),
            "opt_vs_torch_err": _opt_reference_error(opt_values, ref, value_dtype),
            "opt_alg2_vs_torch_err": _opt_reference_error(alg2_values, ref, value_dtype),
            "base_vs_cusparse_err": None,
            "opt_vs_cusparse_err": None,
            "opt_alg2_vs_cusparse_err": None,
        },
        "backend_status": {"cusparse_unavailable_reason": _VENDOR_NOTE if run_cusparse else None},
        "samples": {"base_triton": base_values, "opt_triton": opt_values, "opt_alg2_triton": alg2_values,
                    "torch_ref": ref, "cusparse_ref": None},
    }


def comprehensive_spmm_test(n_rows=4096, n_cols=4096, nnz=65536, n_dense_cols=32, dtype=torch.float32,
                            index_dtype=torch.int32, warmup=20, iters=200, block_n=None, block_nnz=None,
                            max_segments=None, run_cusparse=True, op="non"):
    """Full SpMM benchmark entry for one configuration: benchmark_spmm_case
    with the value dtype passed as ``dtype`` (spmm_csr.py:4796-4825)."""
    return benchmark_spmm_case(n_rows=n_rows, n_cols=n_cols, nnz=nnz, n_dense_cols=n_dense_cols, value_dtype=dtype,
                               index_dtype=index_dtype, warmup=warmup, iters=iters, block_n=block_n,
                               block_nnz=block_nnz, max_segments=max_segments, run_cusparse=run_cusparse, op=op)
