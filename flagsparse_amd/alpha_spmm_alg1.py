"""AlphaSparse-ALG1 CSR SpMM entry points (reference: alpha_spmm_alg1.py).

The reference carries four routes of AlphaSparse's CSR ALG1 SpMM,
C = A @ B with row-major B: the baseline (``flagsparse_alpha_spmm_alg1``),
a FlagTree TLE-struct kernel (``_tle``) and two launch-tuned variants of it
(``_tle_opt``, ``_tle_opt2``).  They share one prepared type, one validation
contract and one argument convention (raw CSR or ``prepared``, never both).

Here the same names, checks, messages and returns run the HIP row-bin SpMM
kernel (``fsp_spmm_csr``): the baseline and the first two TLE routes take the
default launch for the dense width, ``_tle_opt2`` (the reference's
shape-bucket-tuned route) takes the autotuner's winner for the matrix's
shape class, as opt-alg2 does.  No FlagTree runtime is involved, so the
availability queries report True.  ``meta`` describes the HIP launch that
runs (lanes per row along N, vector width, row bins); the reference's Triton
warp/factor figures have no counterpart in it.  Complex values pass the
reference's dtype check and then raise TypeError at launch (no HIP kernel).
"""

import torch

from . import _plan, autotune
from ._common import require_kernel_dtype
from ._matrix import CsrHandle
from .spmm_csr import _prepare_spmm_csr_matrix

SUPPORTED_ALPHA_SPMM_ALG1_DTYPES = (torch.float32, torch.float64, torch.complex64, torch.complex128)


class PreparedAlphaSpmmAlg1:
    """Cached CSR metadata for repeated alpha_spmm_alg1 calls
    (alpha_spmm_alg1.py:60-92), plus the HIP handle with its row-bin plans."""

    __slots__ = (
        "data",
        "kernel_indices",
        "kernel_indptr",
        "shape",
        "n_rows",
        "n_cols",
        "row_lengths",
        "max_row_nnz",
        "_handle",
        "_params",
    )

    def __init__(self, data, kernel_indices, kernel_indptr, shape, n_rows, n_cols, row_lengths, max_row_nnz):
        self.data = data
        self.kernel_indices = kernel_indices
        self.kernel_indptr = kernel_indptr
        self.shape = (int(shape[0]), int(shape[1]))
        self.n_rows = int(n_rows)
        self.n_cols = int(n_cols)
        self.row_lengths = row_lengths
        self.max_row_nnz = int(max_row_nnz)
        self._handle = CsrHandle(kernel_indices, kernel_indptr, self.shape)
        self._params = {}  # n_dense -> TuneParams of the tuned route


def _prepare_alpha_spmm_alg1_common(data, indices, indptr, shape):
    """alpha_spmm_alg1.py:95-119: the CSR SpMM validation, then the dtype check."""
    (data, kernel_indices, kernel_indptr, n_rows, n_cols, row_lengths, max_row_nnz) = _prepare_spmm_csr_matrix(
        data, indices, indptr, shape
    )
    if data.dtype not in SUPPORTED_ALPHA_SPMM_ALG1_DTYPES:
        raise TypeError("alpha_spmm_alg1 TLE routes only support float32, float64, complex64, and complex128")
    return PreparedAlphaSpmmAlg1(data, kernel_indices, kernel_indptr, shape, n_rows, n_cols, row_lengths,
                                 max_row_nnz)


def prepare_alpha_spmm_alg1(data, indices, indptr, shape):
    """Prepare CSR metadata for the baseline ALG1 route."""
    return _prepare_alpha_spmm_alg1_common(data, indices, indptr, shape)


def prepare_alpha_spmm_alg1_tle(data, indices, indptr, shape):
    """Prepare CSR metadata for the TLE-struct ALG1 route."""
    return _prepare_alpha_spmm_alg1_common(data, indices, indptr, shape)


def prepare_alpha_spmm_alg1_tle_opt(data, indices, indptr, shape):
    """Prepare CSR metadata for the TLE-struct ALG1 optimised route."""
    return _prepare_alpha_spmm_alg1_common(data, indices, indptr, shape)


def prepare_alpha_spmm_alg1_tle_opt2(data, indices, indptr, shape):
    """Prepare CSR metadata for the shape-bucket-tuned ALG1 route."""
    return _prepare_alpha_spmm_alg1_common(data, indices, indptr, shape)


def _validate_alpha_spmm_alg1_runtime_inputs(prepared, B, out):
    """alpha_spmm_alg1.py:154-181 (same checks, order and messages)."""
    if B is None:
        raise ValueError("B is required")
    if B.ndim != 2:
        raise ValueError("B must be a 2D dense tensor")
    if not B.is_cuda:
        raise ValueError("B must be a CUDA tensor")
    if B.device != prepared.data.device:
        raise ValueError("B must be on the same CUDA device as sparse matrix data")
    if B.dtype != prepared.data.dtype:
        raise TypeError("B dtype must match sparse matrix dtype")
    if B.shape[0] != prepared.n_cols:
        raise ValueError(f"B.shape[0] must be n_cols={prepared.n_cols}, got {B.shape[0]}")
    if out is not None:
        if not out.is_cuda:
            raise ValueError("out must be a CUDA tensor")
        if out.device != prepared.data.device:
            raise ValueError("out must be on the same CUDA device as sparse matrix data")
        if tuple(out.shape) != (prepared.n_rows, int(B.shape[1])) or out.dtype != prepared.data.dtype:
            raise ValueError("out shape/dtype must match result")
    return B.contiguous()


def _launch(prepared, n_dense, route):
    """(lanes along N, plan config, TuneParams or None) of a route's launch."""
    elem = prepared.data.element_size()
    if route == "alpha_spmm_alg1_tle_opt2" and prepared.data.dtype in (torch.float32, torch.float64):
        params = prepared._params.get(n_dense)
        if params is None:
            params = autotune.tune_spmm(prepared._handle, prepared.data, n_dense)
            prepared._params[n_dense] = params
        lanes = params.lanes_n or _plan.lanes_for_dense(n_dense, elem)
        return lanes, params.spmm_config(lanes), params
    lanes = _plan.lanes_for_dense(n_dense, elem)
    return lanes, _plan.spmm_config(lanes), None


def _runtime_meta(prepared, B, route):
    """The HIP launch of `route` for this B (the reference's meta keys that
    have a meaning here keep their names)."""
    require_kernel_dtype(prepared.data.dtype)
    n_dense = int(B.shape[1])
    device = prepared.data.device
    lanes, cfg, params = _launch(prepared, n_dense, route)
    vec = 16 // B.element_size() if (n_dense * B.element_size()) % 16 == 0 else 1
    props = torch.cuda.get_device_properties(device)
    meta = {
        "device_name": str(props.name),
        "sm_count": int(props.multi_processor_count),
        "warp_size": 64,
        "lanes_n": int(lanes),
        "vec": int(vec),
        "block_cols": int(lanes * vec),
        "n_dense_cols": n_dense,
        "row_bin_lanes": list(cfg.lanes),
        "row_bin_uppers": list(cfg.uppers),
        "unroll": int(cfg.unroll),
        "order_row_path": True,
        "alpha": 1,
        "beta": 0,
        "max_row_nnz": int(prepared.max_row_nnz),
        "loop_strategy": "row_bins",
        "launch_version": "hip_row_bins_tuned" if params is not None else "hip_row_bins",
        "backend": "hip",
    }
    if params is not None:
        meta["launch_configs"] = {"unroll": int(params.unroll), "slot_scale": int(params.slot_scale),
                                  "blocks_per_cu": int(params.blocks_per_cu), "xcd_groups": int(params.xcd_groups)}
    return _with_route(meta, route)


def _with_route(meta, route):
    out = dict(meta)
    out["route"] = route
    return out


def _run(prepared, B, meta, route, out):
    """Launch the route; a caller's meta may fix the lanes along N (a power
    of two <= 64, as build_*_meta reports them)."""
    require_kernel_dtype(prepared.data.dtype)
    n_dense = int(B.shape[1])
    lanes, cfg, params = _launch(prepared, n_dense, route)
    want = meta.get("lanes_n") if isinstance(meta, dict) else None
    if want is not None and int(want) != lanes:
        want = int(want)
        if want < 1 or want > 64 or want & (want - 1):
            raise ValueError(f"meta lanes_n must be a power of two in [1, 64], got {want}")
        lanes = want
        cfg = params.spmm_config(lanes) if params is not None else _plan.spmm_config(lanes)
    direct_out = out if (out is not None and out.is_contiguous()) else None
    C = prepared._handle.spmm(
        prepared.data, B, out=direct_out, lanes_n=lanes, config=cfg,
        max_blocks=params.max_blocks(prepared.data.device) if (params is not None and prepared.n_rows) else 0,
        xcd_groups=params.xcd_groups if params is not None else 0,
    )
    if out is not None and C is not out:
        out.copy_(C)
        C = out
    return C


def _entry(route, prepare, data, indices, indptr, B, shape, prepared, out, return_meta, meta=None):
    """The argument contract shared by the four public routes
    (alpha_spmm_alg1.py:1277-1425)."""
    if prepared is not None and not isinstance(prepared, PreparedAlphaSpmmAlg1):
        raise TypeError("prepared must be a PreparedAlphaSpmmAlg1 instance")
    raw_inputs_provided = any(arg is not None for arg in (data, indices, indptr, shape))
    if prepared is not None and raw_inputs_provided:
        raise ValueError("Pass either raw CSR inputs or prepared, not both")
    if prepared is None:
        if any(arg is None for arg in (data, indices, indptr, shape)):
            raise ValueError("data, indices, indptr, and shape are required when prepared is not provided")
        prepared = prepare(data, indices, indptr, shape)
    B = _validate_alpha_spmm_alg1_runtime_inputs(prepared, B, out)
    if meta is not None and not isinstance(meta, dict):
        raise TypeError("meta must be a dict from build_alpha_spmm_alg1_tle_opt[2]_meta")
    meta = _runtime_meta(prepared, B, route) if meta is None else _with_route(meta, route)
    C = _run(prepared, B, meta, route, out)
    if return_meta:
        return C, meta
    return C


def flagsparse_alpha_spmm_alg1(data=None, indices=None, indptr=None, B=None, shape=None, prepared=None, out=None,
                               return_meta=False):
    """C = A @ B, the ALG1 baseline route (alpha_spmm_alg1.py:1277-1311)."""
    return _entry("alpha_spmm_alg1", prepare_alpha_spmm_alg1, data, indices, indptr, B, shape, prepared, out,
                  return_meta)


def flagsparse_alpha_spmm_alg1_tle(data=None, indices=None, indptr=None, B=None, shape=None, prepared=None,
                                   out=None, return_meta=False):
    """C = A @ B, the TLE-struct route (alpha_spmm_alg1.py:1314-1348)."""
    return _entry("alpha_spmm_alg1_tle", prepare_alpha_spmm_alg1_tle, data, indices, indptr, B, shape, prepared,
                  out, return_meta)


def flagsparse_alpha_spmm_alg1_tle_opt(data=None, indices=None, indptr=None, B=None, shape=None, prepared=None,
                                       out=None, return_meta=False, meta=None):
    """C = A @ B, the optimised TLE route (alpha_spmm_alg1.py:1351-1385);
    `meta` from build_alpha_spmm_alg1_tle_opt_meta fixes the launch."""
    return _entry("alpha_spmm_alg1_tle_opt", prepare_alpha_spmm_alg1_tle_opt, data, indices, indptr, B, shape,
                  prepared, out, return_meta, meta)


def flagsparse_alpha_spmm_alg1_tle_opt2(data=None, indices=None, indptr=None, B=None, shape=None, prepared=None,
                                        out=None, return_meta=False, meta=None):
    """C = A @ B, the shape-bucket-tuned TLE route (alpha_spmm_alg1.py:1388-1425)."""
    return _entry("alpha_spmm_alg1_tle_opt2", prepare_alpha_spmm_alg1_tle_opt2, data, indices, indptr, B, shape,
                  prepared, out, return_meta, meta)


def build_alpha_spmm_alg1_tle_opt_meta(prepared, B):
    """The launch meta of the optimised route for this B (alpha_spmm_alg1.py:998-1003)."""
    B = _validate_alpha_spmm_alg1_runtime_inputs(prepared, B, None)
    return _runtime_meta(prepared, B, "alpha_spmm_alg1_tle_opt")


def build_alpha_spmm_alg1_tle_opt2_meta(prepared, B):
    """The launch meta of the tuned route for this B (alpha_spmm_alg1.py:1006-1012)."""
    B = _validate_alpha_spmm_alg1_runtime_inputs(prepared, B, None)
    return _runtime_meta(prepared, B, "alpha_spmm_alg1_tle_opt2")


def is_alpha_spmm_alg1_tle_available():
    """True: the route runs on the HIP kernels, which need no FlagTree runtime."""
    return True


def alpha_spmm_alg1_tle_unavailable_reason():
    """Why the TLE route is unavailable: never, so the reason is empty."""
    return ""


def is_alpha_spmm_alg1_tle_opt_available():
    return is_alpha_spmm_alg1_tle_available()


def alpha_spmm_alg1_tle_opt_unavailable_reason():
    return alpha_spmm_alg1_tle_unavailable_reason()


def is_alpha_spmm_alg1_tle_opt2_available():
    return is_alpha_spmm_alg1_tle_available()


def alpha_spmm_alg1_tle_opt2_unavailable_reason():
    return alpha_spmm_alg1_tle_unavailable_reason()
