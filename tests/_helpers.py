"""Shared helpers of the GPU parity tests: seeded CSR inputs on the device and
the oracle comparison with the north-star tolerances (max|d|/max|ref|)."""

import numpy as np
import torch

from flagsparse_amd import synthetic
from oracle import oracle

# relative error bounds (max|cand - ref| / max|ref|) against the FP64 oracle
REL_TOL = {
    torch.float32: 1e-5,
    torch.float64: 1e-12,
    torch.bfloat16: 1e-2,
    torch.float16: 5e-3,
}


def np_dtype_round(a, dtype):
    """Values as the kernel sees them: rounded to `dtype`, back in float64."""
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).to(torch.float64).numpy()


def device_csr(ip, ix, v, device, dtype=torch.float32, index_dtype=torch.int32, indptr_dtype=None):
    return synthetic.to_torch(ip, ix, v, device, dtype, index_dtype, indptr_dtype)


def rel_err(cand, ref):
    c = cand.detach().to(torch.float64).cpu().numpy() if torch.is_tensor(cand) else np.asarray(cand)
    return oracle.max_relative_error(c, ref)


def assert_rel(cand, ref, dtype, what=""):
    err = rel_err(cand, ref)
    tol = REL_TOL[dtype]
    assert err <= tol, f"{what}: rel err {err:.3e} > {tol:.1e}"
    return err


def long_row_matrix(n_rows, n_cols, long_rows, long_len, short_len, seed):
    """Rows of `short_len` nonzeros plus a few rows of `long_len` (distinct
    sorted columns)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lengths = np.full(n_rows, short_len, dtype=np.int64)
    lengths[list(long_rows)] = long_len
    ip, ix = synthetic.distinct_sorted_columns(rng, lengths, n_cols)
    v = rng.uniform(-1, 1, ix.size)
    return ip, ix, v
