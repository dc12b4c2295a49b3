"""Operands past 4 GiB: the kernels switch from 32-bit byte offsets to 64-bit
ones when an operand they address does not fit in 4 GiB, even with int32
indptr / indices.  Each case pairs a small A (256 rows) with a dense operand of
2**24 + 1024 rows of 64 float32 (4 GiB + 256 KiB) and puts half of A's
columns in its last 4096 rows, so the addressed bytes lie beyond 2**32.

The oracle sees only the columns A references: those rows of the big operand
are gathered to the host and A's column ids renumbered into them, which leaves
every product unchanged.
"""

import numpy as np
import pytest
import torch

import flagsparse_amd as fs
from oracle import oracle

from ._helpers import assert_rel

pytestmark = pytest.mark.gpu

BIG = (1 << 24) + 1024  # rows of the big operand: BIG * 64 * 4 bytes > 2**32
M, N_DENSE, PER_ROW = 256, 64, 8


def _pattern(seed):
    """256 rows of 8 distinct sorted columns: 4 anywhere, 4 in the last 4096."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cols = []
    for _ in range(M):
        lo = rng.choice(BIG - 4096, 4, replace=False)
        hi = BIG - 4096 + rng.choice(4096, 4, replace=False)
        cols.append(np.sort(np.concatenate([lo, hi])))
    ix = np.concatenate(cols).astype(np.int64)
    ip = np.arange(0, M * PER_ROW + 1, PER_ROW, dtype=np.int64)
    v = rng.uniform(-1, 1, ix.size)
    return ip, ix, v


def _compressed(ix):
    """Sorted distinct referenced columns and A's column ids renumbered into them."""
    uniq = np.unique(ix)
    return uniq, np.searchsorted(uniq, ix).astype(np.int64)


def _device_csr(ip, ix, v, device):
    return (torch.tensor(v, dtype=torch.float32, device=device), torch.tensor(ix, dtype=torch.int32, device=device),
            torch.tensor(ip, dtype=torch.int32, device=device))


def _big(device, seed):
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    return torch.randn((BIG, N_DENSE), device=device, generator=g)


def _f32(v):
    return np.asarray(v, dtype=np.float32).astype(np.float64)


def test_spmm_b_past_4gib(device):
    ip, ix, v = _pattern(11)
    data, idx, ptr = _device_csr(ip, ix, v, device)
    B = _big(device, 12)
    C = fs.flagsparse_spmm_csr(data, idx, ptr, B, (M, BIG))
    uniq, rix = _compressed(ix)
    b_sub = B[torch.tensor(uniq, device=device)].double().cpu().numpy()
    del B
    ref = oracle.spmm(ip, rix, _f32(v), b_sub, (M, uniq.size))
    assert C.shape == (M, N_DENSE)
    assert_rel(C, ref, torch.float32, "B of 4 GiB + 256 KiB")
    torch.cuda.empty_cache()


def test_spmm_trans_c_past_4gib(device):
    ip, ix, v = _pattern(13)
    data, idx, ptr = _device_csr(ip, ix, v, device)
    rng = np.random.default_rng(14)
    b = rng.standard_normal((M, N_DENSE))
    C = fs.flagsparse_spmm_csr(data, idx, ptr, torch.tensor(b, dtype=torch.float32, device=device), (M, BIG),
                               op="trans")
    assert C.shape == (BIG, N_DENSE)
    uniq, rix = _compressed(ix)
    ref = oracle.spmm(ip, rix, _f32(v), _f32(b), (M, uniq.size), op="trans")
    rows = torch.tensor(uniq, device=device)
    assert_rel(C[rows], ref, torch.float32, "C of 4 GiB + 256 KiB (trans)")
    # every row A's columns do not reach is exactly zero
    assert int(torch.count_nonzero(C)) == int(torch.count_nonzero(C[rows]))
    del C
    torch.cuda.empty_cache()


def test_sddmm_y_past_4gib(device):
    ip, ix, v = _pattern(15)
    data, idx, ptr = _device_csr(ip, ix, v, device)
    rng = np.random.default_rng(16)
    X = rng.standard_normal((M, N_DENSE))
    Y = _big(device, 17)
    out = fs.flagsparse_sddmm_csr(data=data, indices=idx, indptr=ptr, x=torch.tensor(X, dtype=torch.float32,
                                                                                    device=device),
                                  y=Y, shape=(M, BIG), alpha=0.75, beta=0.5)
    uniq, rix = _compressed(ix)
    y_sub = Y[torch.tensor(uniq, device=device)].double().cpu().numpy()
    del Y
    ref = oracle.sddmm(ip, rix, _f32(X), y_sub, _f32(v), 0.75, 0.5)
    assert out.shape == (ix.size,)
    assert_rel(out, ref, torch.float32, "Y of 4 GiB + 256 KiB")
    torch.cuda.empty_cache()


def test_spmv_x_past_4gib(device):
    """x of 2**30 + 1024 float32 (4 GiB + 4 KiB); columns in its last 4096."""
    n = (1 << 30) + 1024
    rng = np.random.Generator(np.random.PCG64(18))
    ix = np.concatenate([np.sort(n - 4096 + rng.choice(4096, PER_ROW, replace=False)) for _ in range(M)])
    ip = np.arange(0, M * PER_ROW + 1, PER_ROW, dtype=np.int64)
    v = rng.uniform(-1, 1, ix.size)
    data, idx, ptr = _device_csr(ip, ix, v, device)
    g = torch.Generator(device=device)
    g.manual_seed(19)
    x = torch.randn(n, device=device, generator=g)
    y = fs.flagsparse_spmv_csr(data, idx, ptr, x, (M, n))
    uniq, rix = _compressed(ix)
    x_sub = x[torch.tensor(uniq, device=device)].double().cpu().numpy()
    del x
    ref = oracle.spmv(ip, rix, _f32(v), x_sub, (M, uniq.size))
    assert_rel(y, ref, torch.float32, "x of 4 GiB + 4 KiB")
    torch.cuda.empty_cache()
