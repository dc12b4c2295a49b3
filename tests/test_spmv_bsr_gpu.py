"""BSR SpMV (non / trans / conj, block sizes 2..32) against the FP64 oracle
(the C BSR product for non, the CSR expansion's transposed SpMV for trans),
the reference's argument contract (padded output, logical or padded x,
algorithm and op checks) and determinism."""

import numpy as np
import pytest
import torch

import flagsparse_amd as fs
from flagsparse_amd import synthetic as S
from oracle import oracle as O

from ._helpers import REL_TOL, rel_err

pytestmark = pytest.mark.gpu


def _bsr(m, n, d, density, seed):
    nbr, nbc = -(-m // d), -(-n // d)
    rng = np.random.default_rng(seed)
    lengths = rng.binomial(nbc, density, nbr)
    bp, bc, bd = S.bsr_blocks(nbr, nbc, lengths, block=d, seed=seed)
    return bp, bc, bd


def _dev(bp, bc, bd, device, dtype, index_dtype=torch.int32):
    return (torch.from_numpy(bd).to(device, dtype), torch.from_numpy(bc).to(device, index_dtype),
            torch.from_numpy(bp).to(device, index_dtype))


@pytest.mark.parametrize("d", [2, 3, 4, 5, 8, 12, 16, 32])
@pytest.mark.parametrize("op", ["non", "trans", "conj"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_bsr_spmv_matches_oracle(device, d, op, dtype):
    m, n = 7 * d + 3, 5 * d + 1  # ragged: the last block row / column is padding
    bp, bc, bd = _bsr(m, n, d, 0.5, seed=d)
    nbr, nbc = bp.size - 1, -(-n // d)
    data, indices, indptr = _dev(bp, bc, bd, device, dtype)
    trans = op != "non"
    xlen = m if trans else n
    x = np.random.default_rng(d + 1).uniform(-1, 1, xlen)
    y = fs.flagsparse_spmv_bsr(data, indices, indptr, S.dense(x, device, dtype), (m, n), block_dim=d, op=op)
    assert y.numel() == (nbc if trans else nbr) * d  # padded output, as the reference
    ip, ix, v = S.bsr_to_csr(bp, bc, bd)  # full padded CSR
    xp = np.zeros((nbr if trans else nbc) * d)
    xp[:xlen] = x
    want = O.spmv(ip, ix, v, xp, (nbr * d, nbc * d), "trans" if trans else "non")
    assert rel_err(y, want) <= REL_TOL[dtype]
    if not trans:
        np.testing.assert_allclose(O.bsr_spmm(bp, bc, bd, xp[:, None], bs=d)[:, 0], want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64])
def test_bsr_spmv_prepared_padded_x_out_and_meta(device, index_dtype):
    d, m, n = 4, 50, 38
    bp, bc, bd = _bsr(m, n, d, 0.4, seed=3)
    data, indices, indptr = _dev(bp, bc, bd, device, torch.float32, index_dtype)
    p = fs.prepare_spmv_bsr(data, indices, indptr, (m, n), d, use_opt=True)
    x = torch.rand(n, device=device)
    xp = torch.zeros(p.padded_n_cols, device=device)
    xp[:n] = x
    y1 = fs.flagsparse_spmv_bsr(x=x, prepared=p)
    y2 = fs.flagsparse_spmv_bsr(x=xp, prepared=p, use_opt=True)
    assert torch.equal(y1, y2)
    out = torch.empty(p.padded_n_rows, device=device)
    y3, meta = fs.flagsparse_spmv_bsr(x=x, prepared=p, out=out, return_meta=True)
    assert y3 is out and torch.equal(out, y1)
    assert meta["padded_shape"] == (p.padded_n_rows, p.padded_n_cols) and meta["nnzb"] == bc.size
    for _ in range(3):  # deterministic
        assert torch.equal(fs.flagsparse_spmv_bsr(x=x, prepared=p), y1)


def test_bsr_spmv_contract_errors(device):
    d, m, n = 4, 20, 20
    bp, bc, bd = _bsr(m, n, d, 0.5, seed=5)
    data, indices, indptr = _dev(bp, bc, bd, device, torch.float32)
    x = torch.rand(n, device=device)
    with pytest.raises(ValueError, match="only supports op='non'"):
        fs.flagsparse_spmv_bsr(data, indices, indptr, x, (m, n), block_dim=d, op="trans", use_opt=True)
    with pytest.raises(ValueError, match="block_dim values 2, 4, 8, and 16"):
        fs.prepare_spmv_bsr(torch.rand(0, 3, 3, device=device), indices[:0], torch.zeros(8, dtype=torch.int32,
                            device=device), (20, 20), 3, use_opt=True)
    with pytest.raises(ValueError, match="greater than 1"):
        fs.prepare_spmv_bsr(data, indices, indptr, (m, n), 1)
    with pytest.raises(ValueError, match="x length must be"):
        fs.flagsparse_spmv_bsr(data, indices, indptr, torch.rand(n + 1, device=device), (m, n), block_dim=d)
    with pytest.raises(ValueError, match="transpose conflicts with op"):
        fs.flagsparse_spmv_bsr(data, indices, indptr, x, (m, n), block_dim=d, op="non", transpose=True)
    bad = indices.clone()
    bad[0] = 5
    with pytest.raises(IndexError, match="n_block_cols"):
        fs.flagsparse_spmv_bsr(data, bad, indptr, x, (m, n), block_dim=d)
    with pytest.raises(TypeError, match="BSR SpMV supports"):
        fs.flagsparse_spmv_bsr(data.half(), indices, indptr, x.half(), (m, n), block_dim=d)
