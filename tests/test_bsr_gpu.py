"""GPU parity of the BSR path against the oracle.

* CSR -> BSR conversion: bit-exact against oracle.csr_to_bsr (itself pinned
  to scipy ``tobsr`` + ``sort_indices``), every value/index dtype, ragged
  shapes, empty matrices, and the BSR -> CSR -> BSR round trip at the full
  65536 x 65536 size.
* B operand packing: the MFMA layout Bp[k / 8][n][k % 8] checked element-wise.
* FP8 quantisation: codes and scales bit-exact against the oracle's float32
  restatement.
* MFMA BSR SpMM: bf16 / f16 against the FP64 oracle on the rounded inputs
  (float32 accumulation, rel 1e-5) and on the unrounded inputs (the bf16 /
  f16 storage bound); fp8 against the oracle on the dequantised operands
  (rel 1e-4, see FP8_VS_DEQUANT_TOL) with the error against unquantised FP64
  reported.
"""

import numpy as np
import pytest
import torch

from flagsparse_amd import bsr as FB
from flagsparse_amd import synthetic as S
from flagsparse_amd.spmm_csr import flagsparse_spmm_csr
from oracle import oracle as O

from ._helpers import REL_TOL, np_dtype_round, rel_err

pytestmark = pytest.mark.gpu

BS = 16
# The fp8 MFMA (v_mfma_f32_16x16x32_fp8_fp8) handles e4m3 subnormals exactly
# but sums its 32 products below float32 precision: on random full-range
# codes |D - exact| reaches ~6e-5 of max|D| (measured on MI355X).  Against the
# oracle on the dequantised operands the fp8 path is therefore held to 1e-4,
# the bf16 / f16 paths (exact products, float32 sums) to 1e-5.
FP8_VS_DEQUANT_TOL = 1e-4


def _dev_bsr(bp, bc, bd, device, dtype, indptr_dtype=torch.int32):
    data = torch.from_numpy(np.ascontiguousarray(bd)).to(device=device, dtype=dtype)
    return (data, torch.from_numpy(bc).to(device=device, dtype=torch.int32),
            torch.from_numpy(bp).to(device=device, dtype=indptr_dtype))


def _padded_b(B, n_block_cols):
    out = np.zeros((n_block_cols * BS, B.shape[1]))
    out[: B.shape[0]] = B
    return out


# ------------------------------------------------------------ CSR -> BSR
CONV_SHAPES = [(40, 33, 0.2), (64, 64, 0.05), (1, 100, 0.5), (100, 17, 0.3), (300, 260, 0.02), (16, 16, 1.0)]
CONV_DTYPES = [torch.float32, torch.float64, torch.bfloat16, torch.float16]
CONV_INDEX = [(torch.int32, torch.int32), (torch.int64, torch.int64), (torch.int32, torch.int64)]


@pytest.mark.parametrize("m,n,density", CONV_SHAPES)
@pytest.mark.parametrize("vdtype", CONV_DTYPES)
@pytest.mark.parametrize("index_dtype,indptr_dtype", CONV_INDEX)
def test_csr_to_bsr_bit_exact(device, m, n, density, vdtype, index_dtype, indptr_dtype):
    ip, ix, v = S.random_csr(m, n, density, seed=m * 7 + n)
    data, indices, indptr = S.to_torch(ip, ix, v, device, vdtype, index_dtype, indptr_dtype)
    bd, bc, bp = FB.csr_to_bsr(data, indices, indptr, (m, n))
    stored = data.to(torch.float64).cpu().numpy()
    rp, rc, rd = O.csr_to_bsr(ip, ix, stored, (m, n))
    assert bp.dtype == indptr_dtype and bc.dtype == torch.int32 and bd.dtype == vdtype
    np.testing.assert_array_equal(bp.cpu().numpy(), rp)
    np.testing.assert_array_equal(bc.cpu().numpy(), rc)
    np.testing.assert_array_equal(bd.to(torch.float64).cpu().numpy(), rd)


@pytest.mark.parametrize("m,n", [(0, 0), (0, 40), (40, 0), (37, 20)])
def test_csr_to_bsr_empty(device, m, n):
    ip, ix, v = S.random_csr(m, n, 0.0, seed=1, ensure_nonempty=False)
    data, indices, indptr = S.to_torch(ip, ix, v, device)
    bd, bc, bp = FB.csr_to_bsr(data, indices, indptr, (m, n))
    assert bd.shape == (0, BS, BS) and bc.numel() == 0
    np.testing.assert_array_equal(bp.cpu().numpy(), np.zeros(-(-m // BS) + 1, dtype=np.int64))


def test_csr_to_bsr_rejects_bad_input(device):
    ip, ix, v = S.random_csr(20, 20, 0.2, seed=2)
    data, indices, indptr = S.to_torch(ip, ix, v, device)
    with pytest.raises(ValueError, match="block_dim"):
        FB.csr_to_bsr(data, indices, indptr, (20, 20), block_dim=8)
    bad = indices.clone()
    bad[0] = 20
    with pytest.raises(IndexError, match="out of range"):
        FB.csr_to_bsr(data, bad, indptr, (20, 20))


# --------------------------------------------------------- operand packing
@pytest.mark.parametrize("K,N", [(16, 1), (90, 130), (256, 128), (33, 7)])
def test_pack_dense_operand_layout(device, K, N):
    B = np.random.default_rng(K + N).standard_normal((K, N))
    Bt = S.dense(B, device)
    Bp, absmax = FB.pack_dense_operand(Bt, torch.bfloat16)
    assert absmax is None
    k8, npad = -(-K // BS) * 2, -(-N // 128) * 128
    got = Bp.view(k8, npad, 8).permute(0, 2, 1).reshape(k8 * 8, npad).to(torch.float64).cpu().numpy()
    want = np.zeros((k8 * 8, npad))
    want[:K, :N] = np_dtype_round(B, torch.bfloat16)
    np.testing.assert_array_equal(got, want)


def test_pack_dense_operand_fp8_codes(device):
    B = np.random.default_rng(5).standard_normal((70, 40)) * 3.0
    Bp, absmax = FB.pack_dense_operand(S.dense(B, device), torch.uint8)
    k8, npad = -(-70 // BS) * 2, 128
    got = Bp.view(k8, npad, 8).permute(0, 2, 1).reshape(k8 * 8, npad).cpu().numpy()
    codes, scale = O.fp8_tensor_quantize_f32(B)
    want = np.zeros((k8 * 8, npad), dtype=np.uint8)
    want[:70, :40] = codes
    np.testing.assert_array_equal(got, want)
    assert float(absmax.item()) == float(np.abs(B.astype(np.float32)).max())
    assert np.float32(absmax.item()) / np.float32(448.0) == scale


# ------------------------------------------------------- FP8 quantisation
@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16, torch.float64])
def test_quantize_fp8_bit_exact(device, dtype):
    rng = np.random.default_rng(11)
    bd = rng.standard_normal((37, BS, BS)) * np.exp(rng.uniform(-8, 8, (37, 1, 1)))
    bd[5] = 0.0                        # all-zero block: scale 1, codes 0
    bd[6, 3, 4] = 1e-30                # tiny entry flushes to a signed zero code
    bd[7] = -bd[7]
    dev = S.dense(bd, device, dtype)
    codes, scales = FB.quantize_bsr_fp8(dev)
    src = dev.to(torch.float32).cpu().numpy()  # what the kernel reads, as f32
    want_codes, want_scales = O.fp8_block_quantize_f32(src)
    np.testing.assert_array_equal(scales.cpu().numpy(), want_scales)
    np.testing.assert_array_equal(codes.cpu().numpy(), want_codes)


# ------------------------------------------------------------ BSR SpMM
# (n_rows, n_cols, block-row lengths, N): ragged shapes, empty block rows,
# odd block counts (an unpaired last block), N across 128-column chunks.
SPMM_CASES = [
    (96, 144, [3, 1, 0, 5, 2, 9], 5),
    (120, 180, [4] * 8, 128),
    (256, 256, [16] * 16, 200),
    (520, 640, [7] * 33, 64),
    (80, 112, [1, 0, 1, 1, 0], 16),
    (16, 16, [1], 1),
]


def _spmm_case(case, seed):
    n_rows, n_cols, lengths, N = case
    nbr, nbc = -(-n_rows // BS), -(-n_cols // BS)
    assert len(lengths) == nbr
    bp, bc, bd = S.bsr_blocks(nbr, nbc, np.array(lengths), seed=seed)
    B = np.random.default_rng(seed + 1).standard_normal((n_cols, N))
    return n_rows, n_cols, nbc, bp, bc, bd, B


@pytest.mark.parametrize("case", SPMM_CASES)
@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("b_dtype", [torch.float32, "same"])
def test_bsr_spmm_bf16_f16(device, case, dtype, b_dtype):
    n_rows, n_cols, nbc, bp, bc, bd, B = _spmm_case(case, seed=len(case[2]) + case[3])
    data, indices, indptr = _dev_bsr(bp, bc, bd, device, dtype)
    Bt = S.dense(B, device, dtype if b_dtype == "same" else b_dtype)
    C = FB.flagsparse_spmm_bsr(data, indices, indptr, Bt, (n_rows, n_cols))
    assert C.dtype == torch.float32 and tuple(C.shape) == (n_rows, B.shape[1])
    rounded = O.bsr_spmm(bp, bc, np_dtype_round(bd, dtype), _padded_b(np_dtype_round(B, dtype), nbc))[:n_rows]
    exact = O.bsr_spmm(bp, bc, bd, _padded_b(B, nbc))[:n_rows]
    err = rel_err(C, rounded)
    assert err <= 1e-5, f"vs oracle on rounded inputs: {err:.3e}"
    err64 = rel_err(C, exact)
    assert err64 <= REL_TOL[dtype], f"vs FP64 oracle: {err64:.3e}"


@pytest.mark.parametrize("case", SPMM_CASES)
def test_bsr_spmm_fp8(device, case):
    n_rows, n_cols, nbc, bp, bc, bd, B = _spmm_case(case, seed=3 * len(case[2]) + case[3])
    data, indices, indptr = _dev_bsr(bp, bc, bd, device, torch.float32)
    codes, scales = FB.quantize_bsr_fp8(data)
    C = FB.flagsparse_spmm_bsr(codes, indices, indptr, S.dense(B, device), (n_rows, n_cols), scales=scales)
    a_codes, a_scales = O.fp8_block_quantize_f32(bd)
    b_codes, b_scale = O.fp8_tensor_quantize_f32(B)
    a_deq = O.e4m3_decode(a_codes) * a_scales.astype(np.float64)[:, None, None]
    b_deq = O.e4m3_decode(b_codes) * np.float64(b_scale)
    want = O.bsr_spmm(bp, bc, a_deq, _padded_b(b_deq, nbc))[:n_rows]
    err = rel_err(C, want)
    assert err <= FP8_VS_DEQUANT_TOL, f"vs oracle on the dequantised operands: {err:.3e}"
    err64 = rel_err(C, O.bsr_spmm(bp, bc, bd, _padded_b(B, nbc))[:n_rows])
    print(f"[fp8] rel err vs unquantised FP64: {err64:.3e}")
    assert err64 <= 0.1


ABSMAX_SHAPES = [(1, 1), (3, 5), (1000, 1), (257, 129), (65536, 128), (4097, 1000)]


@pytest.mark.parametrize("rows,cols", ABSMAX_SHAPES)
@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16, torch.float16, torch.float64])
@pytest.mark.parametrize("strided", [False, True])
def test_absmax_exact(device, rows, cols, dtype, strided):
    from flagsparse_amd import _lib

    g = torch.Generator(device="cpu").manual_seed(rows * 7 + cols)
    full = (torch.randn(rows, cols + (3 if strided else 0), generator=g) * 5).to(dtype).to(device)
    x = full[:, :cols]
    out = torch.full((1,), -1.0, dtype=torch.float32, device=device)
    _lib.check(_lib.load().fsp_absmax(_lib.value_code(dtype), x.data_ptr(), rows, cols, int(full.stride(0)),
                                      out.data_ptr(), _lib.stream_ptr(device)), "fsp_absmax")
    assert float(out.item()) == float(x.abs().max().float().item())


@pytest.mark.parametrize("bad", [float("inf"), float("-inf"), float("nan")])
def test_non_finite_b_gives_nan_fp8_product(device, bad):
    """A non-finite B makes the fp8 product NaN instead of silently zero
    (absmax keeps Inf / NaN; the B scale is NaN)."""
    n_rows, n_cols, nbc, bp, bc, bd, B = _spmm_case(SPMM_CASES[0], seed=5)
    data, indices, indptr = _dev_bsr(bp, bc, bd, device, torch.float32)
    codes, scales = FB.quantize_bsr_fp8(data)
    Bt = S.dense(B, device).clone()
    Bt[1, 2] = bad
    _, absmax = FB.pack_dense_operand(Bt, torch.uint8)
    got = float(absmax.item())
    assert (np.isnan(got) if np.isnan(bad) else got == float("inf"))
    C = FB.flagsparse_spmm_bsr(codes, indices, indptr, Bt, (n_rows, n_cols), scales=scales)
    assert torch.isnan(C).all()
    # a finite B through the same call is finite
    C_ok = FB.flagsparse_spmm_bsr(codes, indices, indptr, S.dense(B, device), (n_rows, n_cols), scales=scales)
    assert torch.isfinite(C_ok).all()


def test_non_finite_block_gets_nan_scale(device):
    n_rows, n_cols, nbc, bp, bc, bd, B = _spmm_case(SPMM_CASES[0], seed=6)
    data, indices, indptr = _dev_bsr(bp, bc, bd, device, torch.float32)
    data[1, 3, 4] = float("inf")
    codes, scales = FB.quantize_bsr_fp8(data)
    assert torch.isnan(scales[1]) and torch.isfinite(scales[0]) and torch.isfinite(scales[2:]).all()
    C = FB.flagsparse_spmm_bsr(codes, indices, indptr, S.dense(B, device), (n_rows, n_cols), scales=scales)
    assert not torch.isfinite(C).all()


def test_bsr_spmm_validation(device):
    n_rows, n_cols, nbc, bp, bc, bd, B = _spmm_case(SPMM_CASES[0], seed=1)
    data, indices, indptr = _dev_bsr(bp, bc, bd, device, torch.bfloat16)
    Bt = S.dense(B, device)
    with pytest.raises(ValueError, match="indptr length"):
        FB.flagsparse_spmm_bsr(data, indices, indptr[:-1], Bt, (n_rows, n_cols))
    bad = indices.clone()
    bad[0] = nbc
    with pytest.raises(IndexError, match="n_block_cols"):
        FB.flagsparse_spmm_bsr(data, bad, indptr, Bt, (n_rows, n_cols))
    with pytest.raises(TypeError):
        FB.flagsparse_spmm_bsr(data.float(), indices, indptr, Bt, (n_rows, n_cols))
    with pytest.raises(ValueError, match="scales"):
        FB.flagsparse_spmm_bsr(data.view(torch.uint8)[:, :, :16], indices, indptr, Bt, (n_rows, n_cols))
    with pytest.raises(ValueError, match="B.shape"):
        FB.flagsparse_spmm_bsr(data, indices, indptr, Bt[1:], (n_rows, n_cols))


def test_bsr_spmm_deterministic(device):
    n_rows, n_cols, nbc, bp, bc, bd, B = _spmm_case(SPMM_CASES[3], seed=9)
    data, indices, indptr = _dev_bsr(bp, bc, bd, device, torch.bfloat16)
    Bt = S.dense(B, device, torch.bfloat16)
    c0 = FB.flagsparse_spmm_bsr(data, indices, indptr, Bt, (n_rows, n_cols))
    for _ in range(3):
        assert torch.equal(FB.flagsparse_spmm_bsr(data, indices, indptr, Bt, (n_rows, n_cols)), c0)


# --------------------------------------- full size: 65536 x 65536, 40 / row
@pytest.fixture(scope="module")
def full_bsr():
    bp, bc, bd = S.bsr_blocks()            # 4096 x 4096 block rows/cols, 40 per row, seed 2
    B = np.random.default_rng(2).standard_normal((65536, 128))
    return bp, bc, bd, B


def test_full_bsr_csr_round_trip(device, full_bsr):
    bp, bc, bd, _ = full_bsr
    ip, ix, v = S.bsr_to_csr(bp, bc, bd)
    assert ix.size == 41_943_040
    data, indices, indptr = S.to_torch(ip, ix, v, device, torch.float32, torch.int32, torch.int64)
    del ip, ix, v
    gd, gc, gp = FB.csr_to_bsr(data, indices, indptr, (65536, 65536))
    np.testing.assert_array_equal(gp.cpu().numpy(), bp)
    np.testing.assert_array_equal(gc.cpu().numpy(), bc)
    assert torch.equal(gd, torch.from_numpy(bd).to(device=device, dtype=torch.float32))


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16, "fp8"])
def test_full_bsr_spmm(device, full_bsr, dtype):
    bp, bc, bd, B = full_bsr
    Bt = S.dense(B, device)
    if dtype == "fp8":
        data, indices, indptr = _dev_bsr(bp, bc, bd, device, torch.float32)
        codes, scales = FB.quantize_bsr_fp8(data)
        C = FB.flagsparse_spmm_bsr(codes, indices, indptr, Bt, (65536, 65536), scales=scales)
        a_codes, a_scales = O.fp8_block_quantize_f32(bd[: bp[64]])
        b_codes, b_scale = O.fp8_tensor_quantize_f32(B)
        a_use = O.e4m3_decode(a_codes) * a_scales.astype(np.float64)[:, None, None]
        b_use = O.e4m3_decode(b_codes) * np.float64(b_scale)
    else:
        data, indices, indptr = _dev_bsr(bp, bc, bd, device, dtype)
        C = FB.flagsparse_spmm_bsr(data, indices, indptr, Bt, (65536, 65536))
        a_use, b_use = np_dtype_round(bd[: bp[64]], dtype), np_dtype_round(B, dtype)
    # first 64 block rows (1024 rows) against the oracle
    sample = C[:1024]
    want = O.bsr_spmm(bp[:65], bc[: bp[64]], a_use, b_use)
    assert rel_err(sample, want) <= (FP8_VS_DEQUANT_TOL if dtype == "fp8" else 1e-5)
    exact = O.bsr_spmm(bp[:65], bc[: bp[64]], bd[: bp[64]], B)
    err64 = rel_err(sample, exact)
    print(f"[{dtype}] rel err vs FP64 (1024 rows): {err64:.3e}")
    assert err64 <= (0.1 if dtype == "fp8" else REL_TOL[dtype])
    # whole matrix against the FP32 CSR plain form of the same matrix
    ip, ix, v = S.bsr_to_csr(bp, bc, bd)
    cd, ci, cp = S.to_torch(ip, ix, v, device, torch.float32, torch.int32, torch.int64)
    del ip, ix, v
    ref32 = flagsparse_spmm_csr(cd, ci, cp, Bt, (65536, 65536))
    num = float((C - ref32).abs().max())
    den = float(ref32.abs().max())
    print(f"[{dtype}] rel err vs FP32 CSR (all rows): {num / den:.3e}")
    assert num / den <= (0.1 if dtype == "fp8" else REL_TOL[dtype])
