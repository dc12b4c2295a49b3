"""Pins the C oracle (CPU only).

Each oracle function is checked against scipy.sparse (an independent
implementation), dense numpy brute force on small inputs and closed-form
invariants, so that a dropped term, a wrong index or a transposed operand in
the oracle fails here.  The reference ships no golden vectors for this path.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from flagsparse_amd import synthetic as S
from oracle import oracle as O


def _mat(m, n, density, seed):
    ip, ix, v = S.random_csr(m, n, density, seed)
    return ip, ix, v, sp.csr_matrix((v, ix, ip), shape=(m, n))


SHAPES = [(1, 32), (160, 1024), (128, 256), (8, 4), (64, 16), (37, 53), (0, 5), (5, 0)]


@pytest.mark.parametrize("m,n", SHAPES)
def test_spmv_matches_scipy_and_dense(m, n):
    ip, ix, v, A = _mat(m, n, 0.1, seed=m * 131 + n)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(n)
    xt = rng.standard_normal(m)
    np.testing.assert_allclose(O.spmv(ip, ix, v, x, (m, n)), A @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(O.spmv(ip, ix, v, xt, (m, n), "trans"), A.T @ xt, rtol=1e-13, atol=1e-13)
    dense = A.toarray()
    np.testing.assert_allclose(O.spmv(ip, ix, v, x, (m, n)), dense @ x, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("m,n,k", [(8, 4, 16), (64, 16, 96), (33, 70, 5), (1, 1, 1), (10, 12, 0)])
def test_spmm_matches_scipy_and_dense(m, n, k):
    ip, ix, v, A = _mat(m, n, 0.2, seed=m + 7 * n + k)
    rng = np.random.default_rng(2)
    B = rng.standard_normal((n, k))
    Bt = rng.standard_normal((m, k))
    np.testing.assert_allclose(O.spmm(ip, ix, v, B, (m, n)), A @ B, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(O.spmm(ip, ix, v, Bt, (m, n), "trans"), A.T @ Bt, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(O.spmm(ip, ix, v, B, (m, n)), A.toarray() @ B, rtol=1e-13, atol=1e-13)


def test_spmm_columns_equal_spmv():
    ip, ix, v, A = _mat(50, 40, 0.15, seed=3)
    B = np.random.default_rng(3).standard_normal((40, 6))
    C = O.spmm(ip, ix, v, B, (50, 40))
    for j in range(6):
        np.testing.assert_allclose(C[:, j], O.spmv(ip, ix, v, B[:, j], (50, 40)), rtol=0, atol=0)


def test_spmm_is_linear_in_values_and_dense():
    ip, ix, v, _ = _mat(30, 25, 0.2, seed=4)
    rng = np.random.default_rng(4)
    B1, B2 = rng.standard_normal((25, 3)), rng.standard_normal((25, 3))
    v2 = rng.standard_normal(v.size)
    lhs = O.spmm(ip, ix, 2.0 * v + v2, B1 - 3.0 * B2, (30, 25))
    rhs = (2 * O.spmm(ip, ix, v, B1, (30, 25)) - 6 * O.spmm(ip, ix, v, B2, (30, 25))
           + O.spmm(ip, ix, v2, B1, (30, 25)) - 3 * O.spmm(ip, ix, v2, B2, (30, 25)))
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12, atol=1e-12)


def test_transpose_identity_xT_A_y():
    # <u, A w> == <A^T u, w> pins the transposed operand.
    ip, ix, v, _ = _mat(45, 31, 0.2, seed=5)
    rng = np.random.default_rng(5)
    u, w = rng.standard_normal(45), rng.standard_normal(31)
    lhs = u @ O.spmv(ip, ix, v, w, (45, 31))
    rhs = O.spmv(ip, ix, v, u, (45, 31), "trans") @ w
    assert abs(lhs - rhs) < 1e-12


@pytest.mark.parametrize("m,n,k", [(8, 10, 8), (64, 96, 16), (20, 7, 1), (5, 9, 0)])
def test_sddmm_matches_dense_closed_form(m, n, k):
    ip, ix, v, A = _mat(m, n, 0.25, seed=m + n + k)
    rng = np.random.default_rng(6)
    X, Y = rng.standard_normal((m, k)), rng.standard_normal((n, k))
    full = X @ Y.T
    rows = np.repeat(np.arange(m), np.diff(ip))
    want = 1.25 * full[rows, ix] + 0.5 * v
    np.testing.assert_allclose(O.sddmm(ip, ix, X, Y, v, 1.25, 0.5), want, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(O.sddmm(ip, ix, X, Y), full[rows, ix], rtol=1e-12, atol=1e-12)


def test_sddmm_is_the_spmm_value_gradient():
    # d/dvalues of sum(G * (A B)) equals SDDMM(G, B): finite differences.
    ip, ix, v, _ = _mat(12, 9, 0.3, seed=7)
    rng = np.random.default_rng(7)
    B, G = rng.standard_normal((9, 4)), rng.standard_normal((12, 4))
    grad = O.sddmm(ip, ix, G, B)
    eps = 1e-6
    for p in range(v.size):
        vp, vm = v.copy(), v.copy()
        vp[p] += eps
        vm[p] -= eps
        fd = (np.sum(G * O.spmm(ip, ix, vp, B, (12, 9))) - np.sum(G * O.spmm(ip, ix, vm, B, (12, 9)))) / (2 * eps)
        assert abs(fd - grad[p]) < 1e-6


@pytest.mark.parametrize("m,n,density", [(40, 33, 0.2), (1, 100, 0.5), (100, 1, 0.5), (17, 64, 0.0), (64, 64, 0.9)])
def test_csc_matches_scipy_tocsc_bit_exact(m, n, density):
    ip, ix, v, A = _mat(m, n, density, seed=m * n)
    colptr, rowind, perm = O.csr_to_csc(ip, ix, (m, n))
    At = A.tocsc()
    At.has_sorted_indices  # scipy keeps rows ascending inside columns for canonical CSR
    np.testing.assert_array_equal(colptr, At.indptr)
    np.testing.assert_array_equal(rowind, At.indices)
    np.testing.assert_array_equal(v[perm], At.data)
    assert np.array_equal(np.sort(perm), np.arange(ix.size))  # bijection
    # stable: inside each column, CSR positions ascend
    for c in range(n):
        seg = perm[colptr[c]: colptr[c + 1]]
        assert np.all(np.diff(seg) > 0)


def test_csc_is_stable_argsort_by_column():
    ip, ix, v, _ = _mat(70, 20, 0.3, seed=9)
    _, _, perm = O.csr_to_csc(ip, ix, (70, 20))
    np.testing.assert_array_equal(perm, np.argsort(ix, kind="stable"))


def test_relative_error_metric():
    assert O.max_relative_error([1.0, 2.0], [1.0, 2.0]) == 0.0
    assert O.max_relative_error([1.0, 2.5], [1.0, 2.0]) == pytest.approx(0.25)
    assert O.max_relative_error([], []) == 0.0
    assert O.max_relative_error([0.0], [0.0]) == 0.0
    assert O.max_relative_error([1.0], [0.0]) == float("inf")


def test_planted_mistakes_are_caught():
    # A plausible wrong restatement (column index used as row) must fail the
    # scipy check that pins the real oracle.
    ip, ix, v, A = _mat(30, 30, 0.2, seed=10)
    x = np.random.default_rng(10).standard_normal(30)
    rows = np.repeat(np.arange(30), np.diff(ip))
    wrong = np.zeros(30)
    np.add.at(wrong, ix, v * x[rows])  # transposed operand
    assert not np.allclose(wrong, A @ x)
    dropped = O.spmv(ip, ix, np.where(np.arange(v.size) == 0, 0.0, v), x, (30, 30))  # dropped term
    assert not np.allclose(dropped, A @ x)


@pytest.mark.parametrize("m,n,density", [(40, 33, 0.2), (64, 64, 0.05), (1, 100, 0.5), (100, 17, 0.3), (7, 9, 0.0)])
def test_csr_to_bsr_matches_scipy_tobsr(m, n, density):
    ip, ix, v, A = _mat(m, n, density, seed=m + n)
    bp, bc, bd = O.csr_to_bsr(ip, ix, v, (m, n), bs=16)
    mp, np_ = -(-m // 16) * 16, -(-n // 16) * 16
    Ap = sp.csr_matrix((v, ix, np.concatenate([ip, np.full(mp - m, ip[-1])])), shape=(mp, np_))
    S = Ap.tobsr(blocksize=(16, 16))
    S.sort_indices()
    # scipy keeps explicit blocks only where entries exist, like the oracle
    np.testing.assert_array_equal(bp, S.indptr)
    np.testing.assert_array_equal(bc, S.indices)
    np.testing.assert_array_equal(bd, S.data)


def test_bsr_spmm_matches_dense():
    from flagsparse_amd import synthetic as S

    bp, bc, bd = S.bsr_blocks(6, 9, 3, seed=1)
    rng = np.random.default_rng(1)
    B = rng.standard_normal((9 * 16, 5))
    dense = np.zeros((6 * 16, 9 * 16))
    for br in range(6):
        for j in range(bp[br], bp[br + 1]):
            dense[br * 16:(br + 1) * 16, bc[j] * 16:(bc[j] + 1) * 16] = bd[j]
    np.testing.assert_allclose(O.bsr_spmm(bp, bc, bd, B), dense @ B, rtol=1e-12, atol=1e-12)
    ip, ix, v = S.bsr_to_csr(bp, bc, bd)
    np.testing.assert_allclose(O.spmm(ip, ix, v, B, (96, 144)), dense @ B, rtol=1e-12, atol=1e-12)


def test_e4m3_codec_matches_torch_and_spec():
    import torch

    x = np.concatenate([np.random.default_rng(0).standard_normal(20000) * 40, np.linspace(-448, 448, 4001),
                        2.0 ** -np.arange(0, 12), [0.0, -0.0, 2 ** -9, 2 ** -10, 1.5 * 2 ** -9, 448.0, 464.0 - 1e-9]])
    codes = O.e4m3_encode(x)
    ref = torch.tensor(x, dtype=torch.float64).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    np.testing.assert_array_equal(codes, ref)
    # spec anchors: 448 = 0x7E, smallest subnormal 2^-9 = 0x01, 1.0 = 0x38
    assert O.e4m3_encode([448.0, 2 ** -9, 1.0]).tolist() == [0x7E, 0x01, 0x38]
    dec = O.e4m3_decode(np.arange(256, dtype=np.uint8))
    assert np.isnan(dec[0x7F]) and np.isnan(dec[0xFF]) and dec[0x7E] == 448.0
    finite = np.isfinite(dec)
    # every finite code round-trips (including -0.0 -> 0x80)
    np.testing.assert_array_equal(O.e4m3_encode(dec[finite]), np.arange(256, dtype=np.uint8)[finite])


def test_fp8_block_quantize_scale_rule():
    bd = np.random.default_rng(2).standard_normal((5, 16, 16))
    bd[3] = 0.0
    codes, scale = O.fp8_block_quantize(bd)
    np.testing.assert_allclose(scale[[0, 1, 2, 4]], np.abs(bd[[0, 1, 2, 4]]).reshape(4, -1).max(1) / 448.0)
    assert scale[3] == 1.0 and not codes[3].any()
    deq = O.e4m3_decode(codes) * scale[:, None, None]
    # e4m3 keeps 3 mantissa bits: relative error <= 2^-4 for normals
    big = np.abs(bd) > 2 ** -6 * scale[:, None, None]
    assert np.all(np.abs(deq - bd)[big] <= 2 ** -4 * np.abs(bd)[big] + 1e-15)


def test_bsr_to_csr_expansion_matches_scipy():
    bp, bc, bd = S.bsr_blocks(7, 9, np.array([0, 3, 1, 9, 2, 0, 4]), seed=3)
    A = sp.bsr_matrix((bd, bc, bp), shape=(112, 144)).toarray()
    ip, ix, v = S.bsr_to_csr(bp, bc, bd)
    C = sp.csr_matrix((v, ix, ip), shape=(112, 144))
    assert C.has_sorted_indices and C.nnz == bd.size
    np.testing.assert_array_equal(C.toarray(), A)
    ip, ix, v = S.bsr_to_csr(bp, bc, bd, shape=(100, 130))
    np.testing.assert_array_equal(sp.csr_matrix((v, ix, ip), shape=(100, 130)).toarray(), A[:100, :130])
    # and back: the oracle's CSR -> BSR of the expansion is the identity
    rp, rc, rd = O.csr_to_bsr(*S.bsr_to_csr(bp, bc, bd), (112, 144))
    np.testing.assert_array_equal(rp, bp)
    np.testing.assert_array_equal(rc, bc)
    np.testing.assert_array_equal(rd, bd)


def test_fp8_float32_quantizers():
    rng = np.random.default_rng(3)
    bd = rng.standard_normal((9, 16, 16)).astype(np.float32)
    bd[2] = 0
    c32, s32 = O.fp8_block_quantize_f32(bd)
    c64, s64 = O.fp8_block_quantize(bd)
    assert s32.dtype == np.float32
    np.testing.assert_allclose(s32, s64, rtol=2 ** -23)
    # the two roundings differ at most by one code step, on rare ties
    assert np.mean(c32 == c64) > 0.99
    deq = O.e4m3_decode(c32) * s32[:, None, None]
    big = np.abs(bd) > 2 ** -6 * s32[:, None, None]
    assert np.all(np.abs(deq - bd)[big] <= 2 ** -4 * np.abs(bd)[big] * (1 + 1e-6))
    B = rng.standard_normal((30, 7)) * 5
    cb, sb = O.fp8_tensor_quantize_f32(B)
    assert sb == np.float32(np.abs(B.astype(np.float32)).max()) / np.float32(448)
    assert np.abs(O.e4m3_decode(cb)).max() == 448.0


def test_coo_to_csr_matches_scipy():
    rng = np.random.default_rng(12)
    m, n, k = 30, 17, 400
    row, col = rng.integers(0, m, k), rng.integers(0, n, k)
    row[:50], col[:50] = row[-50:], col[-50:]  # duplicates
    v = rng.standard_normal(k)
    ip, ix, dv = O.coo_to_csr(row, col, v, (m, n))
    mine = sp.csr_matrix((dv, ix, ip), shape=(m, n))
    ref = sp.coo_matrix((v, (row, col)), shape=(m, n)).tocsr()  # sums duplicates
    np.testing.assert_allclose(mine.toarray(), ref.toarray(), rtol=1e-13, atol=1e-13)
    keys = ip.searchsorted(np.arange(ix.size), side="right") - 1  # row of each sorted nonzero
    assert np.all(np.diff(keys * n + ix) >= 0)  # ordered by (row, col)
    # stable: equal keys keep input order
    order = np.lexsort((col, row))
    np.testing.assert_array_equal(dv, v[order])
