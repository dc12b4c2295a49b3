"""SpMV parity on the GPU: HIP kernels vs the C oracle (FP64 golden).

Mirrors the reference's tests/pytest/test_spmv_csr_accuracy.py grid
(shapes (1,32), (160,1024), (128,256); int32/int64; non/trans/conj;
prepared mismatch errors; int64 -> int32 fallback policy) with fixed seeds,
plus the 16384 x 16384 x 16 workload, long rows (segment split), empty
rows/matrices and bitwise determinism.  Bound: max|d|/max|ref| <= 1e-5 (f32),
1e-12 (f64), 1e-2 (bf16), 5e-3 (f16).
"""

import numpy as np
import pytest
import torch

import flagsparse_amd as fs
from flagsparse_amd import spmv_csr as spmv_mod
from flagsparse_amd import synthetic
from oracle import oracle

from ._helpers import assert_rel, device_csr, long_row_matrix, np_dtype_round

pytestmark = pytest.mark.gpu

SHAPES = [(1, 32), (160, 1024), (128, 256)]


@pytest.mark.parametrize("M,N", SHAPES)
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["float32", "float64"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["int32", "int64"])
@pytest.mark.parametrize("op", ["non", "trans", "conj"])
def test_spmv_matches_oracle(device, M, N, dtype, index_dtype, op):
    ip, ix, v = synthetic.random_csr(M, N, min(0.25, max(0.06, 32.0 / (M * N))), seed=M * 1000 + N)
    data, idx, ptr = device_csr(ip, ix, v, device, dtype, index_dtype)
    x_len = M if op != "non" else N
    x = np.random.default_rng(M + N).standard_normal(x_len)
    xt = torch.tensor(x, dtype=dtype, device=device)
    y = fs.flagsparse_spmv_csr(data, idx, ptr, xt, shape=(M, N), op=op, index_fallback_policy="auto")
    assert y.dtype == dtype and y.shape == (N if op != "non" else M,)
    ref = oracle.spmv(ip, ix, np_dtype_round(v, dtype), np_dtype_round(x, dtype), (M, N), op)
    assert_rel(y, ref, dtype, f"spmv {op}")


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16], ids=["bfloat16", "float16"])
@pytest.mark.parametrize("op", ["non", "trans"])
def test_spmv_half_storage(device, dtype, op):
    ip, ix, v = synthetic.random_csr(300, 200, 0.1, seed=7)
    data, idx, ptr = device_csr(ip, ix, v, device, dtype)
    x = np.random.default_rng(7).standard_normal(300 if op == "trans" else 200)
    y = fs.flagsparse_spmv_csr(data, idx, ptr, torch.tensor(x, dtype=dtype, device=device), shape=(300, 200), op=op)
    ref = oracle.spmv(ip, ix, np_dtype_round(v, dtype), np_dtype_round(x, dtype), (300, 200), op)
    assert_rel(y, ref, dtype, "spmv half")


def test_spmv_uniform_16384x16_full_size(device):
    ip, ix, v, x = synthetic.uniform_rows(16384, 16384, 16, seed=0)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    prepared = fs.prepare_spmv_csr(data, idx, ptr, (16384, 16384))
    y = fs.flagsparse_spmv_csr(x=torch.tensor(x, dtype=torch.float32, device=device), prepared=prepared)
    ref = oracle.spmv(ip, ix, np_dtype_round(v, torch.float32), np_dtype_round(x, torch.float32), (16384, 16384))
    err = assert_rel(y, ref, torch.float32, "uniform 16x")
    # f64 accumulation + one rounding: every entry equals the correctly rounded oracle or is 1 ulp off
    y64 = y.double().cpu().numpy()
    assert np.max(np.abs(y64 - ref.astype(np.float32).astype(np.float64))) <= 1.2e-7 * max(1.0, np.abs(ref).max())
    assert err < 1e-6


@pytest.mark.parametrize("long_len", [4097, 9000, 40000])
@pytest.mark.parametrize("op", ["non", "trans"])
def test_spmv_long_rows_split_into_segments(device, long_len, op):
    n_cols = 50000
    ip, ix, v = long_row_matrix(64, n_cols, [0, 5, 63], long_len, 9, seed=long_len)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float64, torch.int64)
    prepared = fs.prepare_spmv_csr(data, idx, ptr, (64, n_cols), op=op)
    x = np.random.default_rng(1).standard_normal(64 if op == "trans" else n_cols)
    y = fs.flagsparse_spmv_csr(x=torch.tensor(x, device=device), prepared=prepared, op=op)
    ref = oracle.spmv(ip, ix, v, x, (64, n_cols), op)
    assert_rel(y, ref, torch.float64, "long rows")
    if op == "non":
        assert prepared.plan().n_long_rows == 3 and prepared.plan().n_segments >= 3


def test_spmv_all_row_bins_and_empty_rows(device):
    # lengths 0..300 cover every lane bin; zero-length rows must produce 0
    rng = np.random.Generator(np.random.PCG64(5))
    lengths = np.concatenate([np.zeros(50, np.int64), rng.integers(0, 300, 2000)])
    rng.shuffle(lengths)
    ip, ix = synthetic.distinct_sorted_columns(rng, lengths, 1000)
    v = rng.uniform(-1, 1, ix.size)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    x = rng.uniform(-1, 1, 1000)
    y = fs.flagsparse_spmv_csr(data, idx, ptr, torch.tensor(x, dtype=torch.float32, device=device),
                               shape=(lengths.size, 1000))
    ref = oracle.spmv(ip, ix, np_dtype_round(v, torch.float32), np_dtype_round(x, torch.float32),
                      (lengths.size, 1000))
    assert_rel(y, ref, torch.float32, "bins")
    assert torch.all(y[torch.from_numpy(lengths == 0).to(device)] == 0)


def test_spmv_empty_matrix_and_zero_rows(device):
    ptr = torch.zeros(5, dtype=torch.int32, device=device)
    idx = torch.empty(0, dtype=torch.int32, device=device)
    data = torch.empty(0, dtype=torch.float32, device=device)
    y = fs.flagsparse_spmv_csr(data, idx, ptr, torch.ones(7, device=device), shape=(4, 7))
    assert torch.equal(y, torch.zeros(4, device=device))
    y = fs.flagsparse_spmv_csr(data, idx, ptr, torch.ones(4, device=device), shape=(4, 7), op="trans")
    assert torch.equal(y, torch.zeros(7, device=device))
    y = fs.flagsparse_spmv_csr(data, idx, torch.zeros(1, dtype=torch.int32, device=device),
                               torch.ones(3, device=device), shape=(0, 3))
    assert y.shape == (0,)


def test_spmv_deterministic_bitwise(device):
    ip, ix, v, _ = synthetic.power_law_rows(20000, 20000, 2.1, 24, 4096, n_dense=0, seed=3)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    prepared = fs.prepare_spmv_csr(data, idx, ptr, (20000, 20000))
    x = torch.randn(20000, device=device)
    y1 = fs.flagsparse_spmv_csr(x=x, prepared=prepared)
    y2 = fs.flagsparse_spmv_csr(x=x, prepared=prepared)
    assert torch.equal(y1, y2)


def test_spmv_out_argument(device):
    ip, ix, v = synthetic.random_csr(40, 30, 0.2, seed=2)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    x = torch.randn(30, device=device)
    out = torch.empty(40, device=device)
    y = fs.flagsparse_spmv_csr(data, idx, ptr, x, (40, 30), out=out)
    assert y is out
    with pytest.raises(ValueError, match="out shape/dtype must match result"):
        fs.flagsparse_spmv_csr(data, idx, ptr, x, (40, 30), out=torch.empty(41, device=device))


def test_spmv_return_time_and_meta(device):
    ip, ix, v = synthetic.random_csr(40, 30, 0.2, seed=2)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    x = torch.randn(30, device=device)
    y, ms, meta = fs.flagsparse_spmv_csr(data, idx, ptr, x, (40, 30), return_time=True, return_meta=True,
                                         use_opt=True)
    assert ms == pytest.approx(meta["op_total_ms"]) and meta["bucket_count"] >= 1
    assert set(meta) == {"symbolic_ms", "compute_ms", "op_total_ms", "bucket_count"}


def test_spmv_prepared_transpose_mismatch_rejected(device):
    ip, ix, v = synthetic.random_csr(8, 10, 0.3, seed=1)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    prepared = spmv_mod.prepare_spmv_csr(data, idx, ptr, (8, 10), transpose=True)
    x = torch.randn(8, device=device)
    with pytest.raises(ValueError, match="does not match prepared.transpose"):
        fs.flagsparse_spmv_csr(x=x, prepared=prepared, transpose=False)


def test_spmv_prepared_op_mismatch_rejected(device):
    ip, ix, v = synthetic.random_csr(8, 10, 0.3, seed=1)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    prepared = spmv_mod.prepare_spmv_csr(data, idx, ptr, (8, 10), op="conj")
    with pytest.raises(ValueError, match="does not match prepared.op"):
        fs.flagsparse_spmv_csr(x=torch.randn(8, device=device), prepared=prepared, op="trans")


def test_spmv_validation_errors_on_device(device):
    ip, ix, v = synthetic.random_csr(8, 10, 0.3, seed=1)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32)
    with pytest.raises(ValueError, match="indptr must start at zero"):
        fs.flagsparse_spmv_csr(data, idx, ptr + 1, torch.randn(10, device=device), (8, 10))
    bad = idx.clone()
    bad[0] = 10
    with pytest.raises(IndexError, match="indices out of range for n_cols"):
        fs.flagsparse_spmv_csr(data, bad, ptr, torch.randn(10, device=device), (8, 10))
    with pytest.raises(TypeError, match="x dtype must match"):
        fs.flagsparse_spmv_csr(data, idx, ptr, torch.randn(10, device=device, dtype=torch.float64), (8, 10))
    with pytest.raises(ValueError, match="x length must be n_cols=10, got 9"):
        fs.flagsparse_spmv_csr(data, idx, ptr, torch.randn(9, device=device), (8, 10))


def test_spmv_int64_auto_fallback_to_int32(device, monkeypatch):
    ip, ix, v = synthetic.random_csr(12, 9, 0.3, seed=4)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32, torch.int64)
    x = np.random.default_rng(4).standard_normal(9)
    state = {"forced_once": False}
    original = spmv_mod._hip_spmv_csr_impl_prepared

    def fail_int64_once(prepared, x_in, out=None):
        if prepared.kernel_indices.dtype == torch.int64 and not state["forced_once"]:
            state["forced_once"] = True
            raise RuntimeError("forced int64 launch failure")
        return original(prepared, x_in, out)

    monkeypatch.setattr(spmv_mod, "_hip_spmv_csr_impl_prepared", fail_int64_once)
    y = fs.flagsparse_spmv_csr(data, idx, ptr, torch.tensor(x, dtype=torch.float32, device=device), shape=(12, 9),
                               index_fallback_policy="auto")
    assert state["forced_once"]
    assert_rel(y, oracle.spmv(ip, ix, np_dtype_round(v, torch.float32), np_dtype_round(x, torch.float32), (12, 9)),
               torch.float32)


def test_spmv_int64_strict_no_fallback(device, monkeypatch):
    ip, ix, v = synthetic.random_csr(12, 9, 0.3, seed=4)
    data, idx, ptr = device_csr(ip, ix, v, device, torch.float32, torch.int64)

    def fail_int64(prepared, x_in, out=None):
        raise RuntimeError("forced int64 launch failure")

    monkeypatch.setattr(spmv_mod, "_hip_spmv_csr_impl_prepared", fail_int64)
    with pytest.raises(RuntimeError, match="forced int64 launch failure"):
        fs.flagsparse_spmv_csr(data, idx, ptr, torch.randn(9, device=device), shape=(12, 9),
                               index_fallback_policy="strict")


def test_spmv_int64_auto_does_not_fallback_when_index_exceeds_int32(device, monkeypatch):
    limit = spmv_mod._INDEX_LIMIT_INT32
    data = torch.ones(1, dtype=torch.float32, device=device)
    prepared = spmv_mod.PreparedCsrSpmv(
        data=data,
        kernel_indices=torch.tensor([limit + 1], dtype=torch.int64, device=device),
        kernel_indptr=torch.tensor([0, 1], dtype=torch.int64, device=device),
        shape=(1, limit + 2),
        n_rows=1,
        n_cols=limit + 2,
        block_nnz=256,
        max_segments=1,
        max_row_nnz=1,
        opt_buckets=[],
        transpose=False,
        index_fallback_policy="auto",
    )

    def fail_launch(_prepared, _x, use_opt=False, opt_buckets=None):
        raise RuntimeError("forced native int64 failure")

    monkeypatch.setattr(spmv_mod, "_run_spmv_prepared", fail_launch)
    x = torch.empty(0, dtype=torch.float32, device=device)
    with pytest.raises(RuntimeError, match="int32 fallback is unsafe"):
        spmv_mod._run_spmv_prepared_with_fallback(prepared, x, use_opt=False)


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("max_blocks,xcd", [(0, 0), (512, 0), (0, 8), (1024, 8)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_spmv_loop_variants_and_grids_bitwise_equal(device, variant, max_blocks, xcd, dtype):
    """The pipelined (0) and plain (1) SpMV loops, one-tile-per-workgroup,
    persistent and XCD-region grids all reduce every row in the same order."""
    from flagsparse_amd import _plan
    from flagsparse_amd._matrix import CsrHandle

    ip, ix, v = long_row_matrix(5000, 3000, long_rows=(7, 4000), long_len=2900, short_len=9, seed=21)
    pl_ip, pl_ix, pl_v, _ = synthetic.power_law_rows(20000, 20000, 2.1, 24, 4096, n_dense=0, seed=5)
    for ip_, ix_, v_, shape in ((ip, ix, v, (5000, 3000)), (pl_ip, pl_ix, pl_v, (20000, 20000))):
        data, idx, ptr = device_csr(ip_, ix_, v_, device, dtype)
        h = CsrHandle(idx, ptr, shape)
        x = np.random.default_rng(3).uniform(-1, 1, shape[1])
        xt = synthetic.dense(x, device, dtype)
        cfg = _plan.spmv_config(unroll=4)
        ref = h.spmv(data, xt, config=cfg)
        got = h.spmv(data, xt, config=cfg, max_blocks=max_blocks, xcd_groups=xcd, variant=variant)
        assert torch.equal(got, ref)
        gold = oracle.spmv(ip_, ix_, np_dtype_round(v_, dtype), np_dtype_round(x, dtype), shape)
        assert_rel(got, gold, dtype, "spmv variant")


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.bfloat16, torch.float16],
                         ids=["f32", "f64", "bf16", "f16"])
@pytest.mark.parametrize("unroll", [2, 4, 8])
def test_spmv_x_in_lds_variant_bitwise_equal_to_plain_loop(device, dtype, unroll):
    """Variants 2 and 3 (x staged in LDS; plain and pipelined loops) run the
    same arithmetic per lane as the global-x loops: equal to variant 1 on every
    grid, for CSR and CSC (trans) launches, with long-row segments, and within
    tolerance of the oracle."""
    from flagsparse_amd import _plan
    from flagsparse_amd._matrix import CsrHandle

    ip, ix, v = long_row_matrix(5000, 3000, long_rows=(7, 4000), long_len=2900, short_len=9, seed=21)
    u_ip, u_ix, u_v, _ = synthetic.uniform_rows(4096, 8192 if dtype == torch.float64 else 16384, 16, seed=0)
    cases = [(ip, ix, v, (5000, 3000)), (u_ip, u_ix, u_v, (4096, 8192 if dtype == torch.float64 else 16384))]
    cfg = _plan.spmv_config(unroll=unroll)
    for ip_, ix_, v_, shape in cases:
        data, idx, ptr = device_csr(ip_, ix_, v_, device, dtype)
        h = CsrHandle(idx, ptr, shape)
        for trans in (False, True):
            n_x = shape[0] if trans else shape[1]
            if n_x * data.element_size() > 65536:
                continue
            x = np.random.default_rng(3).uniform(-1, 1, n_x)
            xt = synthetic.dense(x, device, dtype)
            ref = h.spmv(data, xt, transpose=trans, config=cfg, variant=1)
            for variant in (2, 3):
                for max_blocks, xcd in ((0, 0), (256, 0), (37, 8), (0, 8)):
                    got = h.spmv(data, xt, transpose=trans, config=cfg, max_blocks=max_blocks, xcd_groups=xcd,
                                 variant=variant)
                    assert torch.equal(got, ref), (shape, trans, variant, max_blocks, xcd)
            gold = oracle.spmv(ip_, ix_, np_dtype_round(v_, dtype), np_dtype_round(x, dtype), shape,
                               "trans" if trans else "non")
            assert_rel(ref, gold, dtype, f"ldsx {shape} trans={trans}")


def test_spmv_x_in_lds_rejects_large_x_and_default_route(device):
    from flagsparse_amd._matrix import CsrHandle

    ip, ix, v = synthetic.random_csr(64, 20000, 0.01, seed=4)
    data, idx, ptr = device_csr(ip, ix, v, device)
    h = CsrHandle(idx, ptr, (64, 20000))
    for variant in (2, 3):
        with pytest.raises(RuntimeError, match="LDS"):
            h.spmv(data, torch.ones(20000, device=device), variant=variant)
    # the public entry point takes the LDS loop for a large matrix with a small x
    ip, ix, v, x = synthetic.uniform_rows(65536, 16384, 16, seed=0)
    data, idx, ptr = device_csr(ip, ix, v, device)
    from flagsparse_amd import _plan

    assert _plan.spmv_default_variant(16384, 4, int(ix.size)) == 2
    y = fs.flagsparse_spmv_csr(data, idx, ptr, synthetic.dense(x, device), (65536, 16384))
    assert_rel(y, oracle.spmv(ip, ix, v, x, (65536, 16384)), torch.float32, "default ldsx")


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.bfloat16, torch.float16],
                         ids=["f32", "f64", "bf16", "f16"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["i32", "i64"])
@pytest.mark.parametrize("unroll", [4, 8])
def test_spmv_vector_chunk_variant(device, dtype, index_dtype, unroll):
    """Variant 4 (4-element chunks, x in LDS): rows starting anywhere, the
    last chunk at the arrays' end, long-row segments, CSR and CSC (through
    the permutation) launches, every grid; within tolerance of the oracle,
    close to the plain loop and deterministic."""
    from flagsparse_amd import _plan
    from flagsparse_amd._matrix import CsrHandle

    n_cols = 8192 if dtype == torch.float64 else 16384
    ip, ix, v = long_row_matrix(5001, 3000, long_rows=(7, 4000), long_len=2900, short_len=9, seed=21)
    r_ip, r_ix, r_v = synthetic.random_csr(4093, n_cols, 0.0017, seed=9)  # ragged rows, nnz % 4 != 0 likely
    cases = [(ip, ix, v, (5001, 3000)), (r_ip, r_ix, r_v, (4093, n_cols))]
    # rows above 1024 nonzeros run as segments (the 2900-long rows here)
    cfg = _plan.spmv_config(unroll=unroll, long_threshold=1024, seg_len=1024)
    for ip_, ix_, v_, shape in cases:
        data, idx, ptr = device_csr(ip_, ix_, v_, device, dtype, index_dtype)
        assert _plan.chunk_aligned(idx, data)
        assert h_segments(idx, ptr, shape, cfg) > 0 or shape[0] == 4093
        h = CsrHandle(idx, ptr, shape)
        for trans in (False, True):
            n_x = shape[0] if trans else shape[1]
            if n_x * data.element_size() > 65536:
                continue
            x = np.random.default_rng(3).uniform(-1, 1, n_x)
            xt = synthetic.dense(x, device, dtype)
            gold = oracle.spmv(ip_, ix_, np_dtype_round(v_, dtype), np_dtype_round(x, dtype), shape,
                               "trans" if trans else "non")
            plain = h.spmv(data, xt, transpose=trans, config=cfg, variant=1)
            first = None
            for max_blocks, xcd in ((0, 0), (256, 0), (37, 8)):
                got = h.spmv(data, xt, transpose=trans, config=cfg, max_blocks=max_blocks, xcd_groups=xcd, variant=4)
                assert_rel(got, gold, dtype, f"vec {shape} trans={trans} grid={max_blocks},{xcd}")
                assert_rel(got, plain.to(torch.float64).cpu().numpy(), dtype, "vec vs plain")
                if first is None:
                    first = got
                else:
                    assert torch.equal(got, first)  # same per-row order on every grid


def h_segments(idx, ptr, shape, cfg):
    from flagsparse_amd._matrix import CsrHandle

    return CsrHandle(idx, ptr, shape).plan(cfg).n_segments


def test_spmv_vector_chunk_variant_rejects_misaligned_and_unroll(device):
    from flagsparse_amd import _plan
    from flagsparse_amd._matrix import CsrHandle

    ip, ix, v = synthetic.random_csr(500, 700, 0.02, seed=5)
    data, idx, ptr = device_csr(ip, ix, v, device)
    h = CsrHandle(idx, ptr, (500, 700))
    x = torch.ones(700, device=device)
    with pytest.raises(RuntimeError, match="unroll 4 or 8"):
        h.spmv(data, x, config=_plan.spmv_config(unroll=2), variant=4)
    # a values view one element in: not on a 4-element boundary
    big = torch.zeros(data.numel() + 1, device=device)
    big[1:] = data
    shifted = big[1:]
    assert not _plan.chunk_aligned(shifted)
    with pytest.raises(RuntimeError, match="aligned"):
        h.spmv(shifted, x, variant=4)
    # the untuned entry point picks the plain LDS loop for such a view
    assert _plan.spmv_default_variant(700, 4, 1 << 20, _plan.chunk_aligned(idx, shifted)) == 2
    assert _plan.spmv_default_variant(700, 4, 1 << 20, _plan.chunk_aligned(idx, data)) == 4


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.bfloat16], ids=["f32", "f64", "bf16"])
@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("per_row,base", [(16, 0), (16, 40), (24, 0), (13, 0), (16, 2)])
def test_spmv_constant_row_length_computes_bounds(device, dtype, variant, per_row, base):
    """Rows that all hold per_row nonzeros: the plan keeps row_len, and the
    vector-chunk loop (variant 4) computes each row's bounds instead of
    loading them; variants 0-3 load them.  Bitwise equal to the same plan
    with row_len cleared (bounds loaded from indptr), on every loop variant,
    also for an indptr that starts at 40 or 2 (the index and value arrays
    then begin that many elements in).  Variant 4 takes its
    software-pipelined loop for 16 and 24 per row at bases 0 and 40 (24: two
    of a row's eight lanes hold no chunk), and its plain computed-bounds loop
    for 13 per row and base 2."""
    import dataclasses

    from flagsparse_amd import _plan
    from flagsparse_amd._matrix import CsrHandle

    n_rows, n_cols = 5003, 4096
    ip, ix, v, x = synthetic.uniform_rows(n_rows, n_cols, per_row, seed=base + variant)
    rng = np.random.default_rng(base)
    ip_b = ip + base
    ix_b = np.concatenate([rng.integers(0, n_cols, base), ix])
    v_b = np.concatenate([rng.uniform(-1, 1, base), v])
    data, idx, ptr = device_csr(ip_b, ix_b, v_b, device, dtype)
    xt = synthetic.dense(x, device, dtype)
    h = CsrHandle(idx, ptr, (n_rows, n_cols))
    cfg = _plan.spmv_config(unroll=4)
    plan = h.plan(cfg)
    assert plan.identity and plan.row_len == per_row and plan.indptr_base == base
    got = h.spmv(data, xt, config=cfg, variant=variant)
    h._plans[cfg] = dataclasses.replace(plan, row_len=0, indptr_base=0)
    loaded = h.spmv(data, xt, config=cfg, variant=variant)
    assert torch.equal(got, loaded)
    gold = oracle.spmv(ip, ix, np_dtype_round(v, dtype), np_dtype_round(x, dtype), (n_rows, n_cols))
    assert_rel(got, gold, dtype, f"row_len variant {variant}")


def test_plan_row_len_only_for_constant_rows(device):
    from flagsparse_amd import _plan

    ip, ix, v, _ = synthetic.uniform_rows(4096, 4096, 16, seed=0)
    d, i, p = device_csr(ip, ix, v, device, torch.float32, torch.int64, torch.int64)
    stats, plan = _plan.inspect_csr(p, i, 4096, ix.size, _plan.spmv_config())
    assert plan.row_len == 16 and plan.indptr_base == 0
    assert plan.struct().row_len == 16
    ip2 = ip.copy()
    ip2[100] += 1  # rows 99 and 100: 17 and 15 nonzeros
    d, i, p = device_csr(ip2, ix, v, device, torch.float32)
    assert _plan.build_plan(p, 4096, _plan.spmv_config()).row_len == 0
    assert _plan.inspect_csr(p, i, 4096, ix.size, _plan.spmv_config())[1].row_len == 0
    # empty rows only: nothing to compute
    p0 = torch.zeros(11, dtype=torch.int32, device=device)
    assert _plan.build_plan(p0, 10, _plan.spmv_config()).row_len == 0
