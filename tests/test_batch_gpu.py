"""GPU parity of the batched SpMV / SpMM (one launch for a packed batch).

* every matrix of a batch against the FP64 oracle (f32 rel 1e-5, f64 1e-12);
* bitwise equality with the per-matrix launch loop (same parameters: every
  row is reduced in the same order);
* persistent grids (capped workgroups, contiguous tile runs) bitwise equal to
  one workgroup per tile;
* HIP graph replays, each checked against the eager result, with eager work
  and host syncs between replays;
* validation of the packed offset table, the autotuner and its cache;
* the full 4096-matrix batch (log-uniform 64..8192 rows, 4..64 nonzeros per
  row) against the per-matrix loop (bitwise) and sampled matrices against the
  oracle.
"""

import numpy as np
import pytest
import torch

from flagsparse_amd import autotune
from flagsparse_amd import synthetic as S
from flagsparse_amd.batch import CsrBatch
from oracle import oracle as O

from ._helpers import REL_TOL, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _no_tune_file(monkeypatch):
    monkeypatch.setenv("FLAGSPARSE_AMD_TUNE_CACHE", "")  # in-memory cache only
    autotune.clear_memory()


def _small_batch(seed, n_mats=23, max_rows=200, max_len=40, long_rows=(), rect=True):
    """Per-matrix numpy CSRs: ragged sizes, empty rows, optional empty
    matrices and rectangular shapes, plus optional long rows."""
    rng = np.random.default_rng(seed)
    long_of = {r % n_mats: L for r, L in long_rows}
    mats = []
    for m in range(n_mats):
        n_rows = int(rng.integers(0, max_rows)) if m % 7 else 0
        if m in long_of:
            n_rows = max(n_rows, 1)
        n_cols = int(rng.integers(1, max_rows + 1)) if rect else max(n_rows, 1)
        lengths = rng.integers(0, min(max_len, n_cols) + 1, n_rows)
        if m in long_of:
            n_cols = max(n_cols, long_of[m])
            lengths[0] = long_of[m]
        ip, ix = S.distinct_sorted_columns(rng, lengths, n_cols)
        v = rng.uniform(-1, 1, ix.size)
        mats.append((ip, ix, v, (n_rows, n_cols)))
    return mats


def _to_batch(mats, device, dtype=torch.float32, index_dtype=torch.int32):
    dev = [(S.dense(v, device, dtype), torch.from_numpy(ix).to(device, index_dtype),
            torch.from_numpy(ip).to(device, index_dtype), shape) for ip, ix, v, shape in mats]
    return CsrBatch.from_matrices(dev)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64])
def test_batch_spmv_matches_oracle_and_loop(device, dtype, index_dtype):
    mats = _small_batch(1)
    b = _to_batch(mats, device, dtype, index_dtype)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, b.total_cols)
    xt = S.dense(x, device, dtype)
    y = b.spmv(xt)
    for m, (ip, ix, v, shape) in enumerate(mats):
        want = O.spmv(ip, ix, v, x[b.col_offsets[m]: b.col_offsets[m + 1]], shape)
        got = y[b.row_offsets[m]: b.row_offsets[m + 1]]
        assert rel_err(got, want) <= REL_TOL[dtype], f"matrix {m}"
    assert torch.equal(b.spmv_per_matrix(xt), y)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("k", [1, 32, 70])
def test_batch_spmm_matches_oracle_and_loop(device, dtype, k):
    mats = _small_batch(3)
    b = _to_batch(mats, device, dtype)
    B = np.random.default_rng(4).standard_normal((b.total_cols, k))
    Bt = S.dense(B, device, dtype)
    C = b.spmm(Bt)
    for m, (ip, ix, v, shape) in enumerate(mats):
        want = O.spmm(ip, ix, v, B[b.col_offsets[m]: b.col_offsets[m + 1]], shape)
        assert rel_err(C[b.row_offsets[m]: b.row_offsets[m + 1]], want) <= REL_TOL[dtype], f"matrix {m}"
    assert torch.equal(b.spmm_per_matrix(Bt), C)


def test_batch_long_rows_split_into_segments(device):
    # rows above the SpMV (4096) and SpMM (256) long thresholds
    mats = _small_batch(5, n_mats=9, long_rows=((1, 9000), (4, 300), (6, 5000)))
    b = _to_batch(mats, device)
    x = np.random.default_rng(6).uniform(-1, 1, b.total_cols)
    B = np.random.default_rng(7).standard_normal((b.total_cols, 16))
    y = b.spmv(S.dense(x, device))
    C = b.spmm(S.dense(B, device))
    assert torch.equal(b.spmv_per_matrix(S.dense(x, device)), y)
    assert torch.equal(b.spmm_per_matrix(S.dense(B, device)), C)
    for m, (ip, ix, v, shape) in enumerate(mats):
        r0, r1 = b.row_offsets[m], b.row_offsets[m + 1]
        c0, c1 = b.col_offsets[m], b.col_offsets[m + 1]
        assert rel_err(y[r0:r1], O.spmv(ip, ix, v, x[c0:c1], shape)) <= 1e-5
        assert rel_err(C[r0:r1], O.spmm(ip, ix, v, B[c0:c1], shape)) <= 1e-5


@pytest.mark.parametrize("bpc,xg", [(1, 0), (2, 0), (8, 0), (0, 8), (1, 8), (8, 8), (8, 3)])
def test_persistent_and_xcd_grids_are_bitwise_equal(device, bpc, xg):
    ip, ix, v, ro, no = S.csr_batch(n_mats=96, rows_lo=16, rows_hi=2048, seed=11, device=device)
    b = CsrBatch(v, ix, ip, ro, None, no)
    x = torch.rand(b.total_cols, device=device) - 0.5
    B = torch.randn(b.total_cols, 32, device=device)
    base_v = autotune.TuneParams(unroll=4, blocks_per_cu=0)
    pers_v = autotune.TuneParams(unroll=4, blocks_per_cu=bpc, xcd_groups=xg)
    assert torch.equal(b.spmv(x, params=base_v), b.spmv(x, params=pers_v))
    base_m = autotune.TuneParams(unroll=4, slot_scale=2, blocks_per_cu=0)
    pers_m = autotune.TuneParams(unroll=4, slot_scale=2, blocks_per_cu=bpc, xcd_groups=xg)
    assert torch.equal(b.spmm(B, params=base_m), b.spmm(B, params=pers_m))


def test_graph_replays_match_eager(device):
    ip, ix, v, ro, no = S.csr_batch(n_mats=64, rows_lo=16, rows_hi=1024, seed=12, device=device)
    b = CsrBatch(v, ix, ip, ro, None, no)
    x = torch.empty(b.total_cols, device=device)
    y = torch.empty(b.total_rows, device=device)
    B = torch.empty(b.total_cols, 32, device=device)
    C = torch.empty(b.total_rows, 32, device=device)
    pv = autotune.TuneParams(unroll=4, blocks_per_cu=4)
    pm = autotune.TuneParams(unroll=4, slot_scale=4, blocks_per_cu=8)
    b._params.clear()
    x.uniform_(-1, 1)
    B.normal_()
    gv = b._capture(lambda: b.spmv(x, out=y, params=pv))
    gm = b._capture(lambda: b.spmm(B, out=C, params=pm))
    for rep in range(3):
        x.uniform_(-1, 1)  # eager work between replays
        B.normal_()
        torch.cuda.synchronize()
        gv.replay()
        gm.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, b.spmv(x, params=pv)), f"spmv replay {rep}"
        assert torch.equal(C, b.spmm(B, params=pm)), f"spmm replay {rep}"


def test_batch_validation(device):
    mats = _small_batch(8, n_mats=4)
    good = _to_batch(mats, device)
    data, indices, indptr = good.data, good.local_indices, good.local_indptr
    ro, co, no = good.row_offsets, good.col_offsets, good.nnz_offsets
    owner = int(np.searchsorted(no, 0, side="right")) - 1  # matrix of nonzero 0
    bad = indices.clone()
    bad[0] = int(co[owner + 1] - co[owner])  # one past its matrix's last column
    with pytest.raises(IndexError, match="out of range"):
        CsrBatch(data, bad, indptr, ro, co, no)
    with pytest.raises(ValueError, match="indptr length"):
        CsrBatch(data, indices, indptr[:-1], ro, co, no)
    wrong = no.copy()
    wrong[1:] += 1
    wrong[-1] = no[-1]
    with pytest.raises(ValueError):
        CsrBatch(data, indices, indptr, ro, co, wrong)
    with pytest.raises(ValueError, match="nondecreasing"):
        CsrBatch(data, indices, indptr, ro[::-1].copy(), co, no)
    with pytest.raises(ValueError, match="x.shape"):
        good.spmv(torch.zeros(good.total_cols + 1, device=device))
    # replacement values: one per nonzero, on the batch's device, a kernel dtype
    x = torch.zeros(good.total_cols, device=device)
    B = torch.zeros((good.total_cols, 4), device=device)
    for call in (lambda d: good.spmv(x, data=d), lambda d: good.spmm(B, data=d),
                 lambda d: good.spmv_per_matrix(x, data=d), lambda d: good.spmm_per_matrix(B, data=d)):
        with pytest.raises(ValueError, match="nnz="):
            call(data[:-1])
        with pytest.raises(ValueError, match="nnz="):
            call(data.reshape(1, -1))
        with pytest.raises(ValueError, match="batch's device"):
            call(data.cpu())
        with pytest.raises(TypeError, match="data dtype"):
            call(torch.zeros(good.nnz, dtype=torch.int32, device=device))
    # strided replacement values are made contiguous and give the same result
    strided = torch.stack([data, data], 1)[:, 0]
    assert torch.equal(good.spmv(x + 1, data=strided), good.spmv(x + 1))
    with pytest.raises(ValueError, match="output tensor"):
        good.graph_spmv(x, None)
    # one value dtype per batch: no silent promotion when packing
    m0 = (data[:3], indices[:3], torch.tensor([0, 3], dtype=indptr.dtype, device=device), (1, 8))
    m1 = (data[:2].double(), indices[:2], torch.tensor([0, 2], dtype=indptr.dtype, device=device), (1, 8))
    with pytest.raises(TypeError, match="same value dtype"):
        CsrBatch.from_matrices([m0, m1])


def test_autotune_picks_and_caches(device, tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("FLAGSPARSE_AMD_TUNE_CACHE", str(path))
    ip, ix, v, ro, no = S.csr_batch(n_mats=128, rows_lo=64, rows_hi=4096, seed=13, device=device)
    b = CsrBatch(v, ix, ip, ro, None, no)
    p = b.params("spmv")
    from flagsparse_amd import _plan

    cands = autotune.batch_spmv_candidates(4, b.total_cols, _plan.chunk_aligned(b.indices, b.data))
    assert any(c.variant == 5 for c in cands)  # the packed arrays sit on 4-element boundaries
    assert p in cands
    assert len(b.tune_results) == len(cands)
    assert path.exists() and "spmv|float32" in path.read_text()
    autotune.clear_memory()
    b2 = CsrBatch(v, ix, ip, ro, None, no)
    assert b2.params("spmv") == p  # read back from the file, not re-tuned
    assert not hasattr(b2, "tune_results")
    x = torch.rand(b.total_cols, device=device)
    assert torch.equal(b2.spmv(x), b.spmv_per_matrix(x, params=p))


def test_unaligned_values_never_take_the_chunk_loops(device, tmp_path, monkeypatch):
    """A parts-kernel (variant 5) winner of the aligned class is not served to
    replacement values that start off a 4-element boundary."""
    monkeypatch.setenv("FLAGSPARSE_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune.clear_memory()
    ip, ix, v, ro, no = S.csr_batch(n_mats=64, rows_lo=64, rows_hi=2048, seed=14, device=device)
    b = CsrBatch(v, ix, ip, ro, None, no)
    x = torch.rand(b.total_cols, device=device) * 2 - 1
    b.params("spmv")
    p5 = autotune.TuneParams(unroll=4, lanes_n=8, variant=5, part_nnz=1 << 13)
    for key in list(b._params):
        b._params[key] = p5  # the aligned class now holds a parts-kernel winner
    ref = b.spmv(x)
    buf = torch.empty(b.nnz + 1, dtype=b.data.dtype, device=device)
    buf[1:] = b.data
    shifted = buf[1:]
    assert shifted.data_ptr() % 16 != 0
    got = b.spmv(x, data=shifted)  # its own class: tuned without the chunk loops
    assert b.params("spmv", data=shifted).variant not in (4, 5)
    assert rel_err(got, ref.double().cpu().numpy()) <= 1e-6
    # explicit chunk-loop parameters on unaligned values take the pipelined loop
    assert rel_err(b.spmv(x, data=shifted, params=p5), ref.double().cpu().numpy()) <= 1e-6
    assert rel_err(b.spmv_per_matrix(x, data=shifted, params=p5), ref.double().cpu().numpy()) <= 1e-6
    autotune.clear_memory()


# ------------------------------------- full batch: 4096 matrices, seed 4
@pytest.fixture(scope="module")
def full_batch():
    dev = torch.device("cuda:0")
    ip, ix, v, ro, no = S.csr_batch(seed=4, device=dev)
    return CsrBatch(v, ix, ip, ro, None, no), (ip, ix, v)


def _oracle_sample(b, arrays, x, B, y, C):
    ip, ix, v = (a.cpu().numpy() for a in arrays)
    for m in (0, 1, 2047, 4095):
        r0, r1 = int(b.row_offsets[m]), int(b.row_offsets[m + 1])
        n0, n1 = int(b.nnz_offsets[m]), int(b.nnz_offsets[m + 1])
        lp = ip[r0 + m: r1 + m + 1]
        xs = x[r0:r1].cpu().numpy()
        assert rel_err(y[r0:r1], O.spmv(lp, ix[n0:n1], v[n0:n1], xs, (r1 - r0, r1 - r0))) <= 1e-5
        Bs = B[r0:r1].cpu().numpy()
        assert rel_err(C[r0:r1], O.spmm(lp, ix[n0:n1], v[n0:n1], Bs, (r1 - r0, r1 - r0))) <= 1e-5


def test_full_batch(device, full_batch):
    b, arrays = full_batch
    assert b.n_mats == 4096
    x = torch.rand(b.total_cols, device=device) * 2 - 1
    B = torch.randn(b.total_cols, 32, device=device)
    pv = autotune.TuneParams(unroll=4, blocks_per_cu=8)
    pm = autotune.TuneParams(unroll=4, slot_scale=4, blocks_per_cu=8)
    y = b.spmv(x, params=pv)
    C = b.spmm(B, params=pm)
    assert torch.equal(b.spmv_per_matrix(x, params=pv), y)
    assert torch.equal(b.spmm_per_matrix(B, params=pm), C)
    _oracle_sample(b, arrays, x, B, y, C)


def test_full_batch_tuned_graph(device, full_batch, tmp_path, monkeypatch):
    """BASELINE config 5 as specified: parameters from the autotuner, the
    SpMV and the k = 32 SpMM each one launch captured in a HIP graph.  Every
    replay equals the eager launch, which equals the per-matrix launch loop
    bit for bit, and a sample of matrices meets the oracle."""
    monkeypatch.setenv("FLAGSPARSE_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    b, arrays = full_batch
    gen = torch.Generator(device=device).manual_seed(4)
    x = torch.rand(b.total_cols, device=device, generator=gen) * 2 - 1
    B = torch.randn(b.total_cols, 32, device=device, generator=gen)
    pv = b.params("spmv")
    pm = b.params("spmm", 32)
    y = b.spmv(x)
    C = b.spmm(B)
    yg, Cg = torch.empty_like(y), torch.empty_like(C)
    gv = b.graph_spmv(x, yg)
    gm = b.graph_spmm(B, Cg)
    for rep in range(3):
        yg.fill_(float("nan"))
        Cg.fill_(float("nan"))
        gv.replay()
        gm.replay()
        torch.cuda.synchronize()
        assert torch.equal(yg, y), f"spmv replay {rep} (params {pv})"
        assert torch.equal(Cg, C), f"spmm replay {rep} (params {pm})"
        y2 = b.spmv(x)  # eager work between the replays
        torch.cuda.synchronize()
        assert torch.equal(y2, y)
    assert torch.equal(b.spmv_per_matrix(x, params=pv), y)
    assert torch.equal(b.spmm_per_matrix(B, params=pm), C)
    _oracle_sample(b, arrays, x, B, y, C)


def test_tune_spmv_single_matrix(device):
    from flagsparse_amd._matrix import CsrHandle

    ip, ix, v, _ = S.power_law_rows(30000, 30000, 2.1, 24, 4096, n_dense=0, seed=8)
    data, idx, ptr = S.to_torch(ip, ix, v, device)
    h = CsrHandle(idx, ptr, (30000, 30000))
    p = autotune.tune_spmv(h, data)
    assert p in autotune.candidates("spmv")
    assert autotune.tune_spmv(h, data) == p  # cached for the shape class
    x = np.random.default_rng(1).uniform(-1, 1, 30000)
    y = h.spmv(data, S.dense(x, device), config=p.spmv_config(), max_blocks=p.max_blocks(device),
               xcd_groups=p.xcd_groups, variant=p.variant)
    assert rel_err(y, O.spmv(ip, ix, v, x, (30000, 30000))) <= 1e-5


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.bfloat16, torch.float16],
                         ids=["f32", "f64", "bf16", "f16"])
@pytest.mark.parametrize("index_dtype", [torch.int32, torch.int64], ids=["i32", "i64"])
@pytest.mark.parametrize("lanes", [4, 8, 16])
def test_batch_spmv_row_range_parts(device, dtype, index_dtype, lanes):
    """Variant 5 (row-range parts, each matrix's column range of x in LDS):
    the oracle per matrix, matrices cut into several parts (small part size)
    or one, a matrix whose columns exceed the LDS window (global gathers),
    bitwise equal to its per-matrix loop, and graph replays equal to eager."""
    from flagsparse_amd.autotune import TuneParams

    mats = _small_batch(1)
    wide = S.random_csr(300, 20000, 0.01, seed=77)  # 20000 columns > the 64 KiB window for f32 / f64
    mats.append(wide + ((300, 20000),))
    b = _to_batch(mats, device, dtype, index_dtype)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, b.total_cols)
    xt = S.dense(x, device, dtype)
    for part_nnz in (64, 1 << 17):
        p = TuneParams(unroll=4, lanes_n=lanes, variant=5, part_nnz=part_nnz)
        parts, start = b.parts(part_nnz)
        assert int(parts[:, 1].sub(parts[:, 0]).sum()) == b.total_rows  # every row in exactly one part
        y = b.spmv(xt, params=p)
        xr = _rounded(x, dtype)
        for m, (ip, ix, v, shape) in enumerate(mats):
            want = O.spmv(ip, ix, _rounded(v, dtype), xr[b.col_offsets[m]: b.col_offsets[m + 1]], shape)
            got = y[b.row_offsets[m]: b.row_offsets[m + 1]]
            assert rel_err(got, want) <= REL_TOL[dtype], f"matrix {m} part_nnz {part_nnz}"
        assert torch.equal(b.spmv_per_matrix(xt, params=p), y)
        yg = torch.empty_like(y)
        g = b._capture(lambda: b.spmv(xt, out=yg, params=p))
        for rep in range(3):
            yg.zero_()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(yg, y), f"replay {rep}"


def _rounded(a, dtype):
    """Values as the kernel sees them (rounded to dtype), in float64."""
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).to(torch.float64).numpy()


def test_spmv_parts_entry_validation(device):
    from flagsparse_amd import _lib

    lib = _lib.load()
    one = torch.zeros(8, dtype=torch.int32, device=device)
    parts = torch.zeros(1, 4, dtype=torch.int64, device=device)
    y = torch.zeros(4, device=device)
    vals = torch.zeros(9, device=device)

    def call(data_ptr, n_parts, lanes, row_len=0):
        return lib.fsp_spmv_csr_parts(0, 0, 0, one.data_ptr(), one.data_ptr(), data_ptr, y.data_ptr(),
                                      y.data_ptr(), 4, n_parts, parts.data_ptr(), lanes, row_len, 0,
                                      _lib.stream_ptr(device))

    with pytest.raises(RuntimeError, match="lanes"):
        _lib.check(call(y.data_ptr(), 1, 5), "fsp_spmv_csr_parts")
    with pytest.raises(RuntimeError, match="row_len"):
        _lib.check(call(y.data_ptr(), 1, 4, row_len=-1), "fsp_spmv_csr_parts")
    with pytest.raises(RuntimeError, match="aligned"):
        _lib.check(call(vals[1:].data_ptr(), 1, 4), "fsp_spmv_csr_parts")
    assert call(y.data_ptr(), 0, 4) == 0  # no parts: nothing to launch


def test_per_matrix_loops_take_strided_operands(device):
    from flagsparse_amd.autotune import TuneParams

    b = _to_batch(_small_batch(3), device)
    xs = torch.rand(2 * b.total_cols, device=device)[::2]  # strided view
    Bs = torch.rand(b.total_cols, 24, device=device)[:, ::2]  # strided columns
    for p in (TuneParams(), TuneParams(unroll=4, lanes_n=8, variant=5, part_nnz=256)):
        assert torch.equal(b.spmv_per_matrix(xs, params=p), b.spmv(xs.contiguous(), params=p))
    pm = TuneParams()
    assert torch.equal(b.spmm_per_matrix(Bs, params=pm), b.spmm(Bs.contiguous(), params=pm))
    with pytest.raises(ValueError, match="x.shape"):
        b.spmv_per_matrix(xs[1:])
    y_strided = torch.empty(2 * b.total_rows, device=device)[::2]
    with pytest.raises(ValueError, match="contiguous"):
        b.spmv(xs.contiguous(), out=y_strided)
    with pytest.raises(ValueError, match="contiguous"):
        b.spmm(Bs.contiguous(), out=torch.empty(b.total_rows, 13, device=device))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64, torch.bfloat16], ids=["f32", "f64", "bf16"])
@pytest.mark.parametrize("lanes", [4, 8, 16])
def test_pipelined_parts_match_plain_loop(device, dtype, lanes):
    """Small parts (row pointers staged in LDS beside x) run the
    software-pipelined parts loop where their rows fit one trip; a part
    holding a whole 7001^2 matrix cannot stage its row pointers (8 bytes a
    row beside x exceed 64 KiB) and runs the plain loop with bounds read
    from HBM.  Both add each lane's chunks in the same order: bitwise equal.
    The odd size puts matrices 1-3's x windows off 16-byte boundaries (the
    element-wise LDS copy), and the oracle pins the first matrix."""
    from flagsparse_amd.autotune import TuneParams

    ip, ix, v, ro, no = S.csr_batch(n_mats=4, rows_lo=7001, rows_hi=7001, nnz_lo=4, nnz_hi=24, seed=11,
                                    device=device)
    b = CsrBatch(v.to(dtype), ix, ip, ro, None, no)
    assert any(int(c) % 4 for c in b.col_offsets[:-1])  # at least one unaligned window
    x = (torch.rand(b.total_cols, device=device) * 2 - 1).to(dtype)
    small = TuneParams(unroll=4, lanes_n=lanes, variant=5, part_nnz=2048)
    whole = TuneParams(unroll=4, lanes_n=lanes, variant=5, part_nnz=1 << 40)
    assert b.parts(whole.part_nnz)[0].shape[0] == b.n_mats
    y = b.spmv(x, params=small)
    assert torch.equal(y, b.spmv(x, params=whole))
    m0 = int(b.row_offsets[1])
    n0 = int(b.col_offsets[1])
    ip0 = ip[: m0 + 1].cpu().numpy().astype(np.int64)
    nz0 = int(ip0[-1])
    vr = b.data[:nz0].double().cpu().numpy()
    ref = O.spmv(ip0, ix[:nz0].cpu().numpy(), vr, x[:n0].double().cpu().numpy(), (m0, n0))
    assert rel_err(y[:m0], ref) <= REL_TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
@pytest.mark.parametrize("lanes", [4, 8, 16])
def test_constant_row_length_batch_computes_bounds(device, dtype, lanes):
    """A batch whose matrices all have 16 nonzeros per row (two replicas of
    one uniform matrix and a uniform matrix of another shape): the parts
    loop computes the row bounds; bitwise equal to the loaded bounds and to
    the per-matrix launches, and within tolerance of the oracle."""
    from flagsparse_amd.autotune import TuneParams

    mats, golds = [], []
    for k, (m, n) in enumerate(((300, 2048), (300, 2048), (1000, 4096))):
        ip, ix, v, _ = S.uniform_rows(m, n, 16, seed=k % 2)
        d, i, p = S.to_torch(ip, ix, v, device, dtype, torch.int32)
        mats.append((d, i, p, (m, n)))
        golds.append((ip, ix, v, (m, n)))
    b = CsrBatch.from_matrices(mats)
    assert b.row_len == 16
    x = torch.rand(b.total_cols, device=device, dtype=dtype)
    params = TuneParams(unroll=4, lanes_n=lanes, variant=5, part_nnz=1024)
    y = b.spmv(x, params=params)
    assert torch.equal(y, b.spmv_per_matrix(x, params=params))
    b.row_len = 0
    assert torch.equal(y, b.spmv(x, params=params))
    xh = x.double().cpu().numpy()
    r0 = c0 = 0
    for ip, ix, v, (m, n) in golds:
        vr = torch.from_numpy(v).to(dtype).double().numpy()
        ref = O.spmv(ip, ix, vr, xh[c0:c0 + n], (m, n))
        assert rel_err(y[r0:r0 + m], ref) <= REL_TOL[dtype]
        r0, c0 = r0 + m, c0 + n
