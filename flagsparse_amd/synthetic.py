"""Seeded synthetic CSR workloads (numpy PCG64; deterministic per seed).

Matrices are canonical CSR: columns sorted and distinct inside each row.
Arrays come back as numpy (indptr/indices int64, values float64); ``to_torch``
moves them to a device in the requested dtypes.

Workloads (sizes of the headline benchmarks):

* ``uniform_rows``      -- every row has the same number of distinct uniform
                           columns (SpMV 16384 x 16384, 16 per row).
* ``power_law_rows``    -- row lengths n = min(cap, floor(x_m U^(-1/(a-1))))
                           with U ~ U(0,1), exponent a, x_m solved so that
                           E[n] = mean exactly (SpMM 262144 x 262144, a = 2.1,
                           mean 24, cap 4096).
* ``power_law_graph``   -- symmetric Chung-Lu graph (power-law weights, edges
                           sampled proportionally to w_i w_j, symmetrised,
                           deduplicated, no self loops), values
                           D^-1/2 A D^-1/2 (200000 nodes, exponent 2.3,
                           mean degree 14).
* ``bsr_blocks``        -- BSR with distinct uniform block columns per block
                           row and N(0, 1) entries (65536 x 65536, 16 x 16,
                           40 blocks per block row).
* ``csr_batch``         -- a packed batch of square CSR matrices: rows
                           log-uniform, row lengths uniform, distinct sorted
                           columns one per stratum (4096 matrices, 64..8192
                           rows, 4..64 nonzeros per row).
"""

import numpy as np
import torch


def _rng(seed):
    return np.random.Generator(np.random.PCG64(int(seed)))


def distinct_sorted_columns(rng, lengths, n_cols):
    """Per-row distinct uniform columns in [0, n_cols), sorted inside rows."""
    lengths = np.asarray(lengths, dtype=np.int64)
    if np.any(lengths > n_cols):
        raise ValueError("a row is longer than n_cols")
    n_rows = lengths.size
    indptr = np.zeros(n_rows + 1, dtype=np.int64)
    np.cumsum(lengths, out=indptr[1:])
    nnz = int(indptr[-1])
    row_of = np.repeat(np.arange(n_rows, dtype=np.int64), lengths)
    cols = rng.integers(0, n_cols, size=nnz, dtype=np.int64)
    while True:
        key = row_of * np.int64(n_cols) + cols
        order = np.argsort(key, kind="stable")
        key = key[order]
        cols = cols[order]
        dup = np.flatnonzero(key[1:] == key[:-1]) + 1
        if dup.size == 0:
            break
        cols[dup] = rng.integers(0, n_cols, size=dup.size, dtype=np.int64)
    return indptr, cols


def uniform_rows(n_rows=16384, n_cols=16384, per_row=16, seed=0):
    """(indptr, indices, values, x): values and x ~ U[-1, 1]."""
    rng = _rng(seed)
    indptr, indices = distinct_sorted_columns(rng, np.full(n_rows, per_row, dtype=np.int64), n_cols)
    values = rng.uniform(-1.0, 1.0, size=indices.size)
    x = rng.uniform(-1.0, 1.0, size=n_cols)
    return indptr, indices, values, x


def power_law_scale(alpha, mean, cap):
    """x_m with sum_{k=1..cap} min(1, (x_m/k)^(alpha-1)) == mean (bisection)."""
    k = np.arange(1, cap + 1, dtype=np.float64)

    def expect(xm):
        return float(np.minimum(1.0, (xm / k) ** (alpha - 1.0)).sum())

    lo, hi = 1e-9, float(cap)
    if expect(hi) < mean:
        raise ValueError("mean not reachable under the cap")
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if expect(mid) < mean:
            lo = mid
        else:
            hi = mid
    return hi


def power_law_lengths(rng, n_rows, alpha, mean, cap):
    xm = power_law_scale(alpha, mean, cap)
    u = 1.0 - rng.random(n_rows)  # (0, 1]
    lengths = np.floor(xm * u ** (-1.0 / (alpha - 1.0)))
    return np.clip(lengths, 0, cap).astype(np.int64)


def power_law_rows(n_rows=262144, n_cols=262144, alpha=2.1, mean=24, cap=4096, n_dense=64, seed=1):
    """(indptr, indices, values, B): values ~ U[-1, 1], B ~ N(0, 1) [n_cols, n_dense]."""
    rng = _rng(seed)
    lengths = power_law_lengths(rng, n_rows, alpha, mean, min(cap, n_cols))
    indptr, indices = distinct_sorted_columns(rng, lengths, n_cols)
    values = rng.uniform(-1.0, 1.0, size=indices.size)
    B = rng.standard_normal((n_cols, n_dense)) if n_dense else None
    return indptr, indices, values, B


def power_law_graph(n_nodes=200000, alpha=2.3, mean_degree=14, seed=3):
    """(indptr, indices, values) of D^-1/2 A D^-1/2 for a symmetric Chung-Lu graph."""
    rng = _rng(seed)
    u = 1.0 - rng.random(n_nodes)
    w = u ** (-1.0 / (alpha - 1.0))
    w = np.minimum(w, np.sqrt(w.sum()))  # Chung-Lu: keep w_i w_j / W <= 1
    p = w / w.sum()
    m = int(n_nodes * mean_degree // 2)
    src = rng.choice(n_nodes, size=m, p=p)
    dst = rng.choice(n_nodes, size=m, p=p)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    rows = np.concatenate([src, dst]).astype(np.int64)
    cols = np.concatenate([dst, src]).astype(np.int64)
    key = np.unique(rows * np.int64(n_nodes) + cols)
    rows = key // n_nodes
    cols = key % n_nodes
    deg = np.bincount(rows, minlength=n_nodes).astype(np.int64)
    indptr = np.zeros(n_nodes + 1, dtype=np.int64)
    np.cumsum(deg, out=indptr[1:])
    dinv = np.zeros(n_nodes)
    nz = deg > 0
    dinv[nz] = 1.0 / np.sqrt(deg[nz])
    values = dinv[rows] * dinv[cols]
    return indptr, cols, values


def random_csr(n_rows, n_cols, density, seed, ensure_nonempty=True):
    """Small test matrices: Bernoulli(density) pattern, values ~ N(0, 1) * 0.125."""
    rng = _rng(seed)
    mask = rng.random((n_rows, n_cols)) < density
    if ensure_nonempty and n_rows and n_cols and not mask.any():
        mask[0, 0] = True
    rows, cols = np.nonzero(mask)
    deg = np.bincount(rows, minlength=n_rows).astype(np.int64) if n_rows else np.zeros(0, np.int64)
    indptr = np.zeros(n_rows + 1, dtype=np.int64)
    np.cumsum(deg, out=indptr[1:])
    values = rng.standard_normal(rows.size) * 0.125
    return indptr, cols.astype(np.int64), values


def to_torch(indptr, indices, values, device, value_dtype=torch.float32, index_dtype=torch.int32,
             indptr_dtype=None):
    indptr_dtype = indptr_dtype or index_dtype
    return (
        torch.from_numpy(np.ascontiguousarray(values)).to(device=device, dtype=value_dtype),
        torch.from_numpy(np.ascontiguousarray(indices)).to(device=device, dtype=index_dtype),
        torch.from_numpy(np.ascontiguousarray(indptr)).to(device=device, dtype=indptr_dtype),
    )


def dense(array, device, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(array)).to(device=device, dtype=dtype)


def bsr_blocks(n_block_rows=4096, n_block_cols=4096, per_row=40, block=16, seed=2):
    """(block_ptr, block_col, block_data[nnzb, block, block]) with `per_row`
    distinct uniform block columns per block row (ascending) and N(0, 1)
    entries (BSR SpMM 65536 x 65536, 16 x 16 blocks, 40 per block row).
    `per_row` may also be an array of per-block-row counts."""
    rng = _rng(seed)
    lengths = np.broadcast_to(np.asarray(per_row, dtype=np.int64), (n_block_rows,))
    block_ptr, block_col = distinct_sorted_columns(rng, lengths, n_block_cols)
    block_data = rng.standard_normal((block_col.size, block, block))
    return block_ptr, block_col, block_data


def bsr_to_csr(block_ptr, block_col, block_data, shape=None):
    """Expand a BSR matrix into canonical CSR.  Every block entry is kept
    (zeros included); with `shape`, rows >= shape[0] and columns >= shape[1]
    (the padding of the last block row / column) are dropped."""
    block_ptr = np.asarray(block_ptr, dtype=np.int64)
    block_col = np.asarray(block_col, dtype=np.int64)
    bs = block_data.shape[1]
    nbr = block_ptr.size - 1
    per = np.diff(block_ptr)
    n_rows_full = nbr * bs
    indptr = np.zeros(n_rows_full + 1, dtype=np.int64)
    np.cumsum(np.repeat(per * bs, bs), out=indptr[1:])
    indices = np.empty(int(indptr[-1]), dtype=np.int64)
    values = np.empty(int(indptr[-1]), dtype=block_data.dtype)
    lane = np.arange(bs, dtype=np.int64)
    for br in range(nbr):
        b0, b1 = int(block_ptr[br]), int(block_ptr[br + 1])
        if b0 == b1:
            continue
        cols = (block_col[b0:b1, None] * bs + lane[None, :]).reshape(-1)      # one row's columns
        vals = block_data[b0:b1].transpose(1, 0, 2).reshape(bs, -1)          # [row in block, cols]
        p0 = int(indptr[br * bs])
        width = cols.size
        indices[p0:p0 + bs * width] = np.tile(cols, bs)
        values[p0:p0 + bs * width] = vals.reshape(-1)
    if shape is None:
        return indptr, indices, values
    n_rows, n_cols = int(shape[0]), int(shape[1])
    row_of = np.repeat(np.arange(n_rows_full, dtype=np.int64), np.diff(indptr))
    keep = (row_of < n_rows) & (indices < n_cols)
    lengths = np.bincount(row_of[keep], minlength=n_rows_full)[:n_rows]
    out_ptr = np.zeros(n_rows + 1, dtype=np.int64)
    np.cumsum(lengths, out=out_ptr[1:])
    return out_ptr, indices[keep], values[keep]


def csr_batch(n_mats=4096, rows_lo=64, rows_hi=8192, nnz_lo=4, nnz_hi=64, seed=4, device="cpu"):
    """Packed batch of square CSR matrices (float32 values U[-1, 1], int32
    indices), generated with torch's seeded generator on `device` (the full
    4096-matrix batch has ~2.4e8 nonzeros: build it on the GPU).

    Matrix sizes n = floor(exp(U[ln rows_lo, ln(rows_hi + 1)))) clipped to
    [rows_lo, rows_hi]; every row's length is uniform in [nnz_lo, nnz_hi]
    (clipped to n); a row of length L over n columns takes one uniform
    column from each of the L integer strata [floor(j n / L),
    floor((j + 1) n / L)), so its columns are distinct and ascending without
    a sort.  Returns (indptr int32, indices int32, values f32, row_offsets,
    nnz_offsets) as tensors on `device`: the local row pointers of all
    matrices concatenated (matrix m's at row_offsets[m] + m), local columns,
    values, and the int64 offset tables."""
    dev = torch.device(device)
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    i64 = dict(dtype=torch.int64, device=dev)
    u = torch.rand(n_mats, generator=g, dtype=torch.float64, device=dev)
    lo_l, hi_l = float(np.log(rows_lo)), float(np.log(rows_hi + 1))
    n = torch.floor(torch.exp(lo_l + u * (hi_l - lo_l))).to(torch.int64).clamp_(rows_lo, rows_hi)
    row_offsets = torch.zeros(n_mats + 1, **i64)
    torch.cumsum(n, 0, out=row_offsets[1:])
    total_rows = int(row_offsets[-1])
    mat_of_row = torch.repeat_interleave(torch.arange(n_mats, **i64), n)
    lengths = torch.randint(nnz_lo, nnz_hi + 1, (total_rows,), generator=g, **i64)
    lengths = torch.minimum(lengths, n[mat_of_row])
    csum = torch.cumsum(lengths, 0)
    nnz_offsets = torch.zeros(n_mats + 1, **i64)
    nnz_offsets[1:] = csum[row_offsets[1:] - 1]
    total = int(nnz_offsets[-1])
    indptr = torch.zeros(total_rows + n_mats, dtype=torch.int32, device=dev)
    indptr[torch.arange(total_rows, **i64) + mat_of_row + 1] = (csum - nnz_offsets[mat_of_row]).to(torch.int32)
    row_start = csum - lengths
    owner = torch.repeat_interleave(torch.arange(total_rows, **i64), lengths)
    j = torch.arange(total, **i64) - row_start[owner]
    L = lengths[owner]
    nr = n[mat_of_row[owner]]
    del owner
    lo = torch.div(j * nr, L, rounding_mode="floor")
    hi = torch.div((j + 1) * nr, L, rounding_mode="floor")
    del j, L, nr
    pick = torch.rand(total, generator=g, dtype=torch.float64, device=dev)
    indices = (lo + torch.floor(pick * (hi - lo)).to(torch.int64)).to(torch.int32)
    del lo, hi, pick
    values = torch.rand(total, generator=g, dtype=torch.float32, device=dev) * 2.0 - 1.0
    return indptr, indices, values, row_offsets, nnz_offsets


def uniform_rows_device(n_rows, n_cols, per_row, seed, device="cpu"):
    """Large uniform-row CSR built with torch's seeded generator on `device`:
    every row has `per_row` distinct ascending columns, one uniform column
    from each of the integer strata [floor(j n / L), floor((j + 1) n / L));
    values and x ~ U[-1, 1] (float32); int32 indices and row pointer.
    Returns (indptr, indices, values, x) on `device`."""
    dev = torch.device(device)
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    L = int(per_row)
    if L > n_cols:
        raise ValueError("per_row exceeds n_cols")
    nnz = int(n_rows) * L
    indptr = torch.arange(0, nnz + 1, L, dtype=torch.int64, device=dev).to(torch.int32)
    j = torch.arange(L, dtype=torch.int64, device=dev)
    lo = (j * n_cols) // L
    width = ((j + 1) * n_cols) // L - lo
    pick = torch.rand((int(n_rows), L), generator=g, dtype=torch.float64, device=dev)
    indices = (lo + torch.floor(pick * width).to(torch.int64)).to(torch.int32).reshape(-1)
    del pick
    values = torch.rand(nnz, generator=g, dtype=torch.float32, device=dev) * 2.0 - 1.0
    x = torch.rand(int(n_cols), generator=g, dtype=torch.float32, device=dev) * 2.0 - 1.0
    return indptr, indices, values, x
