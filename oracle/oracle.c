/*
 * oracle.c -- CPU restatement of the reference's CSR SpMV / SpMM / SDDMM /
 * transpose semantics.  TEST INFRASTRUCTURE ONLY: it is the checker that
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg compare
 * the HIP kernels against; nothing in flagsparse_amd/ links or calls it.
 *
 * Reference files (FlagSparse, src/flagsparse/sparse_operations/):
 *   spmv_csr.py:179-205   _spmv_csr_real_kernel: y[r] = sum_p data[p] x[col p];
 *                         f32 inputs accumulated in f64 (spmv_csr.py:138-143)
 *   spmv_csr.py:498-536   _transpose_csr_for_spmv: stable bucketing by column
 *   spmm_csr.py:144-200   _spmm_csr_real_kernel: C[r,:] = sum_p data[p] B[col p,:]
 *   spmm_csr.py:407-456   transposed operand for op=trans / conj
 *   sddmm_csr.py:404, 471-483  out[p] = alpha <x[row p], y[col p]> + beta data[p]
 *
 * Every reduction runs in long double (x87 80-bit on x86-64) and is rounded to
 * double once, so the oracle is at least as accurate as an FP64 golden.  All
 * index arrays are int64 and all values double; the Python wrapper converts.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* y[r] = sum_{p in row r} data[p] * x[indices[p]]  (spmv_csr.py:196-205). */
void orc_spmv_csr(int64_t n_rows, const int64_t* indptr, const int64_t* indices, const double* data,
                  const double* x, double* y) {
  for (int64_t r = 0; r < n_rows; ++r) {
    long double acc = 0.0L;
    for (int64_t p = indptr[r]; p < indptr[r + 1]; ++p) {
      acc += (long double)data[p] * (long double)x[indices[p]];
    }
    y[r] = (double)acc;
  }
}

/* y[c] = sum_{p: indices[p] == c} data[p] * x[row(p)]: A^T x by scattering the
 * CSR rows (a different algorithm from the GPU's CSC gather). */
void orc_spmv_csr_trans(int64_t n_rows, int64_t n_cols, const int64_t* indptr, const int64_t* indices,
                        const double* data, const double* x, double* y) {
  long double* acc = (long double*)calloc((size_t)(n_cols > 0 ? n_cols : 1), sizeof(long double));
  for (int64_t r = 0; r < n_rows; ++r) {
    for (int64_t p = indptr[r]; p < indptr[r + 1]; ++p) {
      acc[indices[p]] += (long double)data[p] * (long double)x[r];
    }
  }
  for (int64_t c = 0; c < n_cols; ++c) y[c] = (double)acc[c];
  free(acc);
}

/* C[r, j] = sum_{p in row r} data[p] * B[indices[p], j]  (spmm_csr.py:174-198).
 * B is [*, n_dense] row-major, C is [n_rows, n_dense] row-major. */
void orc_spmm_csr(int64_t n_rows, int64_t n_dense, const int64_t* indptr, const int64_t* indices,
                  const double* data, const double* B, double* C) {
  long double* acc = (long double*)malloc((size_t)(n_dense > 0 ? n_dense : 1) * sizeof(long double));
  for (int64_t r = 0; r < n_rows; ++r) {
    for (int64_t j = 0; j < n_dense; ++j) acc[j] = 0.0L;
    for (int64_t p = indptr[r]; p < indptr[r + 1]; ++p) {
      const long double a = (long double)data[p];
      const double* brow = B + indices[p] * n_dense;
      for (int64_t j = 0; j < n_dense; ++j) acc[j] += a * (long double)brow[j];
    }
    for (int64_t j = 0; j < n_dense; ++j) C[r * n_dense + j] = (double)acc[j];
  }
  free(acc);
}

/* C = A^T B by scattering CSR rows: C[indices[p], :] += data[p] * B[r, :].
 * C is [n_cols, n_dense]. */
void orc_spmm_csr_trans(int64_t n_rows, int64_t n_cols, int64_t n_dense, const int64_t* indptr,
                        const int64_t* indices, const double* data, const double* B, double* C) {
  const size_t n = (size_t)(n_cols * n_dense > 0 ? n_cols * n_dense : 1);
  long double* acc = (long double*)calloc(n, sizeof(long double));
  for (int64_t r = 0; r < n_rows; ++r) {
    const double* brow = B + r * n_dense;
    for (int64_t p = indptr[r]; p < indptr[r + 1]; ++p) {
      const long double a = (long double)data[p];
      long double* crow = acc + indices[p] * n_dense;
      for (int64_t j = 0; j < n_dense; ++j) crow[j] += a * (long double)brow[j];
    }
  }
  for (int64_t i = 0; i < n_cols * n_dense; ++i) C[i] = (double)acc[i];
  free(acc);
}

/* out[p] = alpha * <X[row p, :], Y[indices[p], :]> + beta * data[p]
 * (data may be NULL: then the beta term is absent; sddmm_csr.py:404, 471-483). */
void orc_sddmm_csr(int64_t n_rows, int64_t k, const int64_t* indptr, const int64_t* indices, const double* X,
                   const double* Y, const double* data, double alpha, double beta, double* out) {
  for (int64_t r = 0; r < n_rows; ++r) {
    const double* xrow = X + r * k;
    for (int64_t p = indptr[r]; p < indptr[r + 1]; ++p) {
      const double* yrow = Y + indices[p] * k;
      long double dot = 0.0L;
      for (int64_t j = 0; j < k; ++j) dot += (long double)xrow[j] * (long double)yrow[j];
      long double v = (long double)alpha * dot;
      if (data != NULL) v += (long double)beta * (long double)data[p];
      out[p] = (double)v;
    }
  }
}

/* Stable CSR -> CSC by counting sort (spmv_csr.py:513-528 does a stable
 * argsort by column): colptr[n_cols+1], rowind[nnz], perm[nnz] with CSC slot q
 * holding CSR nonzero perm[q]; inside a column, CSR order is kept. */
void orc_csr_to_csc(int64_t n_rows, int64_t n_cols, const int64_t* indptr, const int64_t* indices,
                    int64_t* colptr, int64_t* rowind, int64_t* perm) {
  const int64_t nnz = indptr[n_rows];
  for (int64_t c = 0; c <= n_cols; ++c) colptr[c] = 0;
  for (int64_t p = 0; p < nnz; ++p) colptr[indices[p] + 1] += 1;
  for (int64_t c = 0; c < n_cols; ++c) colptr[c + 1] += colptr[c];
  int64_t* next = (int64_t*)malloc((size_t)(n_cols > 0 ? n_cols : 1) * sizeof(int64_t));
  for (int64_t c = 0; c < n_cols; ++c) next[c] = colptr[c];
  for (int64_t r = 0; r < n_rows; ++r) {
    for (int64_t p = indptr[r]; p < indptr[r + 1]; ++p) {
      const int64_t q = next[indices[p]]++;
      rowind[q] = r;
      perm[q] = p;
    }
  }
  free(next);
}

/* CSR -> BSR with bs x bs blocks (no reference counterpart on this path; the
 * storage convention data[nnzb][bs][bs], row-major inside a block, block
 * columns ascending inside a block row, follows the reference's BSR SpMV,
 * spmv_bsr.py:451-452).  Pass block_col == NULL to only count: returns nnzb.
 * block_data must be zeroed by the caller (missing entries are zero). */
int64_t orc_csr_to_bsr(int64_t n_rows, int64_t n_cols, int64_t bs, const int64_t* indptr, const int64_t* indices,
                       const double* data, int64_t* block_ptr, int64_t* block_col, double* block_data) {
  const int64_t nbr = (n_rows + bs - 1) / bs, nbc = (n_cols + bs - 1) / bs;
  int64_t* mark = (int64_t*)malloc((size_t)(nbc > 0 ? nbc : 1) * sizeof(int64_t));
  int64_t* slot = (int64_t*)malloc((size_t)(nbc > 0 ? nbc : 1) * sizeof(int64_t));
  for (int64_t c = 0; c < nbc; ++c) mark[c] = -1;
  int64_t nnzb = 0;
  if (block_ptr) block_ptr[0] = 0;
  for (int64_t br = 0; br < nbr; ++br) {
    const int64_t r0 = br * bs, r1 = (r0 + bs < n_rows) ? r0 + bs : n_rows;
    /* collect the distinct block columns of this block row, ascending */
    int64_t cnt = 0;
    for (int64_t r = r0; r < r1; ++r)
      for (int64_t p = indptr[r]; p < indptr[r + 1]; ++p) {
        const int64_t bc = indices[p] / bs;
        if (mark[bc] != br) { mark[bc] = br; ++cnt; }
      }
    if (block_col) {
      int64_t k = nnzb;
      for (int64_t bc = 0; bc < nbc; ++bc)
        if (mark[bc] == br) { block_col[k] = bc; slot[bc] = k; ++k; }
      for (int64_t r = r0; r < r1; ++r)
        for (int64_t p = indptr[r]; p < indptr[r + 1]; ++p) {
          const int64_t bc = indices[p] / bs;
          block_data[(slot[bc] * bs + (r - r0)) * bs + (indices[p] - bc * bs)] += data[p];
        }
    }
    nnzb += cnt;
    if (block_ptr) block_ptr[br + 1] = nnzb;
  }
  free(mark);
  free(slot);
  return nnzb;
}

/* C[n_brows*bs, n_dense] = BSR(A) @ B, long-double accumulation. */
void orc_bsr_spmm(int64_t n_brows, int64_t bs, int64_t n_dense, const int64_t* block_ptr, const int64_t* block_col,
                  const double* block_data, const double* B, double* C) {
  long double* acc = (long double*)malloc((size_t)(bs * (n_dense > 0 ? n_dense : 1)) * sizeof(long double));
  for (int64_t br = 0; br < n_brows; ++br) {
    for (int64_t i = 0; i < bs * n_dense; ++i) acc[i] = 0.0L;
    for (int64_t j = block_ptr[br]; j < block_ptr[br + 1]; ++j) {
      const double* blk = block_data + j * bs * bs;
      for (int64_t r = 0; r < bs; ++r)
        for (int64_t c = 0; c < bs; ++c) {
          const long double a = (long double)blk[r * bs + c];
          if (a == 0.0L) continue;
          const double* brow = B + (block_col[j] * bs + c) * n_dense;
          for (int64_t n = 0; n < n_dense; ++n) acc[r * n_dense + n] += a * (long double)brow[n];
        }
    }
    for (int64_t i = 0; i < bs * n_dense; ++i) C[br * bs * n_dense + i] = (double)acc[i];
  }
  free(acc);
}
