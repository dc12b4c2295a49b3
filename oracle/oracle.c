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
