/*
 * flagsparse_hip.h -- C ABI of libflagsparse_hip.so, the gfx950 (MI355X) kernels
 * behind the flagsparse_amd CSR SpMV / SpMM / SDDMM operators.
 *
 * Conventions
 *   - Every pointer named *_dev is device memory; everything else is host memory.
 *   - Sizes are int64_t.  Value / index types are the FSP_* enums below.
 *   - Every entry point takes the hipStream_t to launch on (passed as void*;
 *     NULL = the default stream), is re-entrant, never allocates device memory
 *     and never synchronises the device unless its comment says so.  Scratch
 *     memory is always supplied by the caller (sizes come from *_workspace()).
 *   - Return value: FSP_OK (0) or a negative fsp_status; fsp_last_error()
 *     returns a thread-local message for the last failure on this thread.
 *
 * The reference (FlagSparse) is pure Python + Triton: its hot path has no
 * native FFI.  Each entry point below replaces the Triton launch site named in
 * its comment (file:line relative to src/flagsparse/sparse_operations/).
 */
#ifndef FLAGSPARSE_HIP_H
#define FLAGSPARSE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSP_ABI_VERSION 15
#define FSP_MAX_BINS 8

typedef enum fsp_value_type {
  FSP_F32 = 0,
  FSP_F64 = 1,
  FSP_F16 = 2,
  FSP_BF16 = 3,
  FSP_F8E4M3 = 4 /* OCP e4m3 (finite-only, bias 7); BSR SpMM operand only */
} fsp_value_type;

typedef enum fsp_index_type { FSP_I32 = 0, FSP_I64 = 1 } fsp_index_type;

typedef enum fsp_status {
  FSP_OK = 0,
  FSP_ERR_ARG = -1,       /* bad size / pointer / combination */
  FSP_ERR_DTYPE = -2,     /* value or index type not supported by this entry */
  FSP_ERR_LAUNCH = -3,    /* the HIP runtime rejected a launch or copy */
  FSP_ERR_WORKSPACE = -4  /* caller-supplied scratch too small */
} fsp_status;

/* One row-length bin of a CSR work plan.  `lanes` is the number of lanes that
 * cooperate on one row: SpMV splits a row's nonzeros over `lanes` lanes; SpMM
 * and SDDMM split them over `lanes` nonzero slots (each slot spans the lanes
 * that cover the dense dimension). */
typedef struct fsp_bin {
  int32_t lanes;
  int32_t reserved;
  int64_t count;         /* rows in this bin */
  const int32_t* rows;   /* device row ids, ascending; NULL = rows 0..count-1 */
} fsp_bin;

/* Work plan of one CSR (or CSC) matrix: rows binned by length, plus the
 * nnz-balanced split of rows longer than the last bin into segments.  All
 * arrays are device memory owned by the caller (built by fsp_csr_row_bins and
 * the host-side segment split). */
typedef struct fsp_csr_plan {
  int32_t n_bins;
  int32_t unroll;                 /* nonzeros per lane per loop trip (1..8) */
  fsp_bin bins[FSP_MAX_BINS];
  int64_t n_segments;             /* long-row segments */
  const int32_t* seg_row_dev;     /* [n_segments] row of each segment */
  const int64_t* seg_begin_dev;   /* [n_segments] first nonzero */
  const int64_t* seg_end_dev;     /* [n_segments] one past the last nonzero */
  int64_t n_long_rows;
  const int32_t* long_rows_dev;   /* [n_long_rows] */
  const int64_t* long_seg_ptr_dev;/* [n_long_rows+1] segment range per long row */
  void* partials_dev;             /* scratch: n_segments * dense_cols * acc bytes */
  int64_t partials_bytes;
  int64_t max_blocks;             /* 0: one workgroup per row tile; > 0: persistent
                                     grid of at most this many workgroups, each
                                     walking a contiguous run of tiles */
  int64_t xcd_groups;             /* > 1: cut each bin's row tiles into this many
                                     contiguous regions, workgroup b serving
                                     region b % xcd_groups (8 = one per XCD
                                     on MI355X: neighbouring rows share an L2) */
  int64_t variant;                /* SpMV loop: 0 = software-pipelined (next trip's
                                     indices load under this trip's x gathers),
                                     1 = plain loop, 2 = plain and 3 = pipelined
                                     loop with x staged in LDS by 1024-thread
                                     workgroups (x_len * element size <= 64 KiB;
                                     persistent grid, max_blocks 0 = two per
                                     CU), 4 = as 2 with the index / value (or
                                     perm) streams read as 4-element chunks
                                     (arrays aligned to 4 elements, unroll 4
                                     or 8); other kernels ignore it */
  int64_t row_len;                /* > 0: every row has exactly row_len nonzeros,
                                     indptr[i] = indptr_base + i * row_len, and
                                     the SpMV kernels compute row bounds instead
                                     of loading them; 0: rows vary (bounds are
                                     loaded).  The caller vouches for it */
  int64_t indptr_base;            /* indptr[0] when row_len > 0 */
} fsp_csr_plan;

/* ---- library ------------------------------------------------------------ */
int fsp_abi_version(void);
const char* fsp_last_error(void);
/* gcnArchName of the current device into buf (e.g. "gfx950:sramecc+:xnack-"). */
int fsp_device_arch(char* buf, int64_t buf_len);

/* ---- plan building -------------------------------------------------------
 * Replaces the per-call torch.nonzero row bucketing of
 * spmv_csr.py:371-420 (_build_spmv_opt_buckets) and the symbolic count /
 * compact kernels of spmm_csr.py:3022-3072.  Rows are stably partitioned by
 * length into n_bins bins: bin b holds rows with
 * upper[b-1] < len <= upper[b] (upper[-1] = -1, upper[n_bins-1] = +inf).
 * rows_out_dev[n_rows] receives the row ids bin after bin (ascending inside
 * each bin); bin_offsets_out_dev[n_bins+1] the bin boundaries. */
/* Validation figures of a CSR in one launch (replaces the per-check host
 * syncs of _prepare_spmv_csr_matrix, spmv_csr.py:539-595, and
 * _prepare_spmm_csr_matrix, spmm_csr.py:342-404): stats_dev[6] receives
 * indptr[0], indptr[n_rows], the number of rows with indptr[i+1] <
 * indptr[i], the smallest and largest column index (INT64_MAX / INT64_MIN
 * when nnz = 0) and the longest row length.  Reads only indptr[0..n_rows]
 * and indices[0..nnz). */
int fsp_csr_inspect(int32_t indptr_type, const void* indptr_dev, int32_t index_type,
                    const void* indices_dev, int64_t n_rows, int64_t nnz,
                    int64_t* stats_dev, void* stream);

int fsp_csr_row_bins_workspace(int64_t n_rows, int32_t n_bins, int64_t* bytes);
int fsp_csr_row_bins(int32_t indptr_type, const void* indptr_dev, int64_t n_rows,
                     int32_t n_bins, const int64_t* upper,
                     int32_t* rows_out_dev, int64_t* bin_offsets_out_dev,
                     void* ws_dev, int64_t ws_bytes, void* stream);

/* ---- CSR -> CSC -----------------------------------------------------------
 * Replaces _transpose_csr_for_spmv (spmv_csr.py:498-536) and
 * _transpose_csr_for_spmm (spmm_csr.py:407-445): a stable bucketing of the
 * nonzeros by column.  Outputs colptr[n_cols+1] (type colptr_type),
 * rowind[nnz] (type rowind_type) and perm[nnz] (type perm_type) with
 * CSC position q holding CSR nonzero perm[q]; entries of one column keep CSR
 * order (ascending row).  Bit-exact with a stable argsort by column. */
int fsp_csr_to_csc_workspace(int64_t n_rows, int64_t n_cols, int64_t nnz,
                             int32_t perm_type, int64_t* bytes);
int fsp_csr_to_csc(int32_t indptr_type, int32_t index_type,
                   const void* indptr_dev, const void* indices_dev,
                   int64_t n_rows, int64_t n_cols, int64_t nnz,
                   int32_t colptr_type, void* colptr_dev,
                   int32_t rowind_type, void* rowind_dev,
                   int32_t perm_type, void* perm_dev,
                   void* ws_dev, int64_t ws_bytes, void* stream);
/* out[q] = in[perm[q]] for q < nnz (value bytes = elem_bytes: 1,2,4,8,16). */
int fsp_gather(int32_t elem_bytes, int32_t perm_type, const void* perm_dev,
               const void* in_dev, void* out_dev, int64_t nnz, void* stream);

/* ---- SpMV ----------------------------------------------------------------
 * y = A x over a CSR (op non) or over the CSC of A (op trans/conj; pass the
 * CSC arrays and perm != NULL so values are read as data[perm[q]]).
 * Replaces _spmv_csr_real_kernel (spmv_csr.py:179-205) as launched by
 * _triton_spmv_csr_impl_prepared (spmv_csr.py:702-746) and the bucketed
 * kernels of the use_opt path (spmv_csr.py:247-368, 423-482).
 * Accumulation: f64 for f32/f64 values (the reference promotes f32 SpMV to
 * f64, spmv_csr.py:138-143), f32 for f16/bf16.  y has value_type.  nnz = 0
 * writes y = 0 (indices / data / x may then be NULL).  x_len = elements of x
 * (columns of the operator; every column index is below it). */
int fsp_spmv_csr(int32_t value_type, int32_t index_type, int32_t indptr_type,
                 const void* indptr_dev, const void* indices_dev,
                 const void* data_dev, int32_t perm_type, const void* perm_dev,
                 const void* x_dev, void* y_dev, int64_t n_rows, int64_t nnz,
                 int64_t x_len, const fsp_csr_plan* plan, void* stream);

/* Row-range SpMV for matrices whose rows fall into groups with narrow column
 * ranges (a batch of matrices is the case it serves; no counterpart in the
 * reference, which runs one matrix per call).  parts_dev[n_parts][4] holds
 * (row_begin, row_end, col_lo, col_hi) per part: rows [row_begin, row_end)
 * use only columns [col_lo, col_hi).  One 1024-thread workgroup per part
 * copies x[col_lo, col_hi) into LDS when it holds at most 64 KiB (else it
 * gathers from global memory), and the part's slice of indptr beside it when
 * both fit, and runs the rows with `lanes` (4, 8 or 16)
 * lanes per row, reading indices and data as 4-element chunks (both aligned
 * to 4 elements).  Parts must not share rows; rows outside every part are not
 * written.  Accumulation as fsp_spmv_csr; nnz = length of indices / data.
 * The table is read on the device and not checked: every row_end must be a
 * valid row of indptr_dev (indptr has row_end + 1 entries), col_lo <= col_hi
 * must lie within x, and the rows' column ids within their part's window.
 * The Python layer builds it from a validated batch (CsrBatch.parts).
 * row_len > 0 states that every row holds exactly row_len nonzeros and
 * indptr[i] = indptr_base + i * row_len (the caller vouches for it): the
 * kernel then computes row bounds instead of loading them; 0 loads them. */
int fsp_spmv_csr_parts(int32_t value_type, int32_t index_type, int32_t indptr_type,
                       const void* indptr_dev, const void* indices_dev, const void* data_dev,
                       const void* x_dev, void* y_dev, int64_t nnz, int64_t n_parts,
                       const int64_t* parts_dev, int32_t lanes, int64_t row_len,
                       int64_t indptr_base, void* stream);

/* ---- SpMM ----------------------------------------------------------------
 * C[n_rows, n_dense] = A B with B row-major (ldb >= n_dense) and C row-major
 * (ldc >= n_dense).  Same CSR / CSC+perm convention as fsp_spmv_csr.
 * Replaces _spmm_csr_real_kernel (spmm_csr.py:144-200) launched by
 * _triton_spmm_csr_impl (spmm_csr.py:3373-3466), and the bucketed /
 * long-row split kernels of the opt path (spmm_csr.py:3469-3794).
 * Accumulation: f32 for f32/f16/bf16, f64 for f64 (spmm_csr.py:3419).
 * `lanes_n` = lanes spanning the dense dimension (power of two, 1..64);
 * `nnz` = nonzeros of A, `b_rows` = rows of B (they bound the kernel's
 * offsets: 32-bit addressing is used when every array fits in 4 GiB). */
int fsp_spmm_csr(int32_t value_type, int32_t index_type, int32_t indptr_type,
                 const void* indptr_dev, const void* indices_dev,
                 const void* data_dev, int32_t perm_type, const void* perm_dev,
                 const void* B_dev, int64_t ldb, void* C_dev, int64_t ldc,
                 int64_t n_rows, int64_t n_dense, int64_t nnz, int64_t b_rows,
                 int32_t lanes_n, const fsp_csr_plan* plan, void* stream);

/* The differentiable SpMM's backward in one pass over A's CSC.  No
 * counterpart in the reference, which has no autograd; it computes what the
 * reference's trans SpMM (spmm_csr.py:3981, op trans) and
 * flagsparse_sddmm_csr (sddmm_csr.py:389) compute for C = A B, from one
 * gather of dC's rows.  With A's CSC (colptr, rowind over n_cols columns,
 * values already in CSC order in data_csc) and pos[q] = the CSR position of
 * CSC nonzero q (the stable CSC permutation):
 *   dB[j, :]        = sum over q of column j: data_csc[q] * dC[rowind[q], :]
 *   dvalues[pos[q]] = <dC[rowind[q], :], B[j, :]>
 * dC is n_rows x n_dense (lddc); B (the forward's dense operand) and dB are
 * n_cols x n_dense (ldb, lddb); dvalues has nnz entries in CSR order.
 * `plan` is the plan of the CSC (long columns as segments, partials as
 * fsp_spmm_csr).  dB is bitwise equal to fsp_spmm_csr over the same CSC and
 * plan (with B's rows 16-byte aligned wherever dC's and dB's are); dvalues
 * are f32 (f64 for f64) dot products in the lanes' order.
 * n_dense must fit one lane group (64 lanes x 16 bytes: 256 f32 / 512
 * bf16 or f16 / 128 f64 columns on 16-byte aligned rows, else 64 columns);
 * wider operands return FSP_ERR_ARG and the caller runs the two kernels. */
int fsp_spmm_csr_trans_sddmm(int32_t value_type, int32_t index_type, int32_t indptr_type,
                             const void* colptr_dev, const void* rowind_dev,
                             const void* data_csc_dev, int32_t pos_type, const void* pos_dev,
                             const void* dC_dev, int64_t lddc, const void* B_dev, int64_t ldb,
                             void* dB_dev, int64_t lddb, void* dvalues_dev,
                             int64_t n_cols, int64_t n_dense, int64_t nnz, int64_t n_rows,
                             int32_t lanes_n, const fsp_csr_plan* plan, void* stream);

/* Size of the static device scratch behind fsp_spmm_csr_noalloc (per GPU). */
#define FSP_NOALLOC_SCRATCH_BYTES (16ll << 20)

/* No counterpart in the reference: fsp_spmm_csr for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and result as
 * fsp_spmm_csr, except that plan->partials_dev / partials_bytes are ignored:
 * the long-row partials live in a file-scope device array of
 * FSP_NOALLOC_SCRATCH_BYTES (16 MiB) that each GPU holds until the process
 * ends.  Calls with plan->n_segments * n_dense * (8 for f64, else 4) above
 * that size return FSP_ERR_WORKSPACE before anything is launched.  The call
 * allocates nothing (no hipMalloc / hipFree, no allocating library routine);
 * its first call on a device resolves the array's address there, and the
 * stream must belong to the current device.  Every call on one GPU shares
 * that GPU's array, which keeps the last call's partials (each call writes
 * every partial it reads, so nothing needs clearing): calls that overlap on
 * one GPU (other streams, other threads) race and must be serialised by the
 * caller. */
int fsp_spmm_csr_noalloc(int32_t value_type, int32_t index_type, int32_t indptr_type,
                         const void* indptr_dev, const void* indices_dev,
                         const void* data_dev, int32_t perm_type, const void* perm_dev,
                         const void* B_dev, int64_t ldb, void* C_dev, int64_t ldc,
                         int64_t n_rows, int64_t n_dense, int64_t nnz, int64_t b_rows,
                         int32_t lanes_n, const fsp_csr_plan* plan, void* stream);

/* ---- SDDMM ---------------------------------------------------------------
 * out[p] = alpha * <X[row(p), :], Y[col(p), :]> + beta * data_in[p]
 * (data_in may be NULL when beta == 0).  X is [n_rows, k] (row stride ldx),
 * Y is [n_cols, k] (row stride ldy).  Replaces _sddmm_csr_real_kernel
 * (sddmm_csr.py:121-164) launched by _run_sddmm_prepared (sddmm_csr.py:291-386).
 * Accumulation f32 (f32/f16/bf16) or f64 (f64).  The plan's segments cut long
 * rows; SDDMM needs no partials.  `nnz` = nonzeros of the pattern, `y_rows` =
 * rows of Y (bounds for 32-bit addressing). */
int fsp_sddmm_csr(int32_t value_type, int32_t index_type, int32_t indptr_type,
                  const void* indptr_dev, const void* indices_dev,
                  const void* X_dev, int64_t ldx, const void* Y_dev, int64_t ldy,
                  const void* data_in_dev, void* out_dev, int64_t n_rows,
                  int64_t k, int64_t nnz, int64_t y_rows, double alpha, double beta,
                  int32_t lanes_k, const fsp_csr_plan* plan, void* stream);

/* ---- BSR (16 x 16 blocks) on the matrix cores ------------------------------
 * No reference counterpart computes BSR SpMM; the block storage follows the
 * reference's BSR SpMV (spmv_bsr.py:448-470: data[nnzb][bd][bd] row-major
 * inside a block, block columns ascending inside a block row).
 *
 * CSR -> BSR (bit-exact vs a host conversion): fsp_csr_to_bsr_count sorts the
 * block keys and writes the block count to *nnzb_dev; the caller reads it,
 * allocates the outputs and calls fsp_csr_to_bsr_fill with the same
 * workspace.  Values of a block that no nonzero touches are zero. */
int fsp_csr_to_bsr_workspace(int64_t n_rows, int64_t n_cols, int64_t nnz, int64_t* bytes);
int fsp_csr_to_bsr_count(int32_t indptr_type, int32_t index_type, const void* indptr_dev,
                         const void* indices_dev, int64_t n_rows, int64_t n_cols, int64_t nnz,
                         void* ws_dev, int64_t ws_bytes, int64_t* nnzb_dev, void* stream);
int fsp_csr_to_bsr_fill(int32_t value_type, int32_t indptr_type, int32_t index_type,
                        const void* indptr_dev, const void* indices_dev, const void* data_dev,
                        int64_t n_rows, int64_t n_cols, int64_t nnz, int64_t nnzb,
                        int32_t block_ptr_type, void* block_ptr_dev, int32_t* block_col_dev,
                        void* block_data_dev, void* ws_dev, int64_t ws_bytes, void* stream);
/* max |x| over a [rows, cols] array (row stride ld) into *out_dev (float). */
int fsp_absmax(int32_t value_type, const void* x_dev, int64_t rows, int64_t cols, int64_t ld,
               float* out_dev, void* stream);
/* Packs B [K, N] (row stride ldb) into the MFMA operand layout
 * Bp[k/8][n][k%8] of out_type (FSP_BF16 / FSP_F16, 16 bytes per (k/8, n);
 * FSP_F8E4M3, 8 bytes, codes of B * 448 / absmax).  Bp holds
 * ceil(K/16)*2 x ceil(N/128)*128 entries of 8 elements; padding is zero. */
int fsp_bsr_pack_b(int32_t in_type, int32_t out_type, const void* B_dev, int64_t ldb, int64_t K,
                   int64_t N, const float* absmax_dev, void* Bp_dev, void* stream);
/* Per-block absmax quantisation to OCP FP8 e4m3: codes[nnzb][16][16] and
 * scales[nnzb] with block = codes * scale (round to nearest even). */
int fsp_bsr_quantize_fp8(int32_t value_type, const void* block_data_dev, int64_t nnzb, uint8_t* q_dev,
                         float* scales_dev, void* stream);
/* C[M, N] (f32, row stride ldc) = BSR(A) @ B with B packed by fsp_bsr_pack_b.
 * value_type FSP_BF16 / FSP_F16 (block_data of that type) or FSP_F8E4M3
 * (block_data = codes, block_scales and the B absmax used for packing). */
int fsp_bsr_spmm(int32_t value_type, int32_t block_ptr_type, const void* block_ptr_dev,
                 const int32_t* block_col_dev, const void* block_data_dev, const float* block_scales_dev,
                 int64_t n_block_rows, int64_t nnzb, const void* Bp_dev, const float* b_absmax_dev,
                 int64_t K, int64_t N, float* C_dev, int64_t ldc, int64_t M, void* stream);

/* ---- BSR SpMV ----------------------------------------------------------------
 * y (padded: n_out_blocks * block_dim) = A x over a BSR matrix with square
 * blocks of block_dim in [1, 32], data [nnzb][block_dim][block_dim]
 * row-major.  op non: ptr = block row pointer, bidx = block columns,
 * perm = NULL, x padded to n_block_cols * block_dim.  op trans / conj (real
 * values): ptr / bidx / perm = block column pointer, block rows and block
 * ids of the block CSC (fsp_csr_to_csc over the block structure), x padded
 * to n_block_rows * block_dim.  Deterministic and atomic-free.  Replaces
 * _spmv_bsr_non_real_kernel / _spmv_bsr_trans_real_kernel and the
 * blockrow_reduce kernels (spmv_bsr.py:183-221, 275-318, 376-405, 603-757). */
int fsp_bsr_spmv(int32_t value_type, int32_t ptr_type, int32_t index_type, const void* ptr_dev, const void* bidx_dev,
                 int32_t perm_type, const void* perm_dev, const void* data_dev, int32_t block_dim, const void* x_dev,
                 void* y_dev, int64_t n_out_blocks, void* stream);

/* ---- COO -> CSR ------------------------------------------------------------
 * Replaces coo_to_csr_for_spmv (spmv_csr.py:958-981: argsort of row * n_cols
 * + col, bincount, cumsum).  Stable: duplicates keep their input order.
 * Outputs: indptr [n_rows + 1], cols [nnz] (column of each sorted nonzero),
 * perm [nnz] (input position of each sorted nonzero; gather the values with
 * fsp_gather).  Row ids outside [0, n_rows) set bit 1 of *err_dev, columns
 * outside [0, n_cols) bit 2 (the outputs are then meaningless but every
 * write stays in bounds). */
int fsp_coo_to_csr_workspace(int64_t n_rows, int64_t n_cols, int64_t nnz, int32_t perm_type, int64_t* bytes);
int fsp_coo_to_csr(int32_t index_type, const void* row_dev, const void* col_dev, int64_t n_rows, int64_t n_cols,
                   int64_t nnz, int32_t indptr_type, void* indptr_dev, int32_t out_index_type, void* cols_dev,
                   int32_t perm_type, void* perm_dev, uint32_t* err_dev, void* ws_dev, int64_t ws_bytes,
                   void* stream);

/* ---- batches of CSR matrices --------------------------------------------
 * No counterpart in the reference (it runs one matrix per call).  A batch of
 * n_mats matrices is a packed offset table over concatenated arrays: matrix m
 * has rows [row_off[m], row_off[m+1]), columns [col_off[m], col_off[m+1]) and
 * nonzeros [nnz_off[m], nnz_off[m+1]); its own row pointer (local offsets from
 * 0, n_m + 1 entries) sits at local_indptr[row_off[m] + m].  fsp_batch_pack
 * writes the block-diagonal CSR of the batch: out_indptr [total_rows + 1]
 * (global offsets) and out_indices [total_nnz] (col_off[m] + local column).
 * The values are shared.  Violations (row pointer not starting at 0 / ending
 * at the matrix's nonzero count: bit 1; decreasing: bit 2; a column outside
 * its matrix: bit 4) are OR-ed into *err_dev.  The block-diagonal matrix then
 * runs through fsp_spmv_csr / fsp_spmm_csr as ONE launch for the batch
 * (plan.max_blocks > 0 makes it persistent). */
int fsp_batch_pack(int32_t indptr_type, int32_t index_type, int64_t n_mats, const int64_t* row_off_dev,
                   const int64_t* col_off_dev, const int64_t* nnz_off_dev, const void* local_indptr_dev,
                   const void* local_indices_dev, int64_t total_rows, int64_t total_nnz, int32_t out_indptr_type,
                   void* out_indptr_dev, int32_t out_index_type, void* out_indices_dev, uint32_t* err_dev,
                   void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FLAGSPARSE_HIP_H */
